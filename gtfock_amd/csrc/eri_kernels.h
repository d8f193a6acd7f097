// eri_kernels.h -- per-class ERI kernels with their epilogues.
//
//  EPI_JK    fused J/K digestion of the screened unique quartets; the
//            multiplicities are those of update_F (pfock/fock_task.c:29-117),
//            the accumulators hold 2J and -K before symmetrisation.
//  EPI_DIAG  Schwarz values max |(mn|mn)| of a shell pair
//            (pfock/screening.c:63-89).
//  EPI_STORE write the contracted integrals (for CInt_computeShellQuartet).
//
// Work is handed out by a device-wide atomic task counter (the GPU form of
// taskq_next, pfock/taskq.c:55-62): each workgroup takes 256 consecutive work
// units (unit = quartet * NSPLIT + bra-component chunk) until none are left.
#pragma once
#include "eri_md.h"

enum { EPI_JK = 0, EPI_DIAG = 1, EPI_STORE = 2 };

struct EriArgs {
    // work decode: either bra rows (row_off over bra_list x ket_list) or qlist
    long nunits;
    const long* row_off;
    int nrows;
    const int* bra_list;
    const int* ket_list;
    const int2* qlist;
    int* counter;   // task counter (reset to 0 before the launch)
    int ntasks;
    PairView pv;
    const int* fstart;
    const double* boys;
    // EPI_JK
    const double* D;
    double* J;
    double* K;
    int ld;
    // EPI_DIAG
    unsigned long long* vbits;
    // EPI_STORE
    double* out;
};

__device__ __forceinline__ int find_row(const long* __restrict__ off, int nrows, long q)
{
    int lo = 0, hi = nrows;  // off[lo] <= q < off[hi]
    while (hi - lo > 1) {
        int mid = (lo + hi) >> 1;
        if (off[mid] <= q) lo = mid;
        else hi = mid;
    }
    return lo;
}

template <int LA, int LB, int LC, int LD, int NC, int EPI>
__device__ __forceinline__ void eri_unit(const EriArgs& a, long u)
{
    using S = EriShape<LA, LB, LC, LD>;
    constexpr int NSPLIT = S::NAB / NC;
    constexpr int NCD = S::NCD, ND = S::ND, NB = S::NB;
    const long q = u / NSPLIT;
    const int c0 = (int)(u - q * NSPLIT) * NC;
    int bra, ket;
    if (a.qlist) {
        const int2 bk = a.qlist[q];
        bra = bk.x;
        ket = bk.y;
    } else {
        const int r = find_row(a.row_off, a.nrows, q);
        bra = a.bra_list[r];
        ket = a.ket_list[q - a.row_off[r]];
    }
    double acc[NC][NCD];
    eri_block<LA, LB, LC, LD, NC, double>(a.pv, bra, ket, c0, a.boys, acc);

    if constexpr (EPI == EPI_STORE) {
        double* o = a.out + q * (long)(S::NAB * NCD) + (long)c0 * NCD;
        static_for<0, NC>([&](auto c_) {
            constexpr int c = decltype(c_)::value;
            static_for<0, NCD>([&](auto d_) {
                constexpr int cd = decltype(d_)::value;
                o[c * NCD + cd] = acc[c][cd];
            });
        });
    } else if constexpr (EPI == EPI_DIAG) {
        double mx = 0.0;
        static_for<0, NC>([&](auto c_) {
            constexpr int c = decltype(c_)::value;
            static_for<0, NCD>([&](auto d_) {
                constexpr int cd = decltype(d_)::value;
                if (cd == c0 + c) mx = fmax(mx, fabs(acc[c][cd]));
            });
        });
        atomicMax(a.vbits + bra, (unsigned long long)__double_as_longlong(mx));
    } else {
        const int M = a.pv.s1[bra], N = a.pv.s2[bra], P = a.pv.s1[ket], Q = a.pv.s2[ket];
        const int f1 = (M != N), f2 = (P != Q), f3 = (bra != ket);
        const double cJ1 = 2.0 * (1 + f1) * (1 + f2);
        const double cJ2 = 2.0 * f3 * (1 + f1) * (1 + f2);
        const double cMP = (1 + f3), cNP = f1 * (1 + f3), cMQ = f2 * (1 + f3), cNQ = f1 * f2 * (1 + f3);
        const int ld = a.ld;
        const int oM = a.fstart[M], oN = a.fstart[N], oP = a.fstart[P], oQ = a.fstart[Q];
        const double* __restrict__ D = a.D;
        double* J = a.J;
        double* K = a.K;
        int mrow[NC], ncol[NC];
        double dmn[NC];
        static_for<0, NC>([&](auto c_) {
            constexpr int c = decltype(c_)::value;
            const int ab = c0 + c;
            mrow[c] = oM + ab / NB;
            ncol[c] = oN + ab % NB;
            dmn[c] = D[(long)mrow[c] * ld + ncol[c]];
        });
        // J_MN += cJ1 * sum_pq D_pq I   and   J_PQ += cJ2 * sum_mn D_mn I
        static_for<0, NC>([&](auto c_) {
            constexpr int c = decltype(c_)::value;
            double s = 0.0;
            static_for<0, NCD>([&](auto d_) {
                constexpr int cd = decltype(d_)::value;
                s += D[(long)(oP + cd / ND) * ld + oQ + cd % ND] * acc[c][cd];
            });
            atomicAdd(&J[(long)mrow[c] * ld + ncol[c]], cJ1 * s);
        });
        if (f3) {
            static_for<0, NCD>([&](auto d_) {
                constexpr int cd = decltype(d_)::value;
                double s = 0.0;
                static_for<0, NC>([&](auto c_) {
                    constexpr int c = decltype(c_)::value;
                    s += dmn[c] * acc[c][cd];
                });
                atomicAdd(&J[(long)(oP + cd / ND) * ld + oQ + cd % ND], cJ2 * s);
            });
        }
        // K_MP -= cMP sum_q D_nq I,  K_NP -= cNP sum_q D_mq I
        // K_MQ -= cMQ sum_p D_np I,  K_NQ -= cNQ sum_p D_mp I
        static_for<0, NC>([&](auto c_) {
            constexpr int c = decltype(c_)::value;
            const long rm = (long)mrow[c] * ld, rn = (long)ncol[c] * ld;
            static_for<0, S::NCc>([&](auto p_) {
                constexpr int p = decltype(p_)::value;
                double smp = 0.0, snp = 0.0;
                static_for<0, ND>([&](auto q_) {
                    constexpr int qq = decltype(q_)::value;
                    const double x = acc[c][p * ND + qq];
                    smp += D[rn + oQ + qq] * x;
                    snp += D[rm + oQ + qq] * x;
                });
                atomicAdd(&K[rm + oP + p], -cMP * smp);
                if (f1) atomicAdd(&K[rn + oP + p], -cNP * snp);
            });
            if (f2) {
                static_for<0, ND>([&](auto q_) {
                    constexpr int qq = decltype(q_)::value;
                    double smq = 0.0, snq = 0.0;
                    static_for<0, S::NCc>([&](auto p_) {
                        constexpr int p = decltype(p_)::value;
                        const double x = acc[c][p * ND + qq];
                        smq += D[rn + oP + p] * x;
                        snq += D[rm + oP + p] * x;
                    });
                    atomicAdd(&K[rm + oQ + qq], -cMQ * smq);
                    if (f1) atomicAdd(&K[rn + oQ + qq], -cNQ * snq);
                });
            }
        });
    }
}

template <int LA, int LB, int LC, int LD, int NC, int EPI>
__global__ __launch_bounds__(256) void eri_kernel(EriArgs a)
{
    if (a.counter) {
        __shared__ int s_task;
        for (;;) {
            if (threadIdx.x == 0) s_task = atomicAdd(a.counter, 1);
            __syncthreads();
            const int task = s_task;
            __syncthreads();
            if (task >= a.ntasks) break;
            const long u = (long)task * 256 + threadIdx.x;
            if (u < a.nunits) eri_unit<LA, LB, LC, LD, NC, EPI>(a, u);
        }
    } else {
        const long u = (long)blockIdx.x * 256 + threadIdx.x;
        if (u < a.nunits) eri_unit<LA, LB, LC, LD, NC, EPI>(a, u);
    }
}

// launch helpers (eri_launch.hip)
typedef void (*eri_kernel_t)(EriArgs);
struct EriClassInfo {
    int la, lb, lc, ld;
    int nc;         // bra components per thread
    int nsplit;     // threads per quartet
    int nints;      // integrals per quartet
    eri_kernel_t jk, diag, store;
};
const EriClassInfo* gtf_eri_class(int cls);
