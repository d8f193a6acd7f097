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

#ifndef GTF_ATOMIC_ADD
#define GTF_ATOMIC_ADD(p, v) atomicAdd((p), (v))
#endif

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
    const double* boys8;  // compact Taylor tables, BOYS8_NM slabs of BOYS8_SLAB
    // EPI_JK: nd densities (and their J, K) dstride elements apart, all
    // digested from each integral (pfock/fock_task.c:46 loops num_dmat)
    const double* D;
    double* J;
    double* K;
    int nd;
    long dstride;
    // EPI_DIAG
    unsigned long long* vbits;
    // EPI_STORE
    double* out;
    // bra-row tiles of the LDS-strip J/K kernel: (bra row, first ket, end ket)
    const int4* tiles;
    int ntiles;
    int nbf;  // LDS strip capacity (largest molecule)
    // exact screening of the block-padded rows
    const double* vbra;
    const double* vket;
    const int* rank_bra;
    const int* rank_ket;
    double tol2;
    int same;
    // mixed precision: FP64 quartets have v_bra * v_ket >= tau2, FP32 the rest
    double tau2;
    // packed per-molecule matrices: block offset and leading dimension per shell
    const long* mbase;
    const int* mld;
    // global-atomic path: bra pair (index into bra_list) and absolute first
    // ket of each row
    const int* row_bra;
    const int* row_kstart;
    // batches: molecule of each shell and, per molecule, whether its J/K are
    // built (0: its quartets are skipped; a molecule whose SCF has ended)
    const int* shell_mol;
    const int* mol_active;
    // tile kernels (optional): per workgroup {clock at start, clock at exit,
    // tiles taken}, three entries at 3 * blockIdx.x -- the load balance of a
    // launch (pfock/pfock.c:1918-2010 reports it across processes)
    long long* wg_stat;
};

// one workgroup's record of its part in a tile launch
__device__ __forceinline__ void record_workgroup(const EriArgs& a, long long t_begin, int taken)
{
    if (a.wg_stat && threadIdx.x == 0) {
        long long* w = a.wg_stat + 3L * blockIdx.x;
        w[0] = t_begin;
        w[1] = (long long)wall_clock64();
        w[2] = taken;
    }
}

// the bra pair's molecule takes part in this build
__device__ __forceinline__ bool bra_active(const EriArgs& a, int bra)
{
    return !a.mol_active || a.mol_active[a.shell_mol[a.pv.s1[bra]]];
}

// the Schwarz test of pfock/fock_task.c:192 and, for a diagonal class, the
// one-orientation rule, for bra row r and ket slot j
__device__ __forceinline__ bool quartet_valid(const EriArgs& a, int r, int j)
{
    return a.vbra[r] * a.vket[j] >= a.tol2 && (!a.same || a.rank_ket[j] <= a.rank_bra[r]);
}

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

template <int LA, int LB, int LC, int LD, int NC, bool STRIP>
__device__ __forceinline__ void digest(const EriArgs& a, int bra, int ket, int c0,
                                       const double (&acc)[NC][EriShape<LA, LB, LC, LD>::NCD], double* jmn,
                                       double* strip, int nbf, double* jpq, double* jmn_reg = nullptr,
                                       long doff = 0);

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
        const int row = find_row(a.row_off, a.nrows, q);
        const int r = a.row_bra[row];
        const int j = a.row_kstart[row] + (int)(q - a.row_off[row]);
        if (!quartet_valid(a, r, j)) return;
        bra = a.bra_list[r];
        ket = a.ket_list[j];
    }
    if constexpr (EPI == EPI_JK) {
        if (!bra_active(a, bra)) return;
    }
    double acc[NC][NCD];
    eri_block<LA, LB, LC, LD, NC, double>(a.pv, bra, ket, c0, a.boys8 + (LA + LB + LC + LD) * BOYS8_SLAB, acc);

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
        for (int d = 0; d < a.nd; d++)
            digest<LA, LB, LC, LD, NC, false>(a, bra, ket, c0, acc, nullptr, nullptr, 0, nullptr, nullptr,
                                              d * a.dstride);
    }
}

// update_F digestion of one thread's NC bra components of quartet (bra|ket).
// STRIP = false: FP64 atomics straight into J and K.
// STRIP = true:  the bra-side blocks (J_MN and the exchange blocks K_MP, K_NP,
// K_MQ, K_NQ, which all lie in the rows of shells M and N) go to LDS
// (jmn[NA*NB], strip[(NA+NB) * nbf], row r < NA is function fM + r, row
// NA + r is fN + r); J_PQ goes to the ket's LDS row jpq[NCD] when one is
// given (the threads sharing a quartet then meet in LDS), else to HBM.  With
// jmn_reg (a thread that owns all bra components of its quartets) J_MN is
// summed in those registers and added to LDS once per tile.
template <int LA, int LB, int LC, int LD, int NC, bool STRIP>
__device__ __forceinline__ void digest(const EriArgs& a, int bra, int ket, int c0,
                                       const double (&acc)[NC][EriShape<LA, LB, LC, LD>::NCD], double* jmn,
                                       double* strip, int nbf, double* jpq, double* jmn_reg, long doff)
{
    using S = EriShape<LA, LB, LC, LD>;
    constexpr int NCD = S::NCD, ND = S::ND, NB = S::NB, NA = S::NA;
    const int M = a.pv.s1[bra], N = a.pv.s2[bra], P = a.pv.s1[ket], Q = a.pv.s2[ket];
    const int f1 = (M != N), f2 = (P != Q), f3 = (bra != ket);
    const double cJ1 = 2.0 * (1 + f1) * (1 + f2);
    const double cJ2 = 2.0 * f3 * (1 + f1) * (1 + f2);
    const double cMP = (1 + f3), cNP = f1 * (1 + f3), cMQ = f2 * (1 + f3), cNQ = f1 * f2 * (1 + f3);
    const int ld = a.mld[M];
    const long base = a.mbase[M];
    const int oM = a.fstart[M], oN = a.fstart[N], oP = a.fstart[P], oQ = a.fstart[Q];
    const double* __restrict__ D = a.D + doff + base;
    double* J = a.J + doff + base;
    double* K = a.K + doff + base;
    int mloc[NC], nloc[NC];
    double dmn[NC];
    static_for<0, NC>([&](auto c_) {
        constexpr int c = decltype(c_)::value;
        const int ab = c0 + c;
        mloc[c] = ab / NB;
        nloc[c] = ab % NB;
        dmn[c] = D[(long)(oM + mloc[c]) * ld + oN + nloc[c]];
    });
    // J_MN += cJ1 * sum_pq D_pq I   and   J_PQ += cJ2 * sum_mn D_mn I
    static_for<0, NC>([&](auto c_) {
        constexpr int c = decltype(c_)::value;
        double s = 0.0;
        static_for<0, NCD>([&](auto d_) {
            constexpr int cd = decltype(d_)::value;
            s += D[(long)(oP + cd / ND) * ld + oQ + cd % ND] * acc[c][cd];
        });
        if constexpr (STRIP) {
            if (jmn_reg) jmn_reg[c] += cJ1 * s;
            else atomicAdd(&jmn[mloc[c] * NB + nloc[c]], cJ1 * s);
        } else {
            GTF_ATOMIC_ADD(&J[(long)(oM + mloc[c]) * ld + oN + nloc[c]], cJ1 * s);
        }
    });
    if (f3) {
        static_for<0, NCD>([&](auto d_) {
            constexpr int cd = decltype(d_)::value;
            double s = 0.0;
            static_for<0, NC>([&](auto c_) {
                constexpr int c = decltype(c_)::value;
                s += dmn[c] * acc[c][cd];
            });
            if (jpq) {
                atomicAdd(&jpq[cd], cJ2 * s);
            } else {
                GTF_ATOMIC_ADD(&J[(long)(oP + cd / ND) * ld + oQ + cd % ND], cJ2 * s);
            }
        });
    }
    // K_MP -= cMP sum_q D_nq I,  K_NP -= cNP sum_q D_mq I
    // K_MQ -= cMQ sum_p D_np I,  K_NQ -= cNQ sum_p D_mp I
    static_for<0, NC>([&](auto c_) {
        constexpr int c = decltype(c_)::value;
        const long rm = (long)(oM + mloc[c]) * ld, rn = (long)(oN + nloc[c]) * ld;
        double* km = STRIP ? strip + (long)mloc[c] * nbf : K + rm;
        double* kn = STRIP ? strip + (long)(NA + nloc[c]) * nbf : K + rn;
        static_for<0, S::NCc>([&](auto p_) {
            constexpr int p = decltype(p_)::value;
            double smp = 0.0, snp = 0.0;
            static_for<0, ND>([&](auto q_) {
                constexpr int qq = decltype(q_)::value;
                const double x = acc[c][p * ND + qq];
                smp += D[rn + oQ + qq] * x;
                snp += D[rm + oQ + qq] * x;
            });
            if constexpr (STRIP) {
                atomicAdd(&km[oP + p], -cMP * smp);
                if (f1) atomicAdd(&kn[oP + p], -cNP * snp);
            } else {
                GTF_ATOMIC_ADD(&km[oP + p], -cMP * smp);
                if (f1) GTF_ATOMIC_ADD(&kn[oP + p], -cNP * snp);
            }
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
                if constexpr (STRIP) {
                    atomicAdd(&km[oQ + qq], -cMQ * smq);
                    if (f1) atomicAdd(&kn[oQ + qq], -cNQ * snq);
                } else {
                    GTF_ATOMIC_ADD(&km[oQ + qq], -cMQ * smq);
                    if (f1) GTF_ATOMIC_ADD(&kn[oQ + qq], -cNQ * snq);
                }
            });
        }
    });
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

// Fused ERI + J/K digestion over bra-row tiles.  A tile is a bounded number
// of work units of one bra row; the workgroup accumulates J_MN and
// the exchange blocks of the bra shells in LDS and flushes them with one HBM
// atomic per touched element when the tile is done.  The J_PQ block of
// every ket of the tile is also kept in LDS and flushed at the end, so no
// HBM atomic (whose completion the next loads would wait for) sits in the
// quartet loop.  Dynamic LDS: GTF_BOYS_LDS (the Boys slab, BLDS only) +
// (NA + NB) * nbf + NA * NB + kmax * NCD doubles, nbf = the largest molecule's size, kmax =
// the most kets a tile of the class may hold (units per tile / NSPLIT).
//
// A tile is (bra row, first ket, end ket, nic): each of its quartets is split
// into nic bra-primitive chunks (sizes differing by at most one) times NSPLIT
// bra-component chunks, so
// that a unit holds about unit_pq primitive quartets however contracted
// the pairs are, and a tile holds at most `units` units and about tile_pq
// primitive quartets (per class, chosen on the host; the defaults below).
// Digestion is linear in the integrals, so every unit digests its partial
// sum on its own.
#define GTF_TILE_UNITS 1024
#define GTF_UNIT_PQ 64
#define GTF_TILE_PQ 65536

// The Boys Taylor slab of order L copied from HBM into the first
// GTF_BOYS_LDS doubles of dynamic LDS.  A persistent workgroup copies it
// once; every primitive quartet of every tile it takes then reads its Taylor
// row from LDS instead of a dependent global load.  The barrier that follows
// each tile's clearing of its LDS work area orders the copy before the first
// read.  It costs 23 KB of LDS per workgroup, which lowers the occupancy of
// classes whose strips are already large, so it is a per-class choice (the
// tile shape's boys_lds).  FP32 tiles use the same FP64 slab: the Boys
// function is FP64 in both precisions.
#define GTF_BOYS_LDS BOYS8_SLAB
template <int L>
__device__ __forceinline__ const double* boys_to_lds(const EriArgs& a, double* smem)
{
    const double* src = a.boys8 + L * BOYS8_SLAB;
    for (int k = threadIdx.x; k < BOYS8_SLAB; k += 256) smem[k] = src[k];
    return smem;
}

// Boys slab of a tile kernel: in LDS (BLDS; the work area then starts after
// it) or read in place from HBM through the caches (the work area starts at
// the LDS base)
template <int L, bool BLDS>
__device__ __forceinline__ const double* tile_boys(const EriArgs& a, double* smem, double*& work)
{
    if constexpr (BLDS) {
        work = smem + GTF_BOYS_LDS;
        return boys_to_lds<L>(a, smem);
    } else {
        work = smem;
        return a.boys8 + L * BOYS8_SLAB;
    }
}

// Minimum waves per SIMD asked of the compiler for a class's tile kernel.
// (ps|ss) in FP64 takes 152 VGPRs (3 waves) when left alone and fits 127
// with none spilled at 4, 8 % faster on MI355X; the same cap makes (pp|ss)
// and (pp|pp) spill and run 2.5x slower, so the others are left alone.
template <int LA, int LB, int LC, int LD>
constexpr int tile_min_waves()
{
    return (LA == 1 && LB == 0 && LC == 0 && LD == 0) ? 4 : 1;
}

template <int LA, int LB, int LC, int LD, int NC, typename Real, bool BLDS>
__global__ __launch_bounds__(256, (tile_min_waves<LA, LB, LC, LD>())) void eri_jk_tile_kernel(EriArgs a)
{
    using S = EriShape<LA, LB, LC, LD>;
    constexpr int NSPLIT = S::NAB / NC;
    constexpr int NA = S::NA, NB = S::NB, NCD = S::NCD;
    constexpr bool LOW = std::is_same<Real, float>::value;
    constexpr bool JPQ_LDS = true;
    constexpr int L = LA + LB + LC + LD;
    constexpr int ND = S::ND;
    extern __shared__ double smem[];
    __shared__ int s_task;
    double* work;
    const double* boys = tile_boys<L, BLDS>(a, smem, work);
    const long long t_begin = (long long)wall_clock64();
    int taken = 0;
    for (;;) {
        if (threadIdx.x == 0) s_task = atomicAdd(a.counter, 1);
        __syncthreads();
        const int task = s_task;
        if (task >= a.ntiles) break;
        taken++;
        const int4 t = a.tiles[task];
        const int bra = a.bra_list[t.x];
        if (!bra_active(a, bra)) {
            __syncthreads();  // every thread has read s_task before it is rewritten
            continue;
        }
        const int nbp = a.pv.nprim[bra];
        const int nic = t.w;  // bra-primitive chunks per quartet
        // strip rows are as wide as the bra's molecule (at most a.nbf)
        const int nbf = a.mld[a.pv.s1[bra]];
        const int nstrip = (NA + NB) * nbf;
        double* strip = work;
        double* jmn = work + nstrip;
        double* jpq = jmn + NA * NB;
        const int njpq = JPQ_LDS ? (t.z - t.y) * NCD : 0;
        // one work area (strip, J_MN, J_PQ) per density
        const int wsz = nstrip + NA * NB + njpq;
        for (int k = threadIdx.x; k < a.nd * wsz; k += 256) work[k] = 0.0;
        __syncthreads();
        // Unit order.  NSPLIT == 1: the nic chunks of one quartet sit in
        // neighbouring lanes (nic is a power of two, so a group never spans
        // two wavefronts); the group adds its partial integrals with lane
        // shuffles and its first lane digests the sum.  NSPLIT > 1: component
        // chunk fastest, then ket, then bra-primitive chunk, so neighbouring
        // lanes digest into different ket columns.
        const int nkt = t.z - t.y;
        const int nunits = nkt * nic * NSPLIT;
        double jreg[NC];
        static_for<0, NC>([&](auto c_) { jreg[decltype(c_)::value] = 0.0; });
        for (int u = threadIdx.x; u < nunits; u += 256) {
            int q, ic, c0;
            if constexpr (NSPLIT == 1) {
                q = u / nic;
                ic = u - q * nic;
                c0 = 0;
            } else {
                const int rest = u / NSPLIT;
                c0 = (u - rest * NSPLIT) * NC;
                ic = rest / nkt;
                q = rest - ic * nkt;
            }
            const int i0 = ic * nbp / nic, i1 = (ic + 1) * nbp / nic;
            const int j = t.y + q;
            if (!quartet_valid(a, t.x, j)) continue;
            // precision split on the Schwarz bound sqrt(v_bra v_ket) vs tau
            if ((a.vbra[t.x] * a.vket[j] >= a.tau2) == LOW) continue;
            const int ket = a.ket_list[j];
            double acc[NC][NCD];
            if constexpr (LOW) {
                float accf[NC][NCD];
                eri_block<LA, LB, LC, LD, NC, float>(a.pv, bra, ket, c0, boys, accf, i0, i1);
                static_for<0, NC>([&](auto c_) {
                    static_for<0, NCD>([&](auto d_) {
                        acc[decltype(c_)::value][decltype(d_)::value] = (double)accf[decltype(c_)::value][decltype(d_)::value];
                    });
                });
            } else {
                eri_block<LA, LB, LC, LD, NC, double>(a.pv, bra, ket, c0, boys, acc, i0, i1);
            }
            if constexpr (NSPLIT == 1) {
                // sum the group's partial integrals into its first lane
                for (int w = 1; w < nic; w <<= 1)
                    static_for<0, NC>([&](auto c_) {
                        static_for<0, NCD>([&](auto d_) {
                            constexpr int c = decltype(c_)::value, cd = decltype(d_)::value;
                            acc[c][cd] += __shfl_xor(acc[c][cd], w, 64);
                        });
                    });
                if (ic != 0) continue;
            }
            for (int d = 0; d < a.nd; d++)
                digest<LA, LB, LC, LD, NC, true>(a, bra, ket, c0, acc, jmn + d * wsz, strip + d * wsz, nbf,
                                                 JPQ_LDS ? jpq + d * wsz + q * NCD : nullptr,
                                                 NSPLIT == 1 && a.nd == 1 ? jreg : nullptr, d * a.dstride);
        }
        if constexpr (NSPLIT == 1) {
            static_for<0, NC>([&](auto c_) {
                constexpr int c = decltype(c_)::value;
                if (jreg[c] != 0.0) atomicAdd(&jmn[(c / NB) * NB + c % NB], jreg[c]);
            });
        }
        __syncthreads();
        const int M = a.pv.s1[bra], N = a.pv.s2[bra];
        const int oM = a.fstart[M], oN = a.fstart[N];
        const long base = a.mbase[M];
        const int ld = a.mld[M];
        for (int d = 0; d < a.nd; d++) {
            const double* sd = strip + d * wsz;
            const double* jd = jmn + d * wsz;
            double* Kd = a.K + d * a.dstride + base;
            double* Jd = a.J + d * a.dstride + base;
            for (int k = threadIdx.x; k < nstrip; k += 256) {
                const double v = sd[k];
                if (v != 0.0) {
                    const int r = k / nbf, col = k - r * nbf;
                    const int row = r < NA ? oM + r : oN + (r - NA);
                    GTF_ATOMIC_ADD(&Kd[(long)row * ld + col], v);
                }
            }
            for (int k = threadIdx.x; k < NA * NB; k += 256) {
                const double v = jd[k];
                if (v != 0.0) GTF_ATOMIC_ADD(&Jd[(long)(oM + k / NB) * ld + oN + k % NB], v);
            }
            if constexpr (JPQ_LDS) {
                const double* pd = jpq + d * wsz;
                for (int k = threadIdx.x; k < njpq; k += 256) {
                    const double v = pd[k];
                    if (v != 0.0) {
                        const int q = k / NCD, cd = k - q * NCD;
                        const int ket = a.ket_list[t.y + q];
                        const int oP = a.fstart[a.pv.s1[ket]], oQ = a.fstart[a.pv.s2[ket]];
                        GTF_ATOMIC_ADD(&Jd[(long)(oP + cd / ND) * ld + oQ + cd % ND], v);
                    }
                }
            }
        }
        __syncthreads();
    }
    record_workgroup(a, t_begin, taken);
}

// update_F digestion of a quartet evaluated with bra and ket exchanged: the
// thread holds (cd | ab) for its NCT ket components cd0.. and every bra
// component ab, acc[k][ab].  Same contributions as digest<..., true>: J_MN
// summed in jmn_reg[NAB] (registers, all bra components), J_PQ to the ket's
// LDS row, the exchange blocks K_MP, K_NP, K_MQ, K_NQ to the bra rows of the
// LDS strip.
template <int LA, int LB, int LC, int LD, int NCT>
__device__ __forceinline__ void digest_t(const EriArgs& a, int bra, int ket, int cd0,
                                         const double (&acc)[NCT][EriShape<LA, LB, LC, LD>::NAB],
                                         double* jmn_reg, double* strip, int nbf, double* jpq, double* jmn_lds,
                                         long doff)
{
    using S = EriShape<LA, LB, LC, LD>;
    constexpr int NAB = S::NAB, NA = S::NA, NB = S::NB, ND = S::ND;
    const int M = a.pv.s1[bra], N = a.pv.s2[bra], P = a.pv.s1[ket], Q = a.pv.s2[ket];
    const int f1 = (M != N), f2 = (P != Q), f3 = (bra != ket);
    const double cJ1 = 2.0 * (1 + f1) * (1 + f2);
    const double cJ2 = 2.0 * f3 * (1 + f1) * (1 + f2);
    const double cMP = (1 + f3), cNP = f1 * (1 + f3), cMQ = f2 * (1 + f3), cNQ = f1 * f2 * (1 + f3);
    const int ld = a.mld[M];
    const long base = a.mbase[M];
    const int oM = a.fstart[M], oN = a.fstart[N], oP = a.fstart[P], oQ = a.fstart[Q];
    const double* __restrict__ D = a.D + doff + base;
    static_for<0, NCT>([&](auto k_) {
        constexpr int k = decltype(k_)::value;
        const int cd = cd0 + k;
        const int p = cd / ND, q = cd - (cd / ND) * ND;
        const double dpq = D[(long)(oP + p) * ld + oQ + q];
        // J_MN += cJ1 D_pq I (registers, or LDS with several densities) and
        // J_PQ += cJ2 sum_mn D_mn I
        double jpq_s = 0.0;
        static_for<0, NAB>([&](auto ab_) {
            constexpr int ab = decltype(ab_)::value;
            constexpr int m = ab / NB, n = ab % NB;
            if (jmn_lds) atomicAdd(&jmn_lds[ab], cJ1 * dpq * acc[k][ab]);
            else jmn_reg[ab] += cJ1 * dpq * acc[k][ab];
            jpq_s += D[(long)(oM + m) * ld + oN + n] * acc[k][ab];
        });
        if (f3) atomicAdd(&jpq[cd], cJ2 * jpq_s);
        // K_MP[m][p] -= cMP sum_n D_nq I,  K_MQ[m][q] -= cMQ sum_n D_np I
        static_for<0, NA>([&](auto m_) {
            constexpr int m = decltype(m_)::value;
            double smp = 0.0, smq = 0.0;
            static_for<0, NB>([&](auto n_) {
                constexpr int n = decltype(n_)::value;
                const double x = acc[k][m * NB + n];
                smp += D[(long)(oN + n) * ld + oQ + q] * x;
                smq += D[(long)(oN + n) * ld + oP + p] * x;
            });
            atomicAdd(&strip[m * nbf + oP + p], -cMP * smp);
            if (f2) atomicAdd(&strip[m * nbf + oQ + q], -cMQ * smq);
        });
        // K_NP[n][p] -= cNP sum_m D_mq I,  K_NQ[n][q] -= cNQ sum_m D_mp I
        if (f1) {
            static_for<0, NB>([&](auto n_) {
                constexpr int n = decltype(n_)::value;
                double snp = 0.0, snq = 0.0;
                static_for<0, NA>([&](auto m_) {
                    constexpr int m = decltype(m_)::value;
                    const double x = acc[k][m * NB + n];
                    snp += D[(long)(oM + m) * ld + oQ + q] * x;
                    snq += D[(long)(oM + m) * ld + oP + p] * x;
                });
                atomicAdd(&strip[(NA + n) * nbf + oP + p], -cNP * snp);
                if (f2) atomicAdd(&strip[(NA + n) * nbf + oQ + q], -cNQ * snq);
            });
        }
    });
}

// The bra-row tile kernel for classes evaluated with bra and ket exchanged:
// a thread owns NCT ket components and every bra component of its quartet,
// so the Boys function and the R recursion of a primitive quartet are
// computed by NCD / NCT threads instead of NAB / NC.  Units: (ket, ket
// component chunk), no bra-primitive chunks.  LDS as eri_jk_tile_kernel.
template <int LA, int LB, int LC, int LD, int NCT, typename Real, bool BLDS>
__global__ __launch_bounds__(256) void eri_jk_tile_t_kernel(EriArgs a)
{
    using S = EriShape<LA, LB, LC, LD>;
    constexpr int NSPLIT = S::NCD / NCT;
    constexpr int NA = S::NA, NB = S::NB, NCD = S::NCD, NAB = S::NAB, ND = S::ND;
    constexpr bool LOW = std::is_same<Real, float>::value;
    constexpr int L = LA + LB + LC + LD;
    extern __shared__ double smem[];
    __shared__ int s_task;
    double* work;
    const double* boys = tile_boys<L, BLDS>(a, smem, work);
    const long long t_begin = (long long)wall_clock64();
    int taken = 0;
    for (;;) {
        if (threadIdx.x == 0) s_task = atomicAdd(a.counter, 1);
        __syncthreads();
        const int task = s_task;
        if (task >= a.ntiles) break;
        taken++;
        const int4 t = a.tiles[task];
        const int bra = a.bra_list[t.x];
        if (!bra_active(a, bra)) {
            __syncthreads();  // every thread has read s_task before it is rewritten
            continue;
        }
        const int nbf = a.mld[a.pv.s1[bra]];
        const int nstrip = (NA + NB) * nbf;
        double* strip = work;
        double* jmn = work + nstrip;
        double* jpq = jmn + NA * NB;
        const int nkt = t.z - t.y;
        const int njpq = nkt * NCD;
        const int wsz = nstrip + NA * NB + njpq;  // work area per density
        for (int k = threadIdx.x; k < a.nd * wsz; k += 256) work[k] = 0.0;
        __syncthreads();
        double jreg[NAB];
        static_for<0, NAB>([&](auto c_) { jreg[decltype(c_)::value] = 0.0; });
        const int nunits = nkt * NSPLIT;
        for (int u = threadIdx.x; u < nunits; u += 256) {
            const int q = u / NSPLIT;
            const int cd0 = (u - q * NSPLIT) * NCT;
            const int j = t.y + q;
            if (!quartet_valid(a, t.x, j)) continue;
            if ((a.vbra[t.x] * a.vket[j] >= a.tau2) == LOW) continue;
            const int ket = a.ket_list[j];
            double acc[NCT][NAB];
            if constexpr (LOW) {
                float accf[NCT][NAB];
                eri_block<LC, LD, LA, LB, NCT, float>(a.pv, ket, bra, cd0, boys, accf);
                static_for<0, NCT>([&](auto c_) {
                    static_for<0, NAB>([&](auto d_) {
                        acc[decltype(c_)::value][decltype(d_)::value] =
                            (double)accf[decltype(c_)::value][decltype(d_)::value];
                    });
                });
            } else {
                eri_block<LC, LD, LA, LB, NCT, double>(a.pv, ket, bra, cd0, boys, acc);
            }
            for (int d = 0; d < a.nd; d++)
                digest_t<LA, LB, LC, LD, NCT>(a, bra, ket, cd0, acc, jreg, strip + d * wsz, nbf,
                                              jpq + d * wsz + q * NCD, a.nd == 1 ? nullptr : jmn + d * wsz,
                                              d * a.dstride);
        }
        if (a.nd == 1)
            static_for<0, NAB>([&](auto c_) {
                constexpr int c = decltype(c_)::value;
                if (jreg[c] != 0.0) atomicAdd(&jmn[c], jreg[c]);
            });
        __syncthreads();
        const int M = a.pv.s1[bra], N = a.pv.s2[bra];
        const int oM = a.fstart[M], oN = a.fstart[N];
        const long base = a.mbase[M];
        const int ld = a.mld[M];
        for (int d = 0; d < a.nd; d++) {
            const double* sd = strip + d * wsz;
            const double* jd = jmn + d * wsz;
            const double* pd = jpq + d * wsz;
            double* Kd = a.K + d * a.dstride + base;
            double* Jd = a.J + d * a.dstride + base;
            for (int k = threadIdx.x; k < nstrip; k += 256) {
                const double v = sd[k];
                if (v != 0.0) {
                    const int r = k / nbf, col = k - r * nbf;
                    const int row = r < NA ? oM + r : oN + (r - NA);
                    GTF_ATOMIC_ADD(&Kd[(long)row * ld + col], v);
                }
            }
            for (int k = threadIdx.x; k < NA * NB; k += 256) {
                const double v = jd[k];
                if (v != 0.0) GTF_ATOMIC_ADD(&Jd[(long)(oM + k / NB) * ld + oN + k % NB], v);
            }
            for (int k = threadIdx.x; k < njpq; k += 256) {
                const double v = pd[k];
                if (v != 0.0) {
                    const int q = k / NCD, cd = k - q * NCD;
                    const int ket = a.ket_list[t.y + q];
                    const int oP = a.fstart[a.pv.s1[ket]], oQ = a.fstart[a.pv.s2[ket]];
                    GTF_ATOMIC_ADD(&Jd[(long)(oP + cd / ND) * ld + oQ + cd % ND], v);
                }
            }
        }
        __syncthreads();
    }
    record_workgroup(a, t_begin, taken);
}

// launch helpers (eri_table.cpp)
typedef void (*eri_kernel_t)(EriArgs);
struct EriClassInfo {
    int la, lb, lc, ld;
    int nc;         // bra components per thread
    int nsplit;     // threads per quartet
    int nints;      // integrals per quartet
    eri_kernel_t jk, diag, store;
    // tile kernels, [0] Boys slab read from HBM, [1] staged in LDS
    eri_kernel_t jk_tile[2];     // FP64 quartets (v_bra v_ket >= tau2)
    eri_kernel_t jk_tile_lo[2];  // FP32 quartets (v_bra v_ket < tau2)
    // classes evaluated with bra and ket exchanged (nsplit_t > 0): threads
    // per quartet and the two tile kernels
    int nsplit_t;
    eri_kernel_t jk_tile_t[2], jk_tile_t_lo[2];
};
const EriClassInfo* gtf_eri_class(int cls);
