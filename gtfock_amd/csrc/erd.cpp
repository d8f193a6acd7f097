// erd.cpp -- the CInt ERD / OED handles (libcint/include/CInt.h:90-181) on
// the GPU: shell-quartet integrals come from the same per-class McMurchie-
// Davidson kernels as the Fock build (EPI_STORE epilogue), one-electron blocks
// from the one-electron kernel.
#include <algorithm>
#include <cstring>
#include <new>
#include <vector>

#include "../../include/CInt.h"
#include "../../include/gtfock_amd.h"
#include "engine.h"
#include "oneel.h"
#include "screen.h"

struct ERD {
    BasisSet* basis = nullptr;
    GtfSystem sys;
    hipStream_t stream = nullptr;
    std::vector<std::vector<double>> bufs;  // per thread id
    double* d_out = nullptr;
    long out_cap = 0;
    int2* d_q = nullptr;
    long q_cap = 0;
};

struct OED {
    int nbf = 0;
    std::vector<int> fstart;
    std::vector<double> S, T, V;
    std::vector<double> buf;
};

static int erd_reserve(ERD* e, long nq, long nout)
{
    if (nq > e->q_cap) {
        if (e->d_q) (void)hipFree(e->d_q);
        e->d_q = nullptr;
        if (hipMalloc(&e->d_q, sizeof(int2) * nq) != hipSuccess) return -2;
        e->q_cap = nq;
    }
    if (nout > e->out_cap) {
        if (e->d_out) (void)hipFree(e->d_out);
        e->d_out = nullptr;
        if (hipMalloc(&e->d_out, sizeof(double) * nout) != hipSuccess) return -2;
        e->out_cap = nout;
    }
    return 0;
}

// Integrals of nq quartets (M,N,P,Q), each in the reference layout
// iM + dM*(iN + dN*(iP + dP*iQ)), concatenated in input order.
static int erd_batch(ERD* e, int nq, const int* q, double* out)
{
    GtfSystem& s = e->sys;
    const BasisSet* b = e->basis;
    std::vector<long> outoff(nq + 1, 0);
    for (int i = 0; i < nq; i++) {
        long d = 1;
        for (int k = 0; k < 4; k++) {
            const int sh = q[4 * i + k];
            if (sh < 0 || sh >= s.nshells) return -3;
            d *= b->fstart[sh + 1] - b->fstart[sh];
        }
        outoff[i + 1] = outoff[i] + d;
    }
    std::vector<std::vector<int>> members(GTF_NCLASSES);
    std::vector<int2> canon(nq);
    std::vector<char> swapped(nq);
    for (int i = 0; i < nq; i++) {
        int p1 = s.pair_id(q[4 * i], q[4 * i + 1]), p2 = s.pair_id(q[4 * i + 2], q[4 * i + 3]);
        int t1 = s.h_type[p1], t2 = s.h_type[p2];
        swapped[i] = t1 < t2;
        if (swapped[i]) { std::swap(p1, p2); std::swap(t1, t2); }
        canon[i] = make_int2(p1, p2);
        members[class_index(t1, t2)].push_back(i);
    }
    for (int cls = 0; cls < GTF_NCLASSES; cls++) {
        const std::vector<int>& mem = members[cls];
        if (mem.empty()) continue;
        const EriClassInfo* ci = gtf_eri_class(cls);
        const long n = (long)mem.size(), nint = ci->nints;
        if (erd_reserve(e, n, n * nint)) return -2;
        std::vector<int2> ql(n);
        for (long k = 0; k < n; k++) ql[k] = canon[mem[k]];
        if (hipMemcpyAsync(e->d_q, ql.data(), sizeof(int2) * n, hipMemcpyHostToDevice, e->stream) != hipSuccess)
            return -1;
        if (s.store(cls, (int)n, e->d_q, e->d_out)) return -1;
        std::vector<double> res(n * nint);
        if (hipMemcpyAsync(res.data(), e->d_out, sizeof(double) * n * nint, hipMemcpyDeviceToHost, e->stream) !=
                hipSuccess ||
            hipStreamSynchronize(e->stream) != hipSuccess)
            return -1;
        const int NB = 2 * ci->lb + 1, NCc = 2 * ci->lc + 1, ND = 2 * ci->ld + 1;
        const int NCD = NCc * ND;
        for (long k = 0; k < n; k++) {
            const int i = mem[k];
            const int M = q[4 * i], N = q[4 * i + 1], P = q[4 * i + 2], Q = q[4 * i + 3];
            const int dM = b->fstart[M + 1] - b->fstart[M], dN = b->fstart[N + 1] - b->fstart[N];
            const int dP = b->fstart[P + 1] - b->fstart[P], dQ = b->fstart[Q + 1] - b->fstart[Q];
            const int bra = canon[i].x, ket = canon[i].y;
            const double* src = res.data() + k * nint;
            double* dst = out + outoff[i];
            for (int iQ = 0; iQ < dQ; iQ++)
                for (int iP = 0; iP < dP; iP++)
                    for (int iN = 0; iN < dN; iN++)
                        for (int iM = 0; iM < dM; iM++) {
                            // requested bra/ket pairs in terms of canonical pairs
                            int x1 = M, y1 = N, ix1 = iM, iy1 = iN, x2 = P, y2 = Q, ix2 = iP, iy2 = iQ;
                            if (swapped[i]) {
                                std::swap(x1, x2); std::swap(y1, y2);
                                std::swap(ix1, ix2); std::swap(iy1, iy2);
                            }
                            const int a = (s.h_s1[bra] == x1) ? ix1 : iy1;
                            const int bb = (s.h_s1[bra] == x1) ? iy1 : ix1;
                            const int c = (s.h_s1[ket] == x2) ? ix2 : iy2;
                            const int d = (s.h_s1[ket] == x2) ? iy2 : ix2;
                            (void)y1; (void)y2;
                            dst[iM + dM * (iN + dN * (iP + dP * iQ))] = src[(a * NB + bb) * NCD + c * ND + d];
                        }
        }
    }
    return 0;
}

extern "C" {

CIntStatus_t CInt_createERD(BasisSet_t basis, ERD_t* erd, int nthreads)
{
    if (!basis || !erd) return CINT_STATUS_INVALID_VALUE;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
        fprintf(stderr, "CInt_createERD: no HIP device is visible\n");
        return CINT_STATUS_NOT_INITIALIZED;
    }
    ERD* e = new (std::nothrow) ERD();
    if (!e) return CINT_STATUS_ALLOC_FAILED;
    e->basis = basis;
    e->bufs.resize(std::max(1, nthreads));
    if (hipStreamCreate(&e->stream) != hipSuccess || e->sys.init(basis, e->stream) != 0) {
        CInt_destroyERD(e);
        return CINT_STATUS_EXECUTION_FAILED;
    }
    *erd = e;
    return CINT_STATUS_SUCCESS;
}

CIntStatus_t CInt_destroyERD(ERD_t e)
{
    if (!e) return CINT_STATUS_SUCCESS;
    if (e->stream) (void)hipStreamSynchronize(e->stream);
    if (e->d_q) (void)hipFree(e->d_q);
    if (e->d_out) (void)hipFree(e->d_out);
    e->sys.release();
    if (e->stream) (void)hipStreamDestroy(e->stream);
    delete e;
    return CINT_STATUS_SUCCESS;
}

CIntStatus_t CInt_computeShellQuartet(BasisSet_t basis, ERD_t e, int tid, int A, int B, int C, int D,
                                      double** integrals, int* nints)
{
    (void)basis;
    if (!e || tid < 0 || tid >= (int)e->bufs.size()) return CINT_STATUS_INVALID_VALUE;
    const BasisSet* b = e->basis;
    const int q[4] = {A, B, C, D};
    long n = 1;
    for (int k = 0; k < 4; k++) {
        if (q[k] < 0 || q[k] >= b->nshells) return CINT_STATUS_INVALID_VALUE;
        n *= b->fstart[q[k] + 1] - b->fstart[q[k]];
    }
    std::vector<double>& buf = e->bufs[tid];
    buf.resize(n);
    int rc = erd_batch(e, 1, q, buf.data());
    if (rc) return rc == -2 ? CINT_STATUS_ALLOC_FAILED : CINT_STATUS_EXECUTION_FAILED;
    *integrals = buf.data();
    *nints = (int)n;
    return CINT_STATUS_SUCCESS;
}

CIntStatus_t CInt_computeShellQuartets(BasisSet_t basis, ERD_t e, uint32_t threadId, uint32_t A,
                                       const uint32_t* Bs, uint32_t C, const uint32_t* Ds, uint32_t count,
                                       double** integrals, int* integralsCount)
{
    (void)basis;
    if (!e || threadId >= e->bufs.size()) return CINT_STATUS_INVALID_VALUE;
    std::vector<int> q(4 * (size_t)count);
    long n = 0;
    const BasisSet* b = e->basis;
    for (uint32_t i = 0; i < count; i++) {
        q[4 * i] = (int)A;
        q[4 * i + 1] = (int)Bs[i];
        q[4 * i + 2] = (int)C;
        q[4 * i + 3] = (int)Ds[i];
        long d = 1;
        for (int k = 0; k < 4; k++) {
            if (q[4 * i + k] < 0 || q[4 * i + k] >= b->nshells) return CINT_STATUS_INVALID_VALUE;
            d *= b->fstart[q[4 * i + k] + 1] - b->fstart[q[4 * i + k]];
        }
        n += d;
    }
    std::vector<double>& buf = e->bufs[threadId];
    buf.resize(std::max(1L, n));
    int rc = erd_batch(e, (int)count, q.data(), buf.data());
    if (rc) return rc == -2 ? CINT_STATUS_ALLOC_FAILED : CINT_STATUS_EXECUTION_FAILED;
    *integrals = buf.data();
    *integralsCount = (int)n;
    return CINT_STATUS_SUCCESS;
}

void CInt_getMaxMemory(ERD_t e, double* memsize) { *memsize = e ? (double)e->sys.dev_bytes : 0.0; }

// batched quartets for callers that want many at once (include/gtfock_amd.h)
int gtf_erd_quartets(ERD_t e, int nq, const int* quartets, double* out) { return erd_batch(e, nq, quartets, out); }

CIntStatus_t CInt_createOED(BasisSet_t basis, OED_t* oed)
{
    if (!basis || !oed) return CINT_STATUS_INVALID_VALUE;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return CINT_STATUS_NOT_INITIALIZED;
    OED* o = new (std::nothrow) OED();
    if (!o) return CINT_STATUS_ALLOC_FAILED;
    hipStream_t st;
    if (hipStreamCreate(&st) != hipSuccess) { delete o; return CINT_STATUS_EXECUTION_FAILED; }
    int rc;
    {
        GtfSystem sys;
        rc = sys.init(basis, st);
        const int n = basis->nbf;
        const long nn = (long)n * n;
        double *S = nullptr, *T = nullptr, *V = nullptr;
        if (rc == 0) {
            S = sys.dalloc<double>(nn);
            T = sys.dalloc<double>(nn);
            V = sys.dalloc<double>(nn);
            if (!S || !T || !V) rc = -2;
        }
        if (rc == 0) {
            BasisView bv{sys.nshells, sys.nbf, sys.natoms, sys.d_L, sys.d_nprim, sys.d_poff, sys.d_fstart,
                         sys.d_atom, sys.d_alpha, sys.d_coef, sys.d_center, sys.d_atom_xyz, sys.d_atom_Z};
            rc = gtf_oneel(st, bv, sys.d_boys, n, S, T, V);
        }
        if (rc == 0) {
            o->nbf = n;
            o->fstart = basis->fstart;
            o->S.resize(nn);
            o->T.resize(nn);
            o->V.resize(nn);
            if (hipMemcpyAsync(o->S.data(), S, sizeof(double) * nn, hipMemcpyDeviceToHost, st) != hipSuccess ||
                hipMemcpyAsync(o->T.data(), T, sizeof(double) * nn, hipMemcpyDeviceToHost, st) != hipSuccess ||
                hipMemcpyAsync(o->V.data(), V, sizeof(double) * nn, hipMemcpyDeviceToHost, st) != hipSuccess ||
                hipStreamSynchronize(st) != hipSuccess)
                rc = -1;
        }
        sys.release();
    }
    (void)hipStreamDestroy(st);
    if (rc) { delete o; return rc == -2 ? CINT_STATUS_ALLOC_FAILED : CINT_STATUS_EXECUTION_FAILED; }
    *oed = o;
    return CINT_STATUS_SUCCESS;
}

CIntStatus_t CInt_destroyOED(OED_t oed)
{
    delete oed;
    return CINT_STATUS_SUCCESS;
}

// block of shell pair (A,B), row index fastest (pfock/one_electron.c:14-25)
static CIntStatus_t oed_block(OED_t o, int A, int B, int which, double** integrals, int* nints)
{
    if (!o || A < 0 || B < 0 || A + 1 >= (int)o->fstart.size() || B + 1 >= (int)o->fstart.size())
        return CINT_STATUS_INVALID_VALUE;
    const int r0 = o->fstart[A], nr = o->fstart[A + 1] - r0, c0 = o->fstart[B], nc = o->fstart[B + 1] - c0;
    o->buf.resize((size_t)nr * nc);
    for (int k = 0; k < nr; k++)
        for (int l = 0; l < nc; l++) {
            const long idx = (long)(r0 + k) * o->nbf + c0 + l;
            double v;
            switch (which) {
            case 0: v = o->S[idx]; break;
            case 1: v = o->T[idx]; break;
            case 2: v = o->V[idx]; break;
            default: v = o->T[idx] + o->V[idx]; break;
            }
            o->buf[k + (size_t)nr * l] = v;
        }
    *integrals = o->buf.data();
    *nints = nr * nc;
    return CINT_STATUS_SUCCESS;
}

CIntStatus_t CInt_computePairOvl(BasisSet_t, OED_t o, int A, int B, double** I, int* n) { return oed_block(o, A, B, 0, I, n); }
CIntStatus_t CInt_computePairKin(BasisSet_t, OED_t o, int A, int B, double** I, int* n) { return oed_block(o, A, B, 1, I, n); }
CIntStatus_t CInt_computePairPot(BasisSet_t, OED_t o, int A, int B, double** I, int* n) { return oed_block(o, A, B, 2, I, n); }
CIntStatus_t CInt_computePairCoreH(BasisSet_t, OED_t o, int A, int B, double** I, int* n) { return oed_block(o, A, B, 3, I, n); }

}  // extern "C"
