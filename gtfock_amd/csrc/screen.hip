// screen.hip -- shell-pair data, Schwarz screening support and the
// prefix-scan compaction of surviving quartets.
//
// Shell pairs (canonical, L(s1) >= L(s2)) carry, per primitive pair,
// (p, P) and the McMurchie-Davidson Hermite coefficients
//   E^{ab}_{tuv} = E^x_{ax bx, t} E^y_{ay by, u} E^z_{az bz, v}
// with E^{i+1,j}_t = E^{ij}_{t-1}/(2p) + X_PA E^{ij}_t + (t+1) E^{ij}_{t+1}
// and  E^{i,j+1}_t = E^{ij}_{t-1}/(2p) + X_PB E^{ij}_t + (t+1) E^{ij}_{t+1}
// (Helgaker, Jorgensen, Olsen, Molecular Electronic-Structure Theory 9.5),
// transformed to real solid harmonics and scaled by c_a c_b exp(-mu AB^2)/p.
#include "gtf_internal.h"
#include "screen.h"

// Primitive pairs whose largest scaled coefficient |E| is below prim_eps are
// dropped (their quartets contribute < 1e-15 to any Fock element); the
// survivors are packed at the front of the pair's slots, in primitive order,
// and nprim_out receives their count.
//
// One wavefront per shell pair, one lane per primitive pair (chunks of 64):
// a first pass computes each primitive pair's largest |E|, a ballot gives
// every survivor its packed slot, and a second pass writes its coefficients
// there.

// Hermite coefficients of primitive pair (i, j) of shells A, B: returns the
// largest scaled |E|; with out != null also writes (p, P), the full and the
// compact coefficient blocks (out, outz, prim_out)
__device__ double pair_prim(const BasisView& b, int A, int B, int i, int j, double* out, double* outz,
                            double4* prim_out)
{
    const int la = b.L[A], lb = b.L[B], lab = la + lb;
    const int na = nsph(la), nb = nsph(lb), H = nherm(lab);
    const double* CA = b.center + 3 * A;
    const double* CB = b.center + 3 * B;
    double AB2 = 0.0;
    for (int d = 0; d < 3; d++) AB2 += (CA[d] - CB[d]) * (CA[d] - CB[d]);
    const double a = b.alpha[b.poff[A] + i], bb = b.alpha[b.poff[B] + j];
    const double p = a + bb, mu = a * bb / p;
    double P[3];
    double E1[3][GTF_LMAX + 1][GTF_LMAX + 1][2 * GTF_LMAX + 2];
    for (int d = 0; d < 3; d++) {
        P[d] = (a * CA[d] + bb * CB[d]) / p;
        const double XPA = P[d] - CA[d], XPB = P[d] - CB[d];
        for (int x = 0; x <= GTF_LMAX; x++)
            for (int y = 0; y <= GTF_LMAX; y++)
                for (int t = 0; t < 2 * GTF_LMAX + 2; t++) E1[d][x][y][t] = 0.0;
        E1[d][0][0][0] = 1.0;
        for (int x = 0; x < la; x++)
            for (int t = 0; t <= x + 1; t++) {
                double v = XPA * E1[d][x][0][t] + (t + 1) * E1[d][x][0][t + 1];
                if (t > 0) v += E1[d][x][0][t - 1] / (2.0 * p);
                E1[d][x + 1][0][t] = v;
            }
        for (int y = 0; y < lb; y++)
            for (int x = 0; x <= la; x++)
                for (int t = 0; t <= x + y + 1; t++) {
                    double v = XPB * E1[d][x][y][t] + (t + 1) * E1[d][x][y][t + 1];
                    if (t > 0) v += E1[d][x][y][t - 1] / (2.0 * p);
                    E1[d][x][y + 1][t] = v;
                }
    }
    const double scale = b.coef[b.poff[A] + i] * b.coef[b.poff[B] + j] * exp(-mu * AB2) / p;
    int nzpos = 0;
    double emax = 0.0;
    const int ca0 = la == 0 ? 0 : nherm(la - 1), cb0 = lb == 0 ? 0 : nherm(lb - 1);
    for (int as = 0; as < na; as++)
        for (int bs = 0; bs < nb; bs++)
            for (int h = 0; h < H; h++) {
                const int t = herm_t(h), u = herm_u(h), v = herm_v(h);
                double s = 0.0;
                for (int ca = 0; ca < ncart(la); ca++) {
                    const double ta = sph_coef(la, as, ca);
                    if (ta == 0.0) continue;
                    const int ax = herm_t(ca0 + ca), ay = herm_u(ca0 + ca), az = herm_v(ca0 + ca);
                    for (int cb = 0; cb < ncart(lb); cb++) {
                        const double tb = sph_coef(lb, bs, cb);
                        if (tb == 0.0) continue;
                        const int bx = herm_t(cb0 + cb), by = herm_u(cb0 + cb), bz = herm_v(cb0 + cb);
                        if (t > ax + bx || u > ay + by || v > az + bz) continue;
                        s += ta * tb * E1[0][ax][bx][t] * E1[1][ay][by][u] * E1[2][az][bz][v];
                    }
                }
                emax = fmax(emax, fabs(scale * s));
                if (out) {
                    out[(as * nb + bs) * H + h] = scale * s;
                    if (herm_nz(la, lb, as, bs, h)) outz[nzpos++] = scale * s;
                }
            }
    if (prim_out) *prim_out = make_double4(p, P[0], P[1], P[2]);
    return emax;
}

__global__ __launch_bounds__(256) void build_pairs_kernel(BasisView b, int npairs, const int* __restrict__ s1,
                                                          const int* __restrict__ s2, const int* __restrict__ poff,
                                                          const long* __restrict__ eoff, double4* __restrict__ prim,
                                                          double* __restrict__ E, const long* __restrict__ zoff,
                                                          double* __restrict__ Ez, int* __restrict__ nprim_out,
                                                          double prim_eps)
{
    const int pid = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (pid >= npairs) return;
    const int A = s1[pid], B = s2[pid];
    const int la = b.L[A], lb = b.L[B];
    const int esz = nsph(la) * nsph(lb) * nherm(la + lb), ez = herm_nnz(la, lb);
    const int nA = b.nprim[A], nB = b.nprim[B], nraw = nA * nB;
    int k = 0;  // survivors so far (wave-uniform)
    for (int c0 = 0; c0 < nraw; c0 += 64) {
        const int r = c0 + lane;
        const bool live = r < nraw;
        const int i = live ? r / nB : 0, j = live ? r - (r / nB) * nB : 0;
        const bool keep = live && pair_prim(b, A, B, i, j, nullptr, nullptr, nullptr) >= prim_eps;
        const unsigned long long mask = __ballot(keep);
        if (keep) {
            const int slot = k + __popcll(mask & ((1ull << lane) - 1ull));
            pair_prim(b, A, B, i, j, E + eoff[pid] + (long)slot * esz, Ez + zoff[pid] + (long)slot * ez,
                      prim + poff[pid] + slot);
        }
        k += __popcll(mask);
    }
    if (lane == 0) nprim_out[pid] = k;
}

// Row k = (bra pair r, ket segment s).  exact[k] = number of kets j of the
// segment (value order) with v_r * v_j >= tol2 (the test of
// pfock/fock_task.c:192); for a diagonal class also rank(j) <= rank(r), so
// each unordered pair of pairs appears once (ranks rise along a segment's
// value order, so both conditions select prefixes).  counts[k] rounds
// exact[k] up to whole ket blocks within the segment (the kernels re-apply
// the exact test to the padding).
__global__ void count_rows_kernel(int nrows, const int* __restrict__ row_bra, const int* __restrict__ row_seg,
                                  const double* __restrict__ vbra, const int* __restrict__ rank_bra,
                                  const double* __restrict__ vket_sorted, const int* __restrict__ rank_sorted,
                                  const int* __restrict__ seg_start, const int* __restrict__ seg_len, double tol2,
                                  int same, long* __restrict__ exact, long* __restrict__ counts)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nrows) return;
    const int r = row_bra[k], s = row_seg[k];
    const double vk = vbra[r];
    const double* vs = vket_sorted + seg_start[s];
    const int nket = seg_len[s];
    int lo = 0, hi = nket;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (vk * vs[mid] >= tol2) lo = mid + 1;
        else hi = mid;
    }
    long cnt = lo;
    if (same) {
        const int* rs = rank_sorted + seg_start[s];
        const int rk = rank_bra[r];
        int a = 0, b = nket;  // number of entries with rank <= rk
        while (a < b) {
            const int mid = (a + b) >> 1;
            if (rs[mid] <= rk) a = mid + 1;
            else b = mid;
        }
        if (cnt > a) cnt = a;
    }
    exact[k] = cnt;
    long pad = cnt == 0 ? 0 : ((cnt + GTF_KET_BLOCK - 1) / GTF_KET_BLOCK) * GTF_KET_BLOCK;
    counts[k] = pad < nket ? pad : nket;
}

// Exclusive prefix sum of n longs in one workgroup of 1024 threads:
// per-thread serial chunk sums, a wavefront/LDS scan of the 1024 partial
// sums, then the per-thread chunk rescan.  out has n + 1 entries.
__global__ __launch_bounds__(1024) void scan_kernel(const long* __restrict__ in, long* __restrict__ out, int n)
{
    __shared__ long part[1024];
    const int t = threadIdx.x;
    const int chunk = (n + 1023) / 1024;
    const int beg = min(n, t * chunk), end = min(n, beg + chunk);
    long s = 0;
    for (int i = beg; i < end; i++) s += in[i];
    part[t] = s;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        long v = t >= off ? part[t - off] : 0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    long run = part[t] - s;
    for (int i = beg; i < end; i++) {
        out[i] = run;
        run += in[i];
    }
    if (t == 1023) out[n] = part[1023];
}

__global__ void fill_qlist_diag_kernel(int n, const int* __restrict__ ids, int2* __restrict__ q)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) q[k] = make_int2(ids[k], ids[k]);
}

// ---- host launchers -------------------------------------------------------

int gtf_build_pairs(hipStream_t st, const BasisView& b, int npairs, const int* s1, const int* s2,
                    const int* poff, const long* eoff, double4* prim, double* E, const long* zoff, double* Ez,
                    int* nprim_out, double prim_eps)
{
    if (npairs <= 0) return 0;
    build_pairs_kernel<<<(npairs + 3) / 4, 256, 0, st>>>(b, npairs, s1, s2, poff, eoff, prim, E, zoff, Ez, nprim_out,
                                                       prim_eps);
    GTF_CHECK(hipGetLastError());
    return 0;
}

int gtf_count_rows(hipStream_t st, int nrows, const int* row_bra, const int* row_seg, const double* vbra,
                   const int* rank_bra, const double* vket_sorted, const int* rank_sorted, const int* seg_start,
                   const int* seg_len, double tol2, int same, long* exact, long* counts, long* row_off)
{
    if (nrows <= 0) return 0;
    count_rows_kernel<<<(nrows + 255) / 256, 256, 0, st>>>(nrows, row_bra, row_seg, vbra, rank_bra, vket_sorted,
                                                           rank_sorted, seg_start, seg_len, tol2, same, exact, counts);
    GTF_CHECK(hipGetLastError());
    scan_kernel<<<1, 1024, 0, st>>>(counts, row_off, nrows);
    GTF_CHECK(hipGetLastError());
    return 0;
}

int gtf_fill_qlist_diag(hipStream_t st, int n, const int* ids, int2* q)
{
    if (n <= 0) return 0;
    fill_qlist_diag_kernel<<<(n + 255) / 256, 256, 0, st>>>(n, ids, q);
    GTF_CHECK(hipGetLastError());
    return 0;
}
