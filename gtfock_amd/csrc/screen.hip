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
// One 64-lane workgroup per shell pair.  Per shell pair the lanes tabulate,
// in LDS, the Hermite index triples, the cartesian terms of every spherical
// component pair and the compact position of every structurally nonzero
// (component pair, Hermite index) entry.  Then for each primitive pair in
// turn three lanes run the one-dimensional recursions E^x, E^y, E^z into LDS
// and all lanes form the coefficient block from them, one entry per lane;
// a wavefront max decides whether the primitive pair is kept.
#define BP_MAXAB (nsph(GTF_LMAX) * nsph(GTF_LMAX))
#define BP_MAXH nherm(2 * GTF_LMAX)
#define BP_MAXTERM 9  // cartesian pairs of one spherical pair (d: at most 3 x 3)
#define BP_NT (2 * GTF_LMAX + 2)

__global__ __launch_bounds__(64) void build_pairs_kernel(BasisView b, int npairs, const int* __restrict__ s1,
                                                         const int* __restrict__ s2, const int* __restrict__ poff,
                                                         const long* __restrict__ eoff, double4* __restrict__ prim,
                                                         double* __restrict__ E, const long* __restrict__ zoff,
                                                         double* __restrict__ Ez, int* __restrict__ nprim_out,
                                                         double prim_eps)
{
    __shared__ double e1[3][GTF_LMAX + 1][GTF_LMAX + 1][BP_NT];
    __shared__ double ent[BP_MAXAB * BP_MAXH];
    __shared__ int zpos[BP_MAXAB * BP_MAXH];
    __shared__ int hcode[BP_MAXH];
    __shared__ double tcoef[BP_MAXAB][BP_MAXTERM];
    __shared__ int tcode[BP_MAXAB][BP_MAXTERM];
    __shared__ int nterm[BP_MAXAB];
    __shared__ double pp[5];  // p, P, scale of the current primitive pair
    const int pid = blockIdx.x, lane = threadIdx.x;
    const int A = s1[pid], B = s2[pid];
    const int la = b.L[A], lb = b.L[B];
    const int na = nsph(la), nb = nsph(lb), H = nherm(la + lb), nab = na * nb, esz = nab * H;
    const int nA = b.nprim[A], nB = b.nprim[B], nraw = nA * nB;
    // ---- per shell pair tables
    for (int h = lane; h < H; h += 64) hcode[h] = herm_t(h) | herm_u(h) << 4 | herm_v(h) << 8;
    const int ca0 = la == 0 ? 0 : nherm(la - 1), cb0 = lb == 0 ? 0 : nherm(lb - 1);
    for (int ab = lane; ab < nab; ab += 64) {
        const int as = ab / nb, bs = ab - (ab / nb) * nb;
        int n = 0;
        for (int ca = 0; ca < ncart(la); ca++) {
            const double ta = sph_coef(la, as, ca);
            if (ta == 0.0) continue;
            for (int cb = 0; cb < ncart(lb); cb++) {
                const double tb = sph_coef(lb, bs, cb);
                if (tb == 0.0) continue;
                tcoef[ab][n] = ta * tb;
                tcode[ab][n] = herm_t(ca0 + ca) | herm_u(ca0 + ca) << 2 | herm_v(ca0 + ca) << 4 |
                               herm_t(cb0 + cb) << 6 | herm_u(cb0 + cb) << 8 | herm_v(cb0 + cb) << 10;
                n++;
            }
        }
        nterm[ab] = n;
    }
    for (int k = lane; k < 3 * (GTF_LMAX + 1) * (GTF_LMAX + 1) * BP_NT; k += 64) (&e1[0][0][0][0])[k] = 0.0;
    // structural nonzeros in (component pair, Hermite index) order
    int ez = 0;
    for (int c0 = 0; c0 < esz; c0 += 64) {
        const int idx = c0 + lane;
        const bool nz = idx < esz && herm_nz(la, lb, (idx / H) / nb, (idx / H) % nb, idx % H);
        const unsigned long long m = __ballot(nz);
        if (idx < esz) zpos[idx] = nz ? ez + __popcll(m & ((1ull << lane) - 1ull)) : -1;
        ez += __popcll(m);
    }
    const double* CA = b.center + 3 * A;
    const double* CB = b.center + 3 * B;
    const double AB2 = (CA[0] - CB[0]) * (CA[0] - CB[0]) + (CA[1] - CB[1]) * (CA[1] - CB[1]) +
                       (CA[2] - CB[2]) * (CA[2] - CB[2]);
    __syncthreads();
    // ---- primitive pairs
    int kept = 0;
    for (int r = 0; r < nraw; r++) {
        const int i = r / nB, j = r - (r / nB) * nB;
        const double a = b.alpha[b.poff[A] + i], bb = b.alpha[b.poff[B] + j];
        const double p = a + bb;
        if (lane < 3) {
            const int d = lane;
            const double P = (a * CA[d] + bb * CB[d]) / p;
            const double XPA = P - CA[d], XPB = P - CB[d];
            double(*e)[GTF_LMAX + 1][BP_NT] = e1[d];
            e[0][0][0] = 1.0;
            for (int x = 0; x < la; x++)
                for (int t = 0; t <= x + 1; t++) {
                    double v = XPA * e[x][0][t] + (t + 1) * e[x][0][t + 1];
                    if (t > 0) v += e[x][0][t - 1] / (2.0 * p);
                    e[x + 1][0][t] = v;
                }
            for (int y = 0; y < lb; y++)
                for (int x = 0; x <= la; x++)
                    for (int t = 0; t <= x + y + 1; t++) {
                        double v = XPB * e[x][y][t] + (t + 1) * e[x][y][t + 1];
                        if (t > 0) v += e[x][y][t - 1] / (2.0 * p);
                        e[x][y + 1][t] = v;
                    }
            pp[1 + d] = P;
        }
        if (lane == 3) {
            pp[0] = p;
            pp[4] = b.coef[b.poff[A] + i] * b.coef[b.poff[B] + j] * exp(-(a * bb / p) * AB2) / p;
        }
        __syncthreads();
        const double scale = pp[4];
        double lmax = 0.0;
        for (int idx = lane; idx < esz; idx += 64) {
            const int ab = idx / H, hc = hcode[idx - ab * H];
            const int t = hc & 15, u = (hc >> 4) & 15, v = hc >> 8;
            double sum = 0.0;
            for (int n = 0; n < nterm[ab]; n++) {
                const int c = tcode[ab][n];
                sum += tcoef[ab][n] * e1[0][c & 3][(c >> 6) & 3][t] * e1[1][(c >> 2) & 3][(c >> 8) & 3][u] *
                       e1[2][(c >> 4) & 3][(c >> 10) & 3][v];
            }
            const double val = scale * sum;
            ent[idx] = val;
            lmax = fmax(lmax, fabs(val));
        }
        for (int off = 32; off > 0; off >>= 1) lmax = fmax(lmax, __shfl_xor(lmax, off));
        if (lmax >= prim_eps) {  // the same on every lane
            double* Eo = E + eoff[pid] + (long)kept * esz;
            double* Zo = Ez + zoff[pid] + (long)kept * ez;
            for (int idx = lane; idx < esz; idx += 64) {
                Eo[idx] = ent[idx];
                if (zpos[idx] >= 0) Zo[zpos[idx]] = ent[idx];
            }
            if (lane == 0) prim[poff[pid] + kept] = make_double4(pp[0], pp[1], pp[2], pp[3]);
            kept++;
        }
        __syncthreads();  // e1 and pp are rewritten for the next primitive pair
    }
    if (lane == 0) nprim_out[pid] = kept;
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
    build_pairs_kernel<<<npairs, 64, 0, st>>>(b, npairs, s1, s2, poff, eoff, prim, E, zoff, Ez, nprim_out, prim_eps);
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
