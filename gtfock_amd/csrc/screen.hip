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
// dropped (their quartets contribute < 1e-15 to any Fock element).  The
// survivors are stored in descending order of their Schwarz bound
//   q = sqrt(max over component pairs ab of (ab|ab)),
// the largest self-repulsion of the primitive pair's product functions (the
// same component pair on both sides, P = Q): any primitive quartet of the
// pair obeys |(ab|cd)| <= q_bra q_ket (Schwarz inequality), and with the
// bounds descending the ERI kernels end their primitive loops at the first
// pair of bounds whose product is below the primitive threshold.  Ties keep
// primitive order.  nprim_out receives the survivor count, qout the bounds.
//
// One 64-lane workgroup per shell pair.  Per shell pair the lanes tabulate,
// in LDS, the Hermite index triples, the cartesian terms of every spherical
// component pair and the compact position of every structurally nonzero
// (component pair, Hermite index) entry.  Then for each primitive pair in
// turn three lanes run the one-dimensional recursions E^x, E^y, E^z into LDS
// and all lanes form the coefficient block from them, one entry per lane; a
// wavefront max decides whether the primitive pair is kept.  A first pass
// over the primitive pairs finds the kept ones and their bounds, the lanes
// rank them, and a second pass forms the kept blocks again (the same
// arithmetic, so the same values) and writes each at its rank.
#define BP_MAXAB (nsph(GTF_LMAX) * nsph(GTF_LMAX))
#define BP_MAXH nherm(2 * GTF_LMAX)
#define BP_MAXTERM 9  // cartesian pairs of one spherical pair (d: at most 3 x 3)
#define BP_NT (2 * GTF_LMAX + 2)
#define BP_RL (4 * GTF_LMAX)  // highest Hermite order of a self-repulsion
#define BP_MAXRAW GTF_MAX_RAW_PRIM_PAIRS  // primitive pairs of one shell pair
#define BP_TWO_PI_2_5 34.986836655249725693  // 2 pi^(5/2)

__global__ __launch_bounds__(64) void build_pairs_kernel(BasisView b, int npairs, const int* __restrict__ s1,
                                                         const int* __restrict__ s2, const int* __restrict__ poff,
                                                         const long* __restrict__ eoff, double4* __restrict__ prim,
                                                         double* __restrict__ E, const long* __restrict__ zoff,
                                                         double* __restrict__ Ez, int* __restrict__ nprim_out,
                                                         double* __restrict__ qout, double prim_eps)
{
    __shared__ double e1[3][GTF_LMAX + 1][GTF_LMAX + 1][BP_NT];
    __shared__ double ent[BP_MAXAB * BP_MAXH];
    __shared__ int zpos[BP_MAXAB * BP_MAXH];
    __shared__ int hcode[nherm(BP_RL)];
    __shared__ double tcoef[BP_MAXAB][BP_MAXTERM];
    __shared__ int tcode[BP_MAXAB][BP_MAXTERM];
    __shared__ int nterm[BP_MAXAB];
    __shared__ double pp[5];  // p, P, scale of the current primitive pair
    __shared__ double rz[nherm(BP_RL)];  // R_tuv(p / 2, 0) of the current primitive pair
    __shared__ double qraw[BP_MAXRAW];   // bound of each primitive pair, -1 if dropped
    __shared__ int rnk[BP_MAXRAW];
    const int pid = blockIdx.x, lane = threadIdx.x;
    const int A = s1[pid], B = s2[pid];
    const int la = b.L[A], lb = b.L[B];
    const int na = nsph(la), nb = nsph(lb), H = nherm(la + lb), nab = na * nb, esz = nab * H;
    const int nA = b.nprim[A], nB = b.nprim[B], nraw = nA * nB;
    const int L2 = 2 * (la + lb), HR = nherm(L2);
    // ---- per shell pair tables
    for (int h = lane; h < HR; h += 64) hcode[h] = herm_t(h) | herm_u(h) << 4 | herm_v(h) << 8;
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
    // the coefficient block of primitive pair r into ent (and p, P, scale
    // into pp); returns the wavefront max of |E|, the same on every lane
    auto block = [&](int r) {
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
        return lmax;
    };
    // ---- pass 1: which primitive pairs are kept, and their bounds
    for (int r = 0; r < nraw; r++) {
        const double lmax = block(r);
        double q = -1.0;
        if (lmax >= prim_eps) {  // the same on every lane
            const double p = pp[0];
            // R_tuv(alpha = p / 2, 0) for t + u + v <= L2, stage by stage as
            // in hermite_R (the X R terms vanish at zero distance;
            // F_n(0) = 1 / (2n + 1))
            if (lane == 0) {
                const double m2a = -p;  // -2 alpha
                double pw = 1.0;
                for (int n = 0; n < L2; n++) pw *= m2a;
                rz[0] = pw / (2 * L2 + 1);
                for (int n = L2 - 1; n >= 0; n--) {
                    pw /= m2a;
                    for (int h = nherm(L2 - n) - 1; h >= 1; h--) {
                        const int hc = hcode[h];
                        const int t = hc & 15, u = (hc >> 4) & 15, v = hc >> 8;
                        double rr;
                        if (t > 0) rr = t > 1 ? (t - 1) * rz[hidx(t - 2, u, v)] : 0.0;
                        else if (u > 0) rr = u > 1 ? (u - 1) * rz[hidx(t, u - 2, v)] : 0.0;
                        else rr = v > 1 ? (v - 1) * rz[hidx(t, u, v - 2)] : 0.0;
                        rz[h] = rr;
                    }
                    rz[0] = pw / (2 * n + 1);
                }
            }
            __syncthreads();
            // (ab|ab) / prefactor, one component pair per lane
            double dmax = 0.0;
            for (int ab = lane; ab < nab; ab += 64) {
                const double* eab = ent + ab * H;
                double d = 0.0;
                for (int h1 = 0; h1 < H; h1++) {
                    const double x = eab[h1];
                    if (x == 0.0) continue;
                    const int c1 = hcode[h1];
                    double sm = 0.0;
                    for (int h2 = 0; h2 < H; h2++) {
                        const double y = eab[h2];
                        if (y == 0.0) continue;
                        const int c2 = hcode[h2];
                        const int t2 = c2 & 15, u2 = (c2 >> 4) & 15, v2 = c2 >> 8;
                        const double rv = rz[hidx((c1 & 15) + t2, ((c1 >> 4) & 15) + u2, (c1 >> 8) + v2)];
                        sm += ((t2 + u2 + v2) & 1) ? -y * rv : y * rv;
                    }
                    d += x * sm;
                }
                dmax = fmax(dmax, d);
            }
            for (int off = 32; off > 0; off >>= 1) dmax = fmax(dmax, __shfl_xor(dmax, off));
            q = sqrt(BP_TWO_PI_2_5 * dmax / sqrt(2.0 * p));
        }
        if (lane == 0) qraw[r] = q;
        __syncthreads();  // e1, pp, ent and rz are rewritten for the next primitive pair
    }
    // ---- ranks of the kept primitive pairs by descending bound
    int kept = 0;
    for (int r = lane; r < nraw; r += 64) {
        const double q = qraw[r];
        int k = 0;
        if (q >= 0.0)
            for (int x = 0; x < nraw; x++) {
                const double qx = qraw[x];
                k += qx > q || (qx == q && x < r);
            }
        rnk[r] = k;
        kept += q >= 0.0;
    }
    for (int off = 32; off > 0; off >>= 1) kept += __shfl_xor(kept, off);
    __syncthreads();
    // ---- pass 2: the kept blocks at their ranks
    for (int r = 0; r < nraw; r++) {
        if (qraw[r] < 0.0) continue;  // the same on every lane
        block(r);
        const int k = rnk[r];
        double* Eo = E + eoff[pid] + (long)k * esz;
        double* Zo = Ez + zoff[pid] + (long)k * ez;
        for (int idx = lane; idx < esz; idx += 64) {
            Eo[idx] = ent[idx];
            if (zpos[idx] >= 0) Zo[zpos[idx]] = ent[idx];
        }
        if (lane == 0) {
            prim[poff[pid] + k] = make_double4(pp[0], pp[1], pp[2], pp[3]);
            qout[poff[pid] + k] = qraw[r];
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

// Primitive quartets the J/K kernels evaluate over a list of tiles (bra
// row, first ket, end ket): for each quartet of a tile that passes the
// kernels' tests (Schwarz, one orientation in a diagonal class, the
// precision split), the pairs (i, j) of bra and ket primitives with
// q_i q_j >= peps when the class screens primitives, else all of them.
// With both lists in descending order of q that is what the kernels' loops
// visit, whichever pair they take as the bra.  One workgroup per tile, one
// thread per ket; the total is added to *out.
__global__ __launch_bounds__(256) void prim_census_kernel(CensusArgs c)
{
    const int4 t = c.tiles[blockIdx.x];
    const int bra = c.bra_list[t.x];
    const int nbp = c.pv.nprim[bra];
    const double* __restrict__ qb = c.pv.q + c.pv.poff[bra];
    const double vb = c.vbra[t.x];
    unsigned long long cnt = 0;
    for (int j = t.y + threadIdx.x; j < t.z; j += blockDim.x) {
        if (!(vb * c.vket[j] >= c.tol2) || (c.same && c.rank_ket[j] > c.rank_bra[t.x])) continue;
        if ((vb * c.vket[j] >= c.tau2) == (c.low != 0)) continue;
        const int ket = c.ket_list[j];
        const int nk = c.pv.nprim[ket];
        if (!c.screened || !(c.pv.peps > 0.0)) {
            cnt += (unsigned long long)nbp * nk;
            continue;
        }
        const double* __restrict__ qk = c.pv.q + c.pv.poff[ket];
        for (int i = 0; i < nbp; i++) {
            const double qi = qb[i];
            int lo = 0, hi = nk;
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if (qi * qk[mid] >= c.pv.peps) lo = mid + 1;
                else hi = mid;
            }
            if (lo == 0) break;
            cnt += lo;
        }
    }
    for (int off = 32; off > 0; off >>= 1) cnt += __shfl_xor(cnt, off);
    if ((threadIdx.x & 63) == 0 && cnt) atomicAdd(c.out, cnt);
}

// ---- host launchers -------------------------------------------------------

int gtf_build_pairs(hipStream_t st, const BasisView& b, int npairs, const int* s1, const int* s2,
                    const int* poff, const long* eoff, double4* prim, double* E, const long* zoff, double* Ez,
                    int* nprim_out, double* qout, double prim_eps)
{
    if (npairs <= 0) return 0;
    build_pairs_kernel<<<npairs, 64, 0, st>>>(b, npairs, s1, s2, poff, eoff, prim, E, zoff, Ez, nprim_out, qout,
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

int gtf_prim_census(hipStream_t st, const CensusArgs& c)
{
    if (c.ntiles <= 0) return 0;
    prim_census_kernel<<<c.ntiles, 256, 0, st>>>(c);
    GTF_CHECK(hipGetLastError());
    return 0;
}
