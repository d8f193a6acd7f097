// batch_linalg.hip -- the dense steps of an SCF iteration (pscf/purif.c,
// pscf/scf.c:80-102) for many molecules at once.
//
// Two per-molecule layouts share one device array each:
//   packed  molecule m at off[m], row-major n_m x n_m (the Fock engine's)
//   padded  molecule m at poff[m], row-major n_m x pld[m], pld = n rounded
//           up to 64 with zero padding (the MFMA GEMM's)
// Every kernel indexes molecules by blockIdx.y (blockIdx.z for the GEMM) and
// skips a molecule whose flag in `skip` is set.  Reductions add fixed-order
// partials, so a molecule's result does not depend on which other molecules
// share the batch.
#include "batch_linalg.h"

#include "gemm_tile.h"
#include "ozaki_launch.h"

// C_m = A_m * B_m (TRANSB = 0) or A_m * B_m^T (TRANSB = 1) on padded blocks
// (gemm_tile.h).  A molecule is skipped when skip[m] != 0 or, during
// purification, when it has already converged (done[m] >= 0); fp32[m] ==
// GTF_LOW_FP32 selects the FP32 MFMA tile for it, GTF_LOW_SPLIT leaves it to
// the split-BF16 products.
template <int TRANSB>
__global__ __launch_bounds__(256) void bgemm_kernel(const long* __restrict__ poff, const int* __restrict__ pld,
                                                    const int* __restrict__ skip, const int* __restrict__ done,
                                                    const int* __restrict__ fp32, const double* __restrict__ A0,
                                                    const double* __restrict__ B0, double* __restrict__ C0)
{
    const int m = blockIdx.z;
    const int ld = pld[m];
    const int row0 = blockIdx.y * GEMM_GT, col0 = blockIdx.x * GEMM_GT;
    if (row0 >= ld || col0 >= ld || (skip && skip[m]) || (done && done[m] >= 0)) return;
    const long o = poff[m];
    if (fp32 && fp32[m] == GTF_LOW_SPLIT) return;  // the split-BF16 products (ozaki.hip) cover it
    if (fp32 && fp32[m] == GTF_LOW_FP32) gemm_tile<TRANSB, true>(ld, A0 + o, B0 + o, C0 + o, row0, col0);
    else gemm_tile<TRANSB, false>(ld, A0 + o, B0 + o, C0 + o, row0, col0);
}

// packed -> padded (zero padding) and padded -> packed
__global__ void bpad_kernel(const long* __restrict__ off, const int* __restrict__ nn, const long* __restrict__ poff,
                            const int* __restrict__ pld, const int* __restrict__ skip, const double* __restrict__ src,
                            double* __restrict__ dst)
{
    const int m = blockIdx.y;
    if (skip && skip[m]) return;
    const int n = nn[m], ld = pld[m];
    const double* s = src + off[m];
    double* d = dst + poff[m];
    for (long k = (long)blockIdx.x * blockDim.x + threadIdx.x; k < (long)ld * ld; k += (long)gridDim.x * blockDim.x) {
        const int i = (int)(k / ld), j = (int)(k - (long)i * ld);
        d[k] = (i < n && j < n) ? s[(long)i * n + j] : 0.0;
    }
}
__global__ void bunpad_kernel(const long* __restrict__ off, const int* __restrict__ nn, const long* __restrict__ poff,
                              const int* __restrict__ pld, const int* __restrict__ skip,
                              const double* __restrict__ src, double* __restrict__ dst)
{
    const int m = blockIdx.y;
    if (skip && skip[m]) return;
    const int n = nn[m], ld = pld[m];
    const double* s = src + poff[m];
    double* d = dst + off[m];
    for (long k = (long)blockIdx.x * blockDim.x + threadIdx.x; k < (long)n * n; k += (long)gridDim.x * blockDim.x) {
        const int i = (int)(k / n), j = (int)(k - (long)i * n);
        d[k] = s[(long)i * ld + j];
    }
}

// fixed-order tree reduction of one value per thread of a 256-thread block
template <typename Op>
__device__ __forceinline__ double block_reduce(double v, double* sh, Op op)
{
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) sh[threadIdx.x] = op(sh[threadIdx.x], sh[threadIdx.x + s]);
        __syncthreads();
    }
    const double r = sh[0];
    __syncthreads();
    return r;
}

// Reductions over one molecule's matrix run in BPART workgroups per molecule
// (grid BPART x nmol), each writing its partial to part[(m * BPART + p) * w
// + v]; a finishing kernel (one thread per molecule) adds the partials in
// order p = 0 .. BPART-1.  The result depends only on the molecule's data.

// Gershgorin bounds and trace of F_m (pscf/purif.c:302-358), partials
__global__ __launch_bounds__(256) void bgersh_part_kernel(const long* __restrict__ poff, const int* __restrict__ nn,
                                                          const int* __restrict__ pld, const int* __restrict__ skip,
                                                          const double* __restrict__ F0, double* __restrict__ part)
{
    __shared__ double sh[256];
    const int m = blockIdx.y, p = blockIdx.x;
    if (skip && skip[m]) return;
    const int n = nn[m], ld = pld[m];
    const double* F = F0 + poff[m];
    double mn = 1e300, mx = -1e300, tr = 0.0;
    for (int i = p * 256 + threadIdx.x; i < n; i += BPART * 256) {
        double off = 0.0;
        for (int j = 0; j < n; j++) off += fabs(F[(long)j * ld + i]);
        const double dg = F[(long)i * ld + i];
        off -= fabs(dg);
        mn = fmin(mn, dg - off);
        mx = fmax(mx, dg + off);
        tr += dg;
    }
    mn = block_reduce(mn, sh, [](double a, double b) { return fmin(a, b); });
    mx = block_reduce(mx, sh, [](double a, double b) { return fmax(a, b); });
    tr = block_reduce(tr, sh, [](double a, double b) { return a + b; });
    if (threadIdx.x == 0) {
        double* o = part + ((long)m * BPART + p) * 3;
        o[0] = mn;
        o[1] = mx;
        o[2] = tr;
    }
}

// coef[2m] = -lambda / n, coef[2m + 1] = lambda mu / n + nocc / n with
// mu = tr F / n, lambda = min(nocc / (hmax - mu), (n - nocc) / (mu - hmin))
// (pscf/purif.c:347-373)
__global__ void bgersh_fin_kernel(int nmol, const int* __restrict__ nn, const int* __restrict__ nocc,
                                  const int* __restrict__ skip, const double* __restrict__ part,
                                  double* __restrict__ coef)
{
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= nmol || (skip && skip[m])) return;
    double hmin = 1e300, hmax = -1e300, tr = 0.0;
    for (int p = 0; p < BPART; p++) {
        const double* o = part + ((long)m * BPART + p) * 3;
        hmin = fmin(hmin, o[0]);
        hmax = fmax(hmax, o[1]);
        tr += o[2];
    }
    const int n = nn[m];
    const double mu = tr / n, no = nocc[m];
    const double lambda = fmin(no / (hmax - mu), (double)(n - nocc[m]) / (mu - hmin));
    coef[2 * m] = -lambda / n;
    coef[2 * m + 1] = lambda * mu / n + no / n;
}

// D0 = coef0 F + coef1 I on the n x n block, zero padding
__global__ void bpurif_init_kernel(const long* __restrict__ poff, const int* __restrict__ nn,
                                   const int* __restrict__ pld, const int* __restrict__ skip,
                                   const double* __restrict__ coef, const double* __restrict__ F0,
                                   double* __restrict__ D0)
{
    const int m = blockIdx.y;
    if (skip && skip[m]) return;
    const int n = nn[m], ld = pld[m];
    const double a = coef[2 * m], dshift = coef[2 * m + 1];
    const double* F = F0 + poff[m];
    double* D = D0 + poff[m];
    for (long k = (long)blockIdx.x * blockDim.x + threadIdx.x; k < (long)ld * ld; k += (long)gridDim.x * blockDim.x) {
        const int i = (int)(k / ld), j = (int)(k - (long)i * ld);
        D[k] = (i < n && j < n) ? a * F[k] + (i == j ? dshift : 0.0) : 0.0;
    }
}

// partials of {||D - D2||_F^2, tr(D2 - D3), tr(D - D2)} of the molecules
// still iterating
__global__ __launch_bounds__(256) void bstats_part_kernel(const long* __restrict__ poff, const int* __restrict__ nn,
                                                          const int* __restrict__ pld, const int* __restrict__ skip,
                                                          const int* __restrict__ iters,
                                                          const double* __restrict__ D0,
                                                          const double* __restrict__ D20,
                                                          const double* __restrict__ D30, double* __restrict__ part)
{
    __shared__ double sh[256];
    const int m = blockIdx.y, p = blockIdx.x;
    if ((skip && skip[m]) || iters[m] >= 0) return;
    const int n = nn[m], ld = pld[m];
    const double* D = D0 + poff[m];
    const double* D2 = D20 + poff[m];
    const double* D3 = D30 + poff[m];
    double e = 0.0, t1 = 0.0, t2 = 0.0;
    for (long k = (long)p * 256 + threadIdx.x; k < (long)ld * ld; k += BPART * 256) {
        const double d = D[k] - D2[k];
        e += d * d;
    }
    for (int i = p * 256 + threadIdx.x; i < n; i += BPART * 256) {
        const long k = (long)i * ld + i;
        t1 += D2[k] - D3[k];
        t2 += D[k] - D2[k];
    }
    auto add = [](double a, double b) { return a + b; };
    e = block_reduce(e, sh, add);
    t1 = block_reduce(t1, sh, add);
    t2 = block_reduce(t2, sh, add);
    if (threadIdx.x == 0) {
        double* o = part + ((long)m * BPART + p) * 3;
        o[0] = e;
        o[1] = t1;
        o[2] = t2;
    }
}

// stats[3m..] from the partials; a molecule whose error drops below tol is
// marked done at this iteration (iters[m] = it) and still takes this
// iteration's update
__global__ void bstats_fin_kernel(int nmol, const int* __restrict__ skip, const double* __restrict__ part,
                                  int* __restrict__ iters, double* __restrict__ stats, int it, double tol,
                                  int* __restrict__ fp32, double fp32_until, int* __restrict__ low_iters)
{
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= nmol || (skip && skip[m]) || iters[m] >= 0) return;
    double e = 0.0, t1 = 0.0, t2 = 0.0;
    for (int p = 0; p < BPART; p++) {
        const double* o = part + ((long)m * BPART + p) * 3;
        e += o[0];
        t1 += o[1];
        t2 += o[2];
    }
    stats[3 * m + 0] = e;
    stats[3 * m + 1] = t1;
    stats[3 * m + 2] = t2;
    if (sqrt(e) < tol) iters[m] = it;
    // FP32 GEMMs only while the error is far above FP32 resolution
    if (fp32 && fp32[m] && sqrt(e) < fp32_until) {
        fp32[m] = 0;
        if (low_iters) low_iters[m] = it + 1;  // iterations 0..it ran in the reduced precision
    }
}

// D <- ((1-2c) D + (1+c) D2 - D3)/(1-c) if c < 1/2, else ((1+c) D2 - D3)/c
// (pscf/purif.c:432-453), for molecules iterating at `it`
__global__ void bpurif_update_kernel(const long* __restrict__ poff, const int* __restrict__ pld,
                                     const int* __restrict__ skip, const int* __restrict__ iters,
                                     const double* __restrict__ stats, double* __restrict__ D0,
                                     const double* __restrict__ D20, const double* __restrict__ D30, int it)
{
    const int m = blockIdx.y;
    if ((skip && skip[m]) || (iters[m] >= 0 && iters[m] != it)) return;
    const int ld = pld[m];
    const double c = stats[3 * m + 1] / stats[3 * m + 2];
    double* D = D0 + poff[m];
    const double* D2 = D20 + poff[m];
    const double* D3 = D30 + poff[m];
    for (long k = (long)blockIdx.x * blockDim.x + threadIdx.x; k < (long)ld * ld; k += (long)gridDim.x * blockDim.x) {
        if (c < 0.5) D[k] = ((1.0 - 2.0 * c) * D[k] + (1.0 + c) * D2[k] - D3[k]) / (1.0 - c);
        else D[k] = ((1.0 + c) * D2[k] - D3[k]) / c;
    }
}

// E_m = W_m - W_m^T
__global__ void bantisym_kernel(const long* __restrict__ poff, const int* __restrict__ pld,
                                const int* __restrict__ skip, const double* __restrict__ W0, double* __restrict__ E0)
{
    const int m = blockIdx.y;
    if (skip && skip[m]) return;
    const int ld = pld[m];
    const double* W = W0 + poff[m];
    double* E = E0 + poff[m];
    for (long k = (long)blockIdx.x * blockDim.x + threadIdx.x; k < (long)ld * ld; k += (long)gridDim.x * blockDim.x) {
        const int i = (int)(k / ld), j = (int)(k - (long)i * ld);
        E[k] = W[k] - W[(long)j * ld + i];
    }
}

// dots[m * BMAX_DOTS + v] = <x_m, vecs[v]_m> for v < nvec[m]; vecs[v] at
// v * stride; partials then a finishing pass
__global__ __launch_bounds__(256) void bdot_part_kernel(const long* __restrict__ poff, const int* __restrict__ pld,
                                                        const int* __restrict__ skip, const int* __restrict__ nvec,
                                                        const double* __restrict__ x0, const double* __restrict__ vecs,
                                                        long stride, double* __restrict__ part)
{
    __shared__ double sh[256];
    const int m = blockIdx.y, p = blockIdx.x;
    if (skip && skip[m]) return;
    const int ld = pld[m];
    const double* x = x0 + poff[m];
    for (int v = 0; v < nvec[m]; v++) {
        const double* y = vecs + (long)v * stride + poff[m];
        double s = 0.0;
        for (long k = (long)p * 256 + threadIdx.x; k < (long)ld * ld; k += BPART * 256) s += x[k] * y[k];
        s = block_reduce(s, sh, [](double a, double b) { return a + b; });
        if (threadIdx.x == 0) part[((long)m * BPART + p) * BMAX_DOTS + v] = s;
    }
}

__global__ void bdot_fin_kernel(int nmol, const int* __restrict__ skip, const int* __restrict__ nvec,
                                const double* __restrict__ part, double* __restrict__ dots)
{
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= nmol || (skip && skip[m])) return;
    for (int v = 0; v < nvec[m]; v++) {
        double s = 0.0;
        for (int p = 0; p < BPART; p++) s += part[((long)m * BPART + p) * BMAX_DOTS + v];
        dots[m * BMAX_DOTS + v] = s;
    }
}

// F_m = sum_v coef[m, v] vecs[v]_m   (pscf/purif.c:629-635)
__global__ void blincomb_kernel(const long* __restrict__ poff, const int* __restrict__ pld,
                                const int* __restrict__ skip, const int* __restrict__ nvec,
                                const double* __restrict__ coef, const double* __restrict__ vecs, long stride,
                                double* __restrict__ F0)
{
    const int m = blockIdx.y;
    if (skip && skip[m]) return;
    const int ld = pld[m];
    const int nv = nvec[m];
    double* F = F0 + poff[m];
    for (long k = (long)blockIdx.x * blockDim.x + threadIdx.x; k < (long)ld * ld; k += (long)gridDim.x * blockDim.x) {
        double s = 0.0;
        for (int v = 0; v < nv; v++) s += coef[m * BMAX_DOTS + v] * vecs[(long)v * stride + poff[m] + k];
        F[k] = s;
    }
}

// F = H + G and E_m = sum_ij D_ij (H_ij + F_ij) on packed blocks
// (pscf/scf.c:80-102, 339): partials, then a finishing pass
__global__ __launch_bounds__(256) void bfe_part_kernel(const long* __restrict__ off, const int* __restrict__ nn,
                                                       const double* __restrict__ H0, const double* __restrict__ G0,
                                                       const double* __restrict__ D0, double* __restrict__ F0,
                                                       double* __restrict__ part)
{
    __shared__ double sh[256];
    const int m = blockIdx.y, p = blockIdx.x;
    const long nm = (long)nn[m] * nn[m];
    const double* H = H0 + off[m];
    const double* G = G0 + off[m];
    const double* D = D0 + off[m];
    double* F = F0 + off[m];
    double e = 0.0;
    for (long k = (long)p * 256 + threadIdx.x; k < nm; k += BPART * 256) {
        const double f = H[k] + G[k];
        F[k] = f;
        e += D[k] * (H[k] + f);
    }
    e = block_reduce(e, sh, [](double a, double b) { return a + b; });
    if (threadIdx.x == 0) part[(long)m * BPART + p] = e;
}

__global__ void bfe_fin_kernel(int nmol, const double* __restrict__ part, double* __restrict__ energy)
{
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= nmol) return;
    double e = 0.0;
    for (int p = 0; p < BPART; p++) e += part[(long)m * BPART + p];
    energy[m] = e;
}

// ---- launchers ----------------------------------------------------------------

static dim3 per_mol_grid(int maxelems, int nmol)
{
    return dim3((unsigned)std::max(1, std::min((maxelems + 255) / 256, 256)), (unsigned)nmol);
}

int gtf_bgemm(hipStream_t st, const BatchLayout& L, const int* skip, const int* done, const double* A,
              const double* B, double* C, int transB, const int* fp32)
{
    if (L.nmol <= 0) return 0;
    dim3 grid(L.maxpld / GEMM_GT, L.maxpld / GEMM_GT, L.nmol);
    if (transB) bgemm_kernel<1><<<grid, 256, 0, st>>>(L.poff, L.pld, skip, done, fp32, A, B, C);
    else bgemm_kernel<0><<<grid, 256, 0, st>>>(L.poff, L.pld, skip, done, fp32, A, B, C);
    GTF_CHECK(hipGetLastError());
    return 0;
}

int gtf_bpad(hipStream_t st, const BatchLayout& L, const int* skip, const double* packed, double* padded)
{
    bpad_kernel<<<per_mol_grid(L.maxpld * L.maxpld, L.nmol), 256, 0, st>>>(L.off, L.n, L.poff, L.pld, skip, packed,
                                                                          padded);
    GTF_CHECK(hipGetLastError());
    return 0;
}

int gtf_bunpad(hipStream_t st, const BatchLayout& L, const int* skip, const double* padded, double* packed)
{
    bunpad_kernel<<<per_mol_grid(L.maxn * L.maxn, L.nmol), 256, 0, st>>>(L.off, L.n, L.poff, L.pld, skip, padded,
                                                                       packed);
    GTF_CHECK(hipGetLastError());
    return 0;
}

int gtf_bpurif_init(hipStream_t st, const BatchLayout& L, const int* nocc, const int* skip, const double* F, double* D)
{
    bgersh_part_kernel<<<dim3(BPART, L.nmol), 256, 0, st>>>(L.poff, L.n, L.pld, skip, F, L.part);
    GTF_CHECK(hipGetLastError());
    bgersh_fin_kernel<<<(L.nmol + 63) / 64, 64, 0, st>>>(L.nmol, L.n, nocc, skip, L.part, L.coef);
    GTF_CHECK(hipGetLastError());
    bpurif_init_kernel<<<per_mol_grid(L.maxpld * L.maxpld, L.nmol), 256, 0, st>>>(L.poff, L.n, L.pld, skip, L.coef,
                                                                                F, D);
    GTF_CHECK(hipGetLastError());
    return 0;
}

int gtf_bpurif_stats(hipStream_t st, const BatchLayout& L, const int* skip, const double* D, const double* D2,
                     const double* D3, int* iters, double* stats, int it, double tol, int* fp32, double fp32_until,
                     int* low_iters)
{
    bstats_part_kernel<<<dim3(BPART, L.nmol), 256, 0, st>>>(L.poff, L.n, L.pld, skip, iters, D, D2, D3, L.part);
    GTF_CHECK(hipGetLastError());
    bstats_fin_kernel<<<(L.nmol + 63) / 64, 64, 0, st>>>(L.nmol, skip, L.part, iters, stats, it, tol, fp32,
                                                          fp32_until, low_iters);
    GTF_CHECK(hipGetLastError());
    return 0;
}

int gtf_bpurif_update(hipStream_t st, const BatchLayout& L, const int* skip, const int* iters, const double* stats,
                      double* D, const double* D2, const double* D3, int it)
{
    bpurif_update_kernel<<<per_mol_grid(L.maxpld * L.maxpld, L.nmol), 256, 0, st>>>(L.poff, L.pld, skip, iters, stats,
                                                                                  D, D2, D3, it);
    GTF_CHECK(hipGetLastError());
    return 0;
}

int gtf_bantisym(hipStream_t st, const BatchLayout& L, const int* skip, const double* W, double* E)
{
    bantisym_kernel<<<per_mol_grid(L.maxpld * L.maxpld, L.nmol), 256, 0, st>>>(L.poff, L.pld, skip, W, E);
    GTF_CHECK(hipGetLastError());
    return 0;
}

int gtf_bmultidot(hipStream_t st, const BatchLayout& L, const int* skip, const int* nvec, const double* x,
                  const double* vecs, long stride, double* dots)
{
    bdot_part_kernel<<<dim3(BPART, L.nmol), 256, 0, st>>>(L.poff, L.pld, skip, nvec, x, vecs, stride, L.part);
    GTF_CHECK(hipGetLastError());
    bdot_fin_kernel<<<(L.nmol + 63) / 64, 64, 0, st>>>(L.nmol, skip, nvec, L.part, dots);
    GTF_CHECK(hipGetLastError());
    return 0;
}

int gtf_blincomb(hipStream_t st, const BatchLayout& L, const int* skip, const int* nvec, const double* coef,
                 const double* vecs, long stride, double* F)
{
    blincomb_kernel<<<per_mol_grid(L.maxpld * L.maxpld, L.nmol), 256, 0, st>>>(L.poff, L.pld, skip, nvec, coef, vecs,
                                                                             stride, F);
    GTF_CHECK(hipGetLastError());
    return 0;
}

int gtf_bfock_energy(hipStream_t st, const BatchLayout& L, const double* H, const double* G, const double* D,
                     double* F, double* energy)
{
    bfe_part_kernel<<<dim3(BPART, L.nmol), 256, 0, st>>>(L.off, L.n, H, G, D, F, L.part);
    GTF_CHECK(hipGetLastError());
    bfe_fin_kernel<<<(L.nmol + 63) / 64, 64, 0, st>>>(L.nmol, L.part, energy);
    GTF_CHECK(hipGetLastError());
    return 0;
}
