// batch_linalg.hip -- the dense steps of an SCF iteration (pscf/purif.c,
// pscf/scf.c:80-102) for many molecules at once.
//
// Two per-molecule layouts share one device array each:
//   packed  molecule m at off[m], row-major n_m x n_m (the Fock engine's)
//   padded  molecule m at poff[m], row-major n_m x pld[m], pld = n rounded
//           up to 64 with zero padding (the MFMA GEMM's)
// Every kernel indexes molecules by blockIdx.y (blockIdx.z for the GEMM) and
// skips a molecule whose flag in `skip` is set.  Reductions run in one
// workgroup per molecule in a fixed order, so a molecule's result does not
// depend on which other molecules share the batch.
#include "batch_linalg.h"

typedef double double4_t __attribute__((ext_vector_type(4)));

#define BGT 64
#define BKT 16

// C_m = A_m * B_m (TRANSB = 0) or A_m * B_m^T (TRANSB = 1) on padded blocks;
// 64 x 64 C tile per workgroup as 2 x 2 v_mfma_f64_16x16x4_f64 tiles per wave
// A molecule is skipped when skip[m] != 0 or, during purification, when it
// has already converged (done[m] >= 0).
template <int TRANSB>
__global__ __launch_bounds__(256) void bgemm_kernel(const long* __restrict__ poff, const int* __restrict__ pld,
                                                    const int* __restrict__ skip, const int* __restrict__ done,
                                                    const double* __restrict__ A0, const double* __restrict__ B0,
                                                    double* __restrict__ C0)
{
    const int m = blockIdx.z;
    const int ld = pld[m];
    const int row0 = blockIdx.y * BGT, col0 = blockIdx.x * BGT;
    if (row0 >= ld || col0 >= ld || (skip && skip[m]) || (done && done[m] >= 0)) return;
    const double* A = A0 + poff[m];
    const double* B = B0 + poff[m];
    double* C = C0 + poff[m];
    __shared__ double As[BGT][BKT + 1];
    __shared__ double Bs[BKT][BGT + 1];
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int wr = wave >> 1, wc = wave & 1;
    double4_t acc[2][2];
    for (int i = 0; i < 2; i++)
        for (int j = 0; j < 2; j++) acc[i][j] = (double4_t){0.0, 0.0, 0.0, 0.0};
    const int li = lane & 15, lk = lane >> 4;
    for (int k0 = 0; k0 < ld; k0 += BKT) {
        for (int r = 0; r < 4; r++) {
            const int rr = (tid >> 4) + 16 * r, kk = tid & 15;
            As[rr][kk] = A[(long)(row0 + rr) * ld + k0 + kk];
        }
        if (TRANSB) {
            for (int r = 0; r < 4; r++) {
                const int jj = (tid >> 4) + 16 * r, kk = tid & 15;
                Bs[kk][jj] = B[(long)(col0 + jj) * ld + k0 + kk];
            }
        } else {
            for (int r = 0; r < 4; r++) {
                const int kk = (tid >> 6) + 4 * r, jj = tid & 63;
                Bs[kk][jj] = B[(long)(k0 + kk) * ld + col0 + jj];
            }
        }
        __syncthreads();
        for (int kk = 0; kk < BKT; kk += 4) {
            double a[2], b[2];
            for (int t = 0; t < 2; t++) {
                a[t] = As[wr * 32 + t * 16 + li][kk + lk];
                b[t] = Bs[kk + lk][wc * 32 + t * 16 + li];
            }
            for (int i = 0; i < 2; i++)
                for (int j = 0; j < 2; j++)
                    acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], b[j], acc[i][j], 0, 0, 0);
        }
        __syncthreads();
    }
    for (int i = 0; i < 2; i++)
        for (int j = 0; j < 2; j++)
            for (int r = 0; r < 4; r++) {
                // lane l, register r of the 16 x 16 result: row l/16 + 4r, column l%16
                const int row = row0 + wr * 32 + i * 16 + lk + 4 * r;
                const int col = col0 + wc * 32 + j * 16 + li;
                C[(long)row * ld + col] = acc[i][j][r];
            }
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

// Gershgorin bounds and trace of F_m, then the initial density
//   D0 = (lambda mu / n + nocc / n) I - (lambda / n) F     (pscf/purif.c:302-373)
__global__ __launch_bounds__(256) void bpurif_init_kernel(const long* __restrict__ poff, const int* __restrict__ nn,
                                                          const int* __restrict__ pld, const int* __restrict__ nocc,
                                                          const int* __restrict__ skip, const double* __restrict__ F0,
                                                          double* __restrict__ D0)
{
    __shared__ double sh[256];
    const int m = blockIdx.x;
    if (skip && skip[m]) return;
    const int n = nn[m], ld = pld[m];
    const double* F = F0 + poff[m];
    double* D = D0 + poff[m];
    double mn = 1e300, mx = -1e300, tr = 0.0;
    for (int i = threadIdx.x; i < n; i += 256) {
        double off = 0.0;
        for (int j = 0; j < n; j++) off += fabs(F[(long)j * ld + i]);
        const double dg = F[(long)i * ld + i];
        off -= fabs(dg);
        mn = fmin(mn, dg - off);
        mx = fmax(mx, dg + off);
        tr += dg;
    }
    const double hmin = block_reduce(mn, sh, [](double a, double b) { return fmin(a, b); });
    const double hmax = block_reduce(mx, sh, [](double a, double b) { return fmax(a, b); });
    const double trF = block_reduce(tr, sh, [](double a, double b) { return a + b; });
    const double mu = trF / n;
    const double no = nocc[m];
    const double lambda = fmin(no / (hmax - mu), (double)(n - nocc[m]) / (mu - hmin));
    const double a = -lambda / n, dshift = lambda * mu / n + no / n;
    for (long k = threadIdx.x; k < (long)ld * ld; k += 256) {
        const int i = (int)(k / ld), j = (int)(k - (long)i * ld);
        D[k] = (i < n && j < n) ? a * F[k] + (i == j ? dshift : 0.0) : 0.0;
    }
}

// stats[3m..] = {||D - D2||_F^2, tr(D2 - D3), tr(D - D2)} of the molecules
// still iterating; a molecule whose error drops below tol is marked done at
// this iteration (iters[m] = it) and still takes this iteration's update
__global__ __launch_bounds__(256) void bpurif_stats_kernel(const long* __restrict__ poff, const int* __restrict__ nn,
                                                           const int* __restrict__ pld, const int* __restrict__ skip,
                                                           const double* __restrict__ D0, const double* __restrict__ D20,
                                                           const double* __restrict__ D30, int* __restrict__ iters,
                                                           double* __restrict__ stats, int it, double tol)
{
    __shared__ double sh[256];
    const int m = blockIdx.x;
    if ((skip && skip[m]) || iters[m] >= 0) return;
    const int n = nn[m], ld = pld[m];
    const double* D = D0 + poff[m];
    const double* D2 = D20 + poff[m];
    const double* D3 = D30 + poff[m];
    double e = 0.0, t1 = 0.0, t2 = 0.0;
    for (long k = threadIdx.x; k < (long)ld * ld; k += 256) {
        const double d = D[k] - D2[k];
        e += d * d;
    }
    for (int i = threadIdx.x; i < n; i += 256) {
        const long k = (long)i * ld + i;
        t1 += D2[k] - D3[k];
        t2 += D[k] - D2[k];
    }
    auto add = [](double a, double b) { return a + b; };
    e = block_reduce(e, sh, add);
    t1 = block_reduce(t1, sh, add);
    t2 = block_reduce(t2, sh, add);
    if (threadIdx.x == 0) {
        stats[3 * m + 0] = e;
        stats[3 * m + 1] = t1;
        stats[3 * m + 2] = t2;
        if (sqrt(e) < tol) iters[m] = it;
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

// dots[m * BMAX_DOTS + v] = <x_m, vecs[v]_m> for v < nvec[m]; vecs[v] at v * stride
__global__ __launch_bounds__(256) void bmultidot_kernel(const long* __restrict__ poff, const int* __restrict__ pld,
                                                        const int* __restrict__ skip, const int* __restrict__ nvec,
                                                        const double* __restrict__ x0, const double* __restrict__ vecs,
                                                        long stride, double* __restrict__ dots)
{
    __shared__ double sh[256];
    const int m = blockIdx.x;
    if (skip && skip[m]) return;
    const int ld = pld[m];
    const double* x = x0 + poff[m];
    for (int v = 0; v < nvec[m]; v++) {
        const double* y = vecs + (long)v * stride + poff[m];
        double s = 0.0;
        for (long k = threadIdx.x; k < (long)ld * ld; k += 256) s += x[k] * y[k];
        s = block_reduce(s, sh, [](double a, double b) { return a + b; });
        if (threadIdx.x == 0) dots[m * BMAX_DOTS + v] = s;
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
// (pscf/scf.c:80-102, 339)
__global__ __launch_bounds__(256) void bfock_energy_kernel(const long* __restrict__ off, const int* __restrict__ nn,
                                                           const double* __restrict__ H0, const double* __restrict__ G0,
                                                           const double* __restrict__ D0, double* __restrict__ F0,
                                                           double* __restrict__ energy)
{
    __shared__ double sh[256];
    const int m = blockIdx.x;
    const long nm = (long)nn[m] * nn[m];
    const double* H = H0 + off[m];
    const double* G = G0 + off[m];
    const double* D = D0 + off[m];
    double* F = F0 + off[m];
    double e = 0.0;
    for (long k = threadIdx.x; k < nm; k += 256) {
        const double f = H[k] + G[k];
        F[k] = f;
        e += D[k] * (H[k] + f);
    }
    e = block_reduce(e, sh, [](double a, double b) { return a + b; });
    if (threadIdx.x == 0) energy[m] = e;
}

// ---- launchers ----------------------------------------------------------------

static dim3 per_mol_grid(int maxelems, int nmol)
{
    return dim3((unsigned)std::max(1, std::min((maxelems + 255) / 256, 256)), (unsigned)nmol);
}

int gtf_bgemm(hipStream_t st, const BatchLayout& L, const int* skip, const int* done, const double* A,
              const double* B, double* C, int transB)
{
    if (L.nmol <= 0) return 0;
    dim3 grid(L.maxpld / BGT, L.maxpld / BGT, L.nmol);
    if (transB) bgemm_kernel<1><<<grid, 256, 0, st>>>(L.poff, L.pld, skip, done, A, B, C);
    else bgemm_kernel<0><<<grid, 256, 0, st>>>(L.poff, L.pld, skip, done, A, B, C);
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
    bpurif_init_kernel<<<L.nmol, 256, 0, st>>>(L.poff, L.n, L.pld, nocc, skip, F, D);
    GTF_CHECK(hipGetLastError());
    return 0;
}

int gtf_bpurif_stats(hipStream_t st, const BatchLayout& L, const int* skip, const double* D, const double* D2,
                     const double* D3, int* iters, double* stats, int it, double tol)
{
    bpurif_stats_kernel<<<L.nmol, 256, 0, st>>>(L.poff, L.n, L.pld, skip, D, D2, D3, iters, stats, it, tol);
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
    bmultidot_kernel<<<L.nmol, 256, 0, st>>>(L.poff, L.pld, skip, nvec, x, vecs, stride, dots);
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
    bfock_energy_kernel<<<L.nmol, 256, 0, st>>>(L.off, L.n, H, G, D, F, energy);
    GTF_CHECK(hipGetLastError());
    return 0;
}
