// linalg.hip -- dense FP64 kernels: MFMA GEMM, symmetrisation, fills and the
// deterministic reductions used by purification and DIIS.
//
// GEMM: C = A * B or C = A * B^T for row-major matrices of order n padded to
// ld (a multiple of 64, zero padding), one 64 x 64 tile per workgroup
// (gemm_tile.h), FP64 MFMA or FP32 MFMA on rounded operands.
#include "gtf_internal.h"
#include "linalg.h"
#include "screen.h"

#include "gemm_tile.h"

template <int TRANSB, bool F32>
__global__ __launch_bounds__(256) void dgemm_mfma_kernel(int ld, const double* __restrict__ A,
                                                         const double* __restrict__ B, double* __restrict__ C)
{
    gemm_tile<TRANSB, F32>(ld, A, B, C, blockIdx.y * GEMM_GT, blockIdx.x * GEMM_GT);
}

int gtf_gemm(hipStream_t st, int n, int ld, const double* A, const double* B, double* C, int transB, int fp32)
{
    (void)n;
    if (ld % GEMM_GT != 0) {
        fprintf(stderr, "gtf_gemm: leading dimension %d is not a multiple of %d\n", ld, GEMM_GT);
        return -1;
    }
    dim3 grid(ld / GEMM_GT, ld / GEMM_GT);
    if (fp32) {
        if (transB) dgemm_mfma_kernel<1, true><<<grid, 256, 0, st>>>(ld, A, B, C);
        else dgemm_mfma_kernel<0, true><<<grid, 256, 0, st>>>(ld, A, B, C);
    } else {
        if (transB) dgemm_mfma_kernel<1, false><<<grid, 256, 0, st>>>(ld, A, B, C);
        else dgemm_mfma_kernel<0, false><<<grid, 256, 0, st>>>(ld, A, B, C);
    }
    GTF_CHECK(hipGetLastError());
    return 0;
}

int gtf_dgemm(hipStream_t st, int n, int ld, const double* A, const double* B, double* C, int transB)
{
    return gtf_gemm(st, n, ld, A, B, C, transB, 0);
}

// ---- elementwise ------------------------------------------------------------

__global__ void fill_kernel(double* x, long n, double v)
{
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) x[i] = v;
}
__global__ void fill_int_kernel(int* x, long n, int v)
{
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) x[i] = v;
}
__global__ void add_kernel(const double* A, const double* B, double* C, long n)
{
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        C[i] = A[i] + B[i];
}
// A <- (A + A^T)/2 over the strict lower triangle (GA_Symmetrize, pfock/pfock.c:1721-1724)
__global__ void symmetrize_kernel(double* A, int n, int ld)
{
    const int i = blockIdx.y * 16 + threadIdx.y, j = blockIdx.x * 16 + threadIdx.x;
    if (i >= n || j >= i) return;
    const double s = 0.5 * (A[(long)i * ld + j] + A[(long)j * ld + i]);
    A[(long)i * ld + j] = s;
    A[(long)j * ld + i] = s;
}

static int grid_for(long n) { return (int)std::min<long>((n + 255) / 256, 4096); }

// per-molecule blocks of a packed array: A_m <- (A_m + A_m^T)/2
__global__ void symmetrize_packed_kernel(double* A, const long* __restrict__ off, const int* __restrict__ nn)
{
    const int m = blockIdx.y;
    const int n = nn[m];
    double* B = A + off[m];
    for (long k = (long)blockIdx.x * blockDim.x + threadIdx.x; k < (long)n * n; k += (long)gridDim.x * blockDim.x) {
        const int i = (int)(k / n), j = (int)(k - (long)i * n);
        if (j >= i) continue;
        const double s = 0.5 * (B[(long)i * n + j] + B[(long)j * n + i]);
        B[(long)i * n + j] = s;
        B[(long)j * n + i] = s;
    }
}

int gtf_symmetrize_packed(hipStream_t st, double* A, int nmol, const long* off, const int* n, int maxn)
{
    if (nmol <= 0) return 0;
    const int gx = (int)std::min<long>(((long)maxn * maxn + 255) / 256, 1024);
    symmetrize_packed_kernel<<<dim3(gx, nmol), 256, 0, st>>>(A, off, n);
    GTF_CHECK(hipGetLastError());
    return 0;
}

int gtf_fill(hipStream_t st, double* x, long n, double v)
{
    if (n <= 0) return 0;
    fill_kernel<<<grid_for(n), 256, 0, st>>>(x, n, v);
    GTF_CHECK(hipGetLastError());
    return 0;
}
int gtf_fill_int(hipStream_t st, int* x, long n, int v)
{
    if (n <= 0) return 0;
    fill_int_kernel<<<grid_for(n), 256, 0, st>>>(x, n, v);
    GTF_CHECK(hipGetLastError());
    return 0;
}
int gtf_add(hipStream_t st, const double* A, const double* B, double* C, long n)
{
    if (n <= 0) return 0;
    add_kernel<<<grid_for(n), 256, 0, st>>>(A, B, C, n);
    GTF_CHECK(hipGetLastError());
    return 0;
}
int gtf_symmetrize(hipStream_t st, double* A, int n, int ld)
{
    dim3 g((n + 15) / 16, (n + 15) / 16), b(16, 16);
    symmetrize_kernel<<<g, b, 0, st>>>(A, n, ld);
    GTF_CHECK(hipGetLastError());
    return 0;
}

// ---- purification (pscf/purif.c:257-488) --------------------------------------

// Gershgorin bounds and trace: out = {hmin, hmax, trF} (pscf/purif.c:302-358)
__global__ __launch_bounds__(1024) void gershgorin_kernel(const double* F, int n, int ld, double* out)
{
    __shared__ double smin[1024], smax[1024], ssum[1024];
    double mn = 1e300, mx = -1e300, tr = 0.0;
    for (int i = threadIdx.x; i < n; i += 1024) {
        double off = 0.0;
        for (int j = 0; j < n; j++) off += fabs(F[(long)j * ld + i]);
        const double dg = F[(long)i * ld + i];
        off -= fabs(dg);
        mn = fmin(mn, dg - off);
        mx = fmax(mx, dg + off);
        tr += dg;
    }
    smin[threadIdx.x] = mn;
    smax[threadIdx.x] = mx;
    ssum[threadIdx.x] = tr;
    __syncthreads();
    for (int s = 512; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            smin[threadIdx.x] = fmin(smin[threadIdx.x], smin[threadIdx.x + s]);
            smax[threadIdx.x] = fmax(smax[threadIdx.x], smax[threadIdx.x + s]);
            ssum[threadIdx.x] += ssum[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[0] = smin[0];
        out[1] = smax[0];
        out[2] = ssum[0];
    }
}

// D = (lambda mu / n + nocc / n) I - (lambda / n) F   (pscf/purif.c:367-373)
__global__ void purif_init_kernel(const double* F, double* D, int n, int ld, double a, double dshift)
{
    const long total = (long)ld * ld;
    for (long k = (long)blockIdx.x * blockDim.x + threadIdx.x; k < total; k += (long)gridDim.x * blockDim.x) {
        const int i = (int)(k / ld), j = (int)(k % ld);
        double v = 0.0;
        if (i < n && j < n) v = a * F[k] + (i == j ? dshift : 0.0);
        D[k] = v;
    }
}

// per-block partial sums of ||D - D2||_F^2, tr(D2 - D3), tr(D - D2)
__global__ __launch_bounds__(256) void purif_stats_kernel(const double* D, const double* D2, const double* D3, int n,
                                                          int ld, double* partial)
{
    __shared__ double s0[256], s1[256], s2[256];
    double e = 0.0, t1 = 0.0, t2 = 0.0;
    const long total = (long)ld * ld;
    for (long k = (long)blockIdx.x * 256 + threadIdx.x; k < total; k += (long)gridDim.x * 256) {
        const double d = D[k] - D2[k];
        e += d * d;
    }
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        const long k = (long)i * ld + i;
        t1 += D2[k] - D3[k];
        t2 += D[k] - D2[k];
    }
    s0[threadIdx.x] = e;
    s1[threadIdx.x] = t1;
    s2[threadIdx.x] = t2;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            s0[threadIdx.x] += s0[threadIdx.x + s];
            s1[threadIdx.x] += s1[threadIdx.x + s];
            s2[threadIdx.x] += s2[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        partial[3 * blockIdx.x + 0] = s0[0];
        partial[3 * blockIdx.x + 1] = s1[0];
        partial[3 * blockIdx.x + 2] = s2[0];
    }
}

// sum nparts rows of width w in a fixed order
__global__ void sum_partials_kernel(const double* partial, int nparts, int w, double* out)
{
    const int c = threadIdx.x;
    if (c >= w) return;
    double s = 0.0;
    for (int b = 0; b < nparts; b++) s += partial[(long)b * w + c];
    out[c] = s;
}

// McWeeny / Palser-Manolopoulos update (pscf/purif.c:432-453)
__global__ void purif_update_kernel(double* D, const double* D2, const double* D3, long total, double c)
{
    for (long k = (long)blockIdx.x * blockDim.x + threadIdx.x; k < total; k += (long)gridDim.x * blockDim.x) {
        if (c < 0.5) D[k] = ((1.0 - 2.0 * c) * D[k] + (1.0 + c) * D2[k] - D3[k]) / (1.0 - c);
        else D[k] = ((1.0 + c) * D2[k] - D3[k]) / c;
    }
}

// E = W - W^T  (pscf/purif.c:548-560)
__global__ void antisym_kernel(const double* W, double* E, int ld)
{
    const long total = (long)ld * ld;
    for (long k = (long)blockIdx.x * blockDim.x + threadIdx.x; k < total; k += (long)gridDim.x * blockDim.x) {
        const int i = (int)(k / ld), j = (int)(k % ld);
        E[k] = W[k] - W[(long)j * ld + i];
    }
}

// partial dot products of x with m vectors (each of length total, stride total)
__global__ __launch_bounds__(256) void multidot_kernel(const double* x, const double* vecs, int m, long total,
                                                       double* partial)
{
    __shared__ double s[MAX_DOTS][256];
    double acc[MAX_DOTS];
    for (int v = 0; v < MAX_DOTS; v++) acc[v] = 0.0;
    for (long k = (long)blockIdx.x * 256 + threadIdx.x; k < total; k += (long)gridDim.x * 256) {
        const double xv = x[k];
        for (int v = 0; v < m; v++) acc[v] += xv * vecs[(long)v * total + k];
    }
    for (int v = 0; v < MAX_DOTS; v++) s[v][threadIdx.x] = acc[v];
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if (threadIdx.x < st)
            for (int v = 0; v < m; v++) s[v][threadIdx.x] += s[v][threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int v = 0; v < m; v++) partial[(long)blockIdx.x * MAX_DOTS + v] = s[v][0];
}

// F = sum_i c_i Fvecs_i  (pscf/purif.c:629-635)
__global__ void lincomb_kernel(double* F, const double* vecs, const double* coef, int m, long total)
{
    for (long k = (long)blockIdx.x * blockDim.x + threadIdx.x; k < total; k += (long)gridDim.x * blockDim.x) {
        double s = 0.0;
        for (int v = 0; v < m; v++) s += coef[v] * vecs[(long)v * total + k];
        F[k] = s;
    }
}

#define NPART 256

int gtf_gershgorin(hipStream_t st, const double* F, int n, int ld, double* dout3)
{
    gershgorin_kernel<<<1, 1024, 0, st>>>(F, n, ld, dout3);
    GTF_CHECK(hipGetLastError());
    return 0;
}
int gtf_purif_init(hipStream_t st, const double* F, double* D, int n, int ld, double a, double dshift)
{
    purif_init_kernel<<<grid_for((long)ld * ld), 256, 0, st>>>(F, D, n, ld, a, dshift);
    GTF_CHECK(hipGetLastError());
    return 0;
}
int gtf_purif_stats(hipStream_t st, const double* D, const double* D2, const double* D3, int n, int ld,
                    double* partial, double* dout3)
{
    purif_stats_kernel<<<NPART, 256, 0, st>>>(D, D2, D3, n, ld, partial);
    GTF_CHECK(hipGetLastError());
    sum_partials_kernel<<<1, 64, 0, st>>>(partial, NPART, 3, dout3);
    GTF_CHECK(hipGetLastError());
    return 0;
}
int gtf_purif_update(hipStream_t st, double* D, const double* D2, const double* D3, long total, double c)
{
    purif_update_kernel<<<grid_for(total), 256, 0, st>>>(D, D2, D3, total, c);
    GTF_CHECK(hipGetLastError());
    return 0;
}
int gtf_antisym(hipStream_t st, const double* W, double* E, int ld)
{
    antisym_kernel<<<grid_for((long)ld * ld), 256, 0, st>>>(W, E, ld);
    GTF_CHECK(hipGetLastError());
    return 0;
}
int gtf_multidot(hipStream_t st, const double* x, const double* vecs, int m, long total, double* partial,
                 double* dout)
{
    if (m <= 0) return 0;
    multidot_kernel<<<NPART, 256, 0, st>>>(x, vecs, m, total, partial);
    GTF_CHECK(hipGetLastError());
    sum_partials_kernel<<<1, 64, 0, st>>>(partial, NPART, MAX_DOTS, dout);
    GTF_CHECK(hipGetLastError());
    return 0;
}
int gtf_lincomb(hipStream_t st, double* F, const double* vecs, const double* coef, int m, long total)
{
    lincomb_kernel<<<grid_for(total), 256, 0, st>>>(F, vecs, coef, m, total);
    GTF_CHECK(hipGetLastError());
    return 0;
}
size_t gtf_partial_bytes() { return sizeof(double) * NPART * MAX_DOTS; }

// host-array GEMM for tests: C = A * B (transB = 0) or A * B^T (transB = 1),
// n x n row-major, padded internally to a multiple of 64
extern "C" int gtf_gemm_host(int n, const double* A, const double* B, double* C, int transB, int fp32);

extern "C" int gtf_dgemm_host(int n, const double* A, const double* B, double* C, int transB)
{
    return gtf_gemm_host(n, A, B, C, transB, 0);
}

extern "C" int gtf_gemm_host(int n, const double* A, const double* B, double* C, int transB, int fp32)
{
    const int ld = ((n + 63) / 64) * 64;
    const size_t m = (size_t)ld * ld;
    double *dA = nullptr, *dB = nullptr, *dC = nullptr;
    if (hipMalloc(&dA, m * 8) != hipSuccess || hipMalloc(&dB, m * 8) != hipSuccess ||
        hipMalloc(&dC, m * 8) != hipSuccess) {
        (void)hipFree(dA); (void)hipFree(dB); (void)hipFree(dC);
        return -2;
    }
    int rc = 0;
    if (hipMemset(dA, 0, m * 8) != hipSuccess || hipMemset(dB, 0, m * 8) != hipSuccess ||
        hipMemcpy2D(dA, ld * 8, A, n * 8, n * 8, n, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy2D(dB, ld * 8, B, n * 8, n * 8, n, hipMemcpyHostToDevice) != hipSuccess)
        rc = -1;
    if (!rc) rc = gtf_gemm(nullptr, n, ld, dA, dB, dC, transB, fp32);
    if (!rc && (hipDeviceSynchronize() != hipSuccess ||
                hipMemcpy2D(C, n * 8, dC, ld * 8, n * 8, n, hipMemcpyDeviceToHost) != hipSuccess))
        rc = -1;
    (void)hipFree(dA); (void)hipFree(dB); (void)hipFree(dC);
    return rc;
}
