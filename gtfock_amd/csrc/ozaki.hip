// ozaki.hip -- split-BF16 MFMA products of FP64 matrices (ozaki.h) for the
// purification, for one matrix (poff == nullptr, leading dimension ld0) or
// for every molecule of a batch (padded blocks at poff[m], leading dimension
// pld[m], molecule = blockIdx.z).
//
// Per-molecule scratch, at offsets derived from poff[m]: slices at
// ns * poff[m] (BF16), exponents and max partials at poff[m] / 64.
#include "ozaki.h"
#include "ozaki_launch.h"

__device__ __forceinline__ bool oz_skip(int m, const int* __restrict__ skip, const int* __restrict__ done,
                                        const int* __restrict__ sel)
{
    return (skip && skip[m]) || (done && done[m] >= 0) || (sel && sel[m] != GTF_LOW_SPLIT);
}

__global__ __launch_bounds__(256) void oz_absmax_kernel(const long* __restrict__ poff, const int* __restrict__ pld,
                                                        int ld0, const int* __restrict__ skip,
                                                        const int* __restrict__ done, const int* __restrict__ sel,
                                                        const double* __restrict__ X,
                                                        double* __restrict__ prow, double* __restrict__ pcol)
{
    const int m = blockIdx.z;
    if (oz_skip(m, skip, done, sel)) return;
    const OzView v = oz_view(m, poff, pld, ld0);
    if ((int)blockIdx.y * 64 >= v.ld || (int)blockIdx.x * 64 >= v.ld) return;
    oz_absmax_tile(v.ld, X + v.off, prow + v.off / 64, pcol + v.off / 64, blockIdx.y, blockIdx.x);
}

template <int TRANS>
__global__ __launch_bounds__(256) void oz_split_kernel(const long* __restrict__ poff, const int* __restrict__ pld,
                                                       int ld0, const int* __restrict__ skip,
                                                       const int* __restrict__ done, const int* __restrict__ sel,
                                                       const double* __restrict__ X,
                                                       const double* __restrict__ part, int ns,
                                                       __bf16* __restrict__ Q, int* __restrict__ E)
{
    const int m = blockIdx.z;
    if (oz_skip(m, skip, done, sel)) return;
    const OzView v = oz_view(m, poff, pld, ld0);
    if ((int)blockIdx.y * 64 >= v.ld || (int)blockIdx.x * 64 >= v.ld) return;
    oz_split_tile<TRANS>(v.ld, X + v.off, part + v.off / 64, ns, Q + (long)ns * v.off, E + v.off / 64, blockIdx.y,
                         blockIdx.x);
}

template <int NS>
__global__ __launch_bounds__(256) void oz_gemm_kernel(const long* __restrict__ poff, const int* __restrict__ pld,
                                                      int ld0, const int* __restrict__ skip,
                                                      const int* __restrict__ done, const int* __restrict__ sel,
                                                      const __bf16* __restrict__ QA,
                                                      const int* __restrict__ EA, const __bf16* __restrict__ QB,
                                                      const int* __restrict__ EB, double* __restrict__ C)
{
    const int m = blockIdx.z;
    if (oz_skip(m, skip, done, sel)) return;
    const OzView v = oz_view(m, poff, pld, ld0);
    if ((int)blockIdx.y * 64 >= v.ld || (int)blockIdx.x * 64 >= v.ld) return;
    oz_gemm_tile<NS>(v.ld, QA + (long)NS * v.off, EA + v.off / 64, QB + (long)NS * v.off, EB + v.off / 64, C + v.off,
                     blockIdx.y * 64, blockIdx.x * 64);
}

static bool oz_shape_ok(const OzOperands& o)
{
    if (o.nmol <= 0 || o.maxld <= 0 || o.maxld % 64 != 0 || (!o.poff && o.nmol != 1)) {
        fprintf(stderr, "gtf_oz: invalid shape (nmol %d, ld %d)\n", o.nmol, o.maxld);
        return false;
    }
    return true;
}

int gtf_oz_split(hipStream_t st, const OzOperands& o, const double* X, int trans, int ns, double* prow,
                 double* pcol, __bf16* Q, int* E, bool maxima_ready)
{
    if (!oz_shape_ok(o) || ns < 1 || ns > OZ_MAX_SLICES) return -1;
    const dim3 grid(o.maxld / 64, o.maxld / 64, o.nmol);
    const int ld0 = o.poff ? 0 : o.maxld;
    if (!maxima_ready) {
        oz_absmax_kernel<<<grid, 256, 0, st>>>(o.poff, o.pld, ld0, o.skip, o.done, o.sel, X, prow, pcol);
        GTF_CHECK(hipGetLastError());
    }
    if (trans) oz_split_kernel<1><<<grid, 256, 0, st>>>(o.poff, o.pld, ld0, o.skip, o.done, o.sel, X, pcol, ns, Q, E);
    else oz_split_kernel<0><<<grid, 256, 0, st>>>(o.poff, o.pld, ld0, o.skip, o.done, o.sel, X, prow, ns, Q, E);
    GTF_CHECK(hipGetLastError());
    return 0;
}

int gtf_oz_gemm(hipStream_t st, const OzOperands& o, int ns, const __bf16* QA, const int* EA, const __bf16* QB,
                const int* EB, double* C)
{
    if (!oz_shape_ok(o)) return -1;
    const dim3 grid(o.maxld / 64, o.maxld / 64, o.nmol);
    const int ld0 = o.poff ? 0 : o.maxld;
    switch (ns) {
    case 3: oz_gemm_kernel<3><<<grid, 256, 0, st>>>(o.poff, o.pld, ld0, o.skip, o.done, o.sel, QA, EA, QB, EB, C); break;
    case 4: oz_gemm_kernel<4><<<grid, 256, 0, st>>>(o.poff, o.pld, ld0, o.skip, o.done, o.sel, QA, EA, QB, EB, C); break;
    case 5: oz_gemm_kernel<5><<<grid, 256, 0, st>>>(o.poff, o.pld, ld0, o.skip, o.done, o.sel, QA, EA, QB, EB, C); break;
    case 6: oz_gemm_kernel<6><<<grid, 256, 0, st>>>(o.poff, o.pld, ld0, o.skip, o.done, o.sel, QA, EA, QB, EB, C); break;
    default: fprintf(stderr, "gtf_oz_gemm: %d slices (3..6 supported)\n", ns); return -1;
    }
    GTF_CHECK(hipGetLastError());
    return 0;
}

int gtf_oz_product(hipStream_t st, const OzOperands& o, OzScratch& s, int ns, const double* A, const double* B,
                   double* C, int transB)
{
    // A by rows; B^T by rows: the columns of B (transB = 0) or the rows of B
    if (gtf_oz_split(st, o, A, 0, ns, s.prow, s.pcol, s.QA, s.EA, false)) return -1;
    const bool same = (A == B);
    if (gtf_oz_split(st, o, B, transB ? 0 : 1, ns, s.prow, s.pcol, s.QB, s.EB, same)) return -1;
    return gtf_oz_gemm(st, o, ns, s.QA, s.EA, s.QB, s.EB, C);
}

int gtf_oz_square_cube(hipStream_t st, const OzOperands& o, OzScratch& s, const double* D, double* D2, double* D3)
{
    // D2 = D * D, D3 = D * D2 (pscf/pdgemm.c:53-161): D's row slices serve
    // both products; the second B^T operand is D2's columns
    const int ns = s.ns;
    if (gtf_oz_split(st, o, D, 0, ns, s.prow, s.pcol, s.QA, s.EA, false) ||
        gtf_oz_split(st, o, D, 1, ns, s.prow, s.pcol, s.QB, s.EB, true) ||
        gtf_oz_gemm(st, o, ns, s.QA, s.EA, s.QB, s.EB, D2) ||
        gtf_oz_split(st, o, D2, 1, ns, s.prow, s.pcol, s.QB, s.EB, false) ||
        gtf_oz_gemm(st, o, ns, s.QA, s.EA, s.QB, s.EB, D3))
        return -1;
    return 0;
}

size_t gtf_oz_scratch_bytes(long padded_total, int ns)
{
    // QA, QB: ns slices each; EA, EB ints and prow, pcol doubles at 1/64
    return 2 * sizeof(__bf16) * (size_t)ns * padded_total + 2 * sizeof(int) * (size_t)(padded_total / 64 + 64) +
           2 * sizeof(double) * (size_t)(padded_total / 64 + 64) + 256 * 6;
}

int gtf_oz_scratch_carve(void* base, long padded_total, int ns, OzScratch& s)
{
    char* p = (char*)base;
    auto take = [&](size_t bytes) {
        char* r = p;
        p += (bytes + 255) / 256 * 256;
        return (void*)r;
    };
    s.QA = (__bf16*)take(sizeof(__bf16) * (size_t)ns * padded_total);
    s.QB = (__bf16*)take(sizeof(__bf16) * (size_t)ns * padded_total);
    s.EA = (int*)take(sizeof(int) * (size_t)(padded_total / 64 + 64));
    s.EB = (int*)take(sizeof(int) * (size_t)(padded_total / 64 + 64));
    s.prow = (double*)take(sizeof(double) * (size_t)(padded_total / 64 + 64));
    s.pcol = (double*)take(sizeof(double) * (size_t)(padded_total / 64 + 64));
    s.ns = ns;
    return 0;
}

// host-array product for tests: C = A * B (transB = 0) or A * B^T (1),
// n x n row-major, with ns slices
extern "C" int gtf_gemm_split_host(int n, const double* A, const double* B, double* C, int transB, int ns)
{
    if (n <= 0 || ns < 3 || ns > OZ_MAX_SLICES) return -3;
    const int ld = ((n + 63) / 64) * 64;
    const size_t m = (size_t)ld * ld;
    double *dA = nullptr, *dB = nullptr, *dC = nullptr;
    void* scratch = nullptr;
    const size_t sb = gtf_oz_scratch_bytes((long)m, ns);
    if (hipMalloc(&dA, m * 8) != hipSuccess || hipMalloc(&dB, m * 8) != hipSuccess ||
        hipMalloc(&dC, m * 8) != hipSuccess || hipMalloc(&scratch, sb) != hipSuccess) {
        (void)hipFree(dA); (void)hipFree(dB); (void)hipFree(dC); (void)hipFree(scratch);
        return -2;
    }
    int rc = 0;
    if (hipMemset(dA, 0, m * 8) != hipSuccess || hipMemset(dB, 0, m * 8) != hipSuccess ||
        hipMemcpy2D(dA, ld * 8, A, n * 8, n * 8, n, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy2D(dB, ld * 8, B, n * 8, n * 8, n, hipMemcpyHostToDevice) != hipSuccess)
        rc = -1;
    OzOperands o;
    o.nmol = 1;
    o.maxld = ld;
    OzScratch s;
    gtf_oz_scratch_carve(scratch, (long)m, ns, s);
    if (!rc) rc = gtf_oz_product(nullptr, o, s, ns, dA, dB, dC, transB);
    if (!rc && (hipDeviceSynchronize() != hipSuccess ||
                hipMemcpy2D(C, n * 8, dC, ld * 8, n * 8, n, hipMemcpyDeviceToHost) != hipSuccess))
        rc = -1;
    (void)hipFree(dA); (void)hipFree(dB); (void)hipFree(dC); (void)hipFree(scratch);
    return rc;
}
