// noalloc.hip -- the split-BF16 product (ozaki.hip) without allocation.
#include <mutex>

#include "ozaki_launch.h"

// gtf_gemm_split_noalloc multiplies caller-owned device matrices with its
// scratch (operand slices, exponents, max partials) in one static device
// array per GPU, for callers that cannot allocate (code under graph capture,
// serving loops, harnesses that forbid hipMalloc).  Its size covers
// ld <= OZ_NOALLOC_MAX_LD with OZ_SLICES_MAX slices.  Host code never passes
// the array's name to a kernel: it uses the device address that
// hipGetSymbolAddress returns, looked up once per device.
#define OZ_NOALLOC_MAX_LD 1536
#define OZ_NOALLOC_BYTES (2 * 2 * OZ_SLICES_MAX * OZ_NOALLOC_MAX_LD * OZ_NOALLOC_MAX_LD + (1 << 20))

static __device__ __align__(256) char scratch[OZ_NOALLOC_BYTES];

#define OZ_MAX_DEVICES 64
static void* g_scratch_addr[OZ_MAX_DEVICES];
static std::mutex g_scratch_mu;

static void* scratch_address(int dev)
{
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_addr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            fprintf(stderr, "gtf_gemm_split_noalloc: hipGetSymbolAddress failed (%s)\n", hipGetErrorString(e));
            return nullptr;
        }
        g_scratch_addr[dev] = p;
    }
    return g_scratch_addr[dev];
}

extern "C" int gtf_gemm_split_noalloc(int n, int ld, const double* A, const double* B, double* C, int transB, int ns,
                                      void* stream)
{
    if (n <= 0 || ld < n || ld % 64 != 0 || !A || !B || !C || ns < OZ_SLICES_MIN || ns > OZ_SLICES_MAX) return -3;
    if (ld > OZ_NOALLOC_MAX_LD) {
        fprintf(stderr, "gtf_gemm_split_noalloc: leading dimension %d exceeds the static scratch (at most %d)\n",
                ld, OZ_NOALLOC_MAX_LD);
        return -3;
    }
    const long total = (long)ld * ld;
    if (gtf_oz_scratch_bytes(total, ns) > (size_t)OZ_NOALLOC_BYTES) return -3;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= OZ_MAX_DEVICES) return -1;
    void* base = scratch_address(dev);
    if (!base) return -1;
    OzScratch s;
    gtf_oz_scratch_carve(base, total, ns, s);
    OzOperands o;
    o.maxld = ld;
    return gtf_oz_product((hipStream_t)stream, o, s, ns, A, B, C, transB) ? -1 : 0;
}
