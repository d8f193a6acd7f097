// hostcopy.h -- row blocks between caller (pageable) host memory and device
// matrices through a pinned staging buffer: the host copies the rows into
// pinned memory and the DMA engine moves them in one pitched copy (a pitched
// copy straight from pageable memory is staged row by row by the runtime,
// about ten times slower for an 846 x 846 block).  Both directions are
// synchronous on the stream.
#pragma once
#include <hip/hip_runtime.h>

#include <cstring>

struct PinnedStage {
    double* buf = nullptr;
    size_t cap = 0;  // doubles
    PinnedStage() = default;
    PinnedStage(const PinnedStage&) = delete;
    PinnedStage& operator=(const PinnedStage&) = delete;
    ~PinnedStage()
    {
        if (buf) (void)hipHostFree(buf);
    }
    bool reserve(size_t n)
    {
        if (n <= cap) return true;
        if (buf) (void)hipHostFree(buf);
        buf = nullptr;
        cap = 0;
        if (hipHostMalloc((void**)&buf, n * sizeof(double), hipHostMallocDefault) != hipSuccess) {
            buf = nullptr;
            return false;
        }
        cap = n;
        return true;
    }
};

// dst (device, row pitch dld doubles) <- src (host, row pitch sld): rows x cols
inline hipError_t stage_h2d(PinnedStage& ps, hipStream_t st, double* dst, long dld, const double* src, long sld,
                            int rows, int cols)
{
    if (rows <= 0 || cols <= 0) return hipSuccess;
    if (!ps.reserve((size_t)rows * cols)) return hipErrorOutOfMemory;
    for (int r = 0; r < rows; r++) memcpy(ps.buf + (size_t)r * cols, src + (size_t)r * sld, sizeof(double) * cols);
    const hipError_t e = hipMemcpy2DAsync(dst, sizeof(double) * dld, ps.buf, sizeof(double) * cols,
                                          sizeof(double) * cols, rows, hipMemcpyHostToDevice, st);
    if (e != hipSuccess) return e;
    return hipStreamSynchronize(st);
}

// dst (host, row pitch dld doubles) <- src (device, row pitch sld): rows x cols
inline hipError_t stage_d2h(PinnedStage& ps, hipStream_t st, double* dst, long dld, const double* src, long sld,
                            int rows, int cols)
{
    if (rows <= 0 || cols <= 0) return hipSuccess;
    if (!ps.reserve((size_t)rows * cols)) return hipErrorOutOfMemory;
    hipError_t e = hipMemcpy2DAsync(ps.buf, sizeof(double) * cols, src, sizeof(double) * sld, sizeof(double) * cols,
                                    rows, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return e;
    for (int r = 0; r < rows; r++) memcpy(dst + (size_t)r * dld, ps.buf + (size_t)r * cols, sizeof(double) * cols);
    return hipSuccess;
}
