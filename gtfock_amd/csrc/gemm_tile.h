// gemm_tile.h -- one 64 x 64 tile of C = A * B (TRANSB = 0) or A * B^T
// (TRANSB = 1) for row-major matrices of leading dimension ld (a multiple of
// 64, zero padding), by a 256-thread workgroup: each of the four waves owns a
// 32 x 32 quadrant as 2 x 2 MFMA tiles of 16 x 16, fed from double-buffered
// LDS tiles of depth 16.  F32 = false: v_mfma_f64_16x16x4_f64 on FP64 data.  F32 = true: the
// operands are rounded to FP32 in LDS and multiplied with
// v_mfma_f32_16x16x4_f32 (FP32 accumulation); C is stored as FP64.
//
// Operand lanes (both instructions): A[i][k] at lane i + 16 k, B[k][j] at
// lane j + 16 k.  Result of lane l, register r: FP64 row l/16 + 4 r, FP32 row
// 4 (l/16) + r; column l % 16 in both.
#pragma once
#include "gtf_internal.h"

typedef double gtf_double4_t __attribute__((ext_vector_type(4)));
typedef float gtf_float4_t __attribute__((ext_vector_type(4)));

#define GEMM_GT 64
#define GEMM_KT 16

// Software pipelined: the global loads of K block kb + 1 are issued into
// registers before block kb is multiplied out of LDS, and stored into the
// other LDS buffer after it, so their latency overlaps the MFMAs and one
// barrier per block suffices (a buffer is rewritten only after the barrier
// that follows its last read).
template <int TRANSB, bool F32>
__device__ __forceinline__ void gemm_tile(int ld, const double* __restrict__ A, const double* __restrict__ B,
                                          double* __restrict__ C, int row0, int col0)
{
    using T = typename std::conditional<F32, float, double>::type;
    __shared__ T As[2][GEMM_GT][GEMM_KT + 1];
    __shared__ T Bs[2][GEMM_KT][GEMM_GT + 1];
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int wr = wave >> 1, wc = wave & 1;
    const int li = lane & 15, lk = lane >> 4;
    typename std::conditional<F32, gtf_float4_t, gtf_double4_t>::type acc[2][2];
    for (int i = 0; i < 2; i++)
        for (int j = 0; j < 2; j++)
            for (int r = 0; r < 4; r++) acc[i][j][r] = T(0);
    // this thread's four A and four B elements of a K block
    double ra[4], rb[4];
    auto load = [&](int k0) {
        for (int r = 0; r < 4; r++) {
            const int rr = (tid >> 4) + 16 * r, kk = tid & 15;
            ra[r] = A[(long)(row0 + rr) * ld + k0 + kk];
            if (TRANSB) rb[r] = B[(long)(col0 + rr) * ld + k0 + kk];
            else rb[r] = B[(long)(k0 + (tid >> 6) + 4 * r) * ld + col0 + (tid & 63)];
        }
    };
    auto store = [&](int buf) {
        for (int r = 0; r < 4; r++) {
            const int rr = (tid >> 4) + 16 * r, kk = tid & 15;
            As[buf][rr][kk] = T(ra[r]);
            if (TRANSB) Bs[buf][kk][rr] = T(rb[r]);
            else Bs[buf][(tid >> 6) + 4 * r][tid & 63] = T(rb[r]);
        }
    };
    load(0);
    store(0);
    __syncthreads();
    const int nk = ld / GEMM_KT;
    for (int kb = 0; kb < nk; kb++) {
        const int cur = kb & 1;
        if (kb + 1 < nk) load((kb + 1) * GEMM_KT);
        for (int kk = 0; kk < GEMM_KT; kk += 4) {
            T a[2], b[2];
            for (int t = 0; t < 2; t++) {
                a[t] = As[cur][wr * 32 + t * 16 + li][kk + lk];
                b[t] = Bs[cur][kk + lk][wc * 32 + t * 16 + li];
            }
            for (int i = 0; i < 2; i++)
                for (int j = 0; j < 2; j++) {
                    if constexpr (F32) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[i], b[j], acc[i][j], 0, 0, 0);
                    else acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], b[j], acc[i][j], 0, 0, 0);
                }
        }
        if (kb + 1 < nk) store(cur ^ 1);
        __syncthreads();
    }
    for (int i = 0; i < 2; i++)
        for (int j = 0; j < 2; j++)
            for (int r = 0; r < 4; r++) {
                const int row = row0 + wr * 32 + i * 16 + (F32 ? 4 * lk + r : lk + 4 * r);
                const int col = col0 + wc * 32 + j * 16 + li;
                C[(long)row * ld + col] = (double)acc[i][j][r];
            }
}
