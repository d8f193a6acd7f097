// gemm_tile.h -- one 64 x 64 tile of C = A * B (TRANSB = 0) or A * B^T
// (TRANSB = 1) for row-major matrices of leading dimension ld (a multiple of
// 64, zero padding), by a 256-thread workgroup: each of the four waves owns a
// 32 x 32 quadrant as 2 x 2 MFMA tiles of 16 x 16, fed from LDS tiles of
// depth 16.  F32 = false: v_mfma_f64_16x16x4_f64 on FP64 data, three LDS
// stages filled by direct global -> LDS loads.  F32 = true: the operands are
// rounded to FP32 in double-buffered LDS tiles and multiplied with
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

// FP64 form: the K blocks stream through three LDS stages filled by direct
// global -> LDS loads (global_load_lds_dwordx4, two doubles per lane), issued
// two blocks ahead of the block being multiplied.  A stage holds the 64 x 16 A
// block (rows of 128 B whose 16-B chunks sit at position chunk ^ (row & 7),
// so that the sixteen rows a fragment read touches spread over the banks)
// and the B block: 16 x 64 as is (TRANSB = 0), or 64 x 16 like A (B^T rows).
typedef __attribute__((address_space(3))) void gemm_lds_void;

struct GemmStage {
    double a[GEMM_GT][GEMM_KT];
    double b[GEMM_GT * GEMM_KT];
};

template <int TRANSB>
__device__ __forceinline__ void gemm_tile_f64(int ld, const double* __restrict__ A, const double* __restrict__ B,
                                              double* __restrict__ C, int row0, int col0)
{
    __shared__ __align__(16) GemmStage st0, st1, st2;
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int wr = wave >> 1, wc = wave & 1;
    const int li = lane & 15, lk = lane >> 4;
    const int nblk = ld / GEMM_KT;
    gtf_double4_t acc[2][2];
    for (int i = 0; i < 2; i++)
        for (int j = 0; j < 2; j++)
            for (int r = 0; r < 4; r++) acc[i][j][r] = 0.0;
    // each wave loads two 1-KB groups of each operand block per stage
    auto issue = [&](GemmStage& S, int k0) {
        for (int q = 0; q < 2; q++) {
            const int g = wave * 2 + q;
            const int row = g * 8 + (lane >> 3);
            const int ch = (lane & 7) ^ (row & 7);
            __builtin_amdgcn_global_load_lds(A + (long)(row0 + row) * ld + k0 + 2 * ch, (gemm_lds_void*)&S.a[g * 8][0],
                                             16, 0, 0);
            if (TRANSB)
                __builtin_amdgcn_global_load_lds(B + (long)(col0 + row) * ld + k0 + 2 * ch,
                                                 (gemm_lds_void*)&S.b[g * 8 * GEMM_KT], 16, 0, 0);
            else
                __builtin_amdgcn_global_load_lds(B + (long)(k0 + g * 2 + (lane >> 5)) * ld + col0 + 2 * (lane & 31),
                                                 (gemm_lds_void*)&S.b[g * 2 * GEMM_GT], 16, 0, 0);
        }
    };
    auto compute = [&](const GemmStage& S) {
        for (int kk = 0; kk < GEMM_KT; kk += 4) {
            const int k = kk + lk, c = k >> 1, h = k & 1;
            double a[2], b[2];
            for (int t = 0; t < 2; t++) {
                const int ra = wr * 32 + t * 16 + li, rb = wc * 32 + t * 16 + li;
                a[t] = S.a[ra][2 * (c ^ (ra & 7)) + h];
                b[t] = TRANSB ? S.b[rb * GEMM_KT + 2 * (c ^ (rb & 7)) + h] : S.b[k * GEMM_GT + rb];
            }
            for (int i = 0; i < 2; i++)
                for (int j = 0; j < 2; j++) acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], b[j], acc[i][j], 0, 0, 0);
        }
    };
    // block kb from stage cur, block kb + 2 into stage nxt (as in ozaki.h).
    // __syncthreads: every wave's loads have landed (it waits for all of
    // them, kb + 1's included) and every wave's LDS reads of the previous
    // step are done before any wave refills that step's stage; a bare
    // s_barrier would let the compiler move LDS reads across it
    auto step = [&](const GemmStage& cur, GemmStage& nxt, int kb) {
        __syncthreads();
        issue(nxt, (kb + 2) * GEMM_KT);
        compute(cur);
    };
    auto step_tail = [&](const GemmStage& cur, GemmStage& nxt, int kb) {
        __syncthreads();
        if (kb + 2 < nblk) issue(nxt, (kb + 2) * GEMM_KT);
        compute(cur);
    };
    issue(st0, 0);
    if (nblk > 1) issue(st1, GEMM_KT);
    int kb = 0;
    if (nblk >= 5) {  // the stages in a fixed rotation, three blocks per pass; at most four left for the tail
        step(st0, st2, 0);
        step(st1, st0, 1);
        step(st2, st1, 2);
        for (kb = 3; kb + 5 <= nblk; kb += 3) {
            step(st0, st2, kb);
            step(st1, st0, kb + 1);
            step(st2, st1, kb + 2);
        }
    }
    if (kb < nblk) step_tail(st0, st2, kb);
    if (kb + 1 < nblk) step_tail(st1, st0, kb + 1);
    if (kb + 2 < nblk) step_tail(st2, st1, kb + 2);
    if (kb + 3 < nblk) step_tail(st0, st2, kb + 3);
    for (int i = 0; i < 2; i++)
        for (int j = 0; j < 2; j++)
            for (int r = 0; r < 4; r++) {
                const int row = row0 + wr * 32 + i * 16 + lk + 4 * r;
                const int col = col0 + wc * 32 + j * 16 + li;
                C[(long)row * ld + col] = acc[i][j][r];
            }
}

// FP32 form, software pipelined: the global loads of K block kb + 1 are
// issued into registers before block kb is multiplied out of LDS, and stored
// (rounded) into the other LDS buffer after it, so their latency overlaps the
// MFMAs and one barrier per block suffices (a buffer is rewritten only after
// the barrier that follows its last read).
template <int TRANSB>
__device__ __forceinline__ void gemm_tile_f32(int ld, const double* __restrict__ A, const double* __restrict__ B,
                                              double* __restrict__ C, int row0, int col0)
{
    using T = float;
    __shared__ T As[2][GEMM_GT][GEMM_KT + 1];
    __shared__ T Bs[2][GEMM_KT][GEMM_GT + 1];
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int wr = wave >> 1, wc = wave & 1;
    const int li = lane & 15, lk = lane >> 4;
    gtf_float4_t acc[2][2];
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
                for (int j = 0; j < 2; j++) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[i], b[j], acc[i][j], 0, 0, 0);
        }
        if (kb + 1 < nk) store(cur ^ 1);
        __syncthreads();
    }
    for (int i = 0; i < 2; i++)
        for (int j = 0; j < 2; j++)
            for (int r = 0; r < 4; r++) {
                const int row = row0 + wr * 32 + i * 16 + 4 * lk + r;
                const int col = col0 + wc * 32 + j * 16 + li;
                C[(long)row * ld + col] = (double)acc[i][j][r];
            }
}

template <int TRANSB, bool F32>
__device__ __forceinline__ void gemm_tile(int ld, const double* __restrict__ A, const double* __restrict__ B,
                                          double* __restrict__ C, int row0, int col0)
{
    if constexpr (F32) gemm_tile_f32<TRANSB>(ld, A, B, C, row0, col0);
    else gemm_tile_f64<TRANSB>(ld, A, B, C, row0, col0);
}
