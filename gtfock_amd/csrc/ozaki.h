// ozaki.h -- C = A * B (or A * B^T) of FP64 matrices on the BF16 matrix
// cores, accurate to about 2^(-7 NS) of the operand rows, by exact slicing
// (the error-free splitting of Ozaki, Ogita, Oishi and Rump, Numer. Algor. 59,
// 95 (2012), with integer slices).
//
// Splitting.  Row r of an operand (a row of A, or a row of B^T = a column of
// B) with max |x| < 2^e is written as
//     x = 2^e sum_{i < NS} q_i 2^(-7 (i + 1))  +  residual,
//     |residual| <= 2^(e - 7 NS - 1),
// where q_0 is in [-128, 128] and q_i (i > 0) in [-64, 64]: integers, exact
// in BF16.  Each q_i is the nearest integer to the FP64 residual left by the
// slices before it, scaled by 2^7, so the split carries the FP64 residual of
// every slice into the next one.
//
// Products.  C[m][n] = 2^(eA_m + eB_n - 14) sum_d 2^(-7 d) S_d[m][n] with
//     S_d = sum_{i + j = d, i, j < NS} QA_i * QBt_j^T,   d < NS.
// Every S_d is a sum of integer products; over a K chunk of at most 256 its
// partial sums stay below 2^23 (|q_0 q_0| <= 2^14, |q_0 q_j| <= 2^13,
// |q_i q_j| <= 2^12, at most NS - 1 <= 5 pairs per d), so the MFMA's FP32
// accumulation is exact.  After each chunk (96 of K here) the FP32 sums are
// added, scaled by 2^(-7 d), into an FP64 accumulator.  The only error is the slicing residual and the dropped
// pairs with i + j >= NS: about 2^(-7 NS) relative to the row maxima.
//
// Tile: a 256-thread workgroup computes a 64 x 64 block of C; wave w owns the
// 32 x 32 quadrant (w / 2, w % 2) with one v_mfma_f32_32x32x16_bf16 per
// slice pair and 16-deep K step.  Operand lanes: A[i][k] at lane i + 32 (k / 8),
// element k % 8; B^T[j][k] the same; result register r of lane l is
// C[8 (r / 4) + 4 (l / 32) + r % 4][l % 32].  The K permutation inside a
// 16-deep step is the same for A and B, so it does not change the sum.
//
// Layouts: operand slices are row-major ld x ld BF16 (slice i at i * ld * ld);
// exponents one int per slice row; max partials one double per (64-block,
// row) in either direction.  ld is a multiple of 64 with zero padding.
#pragma once
#include "gtf_internal.h"

typedef __bf16 oz_bf16x8_t __attribute__((ext_vector_type(8)));
typedef float oz_f32x16_t __attribute__((ext_vector_type(16)));

#define OZ_MAX_SLICES 6
#define OZ_KB 32       // K depth staged in LDS per step
#define OZ_KCHUNK 256  // most K depth between FP32 -> FP64 flushes

// One molecule's operand: base pointer and leading dimension.
struct OzView {
    long off;  // element offset of the padded block
    int ld;
};

__device__ __forceinline__ OzView oz_view(int m, const long* __restrict__ poff, const int* __restrict__ pld,
                                          int ld0)
{
    OzView v;
    if (poff) {
        v.off = poff[m];
        v.ld = pld[m];
    } else {
        v.off = 0;
        v.ld = ld0;
    }
    return v;
}

// Partial maxima of |X| over 64 x 64 blocks: prow[cb * ld + r] = max over the
// block's columns (column block cb) of row r, pcol[rb * ld + c] = max over
// the block's rows (row block rb) of column c.  One workgroup per block.
__device__ __forceinline__ void oz_absmax_tile(int ld, const double* __restrict__ X, double* __restrict__ prow,
                                               double* __restrict__ pcol, int rb, int cb)
{
    __shared__ double t[64][65];
    const int tid = threadIdx.x;
    for (int s = 0; s < 16; s++) {
        const int r = s * 4 + (tid >> 6), c = tid & 63;
        t[r][c] = fabs(X[(long)(rb * 64 + r) * ld + cb * 64 + c]);
    }
    __syncthreads();
    if (tid < 64) {
        double mr = 0.0, mc = 0.0;
        for (int k = 0; k < 64; k++) {
            mr = fmax(mr, t[tid][k]);
            mc = fmax(mc, t[k][tid]);
        }
        prow[(long)cb * ld + rb * 64 + tid] = mr;
        pcol[(long)rb * ld + cb * 64 + tid] = mc;
    }
}

// exponent e with max |x| < 2^e (0 for an all-zero row)
__device__ __forceinline__ int oz_exponent(const double* __restrict__ part, int ld, int r)
{
    double mx = 0.0;
    for (int b = 0; b < ld / 64; b++) mx = fmax(mx, part[(long)b * ld + r]);
    int e = 0;
    if (mx > 0.0) (void)frexp(mx, &e);
    return e;
}

// Slices of one 64 x 64 block of the operand rows: TRANS = 0 slices rows of X
// (the A operand, or B^T when B is given transposed), TRANS = 1 slices
// columns of X (B^T of a B given as is).  part is prow (TRANS = 0) or pcol
// (TRANS = 1).  Q holds ns slices of ld x ld; E[ld] the row exponents.
template <int TRANS>
__device__ __forceinline__ void oz_split_tile(int ld, const double* __restrict__ X, const double* __restrict__ part,
                                              int ns, __bf16* __restrict__ Q, int* __restrict__ E, int rb, int cb)
{
    __shared__ double t[64][65];
    __shared__ int es[64];
    const int tid = threadIdx.x;
    // slice rows of this tile: rb (TRANS = 0) or cb (TRANS = 1) block
    const int sr0 = TRANS ? cb * 64 : rb * 64, sc0 = TRANS ? rb * 64 : cb * 64;
    if (tid < 64) {
        const int e = oz_exponent(part, ld, sr0 + tid);
        es[tid] = e;
        if ((TRANS ? rb : cb) == 0) E[sr0 + tid] = e;
    }
    for (int s = 0; s < 16; s++) {
        const int r = s * 4 + (tid >> 6), c = tid & 63;
        const double x = X[(long)(rb * 64 + r) * ld + cb * 64 + c];
        if (TRANS) t[c][r] = x;
        else t[r][c] = x;
    }
    __syncthreads();
    const long sl = (long)ld * ld;
    for (int s = 0; s < 16; s++) {
        const int r = s * 4 + (tid >> 6), c = tid & 63;  // slice row r, column c of the tile
        double v = ldexp(t[r][c], -es[r]);                 // |v| < 1
        __bf16* q = Q + (long)(sr0 + r) * ld + sc0 + c;
        for (int i = 0; i < ns; i++) {
            const double w = v * 128.0;
            const double qi = rint(w);
            v = w - qi;  // exact
            q[i * sl] = (__bf16)(float)qi;
        }
    }
}

typedef __attribute__((address_space(3))) void oz_lds_void;

// One K block of a 64 x 64 tile in LDS: NS slices of 64 operand rows x
// OZ_KB, 64 B per row, the four 16-B chunks of row r stored at chunk
// position (chunk ^ (r & 3)) so that the fragment reads of 32 rows spread
// over the banks.
template <int NS>
struct OzStage {
    __bf16 a[NS][64][OZ_KB];
    __bf16 b[NS][64][OZ_KB];
};

// 64 x 64 block of C from NS slices of A (rows) and of B^T (rows = columns of
// C), scaled back by the row exponents.  The K blocks stream through three
// LDS stages filled by direct global -> LDS loads (global_load_lds_dwordx4),
// issued two blocks ahead of the block being multiplied, so that the load
// latency of the slices (read once per tile, mostly from the last-level
// cache) hides behind MFMA work with one wave per SIMD and no registers
// spent on staging.
template <int NS>
__device__ __forceinline__ void oz_gemm_tile(int ld, const __bf16* __restrict__ QA, const int* __restrict__ EA,
                                             const __bf16* __restrict__ QB, const int* __restrict__ EB,
                                             double* __restrict__ C, int row0, int col0)
{
    __shared__ __align__(16) OzStage<NS> st0, st1, st2;
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int wr = wave >> 1, wc = wave & 1;
    const long sl = (long)ld * ld;
    const int nblk = ld / OZ_KB;
    oz_f32x16_t acc[NS];
    double acc64[16];
    for (int r = 0; r < 16; r++) acc64[r] = 0.0;
    for (int d = 0; d < NS; d++)
        for (int r = 0; r < 16; r++) acc[d][r] = 0.f;
    // loads: wave w fills rows 16 w .. 16 w + 15 of every slice of both
    // operands; lane l writes LDS chunk position l % 4 of row 16 w + l / 4
    const int lrow = wave * 16 + (lane >> 2);
    const int lchunk = (lane & 3) ^ (lrow & 3);
    const __bf16* ga = QA + (long)(row0 + lrow) * ld + 8 * lchunk;
    const __bf16* gb = QB + (long)(col0 + lrow) * ld + 8 * lchunk;
    auto issue = [&](OzStage<NS>& S, int k0) {
        for (int i = 0; i < NS; i++) {
            __builtin_amdgcn_global_load_lds(ga + i * sl + k0, (oz_lds_void*)&S.a[i][wave * 16][0], 16, 0, 0);
            __builtin_amdgcn_global_load_lds(gb + i * sl + k0, (oz_lds_void*)&S.b[i][wave * 16][0], 16, 0, 0);
        }
    };
    // fragment rows of this wave's 32 x 32 quadrant
    const int ra = wr * 32 + (lane & 31), rb = wc * 32 + (lane & 31);
    auto compute = [&](const OzStage<NS>& S) {
        for (int kk = 0; kk < OZ_KB; kk += 16) {
            const int c = (kk >> 3) + (lane >> 5);  // logical 16-B chunk of the k range
            oz_bf16x8_t a[NS], b[NS];
            for (int i = 0; i < NS; i++) {
                a[i] = *reinterpret_cast<const oz_bf16x8_t*>(&S.a[i][ra][8 * (c ^ (ra & 3))]);
                b[i] = *reinterpret_cast<const oz_bf16x8_t*>(&S.b[i][rb][8 * (c ^ (rb & 3))]);
            }
            // slice pair (i, j) adds into S_(i+j); i outermost, so that
            // consecutive MFMAs update different accumulators; the FP32 sums
            // are exact, so the order does not change them
            static_for<0, NS>([&](auto i_) {
                constexpr int i = decltype(i_)::value;
                static_for<0, NS - i>([&](auto j_) {
                    constexpr int j = decltype(j_)::value;
                    acc[i + j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[i], b[j], acc[i + j], 0, 0, 0);
                });
            });
        }
    };
    auto flush = [&]() {
        for (int r = 0; r < 16; r++) {
            double s = 0.0;
            for (int d = NS - 1; d >= 0; d--) s = s * (1.0 / 128.0) + (double)acc[d][r];
            acc64[r] += s;
        }
        for (int d = 0; d < NS; d++)
            for (int r = 0; r < 16; r++) acc[d][r] = 0.f;
    };
    // block kb from stage cur; block kb + 2 loads into stage nxt (the one
    // block kb - 1 used).  __syncthreads: every wave's loads have landed (it
    // waits for all of them, block kb + 1's included) and every wave's LDS
    // reads of the previous step are done before any wave refills that
    // step's stage; a bare s_barrier would let the compiler move LDS reads
    // across it.
    auto step = [&](const OzStage<NS>& cur, OzStage<NS>& nxt, int kb) {
        __syncthreads();
        issue(nxt, (kb + 2) * OZ_KB);
        compute(cur);
    };
    // the last blocks, with fewer (or no) loads ahead
    auto step_tail = [&](const OzStage<NS>& cur, OzStage<NS>& nxt, int kb) {
        __syncthreads();
        if (kb + 2 < nblk) issue(nxt, (kb + 2) * OZ_KB);
        compute(cur);
    };
    issue(st0, 0);
    if (nblk > 1) issue(st1, OZ_KB);
    // three blocks (96 of K) per pass, the stages in a fixed rotation, the
    // FP32 sums flushed after each pass (exact: below OZ_KCHUNK); the first
    // pass is peeled and at most four blocks are left for the tail
    int kb = 0;
    if (nblk >= 5) {
        step(st0, st2, 0);
        step(st1, st0, 1);
        step(st2, st1, 2);
        flush();
        for (kb = 3; kb + 5 <= nblk; kb += 3) {
            step(st0, st2, kb);
            step(st1, st0, kb + 1);
            step(st2, st1, kb + 2);
            flush();
        }
    }
    if (kb < nblk) step_tail(st0, st2, kb);
    if (kb + 1 < nblk) step_tail(st1, st0, kb + 1);
    if (kb + 2 < nblk) step_tail(st2, st1, kb + 2);
    if (kb + 3 < nblk) step_tail(st0, st2, kb + 3);
    flush();
    const int col = col0 + wc * 32 + (lane & 31);
    const int eb = EB[col];
    for (int r = 0; r < 16; r++) {
        const int row = row0 + wr * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
        C[(long)row * ld + col] = ldexp(acc64[r], EA[row] + eb - 14);
    }
}
