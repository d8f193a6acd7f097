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
// Every S_d is a sum of integer products; over a K chunk of 256 its partial
// sums stay below 2^23 (|q_0 q_0| <= 2^14, |q_0 q_j| <= 2^13, |q_i q_j| <= 2^12,
// at most NS - 1 <= 5 pairs per d), so the MFMA's FP32 accumulation is exact.
// After each chunk the FP32 sums are added, scaled by 2^(-7 d), into an FP64
// accumulator.  The only error is the slicing residual and the dropped
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
#define OZ_KCHUNK 256  // K depth between FP32 -> FP64 flushes
#define OZ_LDS_ROW 40  // BF16 per LDS row (32 + 8 padding)

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

// 64 x 64 block of C from NS slices of A (rows) and of B^T (rows = columns of
// C), scaled back by the row exponents.
template <int NS>
__device__ __forceinline__ void oz_gemm_tile(int ld, const __bf16* __restrict__ QA, const int* __restrict__ EA,
                                             const __bf16* __restrict__ QB, const int* __restrict__ EB,
                                             double* __restrict__ C, int row0, int col0)
{
    __shared__ __align__(16) __bf16 As[NS][64][OZ_LDS_ROW];
    __shared__ __align__(16) __bf16 Bs[NS][64][OZ_LDS_ROW];
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int wr = wave >> 1, wc = wave & 1;
    const long sl = (long)ld * ld;
    oz_f32x16_t acc[NS];
    double acc64[16];
    for (int r = 0; r < 16; r++) acc64[r] = 0.0;
    for (int d = 0; d < NS; d++)
        for (int r = 0; r < 16; r++) acc[d][r] = 0.f;
    // global -> LDS: thread t copies 8 BF16 of row t / 4 at k offset 8 (t % 4)
    const int lr = tid >> 2, lk = (tid & 3) * 8;
    const __bf16* ga = QA + (long)(row0 + lr) * ld + lk;
    const __bf16* gb = QB + (long)(col0 + lr) * ld + lk;
    // the next K block's slices are loaded into registers while the current
    // one is multiplied out of LDS
    uint4 va[NS], vb[NS];
    auto load = [&](int k0) {
        for (int i = 0; i < NS; i++) {
            va[i] = *reinterpret_cast<const uint4*>(ga + i * sl + k0);
            vb[i] = *reinterpret_cast<const uint4*>(gb + i * sl + k0);
        }
    };
    auto store = [&]() {
        for (int i = 0; i < NS; i++) {
            *reinterpret_cast<uint4*>(&As[i][lr][lk]) = va[i];
            *reinterpret_cast<uint4*>(&Bs[i][lr][lk]) = vb[i];
        }
    };
    load(0);
    store();
    __syncthreads();
    int kc = 0;
    for (int k0 = 0; k0 < ld; k0 += OZ_KB) {
        if (k0 + OZ_KB < ld) load(k0 + OZ_KB);
        for (int kk = 0; kk < OZ_KB; kk += 16) {
            oz_bf16x8_t a[NS], b[NS];
            for (int i = 0; i < NS; i++) {
                a[i] = *reinterpret_cast<const oz_bf16x8_t*>(&As[i][wr * 32 + (lane & 31)][kk + 8 * (lane >> 5)]);
                b[i] = *reinterpret_cast<const oz_bf16x8_t*>(&Bs[i][wc * 32 + (lane & 31)][kk + 8 * (lane >> 5)]);
            }
            static_for<0, NS>([&](auto d_) {
                constexpr int d = decltype(d_)::value;
                static_for<0, d + 1>([&](auto i_) {
                    constexpr int i = decltype(i_)::value;
                    acc[d] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[i], b[d - i], acc[d], 0, 0, 0);
                });
            });
        }
        __syncthreads();  // every wave is done with this block
        if (k0 + OZ_KB < ld) {
            store();
            __syncthreads();
        }
        kc += OZ_KB;
        if (kc == OZ_KCHUNK || k0 + OZ_KB >= ld) {
            for (int r = 0; r < 16; r++) {
                double s = 0.0;
                for (int d = NS - 1; d >= 0; d--) s = s * (1.0 / 128.0) + (double)acc[d][r];
                acc64[r] += s;
            }
            for (int d = 0; d < NS; d++)
                for (int r = 0; r < 16; r++) acc[d][r] = 0.f;
            kc = 0;
        }
    }
    const int col = col0 + wc * 32 + (lane & 31);
    const int eb = EB[col];
    for (int r = 0; r < 16; r++) {
        const int row = row0 + wr * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
        C[(long)row * ld + col] = ldexp(acc64[r], EA[row] + eb - 14);
    }
}
