// ozaki_launch.h -- launchers of ozaki.hip (split-BF16 MFMA products).
#pragma once
#include "gtf_internal.h"

// slice counts the product kernels are built for
#define OZ_SLICES_MIN 3
#define OZ_SLICES_MAX 6
// purification defaults: slices per operand, and the DIIS residual
// ||X (F D S - S D F) X||_F below which an SCF's purifications run FP64
// throughout (the split products' rounding would otherwise set the floor
// of the SCF's energy changes)
#define OZ_SLICES_DEFAULT 5
#define OZ_SPLIT_RESID 1e-3
// purification product modes (per molecule in a batch): FP64, FP32 MFMA,
// split BF16
#define GTF_LOW_FP64 0
#define GTF_LOW_FP32 1
#define GTF_LOW_SPLIT 2

// The matrices a launch covers: one (poff == nullptr, leading dimension
// maxld) or the padded blocks of a batch (poff/pld device arrays, maxld the
// largest pld); molecules with skip[m] != 0 or done[m] >= 0 are left alone.
struct OzOperands {
    int nmol = 1;
    int maxld = 0;
    const long* poff = nullptr;
    const int* pld = nullptr;
    const int* skip = nullptr;
    const int* done = nullptr;
    const int* sel = nullptr;
};

// Device scratch of one product: slices of both operands, their row
// exponents and the max partials.
struct OzScratch {
    __bf16* QA = nullptr;
    __bf16* QB = nullptr;
    int* EA = nullptr;
    int* EB = nullptr;
    double* prow = nullptr;
    double* pcol = nullptr;
    int ns = 0;
};

// Slices of X's rows (trans = 0) or columns (trans = 1) into Q / E; the
// max partials are recomputed unless maxima_ready (they depend on X only).
int gtf_oz_split(hipStream_t st, const OzOperands& o, const double* X, int trans, int ns, double* prow,
                 double* pcol, __bf16* Q, int* E, bool maxima_ready);
// C = (slices of A) * (slices of B^T)^T
int gtf_oz_gemm(hipStream_t st, const OzOperands& o, int ns, const __bf16* QA, const int* EA, const __bf16* QB,
                const int* EB, double* C);
// C = A * B (transB = 0) or A * B^T (transB = 1): split both, multiply
int gtf_oz_product(hipStream_t st, const OzOperands& o, OzScratch& s, int ns, const double* A, const double* B,
                   double* C, int transB);
// D2 = D * D and D3 = D * D2 with s.ns slices
int gtf_oz_square_cube(hipStream_t st, const OzOperands& o, OzScratch& s, const double* D, double* D2, double* D3);
// bytes of an OzScratch for padded_total elements (sum of pld^2) and ns slices
size_t gtf_oz_scratch_bytes(long padded_total, int ns);
// lay an OzScratch out in base (gtf_oz_scratch_bytes bytes)
int gtf_oz_scratch_carve(void* base, long padded_total, int ns, OzScratch& s);
