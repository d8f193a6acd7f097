// screen.h -- host launchers of screen.hip and linalg.hip.
#pragma once
#include "gtf_internal.h"

int gtf_build_pairs(hipStream_t st, const BasisView& b, int npairs, const int* s1, const int* s2,
                    const int* poff, const long* eoff, double4* prim, double* E, const long* zoff, double* Ez,
                    int* nprim_out, double prim_eps);
int gtf_count_rows(hipStream_t st, int nrows, const int* row_bra, const int* row_seg, const double* vbra,
                   const int* rank_bra, const double* vket_sorted, const int* rank_sorted, const int* seg_start,
                   const int* seg_len, double tol2, int same, long* exact, long* counts, long* row_off);
int gtf_symmetrize_packed(hipStream_t st, double* A, int nmol, const long* off, const int* n, int maxn);
int gtf_fill_qlist_diag(hipStream_t st, int n, const int* ids, int2* q);

// linalg.hip
int gtf_fill(hipStream_t st, double* x, long n, double v);
int gtf_fill_int(hipStream_t st, int* x, long n, int v);
int gtf_symmetrize(hipStream_t st, double* A, int n, int ld);
int gtf_add(hipStream_t st, const double* A, const double* B, double* C, long n);
// C = alpha * op(A) * op(B)^T-convention GEMM on row-major n x n matrices with leading dimension ld
// transB: 1 -> C = A * B^T (pdgemm3D_2), 0 -> C = A * B
int gtf_dgemm(hipStream_t st, int n, int ld, const double* A, const double* B, double* C, int transB);
// the same with FP32 MFMA on operands rounded to FP32 when fp32 != 0
int gtf_gemm(hipStream_t st, int n, int ld, const double* A, const double* B, double* C, int transB, int fp32);
