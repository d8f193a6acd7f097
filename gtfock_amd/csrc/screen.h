// screen.h -- host launchers of screen.hip and linalg.hip.
#pragma once
#include "gtf_internal.h"

// primitive pairs of every shell pair (p, P, Hermite blocks), the kept ones
// in descending order of their Schwarz bounds qout; at most BP_MAXRAW
// primitive pairs per shell pair (the caller checks)
#define GTF_MAX_RAW_PRIM_PAIRS 512
int gtf_build_pairs(hipStream_t st, const BasisView& b, int npairs, const int* s1, const int* s2,
                    const int* poff, const long* eoff, double4* prim, double* E, const long* zoff, double* Ez,
                    int* nprim_out, double* qout, double prim_eps);
// primitive quartets evaluated over a tile list (prim_census_kernel)
struct CensusArgs {
    const int4* tiles;
    int ntiles;
    const int* bra_list;
    const int* ket_list;
    const double* vbra;
    const double* vket;
    const int* rank_bra;
    const int* rank_ket;
    double tol2;
    int same;
    double tau2;  // precision split: FP64 tiles keep v_bra v_ket >= tau2, FP32 tiles the rest
    int low;
    int screened;  // the class applies primitive screening
    PairView pv;   // pv.peps: the threshold
    unsigned long long* out;
};
int gtf_prim_census(hipStream_t st, const CensusArgs& c);
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

// sigsort.hip: significant pairs of one type compacted and ordered on the
// device.  Scratch of n kept pairs (n from gtf_sig_count's total):
struct SigWork {
    int *ids_a, *ids_b, *pos_a, *pos_b, *rank;
    unsigned long long *k64_a, *k64_b;
    unsigned *k32_a, *k32_b;
    void* temp;
    size_t temp_bytes;
};
// outputs: the segment lists in value order (s_id, sorted, sortrank,
// sortnprim), seg_start / seg_len per (molecule, bucket) (zeroed by the
// caller), and the ket-list order (sig, sigv, rank, sigmol)
struct SigOut {
    int *s_id, *sortrank, *sortnprim, *seg_start, *seg_len, *sig, *rank, *sigmol;
    double *sorted, *sigv;
};
int gtf_sig_count(hipStream_t st, int npairs, const int* type, int t, const unsigned long long* vbits, const int* pmol,
                  const double* eta, int* block_count, int* block_off, int* total);
int gtf_sig_order(hipStream_t st, int npairs, const int* type, int t, const unsigned long long* vbits,
                  const int* pmol, const double* eta, const int* nprim, int nbucket, const int* block_off, int n,
                  SigWork& w, SigOut& o);
size_t gtf_sig_temp_bytes(int n);
