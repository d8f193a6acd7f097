// batch_linalg.h -- launchers of batch_linalg.hip (dense SCF steps over a
// batch of molecules).
#pragma once
#include <algorithm>

#include "gtf_internal.h"

// per-molecule DIIS dot products / coefficients (MAX_DIIS + 1 rounded up)
#define BMAX_DOTS 12
// workgroups per molecule of the reductions
#define BPART 64

// device arrays describing the two per-molecule layouts
struct BatchLayout {
    int nmol;
    int maxn, maxpld;
    const long* off;   // packed block offsets (n x n)
    const int* n;
    const long* poff;  // padded block offsets (pld x pld)
    const int* pld;
    double* part;      // reduction scratch, nmol * BPART * BMAX_DOTS doubles
    double* coef;      // 2 per molecule (purification start)
};

// skip[m] != 0 or done[m] >= 0 (either array may be null) leaves molecule m
// alone; fp32[m] (fp32 may be null) == GTF_LOW_FP32 multiplies its operands
// in FP32, == GTF_LOW_SPLIT leaves it to the split-BF16 products
int gtf_bgemm(hipStream_t st, const BatchLayout& L, const int* skip, const int* done, const double* A,
              const double* B, double* C, int transB, const int* fp32 = nullptr);
int gtf_bpad(hipStream_t st, const BatchLayout& L, const int* skip, const double* packed, double* padded);
int gtf_bunpad(hipStream_t st, const BatchLayout& L, const int* skip, const double* padded, double* packed);
int gtf_bpurif_init(hipStream_t st, const BatchLayout& L, const int* nocc, const int* skip, const double* F, double* D);
// also sets fp32[m] (may be null) to GTF_LOW_FP64 once ||D - D^2||_F < fp32_until
// and records in low_iters[m] (may be null) how many iterations ran in the
// reduced precision
int gtf_bpurif_stats(hipStream_t st, const BatchLayout& L, const int* skip, const double* D, const double* D2,
                     const double* D3, int* iters, double* stats, int it, double tol, int* fp32 = nullptr,
                     double fp32_until = 0.0, int* low_iters = nullptr);
int gtf_bpurif_update(hipStream_t st, const BatchLayout& L, const int* skip, const int* iters, const double* stats,
                      double* D, const double* D2, const double* D3, int it);
int gtf_bantisym(hipStream_t st, const BatchLayout& L, const int* skip, const double* W, double* E);
int gtf_bmultidot(hipStream_t st, const BatchLayout& L, const int* skip, const int* nvec, const double* x,
                  const double* vecs, long stride, double* dots);
int gtf_blincomb(hipStream_t st, const BatchLayout& L, const int* skip, const int* nvec, const double* coef,
                 const double* vecs, long stride, double* F);
int gtf_bfock_energy(hipStream_t st, const BatchLayout& L, const double* H, const double* G, const double* D,
                     double* F, double* energy);
