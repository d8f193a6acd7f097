// linalg.h -- purification / DIIS helpers of linalg.hip.
#pragma once
#include "gtf_internal.h"

#define MAX_DOTS 12

int gtf_gershgorin(hipStream_t st, const double* F, int n, int ld, double* dout3);
int gtf_purif_init(hipStream_t st, const double* F, double* D, int n, int ld, double a, double dshift);
int gtf_purif_stats(hipStream_t st, const double* D, const double* D2, const double* D3, int n, int ld,
                    double* partial, double* dout3);
int gtf_purif_update(hipStream_t st, double* D, const double* D2, const double* D3, long total, double c);
int gtf_antisym(hipStream_t st, const double* W, double* E, int ld);
int gtf_multidot(hipStream_t st, const double* x, const double* vecs, int m, long total, double* partial,
                 double* dout);
int gtf_lincomb(hipStream_t st, double* F, const double* vecs, const double* coef, int m, long total);
size_t gtf_partial_bytes();
