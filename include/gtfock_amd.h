/*
 * gtfock_amd.h -- entry points of libgtfock_hip.so beyond GTFock's API.
 *
 * None of these has a counterpart in the reference; they expose what the
 * opaque engine keeps on the device (statistics, the screening CSR, the
 * screened quartet lists, device buffers) so that tests and benchmarks can
 * check and time the hot path without going through host copies.
 */
#ifndef GTFOCK_AMD_H
#define GTFOCK_AMD_H

#include "CInt.h"
#include "pfock.h"

#ifdef __cplusplus
extern "C" {
#endif

/* out[8] = {usq, uitl, nnz, max Schwarz value, GPU seconds of the last
 *           PFock_computeFock, wall seconds of it, PFock_create seconds,
 *           device bytes held} */
int gtf_pfock_stats(PFock_t pfock, double *out);

/* The screening CSR in the layout of pfock/screening.c:152-175
 * (shellptr[nshells+1], shellid/shellrid/shellvalue[nnz]); returns nnz. */
int gtf_pfock_csr(PFock_t pfock, int *shellptr, int *shellid, int *shellrid, double *shellvalue);

/* Screened unique quartets per quartet class (21 classes, bra pair type >=
 * ket pair type; pair types ss, ps, pp, ds, dp, dd). */
int gtf_pfock_class_counts(PFock_t pfock, long *counts);

/* The screened unique quartets as (M,N,P,Q) shell ids in canonical pair
 * orientation; writes at most cap quartets, returns the total count. */
long gtf_pfock_quartets(PFock_t pfock, int *quartets, long cap);

/* Borrowed device pointers of D[index], J[index] (2J accumulator) and
 * K[index] (-K accumulator), row-major with leading dimension nbf. */
int gtf_pfock_device_ptrs(PFock_t pfock, int index, double **D, double **J, double **K);

/* Per quartet class: screened quartets and primitive quartets (sum of
 * nprim(bra) * nprim(ket)), the work units of the FLOP accounting. */
int gtf_pfock_class_work(PFock_t pfock, long *nquartets, double *nprimquartets);

/* Mixed precision: quartets whose Schwarz bound sqrt(v_MN v_PQ) is below
 * tau are evaluated in FP32 (FP64 J/K accumulation); tau = 0 restores the
 * plain FP64 build.  Takes effect from the next PFock_computeFock. */
int gtf_pfock_set_precision(PFock_t pfock, double tau);
/* Per class: quartets and primitive quartets evaluated in FP32. */
int gtf_pfock_mixed_counts(PFock_t pfock, long *nquartets_fp32, double *nprimquartets_fp32);

/* Per-class kernel timing with HIP events recorded on the engine's stream
 * around each class launch; enabling resets the accumulators. */
int gtf_pfock_set_timing(PFock_t pfock, int on);
/* Accumulated milliseconds and launch counts per class since enabling. */
int gtf_pfock_class_times(PFock_t pfock, double *ms, long *launches);

/* The engine's HIP stream (hipStream_t). */
void *gtf_pfock_stream(PFock_t pfock);

/* Device memory held by the engine, in bytes. */
double gtf_pfock_device_bytes(PFock_t pfock);

/* Integrals of nq shell quartets (4 shell ids each) in the reference layout
 * iM + dM*(iN + dN*(iP + dP*iQ)), concatenated in input order. */
int gtf_erd_quartets(ERD_t erd, int nq, const int *quartets, double *out);

/* The FP64 MFMA GEMM on host arrays (n x n, row-major): C = A * B when
 * transB = 0, C = A * B^T (the pdgemm3D_2 convention) when transB = 1. */
int gtf_dgemm_host(int n, const double *A, const double *B, double *C, int transB);

#ifdef __cplusplus
}
#endif

#endif /* GTFOCK_AMD_H */
