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
 *           most device bytes held at once} */
int gtf_pfock_stats(PFock_t pfock, double *out);

/* Densities the last PFock_computeFock digested from each integral in one
 * pass (pfock/fock_task.c:46 loops num_dmat inside update_F): num_dmat
 * when the per-density LDS work areas of every class fit, else 1 (one
 * pass per density). */
int gtf_pfock_fused_densities(PFock_t pfock);

/* Load balance of the last PFock_computeFock (the figures
 * PFock_getStatistics prints; pfock/pfock.c:1918-2010 gathers them across
 * processes): out[0] = slowest persistent workgroup over the mean, summed
 * over the tile launches; out[1] = tiles taken from a class counter after a
 * workgroup's first; out[2] = workgroups.  Returns 0, or -1. */
int gtf_pfock_balance(PFock_t pfock, double *out);

/* Wall seconds of PFock_create's set-up phases (the work of
 * pfock/pfock.c:801-960, screening at :922): out[0] basis upload, out[1]
 * primitive shell pairs, out[2] Schwarz diagonal (MN|MN), out[3]
 * significant-pair ordering, out[4] quartet-row counting, out[5] tile
 * cutting.  Returns 6, or -1. */
int gtf_pfock_setup_times(PFock_t pfock, double *out);

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
/* Primitive screening of the J/K builds (the role of ERD's internal
 * primitive screening under pfock/fock_task.c:197-207): primitive quartets
 * whose Schwarz bound sqrt((ij|ij)(kl|kl)) is below tol are skipped, in
 * the quartet classes of total angular momentum <= 3 (whose s and p shells
 * carry the long contractions); 0 evaluates every primitive quartet of a
 * screened quartet.  Default 1e-15.  With mixed precision the FP32 tiles
 * use max(tol, 1e-9 tau) (their quartets are below tau and carry FP32
 * errors of about 1e-7 tau), none with tol = 0.  Takes effect from the
 * next PFock_computeFock.  Returns 0, or -1 for tol < 0. */
int gtf_pfock_set_prim_tol(PFock_t pfock, double tol);
/* Per class (21): primitive quartets the J/K kernels evaluate at the
 * current threshold, in FP64 and in FP32 tiles (with tol = 0, FP64 + FP32
 * = the primitive quartets of gtf_pfock_class_work).  Returns 21, or -1. */
int gtf_pfock_prim_census(PFock_t pfock, double *npq_fp64, double *npq_fp32);

/* Per-class kernel timing with HIP events recorded on the engine's stream
 * around each class launch; enabling resets the accumulators. */
int gtf_pfock_set_timing(PFock_t pfock, int on);
/* 1: launch the class kernels one after another on the engine's stream
 * (each kernel alone on the GPU, for per-kernel measurement); 0 (default):
 * concurrently on four streams. */
int gtf_pfock_set_serial(PFock_t pfock, int on);
/* Accumulated milliseconds and launch counts per class since enabling. */
int gtf_pfock_class_times(PFock_t pfock, double *ms, long *launches);

/* Kernel shape of quartet class cls: info[11] = {La, Lb, Lc, Ld, bra
 * components per thread NC, threads per quartet, FP64 bra-row tiles, FP32
 * bra-row tiles, exchanged, ket components per thread NCT, boys_lds}; with
 * no tiles the class runs eri_kernel<..., EPI_JK>, else
 * eri_jk_tile_kernel<La, Lb, Lc, Ld, NC, double / float, boys_lds>, or
 * eri_jk_tile_t_kernel<La, Lb, Lc, Ld, NCT, double / float, boys_lds> when
 * exchanged = 1. */
int gtf_pfock_class_info(PFock_t pfock, int cls, int *info);

/* The engine's HIP stream (hipStream_t). */
void *gtf_pfock_stream(PFock_t pfock);

/* The most device memory the engine has held at once, in bytes. */
double gtf_pfock_device_bytes(PFock_t pfock);

/* Tile autotuner.  The tile launches of a quartet class have a shape
 * shape[GTF_TILE_SHAPE_FIELDS] = {units, unit_pq, tile_pq, grid, exchanged,
 * boys_lds, fp32}: at most `units` work units and about `tile_pq` primitive
 * quartets per tile, about `unit_pq` primitive quartets per work unit (how
 * finely a quartet's bra primitives are split), `grid` persistent
 * workgroups, exchanged = 1 for the bra/ket-exchanged kernel (only for
 * classes that have one), boys_lds = 1 for the kernels that stage the Boys
 * table in LDS, fp32 = 1 to evaluate the class's quartets below the mixed-
 * precision threshold tau in FP32 (0: the class stays FP64). */
#define GTF_TILE_SHAPE_FIELDS 7
/* gtf_pfock_autotune times the classes, each alone, over a coordinate
 * search of these (best of reps launches per candidate) with density 0 as
 * last put, then keeps the tuned shapes only if the whole build is faster
 * with them.  ms[3] (may be NULL) = {build ms with the shapes it started
 * from, build ms after tuning, candidates timed}.  Results do not depend on
 * the shapes beyond FP64 summation order.  Returns 0, or -1 on failure. */
int gtf_pfock_autotune(PFock_t pfock, int reps, double *ms);
int gtf_pfock_tile_shape(PFock_t pfock, int cls, int *shape);
int gtf_pfock_set_tile_shape(PFock_t pfock, int cls, const int *shape);
/* all 21 classes at once, shapes[21 * GTF_TILE_SHAPE_FIELDS] in class
 * order; nothing changes unless every shape is valid */
int gtf_pfock_set_tile_shapes(PFock_t pfock, const int *shapes);
/* 0 when every one of the 21 shapes is valid (positive sizes, 0/1 flags,
 * exchanged only where the class has that kernel), else -3 */
int gtf_tile_shapes_check(const int *shapes);

/* Integrals of nq shell quartets (4 shell ids each) in the reference layout
 * iM + dM*(iN + dN*(iP + dP*iQ)), concatenated in input order. */
int gtf_erd_quartets(ERD_t erd, int nq, const int *quartets, double *out);

/* The FP64 MFMA GEMM on host arrays (n x n, row-major): C = A * B when
 * transB = 0, C = A * B^T (the pdgemm3D_2 convention) when transB = 1. */
int gtf_dgemm_host(int n, const double *A, const double *B, double *C, int transB);
/* The same; fp32 != 0 rounds A and B to FP32 and uses the FP32 MFMA. */
int gtf_gemm_host(int n, const double *A, const double *B, double *C, int transB, int fp32);

/* The split-BF16 product on host arrays: every row of A and of B^T is cut
 * into ns (3..6) exact BF16 slices of 7 bits below its largest element, the
 * slice products run on v_mfma_f32_32x32x16_bf16 with exact FP32 sums over
 * 256-deep K chunks and are added in FP64; error about 2^(-7 ns) of the
 * row maxima. */
int gtf_gemm_split_host(int n, const double *A, const double *B, double *C, int transB, int ns);

/* gtf_gemm_split_host on caller-owned device matrices with no allocation.
 * No counterpart in the reference: for callers that cannot allocate (code
 * under HIP graph capture, serving loops, harnesses that forbid hipMalloc).
 * A, B, C: row-major ld x ld device arrays, zero beyond n (ld a multiple of
 * 64, at most 1536); stream: a hipStream_t (NULL: the default stream); the
 * call is asynchronous on it.  The scratch is one static device array per
 * GPU (55 MiB, held until the process ends), shared by every call on that
 * GPU: calls that overlap on one GPU race.  Returns 0, -3 (invalid or too
 * large) or -1 (HIP failure). */
int gtf_gemm_split_noalloc(int n, int ld, const double *A, const double *B, double *C, int transB, int ns,
                           void *stream);

/* Mixed-precision purification (compute_purification): mode 1 uses FP32
 * MFMA for the D^2 and D^3 products, mode 2 the split-BF16 product (FP64
 * residuals carried slice to slice), while ||D - D^2||_F >= until, and FP64
 * MFMA from then on, so the FP64 iterations restore idempotency at FP64
 * resolution; mode 0 is FP64 throughout (the default).  The final X D X
 * stays FP64.  Returns 0, -1 (invalid) or -2 (out of device memory). */
struct _purif_t;
int gtf_purif_set_precision(struct _purif_t *purif, int mode, double until);
/* slices per operand row of the split-BF16 products (3..6, default 5) */
int gtf_purif_set_slices(struct _purif_t *purif, int ns);
/* Mode 2 uses the split products only while the SCF's last DIIS residual
 * ||X (F D S - S D F) X||_F (compute_diis) is at least resid (default
 * 1e-3; before the SCF's first residual they are always used); closer to
 * convergence every product is FP64, so the SCF converges onto the FP64
 * fixed point.  0: split in every SCF iteration.  Returns 0 or -1. */
int gtf_purif_set_split_residual(struct _purif_t *purif, double resid);
/* purification iterations of the last compute_purification run in the
 * reduced precision */
int gtf_purif_fp32_iters(struct _purif_t *purif);

/* ---- batches: many molecules per Fock build and per SCF iteration ------
 *
 * A batch screens every molecule on its own (tolerance tolscr relative to
 * that molecule's largest Schwarz value, as pfock/screening.c:111) and
 * evaluates the screened quartets of all molecules in one launch per
 * quartet class, replayed as one HIP graph per build.  Matrices of the
 * batch are packed: molecule m is the row-major nbf[m] x nbf[m] block at
 * offset[m] (gtf_batch_layout); m < 0 in the put/get calls means the whole
 * packed array.  Return codes: 0 success, -1 HIP failure, -2 out of
 * device memory, -3 invalid argument. */
typedef struct GtfBatch *gtf_batch_t;

int gtf_batch_create(BasisSet_t *basis, int nmol, double tolscr, gtf_batch_t *batch);
void gtf_batch_destroy(gtf_batch_t batch);
/* nbf[nmol] and offset[nmol] (either may be NULL); returns nmol */
int gtf_batch_layout(gtf_batch_t batch, int *nbf, long *offset);
/* total doubles of a packed array */
long gtf_batch_total(gtf_batch_t batch);
/* FP32 ERIs for quartets with Schwarz bound below tau (0: FP64 only) */
int gtf_batch_set_precision(gtf_batch_t batch, double tau);
/* primitive screening threshold of the batch's builds, as
 * gtf_pfock_set_prim_tol */
int gtf_batch_set_prim_tol(gtf_batch_t batch, double tol);
/* mixed-precision purification inside gtf_batch_scf, per molecule, as
 * gtf_purif_set_precision (mode 0 FP64, 1 FP32, 2 split BF16) */
int gtf_batch_set_purif_precision(gtf_batch_t batch, int mode, double until);
/* slices of the split-BF16 products (3..6, default 5) */
int gtf_batch_set_purif_slices(gtf_batch_t batch, int ns);
/* per molecule, as gtf_purif_set_split_residual (default 1e-3) */
int gtf_batch_set_purif_residual(gtf_batch_t batch, double resid);
/* The tile autotuner on the batch's packed density, as gtf_pfock_autotune
 * (the next build re-captures its graph); shapes per class as
 * gtf_pfock_tile_shape / gtf_pfock_set_tile_shape, so the shapes tuned on
 * one batch can be applied to another. */
int gtf_batch_autotune(gtf_batch_t batch, int reps, double *ms);
int gtf_batch_tile_shape(gtf_batch_t batch, int cls, int *shape);
int gtf_batch_set_tile_shape(gtf_batch_t batch, int cls, const int *shape);
int gtf_batch_set_tile_shapes(gtf_batch_t batch, const int *shapes);
/* 1 (default): replay the build as a captured HIP graph; 0: eager launches */
int gtf_batch_use_graph(gtf_batch_t batch, int on);
int gtf_batch_put_density(gtf_batch_t batch, int m, const double *D);
/* G_m = 2 J(D_m) - K(D_m) for every molecule (the F of PFock_getMat) */
int gtf_batch_compute_fock(gtf_batch_t batch);
int gtf_batch_get_fock(gtf_batch_t batch, int m, double *G);
int gtf_batch_get_density(gtf_batch_t batch, int m, double *D);
/* borrowed device pointers of the packed D and G arrays */
int gtf_batch_device_ptrs(gtf_batch_t batch, double **D, double **G);
void *gtf_batch_stream(gtf_batch_t batch);
/* out[11] = {quartets, primitive quartets, FP32 quartets, FP32 primitive
 *            quartets, GPU ms of the last build, graph captured (0/1),
 *            most device bytes held at once (setup temporaries
 *            included), create seconds, SCF seconds in Fock builds,
 *            in DIIS, in purification}; returns 11 */
int gtf_batch_stats(gtf_batch_t batch, double *out);
/* Wall seconds of the batch's set-up phases: out[0..5] as
 * gtf_pfock_setup_times (gtf_batch_create), out[6] the SCF arrays and
 * one-electron matrices, out[7] X = S^-1/2 (the first SCF call or
 * gtf_batch_scf_begin).  Returns 8. */
int gtf_batch_setup_times(gtf_batch_t batch, double *out);
/* per class (21): quartets, primitive quartets, and their FP32 parts */
int gtf_batch_class_work(gtf_batch_t batch, long *nquartets, double *nprimquartets, long *nq_fp32,
                         double *npq_fp32);
/* RHF SCF of every molecule (pscf/scf.c:189-474 with a zero initial
 * density): at most max_iters Fock builds; molecule m stops after the build
 * whose electronic energy differs from the previous one by less than
 * ediff.  energy[m] = final total energy, iters[m] = builds used (may be
 * NULL), history[m * max_iters + it] = energy of build it (may be NULL). */
int gtf_batch_scf(gtf_batch_t batch, int max_iters, double ediff, double *energy, int *iters, double *history);
/* gtf_batch_scf that calls done(ctx, m, energy, builds) on the calling
 * thread as soon as molecule m's SCF ends (converged, diverged with builds
 * < 0, or out of iterations); done may be NULL */
typedef void (*gtf_scf_done_fn)(void *ctx, int m, double energy, int builds);
int gtf_batch_scf_cb(gtf_batch_t batch, int max_iters, double ediff, double *energy, int *iters, double *history,
                     gtf_scf_done_fn done, void *ctx);
/* The same SCF one iteration at a time, for serving loops that admit work
 * between iterations.  gtf_batch_scf_begin resets every molecule's SCF and
 * puts its initial density; each gtf_batch_scf_step runs one iteration of
 * the molecules still running (molecules that have ended drop out of the
 * Fock build) and calls done as each one ends; it returns how many are
 * still running (0: all have ended), or < 0 on failure.
 * gtf_batch_scf_results gives energy[m] (last build), iters[m] (builds of a
 * molecule that has ended, 0 while it runs) and history[m * max_iters + k];
 * any output may be NULL; returns nmol. */
int gtf_batch_scf_begin(gtf_batch_t batch, int max_iters, double ediff, gtf_scf_done_fn done, void *ctx);
int gtf_batch_scf_step(gtf_batch_t batch);
int gtf_batch_scf_results(gtf_batch_t batch, double *energy, int *iters, double *history);
/* Continuous batching: a new batch holding the molecules of `batch` whose
 * SCF is still running -- with their SCF state carried over exactly
 * (density, DIIS vectors and B matrix, iteration counts, energies) -- and
 * nadd new molecules that start their SCF in it (initial density, their own
 * iteration count 0), with batch's screening tolerance, precision, tile
 * shapes and callback.  origin[k] (may be NULL; capacity nrun + nadd) tells
 * where molecule k of the new batch came from: its index in `batch`, or
 * -(j + 1) for add[j].  The caller destroys `batch` afterwards, which
 * returns the memory of its finished molecules.  Returns 0, -1, -2 or -3
 * (no SCF running in batch, or nothing to carry or add). */
int gtf_batch_regroup(gtf_batch_t batch, BasisSet_t *add, int nadd, gtf_batch_t *out, int *origin);
/* wall-clock time (gettimeofday seconds) at which each molecule's last
 * gtf_batch_scf ended; returns nmol */
int gtf_batch_done_times(gtf_batch_t batch, double *t);
/* estimated device bytes of one molecule in a batch running SCF */
double gtf_batch_job_bytes(BasisSet_t basis);

/* ---- an SCF request server --------------------------------------------
 *
 * Jobs queue on submission.  A scheduler admits them in arrival order into
 * cohorts: the oldest queued job opens a cohort and takes the queued jobs
 * of its size class (basis functions within the same power-of-two band),
 * oldest first, up to max_jobs_per_cohort and the device-memory budget
 * (gtf_batch_job_bytes per job).  Each cohort runs as one batch
 * (gtf_batch_scf_step, one SCF iteration at a time) on its own thread and
 * streams, up to max_cohorts at once, of which at most half hold molecules
 * of 384 or more basis functions (the oldest job of a smaller class opens
 * the next cohort while those slots are taken).  Between iterations a
 * cohort admits queued jobs of its size class that fit the budget
 * (gtf_batch_regroup: running molecules carry their SCF state over, new
 * ones start, finished ones leave and return their memory); while a
 * small-class cohort has room, its class's queued jobs wait for it instead
 * of opening another cohort.  A job's result is published when its own SCF
 * ends.  The BasisSet of a job must stay alive until the job is done. */
typedef struct GtfServer *gtf_server_t;

int gtf_server_create(double hbm_budget_bytes, int max_cohorts, int max_jobs_per_cohort, double tolscr, double tau,
                      int max_iters, double ediff, gtf_server_t *server);
/* cohorts of molecules of 384 or more basis functions that may run at once
 * (default: half of max_cohorts, at least 1; at most max_cohorts).  Returns
 * 0 or -3. */
int gtf_server_set_large_slots(gtf_server_t server, int n);
/* tile shapes (21 x GTF_TILE_SHAPE_FIELDS, as gtf_batch_set_tile_shapes) for the cohorts that
 * start after the call, e.g. the autotuner's on a sample of the request
 * distribution; NULL: the built-in defaults.  Returns 0 or -3. */
int gtf_server_set_tile_shapes(gtf_server_t server, const int *shapes);
/* shapes for the cohorts (started after the call) whose largest molecule has
 * nbf_min..nbf_max basis functions, on top of the shapes set so far (the
 * last matching range wins; gtf_server_set_tile_shapes clears the ranges).
 * Returns 0 or -3. */
int gtf_server_set_tile_shapes_range(gtf_server_t server, int nbf_min, int nbf_max, const int *shapes);
/* purification products of the cohorts started after the call, as
 * gtf_batch_set_purif_precision (mode 0 FP64, 1 FP32, 2 split BF16) */
int gtf_server_set_purif_precision(gtf_server_t server, int mode, double until);
/* SCF iterations between two admissions into a running cohort (default 1;
 * a negative value: cohorts never admit jobs while they run) */
int gtf_server_set_join_gap(gtf_server_t server, int iters);
/* Small molecules first: while cohorts of molecules below 384 basis
 * functions run, a cohort of larger molecules waits before each SCF
 * iteration for at most factor times its own last iteration (default 2;
 * 0: never waits), so small jobs are not slowed by the large cohorts'
 * ERI kernels and large ones are never starved.  Returns 0, or -3 for
 * factor < 0. */
int gtf_server_set_large_yield(gtf_server_t server, double factor);
/* returns the job id (>= 0) or -3 */
long gtf_server_submit(gtf_server_t server, BasisSet_t basis);
/* status: 0 queued, 1 running, 2 done, < 0 failed (-2: larger than the
 * budget, -4: cancelled); out[7] = {energy, builds, submit time, start
 * time, done time, cohort, estimated bytes} (out may be NULL) */
int gtf_server_poll(gtf_server_t server, long job, double *out);
/* block until no job is queued or running */
int gtf_server_wait(gtf_server_t server);
/* out[8] = {jobs, queued, running cohorts, cohorts started, peak running
 *           cohorts, bytes in use, peak bytes, budget} */
int gtf_server_stats(gtf_server_t server, double *out);
/* jobs admitted into a running cohort, and regroups (each may be NULL) */
int gtf_server_join_stats(gtf_server_t server, long *joins, long *regroups);
/* cancels queued jobs, waits for running cohorts */
void gtf_server_destroy(gtf_server_t server);

#ifdef __cplusplus
}
#endif

#endif /* GTFOCK_AMD_H */
