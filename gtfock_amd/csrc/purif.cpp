// purif.cpp -- create_purif / compute_purification / compute_diis
// (pscf/purif.c:22-641) with every matrix product on the GPU.
//
// The reference distributes n x n matrices over an np2^3 process cube and
// multiplies with a 3-D DGEMM (pscf/pdgemm.c); here the matrices live in HBM,
// padded to a multiple of 64 with zeros, and each product is one FP64 MFMA
// GEMM launch.  pdgemm3D_2's convention C = A * B^T is kept for the DIIS and
// X D X products.
#include <sys/time.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <new>

#include "../../include/purif.h"
#include "engine.h"
#include "hostcopy.h"
#include "batch_linalg.h"
#include "linalg.h"
#include "ozaki_launch.h"
#include "screen.h"

struct purif_dev {
    hipStream_t stream = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    PinnedStage stage;  // F, D in and out
    int n = 0, ld = 0;
    double *F = nullptr, *D = nullptr, *D2 = nullptr, *D3 = nullptr, *X = nullptr, *S = nullptr;
    double *W = nullptr, *W2 = nullptr, *evecs = nullptr, *Fvecs = nullptr;
    double *scal = nullptr, *partial = nullptr, *coef = nullptr;
    bool xs_loaded = false;
    // Reduced-precision purification products while ||D - D^2||_F exceeds
    // low_until, FP64 from there to the end: mode 1 = FP32 MFMA on rounded
    // operands, mode 2 = split-BF16 MFMA with ns exact slices (ozaki.h)
    int mode = 0;
    double low_until = 1e-3;
    int ns = OZ_SLICES_DEFAULT;
    int low_iters = 0;
    // mode 2 runs the split products only while the last DIIS residual
    // ||X (F D S - S D F) X||_F exceeds split_resid (before the SCF's first
    // residual: split); closer to convergence the purification is FP64
    double split_resid = OZ_SPLIT_RESID;
    double last_resid = -1.0;
    void* oz_mem = nullptr;
    OzScratch oz;
    // the purification loop runs on the device through the batch kernels
    // with a one-molecule layout: the trace-correcting coefficient, the
    // stop test and the precision switch are device state, so the host
    // checks for convergence only every PURIF_CHECK iterations
    long *d_off = nullptr, *d_poff = nullptr;
    int *d_n = nullptr, *d_pld = nullptr, *d_nocc = nullptr, *d_iters = nullptr, *d_low = nullptr,
        *d_lowit = nullptr;
    double *d_part = nullptr, *d_pcoef = nullptr, *d_stats = nullptr;
    BatchLayout layout() const { return BatchLayout{1, n, ld, d_off, d_n, d_poff, d_pld, d_part, d_pcoef}; }
};

#define PURIF_CHECK 4

static double wtime()
{
    struct timeval tv;
    gettimeofday(&tv, nullptr);
    return tv.tv_sec + 1e-6 * tv.tv_usec;
}

#define PD_CHECK(call)                                                                        \
    do {                                                                                      \
        hipError_t e_ = (call);                                                               \
        if (e_ != hipSuccess) {                                                               \
            fprintf(stderr, "purif: HIP error %s at %s:%d\n", hipGetErrorString(e_), __FILE__, __LINE__); \
            return -1;                                                                        \
        }                                                                                     \
    } while (0)

static int h2d(purif_dev* d, double* dst, const double* src, int ldx)
{
    PD_CHECK(stage_h2d(d->stage, d->stream, dst, d->ld, src, ldx, d->n, d->n));
    return 0;
}

static int d2h(purif_dev* d, double* dst, const double* src, int ldx)
{
    PD_CHECK(stage_d2h(d->stage, d->stream, dst, ldx, src, d->ld, d->n, d->n));
    return 0;
}

static int load_xs(purif_t* p)
{
    purif_dev* d = p->dev;
    if (d->xs_loaded) return 0;
    if (h2d(d, d->X, p->X_block, p->ldx) || h2d(d, d->S, p->S_block, p->ldx)) return -1;
    d->xs_loaded = true;
    return 0;
}

// two GEMMs timed with events; adds the kernel time to *acc
static int gemm_pair(purif_dev* d, const double* A1, const double* B1, double* C1, int t1, const double* A2,
                     const double* B2, double* C2, int t2, double* acc, int fp32 = 0)
{
    PD_CHECK(hipEventRecord(d->e0, d->stream));
    if (gtf_gemm(d->stream, d->n, d->ld, A1, B1, C1, t1, fp32)) return -1;
    if (A2 && gtf_gemm(d->stream, d->n, d->ld, A2, B2, C2, t2, fp32)) return -1;
    PD_CHECK(hipEventRecord(d->e1, d->stream));
    PD_CHECK(hipEventSynchronize(d->e1));
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, d->e0, d->e1);
    if (acc) *acc += 1e-3 * ms;
    return 0;
}

extern "C" {

purif_t* create_purif(BasisSet_t basis, int nprow_purif, int npcol_purif, int npgrd_purif)
{
    if (!basis || nprow_purif != 1 || npcol_purif != 1 || npgrd_purif != 1) {
        fprintf(stderr, "create_purif: one GPU runs the whole purification; the process grid must be 1x1x1\n");
        return nullptr;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
        fprintf(stderr, "create_purif: no HIP device is visible\n");
        return nullptr;
    }
    purif_t* p = (purif_t*)calloc(1, sizeof(purif_t));
    purif_dev* d = new (std::nothrow) purif_dev();
    if (!p || !d) { free(p); delete d; return nullptr; }
    p->dev = d;
    const int n = CInt_getNumFuncs(basis);
    p->nbf = n;
    p->nobtls = CInt_getNumOccOrb(basis);
    p->np_purif = p->nprow_purif = p->npcol_purif = p->npgrd_purif = 1;
    p->runpurif = p->rundgemm = 1;
    p->srow_purif = p->scol_purif = 0;
    p->nrows_purif = p->ncols_purif = n;
    p->ldx = n;
    p->meshsize = n * n;
    double** blocks[5] = {&p->H_block, &p->X_block, &p->S_block, &p->F_block, &p->D_block};
    for (double** b : blocks) {
        *b = (double*)aligned_alloc(64, ((sizeof(double) * n * n + 63) / 64) * 64);
        if (!*b) { destroy_purif(p); return nullptr; }
        memset(*b, 0, sizeof(double) * n * n);
    }
    p->len_diis = 0;
    p->bmax = DBL_MIN;
    p->bmax_id = -1;
    for (int i = 0; i <= MAX_DIIS; i++)
        for (int j = 0; j <= MAX_DIIS; j++) p->b_mat[i * (MAX_DIIS + 1) + j] = (i != j) ? -1.0 : 0.0;
    d->n = n;
    d->ld = ((n + 63) / 64) * 64;
    const size_t m = (size_t)d->ld * d->ld;
    bool ok = hipStreamCreate(&d->stream) == hipSuccess && hipEventCreate(&d->e0) == hipSuccess &&
              hipEventCreate(&d->e1) == hipSuccess;
    double** bufs[] = {&d->F, &d->D, &d->D2, &d->D3, &d->X, &d->S, &d->W, &d->W2};
    for (double** b : bufs) ok = ok && hipMalloc((void**)b, sizeof(double) * m) == hipSuccess;
    ok = ok && hipMalloc((void**)&d->evecs, sizeof(double) * m * MAX_DIIS) == hipSuccess;
    ok = ok && hipMalloc((void**)&d->Fvecs, sizeof(double) * m * MAX_DIIS) == hipSuccess;
    ok = ok && hipMalloc((void**)&d->scal, sizeof(double) * 64) == hipSuccess;
    ok = ok && hipMalloc((void**)&d->coef, sizeof(double) * 64) == hipSuccess;
    ok = ok && hipMalloc((void**)&d->partial, gtf_partial_bytes()) == hipSuccess;
    ok = ok && hipMalloc((void**)&d->d_off, sizeof(long) * 2) == hipSuccess;
    ok = ok && hipMalloc((void**)&d->d_poff, sizeof(long) * 2) == hipSuccess;
    int** ints[] = {&d->d_n, &d->d_pld, &d->d_nocc, &d->d_iters, &d->d_low, &d->d_lowit};
    for (int** q : ints) ok = ok && hipMalloc((void**)q, sizeof(int) * 2) == hipSuccess;
    ok = ok && hipMalloc((void**)&d->d_part, sizeof(double) * BPART * BMAX_DOTS) == hipSuccess;
    ok = ok && hipMalloc((void**)&d->d_pcoef, sizeof(double) * 2) == hipSuccess;
    ok = ok && hipMalloc((void**)&d->d_stats, sizeof(double) * 3) == hipSuccess;
    if (!ok) { destroy_purif(p); return nullptr; }
    {
        const long zero = 0;
        const int nn = n, ldv = d->ld, no = p->nobtls;
        ok = hipMemcpy(d->d_off, &zero, sizeof zero, hipMemcpyHostToDevice) == hipSuccess &&
             hipMemcpy(d->d_poff, &zero, sizeof zero, hipMemcpyHostToDevice) == hipSuccess &&
             hipMemcpy(d->d_n, &nn, sizeof nn, hipMemcpyHostToDevice) == hipSuccess &&
             hipMemcpy(d->d_pld, &ldv, sizeof ldv, hipMemcpyHostToDevice) == hipSuccess &&
             hipMemcpy(d->d_nocc, &no, sizeof no, hipMemcpyHostToDevice) == hipSuccess;
        if (!ok) { destroy_purif(p); return nullptr; }
    }
    for (double** b : bufs) ok = ok && gtf_fill(d->stream, *b, (long)m, 0.0) == 0;
    ok = ok && gtf_fill(d->stream, d->evecs, (long)m * MAX_DIIS, 0.0) == 0;
    ok = ok && gtf_fill(d->stream, d->Fvecs, (long)m * MAX_DIIS, 0.0) == 0;
    ok = ok && hipStreamSynchronize(d->stream) == hipSuccess;
    if (!ok) { destroy_purif(p); return nullptr; }
    return p;
}

void destroy_purif(purif_t* p)
{
    if (!p) return;
    purif_dev* d = p->dev;
    if (d) {
        if (d->stream) (void)hipStreamSynchronize(d->stream);
        void* bufs[] = {d->F, d->D, d->D2, d->D3, d->X, d->S, d->W, d->W2, d->evecs, d->Fvecs,
                        d->scal, d->coef, d->partial, d->oz_mem, d->d_off, d->d_poff, d->d_n, d->d_pld,
                        d->d_nocc, d->d_iters, d->d_low, d->d_lowit, d->d_part, d->d_pcoef, d->d_stats};
        for (void* b : bufs)
            if (b) (void)hipFree(b);
        if (d->e0) (void)hipEventDestroy(d->e0);
        if (d->e1) (void)hipEventDestroy(d->e1);
        if (d->stream) (void)hipStreamDestroy(d->stream);
        delete d;
    }
    free(p->H_block);
    free(p->X_block);
    free(p->S_block);
    free(p->F_block);
    free(p->D_block);
    free(p);
}

// Palser-Manolopoulos purification (pscf/purif.c:257-488) of F (orthogonal
// basis), then D_AO = X D X.  Each iteration: D2 = D D, D3 = D D2 (FP64
// MFMA, or in the reduced precision until ||D - D^2||_F < low_until), the
// statistics ||D - D2||_F, tr(D2 - D3), tr(D - D2) and the update with
// c = tr(D2 - D3) / tr(D - D2), all on the device; the molecule is marked
// done at the iteration whose error is below 1e-11 and later launches skip
// it, so the host only checks every PURIF_CHECK iterations.
static int purify(purif_t* p, double* F_block, double* D_block, int* iters)
{
    purif_dev* d = p->dev;
    if (load_xs(p) || h2d(d, d->F, F_block, p->ldx)) return -1;
    const BatchLayout L = d->layout();
    if (d->mode == GTF_LOW_SPLIT && !d->oz_mem) return -1;
    const int mode =
        d->mode == GTF_LOW_SPLIT && d->last_resid >= 0.0 && d->last_resid < d->split_resid ? GTF_LOW_FP64 : d->mode;
    PD_CHECK(hipEventRecord(d->e0, d->stream));
    if (gtf_bpurif_init(d->stream, L, d->d_nocc, nullptr, d->F, d->D)) return -1;
    if (gtf_fill_int(d->stream, d->d_iters, 1, -1) || gtf_fill_int(d->stream, d->d_low, 1, mode) ||
        gtf_fill_int(d->stream, d->d_lowit, 1, 0))
        return -1;
    OzOperands o;
    o.maxld = d->ld;
    o.done = d->d_iters;
    o.sel = d->d_low;
    int it_done = -1, it = 0;
    for (it = 0; it < 200; it++) {
        if (gtf_bgemm(d->stream, L, nullptr, d->d_iters, d->D, d->D, d->D2, 0, d->d_low) ||
            gtf_bgemm(d->stream, L, nullptr, d->d_iters, d->D, d->D2, d->D3, 0, d->d_low))
            return -1;
        if (mode == GTF_LOW_SPLIT && gtf_oz_square_cube(d->stream, o, d->oz, d->D, d->D2, d->D3)) return -1;
        if (gtf_bpurif_stats(d->stream, L, nullptr, d->D, d->D2, d->D3, d->d_iters, d->d_stats, it, 1e-11, d->d_low,
                             d->low_until, d->d_lowit))
            return -1;
        if (gtf_bpurif_update(d->stream, L, nullptr, d->d_iters, d->d_stats, d->D, d->D2, d->D3, it)) return -1;
        if (it % PURIF_CHECK == PURIF_CHECK - 1) {
            PD_CHECK(hipMemcpyAsync(&it_done, d->d_iters, sizeof(int), hipMemcpyDeviceToHost, d->stream));
            PD_CHECK(hipStreamSynchronize(d->stream));
            if (it_done >= 0) break;
        }
    }
    int lowit = 0;
    PD_CHECK(hipMemcpyAsync(&it_done, d->d_iters, sizeof(int), hipMemcpyDeviceToHost, d->stream));
    PD_CHECK(hipMemcpyAsync(&lowit, d->d_lowit, sizeof(int), hipMemcpyDeviceToHost, d->stream));
    PD_CHECK(hipEventRecord(d->e1, d->stream));
    PD_CHECK(hipEventSynchronize(d->e1));
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, d->e0, d->e1);
    p->timedgemm += 1e-3 * ms;
    d->low_iters = mode != 0 ? (lowit > 0 ? lowit : (it_done >= 0 ? it_done + 1 : 200)) : 0;
    // D = X D X (pscf/purif.c:469-476)
    if (gemm_pair(d, d->X, d->D, d->W, 1, d->W, d->X, d->D, 1, &p->timedgemm)) return -1;
    if (d2h(d, D_block, d->D, p->ldx)) return -1;
    *iters = it_done >= 0 ? it_done : 200;
    return 0;
}

int compute_purification(purif_t* p, double* F_block, double* D_block)
{
    p->timedgemm = p->timepdgemm = p->timepass = p->timetr = 0.0;
    const double t0 = wtime();
    int it = 0;
    if (purify(p, F_block, D_block, &it)) {
        fprintf(stderr, "compute_purification: GPU execution failed\n");
        return -1;
    }
    p->timepass = wtime() - t0;
    p->timepdgemm = p->timedgemm;
    return it;
}

// Gauss-Jordan inverse with partial pivoting (the dgetrf + dgetri of
// pscf/purif.c:616-619) for the (len+1)^2 DIIS system
static void small_inverse(int m, const double* A, int lda, double* inv)
{
    double a[(MAX_DIIS + 1) * (MAX_DIIS + 1)];
    for (int i = 0; i < m; i++)
        for (int j = 0; j < m; j++) {
            a[i * m + j] = A[i * lda + j];
            inv[i * m + j] = (i == j);
        }
    for (int col = 0; col < m; col++) {
        int piv = col;
        for (int r = col + 1; r < m; r++)
            if (fabs(a[r * m + col]) > fabs(a[piv * m + col])) piv = r;
        if (piv != col)
            for (int k = 0; k < m; k++) {
                std::swap(a[col * m + k], a[piv * m + k]);
                std::swap(inv[col * m + k], inv[piv * m + k]);
            }
        const double dg = a[col * m + col];
        for (int k = 0; k < m; k++) {
            a[col * m + k] /= dg;
            inv[col * m + k] /= dg;
        }
        for (int r = 0; r < m; r++) {
            if (r == col) continue;
            const double f = a[r * m + col];
            if (f == 0.0) continue;
            for (int k = 0; k < m; k++) {
                a[r * m + k] -= f * a[col * m + k];
                inv[r * m + k] -= f * inv[col * m + k];
            }
        }
    }
}

static int diis(purif_t* p, double* D_block, double* F_block, int iter)
{
    purif_dev* d = p->dev;
    const int ld = d->ld;
    const long m = (long)ld * ld;
    const int LB = MAX_DIIS + 1;
    if (load_xs(p) || h2d(d, d->F, F_block, p->ldx) || h2d(d, d->D, D_block, p->ldx)) return -1;
    double* curF = d->F;
    if (iter == 0) d->last_resid = -1.0;  // a new SCF
    if (iter > 1) {
        const int cur = p->len_diis < MAX_DIIS ? p->len_diis++ : p->bmax_id;
        if (cur < 0 || cur >= MAX_DIIS) {
            fprintf(stderr, "compute_diis: no DIIS slot (non-finite error vectors)\n");
            return -1;
        }
        double* e = d->evecs + cur * m;
        curF = d->Fvecs + cur * m;
        // e = X (FDS - (FDS)^T)^T X^T with A*B^T products (pscf/purif.c:524-565)
        if (gemm_pair(d, d->F, d->D, d->W, 1, d->W, d->S, e, 1, nullptr)) return -1;
        if (gtf_antisym(d->stream, e, d->W2, ld)) return -1;
        if (gemm_pair(d, d->X, d->W2, d->W, 1, d->W, d->X, e, 1, nullptr)) return -1;
        double dots[MAX_DOTS];
        if (gtf_multidot(d->stream, e, d->evecs, p->len_diis, m, d->partial, d->scal)) return -1;
        PD_CHECK(hipMemcpyAsync(dots, d->scal, sizeof(double) * MAX_DOTS, hipMemcpyDeviceToHost, d->stream));
        PD_CHECK(hipStreamSynchronize(d->stream));
        for (int i = 0; i < p->len_diis; i++) {
            p->b_mat[cur * LB + i] = dots[i];
            p->b_mat[i * LB + cur] = dots[i];
        }
        d->last_resid = std::sqrt(std::max(dots[cur], 0.0));
        p->bmax = -DBL_MAX;
        for (int i = 0; i < p->len_diis; i++)
            if (p->bmax < p->b_mat[i * LB + i]) {
                p->bmax = p->b_mat[i * LB + i];
                p->bmax_id = i;
            }
    }
    // F_orth = X F X (pscf/purif.c:598-603)
    if (gemm_pair(d, d->X, d->F, d->W, 1, d->W, d->X, curF, 1, nullptr)) return -1;
    if (iter > 1) {
        const int sz = p->len_diis + 1;
        double inv[(MAX_DIIS + 1) * (MAX_DIIS + 1)], coeffs[MAX_DIIS + 1];
        small_inverse(sz, p->b_mat, LB, inv);
        for (int i = 0; i < sz; i++) coeffs[i] = -inv[i * sz + sz - 1];
        PD_CHECK(hipMemcpyAsync(d->coef, coeffs, sizeof(double) * p->len_diis, hipMemcpyHostToDevice, d->stream));
        if (gtf_lincomb(d->stream, d->F, d->Fvecs, d->coef, p->len_diis, m)) return -1;
        curF = d->F;
    }
    return d2h(d, F_block, curF, p->ldx);
}

int gtf_purif_set_precision(purif_t* p, int mode, double until)
{
    if (!p || !p->dev || until <= 0.0 || mode < 0 || mode > 2) return -1;
    purif_dev* d = p->dev;
    if (mode == 2 && !d->oz_mem) {
        const long total = (long)d->ld * d->ld;
        if (hipMalloc(&d->oz_mem, gtf_oz_scratch_bytes(total, OZ_SLICES_MAX)) != hipSuccess) {
            d->oz_mem = nullptr;
            return -2;
        }
        gtf_oz_scratch_carve(d->oz_mem, total, d->ns, d->oz);
    }
    d->mode = mode;
    d->low_until = until;
    return 0;
}

int gtf_purif_set_split_residual(purif_t* p, double resid)
{
    if (!p || !p->dev || !(resid >= 0.0)) return -1;
    p->dev->split_resid = resid;
    return 0;
}

int gtf_purif_set_slices(purif_t* p, int ns)
{
    if (!p || !p->dev || ns < OZ_SLICES_MIN || ns > OZ_SLICES_MAX) return -1;
    purif_dev* d = p->dev;
    d->ns = ns;
    if (d->oz_mem) gtf_oz_scratch_carve(d->oz_mem, (long)d->ld * d->ld, ns, d->oz);
    return 0;
}

int gtf_purif_fp32_iters(purif_t* p) { return p && p->dev ? p->dev->low_iters : -1; }

void compute_diis(PFock_t pfock, purif_t* p, double* D_block, double* F_block, int iter)
{
    (void)pfock;
    if (diis(p, D_block, F_block, iter)) fprintf(stderr, "compute_diis: GPU execution failed\n");
}

}  // extern "C"
