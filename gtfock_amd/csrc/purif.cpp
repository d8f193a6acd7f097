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
#include "linalg.h"
#include "ozaki_launch.h"
#include "screen.h"

struct purif_dev {
    hipStream_t stream = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
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
    int ns = 5;
    int low_iters = 0;
    void* oz_mem = nullptr;
    OzScratch oz;
};

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
    PD_CHECK(hipMemcpy2DAsync(dst, sizeof(double) * d->ld, src, sizeof(double) * ldx, sizeof(double) * d->n, d->n,
                              hipMemcpyHostToDevice, d->stream));
    return 0;
}

static int d2h(purif_dev* d, double* dst, const double* src, int ldx)
{
    PD_CHECK(hipMemcpy2DAsync(dst, sizeof(double) * ldx, src, sizeof(double) * d->ld, sizeof(double) * d->n, d->n,
                              hipMemcpyDeviceToHost, d->stream));
    PD_CHECK(hipStreamSynchronize(d->stream));
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

// D2 = D * D and D3 = D * D2 in split BF16 (ozaki.h), timed like gemm_pair
static int split_square_cube(purif_dev* d, double* acc)
{
    PD_CHECK(hipEventRecord(d->e0, d->stream));
    OzOperands o;
    o.maxld = d->ld;
    if (gtf_oz_square_cube(d->stream, o, d->oz, d->D, d->D2, d->D3)) return -1;
    PD_CHECK(hipEventRecord(d->e1, d->stream));
    PD_CHECK(hipEventSynchronize(d->e1));
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, d->e0, d->e1);
    if (acc) *acc += 1e-3 * ms;
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
    if (!ok) { destroy_purif(p); return nullptr; }
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
        double* bufs[] = {d->F, d->D, d->D2, d->D3, d->X, d->S, d->W, d->W2, d->evecs, d->Fvecs,
                          d->scal, d->coef, d->partial, (double*)d->oz_mem};
        for (double* b : bufs)
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

static int purify(purif_t* p, double* F_block, double* D_block, int* iters)
{
    purif_dev* d = p->dev;
    const int n = d->n, ld = d->ld, nocc = p->nobtls;
    const long m = (long)ld * ld;
    if (load_xs(p) || h2d(d, d->F, F_block, p->ldx)) return -1;
    double h[3];
    if (gtf_gershgorin(d->stream, d->F, n, ld, d->scal)) return -1;
    PD_CHECK(hipMemcpyAsync(h, d->scal, sizeof h, hipMemcpyDeviceToHost, d->stream));
    PD_CHECK(hipStreamSynchronize(d->stream));
    const double hmin = h[0], hmax = h[1], mu = h[2] / n;
    const double lambda = std::min((double)nocc / (hmax - mu), (double)(n - nocc) / (mu - hmin));
    if (gtf_purif_init(d->stream, d->F, d->D, n, ld, -lambda / n, lambda * mu / n + (double)nocc / n)) return -1;
    int it;
    bool low = d->mode != 0;
    d->low_iters = 0;
    for (it = 0; it < 200; it++) {
        if (low && d->mode == 2) {
            if (split_square_cube(d, &p->timedgemm)) return -1;
        } else if (gemm_pair(d, d->D, d->D, d->D2, 0, d->D, d->D2, d->D3, 0, &p->timedgemm, low ? 1 : 0)) {
            return -1;
        }
        d->low_iters += low ? 1 : 0;
        const double t0 = wtime();
        if (gtf_purif_stats(d->stream, d->D, d->D2, d->D3, n, ld, d->partial, d->scal)) return -1;
        double s[3];
        PD_CHECK(hipMemcpyAsync(s, d->scal, sizeof s, hipMemcpyDeviceToHost, d->stream));
        PD_CHECK(hipStreamSynchronize(d->stream));
        const double err = sqrt(s[0]), c = s[1] / s[2];
        if (gtf_purif_update(d->stream, d->D, d->D2, d->D3, m, c)) return -1;
        p->timetr += wtime() - t0;
        if (low && err < d->low_until) low = false;
        if (err < 1e-11) break;
    }
    // D = X D X (pscf/purif.c:469-476)
    if (gemm_pair(d, d->X, d->D, d->W, 1, d->W, d->X, d->D, 1, &p->timedgemm)) return -1;
    if (d2h(d, D_block, d->D, p->ldx)) return -1;
    *iters = it;
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
