// pfock.cpp -- the PFock API (pfock/pfock.h:241-565) on one GPU.
//
// PFock_create      pfock/pfock.c:801-997  -> device basis + shell pairs,
//                   Schwarz screening, quartet rows (GtfSystem)
// PFock_computeFock pfock/pfock.c:1274-1729 -> per-class fused ERI + J/K
//                   kernels, then GA_Symmetrize of J and K on the device
// PFock_getMat      pfock/pfock.c:1134-1188 -> F = J + K, 2-D copy to host
#include <sys/time.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <vector>

#include "../../include/gtfock_amd.h"
#include "../../include/pfock.h"
#include "engine.h"
#include "hostcopy.h"
#include "oneel.h"
#include "screen.h"

#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

struct PFock {
    BasisSet* basis = nullptr;
    GtfSystem sys;
    hipStream_t stream = nullptr;
    PinnedStage stage;  // host blocks in and out
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    int nbf = 0, nshells = 0, natoms = 0, maxnfuncs = 0;
    int nprow = 1, npcol = 1, nbp_p = 4;
    int max_numdmat = 1, num_dmat = 1, nosymm = 0;
    double tolscr = 0, tolscr2 = 0;
    int committed = 0;
    int fused_nd = 0;  // densities the last build digested from each integral
    std::vector<double*> d_D, d_J, d_K;
    double* d_F = nullptr;
    double *d_S = nullptr, *d_H = nullptr, *d_X = nullptr;
    // reference-layout CSR (pfock/pfock.h:22-29)
    int nnz = 0;
    std::vector<int> shellptr, shellid, shellrid;
    std::vector<double> shellvalue;
    // statistics (pfock/pfock.h:152-177)
    double usq = 0, uitl = 0, tsq = 0;
    double timepass = 0, timecomp = 0, timeinit = 0, timescreen = 0;
    double mem_cpu = 0;
    long nbuilds = 0;
    std::vector<std::vector<double>> host_mirror;
};

static double wtime()
{
    struct timeval tv;
    gettimeofday(&tv, nullptr);
    return tv.tv_sec + 1e-6 * tv.tv_usec;
}

static PFockStatus_t hip_fail(hipError_t e, const char* what)
{
    GTF_PRINTF("%s failed: %s\n", what, hipGetErrorString(e));
    return e == hipErrorOutOfMemory ? PFOCK_STATUS_ALLOC_FAILED : PFOCK_STATUS_EXECUTION_FAILED;
}

#define PF_HIP(call, what)                                   \
    do {                                                     \
        hipError_t e__ = (call);                             \
        if (e__ != hipSuccess) return hip_fail(e__, what);   \
    } while (0)

static int gpu_present()
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return 0;
    return 1;
}

extern "C" {

PFockStatus_t PFock_create(BasisSet_t basis, int nprow, int npcol, int ntasks, double tolscr, int max_numdmat,
                           int symm, PFock_t* _pfock)
{
    if (!basis || !_pfock) return PFOCK_STATUS_INVALID_VALUE;
    *_pfock = nullptr;
    if (!gpu_present()) {
        GTF_PRINTF("no HIP device is visible; this engine runs only on the GPU\n");
        return PFOCK_STATUS_INIT_FAILED;
    }
    const int nshells = basis->nshells;
    // one process: nprow x npcol must be 1 x 1 (pfock/pfock.c:838-842)
    if (nprow <= 0 || nprow > nshells || npcol <= 0 || npcol > nshells || nprow * npcol > 1) {
        GTF_PRINTF("Invalid nprow or npcol\n");
        return PFOCK_STATUS_INVALID_VALUE;
    }
    if (tolscr < 0.0) {
        GTF_PRINTF("Invalid screening threshold\n");
        return PFOCK_STATUS_INVALID_VALUE;
    }
    if (max_numdmat <= 0) {
        GTF_PRINTF("Invalid number of density matrices\n");
        return PFOCK_STATUS_INVALID_VALUE;
    }
    if (symm == 0) {
        GTF_PRINTF("non-symmetric density matrices are not supported\n");
        return PFOCK_STATUS_INVALID_VALUE;
    }
    PFock* pf = new (std::nothrow) PFock();
    if (!pf) return PFOCK_STATUS_ALLOC_FAILED;
    pf->basis = basis;
    pf->nbf = basis->nbf;
    pf->nshells = nshells;
    pf->natoms = basis->natoms;
    pf->maxnfuncs = basis->maxdim;
    pf->nprow = nprow;
    pf->npcol = npcol;
    pf->tolscr = tolscr;
    pf->tolscr2 = tolscr * tolscr;
    pf->max_numdmat = max_numdmat;
    pf->num_dmat = max_numdmat;
    // tasks per process (pfock/pfock.c:867-877); kept for the statistics
    int minns = nshells / std::max(nprow, npcol);
    pf->nbp_p = ntasks >= minns ? minns : (ntasks <= 0 ? std::min(4, minns) : ntasks);
    hipError_t e = hipStreamCreate(&pf->stream);
    if (e != hipSuccess) { delete pf; return hip_fail(e, "hipStreamCreate"); }
    (void)hipEventCreate(&pf->ev0);
    (void)hipEventCreate(&pf->ev1);
    double t0 = wtime();
    pf->sys.balance = true;  // workgroup records for PFock_getStatistics
    int rc = pf->sys.init(basis, pf->stream);
    if (rc == 0) rc = pf->sys.screen(tolscr);
    if (rc != 0) {
        GTF_PRINTF("device setup failed (%d)\n", rc);
        PFock_destroy(pf);
        return rc == -2 ? PFOCK_STATUS_ALLOC_FAILED : PFOCK_STATUS_INIT_FAILED;
    }
    pf->timescreen = wtime() - t0;
    pf->sys.csr(pf->shellptr, pf->shellid, pf->shellrid, pf->shellvalue);
    pf->nnz = (int)pf->shellid.size();
    const long nn = (long)pf->nbf * pf->nbf;
    // the densities (and their J, K) one after another, nn apart, so one
    // build digests every committed density from each integral
    double *dd = pf->sys.dalloc<double>(nn * max_numdmat), *jj = pf->sys.dalloc<double>(nn * max_numdmat),
           *kk = pf->sys.dalloc<double>(nn * max_numdmat);
    if (!dd || !jj || !kk) { PFock_destroy(pf); return PFOCK_STATUS_ALLOC_FAILED; }
    if (gtf_fill(pf->stream, dd, nn * max_numdmat, 0.0) || gtf_fill(pf->stream, jj, nn * max_numdmat, 0.0) ||
        gtf_fill(pf->stream, kk, nn * max_numdmat, 0.0)) {
        PFock_destroy(pf);
        return PFOCK_STATUS_EXECUTION_FAILED;
    }
    for (int i = 0; i < max_numdmat; i++) {
        pf->d_D.push_back(dd + i * nn);
        pf->d_J.push_back(jj + i * nn);
        pf->d_K.push_back(kk + i * nn);
    }
    pf->d_F = pf->sys.dalloc<double>(nn);
    if (!pf->d_F) { PFock_destroy(pf); return PFOCK_STATUS_ALLOC_FAILED; }
    pf->host_mirror.resize(4);
    pf->mem_cpu = sizeof(int) * (pf->shellptr.size() + 2.0 * pf->nnz) + sizeof(double) * pf->nnz +
                  sizeof(double) * pf->sys.h_v.size();
    const double ns = pf->nnz;
    pf->tsq = ns * (ns + 1) / 2.0;
    pf->committed = 0;
    if (hipStreamSynchronize(pf->stream) != hipSuccess) { PFock_destroy(pf); return PFOCK_STATUS_EXECUTION_FAILED; }
    *_pfock = pf;
    return PFOCK_STATUS_SUCCESS;
}

PFockStatus_t PFock_destroy(PFock_t pf)
{
    if (!pf) return PFOCK_STATUS_SUCCESS;
    if (pf->stream) (void)hipStreamSynchronize(pf->stream);
    pf->sys.release();
    if (pf->ev0) (void)hipEventDestroy(pf->ev0);
    if (pf->ev1) (void)hipEventDestroy(pf->ev1);
    if (pf->stream) (void)hipStreamDestroy(pf->stream);
    delete pf;
    return PFOCK_STATUS_SUCCESS;
}

PFockStatus_t PFock_setNumDenMat(int numdmat, PFock_t pf)
{
    if (pf->committed == 1) {
        GTF_PRINTF("Can't change number of matrices after PFock_commitDenMats() is called.\n");
        return PFOCK_STATUS_EXECUTION_FAILED;
    }
    if (numdmat <= 0 || numdmat > pf->max_numdmat) {
        GTF_PRINTF("Invalid number of density matrices\n");
        return PFOCK_STATUS_INVALID_VALUE;
    }
    pf->num_dmat = numdmat;
    return PFOCK_STATUS_SUCCESS;
}

static bool bad_block(PFock_t pf, int rs, int re, int cs, int ce, int stride)
{
    return rs < 0 || cs < 0 || re >= pf->nbf || ce >= pf->nbf || re < rs || ce < cs || stride < ce - cs + 1;
}

PFockStatus_t PFock_putDenMat(int rowstart, int rowend, int colstart, int colend, int stride, double* dmat,
                              int index, PFock_t pf)
{
    if (pf->committed == 1) {
        GTF_PRINTF("Can't change density matrix after PFock_commitDenMats() is called.\n");
        return PFOCK_STATUS_EXECUTION_FAILED;
    }
    if (index < 0 || index >= pf->num_dmat) {
        GTF_PRINTF("Invalid index\n");
        return PFOCK_STATUS_INVALID_VALUE;
    }
    if (!dmat || bad_block(pf, rowstart, rowend, colstart, colend, stride)) {
        GTF_PRINTF("Invalid block\n");
        return PFOCK_STATUS_INVALID_VALUE;
    }
    double* dst = pf->d_D[index] + (long)rowstart * pf->nbf + colstart;
    PF_HIP(stage_h2d(pf->stage, pf->stream, dst, pf->nbf, dmat, stride, rowend - rowstart + 1, colend - colstart + 1),
           "stage_h2d");
    return PFOCK_STATUS_SUCCESS;
}

PFockStatus_t PFock_putDenMatGA(int ga, int index, PFock_t pf)
{
    (void)ga; (void)index; (void)pf;
    GTF_PRINTF("Global Arrays handles are not available on this engine\n");
    return PFOCK_STATUS_INVALID_VALUE;
}

PFockStatus_t PFock_fillDenMat(double value, int index, PFock_t pf)
{
    if (pf->committed == 1) {
        GTF_PRINTF("Can't change density matrix after PFock_commitDenMats() is called.\n");
        return PFOCK_STATUS_EXECUTION_FAILED;
    }
    if (index < 0 || index >= pf->num_dmat) {
        GTF_PRINTF("Invalid index\n");
        return PFOCK_STATUS_INVALID_VALUE;
    }
    if (gtf_fill(pf->stream, pf->d_D[index], (long)pf->nbf * pf->nbf, value)) return PFOCK_STATUS_EXECUTION_FAILED;
    return PFOCK_STATUS_SUCCESS;
}

PFockStatus_t PFock_commitDenMats(PFock_t pf)
{
    PF_HIP(hipStreamSynchronize(pf->stream), "hipStreamSynchronize");
    pf->committed = 1;
    return PFOCK_STATUS_SUCCESS;
}

PFockStatus_t PFock_sync(PFock_t pf)
{
    PF_HIP(hipStreamSynchronize(pf->stream), "hipStreamSynchronize");
    return PFOCK_STATUS_SUCCESS;
}

static const double* matrix_ptr(PFock_t pf, PFockMatType_t type, int index, PFockStatus_t* st)
{
    *st = PFOCK_STATUS_SUCCESS;
    switch (type) {
    case PFOCK_MAT_TYPE_D: return pf->d_D[index];
    case PFOCK_MAT_TYPE_J: return pf->d_J[index];
    case PFOCK_MAT_TYPE_K: return pf->d_K[index];
    case PFOCK_MAT_TYPE_F:
        if (gtf_add(pf->stream, pf->d_J[index], pf->d_K[index], pf->d_F, (long)pf->nbf * pf->nbf))
            *st = PFOCK_STATUS_EXECUTION_FAILED;
        return pf->d_F;
    }
    *st = PFOCK_STATUS_INVALID_VALUE;
    return nullptr;
}

PFockStatus_t PFock_getMat(PFock_t pf, PFockMatType_t type, int index, int rowstart, int rowend, int colstart,
                           int colend, int stride, double* mat)
{
    if (index < 0 || index >= pf->max_numdmat) {
        GTF_PRINTF("Invalid index\n");
        return PFOCK_STATUS_INVALID_VALUE;
    }
    if (!mat || bad_block(pf, rowstart, rowend, colstart, colend, stride)) {
        GTF_PRINTF("Invalid block\n");
        return PFOCK_STATUS_INVALID_VALUE;
    }
    PFockStatus_t st;
    const double* src = matrix_ptr(pf, type, index, &st);
    if (st != PFOCK_STATUS_SUCCESS) return st;
    PF_HIP(stage_d2h(pf->stage, pf->stream, mat, stride, src + (long)rowstart * pf->nbf + colstart, pf->nbf,
                     rowend - rowstart + 1, colend - colstart + 1),
           "stage_d2h");
    return PFOCK_STATUS_SUCCESS;
}

PFockStatus_t PFock_getMatGA(PFock_t pf, PFockMatType_t type, int index, int ga)
{
    (void)pf; (void)type; (void)index; (void)ga;
    GTF_PRINTF("Global Arrays handles are not available on this engine\n");
    return PFOCK_STATUS_INVALID_VALUE;
}

PFockStatus_t PFock_getLocalMatInds(PFock_t pf, int* rowstart, int* rowend, int* colstart, int* colend)
{
    *rowstart = 0;
    *rowend = pf->nbf - 1;
    *colstart = 0;
    *colend = pf->nbf - 1;
    return PFOCK_STATUS_SUCCESS;
}

PFockStatus_t PFock_getLocalMatPtr(PFock_t pf, PFockMatType_t type, int index, int* rowstart, int* rowend,
                                   int* colstart, int* colend, int* stride, double** mat)
{
    if (index < 0 || index >= pf->max_numdmat) {
        GTF_PRINTF("Invalid index\n");
        return PFOCK_STATUS_INVALID_VALUE;
    }
    std::vector<double>& h = pf->host_mirror[type];
    h.resize((size_t)pf->nbf * pf->nbf);
    PFockStatus_t st = PFock_getMat(pf, type, index, 0, pf->nbf - 1, 0, pf->nbf - 1, pf->nbf, h.data());
    if (st != PFOCK_STATUS_SUCCESS) return st;
    PFock_getLocalMatInds(pf, rowstart, rowend, colstart, colend);
    *stride = pf->nbf;
    *mat = h.data();
    return PFOCK_STATUS_SUCCESS;
}

PFockStatus_t PFock_getMatGAHandle(PFock_t pf, PFockMatType_t type, int index, int* ga)
{
    (void)pf; (void)type; (void)index;
    *ga = 0;
    GTF_PRINTF("Global Arrays handles are not available on this engine\n");
    return PFOCK_STATUS_INVALID_VALUE;
}

PFockStatus_t PFock_computeFock(BasisSet_t basis, PFock_t pf)
{
    (void)basis;
    pf->committed = 0;
    const double t0 = wtime();
    const long nn = (long)pf->nbf * pf->nbf;
    PF_HIP(hipEventRecord(pf->ev0, pf->stream), "hipEventRecord");
    const int nd = pf->num_dmat;
    if (gtf_fill(pf->stream, pf->d_J[0], nn * nd, 0.0) || gtf_fill(pf->stream, pf->d_K[0], nn * nd, 0.0))
        return PFOCK_STATUS_EXECUTION_FAILED;
    // every density from each integral (update_F's num_dmat loop,
    // pfock/fock_task.c:46) when the per-density LDS work areas fit, else
    // one build per density
    if (nd == 1 || pf->sys.fits_densities(nd)) {
        if (pf->sys.jk(pf->d_D[0], pf->d_J[0], pf->d_K[0], nd, nn)) return PFOCK_STATUS_EXECUTION_FAILED;
        if (pf->sys.collect_times()) return PFOCK_STATUS_EXECUTION_FAILED;
        pf->fused_nd = nd;
    } else {
        pf->fused_nd = 1;
        for (int i = 0; i < nd; i++) {
            if (pf->sys.jk(pf->d_D[i], pf->d_J[i], pf->d_K[i])) return PFOCK_STATUS_EXECUTION_FAILED;
            if (pf->sys.collect_times()) return PFOCK_STATUS_EXECUTION_FAILED;
        }
    }
    for (int i = 0; i < nd; i++)
        if (pf->sys.symmetrize(pf->d_J[i]) || pf->sys.symmetrize(pf->d_K[i])) return PFOCK_STATUS_EXECUTION_FAILED;
    PF_HIP(hipEventRecord(pf->ev1, pf->stream), "hipEventRecord");
    PF_HIP(hipEventSynchronize(pf->ev1), "hipEventSynchronize");
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, pf->ev0, pf->ev1);
    pf->timecomp = ms * 1e-3;
    pf->timepass = wtime() - t0;
    double usq = 0, uitl = 0;
    for (int c = 0; c < GTF_NCLASSES; c++) {
        usq += (double)pf->sys.nq[c];
        uitl += (double)pf->sys.nq[c] * gtf_eri_class(c)->nints;
    }
    pf->usq = usq * pf->num_dmat;
    pf->uitl = uitl * pf->num_dmat;
    pf->nbuilds++;
    return PFOCK_STATUS_SUCCESS;
}

// ---- one-electron matrices ------------------------------------------------------

static PFockStatus_t ensure_oneel(PFock_t pf)
{
    if (pf->d_S && pf->d_H) return PFOCK_STATUS_SUCCESS;
    const long nn = (long)pf->nbf * pf->nbf;
    double* S = pf->sys.dalloc<double>(nn);
    double* T = pf->sys.dalloc<double>(nn);
    double* V = pf->sys.dalloc<double>(nn);
    if (!S || !T || !V) return PFOCK_STATUS_ALLOC_FAILED;
    GtfSystem& s = pf->sys;
    BasisView bv{s.nshells, s.nbf, s.natoms, s.d_L, s.d_nprim, s.d_poff, s.d_fstart, s.d_atom,
                 s.d_alpha, s.d_coef, s.d_center, s.d_atom_xyz, s.d_atom_Z};
    if (gtf_oneel(pf->stream, bv, s.d_boys, pf->nbf, S, T, V)) return PFOCK_STATUS_EXECUTION_FAILED;
    if (gtf_add(pf->stream, T, V, T, nn)) return PFOCK_STATUS_EXECUTION_FAILED;
    pf->d_S = S;
    pf->d_H = T;
    PF_HIP(hipStreamSynchronize(pf->stream), "hipStreamSynchronize");
    return PFOCK_STATUS_SUCCESS;
}

PFockStatus_t PFock_createCoreHMat(PFock_t pf, BasisSet_t basis)
{
    (void)basis;
    return ensure_oneel(pf);
}

PFockStatus_t PFock_destroyCoreHMat(PFock_t pf)
{
    (void)pf;
    return PFOCK_STATUS_SUCCESS;
}

static PFockStatus_t get_block(PFock_t pf, const double* src, int rs, int re, int cs, int ce, int stride, double* mat)
{
    if (!src) {
        GTF_PRINTF("matrix has not been created\n");
        return PFOCK_STATUS_EXECUTION_FAILED;
    }
    if (!mat || bad_block(pf, rs, re, cs, ce, stride)) {
        GTF_PRINTF("Invalid block\n");
        return PFOCK_STATUS_INVALID_VALUE;
    }
    PF_HIP(stage_d2h(pf->stage, pf->stream, mat, stride, src + (long)rs * pf->nbf + cs, pf->nbf, re - rs + 1,
                     ce - cs + 1),
           "stage_d2h");
    return PFOCK_STATUS_SUCCESS;
}

PFockStatus_t PFock_getCoreHMat(PFock_t pf, int rowstart, int rowend, int colstart, int colend, int stride,
                                double* mat)
{
    return get_block(pf, pf->d_H, rowstart, rowend, colstart, colend, stride, mat);
}

// X = U diag(lambda^-1/2) U^T (pfock/pfock.c:1823-1861): rocsolver syevd, then
// X_ij = sum_k W_ki W_kj lambda_k^-1/2 with W = U^T (row k = eigenvector k)
__global__ void inv_sqrt_assemble_kernel(const double* W, const double* lam, int n, double* X)
{
    const int i = blockIdx.y * 16 + threadIdx.y, j = blockIdx.x * 16 + threadIdx.x;
    if (i >= n || j >= n) return;
    double s = 0.0;
    for (int k = 0; k < n; k++) s += W[(long)k * n + i] * W[(long)k * n + j] / sqrt(lam[k]);
    X[(long)i * n + j] = s;
}

PFockStatus_t PFock_createOvlMat(PFock_t pf, BasisSet_t basis)
{
    (void)basis;
    PFockStatus_t st = ensure_oneel(pf);
    if (st != PFOCK_STATUS_SUCCESS) return st;
    if (pf->d_X) return PFOCK_STATUS_SUCCESS;
    const int n = pf->nbf;
    const long nn = (long)n * n;
    double* W = pf->sys.dalloc<double>(nn);
    double* lam = pf->sys.dalloc<double>(n);
    double* work = pf->sys.dalloc<double>(n);
    int* info = pf->sys.dalloc<int>(1);
    double* X = pf->sys.dalloc<double>(nn);
    if (!W || !lam || !work || !info || !X) return PFOCK_STATUS_ALLOC_FAILED;
    PF_HIP(hipMemcpyAsync(W, pf->d_S, sizeof(double) * nn, hipMemcpyDeviceToDevice, pf->stream), "hipMemcpyAsync");
    rocblas_handle h;
    if (rocblas_create_handle(&h) != rocblas_status_success) return PFOCK_STATUS_INIT_FAILED;
    rocblas_set_stream(h, pf->stream);
    rocblas_status rs = rocsolver_dsyevd(h, rocblas_evect_original, rocblas_fill_upper, n, W, n, lam, work, info);
    PF_HIP(hipStreamSynchronize(pf->stream), "hipStreamSynchronize");
    rocblas_destroy_handle(h);
    int hinfo = 0;
    PF_HIP(hipMemcpy(&hinfo, info, sizeof(int), hipMemcpyDeviceToHost), "hipMemcpy");
    if (rs != rocblas_status_success || hinfo != 0) {
        GTF_PRINTF("eigensolver failed (status %d, info %d)\n", (int)rs, hinfo);
        return PFOCK_STATUS_EXECUTION_FAILED;
    }
    dim3 g((n + 15) / 16, (n + 15) / 16), b(16, 16);
    inv_sqrt_assemble_kernel<<<g, b, 0, pf->stream>>>(W, lam, n, X);
    PF_HIP(hipGetLastError(), "inv_sqrt_assemble_kernel");
    PF_HIP(hipStreamSynchronize(pf->stream), "hipStreamSynchronize");
    pf->d_X = X;
    return PFOCK_STATUS_SUCCESS;
}

PFockStatus_t PFock_destroyOvlMat(PFock_t pf)
{
    (void)pf;
    return PFOCK_STATUS_SUCCESS;
}

PFockStatus_t PFock_getOvlMat(PFock_t pf, int rowstart, int rowend, int colstart, int colend, int stride,
                              double* mat)
{
    return get_block(pf, pf->d_S, rowstart, rowend, colstart, colend, stride, mat);
}

PFockStatus_t PFock_getOvlMat2(PFock_t pf, int rowstart, int rowend, int colstart, int colend, int stride,
                               double* mat)
{
    return get_block(pf, pf->d_X, rowstart, rowend, colstart, colend, stride, mat);
}

PFockStatus_t PFock_getMemorySize(PFock_t pf, double* mem_cpu)
{
    *mem_cpu = pf->mem_cpu;
    return PFOCK_STATUS_SUCCESS;
}

// pfock/pfock.c:1918-2020, one process
PFockStatus_t PFock_getStatistics(PFock_t pf)
{
    printf("    PFock Statistics:\n");
    printf("      average totaltime   = %.3g\n"
           "      average timecomp    = %.3g\n"
           "      comp/total = %.3g\n",
           pf->timepass, pf->timecomp, pf->timepass > 0 ? pf->timecomp / pf->timepass : 0.0);
    printf("      usq = %.4g (lb = %.3g)\n"
           "      uitl = %.4g (lb = %.3g)\n"
           "      nsq = %.4g (screening = %.3g)\n",
           pf->usq, 1.0, pf->uitl, 1.0, pf->tsq, pf->tsq > 0 ? pf->usq / pf->tsq : 0.0);
    // One process: the balance is that of the persistent workgroups of the
    // last build's tile launches (slowest workgroup over the mean, summed
    // over the launches); a steal is a tile a workgroup takes from the
    // shared class counter after its first.
    double lb = 1.0;
    long steals = 0, nwg = 0;
    if (pf->sys.balance_stats(&lb, &steals, &nwg)) return PFOCK_STATUS_EXECUTION_FAILED;
    printf("      load blance = %.3lf\n", lb);
    printf("      steals = %.3g (average = %.3g)\n", (double)steals, nwg > 0 ? (double)steals / nwg : 0.0);
    fflush(stdout);
    return PFOCK_STATUS_SUCCESS;
}

// ---- extra entry points (include/gtfock_amd.h) --------------------------------

int gtf_pfock_fused_densities(PFock_t pf) { return pf ? pf->fused_nd : -1; }

int gtf_pfock_balance(PFock_t pf, double* out)
{
    if (!pf || !out) return -1;
    double lb = 1.0;
    long steals = 0, nwg = 0;
    if (pf->sys.balance_stats(&lb, &steals, &nwg)) return -1;
    out[0] = lb;
    out[1] = (double)steals;
    out[2] = (double)nwg;
    return 0;
}

int gtf_pfock_setup_times(PFock_t pf, double* out)
{
    if (!pf || !out) return -1;
    for (int k = 0; k < GtfSystem::NSETUP; k++) out[k] = pf->sys.setup_s[k];
    return GtfSystem::NSETUP;
}

int gtf_pfock_stats(PFock_t pf, double* out)
{
    out[0] = pf->usq;
    out[1] = pf->uitl;
    out[2] = pf->nnz;
    out[3] = pf->sys.vmax;
    out[4] = pf->timecomp;
    out[5] = pf->timepass;
    out[6] = pf->timescreen;
    out[7] = (double)pf->sys.dev_bytes;
    return 0;
}

int gtf_pfock_csr(PFock_t pf, int* shellptr, int* shellid, int* shellrid, double* shellvalue)
{
    std::copy(pf->shellptr.begin(), pf->shellptr.end(), shellptr);
    std::copy(pf->shellid.begin(), pf->shellid.end(), shellid);
    std::copy(pf->shellrid.begin(), pf->shellrid.end(), shellrid);
    std::copy(pf->shellvalue.begin(), pf->shellvalue.end(), shellvalue);
    return pf->nnz;
}

int gtf_pfock_class_counts(PFock_t pf, long* counts)
{
    for (int c = 0; c < GTF_NCLASSES; c++) counts[c] = pf->sys.nq[c];
    return GTF_NCLASSES;
}

// unique screened quartets as canonical (M,N,P,Q) shell ids, class by class
long gtf_pfock_quartets(PFock_t pf, int* out, long cap) { return pf->sys.quartets(out, cap); }

// the engine's device buffers of D / J / K (borrowed; valid until destroy)
int gtf_pfock_device_ptrs(PFock_t pf, int index, double** D, double** J, double** K)
{
    if (index < 0 || index >= pf->max_numdmat) return -1;
    *D = pf->d_D[index];
    *J = pf->d_J[index];
    *K = pf->d_K[index];
    return 0;
}

void* gtf_pfock_stream(PFock_t pf) { return (void*)pf->stream; }

int gtf_pfock_class_info(PFock_t pf, int cls, int* info)
{
    if (!pf || !info || cls < 0 || cls >= GTF_NCLASSES) return -1;
    const EriClassInfo* ci = gtf_eri_class(cls);
    info[0] = ci->la;
    info[1] = ci->lb;
    info[2] = ci->lc;
    info[3] = ci->ld;
    info[4] = ci->nc;
    info[5] = ci->nsplit;
    info[6] = pf->sys.ntiles[cls];
    info[7] = pf->sys.ntiles_lo[cls];
    // the exchanged tile kernel runs the class: ket components per thread
    const bool ex = pf->sys.exchanged(cls);
    info[8] = ex ? 1 : 0;
    info[9] = ex ? (2 * ci->lc + 1) * (2 * ci->ld + 1) / ci->nsplit_t : 0;
    info[10] = pf->sys.shape[cls].boys_lds;
    return 0;
}

int gtf_pfock_class_work(PFock_t pf, long* nquartets, double* nprimquartets)
{
    for (int c = 0; c < GTF_NCLASSES; c++) {
        nquartets[c] = pf->sys.nq[c];
        nprimquartets[c] = pf->sys.nprimq[c];
    }
    return GTF_NCLASSES;
}

int gtf_pfock_prim_census(PFock_t pf, double* npq_fp64, double* npq_fp32)
{
    if (!pf || !npq_fp64 || !npq_fp32) return -1;
    return pf->sys.prim_census(npq_fp64, npq_fp32) ? -1 : GTF_NCLASSES;
}

int gtf_pfock_set_precision(PFock_t pf, double tau)
{
    if (tau < 0.0) return -1;
    if (hipStreamSynchronize(pf->stream) != hipSuccess) return -1;
    pf->sys.tau = tau;
    return pf->sys.build_tiles();
}

int gtf_pfock_set_prim_tol(PFock_t pf, double tol)
{
    if (!pf || !(tol >= 0.0)) return -1;
    pf->sys.prim_tol = tol;
    return 0;
}

int gtf_pfock_mixed_counts(PFock_t pf, long* nq_lo, double* nprimq_lo)
{
    for (int c = 0; c < GTF_NCLASSES; c++) {
        nq_lo[c] = pf->sys.nq_lo[c];
        nprimq_lo[c] = pf->sys.nprimq_lo[c];
    }
    return GTF_NCLASSES;
}

int gtf_pfock_set_timing(PFock_t pf, int on)
{
    GtfSystem& s = pf->sys;
    if (on && !s.timing)
        for (hipEvent_t& e : s.tev)
            if (!e && hipEventCreate(&e) != hipSuccess) return -1;
    s.timing = on != 0;
    for (int c = 0; c < GTF_NCLASSES; c++) {
        s.class_ms[c] = 0.0;
        s.class_launches[c] = 0;
    }
    return 0;
}

int gtf_pfock_set_serial(PFock_t pf, int on)
{
    if (!pf) return -1;
    pf->sys.serial = on != 0;
    return 0;
}

int gtf_pfock_class_times(PFock_t pf, double* ms, long* launches)
{
    for (int c = 0; c < GTF_NCLASSES; c++) {
        ms[c] = pf->sys.class_ms[c];
        launches[c] = pf->sys.class_launches[c];
    }
    return GTF_NCLASSES;
}

double gtf_pfock_device_bytes(PFock_t pf) { return (double)pf->sys.dev_bytes; }

static_assert(GTF_SHAPE_FIELDS == GTF_TILE_SHAPE_FIELDS, "tile shape size");

int gtf_pfock_autotune(PFock_t pf, int reps, double* ms)
{
    if (!pf || reps <= 0) return -1;
    // scratch J / K: the build's own accumulators keep their contents
    const size_t bytes = sizeof(double) * (size_t)pf->sys.mat_total;
    double *J = nullptr, *K = nullptr;
    if (hipMalloc(&J, bytes) != hipSuccess) return -1;
    int rc = -1;
    if (hipMalloc(&K, bytes) == hipSuccess) {
        if (hipStreamSynchronize(pf->stream) == hipSuccess) rc = pf->sys.autotune(pf->d_D[0], J, K, reps, ms);
        (void)hipStreamSynchronize(pf->stream);
        (void)hipFree(K);
    }
    (void)hipFree(J);
    return rc;
}

int gtf_pfock_tile_shape(PFock_t pf, int cls, int* shape)
{
    return pf ? pf->sys.get_shape(cls, shape) : -1;
}

int gtf_pfock_set_tile_shapes(PFock_t pf, const int* shapes)
{
    if (!pf || hipStreamSynchronize(pf->stream) != hipSuccess) return -1;
    return pf->sys.set_shapes(shapes) ? -1 : 0;
}

int gtf_pfock_set_tile_shape(PFock_t pf, int cls, const int* shape)
{
    if (!pf || hipStreamSynchronize(pf->stream) != hipSuccess) return -1;
    return pf->sys.set_shape(cls, shape) ? -1 : 0;
}

}  // extern "C"
