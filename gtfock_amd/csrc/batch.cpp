// batch.cpp -- many molecules per Fock build and per SCF iteration.
//
// A batch is one multi-molecule GtfSystem: the screened quartets of every
// molecule go into the same per-class tile lists, so one launch per quartet
// class (one captured HIP graph per build) serves the whole batch.  The SCF
// driver around it is pscf/scf.c:189-474 per molecule, with the dense steps
// (DIIS, pscf/purif.c:491-641, and purification, pscf/purif.c:257-488) run
// for all molecules by the kernels of batch_linalg.hip; each molecule keeps
// its own iteration count and convergence test.
#include <sys/time.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <new>
#include <vector>

#include "../../include/gtfock_amd.h"
#include "batch_linalg.h"
#include "engine.h"
#include "oneel.h"
#include "ozaki_launch.h"
#include "screen.h"

#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#define B_MAX_DIIS 10

#define B_CHECK(call)                                                                        \
    do {                                                                                     \
        hipError_t e_ = (call);                                                              \
        if (e_ != hipSuccess) {                                                              \
            GTF_PRINTF("HIP error %s (%s)\n", hipGetErrorString(e_), #call);                 \
            return -1;                                                                       \
        }                                                                                    \
    } while (0)

static double bwtime()
{
    struct timeval tv;
    gettimeofday(&tv, nullptr);
    return tv.tv_sec + 1e-6 * tv.tv_usec;
}

struct GtfBatch {
    std::vector<const BasisSet*> bases;
    GtfSystem sys;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    int nmol = 0, maxn = 0, maxpld = 0;
    std::vector<int> n, pld, nocc, atom_off;
    std::vector<long> off, poff;
    long total = 0, ptotal = 0;
    std::vector<double> enuc;
    long *d_off = nullptr, *d_poff = nullptr;
    int *d_n = nullptr, *d_pld = nullptr, *d_nocc = nullptr;
    // Fock build (packed): D in, J / K accumulators, G = J + K = 2J(D) - K(D)
    double *d_D = nullptr, *d_J = nullptr, *d_K = nullptr, *d_G = nullptr;
    // HIP graph of one build
    bool use_graph = true;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    double last_ms = 0.0, create_s = 0.0;
    // SCF set-up: one-electron matrices (and the SCF arrays), X = S^-1/2
    double setup_oneel_s = 0.0, setup_xmat_s = 0.0;
    long nbuilds = 0;
    // SCF state
    bool scf_ready = false;
    double *d_H = nullptr, *d_S = nullptr, *d_F = nullptr;                      // packed
    double *p_X = nullptr, *p_S = nullptr, *p_F = nullptr, *p_D = nullptr;      // padded
    double *p_D2 = nullptr, *p_D3 = nullptr, *p_W = nullptr, *p_W2 = nullptr;  // padded work
    double *evecs = nullptr, *Fvecs = nullptr;                                  // B_MAX_DIIS padded sets
    double *d_stats = nullptr, *d_dots = nullptr, *d_coef = nullptr, *d_energy = nullptr;
    int *d_iters = nullptr, *d_skip = nullptr, *d_nvec = nullptr, *d_slot = nullptr, *d_skip2 = nullptr;
    // host sources of asynchronous copies (alive until the stream syncs)
    std::vector<int> h_skip, h_skip2, h_slot, h_nvec;
    std::vector<double> h_coef;
    // per-molecule DIIS bookkeeping (pscf/purif.h b_mat, len_diis, bmax, bmax_id)
    std::vector<int> len_diis, bmax_id;
    std::vector<double> bmax, b_mat;
    std::vector<int> purif_iters;
    std::vector<int> h_active;      // source of the last copy into sys.d_mol_active
    std::vector<double> done_time;  // wall clock at which each molecule's SCF ended
    double time_fock = 0, time_diis = 0, time_purif = 0;

    double *d_part = nullptr, *d_pcoef = nullptr;
    // reduced-precision purification products while ||D - D^2||_F >
    // purif_until, per molecule: mode 1 FP32 MFMA, mode 2 split BF16 with
    // purif_ns slices (ozaki.h); d_low[m] holds the mode a molecule is in
    int purif_mode = 0;
    double purif_until = 1e-3;
    int purif_ns = OZ_SLICES_DEFAULT;
    // the split products run only while a molecule's last DIIS residual
    // ||X (F D S - S D F) X||_F exceeds purif_resid (before its first: split);
    // closer to convergence its purification is FP64 throughout
    double purif_resid = OZ_SPLIT_RESID;
    std::vector<double> resid;  // per molecule, -1 before the first
    std::vector<int> h_low;     // source of the copy into d_low
    int* d_low = nullptr;
    void* oz_mem = nullptr;
    OzScratch oz;
    BatchLayout layout() const { return BatchLayout{nmol, maxn, maxpld, d_off, d_n, d_poff, d_pld, d_part, d_pcoef}; }
    int enqueue_fock();
    int fock();
    // from / from_idx (may be null): molecule m's one-electron matrices and
    // X are copied from molecule from_idx[m] of batch `from` when >= 0
    int scf_setup(const GtfBatch* from = nullptr, const int* from_idx = nullptr);
    int put_initial_guess();
    int diis(const std::vector<int>& iter, const std::vector<int>& active);
    // step-wise SCF (gtf_batch_scf_begin / _step) and its per-molecule state
    int scf_begin(int max_iters, double ediff, gtf_scf_done_fn done, void* ctx);
    int scf_step();
    int scf_finish();
    void scf_end_molecule(int m, int nit);
    double tolscr = 1e-11;
    int run_max_iters = 0;
    double run_ediff = 0.0;
    gtf_scf_done_fn run_done = nullptr;
    void* run_ctx = nullptr;
    std::vector<int> run_active, run_nit, run_next;  // run_next: index of the molecule's next build
    std::vector<double> run_e0, run_energy, run_hist;
    int purify(const std::vector<int>& active);
    void invalidate_graph()
    {
        if (gexec) { (void)hipGraphExecDestroy(gexec); gexec = nullptr; }
        if (graph) { (void)hipGraphDestroy(graph); graph = nullptr; }
    }
    ~GtfBatch()
    {
        if (stream) (void)hipStreamSynchronize(stream);
        invalidate_graph();
        sys.release();
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

// store src_m into slot slot[m] of a set of padded vectors
__global__ void bstore_slot_kernel(const long* __restrict__ poff, const int* __restrict__ pld,
                                   const int* __restrict__ skip, const int* __restrict__ slot,
                                   const double* __restrict__ src, double* __restrict__ vecs, long stride)
{
    const int m = blockIdx.y;
    if (skip[m] || slot[m] < 0 || slot[m] >= B_MAX_DIIS) return;
    const long nn = (long)pld[m] * pld[m];
    double* dst = vecs + (long)slot[m] * stride + poff[m];
    for (long k = (long)blockIdx.x * blockDim.x + threadIdx.x; k < nn; k += (long)gridDim.x * blockDim.x)
        dst[k] = src[poff[m] + k];
}

// X = U diag(lambda^-1/2) U^T into a padded block (pfock/pfock.c:1823-1861);
// W holds the eigenvectors as rows (row-major n x n)
__global__ void binv_sqrt_kernel(const double* W, const double* lam, int n, int ld, double* X)
{
    const int i = blockIdx.y * 16 + threadIdx.y, j = blockIdx.x * 16 + threadIdx.x;
    if (i >= ld || j >= ld) return;
    double s = 0.0;
    if (i < n && j < n)
        for (int k = 0; k < n; k++) s += W[(long)k * n + i] * W[(long)k * n + j] / sqrt(lam[k]);
    X[(long)i * ld + j] = s;
}

int GtfBatch::enqueue_fock()
{
    if (gtf_fill(stream, d_J, total, 0.0) || gtf_fill(stream, d_K, total, 0.0)) return -1;
    if (sys.jk(d_D, d_J, d_K)) return -1;
    if (sys.symmetrize(d_J) || sys.symmetrize(d_K)) return -1;
    return gtf_add(stream, d_J, d_K, d_G, total);
}

int GtfBatch::fock()
{
    B_CHECK(hipEventRecord(ev0, stream));
    if (!use_graph || sys.timing) {
        if (enqueue_fock()) return -1;
    } else {
        if (!gexec) {
            // the class streams and events are created by the first build
            if (!sys.cstream[0]) {
                if (enqueue_fock()) return -1;
                B_CHECK(hipStreamSynchronize(stream));
            }
            B_CHECK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
            const int rc = enqueue_fock();
            hipGraph_t g = nullptr;
            const hipError_t e = hipStreamEndCapture(stream, &g);
            if (rc || e != hipSuccess) {
                if (g) (void)hipGraphDestroy(g);
                GTF_PRINTF("graph capture of the batched build failed\n");
                return -1;
            }
            graph = g;
            B_CHECK(hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0));
        }
        B_CHECK(hipGraphLaunch(gexec, stream));
    }
    B_CHECK(hipEventRecord(ev1, stream));
    B_CHECK(hipEventSynchronize(ev1));
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, ev0, ev1);
    last_ms = ms;
    if (sys.collect_times()) return -1;
    nbuilds++;
    return 0;
}

// The starting density of pscf/scf.c:34-78 for every molecule: its atomic
// guess blocks (CInt_getInitialGuess; zero for atoms whose basis directory
// has no <Element>.dat) on the block diagonal, times R/2 with
// R = (N_neutral - Q) / N_neutral.  D must already be zero; nothing is
// copied when every block is zero.
int GtfBatch::put_initial_guess()
{
    std::vector<double> h;
    for (int m = 0; m < nmol; m++) {
        const BasisSet* bs = bases[m];
        const int n = this->n[m];
        const int nn = bs->nelec_neutral, q = bs->charge;  // CInt_getNneutral, CInt_getTotalCharge
        const double r2 = 0.5 * ((q != 0 && nn != 0) ? (double)(nn - q) / nn : 1.0);
        for (int a = 0; a < bs->natoms; a++) {
            // the block CInt_getInitialGuess returns, read in place (that
            // call may size a missing block, and several cohort threads can
            // hold one basis)
            const int s0 = bs->fstart[bs->s_startind[a]], s1 = bs->fstart[bs->s_startind[a + 1]] - 1;
            const int ld = s1 - s0 + 1;
            if (bs->Z[a] >= (int)bs->guess.size()) continue;
            const std::vector<double>& gv = bs->guess[bs->Z[a]];
            if ((int)gv.size() != ld * ld) continue;  // no guess: zero block
            const double* g = gv.data();
            for (int i = 0; i < ld; i++)
                for (int j = 0; j < ld; j++) {
                    const double v = g[i * ld + j];
                    if (v == 0.0) continue;
                    if (h.empty()) h.assign((size_t)total, 0.0);
                    h[(size_t)off[m] + (size_t)(s0 + i) * n + s0 + j] = v * r2;
                }
        }
    }
    if (h.empty()) return 0;
    B_CHECK(hipMemcpyAsync(d_D, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice, stream));
    B_CHECK(hipStreamSynchronize(stream));
    return 0;
}

int GtfBatch::scf_setup(const GtfBatch* from, const int* from_idx)
{
    if (scf_ready) return 0;
    auto carried = [&](int m) { return from && from_idx && from_idx[m] >= 0; };
    GtfSystem& s = sys;
    const double t0 = bwtime();
    d_H = s.dalloc<double>(total);
    d_S = s.dalloc<double>(total);
    d_F = s.dalloc<double>(total);
    double* T = s.dalloc<double>(total);
    double* V = s.dalloc<double>(total);
    double** pads[] = {&p_X, &p_S, &p_F, &p_D, &p_D2, &p_D3, &p_W, &p_W2};
    for (double** p : pads) *p = s.dalloc<double>(ptotal);
    evecs = s.dalloc<double>((size_t)ptotal * B_MAX_DIIS);
    Fvecs = s.dalloc<double>((size_t)ptotal * B_MAX_DIIS);
    d_stats = s.dalloc<double>(3 * nmol);
    d_dots = s.dalloc<double>((size_t)BMAX_DOTS * nmol);
    d_coef = s.dalloc<double>((size_t)BMAX_DOTS * nmol);
    d_energy = s.dalloc<double>(nmol);
    d_iters = s.dalloc<int>(nmol);
    d_skip = s.dalloc<int>(nmol);
    d_skip2 = s.dalloc<int>(nmol);
    d_nvec = s.dalloc<int>(nmol);
    d_slot = s.dalloc<int>(nmol);
    d_part = s.dalloc<double>((size_t)nmol * BPART * BMAX_DOTS);
    d_pcoef = s.dalloc<double>(2 * (size_t)nmol);
    d_low = s.dalloc<int>(nmol);
    if (!d_part || !d_pcoef || !d_low) return -2;
    if (!d_H || !d_S || !d_F || !T || !V || !evecs || !Fvecs || !d_stats || !d_dots || !d_coef || !d_energy ||
        !d_iters || !d_skip || !d_skip2 || !d_nvec || !d_slot)
        return -2;
    for (double** p : pads)
        if (!*p) return -2;
    for (double** p : pads)
        if (gtf_fill(stream, *p, ptotal, 0.0)) return -1;
    if (gtf_fill(stream, evecs, ptotal * B_MAX_DIIS, 0.0) || gtf_fill(stream, Fvecs, ptotal * B_MAX_DIIS, 0.0))
        return -1;
    // one-electron matrices of the molecules not carried over, in one
    // launch: each over its slice of the shell and atom arrays
    std::vector<OneelMol> om;
    int nblk = 0;
    for (int m = 0; m < nmol; m++) {
        if (carried(m)) continue;
        const int ns = s.mol_nshells[m];
        om.push_back(OneelMol{nblk, s.mol_shell_off[m], ns, atom_off[m], bases[m]->natoms, n[m], off[m]});
        nblk += (ns * ns + 63) / 64;
    }
    const int nom = (int)om.size();
    om.push_back(OneelMol{nblk, 0, 0, 0, 0, 0, 0});
    OneelMol* d_om = s.dalloc<OneelMol>(om.size());
    if (!d_om) return -2;
    B_CHECK(hipMemcpyAsync(d_om, om.data(), sizeof(OneelMol) * om.size(), hipMemcpyHostToDevice, stream));
    const BasisView ball{s.nshells, s.nbf, s.natoms, s.d_L, s.d_nprim, s.d_poff, s.d_fstart, s.d_atom,
                         s.d_alpha, s.d_coef, s.d_center, s.d_atom_xyz, s.d_atom_Z};
    if (gtf_oneel_batch(stream, ball, s.d_boys, d_om, nom, nblk, d_S, T, V)) return -1;
    if (gtf_add(stream, T, V, d_H, total)) return -1;
    for (int m = 0; m < nmol; m++) {
        if (!carried(m)) continue;
        const int mo = from_idx[m];
        const size_t nn = (size_t)n[m] * n[m];
        if (hipMemcpyAsync(d_H + off[m], from->d_H + from->off[mo], sizeof(double) * nn, hipMemcpyDeviceToDevice,
                           stream) != hipSuccess ||
            hipMemcpyAsync(d_S + off[m], from->d_S + from->off[mo], sizeof(double) * nn, hipMemcpyDeviceToDevice,
                           stream) != hipSuccess ||
            hipMemcpyAsync(p_X + poff[m], from->p_X + from->poff[mo], sizeof(double) * pld[m] * pld[m],
                           hipMemcpyDeviceToDevice, stream) != hipSuccess)
            return -1;
    }
    // X = S^-1/2 per molecule: molecules of one basis size share one
    // strided-batched eigensolver call, the return codes are read once
    std::vector<std::vector<int>> groups;
    {
        std::vector<int> order(nmol);
        for (int m = 0; m < nmol; m++) order[m] = m;
        std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return n[x] < n[y]; });
        int prev = -1;
        for (int k = 0; k < nmol; k++) {
            if (carried(order[k])) continue;
            if (prev < 0 || n[order[k]] != n[prev]) groups.emplace_back();
            groups.back().push_back(order[k]);
            prev = order[k];
        }
    }
    B_CHECK(hipStreamSynchronize(stream));  // om is read by the copy above
    s.dfree(d_om);
    const double t1 = bwtime();
    size_t wmax = 1;
    for (const auto& gr : groups) wmax = std::max(wmax, gr.size() * (size_t)n[gr[0]] * n[gr[0]]);
    double* W = s.dalloc<double>(wmax);
    double* lam = s.dalloc<double>((size_t)nmol * maxn);
    double* work = s.dalloc<double>((size_t)nmol * maxn);
    int* info = s.dalloc<int>(nmol);
    if (!W || !lam || !work || !info) return -2;
    rocblas_handle h;
    if (rocblas_create_handle(&h) != rocblas_status_success) return -1;
    rocblas_set_stream(h, stream);
    int bad = 0, done = 0;
    for (const auto& gr : groups) {
        const int nm = n[gr[0]], cnt = (int)gr.size();
        const size_t nn = (size_t)nm * nm;
        // the group's results go to lam / info slots done .. done + cnt - 1
        for (int k = 0; k < cnt && !bad; k++)
            if (hipMemcpyAsync(W + k * nn, d_S + off[gr[k]], sizeof(double) * nn, hipMemcpyDeviceToDevice, stream) !=
                hipSuccess)
                bad = 1;
        if (bad) break;
        if (rocsolver_dsyevd_strided_batched(h, rocblas_evect_original, rocblas_fill_upper, nm, W, nm, (rocblas_stride)nn,
                                             lam + (size_t)done * maxn, maxn, work + (size_t)done * maxn, maxn,
                                             info + done, cnt) != rocblas_status_success) {
            bad = 2;
            break;
        }
        for (int k = 0; k < cnt; k++) {
            const int m = gr[k];
            dim3 g((pld[m] + 15) / 16, (pld[m] + 15) / 16), b(16, 16);
            binv_sqrt_kernel<<<g, b, 0, stream>>>(W + k * nn, lam + (size_t)(done + k) * maxn, nm, pld[m],
                                                  p_X + poff[m]);
            if (hipGetLastError() != hipSuccess) bad = 4;
        }
        done += cnt;
    }
    std::vector<int> hinfo(nmol, 0);
    if (!bad && done > 0 &&
        (hipMemcpyAsync(hinfo.data(), info, sizeof(int) * done, hipMemcpyDeviceToHost, stream) != hipSuccess ||
         hipStreamSynchronize(stream) != hipSuccess))
        bad = 3;
    for (int k = 0; k < done && !bad; k++)
        if (hinfo[k] != 0) bad = 3;
    rocblas_destroy_handle(h);
    if (bad) {
        GTF_PRINTF("S^-1/2 failed (%d)\n", bad);
        return -1;
    }
    const BatchLayout L = layout();
    if (gtf_bpad(stream, L, nullptr, d_S, p_S)) return -1;
    B_CHECK(hipStreamSynchronize(stream));
    s.dfree(T);
    s.dfree(V);
    s.dfree(W);
    s.dfree(lam);
    s.dfree(work);
    s.dfree(info);
    setup_oneel_s = t1 - t0;
    setup_xmat_s = bwtime() - t1;
    scf_ready = true;
    return 0;
}

static void binverse(int m, const double* A, int lda, double* inv)
{
    double a[(B_MAX_DIIS + 1) * (B_MAX_DIIS + 1)];
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

// RHF SCF of every molecule (pscf/scf.c:189-474), one iteration per step:
// scf_begin resets the DIIS state and puts scf.c's initial density; each
// scf_step builds F = H + G(D) for the molecules still running (the others'
// quartets are skipped), takes their energies, ends those that converged
// (|E_k - E_(k-1)| < ediff), diverged or ran out of iterations -- reporting
// each through the callback at once -- and runs DIIS and purification for
// the rest.  Every quantity is per molecule, so molecules may also join a
// running SCF (gtf_batch_regroup) with their own iteration counts.
int GtfBatch::scf_begin(int max_iters, double ediff, gtf_scf_done_fn done, void* ctx)
{
    if (scf_setup()) return -1;
    const int LB = B_MAX_DIIS + 1;
    len_diis.assign(nmol, 0);
    bmax_id.assign(nmol, -1);
    bmax.assign(nmol, DBL_MIN);
    resid.assign(nmol, -1.0);
    b_mat.assign((size_t)nmol * LB * LB, 0.0);
    for (int m = 0; m < nmol; m++)
        for (int i = 0; i <= B_MAX_DIIS; i++)
            for (int j = 0; j <= B_MAX_DIIS; j++) b_mat[(size_t)m * LB * LB + i * LB + j] = (i != j) ? -1.0 : 0.0;
    time_fock = time_diis = time_purif = 0.0;
    if (gtf_fill(stream, d_D, total, 0.0)) return -1;
    if (put_initial_guess()) return -1;
    run_max_iters = max_iters;
    run_ediff = ediff;
    run_done = done;
    run_ctx = ctx;
    run_active.assign(nmol, 1);
    run_nit.assign(nmol, max_iters);
    run_next.assign(nmol, 0);
    run_e0.assign(nmol, -1.0);
    run_energy.assign(nmol, 0.0);
    run_hist.assign((size_t)nmol * max_iters, NAN);
    B_CHECK(hipStreamSynchronize(stream));
    return 0;
}

void GtfBatch::scf_end_molecule(int m, int nit)
{
    run_active[m] = 0;
    run_nit[m] = nit;
    done_time[m] = bwtime();
    if (run_done) run_done(run_ctx, m, run_energy[m], nit);
}

int GtfBatch::scf_step()
{
    int nactive = 0;
    for (int m = 0; m < nmol; m++) nactive += run_active[m];
    if (nactive == 0) return 0;
    const BatchLayout L = layout();
    double t0 = bwtime();
    // molecules whose SCF has ended drop out of the build (their tiles exit
    // at once; the captured graph stays valid)
    h_active = run_active;
    B_CHECK(hipMemcpyAsync(sys.d_mol_active, h_active.data(), sizeof(int) * nmol, hipMemcpyHostToDevice, stream));
    if (fock()) return -1;
    if (gtf_bfock_energy(stream, L, d_H, d_G, d_D, d_F, d_energy)) return -1;
    std::vector<double> e(nmol);
    B_CHECK(hipMemcpyAsync(e.data(), d_energy, sizeof(double) * nmol, hipMemcpyDeviceToHost, stream));
    B_CHECK(hipStreamSynchronize(stream));
    time_fock += bwtime() - t0;
    std::vector<int> iter(nmol, 0);
    nactive = 0;
    for (int m = 0; m < nmol; m++) {
        if (!run_active[m]) continue;
        const int it = run_next[m]++;
        iter[m] = it;
        run_energy[m] = e[m] + enuc[m];
        run_hist[(size_t)m * run_max_iters + it] = run_energy[m];
        if (!std::isfinite(e[m])) {
            scf_end_molecule(m, -(it + 1));  // diverged: builds < 0
        } else if (it > 0 && fabs(e[m] - run_e0[m]) < run_ediff) {
            scf_end_molecule(m, it + 1);
        } else if (it == run_max_iters - 1) {
            scf_end_molecule(m, run_max_iters);  // out of iterations
        } else {
            run_e0[m] = e[m];
            nactive++;
        }
    }
    if (nactive == 0) return 0;
    h_skip.assign(nmol, 0);
    for (int m = 0; m < nmol; m++) h_skip[m] = !run_active[m];
    B_CHECK(hipMemcpyAsync(d_skip, h_skip.data(), sizeof(int) * nmol, hipMemcpyHostToDevice, stream));
    t0 = bwtime();
    if (diis(iter, run_active)) return -1;
    B_CHECK(hipStreamSynchronize(stream));
    const double t1 = bwtime();
    time_diis += t1 - t0;
    if (purify(run_active)) return -1;
    B_CHECK(hipStreamSynchronize(stream));
    time_purif += bwtime() - t1;
    return nactive;
}

// every molecule's quartets in later plain builds (gtf_batch_compute_fock)
int GtfBatch::scf_finish()
{
    h_active.assign(nmol, 1);
    B_CHECK(hipMemcpyAsync(sys.d_mol_active, h_active.data(), sizeof(int) * nmol, hipMemcpyHostToDevice, stream));
    B_CHECK(hipStreamSynchronize(stream));
    return 0;
}

// compute_diis (pscf/purif.c:491-641) for the active molecules: F (packed,
// AO) and D (packed) in, the orthogonal-basis extrapolated F in p_F.
// iter[m] is molecule m's own iteration index (molecules that joined a
// running batch later count from their own start): the error vector and
// extrapolation apply from iter[m] > 1 on, as scf.c's compute_diis.
int GtfBatch::diis(const std::vector<int>& iter, const std::vector<int>& active)
{
    const BatchLayout L = layout();
    const int LB = B_MAX_DIIS + 1;
    const int* skip = d_skip;
    if (gtf_bpad(stream, L, skip, d_F, p_W2) || gtf_bpad(stream, L, skip, d_D, p_D)) return -1;
    std::vector<int> slot(nmol, -1), nvec(nmol, 0), skip_x(nmol, 1);
    bool any = false;
    for (int m = 0; m < nmol; m++)
        if (active[m] && iter[m] > 1) {
            any = true;
            skip_x[m] = 0;
            slot[m] = len_diis[m] < B_MAX_DIIS ? len_diis[m]++ : bmax_id[m];
            if (slot[m] < 0 || slot[m] >= B_MAX_DIIS) {
                // only a non-finite B matrix leaves no largest error; such a
                // molecule has already been stopped by the energy check
                GTF_PRINTF("molecule %d: no DIIS slot (bmax_id %d)\n", m, bmax_id[m]);
                return -1;
            }
            nvec[m] = len_diis[m];
        }
    // molecules without extrapolation this iteration: skipped by the slot
    // stores and the linear combination
    const int* skip2 = d_skip2;
    if (any) {
        h_skip2 = skip_x;
        h_slot = slot;
        h_nvec = nvec;
        B_CHECK(hipMemcpyAsync(d_skip2, h_skip2.data(), sizeof(int) * nmol, hipMemcpyHostToDevice, stream));
        B_CHECK(hipMemcpyAsync(d_slot, h_slot.data(), sizeof(int) * nmol, hipMemcpyHostToDevice, stream));
        B_CHECK(hipMemcpyAsync(d_nvec, h_nvec.data(), sizeof(int) * nmol, hipMemcpyHostToDevice, stream));
        // e = X (F D S - S D F) X with the A * B^T convention of pdgemm3D_2
        if (gtf_bgemm(stream, L, skip2, nullptr, p_W2, p_D, p_W, 1) ||
            gtf_bgemm(stream, L, skip2, nullptr, p_W, p_S, p_D2, 1) || gtf_bantisym(stream, L, skip2, p_D2, p_D3) ||
            gtf_bgemm(stream, L, skip2, nullptr, p_X, p_D3, p_W, 1) ||
            gtf_bgemm(stream, L, skip2, nullptr, p_W, p_X, p_D2, 1))
            return -1;
        bstore_slot_kernel<<<dim3(256, nmol), 256, 0, stream>>>(d_poff, d_pld, skip2, d_slot, p_D2, evecs, ptotal);
        B_CHECK(hipGetLastError());
        if (gtf_bmultidot(stream, L, skip2, d_nvec, p_D2, evecs, ptotal, d_dots)) return -1;
        std::vector<double> dots((size_t)BMAX_DOTS * nmol);
        B_CHECK(hipMemcpyAsync(dots.data(), d_dots, sizeof(double) * dots.size(), hipMemcpyDeviceToHost, stream));
        B_CHECK(hipStreamSynchronize(stream));
        for (int m = 0; m < nmol; m++) {
            if (skip_x[m]) continue;
            double* bm = &b_mat[(size_t)m * LB * LB];
            const int cur = slot[m];
            for (int i = 0; i < len_diis[m]; i++) {
                bm[cur * LB + i] = dots[(size_t)m * BMAX_DOTS + i];
                bm[i * LB + cur] = dots[(size_t)m * BMAX_DOTS + i];
            }
            bmax[m] = -DBL_MAX;
            for (int i = 0; i < len_diis[m]; i++)
                if (bmax[m] < bm[i * LB + i]) {
                    bmax[m] = bm[i * LB + i];
                    bmax_id[m] = i;
                }
            resid[m] = std::sqrt(std::max(bm[cur * LB + cur], 0.0));
        }
    }
    // F_orth = X F X
    if (gtf_bgemm(stream, L, skip, nullptr, p_X, p_W2, p_W, 1) ||
        gtf_bgemm(stream, L, skip, nullptr, p_W, p_X, p_F, 1))
        return -1;
    if (any) {
        bstore_slot_kernel<<<dim3(256, nmol), 256, 0, stream>>>(d_poff, d_pld, skip2, d_slot, p_F, Fvecs, ptotal);
        B_CHECK(hipGetLastError());
        h_coef.assign((size_t)BMAX_DOTS * nmol, 0.0);
        for (int m = 0; m < nmol; m++) {
            if (skip_x[m]) continue;
            const int sz = len_diis[m] + 1;
            double inv[(B_MAX_DIIS + 1) * (B_MAX_DIIS + 1)];
            binverse(sz, &b_mat[(size_t)m * LB * LB], LB, inv);
            for (int i = 0; i < len_diis[m]; i++) h_coef[(size_t)m * BMAX_DOTS + i] = -inv[i * sz + sz - 1];
        }
        B_CHECK(hipMemcpyAsync(d_coef, h_coef.data(), sizeof(double) * h_coef.size(), hipMemcpyHostToDevice, stream));
        if (gtf_blincomb(stream, L, skip2, d_nvec, d_coef, Fvecs, ptotal, p_F)) return -1;
    }
    return 0;
}

// compute_purification (pscf/purif.c:257-488) for the active molecules:
// p_F (orthogonal basis) in, the AO density in d_D (packed)
int GtfBatch::purify(const std::vector<int>& active)
{
    const BatchLayout L = layout();
    const int* skip = d_skip;
    if (gtf_bpurif_init(stream, L, d_nocc, skip, p_F, p_D)) return -1;
    if (gtf_fill_int(stream, d_iters, nmol, -1)) return -1;
    h_low.assign(nmol, purif_mode);
    if (purif_mode == GTF_LOW_SPLIT)
        for (int m = 0; m < nmol; m++)
            if (resid[m] >= 0.0 && resid[m] < purif_resid) h_low[m] = GTF_LOW_FP64;
    B_CHECK(hipMemcpyAsync(d_low, h_low.data(), sizeof(int) * nmol, hipMemcpyHostToDevice, stream));
    if (purif_mode == GTF_LOW_SPLIT && !oz_mem) {
        oz_mem = sys.dalloc<char>(gtf_oz_scratch_bytes(ptotal, OZ_SLICES_MAX));
        if (!oz_mem) return -2;
        gtf_oz_scratch_carve(oz_mem, ptotal, purif_ns, oz);
    }
    OzOperands o;
    o.nmol = nmol;
    o.maxld = maxpld;
    o.poff = d_poff;
    o.pld = d_pld;
    o.skip = skip;
    o.done = d_iters;
    o.sel = d_low;
    std::vector<int> it_h(nmol, -1);
    int it = 0;
    for (it = 0; it < 200; it++) {
        // molecules already converged skip their GEMMs through d_iters >= 0;
        // those in the split-BF16 phase take the ozaki.h products instead
        if (gtf_bgemm(stream, L, skip, d_iters, p_D, p_D, p_D2, 0, d_low) ||
            gtf_bgemm(stream, L, skip, d_iters, p_D, p_D2, p_D3, 0, d_low))
            return -1;
        if (purif_mode == GTF_LOW_SPLIT && gtf_oz_square_cube(stream, o, oz, p_D, p_D2, p_D3)) return -1;
        if (gtf_bpurif_stats(stream, L, skip, p_D, p_D2, p_D3, d_iters, d_stats, it, 1e-11, d_low, purif_until))
            return -1;
        if (gtf_bpurif_update(stream, L, skip, d_iters, d_stats, p_D, p_D2, p_D3, it)) return -1;
        if (it % 2 == 1) {
            B_CHECK(hipMemcpyAsync(it_h.data(), d_iters, sizeof(int) * nmol, hipMemcpyDeviceToHost, stream));
            B_CHECK(hipStreamSynchronize(stream));
            bool all = true;
            for (int m = 0; m < nmol; m++)
                if (active[m] && it_h[m] < 0) all = false;
            if (all) break;
        }
    }
    B_CHECK(hipMemcpyAsync(it_h.data(), d_iters, sizeof(int) * nmol, hipMemcpyDeviceToHost, stream));
    B_CHECK(hipStreamSynchronize(stream));
    for (int m = 0; m < nmol; m++)
        if (active[m]) purif_iters[m] = it_h[m] < 0 ? 200 : it_h[m];
    // D_AO = X D X
    if (gtf_bgemm(stream, L, skip, nullptr, p_X, p_D, p_W, 1) ||
        gtf_bgemm(stream, L, skip, nullptr, p_W, p_X, p_D2, 1))
        return -1;
    return gtf_bunpad(stream, L, skip, p_D2, d_D);
}

extern "C" {

int gtf_batch_create(BasisSet_t* basis, int nmol, double tolscr, gtf_batch_t* out)
{
    if (!basis || nmol <= 0 || !out || tolscr < 0.0) return -3;
    *out = nullptr;
    const double t0 = bwtime();
    GtfBatch* b = new (std::nothrow) GtfBatch();
    if (!b) return -2;
    b->nmol = nmol;
    int natoms = 0;
    for (int m = 0; m < nmol; m++) {
        if (!basis[m] || basis[m]->nshells == 0) { delete b; return -3; }
        b->bases.push_back(basis[m]);
        const int nm = basis[m]->nbf, pl = ((nm + 63) / 64) * 64;
        b->n.push_back(nm);
        b->pld.push_back(pl);
        b->nocc.push_back(CInt_getNumOccOrb(basis[m]));
        b->atom_off.push_back(natoms);
        natoms += basis[m]->natoms;
        b->off.push_back(b->total);
        b->poff.push_back(b->ptotal);
        b->total += (long)nm * nm;
        b->ptotal += (long)pl * pl;
        b->maxn = std::max(b->maxn, nm);
        b->maxpld = std::max(b->maxpld, pl);
        b->enuc.push_back(CInt_getNucEnergy(basis[m]));
    }
    b->len_diis.assign(nmol, 0);
    b->bmax_id.assign(nmol, -1);
    b->bmax.assign(nmol, DBL_MIN);
    b->resid.assign(nmol, -1.0);
    b->purif_iters.assign(nmol, 0);
    b->done_time.assign(nmol, 0.0);
    if (hipStreamCreate(&b->stream) != hipSuccess || hipEventCreate(&b->ev0) != hipSuccess ||
        hipEventCreate(&b->ev1) != hipSuccess) {
        delete b;
        return -1;
    }
    int rc = b->sys.init(b->bases, b->stream);
    if (rc == 0) rc = b->sys.screen(tolscr);
    b->tolscr = tolscr;
    if (rc) { delete b; return rc; }
    GtfSystem& s = b->sys;
    b->d_off = s.dalloc<long>(nmol);
    b->d_poff = s.dalloc<long>(nmol);
    b->d_n = s.dalloc<int>(nmol);
    b->d_pld = s.dalloc<int>(nmol);
    b->d_nocc = s.dalloc<int>(nmol);
    b->d_D = s.dalloc<double>(b->total);
    b->d_J = s.dalloc<double>(b->total);
    b->d_K = s.dalloc<double>(b->total);
    b->d_G = s.dalloc<double>(b->total);
    if (!b->d_off || !b->d_poff || !b->d_n || !b->d_pld || !b->d_nocc || !b->d_D || !b->d_J || !b->d_K || !b->d_G) {
        delete b;
        return -2;
    }
    if (hipMemcpyAsync(b->d_off, b->off.data(), sizeof(long) * nmol, hipMemcpyHostToDevice, b->stream) != hipSuccess ||
        hipMemcpyAsync(b->d_poff, b->poff.data(), sizeof(long) * nmol, hipMemcpyHostToDevice, b->stream) !=
            hipSuccess ||
        hipMemcpyAsync(b->d_n, b->n.data(), sizeof(int) * nmol, hipMemcpyHostToDevice, b->stream) != hipSuccess ||
        hipMemcpyAsync(b->d_pld, b->pld.data(), sizeof(int) * nmol, hipMemcpyHostToDevice, b->stream) != hipSuccess ||
        hipMemcpyAsync(b->d_nocc, b->nocc.data(), sizeof(int) * nmol, hipMemcpyHostToDevice, b->stream) !=
            hipSuccess ||
        gtf_fill(b->stream, b->d_D, b->total, 0.0) || hipStreamSynchronize(b->stream) != hipSuccess) {
        delete b;
        return -1;
    }
    b->create_s = bwtime() - t0;
    *out = b;
    return 0;
}

void gtf_batch_destroy(gtf_batch_t b) { delete b; }

int gtf_batch_layout(gtf_batch_t b, int* nbf, long* offset)
{
    if (!b) return -3;
    for (int m = 0; m < b->nmol; m++) {
        if (nbf) nbf[m] = b->n[m];
        if (offset) offset[m] = b->off[m];
    }
    return b->nmol;
}

long gtf_batch_total(gtf_batch_t b) { return b ? b->total : -1; }

int gtf_batch_set_precision(gtf_batch_t b, double tau)
{
    if (!b || tau < 0.0) return -3;
    b->sys.tau = tau;
    b->invalidate_graph();
    return b->sys.build_tiles();
}

int gtf_batch_set_prim_tol(gtf_batch_t b, double tol)
{
    if (!b || !(tol >= 0.0)) return -3;
    b->sys.prim_tol = tol;
    b->invalidate_graph();
    return 0;
}

int gtf_batch_autotune(gtf_batch_t b, int reps, double* ms)
{
    if (!b || reps <= 0) return -3;
    const size_t bytes = sizeof(double) * (size_t)b->total;
    double *J = nullptr, *K = nullptr;
    if (hipMalloc(&J, bytes) != hipSuccess) return -2;
    int rc = -2;
    if (hipMalloc(&K, bytes) == hipSuccess) {
        rc = -1;
        if (hipStreamSynchronize(b->stream) == hipSuccess) {
            // the captured build holds the old tile arrays
            b->invalidate_graph();
            rc = b->sys.autotune(b->d_D, J, K, reps, ms) ? -1 : 0;
        }
        (void)hipStreamSynchronize(b->stream);
        (void)hipFree(K);
    }
    (void)hipFree(J);
    return rc;
}

int gtf_batch_tile_shape(gtf_batch_t b, int cls, int* shape)
{
    if (!b) return -3;
    return b->sys.get_shape(cls, shape) < 0 ? -3 : GTF_TILE_SHAPE_FIELDS;
}

int gtf_batch_set_tile_shapes(gtf_batch_t b, const int* shapes)
{
    if (!b) return -3;
    if (hipStreamSynchronize(b->stream) != hipSuccess) return -1;
    b->invalidate_graph();
    return b->sys.set_shapes(shapes);
}

int gtf_batch_set_tile_shape(gtf_batch_t b, int cls, const int* shape)
{
    if (!b) return -3;
    if (hipStreamSynchronize(b->stream) != hipSuccess) return -1;
    b->invalidate_graph();
    return b->sys.set_shape(cls, shape);
}

int gtf_batch_set_purif_precision(gtf_batch_t b, int mode, double until)
{
    if (!b || until <= 0.0 || mode < 0 || mode > GTF_LOW_SPLIT) return -3;
    b->purif_mode = mode;
    b->purif_until = until;
    return 0;
}

int gtf_batch_set_purif_residual(gtf_batch_t b, double resid)
{
    if (!b || !(resid >= 0.0)) return -3;
    b->purif_resid = resid;
    return 0;
}

int gtf_batch_set_purif_slices(gtf_batch_t b, int ns)
{
    if (!b || ns < OZ_SLICES_MIN || ns > OZ_SLICES_MAX) return -3;
    b->purif_ns = ns;
    if (b->oz_mem) gtf_oz_scratch_carve(b->oz_mem, b->ptotal, ns, b->oz);
    return 0;
}

int gtf_batch_use_graph(gtf_batch_t b, int on)
{
    if (!b) return -3;
    b->use_graph = on != 0;
    return 0;
}

int gtf_batch_put_density(gtf_batch_t b, int m, const double* D)
{
    if (!b || !D || m >= b->nmol) return -3;
    const long o = m < 0 ? 0 : b->off[m];
    const long cnt = m < 0 ? b->total : (long)b->n[m] * b->n[m];
    if (hipMemcpyAsync(b->d_D + o, D, sizeof(double) * cnt, hipMemcpyHostToDevice, b->stream) != hipSuccess ||
        hipStreamSynchronize(b->stream) != hipSuccess)
        return -1;
    return 0;
}

int gtf_batch_compute_fock(gtf_batch_t b)
{
    if (!b) return -3;
    return b->fock();
}

int gtf_batch_get_fock(gtf_batch_t b, int m, double* G)
{
    if (!b || !G || m >= b->nmol) return -3;
    const long o = m < 0 ? 0 : b->off[m];
    const long cnt = m < 0 ? b->total : (long)b->n[m] * b->n[m];
    if (hipMemcpyAsync(G, b->d_G + o, sizeof(double) * cnt, hipMemcpyDeviceToHost, b->stream) != hipSuccess ||
        hipStreamSynchronize(b->stream) != hipSuccess)
        return -1;
    return 0;
}

int gtf_batch_device_ptrs(gtf_batch_t b, double** D, double** G)
{
    if (!b) return -3;
    if (D) *D = b->d_D;
    if (G) *G = b->d_G;
    return 0;
}

void* gtf_batch_stream(gtf_batch_t b) { return b ? (void*)b->stream : nullptr; }

int gtf_batch_stats(gtf_batch_t b, double* out)
{
    if (!b || !out) return -3;
    double nq = 0, npq = 0, nq32 = 0, npq32 = 0;
    for (int c = 0; c < GTF_NCLASSES; c++) {
        nq += (double)b->sys.nq[c];
        npq += b->sys.nprimq[c];
        nq32 += (double)b->sys.nq_lo[c];
        npq32 += b->sys.nprimq_lo[c];
    }
    out[0] = nq;
    out[1] = npq;
    out[2] = nq32;
    out[3] = npq32;
    out[4] = b->last_ms;
    out[5] = b->gexec ? 1.0 : 0.0;
    out[6] = (double)b->sys.dev_bytes;
    out[7] = b->create_s;
    out[8] = b->time_fock;
    out[9] = b->time_diis;
    out[10] = b->time_purif;
    return 11;
}

int gtf_batch_setup_times(gtf_batch_t b, double* out)
{
    if (!b || !out) return -3;
    for (int k = 0; k < GtfSystem::NSETUP; k++) out[k] = b->sys.setup_s[k];
    out[GtfSystem::NSETUP] = b->setup_oneel_s;
    out[GtfSystem::NSETUP + 1] = b->setup_xmat_s;
    return GtfSystem::NSETUP + 2;
}

int gtf_batch_class_work(gtf_batch_t b, long* nquartets, double* nprimquartets, long* nq_fp32, double* npq_fp32)
{
    if (!b) return -3;
    for (int c = 0; c < GTF_NCLASSES; c++) {
        if (nquartets) nquartets[c] = b->sys.nq[c];
        if (nprimquartets) nprimquartets[c] = b->sys.nprimq[c];
        if (nq_fp32) nq_fp32[c] = b->sys.nq_lo[c];
        if (npq_fp32) npq_fp32[c] = b->sys.nprimq_lo[c];
    }
    return GTF_NCLASSES;
}

int gtf_batch_scf(gtf_batch_t b, int max_iters, double ediff, double* energy, int* iters, double* history)
{
    return gtf_batch_scf_cb(b, max_iters, ediff, energy, iters, history, nullptr, nullptr);
}

int gtf_batch_scf_cb(gtf_batch_t b, int max_iters, double ediff, double* energy, int* iters, double* history,
                     gtf_scf_done_fn done, void* ctx)
{
    if (!b || max_iters <= 0 || !energy) return -3;
    int rc = b->scf_begin(max_iters, ediff, done, ctx);
    while (rc == 0) {
        const int left = b->scf_step();
        if (left < 0) rc = left;
        if (left <= 0) break;
    }
    if (rc == 0) rc = b->scf_finish();
    if (rc) return rc;
    return gtf_batch_scf_results(b, energy, iters, history) < 0 ? -1 : 0;
}

int gtf_batch_scf_begin(gtf_batch_t b, int max_iters, double ediff, gtf_scf_done_fn done, void* ctx)
{
    if (!b || max_iters <= 0) return -3;
    return b->scf_begin(max_iters, ediff, done, ctx);
}

int gtf_batch_scf_step(gtf_batch_t b)
{
    if (!b || b->run_max_iters <= 0) return -3;
    const int left = b->scf_step();
    if (left == 0 && b->scf_finish()) return -1;
    return left;
}

int gtf_batch_scf_results(gtf_batch_t b, double* energy, int* iters, double* history)
{
    if (!b || b->run_max_iters <= 0) return -3;
    const int nmol = b->nmol, mi = b->run_max_iters;
    for (int m = 0; m < nmol; m++) {
        if (energy) energy[m] = b->run_energy[m];
        if (iters) iters[m] = b->run_active[m] ? 0 : b->run_nit[m];
        if (history)
            for (int k = 0; k < mi; k++) history[(size_t)m * mi + k] = b->run_hist[(size_t)m * mi + k];
    }
    return nmol;
}

int gtf_batch_regroup(gtf_batch_t b, BasisSet_t* add, int nadd, gtf_batch_t* out, int* origin)
{
    if (!b || !out || nadd < 0 || (nadd > 0 && !add) || b->run_max_iters <= 0) return -3;
    *out = nullptr;
    std::vector<int> keep;
    for (int m = 0; m < b->nmol; m++)
        if (b->run_active[m]) keep.push_back(m);
    const int nkeep = (int)keep.size(), nnew = nkeep + nadd;
    if (nnew == 0) return -3;
    std::vector<BasisSet_t> bases;
    for (int m : keep) bases.push_back(const_cast<BasisSet_t>(b->bases[m]));
    for (int j = 0; j < nadd; j++) bases.push_back(add[j]);
    GtfBatch* nb = nullptr;
    int rc = gtf_batch_create(bases.data(), nnew, b->tolscr, &nb);
    if (rc) return rc;
    // the settings of the running batch: tile shapes, precision, graph use
    std::vector<int> shapes((size_t)GTF_NCLASSES * GTF_TILE_SHAPE_FIELDS);
    for (int c = 0; c < GTF_NCLASSES && rc == 0; c++)
        if (b->sys.get_shape(c, &shapes[(size_t)c * GTF_TILE_SHAPE_FIELDS]) < 0) rc = -1;
    nb->sys.tau = b->sys.tau;  // takes effect with the shapes' tile rebuild
    if (rc == 0) rc = nb->sys.set_shapes(shapes.data());
    nb->use_graph = b->use_graph;
    nb->purif_mode = b->purif_mode;
    nb->purif_until = b->purif_until;
    nb->purif_ns = b->purif_ns;
    nb->purif_resid = b->purif_resid;
    // the carried molecules' one-electron matrices and X come from the old batch
    std::vector<int> from(nnew, -1);
    for (int k = 0; k < nkeep; k++) from[k] = keep[k];
    if (rc == 0 && b->scf_ready) rc = nb->scf_setup(b, from.data());
    if (rc == 0) rc = nb->scf_begin(b->run_max_iters, b->run_ediff, b->run_done, b->run_ctx);
    if (rc) {
        delete nb;
        return rc < 0 ? rc : -1;
    }
    // carry the running molecules' SCF state over exactly
    const int LB = B_MAX_DIIS + 1, mi = b->run_max_iters;
    for (int k = 0; k < nkeep; k++) {
        const int mo = keep[k];
        const size_t nn = (size_t)b->n[mo] * b->n[mo], pp = (size_t)b->pld[mo] * b->pld[mo];
        bool ok = hipMemcpyAsync(nb->d_D + nb->off[k], b->d_D + b->off[mo], sizeof(double) * nn,
                                 hipMemcpyDeviceToDevice, nb->stream) == hipSuccess;
        for (int sl = 0; sl < B_MAX_DIIS && ok; sl++) {
            ok = hipMemcpyAsync(nb->evecs + (size_t)sl * nb->ptotal + nb->poff[k],
                                b->evecs + (size_t)sl * b->ptotal + b->poff[mo], sizeof(double) * pp,
                                hipMemcpyDeviceToDevice, nb->stream) == hipSuccess &&
                 hipMemcpyAsync(nb->Fvecs + (size_t)sl * nb->ptotal + nb->poff[k],
                                b->Fvecs + (size_t)sl * b->ptotal + b->poff[mo], sizeof(double) * pp,
                                hipMemcpyDeviceToDevice, nb->stream) == hipSuccess;
        }
        if (!ok) {
            delete nb;
            return -1;
        }
        nb->len_diis[k] = b->len_diis[mo];
        nb->bmax_id[k] = b->bmax_id[mo];
        nb->bmax[k] = b->bmax[mo];
        nb->resid[k] = b->resid[mo];
        std::copy(b->b_mat.begin() + (size_t)mo * LB * LB, b->b_mat.begin() + (size_t)(mo + 1) * LB * LB,
                  nb->b_mat.begin() + (size_t)k * LB * LB);
        nb->run_next[k] = b->run_next[mo];
        nb->run_e0[k] = b->run_e0[mo];
        nb->run_energy[k] = b->run_energy[mo];
        nb->run_nit[k] = b->run_nit[mo];
        std::copy(b->run_hist.begin() + (size_t)mo * mi, b->run_hist.begin() + (size_t)(mo + 1) * mi,
                  nb->run_hist.begin() + (size_t)k * mi);
        nb->purif_iters[k] = b->purif_iters[mo];
    }
    if (hipStreamSynchronize(nb->stream) != hipSuccess) {
        delete nb;
        return -1;
    }
    if (origin) {
        for (int k = 0; k < nkeep; k++) origin[k] = keep[k];
        for (int j = 0; j < nadd; j++) origin[nkeep + j] = -(j + 1);
    }
    *out = nb;
    return 0;
}

int gtf_batch_done_times(gtf_batch_t b, double* t)
{
    if (!b || !t) return -3;
    for (int m = 0; m < b->nmol; m++) t[m] = b->done_time[m];
    return b->nmol;
}

// Device bytes of one molecule in a batch: shell-pair data (primitive pairs,
// full and compact Hermite coefficients), screening lists, the packed and
// padded SCF matrices and the DIIS history.
double gtf_batch_job_bytes(BasisSet_t basis)
{
    if (!basis) return -1.0;
    const BasisSet* b = basis;
    double pairs = 0.0;
    for (int s = 0; s < b->nshells; s++)
        for (int t = 0; t <= s; t++) {
            int la = b->L[s], lb = b->L[t];
            if (lb > la) std::swap(la, lb);
            const double np = (double)b->nprim[s] * b->nprim[t];
            // per primitive pair: (p, P), its Schwarz bound, full and compact E
            pairs += np * (40.0 + 8.0 * (nsph(la) * nsph(lb) * nherm(la + lb) + herm_nnz_lookup(la, lb))) + 96.0;
        }
    const double n = b->nbf, pl = ((b->nbf + 63) / 64) * 64;
    // + the split purification's BF16 slices of both operands at the most
    // slices (and their exponents and row maxima)
    const double split = 2.0 * 2.0 * OZ_SLICES_MAX * pl * pl + 4.0 * 8.0 * pl * pl / 64.0;
    return pairs * 1.6 + 8.0 * (9.0 * n * n + (8.0 + 2.0 * B_MAX_DIIS) * pl * pl) + split + (1 << 20);
}

int gtf_batch_get_density(gtf_batch_t b, int m, double* D)
{
    if (!b || !D || m >= b->nmol) return -3;
    const long o = m < 0 ? 0 : b->off[m];
    const long cnt = m < 0 ? b->total : (long)b->n[m] * b->n[m];
    if (hipMemcpyAsync(D, b->d_D + o, sizeof(double) * cnt, hipMemcpyDeviceToHost, b->stream) != hipSuccess ||
        hipStreamSynchronize(b->stream) != hipSuccess)
        return -1;
    return 0;
}

}  // extern "C"
