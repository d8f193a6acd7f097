// engine.cpp -- GtfSystem: device basis, shell pairs, Schwarz screening,
// quartet compaction and the per-class ERI launches.
#include "engine.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <vector>

#include "screen.h"

void gtf_class_types(int cls, int& A, int& B)
{
    A = 0;
    while ((A + 1) * (A + 2) / 2 <= cls) ++A;
    B = cls - A * (A + 1) / 2;
}

#define SYS_CHECK(call)                                                                      \
    do {                                                                                     \
        hipError_t e_ = (call);                                                              \
        if (e_ != hipSuccess) {                                                              \
            fprintf(stderr, "gtfock_amd: HIP error %s at %s:%d\n", hipGetErrorString(e_), __FILE__, \
                    __LINE__);                                                               \
            return -1;                                                                       \
        }                                                                                    \
    } while (0)

template <typename T>
static int upload(T* d, const std::vector<T>& h, hipStream_t st)
{
    if (h.empty()) return 0;
    SYS_CHECK(hipMemcpyAsync(d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, st));
    return 0;
}

int GtfSystem::init(const BasisSet* b, hipStream_t st)
{
    stream = st;
    nbf = b->nbf;
    nshells = b->nshells;
    natoms = b->natoms;
    std::vector<double> center(3 * nshells), atomZ(natoms);
    for (int s = 0; s < nshells; s++)
        for (int d = 0; d < 3; d++) center[3 * s + d] = b->xyz[3 * b->atom[s] + d];
    for (int a = 0; a < natoms; a++) atomZ[a] = b->Z[a];
    d_L = dalloc<int>(nshells);
    d_nprim = dalloc<int>(nshells);
    d_poff = dalloc<int>(nshells);
    d_fstart = dalloc<int>(nshells + 1);
    d_atom = dalloc<int>(nshells);
    d_alpha = dalloc<double>(b->alpha.size());
    d_coef = dalloc<double>(b->coef.size());
    d_center = dalloc<double>(3 * nshells);
    d_atom_xyz = dalloc<double>(3 * natoms);
    d_atom_Z = dalloc<double>(natoms);
    d_boys = dalloc<double>((size_t)BOYS_NGRID * BOYS_NCOL);
    if (!d_L || !d_nprim || !d_poff || !d_fstart || !d_atom || !d_alpha || !d_coef || !d_center || !d_atom_xyz ||
        !d_atom_Z || !d_boys)
        return -2;
    if (upload(d_L, b->L, st) || upload(d_nprim, b->nprim, st) || upload(d_poff, b->poff, st) ||
        upload(d_fstart, b->fstart, st) || upload(d_atom, b->atom, st) || upload(d_alpha, b->alpha, st) ||
        upload(d_coef, b->coef, st) || upload(d_center, center, st) || upload(d_atom_xyz, b->xyz, st) ||
        upload(d_atom_Z, atomZ, st))
        return -1;
    std::vector<double> boys((size_t)BOYS_NGRID * BOYS_NCOL);
    gtf_boys_table(boys.data(), BOYS_NGRID, BOYS_NCOL, BOYS_DELTA);
    if (upload(d_boys, boys, st)) return -1;

    // all unordered shell pairs (A >= B by index), canonical orientation by L
    npairs = nshells * (nshells + 1) / 2;
    h_s1.resize(npairs);
    h_s2.resize(npairs);
    h_type.resize(npairs);
    h_pnprim.resize(npairs);
    std::vector<int> ppoff(npairs);
    std::vector<long> peoff(npairs);
    long np = 0, ne = 0;
    for (int A = 0; A < nshells; A++)
        for (int B = 0; B <= A; B++) {
            const int id = A * (A + 1) / 2 + B;
            int s1 = A, s2 = B;
            if (b->L[B] > b->L[A]) std::swap(s1, s2);
            const int la = b->L[s1], lb = b->L[s2];
            h_s1[id] = s1;
            h_s2[id] = s2;
            h_type[id] = pair_type(la, lb);
            h_pnprim[id] = b->nprim[s1] * b->nprim[s2];
            ppoff[id] = (int)np;
            peoff[id] = ne;
            np += h_pnprim[id];
            ne += (long)h_pnprim[id] * nsph(la) * nsph(lb) * nherm(la + lb);
        }
    nprimpairs = np;
    nE = ne;
    d_s1 = dalloc<int>(npairs);
    d_s2 = dalloc<int>(npairs);
    d_pnprim = dalloc<int>(npairs);
    d_ppoff = dalloc<int>(npairs);
    d_peoff = dalloc<long>(npairs);
    d_prim = dalloc<double4>(np);
    d_E = dalloc<double>(ne);
    d_counters = dalloc<int>(2 * GTF_NCLASSES);
    if (!d_s1 || !d_s2 || !d_pnprim || !d_ppoff || !d_peoff || !d_prim || !d_E || !d_counters) return -2;
    if (upload(d_s1, h_s1, st) || upload(d_s2, h_s2, st) || upload(d_pnprim, h_pnprim, st) ||
        upload(d_ppoff, ppoff, st) || upload(d_peoff, peoff, st))
        return -1;
    BasisView bv{nshells, nbf, natoms, d_L, d_nprim, d_poff, d_fstart, d_atom,
                 d_alpha, d_coef, d_center, d_atom_xyz, d_atom_Z};
    if (gtf_build_pairs(st, bv, npairs, d_s1, d_s2, d_ppoff, d_peoff, d_prim, d_E, d_pnprim, prim_eps)) return -1;
    // surviving primitive pairs per shell pair
    SYS_CHECK(hipMemcpyAsync(h_pnprim.data(), d_pnprim, sizeof(int) * npairs, hipMemcpyDeviceToHost, st));
    SYS_CHECK(hipStreamSynchronize(st));
    nprimpairs_kept = 0;
    for (int id = 0; id < npairs; id++) nprimpairs_kept += h_pnprim[id];
    return 0;
}

PairView GtfSystem::pair_view() const
{
    return PairView{d_s1, d_s2, d_pnprim, d_ppoff, d_peoff, d_prim, d_E};
}

static int launch(eri_kernel_t fn, EriArgs& a, int blocks, hipStream_t st, size_t lds = 0)
{
    void* args[] = {&a};
    SYS_CHECK(hipLaunchKernel((const void*)fn, dim3(blocks), dim3(256), args, lds, st));
    return 0;
}

int GtfSystem::screen(double tolscr)
{
    if (const char* s = getenv("GTF_JK_TILES")) use_tiles = atoi(s) != 0;
    {
        int dev = 0, lds = 0;
        if (hipGetDevice(&dev) == hipSuccess &&
            hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess && lds > 0)
            max_lds = (size_t)lds;
    }
    tol = tolscr;
    tol2 = tolscr * tolscr;
    const PairView pv = pair_view();
    unsigned long long* d_vbits = dalloc<unsigned long long>(npairs);
    if (!d_vbits) return -2;
    SYS_CHECK(hipMemsetAsync(d_vbits, 0, sizeof(unsigned long long) * npairs, stream));
    // Schwarz values: (MN|MN) diagonal maxima, one launch per pair type
    std::vector<std::vector<int>> bytype(GTF_NPAIRTYPES);
    for (int id = 0; id < npairs; id++) bytype[h_type[id]].push_back(id);
    int2* d_q = dalloc<int2>(npairs);
    int* d_ids = dalloc<int>(npairs);
    if (!d_q || !d_ids) return -2;
    {
        int off = 0;
        for (int t = 0; t < GTF_NPAIRTYPES; t++) {
            const int n = (int)bytype[t].size();
            if (n == 0) continue;
            SYS_CHECK(hipMemcpyAsync(d_ids + off, bytype[t].data(), sizeof(int) * n, hipMemcpyHostToDevice, stream));
            if (gtf_fill_qlist_diag(stream, n, d_ids + off, d_q + off)) return -1;
            const EriClassInfo* ci = gtf_eri_class(class_index(t, t));
            EriArgs a{};
            a.nunits = (long)n * ci->nsplit;
            a.qlist = d_q + off;
            a.pv = pv;
            a.fstart = d_fstart;
            a.boys = d_boys;
            a.vbits = d_vbits;
            if (launch(ci->diag, a, (int)((a.nunits + 255) / 256), stream)) return -1;
            off += n;
        }
    }
    std::vector<unsigned long long> vb(npairs);
    SYS_CHECK(hipMemcpyAsync(vb.data(), d_vbits, sizeof(unsigned long long) * npairs, hipMemcpyDeviceToHost, stream));
    SYS_CHECK(hipStreamSynchronize(stream));
    h_v.resize(npairs);
    vmax = 0.0;
    for (int id = 0; id < npairs; id++) {
        double v;
        memcpy(&v, &vb[id], sizeof v);
        h_v[id] = v;
        vmax = std::max(vmax, v);
    }
    // significant pairs: v > tol^2 / vmax (pfock/screening.c:111,129)
    const double eta = tol2 / vmax;
    for (int t = 0; t < GTF_NPAIRTYPES; t++) {
        std::vector<int>& s = h_sig[t];
        s.clear();
        for (int id : bytype[t])
            if (h_v[id] > eta) s.push_back(id);
        std::sort(s.begin(), s.end(), [&](int x, int y) { return h_v[x] != h_v[y] ? h_v[x] > h_v[y] : x < y; });
        nsig[t] = (int)s.size();
        std::vector<double> sorted(s.size());
        for (size_t k = 0; k < s.size(); k++) sorted[k] = h_v[s[k]];
        // value rank, then each block re-sorted by primitive count (descending)
        std::vector<int> order(s.size());
        std::iota(order.begin(), order.end(), 0);
        for (size_t b0 = 0; b0 < s.size(); b0 += GTF_KET_BLOCK) {
            const size_t b1 = std::min(s.size(), b0 + GTF_KET_BLOCK);
            std::stable_sort(order.begin() + b0, order.begin() + b1,
                             [&](int x, int y) { return h_pnprim[s[x]] > h_pnprim[s[y]]; });
        }
        std::vector<int> perm(s.size());
        h_sigv[t].resize(s.size());
        h_rank[t].resize(s.size());
        for (size_t k = 0; k < s.size(); k++) {
            perm[k] = s[order[k]];
            h_rank[t][k] = order[k];
            h_sigv[t][k] = sorted[order[k]];
        }
        s = perm;
        d_sig[t] = dalloc<int>(s.size());
        d_sigv[t] = dalloc<double>(s.size());
        d_rank[t] = dalloc<int>(s.size());
        d_sigv_sorted[t] = dalloc<double>(s.size());
        if (!d_sig[t] || !d_sigv[t] || !d_rank[t] || !d_sigv_sorted[t]) return -2;
        if (upload(d_sig[t], s, stream) || upload(d_sigv[t], h_sigv[t], stream) ||
            upload(d_rank[t], h_rank[t], stream) || upload(d_sigv_sorted[t], sorted, stream))
            return -1;
    }
    // quartet rows per class: bra type A >= ket type B
    int maxn = 1;
    for (int t = 0; t < GTF_NPAIRTYPES; t++) maxn = std::max(maxn, nsig[t]);
    long* d_counts = dalloc<long>(maxn);
    long* d_exact = dalloc<long>(maxn);
    if (!d_counts || !d_exact) return -2;
    for (int cls = 0; cls < GTF_NCLASSES; cls++) {
        int A, B;
        gtf_class_types(cls, A, B);
        nq[cls] = 0;
        d_rowoff[cls] = dalloc<long>(nsig[A] + 1);
        if (!d_rowoff[cls]) return -2;
        if (nsig[A] == 0 || nsig[B] == 0) continue;
        if (gtf_count_rows(stream, nsig[A], d_sigv[A], d_rank[A], nsig[B], d_sigv_sorted[B], tol2, A == B,
                           d_exact, d_counts, d_rowoff[cls]))
            return -1;
        std::vector<long> off(nsig[A] + 1), exact(nsig[A]);
        SYS_CHECK(hipMemcpyAsync(off.data(), d_rowoff[cls], sizeof(long) * off.size(), hipMemcpyDeviceToHost, stream));
        SYS_CHECK(hipMemcpyAsync(exact.data(), d_exact, sizeof(long) * exact.size(), hipMemcpyDeviceToHost, stream));
        SYS_CHECK(hipStreamSynchronize(stream));
        nslots[cls] = off[nsig[A]];
        // exact quartet and primitive-quartet counts (value order)
        std::vector<double> kp(nsig[B] + 1, 0.0);
        {
            std::vector<int> byrank(nsig[B]);
            for (int j = 0; j < nsig[B]; j++) byrank[h_rank[B][j]] = h_sig[B][j];
            for (int j = 0; j < nsig[B]; j++) kp[j + 1] = kp[j] + h_pnprim[byrank[j]];
        }
        double w = 0.0;
        long qn = 0;
        for (int r = 0; r < nsig[A]; r++) {
            w += (double)h_pnprim[h_sig[A][r]] * kp[exact[r]];
            qn += exact[r];
        }
        nq[cls] = qn;
        nprimq[cls] = w;
        h_exact[cls] = exact;
        h_slots[cls].resize(nsig[A]);
        for (int r = 0; r < nsig[A]; r++) h_slots[cls][r] = off[r + 1] - off[r];
    }
    if (build_tiles()) return -1;
    screened = true;
    return 0;
}

void GtfSystem::csr(std::vector<int>& shellptr, std::vector<int>& shellid, std::vector<int>& shellrid,
                    std::vector<double>& shellvalue) const
{
    const double eta = tol2 / vmax;
    shellptr.assign(nshells + 1, 0);
    shellid.clear();
    shellrid.clear();
    shellvalue.clear();
    for (int A = 0; A < nshells; A++) {
        shellptr[A] = (int)shellid.size();
        for (int B = 0; B < nshells; B++) {
            const double v = h_v[pair_id(A, B)];
            if (!(v > eta)) continue;
            if ((A > B && (A + B) % 2 == 1) || (A < B && (A + B) % 2 == 0)) continue;
            shellvalue.push_back(A == B ? v : -v);
            shellid.push_back(B);
            shellrid.push_back(A);
        }
    }
    shellptr[nshells] = (int)shellid.size();
}

// Tiles of up to GTF_TILE_UNITS work units of one bra row.  With tau > 0 each
// row splits at the first ket whose Schwarz bound sqrt(v_bra v_ket) < tau:
// FP64 tiles cover the block-padded prefix, FP32 tiles start at the block
// holding the split; the kernels re-test the exact predicates per quartet.
int GtfSystem::build_tiles()
{
    const double tau2 = tau * tau;
    for (int cls = 0; cls < GTF_NCLASSES; cls++) {
        int A, B;
        gtf_class_types(cls, A, B);
        const EriClassInfo* ci = gtf_eri_class(cls);
        const int NA = 2 * ci->la + 1, NB = 2 * ci->lb + 1;
        tile_lds[cls] = sizeof(double) * ((size_t)(NA + NB) * nbf + NA * NB);
        ntiles[cls] = ntiles_lo[cls] = 0;
        nq_lo[cls] = 0;
        nprimq_lo[cls] = 0.0;
        if (nq[cls] == 0 || !use_tiles || tile_lds[cls] > max_lds) continue;
        std::vector<double> vs(nsig[B]);
        std::vector<int> byrank(nsig[B]);
        for (int j = 0; j < nsig[B]; j++) {
            vs[h_rank[B][j]] = h_sigv[B][j];
            byrank[h_rank[B][j]] = h_sig[B][j];
        }
        std::vector<double> kp(nsig[B] + 1, 0.0);
        for (int j = 0; j < nsig[B]; j++) kp[j + 1] = kp[j] + h_pnprim[byrank[j]];
        const int tq = std::max(1, GTF_TILE_UNITS / ci->nsplit);
        std::vector<int4> hi, lo;
        for (int r = 0; r < nsig[A]; r++) {
            const long c = h_exact[cls][r];
            if (c == 0) continue;
            long c64 = c;
            if (tau2 > 0.0) {
                const double vk = h_sigv[A][r];
                c64 = std::upper_bound(vs.begin(), vs.begin() + c, vk, [&](double v, double x) {
                          return !(v * x >= tau2);
                      }) - vs.begin();
            }
            const long pad = std::min<long>(nsig[B], ((c + GTF_KET_BLOCK - 1) / GTF_KET_BLOCK) * GTF_KET_BLOCK);
            const long pad64 = std::min<long>(nsig[B], ((c64 + GTF_KET_BLOCK - 1) / GTF_KET_BLOCK) * GTF_KET_BLOCK);
            for (long j0 = 0; j0 < pad64; j0 += tq)
                hi.push_back(make_int4(r, (int)j0, (int)std::min<long>(pad64, j0 + tq), 0));
            if (c > c64) {
                const long s0 = (c64 / GTF_KET_BLOCK) * GTF_KET_BLOCK;
                for (long j0 = s0; j0 < pad; j0 += tq)
                    lo.push_back(make_int4(r, (int)j0, (int)std::min<long>(pad, j0 + tq), 0));
                nq_lo[cls] += c - c64;
                nprimq_lo[cls] += (double)h_pnprim[h_sig[A][r]] * (kp[c] - kp[c64]);
            }
        }
        ntiles[cls] = (int)hi.size();
        ntiles_lo[cls] = (int)lo.size();
        if (!hi.empty()) {
            d_tiles[cls] = dalloc<int4>(hi.size());
            if (!d_tiles[cls] || upload(d_tiles[cls], hi, stream)) return -1;
        }
        if (!lo.empty()) {
            d_tiles_lo[cls] = dalloc<int4>(lo.size());
            if (!d_tiles_lo[cls] || upload(d_tiles_lo[cls], lo, stream)) return -1;
        }
    }
    SYS_CHECK(hipStreamSynchronize(stream));
    return 0;
}

int GtfSystem::jk(const double* D, double* J, double* K, int ld)
{
    if (gtf_fill_int(stream, d_counters, 2 * GTF_NCLASSES, 0)) return -1;
    const PairView pv = pair_view();
    if (!cstream[0]) {
        for (int s = 0; s < NSTREAMS; s++) {
            SYS_CHECK(hipStreamCreate(&cstream[s]));
            SYS_CHECK(hipEventCreateWithFlags(&join_ev[s], hipEventDisableTiming));
        }
        SYS_CHECK(hipEventCreateWithFlags(&fork_ev, hipEventDisableTiming));
    }
    SYS_CHECK(hipEventRecord(fork_ev, stream));
    for (int s = 0; s < NSTREAMS; s++) SYS_CHECK(hipStreamWaitEvent(cstream[s], fork_ev, 0));
    // classes in decreasing order of work, dealt round-robin over the streams
    int order[GTF_NCLASSES];
    for (int c = 0; c < GTF_NCLASSES; c++) order[c] = c;
    std::sort(order, order + GTF_NCLASSES, [&](int x, int y) { return nprimq[x] > nprimq[y]; });
    int slot = 0;
    for (int oi = 0; oi < GTF_NCLASSES; oi++) {
        const int cls = order[oi];
        if (nq[cls] == 0) continue;
        int A, B;
        gtf_class_types(cls, A, B);
        const EriClassInfo* ci = gtf_eri_class(cls);
        hipStream_t cs = cstream[slot++ % NSTREAMS];
        EriArgs a{};
        a.nunits = nslots[cls] * ci->nsplit;
        a.row_off = d_rowoff[cls];
        a.vbra = d_sigv[A];
        a.vket = d_sigv[B];
        a.rank_bra = d_rank[A];
        a.rank_ket = d_rank[B];
        a.tol2 = tol2;
        a.same = A == B;
        a.nrows = nsig[A];
        a.bra_list = d_sig[A];
        a.ket_list = d_sig[B];
        a.counter = d_counters + cls;
        a.ntasks = (int)((a.nunits + 255) / 256);
        a.pv = pv;
        a.fstart = d_fstart;
        a.boys = d_boys;
        a.D = D;
        a.J = J;
        a.K = K;
        a.ld = ld;
        if (timing) SYS_CHECK(hipEventRecord(tev[2 * cls], cs));
        if (ntiles[cls] > 0 || ntiles_lo[cls] > 0) {
            a.nbf = nbf;
            a.tau2 = tau * tau;
            if (ntiles[cls] > 0) {
                a.tiles = d_tiles[cls];
                a.ntiles = ntiles[cls];
                a.counter = d_counters + cls;
                if (launch(ci->jk_tile, a, std::min(a.ntiles, 2048), cs, tile_lds[cls])) return -1;
            }
            if (ntiles_lo[cls] > 0) {
                a.tiles = d_tiles_lo[cls];
                a.ntiles = ntiles_lo[cls];
                a.counter = d_counters + GTF_NCLASSES + cls;
                if (launch(ci->jk_tile_lo, a, std::min(a.ntiles, 2048), cs, tile_lds[cls])) return -1;
            }
        } else if (launch(ci->jk, a, std::min(a.ntasks, 2048), cs)) {
            return -1;
        }
        if (timing) SYS_CHECK(hipEventRecord(tev[2 * cls + 1], cs));
    }
    for (int s = 0; s < NSTREAMS; s++) {
        SYS_CHECK(hipEventRecord(join_ev[s], cstream[s]));
        SYS_CHECK(hipStreamWaitEvent(stream, join_ev[s], 0));
    }
    return 0;
}

int GtfSystem::collect_times()
{
    if (!timing) return 0;
    for (int cls = 0; cls < GTF_NCLASSES; cls++) {
        if (nq[cls] == 0) continue;
        SYS_CHECK(hipEventSynchronize(tev[2 * cls + 1]));
        float ms = 0.f;
        SYS_CHECK(hipEventElapsedTime(&ms, tev[2 * cls], tev[2 * cls + 1]));
        class_ms[cls] += ms;
        class_launches[cls] += 1;
    }
    return 0;
}

int GtfSystem::store(int cls, int nq_, const int2* d_qlist, double* d_out)
{
    if (nq_ <= 0) return 0;
    const EriClassInfo* ci = gtf_eri_class(cls);
    EriArgs a{};
    a.nunits = (long)nq_ * ci->nsplit;
    a.qlist = d_qlist;
    a.pv = pair_view();
    a.fstart = d_fstart;
    a.boys = d_boys;
    a.out = d_out;
    return launch(ci->store, a, (int)((a.nunits + 255) / 256), stream);
}

void GtfSystem::release()
{
    for (int s = 0; s < NSTREAMS; s++) {
        if (cstream[s]) { (void)hipStreamSynchronize(cstream[s]); (void)hipStreamDestroy(cstream[s]); cstream[s] = nullptr; }
        if (join_ev[s]) { (void)hipEventDestroy(join_ev[s]); join_ev[s] = nullptr; }
    }
    if (fork_ev) { (void)hipEventDestroy(fork_ev); fork_ev = nullptr; }
    for (hipEvent_t& e : tev)
        if (e) { (void)hipEventDestroy(e); e = nullptr; }
    for (void* p : allocs) (void)hipFree(p);
    allocs.clear();
    dev_bytes = 0;
}
