// engine.cpp -- GtfSystem: device basis, shell pairs, Schwarz screening,
// quartet compaction and the per-class ERI launches.
#include "engine.h"

#include <sys/time.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <atomic>
#include <numeric>
#include <thread>
#include <vector>

#include "screen.h"

static double sys_wtime()
{
    struct timeval tv;
    gettimeofday(&tv, nullptr);
    return tv.tv_sec + 1e-6 * tv.tv_usec;
}

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

int GtfSystem::init(const std::vector<const BasisSet*>& mols, hipStream_t st)
{
    stream = st;
    nmol = (int)mols.size();
    if (nmol == 0) return -3;
    double t0 = sys_wtime();
    nbf = nshells = natoms = 0;
    mol_nbf.assign(nmol, 0);
    mol_shell_off.assign(nmol, 0);
    mol_nshells.assign(nmol, 0);
    mol_pair_off.assign(nmol, 0);
    mol_npairs.assign(nmol, 0);
    mol_mat_off.assign(nmol, 0);
    mat_total = 0;
    npairs = 0;
    std::vector<int> L, nprim, poff, fstart, atom, mld;
    std::vector<long> mbase;
    std::vector<double> alpha, coef, center, atom_xyz, atomZ;
    h_shell_mol.clear();
    for (int m = 0; m < nmol; m++) {
        const BasisSet* b = mols[m];
        mol_nbf[m] = b->nbf;
        mol_shell_off[m] = nshells;
        mol_nshells[m] = b->nshells;
        mol_pair_off[m] = npairs;
        mol_npairs[m] = b->nshells * (b->nshells + 1) / 2;
        mol_mat_off[m] = mat_total;
        for (int s = 0; s < b->nshells; s++) {
            L.push_back(b->L[s]);
            nprim.push_back(b->nprim[s]);
            poff.push_back((int)alpha.size() + b->poff[s]);
            fstart.push_back(b->fstart[s]);
            atom.push_back(natoms + b->atom[s]);
            mbase.push_back(mat_total);
            mld.push_back(b->nbf);
            h_shell_mol.push_back(m);
            for (int d = 0; d < 3; d++) center.push_back(b->xyz[3 * b->atom[s] + d]);
        }
        alpha.insert(alpha.end(), b->alpha.begin(), b->alpha.end());
        coef.insert(coef.end(), b->coef.begin(), b->coef.end());
        atom_xyz.insert(atom_xyz.end(), b->xyz.begin(), b->xyz.end());
        for (int a = 0; a < b->natoms; a++) atomZ.push_back(b->Z[a]);
        nbf = std::max(nbf, b->nbf);
        nshells += b->nshells;
        natoms += b->natoms;
        npairs += mol_npairs[m];
        mat_total += (long)b->nbf * b->nbf;
    }
    fstart.push_back(mols[nmol - 1]->nbf);
    std::vector<int> moln(mol_nbf);
    d_L = dalloc<int>(nshells);
    d_nprim = dalloc<int>(nshells);
    d_poff = dalloc<int>(nshells);
    d_fstart = dalloc<int>(nshells + 1);
    d_atom = dalloc<int>(nshells);
    d_mbase = dalloc<long>(nshells);
    d_mld = dalloc<int>(nshells);
    d_mol_off = dalloc<long>(nmol);
    d_mol_n = dalloc<int>(nmol);
    d_shell_mol = dalloc<int>(nshells);
    d_mol_active = dalloc<int>(nmol);
    d_alpha = dalloc<double>(alpha.size());
    d_coef = dalloc<double>(coef.size());
    d_center = dalloc<double>(3 * nshells);
    d_atom_xyz = dalloc<double>(3 * natoms);
    d_atom_Z = dalloc<double>(natoms);
    d_boys = dalloc<double>((size_t)BOYS_NGRID * BOYS_NCOL);
    if (!d_L || !d_nprim || !d_poff || !d_fstart || !d_atom || !d_mbase || !d_mld || !d_mol_off || !d_mol_n ||
        !d_alpha || !d_coef || !d_center || !d_atom_xyz || !d_atom_Z || !d_boys || !d_shell_mol || !d_mol_active)
        return -2;
    const std::vector<int> all_active(nmol, 1);
    if (upload(d_shell_mol, h_shell_mol, st) || upload(d_mol_active, all_active, st)) return -1;
    if (upload(d_L, L, st) || upload(d_nprim, nprim, st) || upload(d_poff, poff, st) ||
        upload(d_fstart, fstart, st) || upload(d_atom, atom, st) || upload(d_mbase, mbase, st) ||
        upload(d_mld, mld, st) || upload(d_mol_off, mol_mat_off, st) || upload(d_mol_n, moln, st) ||
        upload(d_alpha, alpha, st) || upload(d_coef, coef, st) || upload(d_center, center, st) ||
        upload(d_atom_xyz, atom_xyz, st) || upload(d_atom_Z, atomZ, st))
        return -1;
    std::vector<double> boys((size_t)BOYS_NGRID * BOYS_NCOL);
    gtf_boys_table(boys.data(), BOYS_NGRID, BOYS_NCOL, BOYS_DELTA);
    if (upload(d_boys, boys, st)) return -1;
    std::vector<double> b8((size_t)BOYS8_NM * BOYS8_SLAB);
    gtf_boys_taylor_tables(b8.data());
    d_boys8 = dalloc<double>(b8.size());
    if (!d_boys8) return -2;
    if (upload(d_boys8, b8, st)) return -1;
    // the uploads read host vectors that go out of scope here
    SYS_CHECK(hipStreamSynchronize(st));
    setup_s[0] = sys_wtime() - t0;
    t0 = sys_wtime();

    // all unordered shell pairs (A >= B by local index) of every molecule,
    // canonical orientation by L
    h_s1.resize(npairs);
    h_s2.resize(npairs);
    h_type.resize(npairs);
    h_pnprim.resize(npairs);
    h_pair_mol.resize(npairs);
    std::vector<int> ppoff(npairs);
    std::vector<long> peoff(npairs), pzoff(npairs);
    long np = 0, ne = 0, nz = 0;
    for (int m = 0; m < nmol; m++) {
        const int s0 = mol_shell_off[m];
        for (int a = 0; a < mol_nshells[m]; a++)
            for (int c = 0; c <= a; c++) {
                const int id = mol_pair_off[m] + a * (a + 1) / 2 + c;
                int s1 = s0 + a, s2 = s0 + c;
                if (L[s2] > L[s1]) std::swap(s1, s2);
                const int la = L[s1], lb = L[s2];
                h_s1[id] = s1;
                h_s2[id] = s2;
                h_type[id] = pair_type(la, lb);
                h_pnprim[id] = nprim[s1] * nprim[s2];
                h_pair_mol[id] = m;
                ppoff[id] = (int)np;
                peoff[id] = ne;
                np += h_pnprim[id];
                ne += (long)h_pnprim[id] * nsph(la) * nsph(lb) * nherm(la + lb);
                pzoff[id] = nz;
                nz += (long)h_pnprim[id] * herm_nnz_lookup(la, lb);
            }
    }
    for (int id = 0; id < npairs; id++)
        if (h_pnprim[id] > GTF_MAX_RAW_PRIM_PAIRS) {
            fprintf(stderr, "gtfock_amd: a shell pair has %d primitive pairs (at most %d supported)\n", h_pnprim[id],
                    GTF_MAX_RAW_PRIM_PAIRS);
            return -3;
        }
    nprimpairs = np;
    nE = ne;
    d_s1 = dalloc<int>(npairs);
    d_s2 = dalloc<int>(npairs);
    d_pnprim = dalloc<int>(npairs);
    d_ppoff = dalloc<int>(npairs);
    d_peoff = dalloc<long>(npairs);
    d_prim = dalloc<double4>(np);
    d_pq = dalloc<double>(np);
    d_E = dalloc<double>(ne);
    d_pzoff = dalloc<long>(npairs);
    d_Ez = dalloc<double>(nz);
    d_counters = dalloc<int>(2 * GTF_NCLASSES);
    if (!d_s1 || !d_s2 || !d_pnprim || !d_ppoff || !d_peoff || !d_prim || !d_E || !d_counters || !d_pzoff || !d_Ez ||
        !d_pq)
        return -2;
    if (upload(d_s1, h_s1, st) || upload(d_s2, h_s2, st) || upload(d_pnprim, h_pnprim, st) ||
        upload(d_ppoff, ppoff, st) || upload(d_peoff, peoff, st) || upload(d_pzoff, pzoff, st))
        return -1;
    BasisView bv{nshells, nbf, natoms, d_L, d_nprim, d_poff, d_fstart, d_atom,
                 d_alpha, d_coef, d_center, d_atom_xyz, d_atom_Z};
    if (gtf_build_pairs(st, bv, npairs, d_s1, d_s2, d_ppoff, d_peoff, d_prim, d_E, d_pzoff, d_Ez, d_pnprim, d_pq,
                        prim_eps))
        return -1;
    // surviving primitive pairs per shell pair
    SYS_CHECK(hipMemcpyAsync(h_pnprim.data(), d_pnprim, sizeof(int) * npairs, hipMemcpyDeviceToHost, st));
    SYS_CHECK(hipStreamSynchronize(st));
    nprimpairs_kept = 0;
    for (int id = 0; id < npairs; id++) nprimpairs_kept += h_pnprim[id];
    setup_s[1] = sys_wtime() - t0;
    return 0;
}

bool GtfSystem::exchanged(int cls) const
{
    return gtf_eri_class(cls)->nsplit_t > 0 && ((exchanged_mask >> cls) & 1u);
}

void GtfSystem::dfree(void* p)
{
    if (!p) return;
    auto it = std::find(allocs.begin(), allocs.end(), p);
    if (it == allocs.end()) return;
    const size_t k = it - allocs.begin();
    live_bytes -= alloc_bytes[k];
    alloc_bytes.erase(alloc_bytes.begin() + k);
    allocs.erase(it);
    (void)hipFree(p);
}

PairView GtfSystem::pair_view() const
{
    return PairView{d_s1, d_s2, d_pnprim, d_ppoff, d_peoff, d_prim, d_E, d_pzoff, d_Ez, d_pq, 0.0};
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
    if (const char* s = getenv("GTF_JK_SERIAL")) serial = atoi(s) != 0;
    if (const char* s = getenv("GTF_JK_EXCHANGED")) exchanged_mask = (unsigned)strtoul(s, nullptr, 0);
    if (const char* s = getenv("GTF_JK_GRID"))
        for (TileShape& t : shape) t.grid = std::max(1, atoi(s));
    {
        int dev = 0, lds = 0;
        if (hipGetDevice(&dev) == hipSuccess &&
            hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess && lds > 0)
            max_lds = (size_t)lds;
    }
    tol = tolscr;
    tol2 = tolscr * tolscr;
    double t0 = sys_wtime();
    const PairView pv = pair_view();
    unsigned long long* d_vbits = dalloc<unsigned long long>(npairs);
    int2* d_q = dalloc<int2>(npairs);
    int* d_ids = dalloc<int>(npairs);
    if (!d_vbits || !d_q || !d_ids) return -2;
    SYS_CHECK(hipMemsetAsync(d_vbits, 0, sizeof(unsigned long long) * npairs, stream));
    // Schwarz values: (MN|MN) diagonal maxima, one launch per pair type
    std::vector<std::vector<int>> bytype(GTF_NPAIRTYPES);
    for (int id = 0; id < npairs; id++) bytype[h_type[id]].push_back(id);
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
            a.boys8 = d_boys8;
                    a.vbits = d_vbits;
            if (launch(ci->diag, a, (int)((a.nunits + 255) / 256), stream)) return -1;
            off += n;
        }
    }
    std::vector<unsigned long long> vb(npairs);
    SYS_CHECK(hipMemcpyAsync(vb.data(), d_vbits, sizeof(unsigned long long) * npairs, hipMemcpyDeviceToHost, stream));
    SYS_CHECK(hipStreamSynchronize(stream));
    dfree(d_q);
    dfree(d_ids);
    h_v.resize(npairs);
    mol_vmax.assign(nmol, 0.0);
    for (int id = 0; id < npairs; id++) {
        double v;
        memcpy(&v, &vb[id], sizeof v);
        h_v[id] = v;
        mol_vmax[h_pair_mol[id]] = std::max(mol_vmax[h_pair_mol[id]], v);
    }
    vmax = mol_vmax[0];
    setup_s[2] = sys_wtime() - t0;
    t0 = sys_wtime();
    // significant pairs of molecule m: v > tol^2 / vmax_m (pfock/screening.c:111,129),
    // in one segment per (molecule, primitive-count bucket), compacted and
    // ordered on the device (sigsort.hip); the host keeps copies for the
    // tile cutting
    std::vector<double> eta(nmol);
    for (int m = 0; m < nmol; m++) eta[m] = mol_vmax[m] > 0.0 ? tol2 / mol_vmax[m] : HUGE_VAL;
    nseg = nmol * NBUCKET;
    double* d_eta = dalloc<double>(nmol);
    int* d_ptype = dalloc<int>(npairs);
    int* d_pmol = dalloc<int>(npairs);
    const int nblocks = (npairs + 255) / 256;
    int* d_bcount = dalloc<int>(nblocks);
    int* d_boff = dalloc<int>(nblocks);
    int* d_total = dalloc<int>(1);
    if (!d_eta || !d_ptype || !d_pmol || !d_bcount || !d_boff || !d_total) return -2;
    if (upload(d_eta, eta, stream) || upload(d_ptype, h_type, stream) || upload(d_pmol, h_pair_mol, stream)) return -1;
    for (int t = 0; t < GTF_NPAIRTYPES; t++) {
        h_listprim[t].clear();
        h_segmax[t].clear();
        int n = 0;
        if (gtf_sig_count(stream, npairs, d_ptype, t, d_vbits, d_pmol, d_eta, d_bcount, d_boff, d_total)) return -1;
        SYS_CHECK(hipMemcpyAsync(&n, d_total, sizeof(int), hipMemcpyDeviceToHost, stream));
        SYS_CHECK(hipStreamSynchronize(stream));
        nsig[t] = n;
        const size_t nn = std::max(n, 1);
        SigWork w{};
        SigOut o{};
        w.ids_a = dalloc<int>(nn);
        w.ids_b = dalloc<int>(nn);
        w.pos_a = dalloc<int>(nn);
        w.pos_b = dalloc<int>(nn);
        w.rank = dalloc<int>(nn);
        w.k64_a = dalloc<unsigned long long>(nn);
        w.k64_b = dalloc<unsigned long long>(nn);
        w.k32_a = dalloc<unsigned>(nn);
        w.k32_b = dalloc<unsigned>(nn);
        w.temp_bytes = gtf_sig_temp_bytes((int)nn);
        w.temp = dalloc<char>(w.temp_bytes);
        o.s_id = dalloc<int>(nn);
        o.sortnprim = dalloc<int>(nn);
        o.sigmol = dalloc<int>(nn);
        d_sig[t] = o.sig = dalloc<int>(nn);
        d_sigv[t] = o.sigv = dalloc<double>(nn);
        d_rank[t] = o.rank = dalloc<int>(nn);
        d_sigv_sorted[t] = o.sorted = dalloc<double>(nn);
        d_sortrank[t] = o.sortrank = dalloc<int>(nn);
        d_seg_start[t] = o.seg_start = dalloc<int>(nseg);
        d_seg_len[t] = o.seg_len = dalloc<int>(nseg);
        if (!w.ids_a || !w.ids_b || !w.pos_a || !w.pos_b || !w.rank || !w.k64_a || !w.k64_b || !w.k32_a ||
            !w.k32_b || !w.temp || !o.s_id || !o.sortnprim || !o.sigmol || !o.sig || !o.sigv || !o.rank ||
            !o.sorted || !o.sortrank || !o.seg_start || !o.seg_len)
            return -2;
        if (gtf_fill_int(stream, o.seg_start, nseg, 0) || gtf_fill_int(stream, o.seg_len, nseg, 0)) return -1;
        if (gtf_sig_order(stream, npairs, d_ptype, t, d_vbits, d_pmol, d_eta, d_pnprim, NBUCKET, d_boff, n, w, o))
            return -1;
        h_sig[t].resize(n);
        h_sigv[t].resize(n);
        h_rank[t].resize(n);
        h_sigmol[t].resize(n);
        h_sorted[t].resize(n);
        h_sortrank[t].resize(n);
        h_sortnprim[t].resize(n);
        seg_len[t].assign(nseg, 0);
        seg_start[t].assign(nseg, 0);
        if (n > 0) {
            SYS_CHECK(hipMemcpyAsync(h_sig[t].data(), o.sig, sizeof(int) * n, hipMemcpyDeviceToHost, stream));
            SYS_CHECK(hipMemcpyAsync(h_sigv[t].data(), o.sigv, sizeof(double) * n, hipMemcpyDeviceToHost, stream));
            SYS_CHECK(hipMemcpyAsync(h_rank[t].data(), o.rank, sizeof(int) * n, hipMemcpyDeviceToHost, stream));
            SYS_CHECK(hipMemcpyAsync(h_sigmol[t].data(), o.sigmol, sizeof(int) * n, hipMemcpyDeviceToHost, stream));
            SYS_CHECK(hipMemcpyAsync(h_sorted[t].data(), o.sorted, sizeof(double) * n, hipMemcpyDeviceToHost, stream));
            SYS_CHECK(hipMemcpyAsync(h_sortrank[t].data(), o.sortrank, sizeof(int) * n, hipMemcpyDeviceToHost,
                                     stream));
            SYS_CHECK(hipMemcpyAsync(h_sortnprim[t].data(), o.sortnprim, sizeof(int) * n, hipMemcpyDeviceToHost,
                                     stream));
            SYS_CHECK(hipMemcpyAsync(seg_len[t].data(), o.seg_len, sizeof(int) * nseg, hipMemcpyDeviceToHost, stream));
        }
        SYS_CHECK(hipStreamSynchronize(stream));
        // segment starts in segment order, empty segments included
        for (int g = 0, at = 0; g < nseg; g++) {
            seg_start[t][g] = at;
            at += seg_len[t][g];
        }
        if (upload(o.seg_start, seg_start[t], stream)) return -1;
        SYS_CHECK(hipStreamSynchronize(stream));
        void* scratch[] = {w.ids_a, w.ids_b, w.pos_a, w.pos_b, w.rank, w.k64_a, w.k64_b, w.k32_a, w.k32_b, w.temp,
                           o.s_id, o.sortnprim, o.sigmol};
        for (void* q : scratch) dfree(q);
    }
    dfree(d_vbits);
    dfree(d_eta);
    dfree(d_ptype);
    dfree(d_pmol);
    dfree(d_bcount);
    dfree(d_boff);
    dfree(d_total);
    setup_s[3] = sys_wtime() - t0;
    t0 = sys_wtime();
    // quartet rows per class: bra type A >= ket type B; a row pairs one bra
    // with one non-empty ket segment of its molecule
    int maxrows = 1;
    for (int cls = 0; cls < GTF_NCLASSES; cls++) {
        int A, B;
        gtf_class_types(cls, A, B);
        h_row_bra[cls].clear();
        h_row_seg[cls].clear();
        for (int r = 0; r < nsig[A]; r++) {
            const int m = h_sigmol[A][r];
            for (int b = 0; b < NBUCKET; b++)
                if (seg_len[B][m * NBUCKET + b] > 0) {
                    h_row_bra[cls].push_back(r);
                    h_row_seg[cls].push_back(m * NBUCKET + b);
                }
        }
        nrows[cls] = (int)h_row_bra[cls].size();
        maxrows = std::max(maxrows, nrows[cls]);
    }
    long* d_counts = dalloc<long>(maxrows);
    long* d_exact = dalloc<long>(maxrows);
    if (!d_counts || !d_exact) return -2;
    for (int cls = 0; cls < GTF_NCLASSES; cls++) {
        int A, B;
        gtf_class_types(cls, A, B);
        const int nr = nrows[cls];
        nq[cls] = 0;
        nslots[cls] = 0;
        nprimq[cls] = 0.0;
        h_exact[cls].assign(nr, 0);
        std::vector<int> kstart(nr);
        for (int k = 0; k < nr; k++) kstart[k] = seg_start[B][h_row_seg[cls][k]];
        d_rowoff[cls] = dalloc<long>(nr + 1);
        d_row_bra[cls] = dalloc<int>(nr);
        d_row_seg[cls] = dalloc<int>(nr);
        d_row_kstart[cls] = dalloc<int>(nr);
        if (!d_rowoff[cls] || !d_row_bra[cls] || !d_row_seg[cls] || !d_row_kstart[cls]) return -2;
        if (nr == 0) continue;
        if (upload(d_row_bra[cls], h_row_bra[cls], stream) || upload(d_row_seg[cls], h_row_seg[cls], stream) ||
            upload(d_row_kstart[cls], kstart, stream))
            return -1;
        if (gtf_count_rows(stream, nr, d_row_bra[cls], d_row_seg[cls], d_sigv[A], d_rank[A], d_sigv_sorted[B],
                           d_sortrank[B], d_seg_start[B], d_seg_len[B], tol2, A == B, d_exact, d_counts,
                           d_rowoff[cls]))
            return -1;
        std::vector<long> off(nr + 1), exact(nr);
        SYS_CHECK(hipMemcpyAsync(off.data(), d_rowoff[cls], sizeof(long) * off.size(), hipMemcpyDeviceToHost, stream));
        SYS_CHECK(hipMemcpyAsync(exact.data(), d_exact, sizeof(long) * exact.size(), hipMemcpyDeviceToHost, stream));
        SYS_CHECK(hipStreamSynchronize(stream));
        nslots[cls] = off[nr];
        // exact quartet and primitive-quartet counts (value order per segment)
        const std::vector<double> kp = ket_prim_prefix(B);
        double w = 0.0;
        long qn = 0;
        for (int k = 0; k < nr; k++) {
            const int s0 = kstart[k];
            w += (double)h_pnprim[h_sig[A][h_row_bra[cls][k]]] * (kp[s0 + exact[k]] - kp[s0]);
            qn += exact[k];
        }
        nq[cls] = qn;
        nprimq[cls] = w;
        h_exact[cls] = exact;
    }
    dfree(d_counts);
    dfree(d_exact);
    setup_s[4] = sys_wtime() - t0;
    if (build_tiles()) return -1;
    screened = true;
    return 0;
}

// prefix sums of primitive-pair counts over a type list in value order
// (segment by segment): kp[s0 + k] - kp[s0] = sum over the k largest of a segment
std::vector<double> GtfSystem::ket_prim_prefix(int B) const
{
    std::vector<double> kp(nsig[B] + 1, 0.0);
    for (int j = 0; j < nsig[B]; j++) kp[j + 1] = kp[j] + h_sortnprim[B][j];
    return kp;
}

// prefix sums of primitive-pair counts over a type list in list order
const std::vector<double>& GtfSystem::list_prim_prefix(int B)
{
    std::vector<double>& lp = h_listprim[B];
    if ((int)lp.size() != nsig[B] + 1) {
        lp.assign(nsig[B] + 1, 0.0);
        for (int j = 0; j < nsig[B]; j++) lp[j + 1] = lp[j] + h_pnprim[h_sig[B][j]];
    }
    return lp;
}

int GtfSystem::seg_nprim_max(int B, int sg)
{
    std::vector<int>& mx = h_segmax[B];
    if ((int)mx.size() != nseg) {
        mx.assign(nseg, 1);
        for (int g = 0; g < nseg; g++)
            for (int j = seg_start[B][g]; j < seg_start[B][g] + seg_len[B][g]; j++)
                mx[g] = std::max(mx[g], h_pnprim[h_sig[B][j]]);
    }
    return mx[sg];
}

long GtfSystem::quartets(int* out, long cap) const
{
    long k = 0;
    for (int cls = 0; cls < GTF_NCLASSES; cls++) {
        if (nq[cls] == 0) continue;
        int A, B;
        gtf_class_types(cls, A, B);
        for (int row = 0; row < nrows[cls]; row++) {
            const int r = h_row_bra[cls][row], sg = h_row_seg[cls][row];
            const int s0 = seg_start[B][sg];
            for (int j = s0; j < s0 + seg_len[B][sg]; j++) {
                if (!(h_sigv[A][r] * h_sigv[B][j] >= tol2) || (A == B && h_rank[B][j] > h_rank[A][r])) continue;
                if (out && k < cap) {
                    const int bra = h_sig[A][r], ket = h_sig[B][j];
                    out[4 * k + 0] = h_s1[bra];
                    out[4 * k + 1] = h_s2[bra];
                    out[4 * k + 2] = h_s1[ket];
                    out[4 * k + 3] = h_s2[ket];
                }
                k++;
            }
        }
    }
    return k;
}

int GtfSystem::prim_census(double* hi, double* lo)
{
    SYS_CHECK(hipStreamSynchronize(stream));
    unsigned long long* d_out = dalloc<unsigned long long>(2 * GTF_NCLASSES);
    if (!d_out) return -2;
    SYS_CHECK(hipMemsetAsync(d_out, 0, sizeof(unsigned long long) * 2 * GTF_NCLASSES, stream));
    PairView pv = pair_view();
    pv.peps = prim_tol;
    std::vector<std::vector<int4>> rows;  // tile lists of the classes that run without tiles
    std::vector<int4*> tmp;
    int rc = 0;
    for (int cls = 0; cls < GTF_NCLASSES && rc == 0; cls++) {
        if (nq[cls] == 0) continue;
        int A, B;
        gtf_class_types(cls, A, B);
        const EriClassInfo* ci = gtf_eri_class(cls);
        CensusArgs c{};
        c.bra_list = d_sig[A];
        c.ket_list = d_sig[B];
        c.vbra = d_sigv[A];
        c.vket = d_sigv[B];
        c.rank_bra = d_rank[A];
        c.rank_ket = d_rank[B];
        c.tol2 = tol2;
        c.same = A == B;
        c.screened = prim_screened_l(ci->la + ci->lb + ci->lc + ci->ld) ? 1 : 0;
        c.pv = pv;
        if (ntiles[cls] > 0 || ntiles_lo[cls] > 0) {
            c.tau2 = class_tau2(cls);
            if (ntiles[cls] > 0) {
                c.tiles = d_tiles[cls];
                c.ntiles = ntiles[cls];
                c.low = 0;
                c.out = d_out + cls;
                if (gtf_prim_census(stream, c)) rc = -1;
            }
            if (ntiles_lo[cls] > 0 && rc == 0) {
                c.tiles = d_tiles_lo[cls];
                c.ntiles = ntiles_lo[cls];
                c.low = 1;
                c.pv.peps = prim_tol_fp32();
                c.out = d_out + GTF_NCLASSES + cls;
                if (gtf_prim_census(stream, c)) rc = -1;
            }
        } else {
            // the global-atomic path evaluates each row's exact ket prefix in FP64
            rows.emplace_back();
            std::vector<int4>& v = rows.back();
            for (int k = 0; k < nrows[cls]; k++)
                if (h_exact[cls][k] > 0) {
                    const int s0 = seg_start[B][h_row_seg[cls][k]];
                    v.push_back(make_int4(h_row_bra[cls][k], s0, s0 + (int)h_exact[cls][k], 1));
                }
            if (v.empty()) continue;
            int4* d = dalloc<int4>(v.size());
            if (!d || upload(d, v, stream)) {
                rc = -1;
                break;
            }
            tmp.push_back(d);
            c.tiles = d;
            c.ntiles = (int)v.size();
            c.tau2 = 0.0;
            c.low = 0;
            c.out = d_out + cls;
            if (gtf_prim_census(stream, c)) rc = -1;
        }
    }
    unsigned long long h[2 * GTF_NCLASSES] = {0};
    if (rc == 0 && (hipMemcpyAsync(h, d_out, sizeof h, hipMemcpyDeviceToHost, stream) != hipSuccess ||
                    hipStreamSynchronize(stream) != hipSuccess))
        rc = -1;
    (void)hipStreamSynchronize(stream);  // the row lists are read by the uploads
    for (int4* d : tmp) dfree(d);
    dfree(d_out);
    if (rc) return rc;
    for (int cls = 0; cls < GTF_NCLASSES; cls++) {
        hi[cls] = (double)h[cls];
        lo[cls] = (double)h[GTF_NCLASSES + cls];
    }
    return 0;
}

void GtfSystem::csr(std::vector<int>& shellptr, std::vector<int>& shellid, std::vector<int>& shellrid,
                    std::vector<double>& shellvalue) const
{
    const double eta = tol2 / mol_vmax[0];
    const int nshells = mol_nshells[0];
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

int GtfSystem::build_tiles()
{
    const double t0 = sys_wtime();
    // the previous tiles may still be read by a launch in flight
    SYS_CHECK(hipStreamSynchronize(stream));
    // fill the shared prefix and segment-maximum caches first: the classes
    // are then cut on host threads that only read shared state
    for (int t = 0; t < GTF_NPAIRTYPES; t++) {
        list_prim_prefix(t);
        seg_nprim_max(t, 0);
    }
    std::vector<int4> hv[GTF_NCLASSES], lv[GTF_NCLASSES];
    {
        const int nth = std::max(1, std::min<int>(16, (int)std::thread::hardware_concurrency()));
        std::atomic<int> next{0};
        auto work = [&]() {
            for (int c; (c = next.fetch_add(1)) < GTF_NCLASSES;) cut_class_tiles(c, hv[c], lv[c]);
        };
        std::vector<std::thread> pool;
        for (int t = 1; t < nth; t++) pool.emplace_back(work);
        work();
        for (std::thread& t : pool) t.join();
    }
    for (int cls = 0; cls < GTF_NCLASSES; cls++)
        if (upload_class_tiles(cls, hv[cls], lv[cls])) return -1;
    // the uploads read hv / lv
    SYS_CHECK(hipStreamSynchronize(stream));
    setup_s[5] = sys_wtime() - t0;
    return 0;
}

int GtfSystem::build_class_tiles(int cls)
{
    int A, B;
    gtf_class_types(cls, A, B);
    SYS_CHECK(hipStreamSynchronize(stream));
    list_prim_prefix(B);
    seg_nprim_max(B, 0);
    std::vector<int4> hv, lv;
    cut_class_tiles(cls, hv, lv);
    if (upload_class_tiles(cls, hv, lv)) return -1;
    SYS_CHECK(hipStreamSynchronize(stream));
    return 0;
}

// replaces the device tiles of a class by hv (FP64) and lv (FP32); the
// stream must have no launch in flight that reads the old ones
int GtfSystem::upload_class_tiles(int cls, const std::vector<int4>& hv, const std::vector<int4>& lv)
{
    dfree(d_tiles[cls]);
    dfree(d_tiles_lo[cls]);
    d_tiles[cls] = d_tiles_lo[cls] = nullptr;
    if (!hv.empty()) {
        d_tiles[cls] = dalloc<int4>(hv.size());
        if (!d_tiles[cls] || upload(d_tiles[cls], hv, stream)) return -1;
    }
    if (!lv.empty()) {
        d_tiles_lo[cls] = dalloc<int4>(lv.size());
        if (!d_tiles_lo[cls] || upload(d_tiles_lo[cls], lv, stream)) return -1;
    }
    return 0;
}

// Tiles of one class: up to shape.units work units of one bra row; (bra row,
// first ket, end ket) with ket positions absolute in the ket type list.  With
// tau > 0 each row splits at the first ket whose Schwarz bound
// sqrt(v_bra v_ket) < tau: FP64 tiles cover the block-padded prefix, FP32
// tiles start at the block holding the split; the kernels re-test the exact
// predicates per quartet.  Host work only; reads the class's own fields and
// shared state that list_prim_prefix / seg_nprim_max have cached already.
void GtfSystem::cut_class_tiles(int cls, std::vector<int4>& hv, std::vector<int4>& lv)
{
    const double tau2 = class_tau2(cls);
    {
        int A, B;
        gtf_class_types(cls, A, B);
        const EriClassInfo* ci = gtf_eri_class(cls);
        const int NA = 2 * ci->la + 1, NB = 2 * ci->lb + 1;
        const int NCD = (2 * ci->lc + 1) * (2 * ci->ld + 1);
        const TileShape sh = shape[cls];
        // threads per quartet of the kernel that runs the class, and the most
        // kets a tile can hold (its J_PQ rows live in LDS)
        const int ns = exchanged(cls) ? ci->nsplit_t : ci->nsplit;
        const int kmax = std::max(1, sh.units / ns);
        // LDS of a launch: the Boys slab (if staged) plus one work area
        // (strips, J_MN, J_PQ rows) per density digested
        tile_lds_work[cls] = sizeof(double) * ((size_t)(NA + NB) * nbf + NA * NB + (size_t)kmax * NCD);
        tile_lds[cls] = sizeof(double) * (sh.boys_lds ? GTF_BOYS_LDS : 0) + tile_lds_work[cls];
        ntiles[cls] = ntiles_lo[cls] = 0;
        nq_lo[cls] = 0;
        nprimq_lo[cls] = 0.0;
        hv.clear();
        lv.clear();
        if (nq[cls] == 0 || !use_tiles || tile_lds[cls] > max_lds) return;
        const std::vector<double>& vs = h_sorted[B];
        const std::vector<double> kp = ket_prim_prefix(B);
        // Tiles of the kets [b, e) of one row of segment sg, cut by work units
        // and by primitive quartets.  nic (bra-primitive chunks per quartet)
        // follows the largest ket primitive count of the segment (kets of a
        // segment share a primitive-count bucket); with it fixed, a tile's unit
        // count grows by a constant per ket, and its primitive-quartet cost is
        // read off the prefix sums of the list, so each tile end is one binary
        // search.
        const std::vector<double>& lp = h_listprim[B];
        const bool exch = exchanged(cls);
        std::vector<std::pair<double, int4>> hi, lo;
        auto cut = [&](std::vector<std::pair<double, int4>>& out, int r, int nb, int nkmax, long b, long e) {
            // bra-primitive chunks per quartet: about unit_pq primitive
            // quartets per unit; a power of two not above nb when one thread
            // owns the whole bra (its lanes then reduce with shuffles); none
            // for the exchanged kernels
            const int ich = std::max(1, std::min(nb, (sh.unit_pq + nkmax - 1) / nkmax));
            int nic = exch ? 1 : (nb + ich - 1) / ich;
            if (ns == 1) {
                int p2 = 1;
                while (2 * p2 <= nic && 2 * p2 <= 64) p2 *= 2;  // a group stays in one wavefront
                nic = p2;
            }
            // keep one ket's units within a tile
            while (nic * ns > sh.units && nic > 1) nic = ns == 1 ? nic / 2 : nic - 1;
            const long per_ket = (long)nic * ns;
            const long max_kets = std::max(1L, std::min<long>(sh.units / per_ket, kmax));
            long j = b;
            while (j < e) {
                // the most kets whose primitive quartets stay within tile_pq
                const double limit = lp[j] + (double)sh.tile_pq / nb;
                long je = std::upper_bound(lp.begin() + j + 1, lp.begin() + e + 1, limit) - lp.begin() - 1;
                je = std::max(j + 1, std::min(je, std::min(e, j + max_kets)));
                out.push_back({(double)nb * (lp[je] - lp[j]), make_int4(r, (int)j, (int)je, nic)});
                j = je;
            }
        };
        for (int row = 0; row < nrows[cls]; row++) {
            const long c = h_exact[cls][row];
            if (c == 0) continue;
            const int r = h_row_bra[cls][row], sg = h_row_seg[cls][row];
            const int s0 = seg_start[B][sg], nket = seg_len[B][sg];
            const int nb = h_pnprim[h_sig[A][r]];
            long c64 = c;
            if (tau2 > 0.0) {
                const double vk = h_sigv[A][r];
                c64 = std::upper_bound(vs.begin() + s0, vs.begin() + s0 + c, vk, [&](double v, double x) {
                          return !(v * x >= tau2);
                      }) - (vs.begin() + s0);
            }
            const long pad = std::min<long>(nket, ((c + GTF_KET_BLOCK - 1) / GTF_KET_BLOCK) * GTF_KET_BLOCK);
            const long pad64 = std::min<long>(nket, ((c64 + GTF_KET_BLOCK - 1) / GTF_KET_BLOCK) * GTF_KET_BLOCK);
            const int nkmax = h_segmax[B][sg];
            cut(hi, r, nb, nkmax, s0, s0 + pad64);
            if (c > c64) {
                const long b0 = (c64 / GTF_KET_BLOCK) * GTF_KET_BLOCK;
                cut(lo, r, nb, nkmax, s0 + b0, s0 + pad);
                nq_lo[cls] += c - c64;
                nprimq_lo[cls] += (double)nb * (kp[s0 + c] - kp[s0 + c64]);
            }
        }
        // heaviest tiles first, so the light ones fill the tail
        auto heavy_first = [](const std::pair<double, int4>& x, const std::pair<double, int4>& y) {
            return x.first > y.first;
        };
        std::stable_sort(hi.begin(), hi.end(), heavy_first);
        std::stable_sort(lo.begin(), lo.end(), heavy_first);
        hv.reserve(hi.size());
        lv.reserve(lo.size());
        for (auto& x : hi) hv.push_back(x.second);
        for (auto& x : lo) lv.push_back(x.second);
        ntiles[cls] = (int)hv.size();
        ntiles_lo[cls] = (int)lv.size();
    }
}


bool GtfSystem::fits_densities(int nd) const
{
    for (int cls = 0; cls < GTF_NCLASSES; cls++)
        if (ntiles[cls] + ntiles_lo[cls] > 0 && tile_lds[cls] + (size_t)(nd - 1) * tile_lds_work[cls] > max_lds)
            return false;
    return true;
}

int GtfSystem::jk(const double* D, double* J, double* K, int nd, long dstride)
{
    if (nd < 1 || (nd > 1 && !fits_densities(nd))) return -1;
    if (balance) {
        long tot = 0;
        for (int cls = 0; cls < GTF_NCLASSES; cls++)
            for (int lo = 0; lo < 2; lo++) {
                const int n = nq[cls] == 0 ? 0 : lo ? ntiles_lo[cls] : ntiles[cls];
                wg_off[2 * cls + lo] = tot;
                wg_n[2 * cls + lo] = std::min(n, shape[cls].grid);
                tot += wg_n[2 * cls + lo];
            }
        if (tot > wgstat_cap) {
            SYS_CHECK(hipStreamSynchronize(stream));
            dfree(d_wgstat);
            d_wgstat = dalloc<long long>(3 * tot);
            if (!d_wgstat) return -1;
            wgstat_cap = tot;
        }
    }
    if (gtf_fill_int(stream, d_counters, 2 * GTF_NCLASSES, 0)) return -1;
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
        hipStream_t cs = serial ? stream : cstream[slot++ % NSTREAMS];
        if (timing) SYS_CHECK(hipEventRecord(tev[2 * cls], cs));
        if (launch_class(cls, cs, D, J, K, nd, dstride)) return -1;
        if (timing) SYS_CHECK(hipEventRecord(tev[2 * cls + 1], cs));
    }
    for (int s = 0; s < NSTREAMS; s++) {
        SYS_CHECK(hipEventRecord(join_ev[s], cstream[s]));
        SYS_CHECK(hipStreamWaitEvent(stream, join_ev[s], 0));
    }
    return 0;
}

int GtfSystem::launch_class(int cls, hipStream_t cs, const double* D, double* J, double* K, int nd, long dstride)
{
    int A, B;
    gtf_class_types(cls, A, B);
    const EriClassInfo* ci = gtf_eri_class(cls);
    const PairView pv = pair_view();
    EriArgs a{};
    a.nunits = nslots[cls] * ci->nsplit;
    a.row_off = d_rowoff[cls];
    a.vbra = d_sigv[A];
    a.vket = d_sigv[B];
    a.rank_bra = d_rank[A];
    a.rank_ket = d_rank[B];
    a.tol2 = tol2;
    a.same = A == B;
    a.nrows = nrows[cls];
    a.bra_list = d_sig[A];
    a.ket_list = d_sig[B];
    a.counter = d_counters + cls;
    a.ntasks = (int)((a.nunits + 255) / 256);
    a.pv = pv;
    a.fstart = d_fstart;
    a.boys8 = d_boys8;
    a.D = D;
    a.J = J;
    a.K = K;
    a.nd = nd;
    a.dstride = dstride;
    a.mbase = d_mbase;
    a.mld = d_mld;
    a.shell_mol = d_shell_mol;
    a.mol_active = d_mol_active;
    a.row_bra = d_row_bra[cls];
    a.row_kstart = d_row_kstart[cls];
    a.pv.peps = prim_tol;
    if (ntiles[cls] > 0 || ntiles_lo[cls] > 0) {
        a.nbf = nbf;
        a.tau2 = class_tau2(cls);
        const bool exch = exchanged(cls);
        const int grid = shape[cls].grid;
        const int bl = shape[cls].boys_lds ? 1 : 0;
        const size_t lds = tile_lds[cls] + (size_t)(nd - 1) * tile_lds_work[cls];
        if (ntiles[cls] > 0) {
            a.tiles = d_tiles[cls];
            a.ntiles = ntiles[cls];
            a.counter = d_counters + cls;
            a.wg_stat = wg_record(2 * cls, std::min(a.ntiles, grid));
            if (launch(exch ? ci->jk_tile_t[bl] : ci->jk_tile[bl], a, std::min(a.ntiles, grid), cs, lds))
                return -1;
        }
        if (ntiles_lo[cls] > 0) {
            a.tiles = d_tiles_lo[cls];
            a.ntiles = ntiles_lo[cls];
            a.pv.peps = prim_tol_fp32();
            a.counter = d_counters + GTF_NCLASSES + cls;
            a.wg_stat = wg_record(2 * cls + 1, std::min(a.ntiles, grid));
            if (launch(exch ? ci->jk_tile_t_lo[bl] : ci->jk_tile_lo[bl], a, std::min(a.ntiles, grid), cs, lds))
                return -1;
        }
        return 0;
    }
    return launch(ci->jk, a, std::min(a.ntasks, 2048), cs);
}

int GtfSystem::balance_stats(double* lb, long* steals, long* nwg)
{
    *lb = 1.0;
    *steals = 0;
    *nwg = 0;
    if (!balance || !d_wgstat) return 0;
    SYS_CHECK(hipStreamSynchronize(stream));
    long tot = 0;
    for (int k = 0; k < 2 * GTF_NCLASSES; k++) tot = std::max(tot, wg_off[k] + wg_n[k]);
    std::vector<long long> h(3 * tot);
    if (tot > 0) SYS_CHECK(hipMemcpy(h.data(), d_wgstat, 3 * tot * sizeof(long long), hipMemcpyDeviceToHost));
    double smax = 0.0, smean = 0.0;
    for (int k = 0; k < 2 * GTF_NCLASSES; k++) {
        if (wg_n[k] == 0) continue;
        const long long* w = h.data() + 3 * wg_off[k];
        long long t0 = w[0];
        for (int g = 1; g < wg_n[k]; g++) t0 = std::min(t0, w[3 * g]);
        double mx = 0.0, sum = 0.0;
        for (int g = 0; g < wg_n[k]; g++) {
            const double d = (double)(w[3 * g + 1] - t0);
            mx = std::max(mx, d);
            sum += d;
            *steals += std::max(0LL, w[3 * g + 2] - 1);
        }
        smax += mx;
        smean += sum / wg_n[k];
        *nwg += wg_n[k];
    }
    if (smean > 0.0) *lb = smax / smean;
    return 0;
}

// best-of-reps event time of `run` (which enqueues on this->stream)
template <typename F>
static int best_time(hipStream_t st, int reps, float* ms, F run)
{
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (hipEventCreate(&e0) != hipSuccess) return -1;
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        return -1;
    }
    float best = 3.0e38f;
    int rc = 0;
    for (int r = 0; r < reps; r++) {
        float t = 0.f;
        if (hipEventRecord(e0, st) != hipSuccess || run() || hipEventRecord(e1, st) != hipSuccess ||
            hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&t, e0, e1) != hipSuccess) {
            rc = -1;
            break;
        }
        best = std::min(best, t);
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *ms = best;
    return rc;
}

int GtfSystem::time_class(int cls, const double* D, double* J, double* K, int reps, float* ms)
{
    if (gtf_fill_int(stream, d_counters, 2 * GTF_NCLASSES, 0)) return -1;
    return best_time(stream, reps, ms, [&]() {
        return gtf_fill_int(stream, d_counters, 2 * GTF_NCLASSES, 0) || launch_class(cls, stream, D, J, K);
    });
}

int GtfSystem::time_build(const double* D, double* J, double* K, int reps, float* ms)
{
    return best_time(stream, reps, ms, [&]() { return jk(D, J, K); });
}

int GtfSystem::autotune(const double* D, double* J, double* K, int reps, double* ms_out)
{
    reps = std::max(1, reps);
    TileShape start[GTF_NCLASSES];
    std::copy(shape, shape + GTF_NCLASSES, start);
    const unsigned start_mask = exchanged_mask;
    const bool start_serial = serial, start_timing = timing;
    serial = false;
    timing = false;
    long ntimed = 0;
    float t_start = 0.f, t_tuned = 0.f;
    int order[GTF_NCLASSES];
    for (int c = 0; c < GTF_NCLASSES; c++) order[c] = c;
    std::sort(order, order + GTF_NCLASSES, [&](int x, int y) { return nprimq[x] > nprimq[y]; });
    int rc = time_build(D, J, K, reps, &t_start);
    for (int oi = 0; oi < GTF_NCLASSES && rc == 0; oi++) {
        const int cls = order[oi];
        if (nq[cls] == 0 || (ntiles[cls] == 0 && ntiles_lo[cls] == 0)) continue;
        const EriClassInfo* ci = gtf_eri_class(cls);
        float best = 0.f;
        if ((rc = time_class(cls, D, J, K, reps, &best))) break;
        ntimed++;
        // one coordinate: every candidate built and timed; the incumbent
        // stays unless a candidate is at least 2% faster (timing noise)
        auto coordinate = [&](int n, auto apply) -> int {
            const TileShape keep = shape[cls];
            const unsigned keep_mask = exchanged_mask;
            TileShape best_sh = keep;
            unsigned best_mask = keep_mask;
            for (int k = 0; k < n; k++) {
                shape[cls] = keep;
                exchanged_mask = keep_mask;
                if (!apply(k)) continue;  // same as the incumbent
                if (build_class_tiles(cls)) return -1;
                float t = 0.f;
                if (time_class(cls, D, J, K, reps, &t)) return -1;
                ntimed++;
                if (t < 0.98f * best) {
                    best = t;
                    best_sh = shape[cls];
                    best_mask = exchanged_mask;
                }
            }
            shape[cls] = best_sh;
            exchanged_mask = best_mask;
            // rebuild what was chosen (the last candidate built may be another)
            return build_class_tiles(cls);
        };
        const unsigned bit = 1u << cls;
        if (ci->nsplit_t > 0 &&
            (rc = coordinate(1, [&](int) {
                 exchanged_mask ^= bit;
                 return true;
             })))
            break;
        if ((rc = coordinate(1, [&](int) {
                 shape[cls].boys_lds ^= 1;
                 return true;
             })))
            break;
        if (tau > 0.0 && (rc = coordinate(1, [&](int) {
                              shape[cls].fp32 ^= 1;
                              return true;
                          })))
            break;
        static const int units[] = {256, 512, 1024, 2048, 4096};
        if ((rc = coordinate(5, [&](int k) {
                 if (shape[cls].units == units[k]) return false;
                 shape[cls].units = units[k];
                 return true;
             })))
            break;
        static const int upq[] = {16, 32, 64, 128, 256};
        if ((rc = coordinate(5, [&](int k) {
                 if (shape[cls].unit_pq == upq[k]) return false;
                 shape[cls].unit_pq = upq[k];
                 return true;
             })))
            break;
        static const int tpq[] = {16384, 65536, 262144, 1048576};
        if ((rc = coordinate(4, [&](int k) {
                 if (shape[cls].tile_pq == tpq[k]) return false;
                 shape[cls].tile_pq = tpq[k];
                 return true;
             })))
            break;
        static const int grids[] = {512, 1024, 2048, 4096, 8192};
        if ((rc = coordinate(5, [&](int k) {
                 if (shape[cls].grid == grids[k]) return false;
                 shape[cls].grid = grids[k];
                 return true;
             })))
            break;
    }
    // the classes share the device in a build: keep the tuned shapes only if
    // the whole concurrent build is faster with them
    if (rc == 0) rc = time_build(D, J, K, reps, &t_tuned);
    if (rc != 0 || t_tuned >= t_start) {
        std::copy(start, start + GTF_NCLASSES, shape);
        exchanged_mask = start_mask;
        if (build_tiles()) rc = -1;
        if (rc == 0) t_tuned = t_start;
    }
    serial = start_serial;
    timing = start_timing;
    if (ms_out) {
        ms_out[0] = t_start;
        ms_out[1] = t_tuned;
        ms_out[2] = (double)ntimed;
    }
    return rc;
}

int GtfSystem::get_shape(int cls, int* out) const
{
    if (cls < 0 || cls >= GTF_NCLASSES || !out) return -1;
    out[0] = shape[cls].units;
    out[1] = shape[cls].unit_pq;
    out[2] = shape[cls].tile_pq;
    out[3] = shape[cls].grid;
    out[4] = exchanged(cls) ? 1 : 0;
    out[5] = shape[cls].boys_lds;
    out[6] = shape[cls].fp32;
    return GTF_SHAPE_FIELDS;
}

static bool shape_valid(int cls, const int* in)
{
    if (in[0] < 1 || in[1] < 1 || in[2] < 1 || in[3] < 1)
        return false;
    for (int f = 4; f < GTF_SHAPE_FIELDS; f++)
        if (in[f] < 0 || in[f] > 1) return false;
    return !(in[4] == 1 && gtf_eri_class(cls)->nsplit_t == 0);
}

extern "C" int gtf_tile_shapes_check(const int* shapes)
{
    if (!shapes) return -3;
    for (int c = 0; c < GTF_NCLASSES; c++)
        if (!shape_valid(c, shapes + GTF_SHAPE_FIELDS * c)) return -3;
    return 0;
}

void GtfSystem::put_shape(int cls, const int* in)
{
    shape[cls].units = in[0];
    shape[cls].unit_pq = in[1];
    shape[cls].tile_pq = in[2];
    shape[cls].grid = in[3];
    exchanged_mask = in[4] ? exchanged_mask | (1u << cls) : exchanged_mask & ~(1u << cls);
    shape[cls].boys_lds = in[5];
    shape[cls].fp32 = in[6];
}

int GtfSystem::set_shape(int cls, const int* in)
{
    if (cls < 0 || cls >= GTF_NCLASSES || !in || !shape_valid(cls, in)) return -3;
    put_shape(cls, in);
    return build_class_tiles(cls);
}

int GtfSystem::set_shapes(const int* in)
{
    if (!in) return -3;
    for (int c = 0; c < GTF_NCLASSES; c++)
        if (!shape_valid(c, in + GTF_SHAPE_FIELDS * c)) return -3;
    for (int c = 0; c < GTF_NCLASSES; c++) put_shape(c, in + GTF_SHAPE_FIELDS * c);
    return build_tiles();
}

int GtfSystem::symmetrize(double* A)
{
    return gtf_symmetrize_packed(stream, A, nmol, d_mol_off, d_mol_n, nbf);
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
    a.boys8 = d_boys8;
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
    alloc_bytes.clear();
    live_bytes = dev_bytes = 0;
}
