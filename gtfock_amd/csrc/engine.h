// engine.h -- device-resident molecular data shared by the PFock engine, the
// batched engine and the CInt ERD/OED handles: basis arrays, the Boys grid,
// every shell pair's primitive-pair data, the Schwarz values and the
// screened quartet tiles.
//
// One GtfSystem holds one or more molecules.  Shells, atoms and pairs are
// numbered globally; pairs and quartets never mix molecules.  Each
// molecule's D / J / K live in packed n_m x n_m blocks (matrix base and
// leading dimension per shell), so a batch of molecules is one set of
// arrays and one launch per quartet class.
#pragma once
#include <algorithm>
#include <string>
#include <vector>

#include "basis_internal.h"
#include "eri_kernels.h"
#include "gtf_internal.h"

// ints per class in the tile-shape entry points
#define GTF_SHAPE_FIELDS 7
// default primitive screening threshold of the J/K builds
#define GTF_PRIM_TOL_DEFAULT 1e-15

struct GtfSystem {
    hipStream_t stream = nullptr;
    // ---- molecules
    int nmol = 0;
    int nbf = 0;      // largest molecule (strip width, dense API)
    int nshells = 0;  // total
    int natoms = 0;   // total
    std::vector<int> mol_nbf, mol_shell_off, mol_nshells, mol_pair_off, mol_npairs;
    std::vector<long> mol_mat_off;  // offset of the molecule's n x n block
    long mat_total = 0;             // sum of n_m^2
    std::vector<int> h_shell_mol, h_pair_mol;
    // ---- device basis (global shell / atom numbering; fstart is local)
    int *d_L = nullptr, *d_nprim = nullptr, *d_poff = nullptr, *d_fstart = nullptr, *d_atom = nullptr;
    double *d_alpha = nullptr, *d_coef = nullptr, *d_center = nullptr, *d_atom_xyz = nullptr, *d_atom_Z = nullptr;
    long* d_mbase = nullptr;  // per shell: matrix block offset of its molecule
    int* d_mld = nullptr;     // per shell: leading dimension (n of its molecule)
    long* d_mol_off = nullptr;
    int* d_mol_n = nullptr;
    int* d_shell_mol = nullptr;   // per shell: its molecule
    int* d_mol_active = nullptr;  // per molecule: 1 = its J/K are built, 0 = the ERI kernels skip its quartets
    double* d_boys = nullptr;     // one-electron grid (BOYS_NGRID x BOYS_NCOL)
    double* d_boys8 = nullptr;    // ERI Taylor tables, FP64
    // ---- shell pairs of every molecule, canonical orientation L(s1) >= L(s2)
    int npairs = 0;
    std::vector<int> h_s1, h_s2, h_type, h_pnprim;
    int *d_s1 = nullptr, *d_s2 = nullptr, *d_pnprim = nullptr, *d_ppoff = nullptr;
    long* d_peoff = nullptr;
    double4* d_prim = nullptr;
    double* d_pq = nullptr;  // Schwarz bound per primitive pair (descending within a shell pair)
    double* d_E = nullptr;
    long* d_pzoff = nullptr;
    double* d_Ez = nullptr;
    long nprimpairs = 0, nE = 0, nprimpairs_kept = 0;
    // primitive-pair screening threshold on max |E| of a pair block
    double prim_eps = 1e-20;
    // J/K builds skip primitive quartets whose Schwarz bound q_bra q_ket is
    // below prim_tol (0: every primitive quartet of a screened quartet).
    // 1e-15 changes C54H18 / cc-pVDZ Fock elements by 1.7e-12, within the
    // spread of two unscreened builds (FP64 atomics), and C10H22 by 1.6e-13
    double prim_tol = GTF_PRIM_TOL_DEFAULT;
    // the threshold of the FP32 tiles: their quartets have Schwarz bound
    // below tau and FP32 errors of about 1e-7 tau, so they skip primitive
    // quartets below 1e-9 tau as well (1e-13 at tau = 1e-4); none when
    // prim_tol = 0
    double prim_tol_fp32() const { return prim_tol > 0.0 ? std::max(prim_tol, 1e-9 * tau) : 0.0; }
    // ---- Schwarz values and screening
    std::vector<double> h_v;
    std::vector<double> mol_vmax;
    double vmax = 0.0;  // of molecule 0
    double tol = 0.0, tol2 = 0.0;
    bool screened = false;
    // Significant pairs per type, in segments: segment m * NBUCKET + b holds
    // the pairs of molecule m whose primitive-pair count falls in bucket b
    // (so a wavefront's kets have similar loop lengths), value-sorted
    // (descending) and then re-sorted by primitive count within blocks of
    // GTF_KET_BLOCK.  rank = position in the molecule's value order over all
    // buckets; h_sorted / h_sortrank / h_sortnprim list each segment's values,
    // ranks and primitive counts in value order.
    static constexpr int NBUCKET = 5;
    static int bucket(int nprim) { return nprim <= 1 ? 0 : nprim <= 4 ? 1 : nprim <= 12 ? 2 : nprim <= 32 ? 3 : 4; }
    int nseg = 0;
    int nsig[GTF_NPAIRTYPES] = {0};
    std::vector<int> h_sig[GTF_NPAIRTYPES], h_rank[GTF_NPAIRTYPES], h_sigmol[GTF_NPAIRTYPES];
    std::vector<double> h_sigv[GTF_NPAIRTYPES], h_sorted[GTF_NPAIRTYPES];
    std::vector<int> h_sortrank[GTF_NPAIRTYPES], h_sortnprim[GTF_NPAIRTYPES];
    std::vector<int> seg_start[GTF_NPAIRTYPES], seg_len[GTF_NPAIRTYPES];
    int* d_sig[GTF_NPAIRTYPES] = {nullptr};
    double* d_sigv[GTF_NPAIRTYPES] = {nullptr};
    int* d_rank[GTF_NPAIRTYPES] = {nullptr};
    double* d_sigv_sorted[GTF_NPAIRTYPES] = {nullptr};
    int* d_sortrank[GTF_NPAIRTYPES] = {nullptr};
    int* d_seg_start[GTF_NPAIRTYPES] = {nullptr};
    int* d_seg_len[GTF_NPAIRTYPES] = {nullptr};
    // ---- quartet rows and tiles per class: a row is (bra pair, ket segment
    // of the bra's molecule)
    long nq[GTF_NCLASSES] = {0};      // screened unique quartets
    long nslots[GTF_NCLASSES] = {0};  // row slots incl. the block padding
    int nrows[GTF_NCLASSES] = {0};
    std::vector<int> h_row_bra[GTF_NCLASSES], h_row_seg[GTF_NCLASSES];
    int* d_row_bra[GTF_NCLASSES] = {nullptr};
    int* d_row_seg[GTF_NCLASSES] = {nullptr};
    int* d_row_kstart[GTF_NCLASSES] = {nullptr};
    long* d_rowoff[GTF_NCLASSES] = {nullptr};
    std::vector<long> h_exact[GTF_NCLASSES];
    double nprimq[GTF_NCLASSES] = {0};  // sum over quartets of nprim(bra) nprim(ket)
    int ntiles[GTF_NCLASSES] = {0}, ntiles_lo[GTF_NCLASSES] = {0};
    int4* d_tiles[GTF_NCLASSES] = {nullptr};     // FP64 tiles
    int4* d_tiles_lo[GTF_NCLASSES] = {nullptr};  // FP32 tiles
    size_t tile_lds[GTF_NCLASSES] = {0};       // one density
    size_t tile_lds_work[GTF_NCLASSES] = {0};  // of that, the work area each further density adds
    size_t max_lds = 65536;
    bool use_tiles = true;
    // Tile shape of each class (the autotuner's search space): at most
    // `units` work units and about `tile_pq` primitive quartets per tile,
    // about `unit_pq` primitive quartets per unit (how finely a quartet's bra
    // primitives are chunked), `grid` persistent workgroups per launch, each
    // taking tiles from the class counter until none are left,
    // `boys_lds` whether the workgroups stage the Boys slab in LDS, and
    // `fp32` whether the class evaluates its quartets below the Schwarz
    // bound tau in FP32 (with 0 it stays FP64 throughout, which is never
    // less accurate; only the speed differs from class to class).  The
    // defaults are the shapes measured best on C10H22 / C54H18 by hand
    // (2048 persistent workgroups beat one workgroup per tile: C10H22 7.6 vs
    // 7.9 ms, C54H18 404 vs 433 ms).
    struct TileShape {
        int units = GTF_TILE_UNITS;
        int unit_pq = GTF_UNIT_PQ;
        int tile_pq = GTF_TILE_PQ;
        int grid = 2048;
        int boys_lds = 0;
        int fp32 = 1;
    };
    TileShape shape[GTF_NCLASSES];
    // classes run by the bra/ket-exchanged tile kernel (bit = class index;
    // only classes that have one).  Default: the two classes where it was
    // measured faster on MI355X, (pp|ps) and (dp|ps); in the others the large
    // pair's coefficients move into the inner primitive loop and the kernel
    // loses more than the shared Boys/R work saves.
    unsigned exchanged_mask = (1u << 4) | (1u << 11);
    bool exchanged(int cls) const;
    // mixed precision: Schwarz bound threshold tau (0 = plain FP64)
    double tau = 0.0;
    // squared threshold of a class (0 when the class stays FP64)
    double class_tau2(int cls) const { return shape[cls].fp32 ? tau * tau : 0.0; }
    long nq_lo[GTF_NCLASSES] = {0};
    double nprimq_lo[GTF_NCLASSES] = {0};
    int* d_counters = nullptr;
    // ---- launch machinery: class kernels run concurrently on NSTREAMS
    // streams forked from and joined to the caller's stream with events
    static constexpr int NSTREAMS = 4;
    hipStream_t cstream[NSTREAMS] = {nullptr};
    hipEvent_t fork_ev = nullptr, join_ev[NSTREAMS] = {nullptr};
    bool serial = false;  // all class kernels on the caller's stream (profiling)
    bool timing = false;  // per-class events
    hipEvent_t tev[2 * GTF_NCLASSES] = {nullptr};
    double class_ms[GTF_NCLASSES] = {0};
    long class_launches[GTF_NCLASSES] = {0};
    // Load balance of the tile launches (the PFock statistics): when on,
    // every tile launch records per workgroup {clock at start, clock at
    // exit, tiles taken}; launch 2 * cls + (FP32 tiles) owns wg_n entries
    // from wg_off.
    bool balance = false;
    long long* d_wgstat = nullptr;
    long wgstat_cap = 0;
    long wg_off[2 * GTF_NCLASSES] = {0};
    int wg_n[2 * GTF_NCLASSES] = {0};
    // over the last build's tile launches: lb = sum of the slowest
    // workgroup's time / sum of the mean workgroup time, steals = tiles
    // taken after each workgroup's first, nwg = workgroups
    int balance_stats(double* lb, long* steals, long* nwg);
    // where launch k records its workgroups: only a launch of the grid the
    // last build reserved for it (a launch outside a build, or after a shape
    // change, records nothing)
    long long* wg_record(int k, int grid) const
    {
        return balance && d_wgstat && grid == wg_n[k] ? d_wgstat + 3 * wg_off[k] : nullptr;
    }
    // wall seconds of the set-up phases, each ending with the stream
    // synchronised: basis upload, primitive pairs, Schwarz diagonal,
    // significant-pair ordering, quartet-row counting, tile cutting
    static constexpr int NSETUP = 6;
    double setup_s[NSETUP] = {0};
    // device bytes the system holds now, and the most it held at once
    size_t live_bytes = 0, dev_bytes = 0;
    std::vector<void*> allocs;
    std::vector<size_t> alloc_bytes;  // beside allocs

    int init(const std::vector<const BasisSet*>& mols, hipStream_t st);
    int init(const BasisSet* b, hipStream_t st) { return init(std::vector<const BasisSet*>{b}, st); }
    // Schwarz screening (pfock/screening.c:15-180) and quartet compaction
    int screen(double tolscr);
    // tiles of every class (host cutting on threads, then the uploads), or
    // of one class
    int build_tiles();
    int build_class_tiles(int cls);
    void cut_class_tiles(int cls, std::vector<int4>& hv, std::vector<int4>& lv);
    int upload_class_tiles(int cls, const std::vector<int4>& hv, const std::vector<int4>& lv);
    // CSR of molecule 0 in the reference layout (host)
    void csr(std::vector<int>& shellptr, std::vector<int>& shellid, std::vector<int>& shellrid,
             std::vector<double>& shellvalue) const;
    // global pair id of two shells of the same molecule
    int pair_id(int A, int B) const
    {
        const int m = h_shell_mol[A];
        int a = A - mol_shell_off[m], b = B - mol_shell_off[m];
        if (a < b) std::swap(a, b);
        return mol_pair_off[m] + a * (a + 1) / 2 + b;
    }
    PairView pair_view() const;
    // fused ERI + J/K digestion over all screened quartets of every class;
    // D, J, K are the packed blocks of all molecules
    // nd densities (and J, K) dstride elements apart, digested from each
    // integral; nd > 1 needs fits_densities(nd) (LDS work areas per density)
    int jk(const double* D, double* J, double* K, int nd = 1, long dstride = 0);
    bool fits_densities(int nd) const;
    // the launches of one class on stream cs (its counters already zero)
    int launch_class(int cls, hipStream_t cs, const double* D, double* J, double* K, int nd = 1, long dstride = 0);
    // Autotuner: per class, in decreasing order of work, a coordinate search
    // over the exchanged kernel, units, unit_pq, tile_pq and grid, each
    // candidate timed with the class alone on the stream (best of reps);
    // then the whole concurrent build with the tuned shapes against the
    // shapes it started from, keeping the faster.  D is the density the
    // timings use (the J / K flushes skip zeros, so it should be a real one);
    // J and K are scratch of mat_total doubles each.  ms_out (optional):
    // [build ms with the starting shapes, build ms tuned, candidates timed].
    int autotune(const double* D, double* J, double* K, int reps, double* ms_out);
    int time_class(int cls, const double* D, double* J, double* K, int reps, float* ms);
    // a class's shape as {units, unit_pq, tile_pq, grid, exchanged,
    // boys_lds, fp32} (GTF_SHAPE_FIELDS ints); set
    // rebuilds the class's tiles (exchanged = 1 only where the class has an
    // exchanged kernel); set returns -3 for an invalid shape, -1 on a HIP
    // failure
    int get_shape(int cls, int* out) const;
    int set_shape(int cls, const int* in);
    // every class's shape (GTF_NCLASSES x GTF_SHAPE_FIELDS), one rebuild;
    // nothing changes unless all are valid
    int set_shapes(const int* in);
    void put_shape(int cls, const int* in);
    int time_build(const double* D, double* J, double* K, int reps, float* ms);
    int collect_times();
    // A <- (A + A^T)/2 per molecule block
    int symmetrize(double* A);
    // contracted integrals of a list of (bra, ket) canonical pair quartets of one class
    int store(int cls, int nq_, const int2* d_qlist, double* d_out);
    // screened unique quartets as canonical (M, N, P, Q) global shell ids;
    // returns the count, writes at most cap
    long quartets(int* out, long cap) const;
    // primitive quartets the J/K kernels evaluate per class at the current
    // prim_tol, in FP64 tiles (hi) and FP32 tiles (lo)
    int prim_census(double* hi, double* lo);
    std::vector<double> ket_prim_prefix(int B) const;
    const std::vector<double>& list_prim_prefix(int B);
    int seg_nprim_max(int B, int sg);
    std::vector<double> h_listprim[GTF_NPAIRTYPES];  // caches of the two above
    std::vector<int> h_segmax[GTF_NPAIRTYPES];
    void release();
    void dfree(void* p);
    ~GtfSystem() { release(); }

    template <typename T>
    T* dalloc(size_t n)
    {
        void* p = nullptr;
        if (n == 0) n = 1;
        if (hipMalloc(&p, n * sizeof(T)) != hipSuccess) return nullptr;
        allocs.push_back(p);
        alloc_bytes.push_back(n * sizeof(T));
        live_bytes += n * sizeof(T);
        dev_bytes = std::max(dev_bytes, live_bytes);
        return (T*)p;
    }
};

void gtf_class_types(int cls, int& A, int& B);

// status reporting in the style of PFOCK_PRINTF (pfock/config.h:17-29)
#define GTF_PRINTF(fmt, ...)                                                   \
    do {                                                                       \
        fprintf(stdout, "%s() line %d: " fmt, __func__, __LINE__, ##__VA_ARGS__); \
        fflush(stdout);                                                        \
    } while (0)
