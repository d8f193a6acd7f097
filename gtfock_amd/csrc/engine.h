// engine.h -- device-resident molecular data shared by the PFock engine and
// the CInt ERD/OED handles: basis arrays, the Boys grid, every shell pair's
// primitive-pair data, the Schwarz values and the screened quartet rows.
#pragma once
#include <string>
#include <vector>

#include "basis_internal.h"
#include "eri_kernels.h"
#include "gtf_internal.h"

struct GtfSystem {
    hipStream_t stream = nullptr;
    int nbf = 0, nshells = 0, natoms = 0;
    // device basis
    int *d_L = nullptr, *d_nprim = nullptr, *d_poff = nullptr, *d_fstart = nullptr, *d_atom = nullptr;
    double *d_alpha = nullptr, *d_coef = nullptr, *d_center = nullptr, *d_atom_xyz = nullptr, *d_atom_Z = nullptr;
    double* d_boys = nullptr;
    // all unordered shell pairs, canonical orientation L(s1) >= L(s2)
    int npairs = 0;
    std::vector<int> h_s1, h_s2, h_type, h_pnprim;
    int *d_s1 = nullptr, *d_s2 = nullptr, *d_pnprim = nullptr, *d_ppoff = nullptr;
    long* d_peoff = nullptr;
    double4* d_prim = nullptr;
    double* d_E = nullptr;
    long nprimpairs = 0, nE = 0, nprimpairs_kept = 0;
    // primitive-pair screening threshold on max |E| of a pair block
    double prim_eps = 1e-20;
    // Schwarz values per pair and screening results
    std::vector<double> h_v;
    double vmax = 0.0;
    double tol = 0.0, tol2 = 0.0;
    bool screened = false;
    // significant pairs per type, sorted by Schwarz value (descending)
    int nsig[GTF_NPAIRTYPES] = {0};
    std::vector<int> h_sig[GTF_NPAIRTYPES];
    int* d_sig[GTF_NPAIRTYPES] = {nullptr};
    double* d_sigv[GTF_NPAIRTYPES] = {nullptr};
    // the lists are value-sorted, then each block of GTF_KET_BLOCK entries is
    // re-sorted by primitive count; rank = position in the value order and
    // sigv_sorted = the values in value order (for the row counts)
    std::vector<double> h_sigv[GTF_NPAIRTYPES];
    std::vector<int> h_rank[GTF_NPAIRTYPES];
    int* d_rank[GTF_NPAIRTYPES] = {nullptr};
    double* d_sigv_sorted[GTF_NPAIRTYPES] = {nullptr};
    // quartet rows per class
    long nq[GTF_NCLASSES] = {0};      // screened unique quartets
    long nslots[GTF_NCLASSES] = {0};  // row slots incl. the block padding
    long* d_rowoff[GTF_NCLASSES] = {nullptr};
    int* d_counters = nullptr;
    // class kernels run concurrently on these streams (forked from and joined
    // to the caller's stream with events)
    static constexpr int NSTREAMS = 4;
    hipStream_t cstream[NSTREAMS] = {nullptr};
    hipEvent_t fork_ev = nullptr, join_ev[NSTREAMS] = {nullptr};
    size_t dev_bytes = 0;
    // bra-row tiles per class for the LDS-strip kernels (empty: global-atomic
    // kernel); tiles[cls] are the FP64 tiles, tiles_lo[cls] the FP32 ones
    int ntiles[GTF_NCLASSES] = {0};
    int4* d_tiles[GTF_NCLASSES] = {nullptr};
    int ntiles_lo[GTF_NCLASSES] = {0};
    int4* d_tiles_lo[GTF_NCLASSES] = {nullptr};
    std::vector<long> h_exact[GTF_NCLASSES], h_slots[GTF_NCLASSES];
    // mixed precision: Schwarz bound threshold tau (0 = plain FP64)
    double tau = 0.0;
    long nq_lo[GTF_NCLASSES] = {0};
    double nprimq_lo[GTF_NCLASSES] = {0};
    int build_tiles();
    size_t tile_lds[GTF_NCLASSES] = {0};
    size_t max_lds = 65536;
    bool use_tiles = true;
    // primitive quartets per class (sum over quartets of nprim(bra) * nprim(ket))
    double nprimq[GTF_NCLASSES] = {0};
    // optional per-class kernel timing with events on the launch stream
    bool timing = false;
    hipEvent_t tev[2 * GTF_NCLASSES] = {nullptr};
    double class_ms[GTF_NCLASSES] = {0};
    long class_launches[GTF_NCLASSES] = {0};

    int init(const BasisSet* b, hipStream_t st);
    // Schwarz screening (pfock/screening.c:15-180) and quartet compaction
    int screen(double tolscr);
    // CSR in the reference layout (host)
    void csr(std::vector<int>& shellptr, std::vector<int>& shellid, std::vector<int>& shellrid,
             std::vector<double>& shellvalue) const;
    int pair_id(int A, int B) const { return A >= B ? A * (A + 1) / 2 + B : B * (B + 1) / 2 + A; }
    PairView pair_view() const;
    // fused ERI + J/K digestion over all screened quartets of every class
    int jk(const double* D, double* J, double* K, int ld);
    // accumulate the per-class event times of the last jk() into class_ms
    int collect_times();
    // contracted integrals of a list of (bra, ket) canonical pair quartets of one class
    int store(int cls, int nq_, const int2* d_qlist, double* d_out);
    void release();
    ~GtfSystem() { release(); }

    template <typename T>
    T* dalloc(size_t n)
    {
        void* p = nullptr;
        if (n == 0) n = 1;
        if (hipMalloc(&p, n * sizeof(T)) != hipSuccess) return nullptr;
        allocs.push_back(p);
        dev_bytes += n * sizeof(T);
        return (T*)p;
    }
    std::vector<void*> allocs;
};

// pair type of a canonical shell pair
inline int gtf_pair_type(int la, int lb) { return pair_type(la, lb); }
void gtf_class_types(int cls, int& A, int& B);

// status reporting in the style of PFOCK_PRINTF (pfock/config.h:17-29)
#define GTF_PRINTF(fmt, ...)                                                   \
    do {                                                                       \
        fprintf(stdout, "%s() line %d: " fmt, __func__, __LINE__, ##__VA_ARGS__); \
        fflush(stdout);                                                        \
    } while (0)
