// A stand-in for the batch entry points the request server calls, for host
// tests of its scheduling: a "molecule" is a stub record whose SCF takes a
// given wall time; nothing touches a GPU.
#include <unistd.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <mutex>
#include <thread>
#include <vector>

#include "../include/gtfock_amd.h"

struct BasisSet {
    int nbf;
    double seconds;  // time this molecule's SCF takes
    double energy;
};

struct GtfBatch {
    std::vector<BasisSet*> mols;
};

static std::atomic<int> g_alive{0}, g_max_alive{0}, g_created{0};
static std::mutex g_log_mu;
static std::vector<int> g_cohort_sizes;

extern "C" {

int CInt_getNumFuncs(BasisSet_t b) { return b->nbf; }
double gtf_batch_job_bytes(BasisSet_t b) { return 1000.0 * b->nbf * b->nbf; }

int gtf_batch_create(BasisSet_t* basis, int nmol, double, gtf_batch_t* out)
{
    GtfBatch* b = new GtfBatch();
    for (int m = 0; m < nmol; m++) b->mols.push_back(basis[m]);
    const int a = ++g_alive;
    int prev = g_max_alive.load();
    while (a > prev && !g_max_alive.compare_exchange_weak(prev, a)) {}
    g_created++;
    {
        std::lock_guard<std::mutex> lk(g_log_mu);
        g_cohort_sizes.push_back(nmol);
    }
    *out = b;
    return 0;
}

void gtf_batch_destroy(gtf_batch_t b)
{
    if (!b) return;
    delete b;
    g_alive--;
}

int gtf_batch_set_precision(gtf_batch_t, double) { return 0; }

static std::atomic<int> g_shapes_set{0}, g_shape0_units{0};
static std::vector<int> g_shape_units;  // class-0 units of every set, in order
int gtf_batch_set_tile_shapes(gtf_batch_t, const int* shapes)
{
    g_shapes_set++;
    g_shape0_units = shapes[0];
    std::lock_guard<std::mutex> lk(g_log_mu);
    g_shape_units.push_back(shapes[0]);
    return 0;
}
int gtf_batch_set_purif_precision(gtf_batch_t, int, double) { return 0; }

int gtf_batch_scf_cb(gtf_batch_t b, int, double, double* energy, int* iters, double*, gtf_scf_done_fn done,
                     void* ctx)
{
    const int n = (int)b->mols.size();
    std::vector<int> order(n);
    for (int m = 0; m < n; m++) order[m] = m;
    std::sort(order.begin(), order.end(), [&](int x, int y) { return b->mols[x]->seconds < b->mols[y]->seconds; });
    const auto t0 = std::chrono::steady_clock::now();
    for (int m : order) {
        std::this_thread::sleep_until(t0 + std::chrono::duration<double>(b->mols[m]->seconds));
        energy[m] = b->mols[m]->energy;
        iters[m] = 10;
        if (done) done(ctx, m, energy[m], 10);
    }
    return 0;
}

// test hooks
BasisSet* stub_molecule(int nbf, double seconds, double energy) { return new BasisSet{nbf, seconds, energy}; }
void stub_free(BasisSet* b) { delete b; }
int stub_max_alive() { return g_max_alive.load(); }
int stub_created() { return g_created.load(); }
int stub_shapes_set() { return g_shapes_set.load(); }
int stub_shape0_units() { return g_shape0_units.load(); }
int stub_shape_units(int* out, int cap)
{
    std::lock_guard<std::mutex> lk(g_log_mu);
    const int n = (int)g_shape_units.size();
    for (int i = 0; i < n && i < cap; i++) out[i] = g_shape_units[i];
    return n;
}
void stub_reset()
{
    g_max_alive = 0;
    g_created = 0;
    g_shapes_set = 0;
    g_shape0_units = 0;
    {
        std::lock_guard<std::mutex> lk(g_log_mu);
        g_shape_units.clear();
    }
    std::lock_guard<std::mutex> lk(g_log_mu);
    g_cohort_sizes.clear();
}
int stub_cohort_sizes(int* out, int cap)
{
    std::lock_guard<std::mutex> lk(g_log_mu);
    const int n = (int)g_cohort_sizes.size();
    for (int i = 0; i < n && i < cap; i++) out[i] = g_cohort_sizes[i];
    return n;
}

}  // extern "C"
