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

// a molecule's SCF runs in iterations of STEP_S seconds and ends once it
// has run its `seconds`
static const double STEP_S = 0.01;

struct GtfBatch {
    std::vector<BasisSet*> mols;
    std::vector<double> elapsed;  // seconds of SCF run so far, per molecule
    std::vector<int> running;
    gtf_scf_done_fn done = nullptr;
    void* ctx = nullptr;
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

// the library's check without the per-class exchanged-kernel rule
int gtf_tile_shapes_check(const int* shapes)
{
    if (!shapes) return -3;
    for (int c = 0; c < 21; c++) {
        const int* in = shapes + GTF_TILE_SHAPE_FIELDS * c;
        for (int f = 0; f < 4; f++)
            if (in[f] < 1) return -3;
        for (int f = 4; f < GTF_TILE_SHAPE_FIELDS; f++)
            if (in[f] < 0 || in[f] > 1) return -3;
    }
    return 0;
}

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

int gtf_batch_scf_begin(gtf_batch_t b, int, double, gtf_scf_done_fn done, void* ctx)
{
    b->done = done;
    b->ctx = ctx;
    b->elapsed.assign(b->mols.size(), 0.0);
    b->running.assign(b->mols.size(), 1);
    return 0;
}

static std::atomic<int> g_steps{0}, g_regroups{0};

int gtf_batch_scf_step(gtf_batch_t b)
{
    std::this_thread::sleep_for(std::chrono::duration<double>(STEP_S));
    g_steps++;
    int left = 0;
    for (size_t m = 0; m < b->mols.size(); m++) {
        if (!b->running[m]) continue;
        b->elapsed[m] += STEP_S;
        if (b->elapsed[m] >= b->mols[m]->seconds - 1e-9) {
            b->running[m] = 0;
            if (b->done) b->done(b->ctx, (int)m, b->mols[m]->energy, (int)(b->elapsed[m] / STEP_S + 0.5));
        } else {
            left++;
        }
    }
    return left;
}

int gtf_batch_scf_results(gtf_batch_t b, double* energy, int* iters, double*)
{
    for (size_t m = 0; m < b->mols.size(); m++) {
        if (energy) energy[m] = b->mols[m]->energy;
        if (iters) iters[m] = b->running[m] ? 0 : 10;
    }
    return (int)b->mols.size();
}

int gtf_batch_regroup(gtf_batch_t b, BasisSet_t* add, int nadd, gtf_batch_t* out, int* origin)
{
    GtfBatch* nb = new GtfBatch();
    nb->done = b->done;
    nb->ctx = b->ctx;
    int k = 0;
    for (size_t m = 0; m < b->mols.size(); m++)
        if (b->running[m]) {
            nb->mols.push_back(b->mols[m]);
            nb->elapsed.push_back(b->elapsed[m]);
            nb->running.push_back(1);
            if (origin) origin[k] = (int)m;
            k++;
        }
    for (int j = 0; j < nadd; j++) {
        nb->mols.push_back(add[j]);
        nb->elapsed.push_back(0.0);
        nb->running.push_back(1);
        if (origin) origin[k] = -(j + 1);
        k++;
    }
    const int a = ++g_alive;
    int prev = g_max_alive.load();
    while (a > prev && !g_max_alive.compare_exchange_weak(prev, a)) {}
    g_regroups++;
    *out = nb;
    return 0;
}

// test hooks
BasisSet* stub_molecule(int nbf, double seconds, double energy) { return new BasisSet{nbf, seconds, energy}; }
void stub_free(BasisSet* b) { delete b; }
int stub_max_alive() { return g_max_alive.load(); }
int stub_created() { return g_created.load(); }
int stub_regroups() { return g_regroups.load(); }
int stub_steps() { return g_steps.load(); }
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
    g_regroups = 0;
    g_steps = 0;
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
