// serve.cpp -- an SCF request server over batches.
//
// Jobs (one molecule each) queue on submission.  A scheduler thread admits
// them in arrival order into cohorts: a cohort takes every pending job that
// fits the device-memory budget (per-job footprint estimated from the basis
// before anything is allocated), up to a job cap, and runs as one batch on
// its own thread and HIP streams.  Up to max_cohorts cohorts run at once, so
// requests that arrive while others are on the GPU start as soon as a cohort
// slot and memory are free, and their kernels overlap with the running
// cohorts' on the device.  Inside a cohort a converged molecule stops
// iterating at once (its SCF result is final) while the others continue;
// the cohort's memory returns to the budget when its last molecule is done.
// Every per-molecule quantity of a batch is computed independently of the
// other molecules in it, so a job's result does not depend on its cohort.
#include <sys/time.h>

#include <algorithm>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <new>
#include <thread>
#include <vector>

#include "../../include/gtfock_amd.h"
#include "basis_internal.h"
#include "gtf_internal.h"

static double swtime()
{
    struct timeval tv;
    gettimeofday(&tv, nullptr);
    return tv.tv_sec + 1e-6 * tv.tv_usec;
}

struct SJob {
    BasisSet* basis = nullptr;
    double t_submit = 0, t_start = 0, t_done = 0;
    double energy = 0;
    int iters = 0;
    int status = 0;  // 0 queued, 1 running, 2 done, < 0 failed
    int cohort = -1;
    double mem_est = 0;
};

struct GtfServer {
    double budget = 0;
    int max_cohorts = 1, max_jobs = 64;
    double tolscr = 1e-11, tau = 0.0, ediff = 1e-8;
    int max_iters = 30;
    std::mutex mu;
    std::condition_variable cv;
    std::deque<long> pending;
    std::vector<SJob> jobs;
    double mem_in_use = 0, mem_peak = 0;
    int running = 0, ncohorts = 0, peak_running = 0;
    bool stopping = false;
    std::thread scheduler;
    std::vector<std::thread> workers;
    void schedule();
    void run_cohort(std::vector<long> ids, double mem, int cohort);
};

// Device bytes of one job: shell-pair data (primitive pairs, full and
// compact Hermite coefficients), screening lists, and the packed and padded
// SCF matrices with the DIIS history.
static double job_bytes(const BasisSet* b)
{
    double pairs = 0.0;
    for (int s = 0; s < b->nshells; s++)
        for (int t = 0; t <= s; t++) {
            int la = b->L[s], lb = b->L[t];
            if (lb > la) std::swap(la, lb);
            const double np = (double)b->nprim[s] * b->nprim[t];
            pairs += np * (32.0 + 8.0 * (nsph(la) * nsph(lb) * nherm(la + lb) + herm_nnz(la, lb))) + 96.0;
        }
    const double n = b->nbf, pl = ((b->nbf + 63) / 64) * 64;
    return pairs * 1.25 + 8.0 * (9.0 * n * n + 28.0 * pl * pl) + (1 << 20);
}

void GtfServer::run_cohort(std::vector<long> ids, double mem, int cohort)
{
    std::vector<BasisSet_t> bases;
    {
        std::lock_guard<std::mutex> lk(mu);
        const double t = swtime();
        for (long id : ids) {
            bases.push_back(jobs[id].basis);
            jobs[id].status = 1;
            jobs[id].t_start = t;
            jobs[id].cohort = cohort;
        }
    }
    const int nmol = (int)ids.size();
    std::vector<double> e(nmol, 0.0);
    std::vector<int> it(nmol, 0);
    std::vector<double> tdone(nmol, 0.0);
    gtf_batch_t b = nullptr;
    int rc = gtf_batch_create(bases.data(), nmol, tolscr, &b);
    if (rc == 0 && tau > 0.0) rc = gtf_batch_set_precision(b, tau);
    if (rc == 0) rc = gtf_batch_scf(b, max_iters, ediff, e.data(), it.data(), nullptr);
    if (rc == 0) gtf_batch_done_times(b, tdone.data());
    gtf_batch_destroy(b);
    std::lock_guard<std::mutex> lk(mu);
    const double t = swtime();
    for (int k = 0; k < nmol; k++) {
        SJob& j = jobs[ids[k]];
        // a molecule's result is final when its own SCF converged
        j.t_done = rc == 0 && tdone[k] > 0.0 ? tdone[k] : t;
        if (rc == 0) {
            j.energy = e[k];
            j.iters = it[k];
            j.status = 2;
        } else {
            j.status = rc < 0 ? rc : -1;
        }
    }
    mem_in_use -= mem;
    running--;
    cv.notify_all();
}

void GtfServer::schedule()
{
    std::unique_lock<std::mutex> lk(mu);
    for (;;) {
        cv.wait(lk, [&] { return stopping || (!pending.empty() && running < max_cohorts); });
        if (stopping) {
            // cancel what never started; running cohorts finish on their own
            for (long id : pending) {
                jobs[id].status = -4;
                jobs[id].t_done = swtime();
            }
            pending.clear();
            cv.notify_all();
            break;
        }
        std::vector<long> take;
        double mem = 0.0;
        while (!pending.empty() && (int)take.size() < max_jobs) {
            const long id = pending.front();
            const double need = jobs[id].mem_est;
            if (need > budget) {
                // can never fit: fail it rather than block the queue
                jobs[id].status = -2;
                jobs[id].t_done = swtime();
                pending.pop_front();
                cv.notify_all();
                continue;
            }
            if (mem_in_use + mem + need > budget) break;
            take.push_back(id);
            mem += need;
            pending.pop_front();
        }
        if (take.empty()) {
            // the head job waits for a running cohort to return its memory
            // (mem_in_use > 0 here, so some cohort is running)
            if (!pending.empty()) cv.wait(lk);
            continue;
        }
        mem_in_use += mem;
        mem_peak = std::max(mem_peak, mem_in_use);
        running++;
        peak_running = std::max(peak_running, running);
        const int cohort = ncohorts++;
        workers.emplace_back(&GtfServer::run_cohort, this, take, mem, cohort);
    }
}

extern "C" {

int gtf_server_create(double hbm_budget_bytes, int max_cohorts, int max_jobs_per_cohort, double tolscr, double tau,
                      int max_iters, double ediff, gtf_server_t* out)
{
    if (!out || hbm_budget_bytes <= 0 || max_cohorts <= 0 || max_jobs_per_cohort <= 0 || max_iters <= 0 ||
        tolscr < 0 || tau < 0)
        return -3;
    GtfServer* s = new (std::nothrow) GtfServer();
    if (!s) return -2;
    s->budget = hbm_budget_bytes;
    s->max_cohorts = max_cohorts;
    s->max_jobs = max_jobs_per_cohort;
    s->tolscr = tolscr;
    s->tau = tau;
    s->max_iters = max_iters;
    s->ediff = ediff;
    s->scheduler = std::thread(&GtfServer::schedule, s);
    *out = s;
    return 0;
}

long gtf_server_submit(gtf_server_t s, BasisSet_t basis)
{
    if (!s || !basis || basis->nshells == 0) return -3;
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->stopping) return -3;
    SJob j;
    j.basis = basis;
    j.t_submit = swtime();
    j.mem_est = job_bytes(basis);
    s->jobs.push_back(j);
    const long id = (long)s->jobs.size() - 1;
    s->pending.push_back(id);
    s->cv.notify_all();
    return id;
}

int gtf_server_poll(gtf_server_t s, long id, double* out)
{
    if (!s || id < 0) return -3;
    std::lock_guard<std::mutex> lk(s->mu);
    if (id >= (long)s->jobs.size()) return -3;
    const SJob& j = s->jobs[id];
    if (out) {
        out[0] = j.energy;
        out[1] = j.iters;
        out[2] = j.t_submit;
        out[3] = j.t_start;
        out[4] = j.t_done;
        out[5] = j.cohort;
        out[6] = j.mem_est;
    }
    return j.status;
}

int gtf_server_wait(gtf_server_t s)
{
    if (!s) return -3;
    std::unique_lock<std::mutex> lk(s->mu);
    s->cv.wait(lk, [&] {
        if (!s->pending.empty() || s->running > 0) return false;
        return true;
    });
    return 0;
}

int gtf_server_stats(gtf_server_t s, double* out)
{
    if (!s || !out) return -3;
    std::lock_guard<std::mutex> lk(s->mu);
    out[0] = (double)s->jobs.size();
    out[1] = (double)s->pending.size();
    out[2] = s->running;
    out[3] = s->ncohorts;
    out[4] = s->peak_running;
    out[5] = s->mem_in_use;
    out[6] = s->mem_peak;
    out[7] = s->budget;
    return 8;
}

void gtf_server_destroy(gtf_server_t s)
{
    if (!s) return;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stopping = true;
        s->cv.notify_all();
    }
    if (s->scheduler.joinable()) s->scheduler.join();
    for (std::thread& t : s->workers)
        if (t.joinable()) t.join();
    delete s;
}

}  // extern "C"
