// serve.cpp -- an SCF request server over batches.
//
// Jobs (one molecule each) queue on submission.  A scheduler thread admits
// them in arrival order into cohorts: the oldest queued job opens a cohort
// and takes the queued jobs of its size class (basis functions in the same
// power-of-two band, so a small molecule does not iterate in lockstep with a
// large one), oldest first, up to a job cap and the device-memory budget
// (per-job footprint estimated from the basis before anything is allocated).
// A cohort runs as one batch on its own thread and HIP streams; up to
// max_cohorts cohorts run at once, so requests that arrive while others are
// on the GPU start as soon as a cohort slot and memory are free, and their
// kernels overlap with the running cohorts' on the device.  A job's result is
// published the moment its own SCF ends; the cohort's memory returns to the
// budget when its last molecule is done.  Every per-molecule quantity of a
// batch is computed independently of the other molecules in it, so a job's
// result does not depend on its cohort.
//
// Only the library's public C API is used here (the batch entry points and
// CInt_getNumFuncs), so the scheduler builds and can be tested on the host
// against a stand-in batch.
#include <sys/time.h>

#include <algorithm>
#include <climits>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <new>
#include <thread>
#include <vector>

#include "../../include/gtfock_amd.h"

static double swtime()
{
    struct timeval tv;
    gettimeofday(&tv, nullptr);
    return tv.tv_sec + 1e-6 * tv.tv_usec;
}

struct SJob {
    BasisSet_t basis = nullptr;
    double t_submit = 0, t_start = 0, t_done = 0;
    double energy = 0;
    int iters = 0;
    int status = 0;  // 0 queued, 1 running, 2 done, < 0 failed
    int cohort = -1;
    int size_class = 0;
    int nbf = 0;
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
    // cohorts of large molecules running at once, and their cap: half the
    // cohort slots, so smaller requests always find a slot
    int running_large = 0, max_large = 1;
    bool stopping = false;
    std::thread scheduler;
    std::vector<std::thread> workers;
    std::vector<int> finished;  // cohorts whose thread has ended, not yet joined
    std::vector<int> worker_cohort;
    // tile shapes of new cohorts by molecule size: a cohort takes the last
    // range that holds its largest basis (none: the built-in defaults)
    struct ShapeRange {
        int nbf_min, nbf_max;
        std::vector<int> shapes;
    };
    std::vector<ShapeRange> shape_ranges;
    void schedule();
    void run_cohort(std::vector<long> ids, double mem, int cohort, bool large);
};

// molecules of at least 384 basis functions count as large
static const int LARGE_CLASS = 4;

// power-of-two band of the basis size: 24..47 -> 0, 48..95 -> 1, ...
static int size_class(int nbf)
{
    int c = 0;
    for (int n = 48; nbf >= n && c < 16; n *= 2) c++;
    return c;
}

struct CohortCtx {
    GtfServer* s;
    const std::vector<long>* ids;
};

// called on the cohort's thread when molecule m's SCF has ended
static void publish(void* ctx, int m, double energy, int builds)
{
    CohortCtx* c = (CohortCtx*)ctx;
    std::lock_guard<std::mutex> lk(c->s->mu);
    SJob& j = c->s->jobs[(*c->ids)[m]];
    j.energy = energy;
    j.iters = builds;
    j.t_done = swtime();
    j.status = builds < 0 ? -5 : 2;
    c->s->cv.notify_all();
}

void GtfServer::run_cohort(std::vector<long> ids, double mem, int cohort, bool large)
{
    std::vector<BasisSet_t> bases;
    std::vector<int> sh;
    {
        std::lock_guard<std::mutex> lk(mu);
        const double t = swtime();
        int nbf_max = 0;
        for (long id : ids) {
            bases.push_back(jobs[id].basis);
            jobs[id].status = 1;
            jobs[id].t_start = t;
            jobs[id].cohort = cohort;
            nbf_max = std::max(nbf_max, jobs[id].nbf);
        }
        for (const ShapeRange& r : shape_ranges)
            if (nbf_max >= r.nbf_min && nbf_max <= r.nbf_max) sh = r.shapes;
    }
    const int nmol = (int)ids.size();
    std::vector<double> e(nmol, 0.0);
    std::vector<int> it(nmol, 0);
    CohortCtx ctx{this, &ids};
    gtf_batch_t b = nullptr;
    int rc = gtf_batch_create(bases.data(), nmol, tolscr, &b);
    // the mixed-precision form: FP32 ERIs below the Schwarz bound tau
    if (rc == 0 && tau > 0.0) rc = gtf_batch_set_precision(b, tau);
    if (rc == 0 && !sh.empty()) rc = gtf_batch_set_tile_shapes(b, sh.data());
    if (rc == 0) rc = gtf_batch_scf_cb(b, max_iters, ediff, e.data(), it.data(), nullptr, publish, &ctx);
    gtf_batch_destroy(b);
    std::lock_guard<std::mutex> lk(mu);
    const double t = swtime();
    for (int k = 0; k < nmol; k++) {
        SJob& j = jobs[ids[k]];
        if (j.status == 1) {
            // a failed cohort (or one that ended without reporting the job)
            j.status = rc < 0 ? rc : -1;
            j.t_done = t;
        }
    }
    running--;
    if (large) running_large--;
    mem_in_use = running == 0 ? 0.0 : mem_in_use - mem;  // no rounding residue once idle
    finished.push_back(cohort);
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
        // jobs larger than the whole budget can never run
        for (auto itp = pending.begin(); itp != pending.end();) {
            if (jobs[*itp].mem_est > budget) {
                jobs[*itp].status = -2;
                jobs[*itp].t_done = swtime();
                itp = pending.erase(itp);
                cv.notify_all();
            } else {
                ++itp;
            }
        }
        if (pending.empty()) continue;
        // the oldest job opens the cohort (the oldest of a smaller class while
        // the large-molecule slots are all taken); same-class jobs join,
        // oldest first
        int cls = -1;
        for (long id : pending)
            if (jobs[id].size_class < LARGE_CLASS || running_large < max_large) {
                cls = jobs[id].size_class;
                break;
            }
        if (cls < 0) {
            // only large jobs are queued and their slots are taken
            cv.wait(lk);
            continue;
        }
        std::vector<long> take;
        double mem = 0.0;
        for (auto itp = pending.begin(); itp != pending.end() && (int)take.size() < max_jobs;) {
            const SJob& j = jobs[*itp];
            if (j.size_class != cls) {
                ++itp;
                continue;
            }
            if (mem_in_use + mem + j.mem_est > budget) {
                if (take.empty()) break;  // the oldest job waits for memory
                ++itp;
                continue;
            }
            take.push_back(*itp);
            mem += j.mem_est;
            itp = pending.erase(itp);
        }
        if (take.empty()) {
            // the oldest job waits for a running cohort to return memory
            // (mem_in_use > 0 here, so some cohort is running)
            cv.wait(lk);
            continue;
        }
        mem_in_use += mem;
        mem_peak = std::max(mem_peak, mem_in_use);
        running++;
        const bool large = cls >= LARGE_CLASS;
        if (large) running_large++;
        peak_running = std::max(peak_running, running);
        const int cohort = ncohorts++;
        // take out the threads of cohorts that have ended (each has already
        // published its results and returned its memory) and join them with
        // the lock released, so a long-running server holds no dead threads
        std::vector<std::thread> ended;
        for (int c : finished)
            for (size_t w = 0; w < workers.size(); w++)
                if (worker_cohort[w] == c) {
                    ended.push_back(std::move(workers[w]));
                    workers.erase(workers.begin() + w);
                    worker_cohort.erase(worker_cohort.begin() + w);
                    break;
                }
        finished.clear();
        workers.emplace_back(&GtfServer::run_cohort, this, take, mem, cohort, large);
        worker_cohort.push_back(cohort);
        if (!ended.empty()) {
            lk.unlock();
            for (std::thread& t : ended)
                if (t.joinable()) t.join();
            lk.lock();
        }
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
    s->max_large = std::max(1, max_cohorts / 2);
    s->max_jobs = max_jobs_per_cohort;
    s->tolscr = tolscr;
    s->tau = tau;
    s->max_iters = max_iters;
    s->ediff = ediff;
    s->scheduler = std::thread(&GtfServer::schedule, s);
    *out = s;
    return 0;
}

int gtf_server_set_large_slots(gtf_server_t s, int n)
{
    if (!s || n < 1) return -3;
    std::lock_guard<std::mutex> lk(s->mu);
    s->max_large = std::min(n, s->max_cohorts);
    s->cv.notify_all();
    return 0;
}

int gtf_server_set_tile_shapes(gtf_server_t s, const int* shapes)
{
    if (!s) return -3;
    std::lock_guard<std::mutex> lk(s->mu);
    s->shape_ranges.clear();
    if (shapes)
        s->shape_ranges.push_back({0, INT_MAX, std::vector<int>(shapes, shapes + 21 * GTF_TILE_SHAPE_FIELDS)});
    return 0;
}

int gtf_server_set_tile_shapes_range(gtf_server_t s, int nbf_min, int nbf_max, const int* shapes)
{
    if (!s || !shapes || nbf_min < 0 || nbf_max < nbf_min) return -3;
    std::lock_guard<std::mutex> lk(s->mu);
    s->shape_ranges.push_back({nbf_min, nbf_max, std::vector<int>(shapes, shapes + 21 * GTF_TILE_SHAPE_FIELDS)});
    return 0;
}

long gtf_server_submit(gtf_server_t s, BasisSet_t basis)
{
    if (!s || !basis) return -3;
    const int nbf = CInt_getNumFuncs(basis);
    const double bytes = gtf_batch_job_bytes(basis);
    if (nbf <= 0 || bytes <= 0) return -3;
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->stopping) return -3;
    SJob j;
    j.basis = basis;
    j.t_submit = swtime();
    j.mem_est = bytes;
    j.size_class = size_class(nbf);
    j.nbf = nbf;
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
    s->cv.wait(lk, [&] { return s->pending.empty() && s->running == 0; });
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
