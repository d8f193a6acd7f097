// serve.cpp -- an SCF request server over batches, with continuous
// batching at SCF-iteration granularity.
//
// Jobs (one molecule each) queue on submission.  A scheduler thread opens
// cohorts in arrival order: the oldest queued job opens a cohort and takes
// the queued jobs of its size class (basis functions in the same
// power-of-two band, so a small molecule does not iterate in lockstep with a
// large one), oldest first, up to a job cap and the device-memory budget
// (per-job footprint estimated from the basis before anything is allocated).
// A cohort runs as one batch on its own thread and HIP streams, one SCF
// iteration at a time (gtf_batch_scf_step); up to max_cohorts cohorts run at
// once and their kernels overlap on the device.
//
// Between two iterations a cohort admits queued jobs of its size class that
// fit the budget: it regroups (gtf_batch_regroup) into a new batch that
// carries its running molecules' SCF state over exactly and starts the new
// ones, and drops the molecules that have finished, so their memory returns
// to the budget then.  While a cohort of a small size class has room, the
// scheduler leaves that class's queued jobs to it rather than opening another
// cohort; large-molecule jobs open cohorts as before and join one when the
// large slots are taken.  A cohort also regroups without newcomers when a
// queued job waits for memory that its finished molecules still hold.
// Regroups cost a batch setup, so a cohort regroups at most every join_gap
// iterations and spends at most about a third of its time on them.  A job's
// result is published the moment its own SCF ends.  Every per-molecule
// quantity of a batch is computed independently of the other molecules in
// it, so a job's result does not depend on its cohort or on when it joined.
//
// Only the library's public C API is used here (the batch entry points and
// CInt_getNumFuncs), so the scheduler builds and can be tested on the host
// against a stand-in batch.
#include <sys/time.h>

#include <algorithm>
#include <atomic>
#include <chrono>
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
    // purification products of the cohorts (gtf_batch_set_purif_precision)
    int purif_mode = 0;
    double purif_until = 1e-3;
    // iterations between two regroups of a cohort (< 0: cohorts never admit
    // queued jobs while they run)
    int join_gap = 1;
    std::mutex mu;
    std::condition_variable cv;
    std::deque<long> pending;
    std::vector<SJob> jobs;
    double mem_in_use = 0, mem_peak = 0;
    int running = 0, ncohorts = 0, peak_running = 0;
    long joins = 0, regroups = 0;
    // cohorts of large molecules running at once, and their cap: half the
    // cohort slots, so smaller requests always find a slot
    int running_large = 0, max_large = 1;
    // small molecules first: cohorts of small molecules running now, and how
    // long a cohort of large molecules defers its next SCF iteration while
    // there are any, in units of its own last iteration (0: never)
    std::atomic<int> small_cohorts{0};
    std::atomic<double> large_yield{2.0};
    bool stopping = false;
    std::thread scheduler;
    std::vector<std::thread> workers;
    std::vector<int> finished;  // cohorts whose thread has ended, not yet joined
    std::vector<int> worker_cohort;
    // per cohort id: size class and free job slots while it admits joiners
    // (-1 when it does not)
    std::vector<int> cohort_class, cohort_room;
    std::vector<double> cohort_keep;  // estimates of its running jobs
    // tile shapes of new cohorts by molecule size: a cohort takes the last
    // range that holds its largest basis (none: the built-in defaults)
    struct ShapeRange {
        int nbf_min, nbf_max;
        std::vector<int> shapes;
    };
    std::vector<ShapeRange> shape_ranges;
    void schedule();
    void run_cohort(std::vector<long> ids, double mem, int cohort, bool large);
    bool class_has_room(int cls, double mem_est) const;
    bool memory_blocked() const;
    void start_job(long id, int cohort, double t);
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
    std::vector<long> ids;  // job id of each molecule of the cohort's current batch
    double ended_mem = 0;   // estimates of jobs that ended in the current batch
};

// called on the cohort's thread when molecule m's SCF has ended
static void publish(void* ctx, int m, double energy, int builds)
{
    CohortCtx* c = (CohortCtx*)ctx;
    std::lock_guard<std::mutex> lk(c->s->mu);
    SJob& j = c->s->jobs[c->ids[m]];
    j.energy = energy;
    j.iters = builds;
    j.t_done = swtime();
    j.status = builds < 0 ? -5 : 2;
    c->ended_mem += j.mem_est;
    c->s->cv.notify_all();
}

// a running cohort of size class cls could take a job of mem_est bytes at
// its next iteration (free slots, and the budget holds both batches of the
// regroup)
bool GtfServer::class_has_room(int cls, double mem_est) const
{
    for (size_t c = 0; c < cohort_class.size(); c++)
        if (cohort_class[c] == cls && cohort_room[c] > 0 && mem_in_use + cohort_keep[c] + mem_est <= budget)
            return true;
    return false;
}

// a queued job waits for memory (it would not fit the budget now)
bool GtfServer::memory_blocked() const
{
    for (long id : pending)
        if (mem_in_use + jobs[id].mem_est > budget) return true;
    return false;
}

void GtfServer::start_job(long id, int cohort, double t)
{
    jobs[id].status = 1;
    jobs[id].t_start = t;
    jobs[id].cohort = cohort;
}

// a batch of the given jobs with the server's settings, its SCF begun
static int open_batch(GtfServer* s, const std::vector<BasisSet_t>& bases, const std::vector<int>& sh, void* ctx,
                      gtf_batch_t* out)
{
    gtf_batch_t b = nullptr;
    int rc = gtf_batch_create(const_cast<BasisSet_t*>(bases.data()), (int)bases.size(), s->tolscr, &b);
    // the mixed-precision form: FP32 ERIs below the Schwarz bound tau
    if (rc == 0 && s->tau > 0.0) rc = gtf_batch_set_precision(b, s->tau);
    if (rc == 0 && s->purif_mode != 0) rc = gtf_batch_set_purif_precision(b, s->purif_mode, s->purif_until);
    if (rc == 0 && !sh.empty()) rc = gtf_batch_set_tile_shapes(b, sh.data());
    if (rc == 0) rc = gtf_batch_scf_begin(b, s->max_iters, s->ediff, publish, ctx);
    *out = b;
    return rc;
}

void GtfServer::run_cohort(std::vector<long> ids, double mem, int cohort, bool large)
{
    std::vector<BasisSet_t> bases;
    std::vector<int> sh;
    int cls = 0;
    {
        std::lock_guard<std::mutex> lk(mu);
        const double t = swtime();
        int nbf_max = 0;
        for (long id : ids) {
            bases.push_back(jobs[id].basis);
            start_job(id, cohort, t);
            nbf_max = std::max(nbf_max, jobs[id].nbf);
            cls = jobs[id].size_class;
        }
        // a cohort keeps its size class, so the shapes chosen here fit the
        // molecules that join it later
        for (const ShapeRange& r : shape_ranges)
            if (nbf_max >= r.nbf_min && nbf_max <= r.nbf_max) sh = r.shapes;
    }
    CohortCtx ctx{this, ids, 0.0};
    gtf_batch_t b = nullptr;
    const bool small = cls < LARGE_CLASS;
    if (small) small_cohorts++;
    int rc = open_batch(this, bases, sh, &ctx, &b);
    int step = 0, last_regroup = -1000000;
    double regroup_s = 0.0, last_iter = 0.0;
    const double t_begin = swtime();
    while (rc == 0) {
        // A large cohort's persistent ERI kernels hold every CU for tens of
        // milliseconds per class launch, so small cohorts beside it slow
        // several-fold: it waits while small cohorts run, at most
        // large_yield times its own last iteration (so it is never
        // starved), and the small jobs, which finish in a few iterations,
        // mostly have the device to themselves.
        if (!small) {
            const double cap = large_yield.load() * last_iter, t0 = swtime();
            while (small_cohorts.load() > 0 && swtime() - t0 < cap)
                std::this_thread::sleep_for(std::chrono::microseconds(500));
        }
        const double t_iter = swtime();
        const int left = gtf_batch_scf_step(b);
        last_iter = swtime() - t_iter;
        if (left < 0) {
            rc = left;
            break;
        }
        step++;
        // admission between iterations
        std::vector<long> join;
        std::vector<BasisSet_t> add;
        double join_mem = 0.0, keep_mem = 0.0;
        bool regroup = false, blocked_here = false;
        {
            std::lock_guard<std::mutex> lk(mu);
            for (long id : ctx.ids)
                if (jobs[id].status == 1) keep_mem += jobs[id].mem_est;
            const double now = swtime();
            const bool paced = step - last_regroup >= join_gap && regroup_s <= 0.5 * (now - t_begin - regroup_s);
            if (join_gap >= 0 && (paced || left == 0) && !stopping) {
                // while a regroup carries running molecules both batches
                // exist, so those count twice against the budget; a cohort
                // whose molecules have all ended frees its batch first
                const double base = left > 0 ? mem_in_use + keep_mem : mem_in_use - ctx.ended_mem;
                for (auto itp = pending.begin(); itp != pending.end() && left + (int)join.size() < max_jobs;) {
                    const SJob& j = jobs[*itp];
                    if (j.size_class != cls) {
                        ++itp;
                        continue;
                    }
                    if (base + join_mem + j.mem_est > budget) {
                        blocked_here = true;
                        ++itp;
                        continue;
                    }
                    join.push_back(*itp);
                    add.push_back(j.basis);
                    join_mem += j.mem_est;
                    itp = pending.erase(itp);
                }
            }
            // without newcomers, a regroup that only drops finished molecules
            // returns their memory to a job that waits for it (here or in the
            // queue), if both batches fit the budget meanwhile
            const bool shrink = left > 0 && ctx.ended_mem > 0 && mem_in_use + keep_mem <= budget &&
                                (blocked_here || memory_blocked());
            regroup = !join.empty() || shrink;
            if (regroup) {
                const double t = swtime();
                for (long id : join) start_job(id, cohort, t);
                mem_in_use += join_mem;
                mem_peak = std::max(mem_peak, left > 0 ? mem_in_use + keep_mem : mem_in_use - ctx.ended_mem);
                joins += (long)join.size();
                regroups++;
            }
            cohort_room[cohort] = (left == 0 && join.empty()) ? -1 : max_jobs - left - (int)join.size();
            cohort_keep[cohort] = keep_mem + join_mem;
        }
        if (!regroup) {
            if (left == 0) break;
            continue;
        }
        const double t0 = swtime();
        gtf_batch_t nb = nullptr;
        std::vector<long> nids;
        if (left == 0) {
            // nothing to carry over: free the finished batch, then open one
            // for the newcomers
            gtf_batch_destroy(b);
            b = nullptr;
            {
                std::lock_guard<std::mutex> lk(mu);
                mem_in_use -= ctx.ended_mem;
                mem -= ctx.ended_mem;
                ctx.ended_mem = 0.0;
                ctx.ids = join;
                cv.notify_all();
            }
            rc = open_batch(this, add, sh, &ctx, &b);
        } else {
            std::vector<int> origin(left + join.size());
            rc = gtf_batch_regroup(b, add.data(), (int)add.size(), &nb, origin.data());
            std::lock_guard<std::mutex> lk(mu);
            if (rc == 0) {
                for (int o : origin) nids.push_back(o >= 0 ? ctx.ids[o] : join[-o - 1]);
                gtf_batch_destroy(b);
                b = nb;
                ctx.ids = nids;
                mem_in_use -= ctx.ended_mem;
                mem -= ctx.ended_mem;
                ctx.ended_mem = 0.0;
                cv.notify_all();
            } else {
                // the newcomers fail with the cohort below
                for (long id : join) ctx.ids.push_back(id);
            }
        }
        mem += join_mem;
        regroup_s += swtime() - t0;
        last_regroup = step;
    }
    if (small) small_cohorts--;
    gtf_batch_destroy(b);
    std::lock_guard<std::mutex> lk(mu);
    const double t = swtime();
    for (long id : ctx.ids) {
        SJob& j = jobs[id];
        if (j.status == 1) {
            // a failed cohort (or one that ended without reporting the job)
            j.status = rc < 0 ? rc : -1;
            j.t_done = t;
        }
    }
    cohort_room[cohort] = -1;
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
        // oldest first.  Small-class jobs are left to a running cohort of
        // their class that has room: it takes them at its next iteration.
        int cls = -1;
        for (long id : pending) {
            const int c = jobs[id].size_class;
            if (c < LARGE_CLASS && join_gap >= 0 && class_has_room(c, jobs[id].mem_est)) continue;
            if (c < LARGE_CLASS || running_large < max_large) {
                cls = c;
                break;
            }
        }
        if (cls < 0) {
            // only jobs that wait to join a cohort, or large jobs whose slots
            // are taken, are queued
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
        cohort_class.push_back(cls);
        cohort_room.push_back(max_jobs - (int)take.size());
        cohort_keep.push_back(mem);
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

int gtf_server_set_purif_precision(gtf_server_t s, int mode, double until)
{
    if (!s || mode < 0 || mode > 2 || until <= 0.0) return -3;
    std::lock_guard<std::mutex> lk(s->mu);
    s->purif_mode = mode;
    s->purif_until = until;
    return 0;
}

int gtf_server_set_large_yield(gtf_server_t s, double factor)
{
    if (!s || !(factor >= 0.0)) return -3;
    s->large_yield.store(factor);
    return 0;
}

int gtf_server_set_join_gap(gtf_server_t s, int iters)
{
    if (!s) return -3;
    std::lock_guard<std::mutex> lk(s->mu);
    s->join_gap = iters;
    s->cv.notify_all();
    return 0;
}

int gtf_server_set_tile_shapes(gtf_server_t s, const int* shapes)
{
    if (!s || (shapes && gtf_tile_shapes_check(shapes))) return -3;
    std::lock_guard<std::mutex> lk(s->mu);
    s->shape_ranges.clear();
    if (shapes)
        s->shape_ranges.push_back({0, INT_MAX, std::vector<int>(shapes, shapes + 21 * GTF_TILE_SHAPE_FIELDS)});
    return 0;
}

int gtf_server_set_tile_shapes_range(gtf_server_t s, int nbf_min, int nbf_max, const int* shapes)
{
    if (!s || !shapes || nbf_min < 0 || nbf_max < nbf_min || gtf_tile_shapes_check(shapes)) return -3;
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

int gtf_server_join_stats(gtf_server_t s, long* joins, long* regroups)
{
    if (!s) return -3;
    std::lock_guard<std::mutex> lk(s->mu);
    if (joins) *joins = s->joins;
    if (regroups) *regroups = s->regroups;
    return 0;
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
