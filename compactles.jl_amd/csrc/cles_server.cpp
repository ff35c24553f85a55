// Serving loop: independent solve requests (a solver descriptor, an initial
// state, tfinal and nmax, as setup + run! of the reference, src/problem.jl:321
// and src/timestep.jl:277) admitted into HBM by their device footprint and run
// concurrently, each on its own context and stream.
//
// Admission: a request is admitted when its footprint (cles_solver_footprint)
// fits in what the budget has left and a runner is free.  Requests are
// considered in arrival order; one that does not fit is passed over for
// smaller ones (continuous batching into free capacity) until it has been
// passed over kAging times, after which nothing behind it is admitted until
// it runs (no starvation).
//
// Same-shape two-dimensional requests (one descriptor, byte for byte) that
// are queued together are admitted together, up to the stack limit and within
// the budget, and run as one stack (cles_stack_*): every launch covers all of
// them, each with its own dt, end time and step cap.
//
// Each admitted request or stack owns its device buffers and stream, and the
// kernels are deterministic and act on each instance's own data, so a
// request's result is bitwise the same alone and in any mix.
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "cles_internal.h"

namespace {

constexpr int kAging = 4;

enum { RQ_QUEUED = 0, RQ_RUNNING = 1, RQ_DONE = 2, RQ_FAILED = 3 };

struct Request {
    int64_t id = 0;
    cles_solver_desc desc{};
    std::vector<double> Q;   // padded host state: input, then result
    int64_t padded[3] = {1, 1, 1};
    int halo[3] = {0, 0, 0};
    double tfinal = 0;
    int64_t nmax = 0;
    size_t footprint = 0;
    int state = RQ_QUEUED;
    int status = CLES_OK;
    std::string msg;
    int passed_over = 0;
    double t_submit = 0, t_admit = 0, t_done = 0;
    double t_sim = 0;
    int64_t steps = 0;
    int group = 1;   // requests of the stack it ran in
};

double now_s(std::chrono::steady_clock::time_point t0)
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace

struct cles_server {
    int device = 0;
    size_t budget = 0, in_use = 0;
    int runners = 1;
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
    std::mutex mu;
    std::condition_variable cv;
    std::deque<int64_t> queue;
    std::map<int64_t, std::unique_ptr<Request>> reqs;
    std::vector<std::thread> workers;
    bool stop = false;
    int64_t next_id = 1;
    int running = 0;
    int run_mode = CLES_RUN_HOST_DT;   // of every solver it creates
    int stack_limit = 8;               // requests per stack (1: no stacks)
    double window = 0.005;             // seconds a stackable request waits for company

    static bool stackable(const cles_solver_desc &d)
    {
        return d.n_global[2] == 1 && (d.n_global[0] > 1 || d.n_global[1] > 1) &&
               d.metric == CLES_METRIC_CARTESIAN && d.n_body_forces == 0;
    }

    // under mu: the next admissible request with the queued requests of the
    // same descriptor that fit beside it (taken off the queue); empty if none
    std::vector<Request *> admit_locked()
    {
        std::vector<Request *> g;
        if (running >= runners) return g;
        auto it = queue.begin();
        for (; it != queue.end(); ++it) {
            const Request *r = reqs[*it].get();
            if (r->footprint <= budget - in_use) break;
            if (r->passed_over >= kAging) return g;   // capacity is held for it
        }
        if (it == queue.end()) return g;
        Request *r = reqs[*it].get();
        if (stack_limit > 1 && stackable(r->desc) && now_s(t0) - r->t_submit < window) {
            // requests submitted together arrive one call apart: a stackable
            // request waits up to `window` for the rest of its stack
            int same = 0;
            for (int64_t id : queue) same += std::memcmp(&reqs[id]->desc, &r->desc, sizeof(cles_solver_desc)) == 0;
            if (same < stack_limit) return g;
        }
        for (auto jt = queue.begin(); jt != it; ++jt) reqs[*jt]->passed_over++;
        queue.erase(it);
        g.push_back(r);
        if (stack_limit > 1 && stackable(r->desc)) {
            size_t fp = r->footprint;
            for (auto jt = queue.begin(); jt != queue.end() && (int)g.size() < stack_limit;) {
                Request *q = reqs[*jt].get();
                if (std::memcmp(&q->desc, &r->desc, sizeof(cles_solver_desc)) == 0 &&
                    fp + q->footprint <= budget - in_use) {
                    fp += q->footprint;
                    g.push_back(q);
                    jt = queue.erase(jt);
                } else {
                    ++jt;
                }
            }
        }
        return g;
    }

    // a stack of same-descriptor requests: cles_stack_* with each request's
    // own state, end time and step cap
    void run_stack(const std::vector<Request *> &g)
    {
        const int K = (int)g.size();
        cles_ctx *ctx = nullptr;
        cles_stack *st = nullptr;
        std::vector<double> tf(K), t(K, 0.0);
        std::vector<int64_t> nm(K), steps(K, 0);
        std::vector<int> status(K, CLES_OK);
        for (int i = 0; i < K; i++) {
            tf[i] = g[i]->tfinal;
            nm[i] = g[i]->nmax;
        }
        int rc = cles_ctx_create(device, &ctx);
        if (rc == CLES_OK) rc = cles_stack_create(ctx, &g[0]->desc, K, &st);
        for (int i = 0; i < K && rc == CLES_OK; i++) rc = cles_stack_upload(st, i, g[i]->Q.data(), g[i]->padded, g[i]->halo);
        if (rc == CLES_OK) rc = cles_stack_run(st, tf.data(), nm.data(), INT64_MAX, nullptr);
        if (rc == CLES_OK) rc = cles_stack_state(st, t.data(), steps.data(), status.data());
        // the states reached (at a failure: the last completed step)
        for (int i = 0; i < K && rc == CLES_OK; i++)
            rc = cles_stack_download(st, i, g[i]->Q.data(), g[i]->padded, g[i]->halo);
        const std::string msg = rc != CLES_OK ? std::string(cles_last_error()) : std::string();
        if (st) cles_stack_destroy(st);
        if (ctx) cles_ctx_destroy(ctx);
        std::lock_guard<std::mutex> lk(mu);
        const double done = now_s(t0);
        for (int i = 0; i < K; i++) {
            Request *r = g[i];
            r->group = K;
            r->status = rc != CLES_OK ? rc : status[i];
            r->msg = rc != CLES_OK ? msg
                                   : (status[i] != CLES_OK ? "run! stopped by check_step (status " +
                                                                 std::to_string(status[i]) + ")"
                                                           : std::string());
            r->t_sim = t[i];
            r->steps = steps[i];
            r->state = r->status == CLES_OK ? RQ_DONE : RQ_FAILED;
            r->t_done = done;
        }
    }

    void run_one(Request *r)
    {
        cles_ctx *ctx = nullptr;
        cles_solver *s = nullptr;
        int rc = cles_ctx_create(device, &ctx);
        if (rc == CLES_OK) rc = cles_solver_create(ctx, &r->desc, &s);
        if (rc == CLES_OK) rc = cles_solver_set_run_mode(s, run_mode);
        if (rc == CLES_OK) rc = cles_solver_upload(s, r->Q.data(), r->padded, r->halo);
        int run_rc = CLES_OK;
        std::string run_msg;
        if (rc == CLES_OK) {
            run_rc = cles_run(s, r->tfinal, r->nmax, nullptr, 0);
            if (run_rc != CLES_OK) run_msg = cles_last_error();
            // the state reached (at a failure: the last completed step)
            rc = cles_solver_download(s, r->Q.data(), r->padded, r->halo);
            if (rc == CLES_OK) rc = cles_solver_get_time(s, &r->t_sim, &r->steps);
        }
        const std::string msg = rc != CLES_OK ? std::string(cles_last_error()) : run_msg;
        if (s) cles_solver_destroy(s);
        if (ctx) cles_ctx_destroy(ctx);
        std::lock_guard<std::mutex> lk(mu);
        r->status = rc != CLES_OK ? rc : run_rc;
        r->msg = msg;
        r->state = r->status == CLES_OK ? RQ_DONE : RQ_FAILED;
        r->t_done = now_s(t0);
    }

    void worker()
    {
        std::unique_lock<std::mutex> lk(mu);
        for (;;) {
            std::vector<Request *> g;
            // a request held for company is admitted when its window ends
            while (!stop && (g = admit_locked()).empty())
                cv.wait_for(lk, std::chrono::milliseconds(1));
            if (g.empty()) return;   // stopping
            size_t fp = 0;
            for (Request *r : g) {
                r->state = RQ_RUNNING;
                r->t_admit = now_s(t0);
                fp += r->footprint;
            }
            in_use += fp;
            running++;
            lk.unlock();
            if (g.size() == 1)
                run_one(g[0]);
            else
                run_stack(g);
            lk.lock();
            in_use -= fp;
            running--;
            cv.notify_all();
        }
    }
};

extern "C" {

int cles_server_create(int device, size_t hbm_budget, int runners, cles_server **out)
{
    CLES_TRY({
        CLES_REQUIRE(out, "NULL argument");
        CLES_REQUIRE(runners >= 1 && runners <= 64, "runners must be 1..64");
        CLES_REQUIRE(hbm_budget > 0, "the HBM budget must be positive");
        auto *sv = new cles_server;
        sv->device = device;
        sv->budget = hbm_budget;
        sv->runners = runners;
        try {
            for (int k = 0; k < runners; k++) sv->workers.emplace_back([sv] { sv->worker(); });
        } catch (...) {
            {
                std::lock_guard<std::mutex> lk(sv->mu);
                sv->stop = true;
            }
            sv->cv.notify_all();
            for (auto &t : sv->workers) t.join();
            delete sv;
            throw cles::Error{CLES_ERR_NOMEM, "could not start the runner threads"};
        }
        *out = sv;
    })
}

int cles_server_destroy(cles_server *sv)
{
    CLES_TRY({
        if (!sv) return CLES_OK;
        {
            std::lock_guard<std::mutex> lk(sv->mu);
            sv->stop = true;
            // queued requests are dropped and end as failed; running ones finish
            for (int64_t qid : sv->queue) {
                Request *r = sv->reqs[qid].get();
                r->state = RQ_FAILED;
                r->status = CLES_ERR_INVALID;
                r->msg = "the server stopped before the request was admitted";
            }
            sv->queue.clear();
        }
        sv->cv.notify_all();
        for (auto &t : sv->workers) t.join();
        delete sv;
    })
}

int cles_server_set_stack_limit(cles_server *sv, int limit, double window_s)
{
    CLES_TRY({
        CLES_REQUIRE(sv, "NULL argument");
        CLES_REQUIRE(limit >= 1 && limit <= 4096, "the stack limit must be 1..4096");
        CLES_REQUIRE(window_s >= 0 && window_s <= 10, "the stacking window must be 0..10 s");
        std::lock_guard<std::mutex> lk(sv->mu);
        sv->stack_limit = limit;
        sv->window = window_s;
    })
}

int cles_server_group(cles_server *sv, int64_t id, int *size)
{
    CLES_TRY({
        CLES_REQUIRE(sv && size, "NULL argument");
        std::lock_guard<std::mutex> lk(sv->mu);
        auto it = sv->reqs.find(id);
        CLES_REQUIRE(it != sv->reqs.end(), "unknown request id");
        CLES_REQUIRE(it->second->state == RQ_DONE || it->second->state == RQ_FAILED, "request has not finished");
        *size = it->second->group;
    })
}

int cles_server_set_run_mode(cles_server *sv, int run_mode)
{
    CLES_TRY({
        CLES_REQUIRE(sv, "NULL argument");
        CLES_REQUIRE(run_mode >= CLES_RUN_HOST_DT && run_mode <= CLES_RUN_STEP_GRAPH, "unknown run mode");
        std::lock_guard<std::mutex> lk(sv->mu);
        sv->run_mode = run_mode;
    })
}

int cles_server_submit(cles_server *sv, const cles_solver_desc *desc, const double *Q_host, const int64_t padded[3],
                       const int halo[3], double tfinal, int64_t nmax, int64_t *id)
{
    CLES_TRY({
        CLES_REQUIRE(sv && desc && Q_host && padded && halo && id, "NULL argument");
        size_t fp = 0;
        const int rc = cles_solver_footprint(desc, &fp);
        if (rc != CLES_OK) throw cles::Error{rc, cles_last_error()};
        if (fp > sv->budget)
            throw cles::Error{CLES_ERR_NOMEM, "request needs " + std::to_string(fp) + " device bytes; the budget is " +
                                                  std::to_string(sv->budget)};
        int64_t np = 1;
        for (int k = 0; k < 3; k++) {
            CLES_REQUIRE(halo[k] >= 0 && padded[k] >= desc->n_global[k] + 2 * (int64_t)halo[k],
                         "padded extents smaller than n + 2*halo");
            np *= padded[k];
        }
        auto r = std::make_unique<Request>();
        r->desc = *desc;
        const size_t ncons = (size_t)desc->n_species + 4;
        r->Q.assign(Q_host, Q_host + (size_t)np * ncons);
        for (int k = 0; k < 3; k++) {
            r->padded[k] = padded[k];
            r->halo[k] = halo[k];
        }
        r->tfinal = tfinal;
        r->nmax = nmax;
        r->footprint = fp;
        {
            std::lock_guard<std::mutex> lk(sv->mu);
            CLES_REQUIRE(!sv->stop, "the server is stopping");
            r->id = sv->next_id++;
            r->t_submit = now_s(sv->t0);
            *id = r->id;
            sv->queue.push_back(r->id);
            sv->reqs[r->id] = std::move(r);
        }
        sv->cv.notify_all();
    })
}

int cles_server_wait(cles_server *sv, int64_t id, double timeout_s, int *state)
{
    CLES_TRY({
        CLES_REQUIRE(sv && state, "NULL argument");
        std::unique_lock<std::mutex> lk(sv->mu);
        auto it = sv->reqs.find(id);
        CLES_REQUIRE(it != sv->reqs.end(), "unknown request id");
        Request *r = it->second.get();
        auto done = [&] { return r->state == RQ_DONE || r->state == RQ_FAILED; };
        if (timeout_s < 0)
            sv->cv.wait(lk, done);
        else
            sv->cv.wait_for(lk, std::chrono::duration<double>(timeout_s), done);
        *state = r->state;
    })
}

int cles_server_result(cles_server *sv, int64_t id, double *Q_host, const int64_t padded[3], const int halo[3],
                       int *status, double *t, int64_t *steps)
{
    CLES_TRY({
        CLES_REQUIRE(sv && status, "NULL argument");
        std::lock_guard<std::mutex> lk(sv->mu);
        auto it = sv->reqs.find(id);
        CLES_REQUIRE(it != sv->reqs.end(), "unknown request id");
        Request *r = it->second.get();
        CLES_REQUIRE(r->state == RQ_DONE || r->state == RQ_FAILED, "request has not finished");
        *status = r->status;
        if (t) *t = r->t_sim;
        if (steps) *steps = r->steps;
        if (Q_host) {
            CLES_REQUIRE(padded && halo, "NULL argument");
            for (int k = 0; k < 3; k++)
                CLES_REQUIRE(padded[k] == r->padded[k] && halo[k] == r->halo[k],
                             "result layout differs from the submitted one");
            std::memcpy(Q_host, r->Q.data(), r->Q.size() * sizeof(double));
        }
        if (r->status != CLES_OK) cles::set_error(r->msg);
    })
}

int cles_server_timing(cles_server *sv, int64_t id, double *t_submit, double *t_admit, double *t_done,
                       size_t *footprint)
{
    CLES_TRY({
        CLES_REQUIRE(sv, "NULL argument");
        std::lock_guard<std::mutex> lk(sv->mu);
        auto it = sv->reqs.find(id);
        CLES_REQUIRE(it != sv->reqs.end(), "unknown request id");
        Request *r = it->second.get();
        if (t_submit) *t_submit = r->t_submit;
        if (t_admit) *t_admit = r->t_admit;
        if (t_done) *t_done = r->t_done;
        if (footprint) *footprint = r->footprint;
    })
}

int cles_server_release(cles_server *sv, int64_t id)
{
    CLES_TRY({
        CLES_REQUIRE(sv, "NULL argument");
        std::lock_guard<std::mutex> lk(sv->mu);
        auto it = sv->reqs.find(id);
        CLES_REQUIRE(it != sv->reqs.end(), "unknown request id");
        CLES_REQUIRE(it->second->state == RQ_DONE || it->second->state == RQ_FAILED, "request has not finished");
        sv->reqs.erase(it);
    })
}

int cles_server_stats(cles_server *sv, size_t *in_use, int *running, int *queued)
{
    CLES_TRY({
        CLES_REQUIRE(sv, "NULL argument");
        std::lock_guard<std::mutex> lk(sv->mu);
        if (in_use) *in_use = sv->in_use;
        if (running) *running = sv->running;
        if (queued) *queued = (int)sv->queue.size();
    })
}

}  // extern "C"
