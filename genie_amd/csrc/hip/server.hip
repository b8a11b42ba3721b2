// Serving form: a continuous-batching request loop over an HBM-resident
// index (C ABI in include/genie_hip.h).
//
// A scheduler thread and E executors (a plan, a non-blocking stream and a
// host thread each).  Whenever an executor is free and requests are pending,
// the scheduler packs the pending requests (FIFO) into a batch -- as many as
// fit the batch limits and the device-memory budget of in-flight requests --
// and hands it to the executor, whose thread uploads it, launches the search
// and an asynchronous copy of the results into pinned host memory, waits for
// its event and returns the results per request.  Requests that arrive while
// other batches are on the device go to the next free executor, so host-side
// batch preparation and device work of different batches overlap.
//
// The reference serves one JSON batch at a time from a polled queue
// (src/dist_genie/online.cu:144-180, scheduler.cc:20-64) and admits whole
// tables by free memory (knn.cu:207-243); here admission is per request.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "engine.h"
#include "genie_hip.h"

using Clock = std::chrono::steady_clock;

namespace {

struct Request {
    uint64_t ticket = 0;
    std::vector<genie_dim_t> items;
    uint32_t k = 0;
    uint64_t mem = 0;
    Clock::time_point t_submit, t_done;
    std::vector<int32_t> ids, counts;
    int32_t* dst_ids = nullptr;     // when set (a replay): the results go straight here
    int32_t* dst_counts = nullptr;
    int rc = GENIE_OK;
    std::string err;
    std::atomic<bool> done{false};  // set (release) after the results and t_done are written
};

struct Executor {
    genie_plan* plan = nullptr;
    hipStream_t stream = nullptr;
    hipEvent_t ev = nullptr;
    int32_t* h_ids = nullptr;
    int32_t* h_counts = nullptr;
    uint32_t* h_ovf = nullptr;
    uint64_t host_cap = 0;
    std::vector<Request*> batch;
    uint64_t mem = 0;
    bool busy = false;      // holds a batch (set by the scheduler, cleared by finish)
    bool assigned = false;  // a batch waits for the executor's thread
    std::condition_variable cv;
    std::thread th;
    // time spent per phase of a batch (ns, this executor's thread): prepare
    // and upload, launch, wait for the results, hand them back
    std::atomic<uint64_t> phase_ns[4] = {};
    std::atomic<uint64_t> phase_max_ns[4] = {};  // the longest single batch per phase
    std::atomic<uint64_t> reruns{0};  // batches run again after a candidate overflow
};

uint64_t ns_since(Clock::time_point t) {
    return (uint64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(Clock::now() - t).count();
}

void note_max(std::atomic<uint64_t>& m, uint64_t d)
{
    uint64_t v = m.load(std::memory_order_relaxed);
    while (d > v && !m.compare_exchange_weak(v, d, std::memory_order_relaxed))
        ;
}

// Adds the time since t to phase i of executor e and to its longest batch.
void note_phase(Executor& e, int i, Clock::time_point t)
{
    const uint64_t d = ns_since(t);
    e.phase_ns[i].fetch_add(d, std::memory_order_relaxed);
    note_max(e.phase_max_ns[i], d);
}

// Device scratch one request holds while in flight: its work items, its
// candidate space as a plan first sizes it (gh::initial_candidates: the dense
// and set forms when the index has forward codes or lists and the plan is
// batched or automatic, else the tile forms) and its output rows.  The
// budget covers these per-request items only: each executor's plan also
// holds buffers sized once for its largest batch (candidate space and
// output rows for max_batch requests of max_k, grown after an overflow),
// which stay allocated while the server lives.
uint64_t request_bytes(uint64_t n_items, uint32_t k, bool dense, uint32_t n_objects)
{
    const uint64_t cand = gh::initial_candidates(k, dense, n_objects);
    return n_items * sizeof(genie_dim_t) + cand * 8 + (uint64_t)k * 8 + 64;
}

}  // namespace

struct genie_server {
    genie_index* ix = nullptr;
    genie_server_config cfg{};
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Request*> pending;
    std::unordered_map<uint64_t, std::unique_ptr<Request>> reqs;
    std::deque<Executor> ex;  // stable addresses: each executor has its own thread
    uint64_t next_ticket = 1, mem_in_use = 0;
    uint64_t n_batches = 0, n_done = 0, max_inflight = 0, inflight = 0;
    uint64_t events = 0;  // submissions and completions: what the scheduler waits for
    std::atomic<uint64_t> queue_max_ns{0};  // longest wait from submission to a batch
    bool stop = false, stop_exec = false;
    std::thread sched;
    int device = 0;

    void run();
    void exec_loop(Executor& e);
    int launch(Executor& e);
    void finish(Executor& e, int rc, const std::string& err);
};

static int fail_s(int code, const std::string& m)
{
    gh::set_error(m);
    return code;
}

int genie_server::launch(Executor& e)
{
    const auto t0 = Clock::now();
    const uint32_t nq = (uint32_t)e.batch.size();
    std::vector<genie_dim_t> items;
    std::vector<uint32_t> kq(nq);
    size_t n_items = 0;
    for (const Request* r : e.batch)
        n_items += r->items.size();
    items.reserve(n_items);
    for (uint32_t q = 0; q < nq; ++q) {
        Request* r = e.batch[q];
        for (genie_dim_t it : r->items) {
            it.query = (int32_t)q;
            items.push_back(it);
        }
        kq[q] = r->k;
    }
    uint64_t tot = 0;
    for (uint32_t k : kq)
        tot += k;
    if (tot > e.host_cap) {
        // once for the largest batch (pinned memory: freeing it waits for the device)
        const uint64_t cap = std::max<uint64_t>(tot, (uint64_t)cfg.max_batch * cfg.max_k);
        genie_hip_host_free(e.h_ids);
        genie_hip_host_free(e.h_counts);
        e.h_ids = e.h_counts = nullptr;
        e.host_cap = 0;
        if (int rc = genie_hip_host_alloc((void**)&e.h_ids, cap * 4))
            return rc;
        if (int rc = genie_hip_host_alloc((void**)&e.h_counts, cap * 4))
            return rc;
        e.host_cap = cap;
    }
    // every request was validated when it was submitted
    if (int rc = gh::plan_upload_validated(e.plan, items.data(), items.size(), nq, kq.data()))
        return rc;
    note_phase(e, 0, t0);
    const auto t1 = Clock::now();
    if (int rc = genie_hip_plan_run(e.plan, e.stream))
        return rc;
    if (int rc = genie_hip_plan_download_async(e.plan, e.stream, e.h_ids, e.h_counts, e.h_ovf))
        return rc;
    if (hipEventRecord(e.ev, e.stream) != hipSuccess)
        return fail_s(GENIE_E_RUNTIME, "hipEventRecord");
    note_phase(e, 1, t1);
    return GENIE_OK;
}

void genie_server::finish(Executor& e, int rc, const std::string& err_in)
{
    std::string err = err_in;
    if (!rc && *e.h_ovf) {  // candidate space too small: grow, rerun, copy synchronously
        e.reruns.fetch_add(1, std::memory_order_relaxed);
        rc = genie_hip_plan_download(e.plan, e.stream, e.h_ids, e.h_counts);
        if (rc)
            err = genie_hip_last_error();
    }
    const auto now = Clock::now();
    // the batch's requests are this executor's until marked done: their
    // results are copied before the lock is taken
    uint64_t off = 0;
    for (Request* r : e.batch) {
        r->rc = rc;
        r->err = err;
        if (!rc && r->dst_ids) {
            std::memcpy(r->dst_ids, e.h_ids + off, r->k * 4ull);
            std::memcpy(r->dst_counts, e.h_counts + off, r->k * 4ull);
        } else if (!rc) {
            r->ids.assign(e.h_ids + off, e.h_ids + off + r->k);
            r->counts.assign(e.h_counts + off, e.h_counts + off + r->k);
        }
        off += r->k;
        r->t_done = now;
    }
    {
        std::lock_guard<std::mutex> lk(mu);
        for (Request* r : e.batch)
            r->done.store(true, std::memory_order_release);
        n_done += e.batch.size();
        inflight -= e.batch.size();
        mem_in_use -= e.mem;
        e.batch.clear();
        e.mem = 0;
        e.busy = false;
        ++events;
    }
    cv_done.notify_all();
    cv_work.notify_one();  // an executor is free
}

// Executor thread: prepare and launch each batch handed over, wait for its
// results (event polling), return them.
void genie_server::exec_loop(Executor& e)
{
    (void)hipSetDevice(device);
    std::unique_lock<std::mutex> lk(mu);
    while (true) {
        e.cv.wait(lk, [&] { return e.assigned || stop_exec; });
        if (!e.assigned)
            break;
        e.assigned = false;
        lk.unlock();
        int rc = launch(e);
        std::string err = rc ? genie_hip_last_error() : "";
        const auto tw = Clock::now();
        if (!rc) {
            hipError_t q;
            while ((q = hipEventQuery(e.ev)) == hipErrorNotReady)
                std::this_thread::sleep_for(std::chrono::microseconds(5));
            if (q != hipSuccess) {
                (void)hipGetLastError();
                rc = GENIE_E_RUNTIME;
                err = hipGetErrorString(q);
            }
        }
        note_phase(e, 2, tw);
        const auto tf = Clock::now();
        finish(e, rc, err);
        note_phase(e, 3, tf);
        lk.lock();
    }
}

// Scheduler thread: FIFO admission of pending requests into free executors.
void genie_server::run()
{
    std::unique_lock<std::mutex> lk(mu);
    while (true) {
        const uint64_t seen = events;
        for (Executor& e : ex) {
            if (e.busy || pending.empty())
                continue;
            uint64_t items = 0, mem = 0;
            while (!pending.empty() && e.batch.size() < cfg.max_batch) {
                Request* r = pending.front();
                const bool first = e.batch.empty() && inflight == 0;  // always admit one when idle
                if (!first && (items + r->items.size() > cfg.max_batch_items ||
                               mem_in_use + mem + r->mem > cfg.memory_budget))
                    break;
                if (first && r->items.size() > cfg.max_batch_items)
                    break;
                items += r->items.size();
                mem += r->mem;
                e.batch.push_back(r);
                pending.pop_front();
            }
            if (e.batch.empty())
                continue;
            note_max(queue_max_ns, ns_since(e.batch.front()->t_submit));  // FIFO: the oldest
            e.mem = mem;
            e.busy = true;
            e.assigned = true;
            mem_in_use += mem;
            inflight += e.batch.size();
            max_inflight = std::max(max_inflight, inflight);
            ++n_batches;
            e.cv.notify_one();
        }
        if (stop && pending.empty() && inflight == 0)
            break;
        // sleep (releasing the lock) until a request arrives or a batch
        // completes; free executors alone are no reason to look again, since
        // the budget may hold the pending requests back until a completion
        cv_work.wait(lk, [&] { return events != seen || (stop && pending.empty() && inflight == 0); });
    }
}

extern "C" {

int genie_hip_server_create(genie_index* ix, const genie_server_config* cfg, genie_server** out)
{
    if (!ix || !cfg || !out)
        return fail_s(GENIE_E_INVALID, "null argument");
    if (cfg->executors == 0 || cfg->executors > 16 || cfg->max_batch == 0 || cfg->max_batch_items == 0 ||
        cfg->max_k == 0 || cfg->max_k > GENIE_MAX_K || (cfg->mode < 0 || cfg->mode > 2))
        return fail_s(GENIE_E_INVALID, "server configuration out of range");
    auto* s = new genie_server();
    s->ix = ix;
    s->cfg = *cfg;
    if (s->cfg.memory_budget == 0)
        s->cfg.memory_budget = 1ull << 30;
    (void)hipGetDevice(&s->device);
    s->ex.resize(cfg->executors);
    int rc = GENIE_OK;
    for (Executor& e : s->ex) {
        if ((rc = genie_hip_plan_create(ix, cfg->max_batch, cfg->max_batch_items, cfg->max_k, 0, cfg->mode,
                                        &e.plan)))
            break;
        if (hipStreamCreateWithFlags(&e.stream, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&e.ev, hipEventDisableTiming) != hipSuccess) {
            rc = fail_s(GENIE_E_RUNTIME, "stream/event creation");
            break;
        }
        if ((rc = genie_hip_host_alloc((void**)&e.h_ovf, 4)))
            break;
        *e.h_ovf = 0;
    }
    if (rc) {
        genie_hip_server_destroy(s);
        return rc;
    }
    for (Executor& e : s->ex) {
        Executor* ep = &e;
        e.th = std::thread([s, ep] { s->exec_loop(*ep); });
    }
    s->sched = std::thread([s] { s->run(); });
    *out = s;
    return GENIE_OK;
}

int genie_hip_server_destroy(genie_server* s)
{
    if (!s)
        return GENIE_OK;
    if (s->sched.joinable()) {
        {
            std::lock_guard<std::mutex> lk(s->mu);
            s->stop = true;
        }
        s->cv_work.notify_all();
        s->sched.join();
    }
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop_exec = true;
    }
    for (Executor& e : s->ex) {
        e.cv.notify_all();
        if (e.th.joinable())
            e.th.join();
    }
    for (Executor& e : s->ex) {
        if (e.stream)
            (void)hipStreamSynchronize(e.stream);
        genie_hip_plan_destroy(e.plan);
        if (e.ev)
            (void)hipEventDestroy(e.ev);
        if (e.stream)
            (void)hipStreamDestroy(e.stream);
        genie_hip_host_free(e.h_ids);
        genie_hip_host_free(e.h_counts);
        genie_hip_host_free(e.h_ovf);
    }
    delete s;
    return GENIE_OK;
}

// Submits one request; *out (optional) receives it, which stays valid until
// its ticket is waited for or erased.
static int submit_request(genie_server* s, const genie_dim_t* items, uint32_t n_items, uint32_t k,
                          uint64_t* ticket, Request** out, int32_t* dst_ids = nullptr,
                          int32_t* dst_counts = nullptr)
{
    if (k == 0 || k > s->cfg.max_k)
        return fail_s(GENIE_E_INVALID, "k outside [1, max_k]");
    if (n_items > s->cfg.max_batch_items)
        return fail_s(GENIE_E_INVALID, "request has more work items than a batch holds");
    auto r = std::make_unique<Request>();
    r->items.assign(items, items + n_items);
    for (genie_dim_t& it : r->items)
        it.query = 0;
    r->k = k;
    r->dst_ids = dst_ids;
    r->dst_counts = dst_counts;
    r->mem = request_bytes(n_items, k, (s->ix->d.dense || s->ix->d.sets) && s->cfg.mode != 0, s->ix->d.n_objects);
    // validate alone so that one bad request cannot fail a whole batch
    if (int rc = gh::validate_items(s->ix->d, r->items.data(), n_items, 1))
        return rc;
    r->t_submit = Clock::now();
    bool wake;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        r->ticket = s->next_ticket++;
        *ticket = r->ticket;
        // with requests already pending the scheduler needs no wake: it has
        // filled every free executor and waits for a completion
        wake = s->pending.empty();
        s->pending.push_back(r.get());
        if (out)
            *out = r.get();
        s->reqs.emplace(r->ticket, std::move(r));
        if (wake)
            ++s->events;
    }
    if (wake)
        s->cv_work.notify_one();
    return GENIE_OK;
}

int genie_hip_server_submit(genie_server* s, const genie_dim_t* items, uint32_t n_items, uint32_t k,
                            uint64_t* ticket)
{
    if (!s || (!items && n_items) || !ticket)
        return fail_s(GENIE_E_INVALID, "null argument");
    return submit_request(s, items, n_items, k, ticket, nullptr);
}

// Waits for a ticket; returns its results and its submit and completion times.
static int wait_ticket(genie_server* s, uint64_t ticket, int32_t* out_ids, int32_t* out_counts,
                       Clock::time_point* t_submit, Clock::time_point* t_done)
{
    std::unique_ptr<Request> r;
    {
        std::unique_lock<std::mutex> lk(s->mu);
        auto it = s->reqs.find(ticket);
        if (it == s->reqs.end())
            return fail_s(GENIE_E_INVALID, "unknown ticket");
        Request* rp = it->second.get();
        s->cv_done.wait(lk, [&] { return rp->done.load(std::memory_order_acquire); });
        r = std::move(it->second);
        s->reqs.erase(it);
    }
    *t_submit = r->t_submit;
    *t_done = r->t_done;
    if (r->rc)
        return fail_s(r->rc, r->err);
    if (out_ids)
        std::memcpy(out_ids, r->ids.data(), r->k * 4);
    if (out_counts)
        std::memcpy(out_counts, r->counts.data(), r->k * 4);
    return GENIE_OK;
}

int genie_hip_server_wait(genie_server* s, uint64_t ticket, int32_t* out_ids, int32_t* out_counts,
                          double* latency_ms)
{
    if (!s)
        return fail_s(GENIE_E_INVALID, "null server");
    Clock::time_point ts, td;
    const int rc = wait_ticket(s, ticket, out_ids, out_counts, &ts, &td);
    if (latency_ms)
        *latency_ms = std::chrono::duration<double, std::milli>(td - ts).count();
    return rc;
}

int genie_hip_server_replay(genie_server* s, const genie_dim_t* items, const uint64_t* item_ptr,
                            const uint32_t* k_per_request, const double* arrival_s, uint32_t n_requests,
                            double* out_latency_ms, int32_t* out_ids, int32_t* out_counts, double* wall_s)
{
    if (!s || !item_ptr || !k_per_request || !arrival_s || !out_latency_ms)
        return fail_s(GENIE_E_INVALID, "null argument");
    std::vector<uint64_t> tickets(n_requests, 0);
    std::vector<Request*> reqp(n_requests, nullptr);
    // each request's results go straight to its rows of the caller's arrays
    const bool direct = out_ids && out_counts;
    std::vector<uint64_t> offs(n_requests + 1, 0);
    for (uint32_t r = 0; r < n_requests; ++r)
        offs[r + 1] = offs[r] + k_per_request[r];
    std::vector<std::atomic<int>> submitted(n_requests);
    for (auto& a : submitted)
        a.store(0);
    std::atomic<int> gen_rc(GENIE_OK);
    std::string gen_err;
    std::mutex gen_mu;
    const auto t0 = Clock::now();
    // open-loop generators: thread g submits requests g, g + G, ... each at its
    // arrival time, so that one thread's submission cost does not cap the load
    constexpr uint32_t G = 4;
    std::vector<std::thread> gens;
    for (uint32_t g = 0; g < G; ++g)
        gens.emplace_back([&, g] {
            for (uint32_t r = g; r < n_requests; r += G) {
                std::this_thread::sleep_until(t0 + std::chrono::duration_cast<Clock::duration>(
                                                       std::chrono::duration<double>(arrival_s[r])));
                const int rc = items || item_ptr[r + 1] == item_ptr[r]
                                   ? submit_request(s, items + item_ptr[r], (uint32_t)(item_ptr[r + 1] - item_ptr[r]),
                                                    k_per_request[r], &tickets[r], &reqp[r],
                                                    direct ? out_ids + offs[r] : nullptr,
                                                    direct ? out_counts + offs[r] : nullptr)
                                   : fail_s(GENIE_E_INVALID, "null items");
                if (rc) {
                    std::lock_guard<std::mutex> lk(gen_mu);
                    gen_rc = rc;
                    gen_err = genie_hip_last_error();
                    submitted[r].store(-1);
                    // later requests are still submitted so the trace keeps its shape
                    continue;
                }
                submitted[r].store(1);
            }
        });
    // results in request order: each request's completion flag is watched
    // without the server's lock, and finished requests are erased in batches
    // (one lock for many), so that collecting does not cap the load
    int rc = GENIE_OK;
    uint64_t off = 0;
    std::vector<uint64_t> finished;
    auto erase_finished = [&] {
        std::lock_guard<std::mutex> lk(s->mu);
        for (uint64_t t : finished)
            s->reqs.erase(t);
        finished.clear();
    };
    for (uint32_t r = 0; r < n_requests; ++r) {
        while (submitted[r].load() == 0)  // the generator has not reached request r yet
            std::this_thread::sleep_for(std::chrono::microseconds(50));
        const uint32_t k = k_per_request[r];
        if (submitted[r].load() < 0) {
            out_latency_ms[r] = -1;
        } else {
            Request* rp = reqp[r];
            while (!rp->done.load(std::memory_order_acquire))
                std::this_thread::sleep_for(std::chrono::microseconds(20));
            // latency from the scheduled arrival: a generator running late
            // counts against the server, not for it
            const auto due = t0 + std::chrono::duration_cast<Clock::duration>(std::chrono::duration<double>(arrival_s[r]));
            out_latency_ms[r] = std::chrono::duration<double, std::milli>(rp->t_done - due).count();
            if (rp->rc) {
                if (!rc)
                    rc = fail_s(rp->rc, rp->err);
            } else if (!direct) {
                if (out_ids)
                    std::memcpy(out_ids + off, rp->ids.data(), k * 4ull);
                if (out_counts)
                    std::memcpy(out_counts + off, rp->counts.data(), k * 4ull);
            }
            finished.push_back(tickets[r]);
            if (finished.size() >= 256)
                erase_finished();
        }
        off += k;
    }
    erase_finished();
    for (std::thread& t : gens)
        t.join();
    if (wall_s)
        *wall_s = std::chrono::duration<double>(Clock::now() - t0).count();
    if (gen_rc != GENIE_OK)
        return fail_s(gen_rc, gen_err);
    return rc;
}

int genie_hip_server_stats(genie_server* s, uint64_t* batches, uint64_t* requests, uint64_t* max_inflight)
{
    if (!s)
        return fail_s(GENIE_E_INVALID, "null server");
    std::lock_guard<std::mutex> lk(s->mu);
    if (batches)
        *batches = s->n_batches;
    if (requests)
        *requests = s->n_done;
    if (max_inflight)
        *max_inflight = s->max_inflight;
    return GENIE_OK;
}

int genie_hip_server_phase_seconds(genie_server* s, double* out)
{
    if (!s || !out)
        return fail_s(GENIE_E_INVALID, "null argument");
    for (int i = 0; i < 4; ++i) {
        uint64_t ns = 0;
        for (const Executor& e : s->ex)
            ns += e.phase_ns[i].load(std::memory_order_relaxed);
        out[i] = (double)ns * 1e-9;
    }
    uint64_t reruns = 0;
    for (const Executor& e : s->ex)
        reruns += e.reruns.load(std::memory_order_relaxed);
    out[4] = (double)reruns;
    return GENIE_OK;
}

int genie_hip_server_phase_max_seconds(genie_server* s, double* out, int reset)
{
    if (!s || !out)
        return fail_s(GENIE_E_INVALID, "null argument");
    for (int i = 0; i < 4; ++i) {
        uint64_t ns = 0;
        for (Executor& e : s->ex)
            ns = std::max(ns, reset ? e.phase_max_ns[i].exchange(0, std::memory_order_relaxed)
                                    : e.phase_max_ns[i].load(std::memory_order_relaxed));
        out[i] = (double)ns * 1e-9;
    }
    out[4] = (double)(reset ? s->queue_max_ns.exchange(0) : s->queue_max_ns.load()) * 1e-9;
    return GENIE_OK;
}

}  // extern "C"
