// lv_server.cpp -- serving loop: continuous admission of independent graphs
// onto one GPU under an HBM budget.
//
// Requests enter a FIFO.  `slots` worker threads each own one device-resident
// graph handle (buffers and HIP stream) that they reuse from request to
// request: a worker takes the first queued request that fits its buffers, or
// whose footprint fits the budget once the worker's own buffers are grown or
// replaced; it uploads the request, runs the full pass loop with the state
// resident on the device and publishes the result.  Requests that arrive
// while others run are admitted as soon as a slot and memory free up, and
// their kernels overlap on the device with the running ones.  Reusing the
// buffers keeps device allocation and release (which synchronise the whole
// device) out of the steady state.  A worker whose buffers block every queued
// request releases them.  Each request's result is the single-graph result
// (lv_run), bit for bit: graphs share nothing but the device.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstdint>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <new>
#include <thread>
#include <vector>

#include "../../include/louvain_hip.h"
#include "lv_host.h"

namespace {

struct Request {
    uint64_t id = 0;
    uint32_t n = 0;
    uint64_t ne = 0;
    int S = 0;  // fixed-point scale bits (from the worker's input checks)
    std::vector<uint32_t> row, col;
    std::vector<float> w;
    double t_submit = 0.0;
};

// one worker's resident graph handle and the budget it holds
struct Slot {
    lv_graph *g = nullptr;
    uint32_t n_cap = 0;
    uint64_t e_cap = 0;
    uint64_t bytes = 0;
};

// what a worker must do with its buffers to take a request
enum class Fit { kNone, kReuse, kGrow, kReplace };

}  // namespace

struct lv_server {
    int device = 0;
    float min_gain = 0.0f;
    uint64_t budget = 0, used = 0;
    bool keep_assignments = false;
    bool stop = false;
    uint32_t running = 0;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Request *> queue;
    std::vector<lv_request_result> done;
    std::map<uint64_t, std::vector<uint32_t>> assignments;
    std::vector<std::thread> workers;
    std::chrono::steady_clock::time_point t0;
    double now_ms() const {
        return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0)
            .count();
    }
};

namespace {

// How slot `sl` can take request r under the budget (caller holds the lock).
Fit fit_of(const lv_server *s, const Slot &sl, const Request *r, uint32_t *n_cap,
           uint64_t *e_cap, uint64_t *bytes) {
    if (sl.g && r->n <= sl.n_cap && r->ne <= sl.e_cap) return Fit::kReuse;
    const uint64_t others = s->used - sl.bytes;
    const uint32_t gn = std::max(sl.n_cap, r->n);
    const uint64_t ge = std::max(sl.e_cap, r->ne);
    const uint64_t grow = lv_graph_footprint(gn, ge);
    if (others + grow <= s->budget) {
        *n_cap = gn;
        *e_cap = ge;
        *bytes = grow;
        return Fit::kGrow;
    }
    const uint64_t exact = lv_graph_footprint(r->n, r->ne);
    if (others + exact <= s->budget) {
        *n_cap = r->n;
        *e_cap = r->ne;
        *bytes = exact;
        return Fit::kReplace;
    }
    return Fit::kNone;
}

void publish(lv_server *s, lv_request_result &res, std::vector<uint32_t> &assign, int rc) {
    res.status = rc;
    res.t_end_ms = s->now_ms();
    std::lock_guard<std::mutex> lk(s->mu);
    s->running--;
    s->done.push_back(res);
    if (s->keep_assignments && rc == LV_OK) s->assignments[res.id] = std::move(assign);
}

void worker(lv_server *s) {
    (void)hipSetDevice(s->device);
    Slot sl;
    for (;;) {
        Request *r = nullptr;
        Fit fit = Fit::kNone;
        uint32_t n_cap = 0;
        uint64_t e_cap = 0, bytes = 0;
        bool release = false;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            s->cv_work.wait(lk, [&] {
                if (s->stop) return true;
                for (auto it = s->queue.begin(); it != s->queue.end(); ++it) {
                    fit = fit_of(s, sl, *it, &n_cap, &e_cap, &bytes);
                    if (fit != Fit::kNone) {  // first admissible in FIFO order
                        r = *it;
                        s->queue.erase(it);
                        return true;
                    }
                }
                // nothing admissible: idle buffers may be what blocks the queue
                if (!s->queue.empty() && sl.g) {
                    release = true;
                    return true;
                }
                return false;
            });
            if (!r && !release) break;  // stopping
            if (r) {
                if (fit != Fit::kReuse) {  // reserve the new size now
                    s->used = s->used - sl.bytes + bytes;
                    sl.bytes = bytes;
                }
                s->running++;
            }
        }
        if (release) {
            lv_graph_destroy(sl.g);
            const uint64_t freed = sl.bytes;
            sl = Slot();
            {
                std::lock_guard<std::mutex> lk(s->mu);
                s->used -= freed;
            }
            s->cv_work.notify_all();
            continue;
        }
        lv_request_result res;
        memset(&res, 0, sizeof(res));
        res.id = r->id;
        res.t_submit_ms = r->t_submit;
        res.t_start_ms = s->now_ms();
        std::vector<uint32_t> assign;
        int rc = LV_OK;
        if (fit != Fit::kReuse) {
            if (sl.g) lv_graph_destroy(sl.g);
            sl.g = nullptr;
            // no pristine copy: every request is uploaded once and run once
            rc = lv_host::graph_new(s->device, n_cap, e_cap, false, &sl.g);
            if (rc == LV_OK) {
                sl.n_cap = n_cap;
                sl.e_cap = e_cap;
            } else {  // give the reservation back
                std::lock_guard<std::mutex> lk(s->mu);
                s->used -= sl.bytes;
                sl = Slot();
            }
        }
        // lv_graph_create's input checks, off the submitting thread
        if (rc == LV_OK)
            rc = lv_host::check_input(r->n, r->row.data(), r->col.data(), r->w.data(), &r->S);
        if (rc == LV_OK)
            rc = lv_host::graph_upload(sl.g, r->n, r->row.data(), r->col.data(), r->w.data(), r->S);
        if (rc == LV_OK) {
            lv_result lr;
            rc = lv_run(sl.g, s->min_gain, &lr);
            if (rc == LV_OK) {
                res.q_final = lr.q_final;
                res.passes = lr.passes;
                res.n_final = lr.n_final;
                res.total_iters = lr.total_iters;
                res.sum_edges = lr.sum_edges;
                if (s->keep_assignments) {
                    assign.resize(r->n);
                    rc = lv_get_assignment(sl.g, assign.data());
                }
            }
        }
        publish(s, res, assign, rc);
        delete r;
        s->cv_work.notify_all();
        s->cv_done.notify_all();
    }
    if (sl.g) lv_graph_destroy(sl.g);
}

}  // namespace

extern "C" {

uint64_t lv_graph_footprint(uint32_t n, uint64_t ne) {
    // every device buffer of a handle, the arenas that grow on demand at their
    // worst case, plus allocation granularity (and a captured sweep graph
    // with LV_GRAPH=1)
    return lv_host::footprint(n, ne) + (4ull << 20);
}

int lv_server_create(int device, uint32_t slots, uint64_t hbm_budget_bytes, float min_gain,
                     int keep_assignments, lv_server **out) {
    if (!out || slots == 0 || hbm_budget_bytes == 0) return LV_ERR_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) {
        (void)hipGetLastError();
        return LV_ERR_NO_DEVICE;
    }
    lv_server *s = new (std::nothrow) lv_server();
    if (!s) return LV_ERR_NOMEM;
    s->device = device;
    s->budget = hbm_budget_bytes;
    s->min_gain = min_gain;
    s->keep_assignments = keep_assignments != 0;
    s->t0 = std::chrono::steady_clock::now();
    for (uint32_t i = 0; i < slots; ++i) s->workers.emplace_back(worker, s);
    *out = s;
    return LV_OK;
}

int lv_server_submit(lv_server *s, uint64_t id, uint32_t n, const uint32_t *row,
                     const uint32_t *col, const float *w) {
    if (!s || !row || n >= 0x80000000u || row[0] != 0) return LV_ERR_ARG;
    const uint64_t ne = row[n];
    if (ne && (!col || !w)) return LV_ERR_ARG;
    if (lv_graph_footprint(n, ne) > s->budget) return LV_ERR_NOMEM;  // never admissible
    Request *r = new (std::nothrow) Request();
    if (!r) return LV_ERR_NOMEM;
    r->id = id;
    r->n = n;
    r->ne = ne;
    r->row.assign(row, row + (size_t)n + 1);
    r->col.assign(col, col + ne);
    r->w.assign(w, w + ne);
    {
        std::lock_guard<std::mutex> lk(s->mu);
        r->t_submit = s->now_ms();
        s->queue.push_back(r);
    }
    s->cv_work.notify_one();
    return LV_OK;
}

int lv_server_poll(lv_server *s, lv_request_result *out, uint32_t max, uint32_t *count) {
    if (!s || !count) return LV_ERR_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    uint32_t k = 0;
    while (k < max && !s->done.empty()) {
        out[k++] = s->done.front();
        s->done.erase(s->done.begin());
    }
    *count = k;
    return LV_OK;
}

int lv_server_get_assignment(lv_server *s, uint64_t id, uint32_t *assignment, uint32_t n) {
    if (!s || !assignment) return LV_ERR_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->assignments.find(id);
    if (it == s->assignments.end() || it->second.size() != n) return LV_ERR_STATE;
    memcpy(assignment, it->second.data(), (size_t)n * sizeof(uint32_t));
    s->assignments.erase(it);
    return LV_OK;
}

int lv_server_drain(lv_server *s) {
    if (!s) return LV_ERR_ARG;
    std::unique_lock<std::mutex> lk(s->mu);
    s->cv_done.wait(lk, [&] { return s->queue.empty() && s->running == 0; });
    return LV_OK;
}

int lv_server_destroy(lv_server *s) {
    if (!s) return LV_ERR_ARG;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    for (auto &t : s->workers) t.join();
    for (Request *r : s->queue) delete r;
    delete s;
    return LV_OK;
}

}  // extern "C"
