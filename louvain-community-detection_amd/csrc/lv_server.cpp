// lv_server.cpp -- serving loop: continuous admission of independent graphs
// onto one GPU under an HBM budget.
//
// Requests enter a FIFO.  `slots` worker threads each own one HIP stream
// (through their lv_graph); a worker takes the first queued request whose
// device footprint fits the remaining budget, uploads it, runs the full pass
// loop with its state resident on the device, and publishes the result.
// Requests that arrive while others run are admitted as soon as a slot and
// memory free up, and their kernels overlap on the device with the running
// ones.  Each request's result is the single-graph result (lv_run), bit for
// bit: graphs share nothing but the device.
#include <hip/hip_runtime.h>

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

namespace {

struct Request {
    uint64_t id = 0;
    uint32_t n = 0;
    std::vector<uint32_t> row, col;
    std::vector<float> w;
    uint64_t bytes = 0;
    double t_submit = 0.0;
};

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

// first queued request that fits the remaining budget (FIFO order otherwise)
Request *take_fitting(lv_server *s) {
    for (auto it = s->queue.begin(); it != s->queue.end(); ++it) {
        if (s->used + (*it)->bytes <= s->budget) {
            Request *r = *it;
            s->queue.erase(it);
            return r;
        }
    }
    return nullptr;
}

void worker(lv_server *s) {
    (void)hipSetDevice(s->device);
    for (;;) {
        Request *r = nullptr;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            s->cv_work.wait(lk, [&] { return s->stop || (r = take_fitting(s)) != nullptr; });
            if (!r) return;  // stopping and nothing admissible
            s->used += r->bytes;
            s->running++;
        }
        lv_request_result res;
        memset(&res, 0, sizeof(res));
        res.id = r->id;
        res.t_submit_ms = r->t_submit;
        res.t_start_ms = s->now_ms();
        std::vector<uint32_t> assign;
        lv_graph *g = nullptr;
        int rc = lv_graph_create(s->device, r->n, r->row.data(), r->col.data(), r->w.data(), &g);
        if (rc == LV_OK) {
            lv_result lr;
            rc = lv_run(g, s->min_gain, &lr);
            if (rc == LV_OK) {
                res.q_final = lr.q_final;
                res.passes = lr.passes;
                res.n_final = lr.n_final;
                res.total_iters = lr.total_iters;
                res.sum_edges = lr.sum_edges;
                if (s->keep_assignments) {
                    assign.resize(r->n);
                    rc = lv_get_assignment(g, assign.data());
                }
            }
            lv_graph_destroy(g);
        }
        res.status = rc;
        res.t_end_ms = s->now_ms();
        {
            std::lock_guard<std::mutex> lk(s->mu);
            s->used -= r->bytes;
            s->running--;
            s->done.push_back(res);
            if (s->keep_assignments && rc == LV_OK) s->assignments[r->id] = std::move(assign);
        }
        delete r;
        s->cv_work.notify_all();  // freed memory may admit a queued request
        s->cv_done.notify_all();
    }
}

}  // namespace

extern "C" {

uint64_t lv_graph_footprint(uint32_t n, uint64_t ne) {
    // every buffer lv_graph_create allocates, plus the largest hash arenas a
    // pass can request (bounded by 4x the entries)
    return 176ull * ((uint64_t)n + 1) + 176ull * ne + (1ull << 20);
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
    if (!s || !row) return LV_ERR_ARG;
    const uint64_t ne = row[n];
    const uint64_t bytes = lv_graph_footprint(n, ne);
    if (bytes > s->budget) return LV_ERR_NOMEM;  // can never be admitted
    Request *r = new (std::nothrow) Request();
    if (!r) return LV_ERR_NOMEM;
    r->id = id;
    r->n = n;
    r->row.assign(row, row + (size_t)n + 1);
    r->col.assign(col, col + ne);
    r->w.assign(w, w + ne);
    r->bytes = bytes;
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
