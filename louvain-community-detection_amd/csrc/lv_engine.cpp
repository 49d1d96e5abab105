// lv_engine.cpp -- host engine behind the C ABI (include/louvain_hip.h).
//
// One lv_graph owns a HIP stream (default priority) and all device buffers
// of one graph, sized once at creation for the original graph.  The pass
// driver mirrors modularityOptimization (modularity.cu:438-520) and
// aggregation (aggregation.cu:387-462); the per-iteration host round trip is
// one small D2H of exact counters (S_in, sum a_c^2 limbs, moved count), from
// which the FP32 modularity is formed exactly as the reference writes it.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstddef>
#include <cstring>
#include <new>
#include <thread>
#include <memory>
#include <vector>

#include "../../include/louvain_hip.h"
#include "lv_device.h"
#include "lv_host.h"

using lv::kNumAggClasses;
using lv::kNumClasses;
typedef unsigned long long ull;
typedef unsigned __int128 u128;

#define LV_HIP(expr)                                   \
    do {                                               \
        hipError_t _e = (expr);                        \
        if (_e != hipSuccess) {                        \
            (void)hipGetLastError();                   \
            return LV_ERR_HIP;                         \
        }                                              \
    } while (0)

namespace {

constexpr uint64_t kFixedLimit = 1ull << 61;
constexpr uint32_t kMaxIters = 100000;

// device scratch counters, one allocation
// Device counters.  The head is what the host reads back each sweep; the
// spread slots behind it are collapsed into `tot` on the device first.
struct Counters {
    uint32_t counts[kNumClasses];
    uint32_t cursor[kNumClasses];
    uint32_t acounts[kNumAggClasses];
    uint32_t acursor[kNumAggClasses];
    uint32_t err;
    uint32_t pad;
    ull degsum[kNumClasses];
    ull tot[6];                 // acc | limbs[4] | moved, collapsed
    // spread slots (not copied back)
    ull acc[lv::kSpread];       // k/s fixed total
    ull limbs[4][lv::kSpread];  // sum a^2 limbs
    ull moved[lv::kSpread];
    lv::SweepTail tail;  // host staging of the device copy a folded pass loop reads
};
constexpr size_t kCountersHead = offsetof(Counters, acc);
static_assert(offsetof(Counters, limbs) == offsetof(Counters, acc) + lv::kSpread * 8, "layout");
static_assert(offsetof(Counters, moved) == offsetof(Counters, acc) + 5 * lv::kSpread * 8, "layout");

struct ProfRec {
    int fam;
    hipEvent_t a, b;
    double bytes;
};

// graph states in batch mode
enum : uint32_t { kGsActive = lv::kGraphActive, kGsWait = 2, kGsDone = 3 };

}  // namespace

// Bump allocator over a caller-provided device region (the fixed scratch
// array of lv_run_noalloc): every buffer of a handle is carved from it with
// 256-byte alignment, nothing is ever freed.
struct Carver {
    char *base = nullptr;
    size_t cap = 0, used = 0;
    bool count_only = false;  // only add up the sizes (no region behind it)
};

struct lv_graph {
    int device = 0;
    // the stream every call on the graph runs on: the handle's own, or one
    // the caller set (lv_graph_set_stream)
    hipStream_t stream = nullptr;
    hipStream_t own_stream = nullptr;
    // non-null: device buffers come from a fixed region, no allocation
    Carver *carve = nullptr;
    // second stream of the handle: bucket step 7 runs the heavy rows' kernel
    // chain beside the block classes (fork and join by events; two branches
    // of the sweep graph when it is captured)
    hipStream_t stream2 = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    // ends of the two groups of sweeps in flight (pipelined pass loop)
    hipEvent_t ev_launch[2] = {nullptr, nullptr};
    // the pass loop's stop-rule arguments on the device (folded decide)
    lv::SweepTail *tail_d = nullptr;
    uint32_t n = 0, n0 = 0;
    uint64_t ne = 0, ne0 = 0;
    uint32_t n_cap = 0;  // buffer capacities (lv_graph_create: the input's size)
    uint64_t e_cap = 0;
    // current and alternate CSR
    uint32_t *row = nullptr, *col = nullptr;
    float *w = nullptr;
    uint32_t *row2 = nullptr, *col2 = nullptr;
    float *w2 = nullptr;
    // pristine copy of the uploaded graph in host memory (lv_graph_reset
    // uploads it again); null on server and carved handles
    uint32_t *row0 = nullptr, *col0 = nullptr;
    float *w0 = nullptr;
    // pass state
    float *k = nullptr;
    uint32_t *comm = nullptr, *comm_sz = nullptr, *new_comm = nullptr;
    ull *A = nullptr;
    // second copy of the community state (passes of small single graphs): a
    // sweep's bucket steps read one copy and apply their moves to the other
    // (both equal between sweeps)
    uint32_t *comm2 = nullptr, *comm_sz2 = nullptr;
    ull *A2 = nullptr;
    uint32_t *init_comm = nullptr;
    uint32_t *order = nullptr;
    // locality order of the current pass (position -> input id and inverse),
    // scratch for the renumbering under it
    uint32_t *lid = nullptr, *pmap = nullptr, *rtmp = nullptr, *rrank = nullptr;
    bool relabeled = false;  // the current pass graph is in locality order
    bool order_ident = false;  // `order` is the identity (classes in locality order)
    // the contraction wrote the graph in its locality order and classified it
    bool prelabeled = false;
    uint32_t cls_heavy_min = 0;  // heavy threshold of the current classes
    lv_tuning tune = {4096u, 512u, -1, 1u << 20, 2u, 1024u, 1280u, 384u, 1u};  // lv_graph_set_tuning
    // classification (host copies)
    uint32_t cls_count[kNumClasses] = {};
    uint32_t cls_off[kNumClasses + 1] = {};
    ull cls_deg[kNumClasses] = {};
    // heavy move scratch (class 8)
    ull *heavy_caps = nullptr, *heavy_off = nullptr;
    uint32_t *heavy_nch = nullptr, *heavy_coff = nullptr;
    uint32_t heavy_chunks = 0;
    // global-table arena, 16-byte slots: heavy rows' tables in sweeps, large
    // communities' tables in the contraction (never live at the same time)
    ull *gtab = nullptr;
    size_t gtab_slots = 0;
    ull *hscr = nullptr;
    size_t hscr_cap = 0;
    // partitioned heavy path
    uint32_t *hp_nbuck = nullptr, *hp_bstart = nullptr, *hp_flag = nullptr;
    uint32_t *hp_bcnt = nullptr, *hp_bpos = nullptr, *hp_bcur = nullptr, *hp_btmp = nullptr;
    size_t hp_bcap = 0;
    uint32_t hp_nbuckets = 0;
    uint32_t *hp_bvtx = nullptr, *hp_cvtx = nullptr;  // bucket / chunk -> heavy index
    uint32_t *hp_flist = nullptr;  // count + chunks of flagged heavy vertices
    size_t hp_ccap = 0;
    uint32_t *dlist = nullptr, *eslot = nullptr;
    // aggregation scratch
    uint32_t *new_id = nullptr, *comm_deg = nullptr, *epos = nullptr, *mstart = nullptr;
    uint32_t *vnid = nullptr;  // new id of each vertex's community
    uint32_t *mcursor = nullptr, *members = nullptr, *cnt = nullptr, *aorder = nullptr;
    uint32_t *tkey = nullptr, *skey = nullptr;
    ull *tval = nullptr, *sval = nullptr;
    ull *agg_caps = nullptr, *agg_off = nullptr;
    uint32_t *agg_mdeg = nullptr, *agg_moff = nullptr, *agg_cnk = nullptr, *agg_coff = nullptr;
    uint32_t *agg_flist = nullptr;  // flagged large communities' chunks, then the communities
    size_t agg_flcap = 0;
    uint32_t *agg_dcnt = nullptr;
    uint32_t agg_nchunks = 0;
    uint32_t agg_nn = 0;  // new vertex count of the aggregation in progress
    // partitioned merge of large communities: per community / per bucket
    uint32_t *ap_nbk = nullptr, *ap_bstart = nullptr, *ap_cflag = nullptr, *ap_nflag = nullptr;
    uint32_t *ap_bcnt = nullptr, *ap_bpos = nullptr, *ap_bcur = nullptr, *ap_bnd = nullptr;
    uint32_t *ap_boff = nullptr, *ap_bowner = nullptr, *ap_btmp = nullptr;
    size_t ap_bcap = 0;

    uint32_t *scan_tmp32 = nullptr;
    ull *scan_tmp64 = nullptr;
    Counters *dc = nullptr;  // device
    Counters *hc = nullptr;  // pinned host
    // the pass loop's state on the device (and its pinned host copy); while
    // halt_on is set (sweeps captured for a device-decided loop) kernels get
    // the halt flag in their view
    lv::PassState *pst = nullptr, *pst_h = nullptr;
    lv::PassState *pst_hd = nullptr;  // device address of the mapped host copy
    bool halt_on = false;
    // scalars
    int S = 0;
    float scale = 1.0f, inv_scale = 1.0f;
    float m = 0.0f;
    bool pass_ready = false;
    uint32_t last_iters = 0;
    float last_q_before = 0.0f, last_q_after = 0.0f;  // of the last pass
    // batch mode (G > 0): G graphs in one segmented CSR, vertex ids global
    uint32_t G = 0;
    uint32_t *gid = nullptr;       // device: vertex -> graph
    float *gm_d = nullptr;         // device: m per graph
    uint32_t *gstate_d = nullptr;  // device: graph state
    uint32_t *voff_d = nullptr;    // device: G+1 current vertex offsets
    uint32_t *vgather = nullptr;   // device: G+1 scratch
    ull *gcnt = nullptr;           // device: per-graph acc[G] | limbs[4][G] | moved[G]
    std::vector<uint32_t> voff, voff0;  // host: current / original vertex offsets
    std::vector<uint32_t> gstate_h;     // host: initial graph states of a run
    // per graph on the device: pass loop variables and run totals; the
    // summary the host reads after each group of sweeps
    lv::GraphState *gst_d = nullptr;
    lv::BatchSummary *bsum_d = nullptr, *bsum_h = nullptr;
    lv::BatchSummary *bsum_hd = nullptr;  // device address of the mapped host copy
    std::vector<int> gstatus0;  // per graph: LV_ERR_RANGE if its weights do not fit
    // profiling
    bool prof_on = false;
    uint32_t prof_mask = ~0u;  // families that get event pairs
    std::vector<ProfRec> *capture_recs = nullptr;  // non-null while capturing a sweep
    std::vector<ProfRec> prof_pending;
    std::vector<hipEvent_t> ev_pool;
    lv_kernel_stats prof = {};
};

namespace {

// ---------------------------------------------------------------------------
// helpers
// ---------------------------------------------------------------------------
template <typename T>
int dmalloc(T **p, size_t n) {
    *p = nullptr;
    if (n == 0) n = 1;
    if (hipMalloc((void **)p, n * sizeof(T)) != hipSuccess) {
        (void)hipGetLastError();
        *p = nullptr;
        return LV_ERR_NOMEM;
    }
    return LV_OK;
}

template <typename T>
void dfree(T *&p) {
    if (p) (void)hipFree(p);
    p = nullptr;
}

// A buffer of the handle: hipMalloc, or the next 256-byte aligned piece of
// the fixed region (LV_ERR_LIMIT once it is exhausted).
template <typename T>
int galloc(lv_graph *g, T **p, size_t n) {
    if (!g->carve) return dmalloc(p, n);
    *p = nullptr;
    if (n == 0) n = 1;
    Carver &c = *g->carve;
    const size_t off = (c.used + 255) & ~(size_t)255;
    if (n > (c.cap - std::min(c.cap, off)) / sizeof(T)) return LV_ERR_LIMIT;
    *p = c.count_only ? nullptr : reinterpret_cast<T *>(c.base + off);
    c.used = off + n * sizeof(T);
    return LV_OK;
}

template <typename T>
void gfree(lv_graph *g, T *&p) {
    if (!g->carve) dfree(p);
    p = nullptr;
}

// The global-table arena holds at least `slots` empty slots.  Its users leave
// every slot they touch empty again, so it is filled only when it grows (on
// the handle's stream, before the kernels that use it).
int ensure_gtab(lv_graph *g, size_t slots) {
    if (slots <= g->gtab_slots) return LV_OK;
    gfree(g, g->gtab);
    g->gtab_slots = 0;
    if (int e = galloc(g, &g->gtab, 2 * slots)) return e;
    g->gtab_slots = slots;
    lv::launch_fill_slots(g->stream, g->gtab, slots);
    return LV_OK;
}

lv::DevView view(lv_graph *g) {
    lv::DevView v;
    v.row = g->row;
    v.col = g->col;
    v.w = g->w;
    v.comm = g->comm;
    v.comm_sz = g->comm_sz;
    v.new_comm = g->new_comm;
    v.A = g->A;
    v.k = g->k;
    v.n = g->n;
    v.S = g->S;
    v.scale = g->scale;
    v.m = g->m;
    v.gid = g->G ? g->gid : nullptr;
    v.gm = g->gm_d;
    v.gstate = g->gstate_d;
    v.G = g->G;
    v.lid = g->relabeled ? g->lid : nullptr;
    v.halt = !g->halt_on ? nullptr : g->G ? &g->bsum_d->halt : &g->pst->done;
    v.comm2 = nullptr;
    v.comm_sz2 = nullptr;
    v.A2 = nullptr;
    v.cu_verts = nullptr;
    v.cu_count = v.cu_nblk = 0;
    v.cu_moved = nullptr;
    v.vbase = -1;
    v.tail = nullptr;
    return v;
}

// exact-total accumulators: spread slots (single graph) or per graph (batch)
ull *acc_ptr(lv_graph *g) { return g->G ? g->gcnt : g->dc->acc; }
ull *limbs_ptr(lv_graph *g) { return g->G ? g->gcnt + g->G : &g->dc->limbs[0][0]; }
ull *moved_ptr(lv_graph *g) { return g->G ? g->gcnt + 5ull * g->G : g->dc->moved; }

hipEvent_t ev_get(lv_graph *g) {
    if (!g->ev_pool.empty()) {
        hipEvent_t e = g->ev_pool.back();
        g->ev_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) {
        (void)hipGetLastError();
        return nullptr;
    }
    return e;
}

// Whether independent kernel chains may run on the handle's second stream.
// While kernel-family timing is on they do not: a family's event window
// would otherwise also cover the other stream's kernels running beside it,
// and its durations would not be those of its own kernels. LV_ONE_STREAM
// keeps every chain on the one stream for profiler runs that must see the
// same kernel durations.
bool use_fork(const lv_graph *g) {
    static const bool off = getenv("LV_ONE_STREAM") != nullptr;
    return !off && !g->prof_on && g->stream2 != nullptr;
}

// RAII-free profiling bracket around a launch sequence of one family
struct ProfScope {
    lv_graph *g;
    int fam;
    double bytes;
    hipStream_t st;
    hipEvent_t a = nullptr;
    ProfScope(lv_graph *g_, int fam_, double bytes_, hipStream_t st_ = nullptr)
        : g(g_), fam(fam_), bytes(bytes_), st(st_ ? st_ : g_->stream) {
        if (g->prof_on && ((g->prof_mask >> fam) & 1u)) {
            a = ev_get(g);
            if (a) record(a);
        }
    }
    // inside a capture the record must become an external event node, or the
    // event carries no timestamp after a replay
    void record(hipEvent_t e) {
        if (g->capture_recs)
            (void)hipEventRecordWithFlags(e, st, hipEventRecordExternal);
        else
            (void)hipEventRecord(e, st);
    }
    ~ProfScope() {
        if (g->prof_on && a) {
            hipEvent_t b = ev_get(g);
            if (b) {
                record(b);
                // inside a captured sweep the pair becomes event-record nodes,
                // read back after every replay
                (g->capture_recs ? *g->capture_recs : g->prof_pending)
                    .push_back({fam, a, b, bytes});
            } else {
                g->ev_pool.push_back(a);
            }
        }
    }
};

void counters_d2h(lv_graph *g);
int sync_counters(lv_graph *g) {
    counters_d2h(g);
    LV_HIP(hipStreamSynchronize(g->stream));
    return LV_OK;
}
// The counters' head (class counts, cursors, error flag) without collapsing
// the spread sums first: for readers of those fields only (one launch fewer).
int sync_head(lv_graph *g) {
    LV_HIP(hipMemcpyAsync(g->hc, g->dc, kCountersHead, hipMemcpyDeviceToHost, g->stream));
    LV_HIP(hipStreamSynchronize(g->stream));
    return LV_OK;
}

int zero_counters(lv_graph *g) {
    // small fill kernels instead of memset nodes (safe under stream capture)
    lv::launch_fill_u32(g->stream, (uint32_t *)g->dc, 0u, sizeof(Counters) / 4);
    if (g->G) lv::launch_fill_u64(g->stream, g->gcnt, 0ull, 6ull * g->G);
    return LV_OK;
}

// D2H of the counters' head (and the per-graph counters in batch mode),
// after collapsing the spread slots on the device
void counters_d2h(lv_graph *g) {
    lv::launch_collapse(g->stream, g->dc->acc, 6, g->dc->tot);
    (void)hipMemcpyAsync(g->hc, g->dc, kCountersHead, hipMemcpyDeviceToHost, g->stream);
}

float fix2f_host(const lv_graph *g, int64_t x) { return (float)x * g->inv_scale; }

// ---------------------------------------------------------------------------
// pass init (modularity.cu:455-463)
// ---------------------------------------------------------------------------
// Degree classes of the current graph: counts, offsets and `order` (the
// vertices grouped by class).
int classify_order(lv_graph *g) {
    g->order_ident = false;
    zero_counters(g);
    {
        ProfScope ps(g, LV_K_CLASSIFY, 8.0 * g->n);
        // a batch lists the vertices of its active graphs only
        lv::launch_classify_count(g->stream, g->row, g->n, g->tune.heavy_min_degree,
                                  g->G ? g->gid : nullptr, g->gstate_d, g->dc->counts,
                                  g->dc->degsum);
    }
    int rc = sync_head(g);
    if (rc) return rc;
    g->cls_off[0] = 0;
    for (int c = 0; c < kNumClasses; ++c) {
        g->cls_count[c] = g->hc->counts[c];
        g->cls_deg[c] = g->hc->degsum[c];
        g->cls_off[c + 1] = g->cls_off[c] + g->cls_count[c];
    }
    // the scatter's cursors, staged in the host counters: the host writes
    // them again only after a later synchronisation of the stream, so the
    // copy needs no wait of its own
    for (int c = 0; c < kNumClasses; ++c) g->hc->cursor[c] = g->cls_off[c];
    LV_HIP(hipMemcpyAsync(g->dc->cursor, g->hc->cursor, sizeof(g->hc->cursor),
                          hipMemcpyHostToDevice, g->stream));
    {
        ProfScope ps(g, LV_K_CLASSIFY, 12.0 * g->n);
        // one graph: runs of up to 256 vertices per thread while at least
        // 1024 workgroups remain (the locality order inherits them: −1 % on
        // R-MAT-24); a batch keeps runs of 16, where longer runs measured
        // slower (+2.6 % at 64 on the 4096-graph batch)
        const uint32_t run = g->G ? 16u : std::min(256u, std::max(16u, g->n / (256u * 1024u)));
        lv::launch_classify_scatter(g->stream, g->row, g->n, g->tune.heavy_min_degree,
                                    g->G ? g->gid : nullptr, g->gstate_d, run, g->dc->cursor,
                                    g->order);
    }
    g->cls_heavy_min = g->tune.heavy_min_degree;
    return LV_OK;
}

// Heavy class: table arena (pow2 >= 2*deg slots per vertex), chunk offsets of
// the split kernels, per-vertex scratch.  Per heavy index, so the same for
// any numbering of the vertices.
int classify_heavy(lv_graph *g) {
    const uint32_t nh = g->cls_count[lv::kClassHeavy];
    g->heavy_chunks = 0;
    if (nh) {
        const uint32_t *hv = g->order + g->cls_off[lv::kClassHeavy];
        lv::launch_heavy_caps(g->stream, g->row, hv, nh, g->heavy_caps);
        lv::launch_scan_u64(g->stream, g->heavy_caps, nh, g->heavy_off, g->scan_tmp64);
        lv::launch_heavy_chunks(g->stream, g->row, hv, nh, g->heavy_nch);
        lv::launch_scan_u32(g->stream, g->heavy_nch, 0, nh, g->heavy_coff, g->scan_tmp32);
        lv::launch_hp_nbuck(g->stream, g->row, hv, nh, g->tune.bucket_entries, g->hp_nbuck);
        lv::launch_scan_u32(g->stream, g->hp_nbuck, 0, nh, g->hp_bstart, g->scan_tmp32);
        ull total = 0;
        uint32_t chunks = 0, nbuckets = 0;
        LV_HIP(hipMemcpyAsync(&total, g->heavy_off + nh, sizeof(ull), hipMemcpyDeviceToHost,
                              g->stream));
        LV_HIP(hipMemcpyAsync(&chunks, g->heavy_coff + nh, sizeof(uint32_t),
                              hipMemcpyDeviceToHost, g->stream));
        LV_HIP(hipMemcpyAsync(&nbuckets, g->hp_bstart + nh, sizeof(uint32_t),
                              hipMemcpyDeviceToHost, g->stream));
        LV_HIP(hipStreamSynchronize(g->stream));
        g->heavy_chunks = chunks;
        g->hp_nbuckets = nbuckets;
        if (nbuckets > g->hp_bcap) {
            gfree(g, g->hp_bcnt);
            gfree(g, g->hp_bpos);
            gfree(g, g->hp_bcur);
            gfree(g, g->hp_btmp);
            gfree(g, g->hp_bvtx);
            int e = galloc(g, &g->hp_bcnt, nbuckets);
            if (!e) e = galloc(g, &g->hp_bpos, (size_t)nbuckets + 1);
            if (!e) e = galloc(g, &g->hp_bcur, nbuckets);
            if (!e) e = galloc(g, &g->hp_btmp, lv::scan_tmp_elems(nbuckets) + 8);
            if (!e) e = galloc(g, &g->hp_bvtx, nbuckets);
            if (e) return e;
            g->hp_bcap = nbuckets;
        }
        if (chunks > g->hp_ccap) {
            gfree(g, g->hp_cvtx);
            gfree(g, g->hp_flist);
            int e = galloc(g, &g->hp_cvtx, chunks);
            if (!e) e = galloc(g, &g->hp_flist, (size_t)chunks + 1);
            if (e) return e;
            g->hp_ccap = chunks;
        }
        // owner of every chunk and bucket, so no kernel searches for it
        lv::launch_owner_fill(g->stream, g->heavy_coff, nh, g->hp_cvtx);
        lv::launch_owner_fill(g->stream, g->hp_bstart, nh, g->hp_bvtx);
        if (int e = ensure_gtab(g, (size_t)total)) return e;
        if ((size_t)nh * 130 > g->hscr_cap) {
            gfree(g, g->hscr);
            if (int e = galloc(g, &g->hscr, (size_t)nh * 130)) return e;
            g->hscr_cap = (size_t)nh * 130;
        }
    }
    return LV_OK;
}

int classify(lv_graph *g) {
    int rc = classify_order(g);
    return rc ? rc : classify_heavy(g);
}

lv::HeavySplit heavy_view(lv_graph *g) {
    lv::HeavySplit h;
    const uint32_t nh = g->cls_count[lv::kClassHeavy];
    h.verts = g->order + g->cls_off[lv::kClassHeavy];
    h.nh = nh;
    h.nchunks = g->heavy_chunks;
    h.coff = g->heavy_coff;
    h.cvtx = g->hp_cvtx;
    h.bvtx = g->hp_bvtx;
    h.slot_off = g->heavy_off;
    h.tab = g->gtab;
    h.dlist = g->dlist;
    h.eslot = g->eslot;
    h.hcnt = g->hscr;
    h.hcew = g->hscr + nh;
    h.hlane = g->hscr + 2 * (size_t)nh;
    h.nbuck = g->hp_nbuck;
    h.bstart = g->hp_bstart;
    h.nbuckets = g->hp_nbuckets;
    h.bucket_entries = g->tune.bucket_entries;
    h.bcnt = g->hp_bcnt;
    h.bpos = g->hp_bpos;
    h.bcur = g->hp_bcur;
    // heavy-entry space buffers share memory with the aggregation scratch
    // (never live at the same time)
    h.pc = g->tkey;
    h.pj = g->skey;
    h.pw = reinterpret_cast<float *>(g->tval);
    h.hflag = g->hp_flag;
    h.flist = g->hp_flist;
    // test hook LV_HEAVY_GLOBAL: 1 every heavy vertex on the global tables,
    // 2 the odd ones after the partitioned path (both paths in one step)
    static const char *hg = getenv("LV_HEAVY_GLOBAL");
    h.force_global = hg ? (atoi(hg) == 2 ? 2 : 1) : 0;
    // test hook LV_HEAVY_DMAX: fewer distinct communities per LDS bucket, so
    // buckets overflow and their vertices fall back to the global tables
    static const char *hd = getenv("LV_HEAVY_DMAX");
    h.dmax = (hd && atoi(hd) > 0) ? (uint32_t)atoi(hd) : 0xffffffffu;
    h.scan_tmp = g->hp_btmp;
    h.err = &g->dc->err;
    return h;
}

// groups of the pinned-order sums: classes 0-1, 2, 3, 4, 5 (rows of at most
// 84 entries, fewer lanes per vertex than the reference's 32 with the same
// bits) and 6-11 (32 lanes, 8 rounds of loads in flight), all in one launch
// (k_r1_groups).
// masked: 0 k_v, 1 s_v, 2 s_v with singleton communities (pass init).  The
// masked sums skip class 0: an empty row adds nothing.
// base: the view to read (two copies: the copy the sweep's last step wrote,
// carrying that step's catch-up into the other copy), view(g) if null.
// acc: the spread slots the sums go to (acc_ptr(g) if null).
int r1_all(lv_graph *g, int masked, int init_state, const lv::DevView *base = nullptr,
           ull *acc = nullptr) {
    if (!acc) acc = acc_ptr(g);
    lv::DevView v = base ? *base : view(g);
    if (init_state) {  // pass init sets both copies of the community state
        v.comm2 = g->comm2;
        v.comm_sz2 = g->comm_sz2;
        v.A2 = g->A2;
    }
    // the heavy rows' staged sums run on the second stream beside the class
    // groups (every group adds its exact total into the counters atomically)
    const bool fork = use_fork(g) && g->cls_count[lv::kClassHeavy] != 0 &&
                      g->cls_off[lv::kClassHeavy] > g->cls_off[masked ? 1 : 0];
    if (fork) {
        (void)hipEventRecord(g->ev_fork, g->stream);
        (void)hipStreamWaitEvent(g->stream2, g->ev_fork, 0);
        ProfScope ps(g, LV_K_R1, (masked ? 16.0 : 8.0) * (double)g->cls_deg[lv::kClassHeavy],
                     g->stream2);
        lv::launch_heavy_r1(g->stream2, v, heavy_view(g), (float *)g->eslot, masked, init_state,
                            acc);
    }
    // classes of each group (contiguous in `order`)
    const int grp[lv::kR1Groups][2] = {{0, 1}, {2, 2}, {3, 3}, {4, 4}, {5, 5},
                                       {6, lv::kClassHeavy - 1}};
    lv::R1Groups L{};
    uint32_t nb = 0;
    double bytes = 0;
    for (int i = lv::kR1Groups - 1; i >= 0; --i) {  // the grid lists the longest rows first
        const int c0 = (masked && grp[i][0] == 0) ? 1 : grp[i][0];
        const uint32_t a = g->cls_off[c0], b = g->cls_off[grp[i][1] + 1];
        double ent = 0;
        for (int c = c0; c <= grp[i][1]; ++c) ent += (double)g->cls_deg[c];
        const uint32_t vpb = lv::kR1GroupVertices[i];
        L.verts[i] = g->order + a;
        L.vbase[i] = g->order_ident ? (int32_t)a : -1;
        L.count[i] = b - a;
        nb += (b - a + vpb - 1) / vpb;
        L.bend[i] = nb;
        bytes += (masked ? 12.0 : 8.0) * ent + 16.0 * (b - a);
    }
    if (masked) {  // the sum of a_c^2 of the same snapshot rides in the same launch
        L.sq_verts = g->order;
        L.sq_vbase = g->order_ident ? 0 : -1;
        L.sq_count = g->cls_off[kNumClasses];  // every listed vertex (a batch: active graphs')
        L.sq_blocks = lv::sumsq_blocks(L.sq_count);
        L.limbs = limbs_ptr(g);
        bytes += 8.0 * g->n;
    }
    {
        ProfScope ps(g, LV_K_R1, bytes + 16.0 * v.cu_count);
        lv::launch_r1_groups(g->stream, v, L, masked, init_state, acc);
    }
    if (fork) {
        (void)hipEventRecord(g->ev_join, g->stream2);
        (void)hipStreamWaitEvent(g->stream, g->ev_join, 0);
    } else if (g->cls_count[lv::kClassHeavy]) {  // heavy rows: staged, LDS-streamed pinned sums
        ProfScope ps(g, LV_K_R1, (masked ? 16.0 : 8.0) * (double)g->cls_deg[lv::kClassHeavy]);
        lv::launch_heavy_r1(g->stream, v, heavy_view(g), (float *)g->eslot, masked, init_state,
                            acc);
    }
    return LV_OK;
}

// FP32 modularity from the exact counters (modularity.cu:428-436)
float q_from_counters(const lv_graph *g) {
    const Counters *c = g->hc;
    float s_in = fix2f_host(g, (int64_t)c->tot[0]);
    u128 sq = 0;
    for (int l = 0; l < 4; ++l) sq += (u128)c->tot[1 + l] << (32 * l);
    float a2 = ldexpf((float)sq, -2 * g->S);
    float m = g->m;
    return (s_in / (2.0f * m)) - (a2 / ((4.0f * m) * m));
}

// launches the S_in and sum a^2 reductions into zeroed counters (the sum of
// a^2 in the masked sums' launch); base as for r1_all
void modularity_launch(lv_graph *g, bool singletons = false,
                       const lv::DevView *base = nullptr) {
    r1_all(g, singletons ? 2 : 1, 0, base);
}

// Locality order for the pass (single-graph mode): position x of the new
// numbering holds input vertex order[x] of the class grouping, so the heavy
// and block classes, whose state most entries gather, are contiguous.
bool want_locality(const lv_graph *g, uint32_t n) {
    if (g->G || n < 2) return false;
    static const char *env = getenv("LV_LOCALITY");
    const int mode = env ? atoi(env) : g->tune.locality;
    if (mode == 0) return false;
    if (mode == 1) return true;
    return n >= g->tune.locality_min_vertices;
}

void relabel(lv_graph *g) {
    hipStream_t s = g->stream;
    const uint32_t n = g->n;
    ProfScope ps(g, LV_K_RELABEL, 20.0 * n + 16.0 * (double)g->ne);
    lv::launch_relabel_map(s, g->order, n, g->lid, g->pmap, g->row, g->rtmp);
    g->order_ident = true;  // relabel_map leaves order as the identity
    lv::launch_scan_u32(s, g->rtmp, 0, n, g->row2, g->scan_tmp32);
    const lv::RelabelView r{g->row, g->col, g->w, g->row2, g->col2, g->w2, g->lid, g->pmap};
    lv::launch_relabel_rows(s, r, g->cls_off[1], g->cls_off[5], 8);
    lv::launch_relabel_rows(s, r, g->cls_off[5], g->cls_off[lv::kClassHeavy], 64);
    lv::launch_relabel_heavy(s, r, g->hp_cvtx, g->heavy_coff, g->heavy_chunks,
                             g->cls_off[lv::kClassHeavy]);
    std::swap(g->row, g->row2);
    std::swap(g->col, g->col2);
    std::swap(g->w, g->w2);
    // the assignment refers to this pass's vertices: input ids -> positions
    lv::launch_gather_u32(s, g->pmap, g->init_comm, g->init_comm, g->n0);
    g->relabeled = true;
}

int pass_init(lv_graph *g, float *q0) {
    int rc = LV_OK;
    // a contraction that wrote the graph in its locality order classified it
    // already (again if the heavy threshold changed since; a graph in a
    // locality order keeps it)
    bool zeroed = false;  // classify_order left every counter but its own head fields zero
    if (!g->prelabeled || g->cls_heavy_min != g->tune.heavy_min_degree) {
        rc = classify(g);
        if (rc) return rc;
        zeroed = true;
        if (!g->relabeled && want_locality(g, g->n)) relabel(g);
    }
    g->prelabeled = false;
    if (!zeroed) zero_counters(g);
    // the k total into the spread slots of the move count (unused until the
    // first sweep), and the singletons' modularity sums (which do not use m)
    // behind it: one collapse and one round trip for both
    r1_all(g, 0, 1, nullptr, moved_ptr(g));
    modularity_launch(g, true);  // every community is a singleton here
    rc = sync_counters(g);
    if (rc) return rc;
    g->m = fix2f_host(g, (int64_t)g->hc->tot[5]) / 2.0f;
    if (!(g->m > 0.0f)) {
        // zero total weight: the reference divides by zero and never stops
        return LV_ERR_RANGE;
    }
    g->pass_ready = true;
    if (q0) *q0 = q_from_counters(g);
    return LV_OK;
}

// classes of buckets 1-5 with at least this many vertices take the narrow
// move kernels (LV_NARROW_MIN overrides the default; 0: always)
uint32_t narrow_min() {
    static const char *env = getenv("LV_NARROW_MIN");
    static const uint32_t n = env ? (uint32_t)strtoul(env, nullptr, 10) : lv::kNarrowMinVertices;
    return n;
}

// Whether the current pass sweeps with two copies of the community state
// (moves applied inside the step kernels, no apply launch per step): single
// graphs of at most 2^22 vertices (co-purchase graph 25.5 -> 22.5 ms per run,
// R-MAT scale 18-22 and SBMs of 0.2-1M vertices 0.6-5.5 % faster).  Larger
// graphs and batches keep one copy and k_apply: a step's gathers then
// alternate between two copies of state too large for the caches (R-MAT-24
// +0.2 %, 4096-graph batch +1.3 % with two copies).  LV_FUSED=0/1 forces one
// way (tests).
bool fused_pass(const lv_graph *g) {
    static const char *env = getenv("LV_FUSED");
    if (env && *env) return atoi(env) != 0;
    return !g->G && g->n <= (1u << 22);
}

// profiling family of each move class
constexpr int kClassFamily[kNumClasses] = {
    0,               LV_K_MOVE_SMALL, LV_K_MOVE_SMALL, LV_K_MOVE_SMALL, LV_K_MOVE_SMALL,
    LV_K_MOVE_SMALL, LV_K_MOVE_WAVE,  LV_K_MOVE_BLOCK, LV_K_MOVE_BLOCK, LV_K_MOVE_BLOCK,
    LV_K_MOVE_BLOCK, LV_K_MOVE_BLOCK, LV_K_MOVE_GLOBAL};

// vertices of bucket step b in `order`: class b, or classes 7..heavy for b = 7
void step_range(const lv_graph *g, int b, uint32_t *a, uint32_t *e) {
    *a = g->cls_off[b];
    *e = g->cls_off[(b == 7 ? lv::kClassHeavy : b) + 1];
}

// The view of a sweep's j-th non-empty bucket step: it reads copy j % 2 of
// the community state (copy 0: comm, comm_sz, A) and applies its moves to
// the other copy.
lv::DevView step_view(lv_graph *g, uint32_t j) {
    lv::DevView v = view(g);
    v.comm2 = g->comm2;
    v.comm_sz2 = g->comm_sz2;
    v.A2 = g->A2;
    if (j & 1u) {
        std::swap(v.comm, v.comm2);
        std::swap(v.comm_sz, v.comm_sz2);
        std::swap(v.A, v.A2);
    }
    v.cu_moved = moved_ptr(g);
    return v;
}

// One bucket step, the sweep's j-th with vertices: every vertex decides from
// the snapshot (copy j % 2) and its move goes to the other copy at once.
// That copy still lacks the previous step's moves (vertices [pa, pe) of
// `order`, none for the sweep's first step): workgroups in front of the
// step's first class kernel apply them.
// front: the previous sweep's stop rule on the device, carried by the
// step's first kernel (null: none).
int step_fused(lv_graph *g, int b, uint32_t j, uint32_t pa, uint32_t pe,
               const lv::SweepTail *front = nullptr) {
    const lv::DevView v = step_view(g, j);
    const int c0 = b, c1 = (b == 7) ? lv::kClassHeavy : b;
    lv::DevView cu = v;  // carries the catch-up to the first class kernel
    cu.cu_verts = g->order + pa;
    cu.cu_count = pe - pa;
    bool caught = pe == pa;
    bool block_cls = false;
    for (int c = c0; c < lv::kClassHeavy && c <= c1; ++c) block_cls |= g->cls_count[c] != 0;
    // only the heavy class has vertices: its chain catches up on its own,
    // and the stop rule gets a launch of its own
    if (!block_cls && front) {
        lv::launch_collapse_decide(g->stream, g->hc->tail, g->S);
        front = nullptr;
    }
    if (!caught && !block_cls) {
        ProfScope ps(g, LV_K_APPLY, 16.0 * (pe - pa));
        lv::launch_catchup(g->stream, cu);
        caught = true;
    }
    // the heavy rows' chain runs on the second stream beside the block
    // classes of the same step (they only read the step-start snapshot, and
    // their moves and the catch-up's are order-free atomic updates of the
    // other copy)
    const bool fork = c1 == lv::kClassHeavy && g->cls_count[lv::kClassHeavy] && block_cls &&
                      use_fork(g);
    if (fork) {
        (void)hipEventRecord(g->ev_fork, g->stream);
        (void)hipStreamWaitEvent(g->stream2, g->ev_fork, 0);
    }
    for (int c = c0; c <= c1; ++c) {
        if (!g->cls_count[c]) continue;
        hipStream_t st = (fork && c == lv::kClassHeavy) ? g->stream2 : g->stream;
        const bool first = !caught && c != lv::kClassHeavy;
        const double cub = first ? 16.0 * (pe - pa) : 0.0;
        ProfScope ps(g, kClassFamily[c],
                     12.0 * (double)g->cls_deg[c] + 28.0 * g->cls_count[c] + cub, st);
        if (c == lv::kClassHeavy)
            lv::launch_heavy_move(st, v, heavy_view(g));
        else {
            lv::DevView vc = first ? cu : v;
            vc.vbase = g->order_ident ? (int32_t)g->cls_off[c] : -1;
            vc.tail = front;
            front = nullptr;
            lv::launch_move_class(st, vc, c, g->order + g->cls_off[c], g->cls_count[c],
                                  lv::MoveShape{g->tune.wave_vertices, g->tune.block_threads,
                                                g->tune.class8_slots, narrow_min()});
        }
        caught |= first;
    }
    if (fork) {
        (void)hipEventRecord(g->ev_join, g->stream2);
        (void)hipStreamWaitEvent(g->stream, g->ev_join, 0);
    }
    return LV_OK;
}

// One bucket step on one copy of the state: the step's kernels write their
// decisions to new_comm, then k_apply applies them in place.
int step_apply(lv_graph *g, int b, const lv::SweepTail *front = nullptr) {
    const lv::DevView v = view(g);
    const int c0 = b, c1 = (b == 7) ? lv::kClassHeavy : b;
    // the heavy rows' chain runs on the second stream beside the block
    // classes of the same step (they only read the step-start snapshot and
    // write new_comm of their own vertices)
    bool fork = false;
    if (c1 == lv::kClassHeavy && g->cls_count[lv::kClassHeavy] && use_fork(g)) {
        for (int c = c0; c < lv::kClassHeavy; ++c) fork |= g->cls_count[c] != 0;
    }
    if (fork) {
        (void)hipEventRecord(g->ev_fork, g->stream);
        (void)hipStreamWaitEvent(g->stream2, g->ev_fork, 0);
    }
    if (front) {  // only the heavy class has vertices: the stop rule on its own
        bool block_cls = false;
        for (int c = c0; c < lv::kClassHeavy && c <= c1; ++c) block_cls |= g->cls_count[c] != 0;
        if (!block_cls) {
            lv::launch_collapse_decide(g->stream, g->hc->tail, g->S);
            front = nullptr;
        }
    }
    for (int c = c0; c <= c1; ++c) {
        if (!g->cls_count[c]) continue;
        hipStream_t st = (fork && c == lv::kClassHeavy) ? g->stream2 : g->stream;
        ProfScope ps(g, kClassFamily[c], 12.0 * (double)g->cls_deg[c] + 28.0 * g->cls_count[c], st);
        if (c == lv::kClassHeavy) {
            lv::launch_heavy_move(st, v, heavy_view(g));
        } else {
            lv::DevView vc = v;
            vc.vbase = g->order_ident ? (int32_t)g->cls_off[c] : -1;
            vc.tail = front;
            front = nullptr;
            lv::launch_move_class(st, vc, c, g->order + g->cls_off[c], g->cls_count[c],
                                  lv::MoveShape{g->tune.wave_vertices, g->tune.block_threads,
                                                g->tune.class8_slots, narrow_min()});
        }
    }
    if (fork) {
        (void)hipEventRecord(g->ev_join, g->stream2);
        (void)hipStreamWaitEvent(g->stream, g->ev_join, 0);
    }
    const uint32_t a = g->cls_off[c0], e = g->cls_off[c1 + 1];
    if (e > a) {
        ProfScope ps(g, LV_K_APPLY, 16.0 * (e - a));
        lv::DevView va = v;
        va.vbase = g->order_ident ? (int32_t)a : -1;
        lv::launch_apply(g->stream, va, g->order + a, e - a, moved_ptr(g));
    }
    return LV_OK;
}

// A sweep's bucket steps 1-7 (those with vertices).  Two copies: each step
// reads the copy its predecessor wrote; the copy the last step wrote holds
// the sweep's result, and the last step's moves still have to be applied to
// the copy it read.  Returns the view of the sweep's result with that
// catch-up attached, for the masked sums' launch (modularity_launch): after
// it both copies hold the sweep's result, as they held its start.  One copy:
// view(g).
// front: the previous sweep's stop rule, carried by the first step with
// vertices (null: none).
lv::DevView sweep_steps(lv_graph *g, const lv::SweepTail *front = nullptr) {
    if (!fused_pass(g)) {
        for (int b = 1; b <= 7; ++b) {
            uint32_t a, e;
            step_range(g, b, &a, &e);
            step_apply(g, b, e > a ? front : nullptr);
            if (e > a) front = nullptr;
        }
        return view(g);
    }
    uint32_t j = 0, pa = 0, pe = 0;
    for (int b = 1; b <= 7; ++b) {
        uint32_t a, e;
        step_range(g, b, &a, &e);
        if (e == a) continue;
        step_fused(g, b, j, pa, pe, j == 0 ? front : nullptr);
        pa = a;
        pe = e;
        ++j;
    }
    lv::DevView v = step_view(g, j);  // reads the copy the last step wrote
    v.cu_verts = g->order + pa;
    v.cu_count = pe - pa;
    return v;
}

// the device work of one sweep, ending with the counters' D2H; no host sync
// (the host-decided loop: LV_HOST_LOOP=1, carved handles, lv_sweep)
void sweep_enqueue(lv_graph *g) {
    zero_counters(g);
    const lv::DevView v = sweep_steps(g);
    modularity_launch(g, false, &v);
    counters_d2h(g);
}

int sweep_finish(lv_graph *g, float *q, uint64_t *moved) {
    LV_HIP(hipStreamSynchronize(g->stream));
    if (g->hc->err) return LV_ERR_OVERFLOW;
    if (q) *q = q_from_counters(g);
    if (moved) *moved = g->hc->tot[5];
    return LV_OK;
}

int sweep(lv_graph *g, float *q, uint64_t *moved) {
    sweep_enqueue(g);
    return sweep_finish(g, q, moved);
}

// One pass's sweep as a HIP graph: the launch sequence is fixed within a
// pass (same classes, same m), so it is captured once and replayed per sweep.
struct SweepGraph {
    lv_graph *g = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<ProfRec> recs;  // profiling event pairs recorded inside the graph
    ~SweepGraph() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        if (g)
            for (auto &r : recs) {
                g->ev_pool.push_back(r.a);
                g->ev_pool.push_back(r.b);
            }
    }
    // accumulate the kernel times of the replay that just finished
    void collect() {
        for (auto &r : recs) {
            float ms = 0.0f;
            if (hipEventElapsedTime(&ms, r.a, r.b) == hipSuccess) {
                g->prof.ms[r.fam] += ms;
                g->prof.launches[r.fam] += 1;
                g->prof.bytes[r.fam] += r.bytes;
            } else {
                (void)hipGetLastError();
            }
        }
    }
};

// Whether a pass's sweeps are captured once into a HIP graph and replayed
// (LV_GRAPH=1) instead of launched kernel by kernel.  Measured on MI355X
// with ROCm 7, replaying the captured sweep is slower than launching its
// kernels from the host, by 0.6 % on R-MAT-24, 1.1 % on the 4096-graph
// batch and 2-4 % on the FEM band and co-purchase graphs (every result
// identical), so kernels are launched directly by default.  A carved handle
// never captures: instantiating a HIP graph allocates.
bool graphs_enabled(const lv_graph *g) {
    static const bool on = [] {
        const char *e = getenv("LV_GRAPH");
        return e && atoi(e) != 0;
    }();
    return on && !g->carve;
}

// Whether the pass loop's stop rule runs on the device (k_collapse_decide,
// read by the host from mapped memory) rather than on the host after a copy
// of the counters (LV_HOST_LOOP=1, and a carved handle, which has no mapped
// host memory).
bool device_loop(const lv_graph *g) {
    static const bool off = [] {
        const char *e = getenv("LV_HOST_LOOP");
        return e && atoi(e) != 0;
    }();
    return !off && !g->carve;
}

// Captures what `enqueue` launches on the graph's stream as one HIP graph;
// on failure the caller falls back to eager launches.
template <typename F>
bool capture(lv_graph *g, SweepGraph &sg, F enqueue) {
    sg.g = g;
    if (hipStreamBeginCapture(g->stream, hipStreamCaptureModeThreadLocal) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    g->capture_recs = &sg.recs;
    enqueue();
    g->capture_recs = nullptr;
    if (hipStreamEndCapture(g->stream, &sg.graph) != hipSuccess) {
        (void)hipGetLastError();
        sg.graph = nullptr;
        return false;
    }
    if (hipGraphInstantiate(&sg.exec, sg.graph, nullptr, nullptr, 0) != hipSuccess) {
        (void)hipGetLastError();
        sg.exec = nullptr;
        return false;
    }
    return true;
}

// Whether the host keeps the next group of sweeps queued behind the one it
// waits for.  The device then never idles through the host's wake-up, read
// of the loop state and next launches (about 25-30 us per group); the group
// queued after the stop returns at entry in every kernel (halt flag).  That
// flag costs every kernel a dependent load, so only pass graphs of at most
// 2^24 entries, whose sweeps take well under a millisecond, are pipelined
// (FEM band -2 %, co-purchase -4 %; R-MAT-24 ran 0.4 % slower).
// LV_PIPELINE=0/1 forces one way.  A profiled run waits for each group: the
// event pairs of a captured sweep are re-recorded by every replay.
bool pipelined(const lv_graph *g) {
    if (g->prof_on) return false;
    static const char *env = getenv("LV_PIPELINE");
    if (env && *env) return atoi(env) != 0;
    return g->ne <= (1ull << 24);
}

// Sweeps per wait of the host in a pass loop decided on the device (the
// tuning's sweeps_per_launch, default 1; LV_SWEEPS_PER_LAUNCH overrides it).
// With more than one, the kernels of the sweeps queued after the stop must
// return at entry, which costs each kernel of every sweep a dependent load
// of the halt flag; measured on MI355X (with captured sweeps) that load
// cost about what the host round trip it saves does, so one sweep per wait
// is the default and the autotuner picks more per graph.  A profiled run
// always takes one, so every timed sweep did its work.
uint32_t sweeps_per_launch(const lv_graph *g) {
    if (g->prof_on) return 1;
    static const char *env = getenv("LV_SWEEPS_PER_LAUNCH");
    if (env && atoi(env) > 0) return (uint32_t)std::min(atoi(env), 16);
    return g->tune.sweeps_per_launch;
}

// The first bucket step of a sweep with vertices (0: none).
int lead_bucket(const lv_graph *g) {
    for (int b = 1; b <= 7; ++b) {
        uint32_t a, e;
        step_range(g, b, &a, &e);
        if (e > a) return b;
    }
    return 0;
}

// the stop rule's arguments for the pass
lv::SweepTail sweep_tail(lv_graph *g, float min_gain) {
    lv::SweepTail t;
    t.spread = g->dc->acc;
    t.tot = g->dc->tot;
    t.err = &g->dc->err;
    t.m = g->m;
    t.min_gain = min_gain;
    t.max_iters = kMaxIters;
    t.ps = g->pst;
    t.ps_host = g->pst_hd;
    return t;
}

// One sweep whose end applies the stop rule on the device and writes the
// loop state to the host's mapped copy; the counters must be zero at its
// start (the collapse leaves them so for the next sweep).  Folded: the stop
// rule of this sweep runs in workgroup 0 of the next sweep's first kernel
// (g->tail_d, staged in g->hc->tail), and this sweep's first kernel carries
// the previous sweep's unless it is the pass's first.
void sweep_enqueue_decide(lv_graph *g, float min_gain, bool fold = false, bool first = true) {
    const lv::DevView v = sweep_steps(g, fold && !first ? g->tail_d : nullptr);
    modularity_launch(g, false, &v);
    if (!fold) lv::launch_collapse_decide(g->stream, sweep_tail(g, min_gain), g->S);
}

int modularity_optimization(lv_graph *g, float min_gain, int *improved, float *q_print,
                            uint32_t *iters) {
    float act = 0.0f, old = 0.0f;
    int rc = pass_init(g, &act);
    if (rc) return rc;
    g->last_q_before = act;
    bool first = true, ret = true;
    uint32_t it = 0;
    // The loop of modularity.cu:464-508 with its stop rule evaluated on the
    // device after every sweep: the host enqueues a group of sweeps (or
    // replays it as a captured graph) and reads the loop state once per group.
    SweepGraph sg;
    if (device_loop(g)) {
        const uint32_t K = sweeps_per_launch(g);
        const bool pipe = pipelined(g);
        // kernels test the halt flag only when sweeps can follow the stop
        const bool halt = K > 1 || pipe;
        // pipelined direct launches fold each sweep's stop rule into the next
        // sweep's first kernel: the kernels already test the halt flag, and
        // the next sweep is queued anyway
        const bool fold = pipe && !graphs_enabled(g) && lead_bucket(g) != 0;
        *g->pst_h = lv::PassState{act, 0.0f, 0.0f, 0u, 1u, 1u, 0u, 0u};
        LV_HIP(hipMemcpyAsync(g->pst, g->pst_h, sizeof(lv::PassState), hipMemcpyHostToDevice,
                              g->stream));
        if (fold) {
            g->hc->tail = sweep_tail(g, min_gain);
            LV_HIP(hipMemcpyAsync(g->tail_d, &g->hc->tail, sizeof(lv::SweepTail),
                                  hipMemcpyHostToDevice, g->stream));
        }
        zero_counters(g);
        uint32_t swept = 0;
        auto enqueue = [&] {
            g->halt_on = halt;
            for (uint32_t k = 0; k < K; ++k) sweep_enqueue_decide(g, min_gain, fold, swept++ == 0);
            g->halt_on = false;
        };
        const bool use_graph = graphs_enabled(g) && capture(g, sg, enqueue);
        auto launch = [&]() -> hipError_t {
            if (use_graph) return hipGraphLaunch(sg.exec, g->stream);
            enqueue();
            return hipSuccess;
        };
        if (pipe) {
            LV_HIP(launch());
            LV_HIP(hipEventRecord(g->ev_launch[0], g->stream));
        }
        for (uint32_t i = 0;; ++i) {
            LV_HIP(launch());
            if (pipe) {
                // launch i + 1 is queued: wait for launch i only
                LV_HIP(hipEventRecord(g->ev_launch[(i + 1) & 1u], g->stream));
                LV_HIP(hipEventSynchronize(g->ev_launch[i & 1u]));
            } else {
                LV_HIP(hipStreamSynchronize(g->stream));
            }
            if (use_graph && g->prof_on) sg.collect();
            lv::PassState p = *g->pst_h;
            if (!p.done && !p.err) continue;
            // stopped: the queued launch returns at entry and leaves the
            // loop state as it is; read it once nothing runs
            if (pipe) {
                LV_HIP(hipStreamSynchronize(g->stream));
                p = *g->pst_h;
            }
            if (p.err) return LV_ERR_OVERFLOW;
            act = p.act;
            it = p.it;
            ret = p.ret != 0u;
            if (!ret && q_print) *q_print = p.q_print;
            break;
        }
        // folded, two copies: the first step of the sweep after the stop moved
        // its vertices in the second copy; put them back as the first copy
        // has them, so both copies hold the pass result again (one copy: that
        // step's apply returned at entry)
        if (fold && fused_pass(g)) {
            uint32_t a, e;
            step_range(g, lead_bucket(g), &a, &e);
            lv::DevView cv = step_view(g, 0);
            cv.cu_verts = g->order + a;
            cv.cu_count = e - a;
            lv::launch_catchup(g->stream, cv);
        }
    } else {  // the stop rule on the host after each sweep
        for (;;) {
            uint64_t moved = 0;
            old = act;
            rc = sweep(g, &act, &moved);
            if (rc) return rc;
            ++it;
            // the loop condition of modularity.cu:508; a sweep without moves
            // is a fixed point and also ends the loop
            bool stop = !(act - old >= min_gain) || moved == 0 || it >= kMaxIters;
            if (first && (act - old < min_gain || moved == 0)) {
                if (q_print) *q_print = old;
                ret = false;
            }
            first = false;
            if (stop) break;
        }
    }
    g->last_iters = it;
    g->last_q_after = act;
    if (improved) *improved = ret ? 1 : 0;
    if (iters) *iters = it;
    return LV_OK;
}

lv::AggView aview(lv_graph *g) {
    lv::AggView a;
    a.kgid = g->G ? g->gid : nullptr;
    a.kbase = g->G ? g->vgather : nullptr;
    a.new_id = g->new_id;
    a.vnid = g->vnid;
    a.comm_deg = g->comm_deg;
    a.epos = g->epos;
    a.mstart = g->mstart;
    a.mcursor = g->mcursor;
    a.members = g->members;
    a.cnt = g->cnt;
    a.tkey = g->tkey;
    a.tval = g->tval;
    a.skey = g->skey;
    a.sval = g->sval;
    a.gtab = g->gtab;
    a.goff = g->agg_off;
    a.mdeg = g->agg_mdeg;
    a.moff = g->agg_moff;
    a.cnk = g->agg_cnk;
    a.coff = g->agg_coff;
    a.dcnt = g->agg_dcnt;
    a.nchunks = g->agg_nchunks;
    a.flist = a.clist = nullptr;
    a.nfl = a.ncl = 0;
    a.nn = g->agg_nn;
    a.err = &g->dc->err;
    return a;
}

// Merge of the large communities (total member degree > 2048): partitioned
// by key range into LDS-sized buckets; a community with a bucket that
// overflows its LDS table is merged again on global tables (LV_AGG_GLOBAL=1
// sends every large community there, a test hook).
int aggregate_large(lv_graph *g, const lv::DevView &v, const uint32_t *lc, uint32_t nl,
                    uint32_t nn, int passes, lv::AggPart *out) {
    hipStream_t s = g->stream;
    lv::launch_agg_large_prep(s, v, aview(g), lc, nl);  // chunks per community
    lv::launch_scan_u32(s, g->agg_cnk, 0, nl, g->agg_coff, g->scan_tmp32);
    lv::launch_aggp_nb(s, aview(g), lc, nl, nn, g->tune.agg_bucket_entries, g->ap_nbk);
    lv::launch_scan_u32(s, g->ap_nbk, 0, nl, g->ap_bstart, g->scan_tmp32);
    uint32_t chunks = 0, NB = 0;
    LV_HIP(hipMemcpyAsync(&chunks, g->agg_coff + nl, 4, hipMemcpyDeviceToHost, s));
    LV_HIP(hipMemcpyAsync(&NB, g->ap_bstart + nl, 4, hipMemcpyDeviceToHost, s));
    LV_HIP(hipStreamSynchronize(s));
    g->agg_nchunks = chunks;
    if (NB > g->ap_bcap) {
        gfree(g, g->ap_bcnt);
        gfree(g, g->ap_bpos);
        gfree(g, g->ap_bcur);
        gfree(g, g->ap_bnd);
        gfree(g, g->ap_boff);
        gfree(g, g->ap_bowner);
        gfree(g, g->ap_btmp);
        int e = galloc(g, &g->ap_bcnt, NB);
        if (!e) e = galloc(g, &g->ap_bpos, (size_t)NB + 1);
        if (!e) e = galloc(g, &g->ap_bcur, NB);
        if (!e) e = galloc(g, &g->ap_bnd, NB);
        if (!e) e = galloc(g, &g->ap_boff, (size_t)NB + 1);
        if (!e) e = galloc(g, &g->ap_bowner, NB);
        if (!e) e = galloc(g, &g->ap_btmp, lv::scan_tmp_elems(NB + 1) + 8);
        if (e) return e;
        g->ap_bcap = NB;
    }
    lv::AggView av = aview(g);
    lv::AggPart &p = *out;
    p.nbk = g->ap_nbk;
    p.bstart = g->ap_bstart;
    p.bowner = g->ap_bowner;
    p.bcnt = g->ap_bcnt;
    p.bpos = g->ap_bpos;
    p.bcur = g->ap_bcur;
    p.bnd = g->ap_bnd;
    p.boff = g->ap_boff;
    p.cflag = g->ap_cflag;
    p.nflag = g->ap_nflag;
    // staging: heavy-move scratch (idle during aggregation), then the next
    // CSR's column and weight arrays (written only after the merges)
    p.fkey = g->dlist;
    p.fw = reinterpret_cast<float *>(g->eslot);
    p.pkey = g->col2;
    p.pw = g->w2;
    p.nn = nn;
    p.NB = NB;
    // LV_AGG_DMAX (test hook) lowers the bucket capacity so some communities
    // overflow and take the global-table merge next to partitioned ones
    static const char *dmax_env = getenv("LV_AGG_DMAX");
    p.dmax = lv::kAggBucketDistinct;
    if (dmax_env && atoi(dmax_env) > 0)
        p.dmax = std::min((int)lv::kAggBucketDistinct, atoi(dmax_env));
    static const bool force_global = getenv("LV_AGG_GLOBAL") != nullptr;
    uint32_t nflag = 0;
    if (force_global) {
        lv::launch_fill_u32(s, g->ap_cflag, 1u, nl);
        p.NB = 0;  // no bucket output
        nflag = nl;
    } else {
        lv::launch_owner_fill(s, g->ap_bstart, nl, g->ap_bowner);
        lv::launch_fill_u32(s, g->ap_bcnt, 0u, NB);
        lv::launch_fill_u32(s, g->ap_cflag, 0u, nl);
        lv::launch_fill_u32(s, g->ap_nflag, 0u, 1);
        lv::launch_aggp(s, v, av, p, lc, nl, 0);
        lv::launch_scan_u32(s, g->ap_bcnt, 0, NB, g->ap_bpos, g->ap_btmp);
        LV_HIP(hipMemcpyAsync(g->ap_bcur, g->ap_bpos, (size_t)NB * 4, hipMemcpyDeviceToDevice, s));
        lv::launch_aggp(s, v, av, p, lc, nl, 1);
        lv::launch_aggp(s, v, av, p, lc, nl, 2);
        lv::launch_scan_u32(s, g->ap_bnd, 0, NB, g->ap_boff, g->ap_btmp);
        LV_HIP(hipMemcpyAsync(&nflag, g->ap_nflag, 4, hipMemcpyDeviceToHost, s));
        LV_HIP(hipStreamSynchronize(s));
        if (nflag) lv::launch_aggp(s, v, av, p, lc, nl, 3);
    }
    if (!nflag) return LV_OK;
    // global tables for the flagged communities only
    lv::launch_aggp_caps(s, av, p, lc, nl, g->agg_caps);
    lv::launch_scan_u64(s, g->agg_caps, nl, g->agg_off, g->scan_tmp64);
    lv::launch_fill_u32(s, g->agg_dcnt, 0u, nl);
    ull total = 0;
    LV_HIP(hipMemcpyAsync(&total, g->agg_off + nl, sizeof(ull), hipMemcpyDeviceToHost, s));
    LV_HIP(hipStreamSynchronize(s));
    if (int e = ensure_gtab(g, (size_t)total)) return e;
    // the global-table kernels take one workgroup per flagged community and
    // per chunk of one, from lists built on the device
    const size_t lcap = (size_t)g->agg_nchunks + nl + 2;
    if (lcap > g->agg_flcap) {
        gfree(g, g->agg_flist);
        if (int e = galloc(g, &g->agg_flist, lcap)) return e;
        g->agg_flcap = lcap;
    }
    uint32_t *flist = g->agg_flist, *clist = g->agg_flist + g->agg_nchunks + 1;
    lv::launch_fill_u32(s, flist, 0u, 1);
    lv::launch_fill_u32(s, clist, 0u, 1);
    lv::launch_agg_flag_list(s, g->ap_cflag, g->agg_coff, nl, flist, clist);
    uint32_t nfl = 0, ncl = 0;
    LV_HIP(hipMemcpyAsync(&nfl, flist, 4, hipMemcpyDeviceToHost, s));
    LV_HIP(hipMemcpyAsync(&ncl, clist, 4, hipMemcpyDeviceToHost, s));
    LV_HIP(hipStreamSynchronize(s));
    av = aview(g);
    av.flist = flist;
    av.clist = clist;
    av.nfl = nfl;
    av.ncl = ncl;
    lv::launch_agg_merge(s, v, av, lv::kAggGlobalClass, lc, nl, passes);
    return LV_OK;
}

// aggregation (aggregation.cu:387-462) + composition of init_comm
// (modularity.cu:515-517, aggregation.cu:459)
int aggregation(lv_graph *g) {
    if (!g->pass_ready) return LV_ERR_STATE;
    hipStream_t s = g->stream;
    lv::DevView v = view(g);
    const uint32_t n = g->n;
    ProfScope ps(g, LV_K_AGGREGATE, 12.0 * g->ne + 24.0 * n);
    // a tiny single graph in its input numbering: one workgroup does it all
    if (!g->G && !g->relabeled && !want_locality(g, n) && n <= lv::kTinyN &&
        g->ne <= lv::kTinyE) {
        lv::launch_agg_tiny(s, v, g->n0, g->init_comm, g->row2, g->col2, g->w2, g->dc->acursor,
                            g->new_id);
        int rc = sync_head(g);
        if (rc) return rc;
        const uint32_t nn = g->hc->acursor[0], ne2 = g->hc->acursor[1];
        if (nn > n || ne2 > g->ne) return LV_ERR_OVERFLOW;
        std::swap(g->row, g->row2);
        std::swap(g->col, g->col2);
        std::swap(g->w, g->w2);
        g->agg_nn = nn;
        g->n = nn;
        g->ne = ne2;
        g->pass_ready = false;
        g->relabeled = false;
        g->prelabeled = false;
        return LV_OK;
    }
    // a single graph that the pass left with few communities: a dense matrix
    // of exact sums in three launches (the matrix in the sums scratch);
    // more communities than the matrix holds fall through to the general path
    if (!g->G && !g->relabeled && !want_locality(g, n) && n <= lv::kFewN &&
        g->e_cap >= (uint64_t)lv::kFewNN * lv::kFewNN * 65 / 64) {
        lv::launch_agg_few(s, v, g->new_id, g->tval, g->n0, g->init_comm, g->row2, g->col2,
                           g->w2, g->dc->acursor);
        int rc = sync_head(g);
        if (rc) return rc;
        const uint32_t nn = g->hc->acursor[0];
        if (nn <= lv::kFewNN) {
            const uint32_t ne2 = g->hc->acursor[1];
            if (nn > n || ne2 > g->ne) return LV_ERR_OVERFLOW;
            std::swap(g->row, g->row2);
            std::swap(g->col, g->col2);
            std::swap(g->w, g->w2);
            g->agg_nn = nn;
            g->n = nn;
            g->ne = ne2;
            g->pass_ready = false;
            g->relabeled = false;
            g->prelabeled = false;
            return LV_OK;
        }
    }
    if (g->G) {  // a batch also clears its per-graph counters
        zero_counters(g);
        lv::launch_fill_u32(s, g->comm_deg, 0u, n);
        lv::launch_fill_u32(s, g->mcursor, 0u, n);
    } else {
        lv::launch_zero3_u32(s, reinterpret_cast<uint32_t *>(g->dc), sizeof(Counters) / 4,
                             g->comm_deg, n, g->mcursor, n);
    }
    lv::launch_agg_prep(s, v, aview(g));
    if (g->relabeled)  // rank in the input numbering
        lv::launch_relabel_new_id(s, g->comm_sz, g->lid, n, g->rtmp, g->rrank, g->new_id,
                                  g->scan_tmp32);
    else
        lv::launch_scan_u32(s, g->comm_sz, 1, n, g->new_id, g->scan_tmp32);
    lv::launch_agg_vnid(s, g->comm, g->new_id, n, g->vnid);
    // a batch: first new id of every graph (the key range of its communities)
    if (g->G) lv::launch_gather_u32(s, g->new_id, g->voff_d, g->vgather, g->G + 1);
    lv::launch_scan_u32(s, g->comm_sz, 0, n, g->mstart, g->scan_tmp32);
    lv::launch_scan_u32(s, g->comm_deg, 0, n, g->epos, g->scan_tmp32);
    lv::launch_agg_members(s, v, aview(g));
    lv::launch_agg_classify_count(s, v, aview(g), g->dc->acounts);
    uint32_t nn = 0;
    LV_HIP(hipMemcpyAsync(&nn, g->new_id + n, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    int rc = sync_head(g);
    if (rc) return rc;
    uint32_t aoff[kNumAggClasses + 1];
    aoff[0] = 0;
    for (int c = 0; c < kNumAggClasses; ++c) aoff[c + 1] = aoff[c] + g->hc->acounts[c];
    if (aoff[kNumAggClasses] != nn) return LV_ERR_OVERFLOW;
    g->agg_nn = nn;
    // staged in the host counters, as classify_order's cursors
    for (int c = 0; c < kNumAggClasses; ++c) g->hc->acursor[c] = aoff[c];
    LV_HIP(hipMemcpyAsync(g->dc->acursor, g->hc->acursor, sizeof(g->hc->acursor),
                          hipMemcpyHostToDevice, s));
    lv::launch_agg_classify_scatter(s, v, aview(g), g->dc->acursor, g->aorder);
    lv::launch_fill_u32(s, g->cnt, 0u, nn);  // edgeless communities keep 0
    int passes = 1;  // radix digits of a new id
    {
        uint32_t bits = 0;
        while (bits < 32 && (n > 1) && ((uint64_t)(n - 1) >> bits)) ++bits;
        passes = std::max(1, (int)((bits + 7) / 8));
    }
    // member entry offsets in member order (merges walk flat member entries)
    lv::launch_agg_mdeg(s, v, aview(g));
    lv::launch_scan_u32(s, g->agg_mdeg, 0, n, g->agg_moff, g->scan_tmp32);
    lv::AggView av = aview(g);
    // large communities: partitioned merge, global tables for overflowing ones;
    // the mid-size classes merge on the second stream meanwhile (disjoint
    // output ranges, no shared scratch)
    const uint32_t nl = g->hc->acounts[lv::kAggGlobalClass];
    const uint32_t *lc = g->aorder + aoff[lv::kAggGlobalClass];
    const bool fork = use_fork(g) && nl && aoff[lv::kAggGlobalClass] > aoff[1];
    hipStream_t sm = fork ? g->stream2 : s;
    if (fork) {
        LV_HIP(hipEventRecord(g->ev_fork, s));
        LV_HIP(hipStreamWaitEvent(g->stream2, g->ev_fork, 0));
    }
    for (int c = 0; c < lv::kAggGlobalClass; ++c)
        lv::launch_agg_merge(sm, v, av, c, g->aorder + aoff[c], aoff[c + 1] - aoff[c], passes);
    lv::AggPart ap = {};
    g->agg_nchunks = 0;
    int rc2 = nl ? aggregate_large(g, v, lc, nl, nn, passes, &ap) : LV_OK;
    if (fork) {  // join, also when the large merge failed
        LV_HIP(hipEventRecord(g->ev_join, g->stream2));
        LV_HIP(hipStreamWaitEvent(s, g->ev_join, 0));
    }
    if (rc2) return rc2;
    if (nl) av = aview(g);
    // new row pointers from distinct counts
    lv::launch_scan_u32(s, g->cnt, 0, nn, g->row2, g->scan_tmp32);
    uint32_t ne2 = 0;
    LV_HIP(hipMemcpyAsync(&ne2, g->row2 + nn, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    lv::EmitTarget t{g->row2, g->row2, nullptr, g->col2, g->w2};
    // The next pass in a locality order: its classes and order come from the
    // new degrees before anything is written, and the rows go straight to
    // their positions (columns mapped), instead of a copy at pass init.  The
    // old row pointers are free once the merges are done.
    const bool pre = want_locality(g, nn);
    uint32_t *const old_row = g->row;
    const uint32_t old_n = g->n;
    const uint64_t old_ne = g->ne;
    if (pre) {
        rc = sync_head(g);  // the merges' overflow flag, before classify clears it
        if (rc) return rc;
        if (g->hc->err) return LV_ERR_OVERFLOW;
        g->row = g->row2;  // by new id: the next pass's input numbering
        g->n = nn;
        g->ne = ne2;
        rc = classify_order(g);
        if (rc) {
            g->row = old_row;
            g->n = old_n;
            g->ne = old_ne;
            return rc;
        }
        lv::launch_relabel_map(s, g->order, nn, g->lid, g->pmap, g->row2, g->rtmp);
        g->order_ident = true;
        lv::launch_scan_u32(s, g->rtmp, 0, nn, old_row, g->scan_tmp32);
        t.rows = old_row;
        t.map = g->pmap;
    }
    // communities without entries (class 0, first in aorder) have empty rows
    lv::launch_agg_emit(s, v, av, g->aorder + aoff[1], aoff[lv::kAggGlobalClass] - aoff[1],
                        nullptr, t);
    if (nl) {
        lv::launch_aggp_emit(s, v, av, ap, lc, t);
        lv::launch_agg_emit(s, v, av, lc, nl, g->ap_cflag, t);
    }
    lv::launch_agg_compose(s, g->comm, g->new_id, t.map, g->init_comm, g->n0);
    rc = sync_head(g);
    if (rc) return rc;
    if (g->hc->err) return LV_ERR_OVERFLOW;
    if (pre) {
        g->row2 = g->row;  // the rows by new id become scratch
        g->row = old_row;  // by position
        rc = classify_heavy(g);
        if (rc) return rc;
    } else {
        std::swap(g->row, g->row2);
    }
    std::swap(g->col, g->col2);
    std::swap(g->w, g->w2);
    g->n = nn;
    g->ne = ne2;
    g->pass_ready = false;
    // in the input numbering, or in the next pass's locality order (lid)
    g->relabeled = pre;
    g->prelabeled = pre;
    return LV_OK;
}

// ---------------------------------------------------------------------------
// batch mode: many independent graphs in one segmented CSR
//
// Every graph keeps the exact arithmetic and stop rules of a single run: its
// own m, its own per-sweep modularity and its own decision to continue, stop
// after an improving pass (aggregate) or finish.  All graphs sweep together;
// a graph that has stopped is frozen (its vertices are skipped by every
// kernel) until the union of the graphs that improved is contracted at once.
// The contraction of the union is the union of the contractions: renumbering
// by rank keeps each graph's vertices contiguous and its rows ordered.  The
// fixed-point scale is the largest over the graphs; exact values and their
// FP32 roundings do not depend on the scale.
// ---------------------------------------------------------------------------
int push_gstate(lv_graph *g) {
    LV_HIP(hipMemcpyAsync(g->gstate_d, g->gstate_h.data(), g->G * sizeof(uint32_t),
                          hipMemcpyHostToDevice, g->stream));
    LV_HIP(hipStreamSynchronize(g->stream));
    return LV_OK;
}

// Pass init of every active graph (modularity.cu:455-463), on the device:
// classes, k_v, each graph's m (a graph without weight finishes), Q0.
int batch_pass_init(lv_graph *g) {
    const uint32_t G = g->G;
    int rc = classify(g);
    if (rc) return rc;
    zero_counters(g);
    r1_all(g, 0, 1);
    lv::launch_batch_m(g->stream, G, g->gcnt, g->S, g->voff_d, g->row, g->gstate_d, g->gm_d,
                       g->gst_d, LV_ERR_RANGE);
    zero_counters(g);
    modularity_launch(g, true);  // every active community is a singleton here
    lv::launch_batch_init(g->stream, G, g->gcnt, g->S, g->gstate_d, g->gst_d);
    g->pass_ready = true;
    return LV_OK;
}

// listed vertices below which a batch pass keeps its classes as they are
// (test hook LV_COMPACT_MIN lowers it so small batches compact too)
uint32_t compact_min() {
    static const char *env = getenv("LV_COMPACT_MIN");
    return env ? (uint32_t)strtoul(env, nullptr, 10) : (1u << 16);
}

// One pass for every active graph; returns in *more whether any graph
// improved (and was contracted).  After every sweep a device kernel applies
// each active graph's stop rule (modularity.cu:503-508) and counts the
// graphs still active; the host reads that summary once per group of
// sweeps.
int batch_pass(lv_graph *g, float min_gain, bool *more) {
    *more = false;
    const uint32_t G = g->G;
    int rc = batch_pass_init(g);
    if (rc) return rc;
    lv::launch_fill_u32(g->stream, reinterpret_cast<uint32_t *>(g->bsum_d), 0u,
                        sizeof(lv::BatchSummary) / 4);
    zero_counters(g);
    const uint32_t K = sweeps_per_launch(g);
    // a sweep starts with zero counters and summary counts: the decide and
    // halt kernels of the one before leave them so (and the halt writes the
    // summary to the host's mapped copy)
    auto enqueue = [&] {
        for (uint32_t k = 0; k < K; ++k) {
            const lv::DevView v = sweep_steps(g);
            modularity_launch(g, false, &v);
            lv::launch_batch_decide(g->stream, G, g->gcnt, &g->dc->err, g->S, min_gain,
                                    kMaxIters, g->voff_d, g->gstate_d, g->gst_d, g->bsum_d);
            lv::launch_batch_halt(g->stream, g->bsum_d, g->bsum_hd);
        }
    };
    auto sg = std::make_unique<SweepGraph>();
    auto recapture = [&] {
        sg = std::make_unique<SweepGraph>();
        g->halt_on = K > 1;
        const bool ok = graphs_enabled(g) && capture(g, *sg, enqueue);
        g->halt_on = false;
        return ok;
    };
    bool use_graph = recapture();
    for (;;) {
        if (use_graph) {
            LV_HIP(hipGraphLaunch(sg->exec, g->stream));
        } else {
            g->halt_on = K > 1;
            enqueue();
            g->halt_on = false;
        }
        LV_HIP(hipStreamSynchronize(g->stream));
        if (use_graph && g->prof_on) sg->collect();
        const lv::BatchSummary bs = *g->bsum_h;
        if (bs.err) return LV_ERR_OVERFLOW;
        if (!bs.active) break;
        // once the active graphs hold at most half of the listed vertices,
        // the classes are rebuilt from the active graphs alone (and the sweeps
        // captured again): frozen graphs' vertices cost a dispatched and
        // idle workgroup share in every kernel of every sweep otherwise
        const uint32_t listed = g->cls_off[kNumClasses];
        if (listed >= compact_min() && 2 * bs.live <= listed) {
            sg.reset();
            rc = classify(g);
            if (rc) return rc;
            use_graph = recapture();
        }
    }
    if (!g->bsum_h->waiting) return LV_OK;
    // contract the graphs that improved; every other graph stays as it is
    lv::launch_batch_identity(g->stream, g->gid, g->gstate_d, kGsWait, g->comm, g->comm_sz, g->n);
    rc = aggregation(g);
    if (rc) return rc;
    // new vertex ranges: rank of each graph's first old vertex
    lv::launch_gather_u32(g->stream, g->new_id, g->voff_d, g->vgather, G + 1);
    LV_HIP(hipMemcpyAsync(g->voff.data(), g->vgather, (G + 1) * sizeof(uint32_t),
                          hipMemcpyDeviceToHost, g->stream));
    LV_HIP(hipStreamSynchronize(g->stream));
    LV_HIP(hipMemcpyAsync(g->voff_d, g->voff.data(), (G + 1) * sizeof(uint32_t),
                          hipMemcpyHostToDevice, g->stream));
    lv::launch_batch_gid(g->stream, g->voff_d, G, g->gid, g->n);
    lv::launch_batch_reactivate(g->stream, G, g->gstate_d, g->gst_d);
    *more = true;
    return LV_OK;
}

// Run f(begin, end) over [0, n) on up to 16 host threads (inputs of
// hundreds of millions of entries are checked before every upload).
template <typename F>
void host_parallel(uint64_t n, F f) {
    unsigned T = std::thread::hardware_concurrency();
    T = std::max(1u, std::min(T, 16u));
    if (n < (1u << 20)) T = 1;
    if (T == 1) {
        f(0, n, 0u);
        return;
    }
    std::vector<std::thread> th;
    for (unsigned t = 0; t < T; ++t)
        th.emplace_back([=] { f(n * t / T, n * (t + 1) / T, t); });
    for (auto &x : th) x.join();
}

// Fixed-point scale of weights w[a, b): *S = the smallest shift making every
// weight an integer (0 if all are zero); LV_ERR_RANGE for a negative,
// infinite or NaN weight or a shift above 62 (the oracle's rule).
int min_shift(const float *w, uint64_t a, uint64_t b, int *S_out) {
    int S = 0;
    for (uint64_t j = a; j < b; ++j) {
        const float x = w[j];
        if (!(x >= 0.0f) || std::isinf(x)) return LV_ERR_RANGE;
        if (x == 0.0f) continue;
        uint32_t bits;
        memcpy(&bits, &x, 4);
        const int ef = (int)((bits >> 23) & 0xff);
        uint32_t man = bits & 0x7fffffu;
        int e;
        if (ef == 0) {
            e = -149;
        } else {
            man |= 0x800000u;
            e = ef - 150;
        }
        e += __builtin_ctz(man);  // man != 0 for a non-zero weight
        if (-e > S) S = -e;
    }
    if (S > 62) return LV_ERR_RANGE;
    *S_out = S;
    return LV_OK;
}

// Whether sum(w[a, b) * 2^S) stays below 2^61 (every exact accumulator then
// fits an int64 with headroom).  A single weight at or above the limit fails
// before its integer conversion, which would not be representable.
bool total_fits(const float *w, uint64_t a, uint64_t b, int S) {
    const float lim = ldexpf(1.0f, 61);
    u128 tot = 0;
    for (uint64_t j = a; j < b; ++j) {
        if (w[j] == 0.0f) continue;
        const float x = ldexpf(w[j], S);
        if (!(x < lim)) return false;
        tot += (u128)(uint64_t)x;
        if (tot >= (u128)kFixedLimit) return false;
    }
    return true;
}

// exact host-side validation and fixed-point scale of one graph (same rule as
// the oracle): S = the smallest shift making every weight an integer; the
// total at that scale must stay below 2^61
int scale_bits(uint64_t ne, const float *w, int *S_out) {
    std::vector<int> sp(16, 0), bad(16, 0);
    host_parallel(ne, [&](uint64_t a, uint64_t b, unsigned t) {
        bad[t] = min_shift(w, a, b, &sp[t]);
    });
    int S = 0;
    for (int t = 0; t < 16; ++t) {
        if (bad[t]) return LV_ERR_RANGE;
        S = std::max(S, sp[t]);
    }
    std::vector<char> fits(16, 1);
    std::vector<u128> part(16, 0);
    host_parallel(ne, [&](uint64_t a, uint64_t b, unsigned t) {
        const float lim = ldexpf(1.0f, 61);
        u128 tot = 0;
        for (uint64_t j = a; j < b && fits[t]; ++j) {
            if (w[j] == 0.0f) continue;
            const float x = ldexpf(w[j], S);
            if (!(x < lim)) fits[t] = 0;
            else tot += (u128)(uint64_t)x;
        }
        part[t] = tot;
    });
    u128 tot = 0;
    for (int t = 0; t < 16; ++t) {
        if (!fits[t]) return LV_ERR_RANGE;
        tot += part[t];
    }
    if (tot >= (u128)kFixedLimit) return LV_ERR_RANGE;
    *S_out = S;
    return LV_OK;
}

void free_all(lv_graph *g) {
    if (g->carve) return;  // carved buffers belong to the fixed region
    if (g->stream) (void)hipStreamSynchronize(g->stream);
    for (auto &r : g->prof_pending) {
        (void)hipEventDestroy(r.a);
        (void)hipEventDestroy(r.b);
    }
    for (auto e : g->ev_pool) (void)hipEventDestroy(e);
    dfree(g->row);
    dfree(g->col);
    dfree(g->w);
    dfree(g->row2);
    dfree(g->col2);
    dfree(g->w2);
    for (void *p : {(void *)g->row0, (void *)g->col0, (void *)g->w0})
        free(p);
    g->row0 = g->col0 = nullptr;
    g->w0 = nullptr;
    dfree(g->k);
    dfree(g->comm);
    dfree(g->comm_sz);
    dfree(g->new_comm);
    dfree(g->A);
    dfree(g->comm2);
    dfree(g->comm_sz2);
    dfree(g->A2);
    dfree(g->init_comm);
    dfree(g->order);
    dfree(g->lid);
    dfree(g->pmap);
    dfree(g->rtmp);
    dfree(g->rrank);
    dfree(g->heavy_caps);
    dfree(g->heavy_off);
    dfree(g->heavy_nch);
    dfree(g->heavy_coff);
    dfree(g->hscr);
    dfree(g->hp_nbuck);
    dfree(g->hp_bstart);
    dfree(g->hp_flag);
    dfree(g->hp_bcnt);
    dfree(g->hp_bpos);
    dfree(g->hp_bcur);
    dfree(g->hp_btmp);
    dfree(g->hp_bvtx);
    dfree(g->hp_cvtx);
    dfree(g->hp_flist);
    dfree(g->ap_nbk);
    dfree(g->ap_bstart);
    dfree(g->ap_cflag);
    dfree(g->ap_nflag);
    dfree(g->ap_bcnt);
    dfree(g->ap_bpos);
    dfree(g->ap_bcur);
    dfree(g->ap_bnd);
    dfree(g->ap_boff);
    dfree(g->ap_bowner);
    dfree(g->ap_btmp);
    dfree(g->gtab);
    dfree(g->dlist);
    dfree(g->eslot);
    dfree(g->new_id);
    dfree(g->vnid);
    dfree(g->comm_deg);
    dfree(g->epos);
    dfree(g->mstart);
    dfree(g->mcursor);
    dfree(g->members);
    dfree(g->cnt);
    dfree(g->aorder);
    dfree(g->tkey);
    dfree(g->skey);
    dfree(g->tval);
    dfree(g->sval);
    dfree(g->agg_caps);
    dfree(g->agg_off);
    dfree(g->agg_mdeg);
    dfree(g->agg_moff);
    dfree(g->agg_cnk);
    dfree(g->agg_flist);
    g->agg_flcap = 0;
    dfree(g->agg_coff);
    dfree(g->agg_dcnt);
    dfree(g->scan_tmp32);
    dfree(g->scan_tmp64);
    dfree(g->dc);
    dfree(g->pst);
    dfree(g->tail_d);
    if (g->hc) (void)hipHostFree(g->hc);
    g->hc = nullptr;
    if (g->pst_h) (void)hipHostFree(g->pst_h);
    g->pst_h = nullptr;
    dfree(g->gid);
    dfree(g->gm_d);
    dfree(g->gstate_d);
    dfree(g->voff_d);
    dfree(g->vgather);
    dfree(g->gcnt);
    dfree(g->gst_d);
    dfree(g->bsum_d);
    if (g->bsum_h) (void)hipHostFree(g->bsum_h);
    g->bsum_h = nullptr;
    if (g->ev_fork) (void)hipEventDestroy(g->ev_fork);
    if (g->ev_join) (void)hipEventDestroy(g->ev_join);
    g->ev_fork = g->ev_join = nullptr;
    for (auto &e : g->ev_launch) {
        if (e) (void)hipEventDestroy(e);
        e = nullptr;
    }
    if (g->stream2) (void)hipStreamDestroy(g->stream2);
    g->stream2 = nullptr;
    if (g->own_stream) (void)hipStreamDestroy(g->own_stream);
    g->own_stream = g->stream = nullptr;
}

// The pass loop of main (main.cu:43-48): modularity optimisation, then
// aggregation while the pass improved; per-run figures into *res (not the
// time).
int run_passes(lv_graph *g, float min_gain, lv_result *res) {
    int rc = LV_OK;
    double bytes = 0;
    for (;;) {
        const double V = g->n, E = (double)g->ne;
        int imp = 0;
        uint32_t it = 0;
        float qp = 0.0f;
        rc = modularity_optimization(g, min_gain, &imp, &qp, &it);
        if (rc) break;
        res->total_iters += it;
        res->sum_edges += g->ne;
        res->sum_edge_visits += (uint64_t)(it + 1) * g->ne;
        bytes += 8 * E + 8 * V + it * (24 * E + 44 * V);
        if (!imp) {
            res->q_final = qp;
            break;
        }
        rc = aggregation(g);
        if (rc) break;
        bytes += 12 * E + 8 * (double)g->ne + 24 * V + 4 * (double)g->n;
        res->passes++;
    }
    res->bytes_alg = bytes;
    res->n_final = g->n;
    res->status = rc;
    return rc;
}

// original vertex -> vertex of the current graph, into host memory
int download_assignment(lv_graph *g, uint32_t *assignment) {
    const uint32_t *src = g->init_comm;
    if (g->relabeled) {  // positions of the current pass -> its input ids
        lv::launch_gather_u32(g->stream, g->lid, g->init_comm, g->rtmp, g->n0);
        src = g->rtmp;
    }
    LV_HIP(hipMemcpyAsync(assignment, src, (size_t)g->n0 * sizeof(uint32_t),
                          hipMemcpyDeviceToHost, g->stream));
    LV_HIP(hipStreamSynchronize(g->stream));
    return LV_OK;
}

}  // namespace

// ===========================================================================
// C ABI
// ===========================================================================
namespace {

// p[lid[x]] = old p[x]
template <typename T>
void permute_by(T *p, const std::vector<uint32_t> &lid) {
    if (!p) return;
    std::vector<T> t(p, p + lid.size());
    for (size_t x = 0; x < lid.size(); ++x) p[lid[x]] = t[x];
}

// positions -> input ids of the current pass (host copy of lid)
int host_lid(lv_graph *g, std::vector<uint32_t> &lid) {
    lid.resize(g->n);
    LV_HIP(hipMemcpyAsync(lid.data(), g->lid, (size_t)g->n * 4, hipMemcpyDeviceToHost, g->stream));
    LV_HIP(hipStreamSynchronize(g->stream));
    return LV_OK;
}

}  // namespace

namespace lv_host {

// Input checks of lv_graph_create: CSR shape, ids in range, weights finite,
// non-negative and within the exact fixed-point range.
int check_input(uint32_t n, const uint32_t *row, const uint32_t *col, const float *w, int *S) {
    if (!row) return LV_ERR_ARG;
    if (n >= 0x80000000u) return LV_ERR_ARG;
    if (row[0] != 0) return LV_ERR_ARG;
    std::vector<char> bad(16, 0);
    host_parallel(n, [&](uint64_t a, uint64_t b, unsigned t) {
        for (uint64_t v = a; v < b; ++v)
            if (row[v + 1] < row[v]) {
                bad[t] = 1;
                return;
            }
    });
    for (char x : bad)
        if (x) return LV_ERR_ARG;
    const uint64_t ne = row[n];
    if (ne && (!col || !w)) return LV_ERR_ARG;
    host_parallel(ne, [&](uint64_t a, uint64_t b, unsigned t) {
        for (uint64_t j = a; j < b; ++j)
            if (col[j] >= n) {
                bad[t] = 1;
                return;
            }
    });
    for (char x : bad)
        if (x) return LV_ERR_ARG;
    return scale_bits(ne, w, S);
}

// Every buffer a pass needs, for up to n_cap vertices and e_cap entries.
// The arenas that depend on the pass graph (heavy tables and buckets, merge
// buckets and tables of large communities) grow on demand, except on a
// carved handle, which reserves them here at their worst case so that no
// pass ever needs more:
//   heavy rows have more than 319 entries, so nh <= E/320;
//   chunks   sum ceil(d/2048)           <= E/2048 + nh       < E/256
//   buckets  sum pow2 >= d/64 (<= 2d/64) <= E/32 + nh        < E/16
//   tables   sum 2d                     <= 2E slots
//   large communities have more than 2048 member entries, so nl <= E/2048;
//   merge buckets  sum (2d/256 + 1)     <= E/128 + nl  (the smallest target)
//   merge tables   sum 2d               <= 2E slots (the heavy rows' arena)
//   chunk list     sum ceil(d/2048) + nl + 2 <= E/1024 + 2 + nl.
// The same walk over a counting region gives lv_graph_footprint.
int graph_buffers(lv_graph *g, uint32_t n_cap, uint64_t e_cap) {
    g->n_cap = n_cap;
    g->e_cap = e_cap;
    const size_t N = (size_t)n_cap + 1, E = (size_t)e_cap;
    int e = LV_OK;
    auto A = [&](auto **p, size_t count) {
        if (!e) e = galloc(g, p, count);
    };
    // 4 entries of padding: 16-byte loads of the masked sums may start
    // inside the last row and end past it
    A(&g->row, N), A(&g->col, E + 4), A(&g->w, E + 4);
    A(&g->row2, N), A(&g->col2, E + 4), A(&g->w2, E + 4);
    A(&g->k, N), A(&g->comm, N), A(&g->comm_sz, N);
    A(&g->new_comm, N), A(&g->A, N), A(&g->init_comm, N);
    A(&g->comm2, N), A(&g->comm_sz2, N), A(&g->A2, N);
    A(&g->order, N), A(&g->heavy_caps, N), A(&g->heavy_off, N + 1);
    A(&g->lid, N), A(&g->pmap, N), A(&g->rtmp, N), A(&g->rrank, N);
    A(&g->heavy_nch, N), A(&g->heavy_coff, N + 1);
    A(&g->hp_nbuck, N), A(&g->hp_bstart, N + 1), A(&g->hp_flag, N);
    A(&g->dlist, E), A(&g->eslot, E);
    A(&g->new_id, N), A(&g->comm_deg, N), A(&g->epos, N), A(&g->vnid, N);
    A(&g->mstart, N), A(&g->mcursor, N), A(&g->members, N);
    A(&g->cnt, N), A(&g->aorder, N);
    A(&g->tkey, E), A(&g->skey, E), A(&g->tval, E), A(&g->sval, E);
    A(&g->agg_caps, N), A(&g->agg_off, N + 1);
    A(&g->agg_mdeg, N), A(&g->agg_moff, N + 1), A(&g->agg_cnk, N);
    A(&g->agg_coff, N + 1), A(&g->agg_dcnt, N);
    A(&g->ap_nbk, N), A(&g->ap_bstart, N + 1), A(&g->ap_cflag, N);
    A(&g->ap_nflag, 1);
    A(&g->scan_tmp32, lv::scan_tmp_elems(n_cap + 1) + 8);
    A(&g->scan_tmp64, lv::scan_tmp_elems(n_cap + 1) + 8);
    A(&g->dc, 1);
    A(&g->pst, 1);
    A(&g->tail_d, 1);
    if (e || !g->carve) return e;
    const size_t hb = E / 16 + 1, hc = E / 256 + 1, slots = 2 * E + 1;
    A(&g->hp_bcnt, hb), A(&g->hp_bpos, hb + 1), A(&g->hp_bcur, hb);
    A(&g->hp_btmp, lv::scan_tmp_elems((uint32_t)hb) + 8), A(&g->hp_bvtx, hb);
    A(&g->hp_cvtx, hc), A(&g->hp_flist, hc + 1);
    A(&g->gtab, 2 * slots);
    A(&g->hscr, (E / 320 + 1) * 130);
    const size_t ab = E / 128 + E / 2048 + 2, lcap = E / 1024 + E / 2048 + 4;
    A(&g->ap_bcnt, ab), A(&g->ap_bpos, ab + 1), A(&g->ap_bcur, ab), A(&g->ap_bnd, ab);
    A(&g->ap_boff, ab + 1), A(&g->ap_bowner, ab);
    A(&g->ap_btmp, lv::scan_tmp_elems((uint32_t)ab + 1) + 8);
    A(&g->agg_flist, lcap);
    if (e) return e;
    g->hp_bcap = hb;
    g->hp_ccap = hc;
    g->gtab_slots = slots;
    g->hscr_cap = (E / 320 + 1) * 130;
    g->ap_bcap = ab;
    g->agg_flcap = lcap;
    return LV_OK;
}

// Device bytes of a handle for up to n vertices and ne entries with every
// on-demand arena at its worst case: graph_buffers over a counting region.
uint64_t footprint(uint32_t n, uint64_t ne) {
    Carver cv;
    cv.cap = SIZE_MAX;
    cv.count_only = true;
    lv_graph g;
    g.carve = &cv;
    (void)graph_buffers(&g, n, ne);
    return cv.used;
}

// A graph handle with its own streams and device buffers for up to n_cap
// vertices and e_cap entries; with `pristine`, also pinned host memory for a
// copy of the input that lv_graph_reset uploads again.
int graph_new(int device, uint32_t n_cap, uint64_t e_cap, bool pristine, lv_graph **out) {
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= device || device < 0) {
        (void)hipGetLastError();
        return LV_ERR_NO_DEVICE;
    }
    lv_graph *g = new (std::nothrow) lv_graph();
    if (!g) return LV_ERR_NOMEM;
    g->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&g->own_stream) != hipSuccess) {
        (void)hipGetLastError();
        delete g;
        return LV_ERR_HIP;
    }
    g->stream = g->own_stream;
    if (hipStreamCreate(&g->stream2) != hipSuccess ||
        hipEventCreateWithFlags(&g->ev_fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&g->ev_join, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&g->ev_launch[0], hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&g->ev_launch[1], hipEventDisableTiming) != hipSuccess) {
        (void)hipGetLastError();
        free_all(g);
        delete g;
        return LV_ERR_HIP;
    }
    int e = graph_buffers(g, n_cap, e_cap);
    if (!e && pristine) {  // pageable: pinning gigabytes costs more than it saves a reset
        g->row0 = static_cast<uint32_t *>(malloc(((size_t)n_cap + 1) * sizeof(uint32_t)));
        g->col0 = static_cast<uint32_t *>(malloc(std::max<size_t>(e_cap, 1) * sizeof(uint32_t)));
        g->w0 = static_cast<float *>(malloc(std::max<size_t>(e_cap, 1) * sizeof(float)));
        if (!g->row0 || !g->col0 || !g->w0) e = LV_ERR_NOMEM;
    }
    // kernels write the pass loop state straight into this copy
    if (!e && (hipHostMalloc((void **)&g->pst_h, sizeof(lv::PassState),
                             hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess ||
               hipHostGetDevicePointer((void **)&g->pst_hd, g->pst_h, 0) != hipSuccess)) {
        (void)hipGetLastError();
        e = LV_ERR_NOMEM;
    }
    if (!e && hipHostMalloc((void **)&g->hc, sizeof(Counters)) != hipSuccess) {
        (void)hipGetLastError();
        e = LV_ERR_NOMEM;
    }
    if (e) {
        free_all(g);
        delete g;
        return LV_ERR_NOMEM;
    }
    *out = g;
    return LV_OK;
}

// The input graph (n0 vertices, ne0 entries, host arrays) as the working CSR,
// with the identity assignment and, for a batch, its first vertex offsets.
int upload_working(lv_graph *g, const uint32_t *row, const uint32_t *col, const float *w) {
    hipStream_t s = g->stream;
    const size_t N = (size_t)g->n0 + 1, E = (size_t)g->ne0;
    LV_HIP(hipMemcpyAsync(g->row, row, N * sizeof(uint32_t), hipMemcpyHostToDevice, s));
    if (E) {
        LV_HIP(hipMemcpyAsync(g->col, col, E * sizeof(uint32_t), hipMemcpyHostToDevice, s));
        LV_HIP(hipMemcpyAsync(g->w, w, E * sizeof(float), hipMemcpyHostToDevice, s));
    }
    lv::launch_iota_u32(s, g->init_comm, g->n0);
    if (g->G) {
        g->voff = g->voff0;
        LV_HIP(hipMemcpyAsync(g->voff_d, g->voff.data(), (g->G + 1) * sizeof(uint32_t),
                              hipMemcpyHostToDevice, s));
        lv::launch_batch_gid(s, g->voff_d, g->G, g->gid, g->n0);
    }
    LV_HIP(hipStreamSynchronize(s));
    g->n = g->n0;
    g->ne = g->ne0;
    g->pass_ready = false;
    g->relabeled = false;
    g->prelabeled = false;
    return LV_OK;
}

// Upload a checked input (n <= n_cap, row[n] <= e_cap) as the graph of the
// handle, with the identity assignment: through the pristine host copy when
// the handle keeps one, straight from the caller's arrays otherwise.
int graph_upload(lv_graph *g, uint32_t n, const uint32_t *row, const uint32_t *col,
                 const float *w, int S) {
    const uint64_t ne = row[n];
    if (n > g->n_cap || ne > g->e_cap || g->G) return LV_ERR_NOMEM;
    LV_HIP(hipSetDevice(g->device));
    g->n = g->n0 = n;
    g->ne = g->ne0 = ne;
    g->S = S;
    g->scale = ldexpf(1.0f, S);
    g->inv_scale = ldexpf(1.0f, -S);
    if (g->row0) {
        // the pristine copy fills on host threads while the upload runs
        std::thread copy([&] {
            memcpy(g->row0, row, ((size_t)n + 1) * sizeof(uint32_t));
            if (ne)
                host_parallel(ne, [&](uint64_t a, uint64_t b, unsigned) {
                    memcpy(g->col0 + a, col + a, (b - a) * sizeof(uint32_t));
                    memcpy(g->w0 + a, w + a, (b - a) * sizeof(float));
                });
        });
        const int rc = upload_working(g, row, col, w);
        copy.join();
        return rc;
    }
    return upload_working(g, row, col, w);
}

static_assert(sizeof(Counters) <= kHostCountersBytes, "host counter buffer");

int run_fixed(int device, char *base, size_t cap, void *host_counters, uint32_t n,
              const uint32_t *row, const uint32_t *col, const float *w, float min_gain,
              void *stream, uint32_t *assignment, lv_result *res) {
    memset(res, 0, sizeof(*res));
    int S = 0;
    int rc = check_input(n, row, col, w, &S);
    if (rc) return rc;
    const uint64_t ne = row[n];
    Carver cv;
    cv.base = base;
    cv.cap = cap;
    lv_graph g;
    g.device = device;
    g.carve = &cv;
    g.stream = static_cast<hipStream_t>(stream);
    g.hc = static_cast<Counters *>(host_counters);
    rc = graph_buffers(&g, n, ne);
    if (rc) return rc;  // LV_ERR_LIMIT: the graph needs more than the region
    hipStream_t s = g.stream;
    LV_HIP(hipMemcpyAsync(g.row, row, ((size_t)n + 1) * sizeof(uint32_t), hipMemcpyHostToDevice,
                          s));
    if (ne) {
        LV_HIP(hipMemcpyAsync(g.col, col, ne * sizeof(uint32_t), hipMemcpyHostToDevice, s));
        LV_HIP(hipMemcpyAsync(g.w, w, ne * sizeof(float), hipMemcpyHostToDevice, s));
    }
    g.n = g.n0 = n;
    g.ne = g.ne0 = ne;
    g.S = S;
    g.scale = ldexpf(1.0f, S);
    g.inv_scale = ldexpf(1.0f, -S);
    // the region keeps the previous call's contents: everything that must
    // start empty is cleared here (the rest is written before it is read)
    lv::launch_iota_u32(s, g.init_comm, n);
    lv::launch_fill_slots(s, g.gtab, g.gtab_slots);
    LV_HIP(hipStreamSynchronize(s));
    const auto t0 = std::chrono::steady_clock::now();
    rc = run_passes(&g, min_gain, res);
    if (!rc)
        rc = download_assignment(&g, assignment);  // synchronizes the stream
    else
        (void)hipStreamSynchronize(s);
    res->alg_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0)
                      .count();
    res->status = rc;
    return rc;
}

}  // namespace lv_host
using lv_host::check_input;
using lv_host::graph_new;
using lv_host::graph_upload;
using lv_host::upload_working;

extern "C" {

const char *lv_status_string(int st) {
    switch (st) {
        case LV_OK: return "ok";
        case LV_ERR_IO: return "Error opening file";
        case LV_ERR_FORMAT: return "malformed .mtx input";
        case LV_ERR_RANGE: return "weights outside the exact fixed-point range";
        case LV_ERR_ARG: return "invalid argument";
        case LV_ERR_NOMEM: return "out of memory";
        case LV_ERR_STATE: return "call out of order";
        case LV_ERR_HIP: return "HIP runtime error";
        case LV_ERR_NO_DEVICE: return "no GPU device";
        case LV_ERR_OVERFLOW: return "internal capacity exceeded";
        case LV_ERR_LIMIT: return "graph larger than the fixed scratch array allows";
        default: return "unknown status";
    }
}

int lv_device_count(int *count) {
    if (!count) return LV_ERR_ARG;
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess) {
        (void)hipGetLastError();
        *count = 0;
        return LV_ERR_NO_DEVICE;
    }
    *count = c;
    return LV_OK;
}

int lv_graph_create(int device, uint32_t n, const uint32_t *row, const uint32_t *col,
                    const float *w, lv_graph **out) {
    if (!out || !row) return LV_ERR_ARG;
    *out = nullptr;
    int S = 0;
    int rc = check_input(n, row, col, w, &S);
    if (rc) return rc;
    lv_graph *g = nullptr;
    rc = graph_new(device, n, row[n], true, &g);
    if (rc) return rc;
    rc = graph_upload(g, n, row, col, w, S);
    if (rc) {
        lv_graph_destroy(g);
        return rc;
    }
    *out = g;
    return LV_OK;
}

int lv_graph_reset(lv_graph *g) {
    if (!g) return LV_ERR_ARG;
    if (!g->row0) return LV_ERR_STATE;  // a handle without a pristine copy
    LV_HIP(hipSetDevice(g->device));
    return upload_working(g, g->row0, g->col0, g->w0);
}

int lv_graph_set_locality(lv_graph *g, int mode) {
    if (!g || mode < -1 || mode > 1) return LV_ERR_ARG;
    g->tune.locality = mode;
    return LV_OK;
}

int lv_tuning_default(lv_tuning *t) {
    if (!t) return LV_ERR_ARG;
    *t = lv_graph().tune;
    return LV_OK;
}

int lv_graph_set_tuning(lv_graph *g, const lv_tuning *t) {
    if (!g || !t) return LV_ERR_ARG;
    const uint32_t b = t->bucket_entries;
    // heavy rows must stay in bucket 7 (deg > 319) and every shorter bucket-7
    // row must fit a block kernel's table
    if (t->heavy_min_degree < lv::kCut[6] || t->heavy_min_degree > lv::kBlockMax) return LV_ERR_ARG;
    if (b < 64 || b > 1024 || (b & (b - 1))) return LV_ERR_ARG;
    if (t->locality < -1 || t->locality > 1) return LV_ERR_ARG;
    if (t->wave_vertices != 1 && t->wave_vertices != 2 && t->wave_vertices != 4) return LV_ERR_ARG;
    if (t->block_threads != 512 && t->block_threads != 1024) return LV_ERR_ARG;
    if (t->class8_slots != 1280 && t->class8_slots != 2048) return LV_ERR_ARG;
    if (t->agg_bucket_entries < 256 || t->agg_bucket_entries > 512) return LV_ERR_ARG;
    if (t->sweeps_per_launch != 1 && t->sweeps_per_launch != 2 && t->sweeps_per_launch != 4)
        return LV_ERR_ARG;
    g->tune = *t;
    return LV_OK;
}

int lv_graph_get_tuning(lv_graph *g, lv_tuning *t) {
    if (!g || !t) return LV_ERR_ARG;
    *t = g->tune;
    return LV_OK;
}

int lv_autotune(lv_graph *g, float min_gain, const lv_tuning *candidates, uint32_t count,
                lv_tuning *best, double *best_ms) {
    if (!g || (!candidates && count)) return LV_ERR_ARG;
    const lv_tuning keep = g->tune;
    // best of two runs of the resident input under t
    auto time_of = [&](const lv_tuning &t, double *ms) -> int {
        int rc = lv_graph_set_tuning(g, &t);
        if (rc) return rc;
        for (int rep = 0; rep < 2; ++rep) {
            rc = lv_graph_reset(g);
            lv_result r;
            if (!rc) rc = g->G ? lv_batch_run(g, min_gain, nullptr, &r) : lv_run(g, min_gain, &r);
            if (rc) return rc;
            if (rep == 0 || r.alg_ms < *ms) *ms = r.alg_ms;
        }
        return LV_OK;
    };
    lv_tuning bt = keep;
    double bms = 0.0;
    if (candidates) {
        if (!count) return LV_ERR_ARG;
        for (uint32_t i = 0; i < count; ++i) {
            double ms = 0.0;
            const int rc = time_of(candidates[i], &ms);
            if (rc) {
                g->tune = keep;
                lv_graph_reset(g);
                return rc;
            }
            if (i == 0 || ms < bms) {
                bms = ms;
                bt = candidates[i];
            }
        }
    } else {
        // one field at a time over its values, the others at the best so far
        int rc = time_of(bt, &bms);
        const bool big = g->n0 >= (1u << 18) && !g->G;
        struct Axis {
            uint32_t lv_tuning::*u;
            std::vector<uint32_t> values;
        };
        const Axis axes[] = {
            {&lv_tuning::heavy_min_degree, {2048u, 4096u, lv::kBlockMax}},
            {&lv_tuning::bucket_entries, {256u, 512u, 1024u}},
            {&lv_tuning::wave_vertices, {1u, 2u, 4u}},
            {&lv_tuning::block_threads, {512u, 1024u}},
            {&lv_tuning::class8_slots, {1280u, 2048u}},
            {&lv_tuning::agg_bucket_entries, {256u, 384u, 512u}},
            {&lv_tuning::sweeps_per_launch, {1u, 2u, 4u}},
        };
        for (const Axis &ax : axes) {
            for (uint32_t val : ax.values) {
                if (rc || bt.*ax.u == val) continue;
                lv_tuning t = bt;
                t.*ax.u = val;
                double ms = 0.0;
                rc = time_of(t, &ms);
                if (!rc && ms < bms) {
                    bms = ms;
                    bt = t;
                }
            }
        }
        for (int loc : {0, 1}) {
            if (rc || !big || bt.locality == loc) continue;
            lv_tuning t = bt;
            t.locality = loc;
            double ms = 0.0;
            rc = time_of(t, &ms);
            if (!rc && ms < bms) {
                bms = ms;
                bt = t;
            }
        }
        if (rc) {
            g->tune = keep;
            lv_graph_reset(g);
            return rc;
        }
    }
    g->tune = bt;
    if (best) *best = bt;
    if (best_ms) *best_ms = bms;
    return lv_graph_reset(g);
}

int lv_graph_destroy(lv_graph *g) {
    if (!g) return LV_ERR_ARG;
    (void)hipSetDevice(g->device);
    free_all(g);
    delete g;
    return LV_OK;
}

int lv_pass_init(lv_graph *g, float *q0) {
    if (!g) return LV_ERR_ARG;
    LV_HIP(hipSetDevice(g->device));
    return pass_init(g, q0);
}

int lv_step(lv_graph *g, int bucket, uint64_t *moved) {
    if (!g || bucket < 1 || bucket > 7) return LV_ERR_ARG;
    if (!g->pass_ready) return LV_ERR_STATE;
    LV_HIP(hipSetDevice(g->device));
    zero_counters(g);
    uint32_t a, e;
    step_range(g, bucket, &a, &e);
    if (!fused_pass(g)) {
        step_apply(g, bucket);
    } else if (e > a) {
        // the step, then its moves applied to the copy it read
        step_fused(g, bucket, 0, 0, 0);
        lv::DevView v = step_view(g, 1);
        v.cu_verts = g->order + a;
        v.cu_count = e - a;
        lv::launch_catchup(g->stream, v);
    }
    int rc = sync_counters(g);
    if (rc) return rc;
    if (g->hc->err) return LV_ERR_OVERFLOW;
    if (moved) *moved = g->hc->tot[5];
    return LV_OK;
}

int lv_sweep(lv_graph *g, float *q, uint64_t *moved) {
    if (!g) return LV_ERR_ARG;
    if (!g->pass_ready) return LV_ERR_STATE;
    LV_HIP(hipSetDevice(g->device));
    return sweep(g, q, moved);
}

int lv_modularity(lv_graph *g, float *q) {
    if (!g || !q) return LV_ERR_ARG;
    if (!g->pass_ready) return LV_ERR_STATE;
    LV_HIP(hipSetDevice(g->device));
    zero_counters(g);
    modularity_launch(g);
    int rc = sync_counters(g);
    if (rc) return rc;
    *q = q_from_counters(g);
    return LV_OK;
}

int lv_modularity_optimization(lv_graph *g, float min_gain, int *improved, float *q_before,
                               float *q_after) {
    if (!g) return LV_ERR_ARG;
    LV_HIP(hipSetDevice(g->device));
    int imp = 0;
    int rc = modularity_optimization(g, min_gain, &imp, nullptr, nullptr);
    if (rc) return rc;
    if (improved) *improved = imp;
    if (q_before) *q_before = g->last_q_before;
    if (q_after) *q_after = g->last_q_after;
    return LV_OK;
}

int lv_pass_stats(lv_graph *g, uint32_t *sweeps, float *q_before, float *q_after) {
    if (!g) return LV_ERR_ARG;
    if (sweeps) *sweeps = g->last_iters;
    if (q_before) *q_before = g->last_q_before;
    if (q_after) *q_after = g->last_q_after;
    return LV_OK;
}

int lv_graph_set_stream(lv_graph *g, void *stream) {
    if (!g || g->carve) return LV_ERR_ARG;
    LV_HIP(hipSetDevice(g->device));
    // work already queued on the previous stream completes first
    LV_HIP(hipStreamSynchronize(g->stream));
    g->stream = stream ? static_cast<hipStream_t>(stream) : g->own_stream;
    return LV_OK;
}

int lv_aggregation(lv_graph *g) {
    if (!g) return LV_ERR_ARG;
    LV_HIP(hipSetDevice(g->device));
    return aggregation(g);
}

int lv_run(lv_graph *g, float min_gain, lv_result *res) {
    if (!g || !res) return LV_ERR_ARG;
    memset(res, 0, sizeof(*res));
    LV_HIP(hipSetDevice(g->device));
    hipEvent_t t0 = nullptr, t1 = nullptr;
    LV_HIP(hipEventCreate(&t0));
    LV_HIP(hipEventCreate(&t1));
    LV_HIP(hipEventRecord(t0, g->stream));
    const int rc = run_passes(g, min_gain, res);
    (void)hipEventRecord(t1, g->stream);
    (void)hipEventSynchronize(t1);
    float ms = 0.0f;
    (void)hipEventElapsedTime(&ms, t0, t1);
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    res->alg_ms = ms;
    return rc;
}

int lv_get_assignment(lv_graph *g, uint32_t *assignment) {
    if (!g || !assignment) return LV_ERR_ARG;
    LV_HIP(hipSetDevice(g->device));
    return download_assignment(g, assignment);
}

int lv_graph_size(lv_graph *g, uint32_t *n, uint64_t *ne, uint32_t *n0) {
    if (!g) return LV_ERR_ARG;
    if (n) *n = g->n;
    if (ne) *ne = g->ne;
    if (n0) *n0 = g->n0;
    return LV_OK;
}

int lv_get_csr(lv_graph *g, uint32_t *row, uint32_t *col, float *w) {
    if (!g) return LV_ERR_ARG;
    LV_HIP(hipSetDevice(g->device));
    hipStream_t s = g->stream;
    const size_t n = g->n, ne = g->ne;
    if (g->relabeled) {  // reported in the input numbering, rows in input order
        std::vector<uint32_t> lid, r(n + 1), c(ne);
        std::vector<float> x(ne);
        int rc = host_lid(g, lid);
        if (rc) return rc;
        LV_HIP(hipMemcpyAsync(r.data(), g->row, (n + 1) * 4, hipMemcpyDeviceToHost, s));
        if (ne) {
            LV_HIP(hipMemcpyAsync(c.data(), g->col, ne * 4, hipMemcpyDeviceToHost, s));
            LV_HIP(hipMemcpyAsync(x.data(), g->w, ne * 4, hipMemcpyDeviceToHost, s));
        }
        LV_HIP(hipStreamSynchronize(s));
        std::vector<uint32_t> pos(n), deg(n + 1, 0);
        for (size_t p = 0; p < n; ++p) {
            pos[lid[p]] = (uint32_t)p;
            deg[lid[p] + 1] = r[p + 1] - r[p];
        }
        for (size_t v = 0; v < n; ++v) deg[v + 1] += deg[v];
        if (row) memcpy(row, deg.data(), (n + 1) * 4);
        for (size_t v = 0; v < n; ++v) {
            const uint32_t p = pos[v], b = r[p], d = r[p + 1] - b;
            for (uint32_t j = 0; j < d; ++j) {
                if (col) col[deg[v] + j] = lid[c[b + j]];
                if (w) w[deg[v] + j] = x[b + j];
            }
        }
        return LV_OK;
    }
    if (row) LV_HIP(hipMemcpyAsync(row, g->row, (n + 1) * 4, hipMemcpyDeviceToHost, s));
    if (col && ne) LV_HIP(hipMemcpyAsync(col, g->col, ne * 4, hipMemcpyDeviceToHost, s));
    if (w && ne) LV_HIP(hipMemcpyAsync(w, g->w, ne * 4, hipMemcpyDeviceToHost, s));
    LV_HIP(hipStreamSynchronize(s));
    return LV_OK;
}

int lv_get_state(lv_graph *g, float *k, uint32_t *comm, uint32_t *comm_sz, int64_t *a) {
    if (!g) return LV_ERR_ARG;
    LV_HIP(hipSetDevice(g->device));
    hipStream_t s = g->stream;
    const size_t n = g->n;
    if (k) LV_HIP(hipMemcpyAsync(k, g->k, n * 4, hipMemcpyDeviceToHost, s));
    if (comm) LV_HIP(hipMemcpyAsync(comm, g->comm, n * 4, hipMemcpyDeviceToHost, s));
    if (comm_sz) LV_HIP(hipMemcpyAsync(comm_sz, g->comm_sz, n * 4, hipMemcpyDeviceToHost, s));
    if (a) LV_HIP(hipMemcpyAsync(a, g->A, n * 8, hipMemcpyDeviceToHost, s));
    LV_HIP(hipStreamSynchronize(s));
    if (g->relabeled) {  // positions -> input ids (vertex and community ids alike)
        std::vector<uint32_t> lid;
        int rc = host_lid(g, lid);
        if (rc) return rc;
        if (comm)
            for (size_t x = 0; x < n; ++x) comm[x] = lid[comm[x]];
        permute_by(k, lid);
        permute_by(comm, lid);
        permute_by(comm_sz, lid);
        permute_by(a, lid);
    }
    return LV_OK;
}

int lv_get_scale(lv_graph *g, int *scale_bits_out, float *m) {
    if (!g) return LV_ERR_ARG;
    if (scale_bits_out) *scale_bits_out = g->S;
    if (m) *m = g->m;
    return LV_OK;
}

int lv_batch_create(int device, uint32_t ngraphs, const uint64_t *graph_off, const uint32_t *row,
                    const uint32_t *col, const float *w, lv_graph **out) {
    if (!out || !graph_off || !row || ngraphs == 0) return LV_ERR_ARG;
    *out = nullptr;
    if (graph_off[0] != 0) return LV_ERR_ARG;
    for (uint32_t gi = 0; gi < ngraphs; ++gi)
        if (graph_off[gi + 1] < graph_off[gi]) return LV_ERR_ARG;
    const uint64_t n64 = graph_off[ngraphs];
    if (n64 >= 0x80000000ull) return LV_ERR_ARG;
    const uint32_t n = (uint32_t)n64;
    if (row[0] != 0) return LV_ERR_ARG;
    for (uint32_t v = 0; v < n; ++v)
        if (row[v + 1] < row[v]) return LV_ERR_ARG;
    const uint64_t ne = row[n];
    if (ne && (!col || !w)) return LV_ERR_ARG;
    const uint32_t G = ngraphs;
    // per graph, on the host threads: every entry stays inside its graph's
    // vertex range, weights are valid, and the graph's own fixed-point shift
    std::vector<int> gS(G, 0), gbad(G, 0);
    host_parallel(G, [&](uint64_t a, uint64_t b, unsigned) {
        for (uint64_t gi = a; gi < b; ++gi) {
            const uint64_t e0 = row[graph_off[gi]], e1 = row[graph_off[gi + 1]];
            for (uint64_t j = e0; j < e1; ++j)
                if (col[j] < graph_off[gi] || col[j] >= graph_off[gi + 1]) {
                    gbad[gi] = LV_ERR_ARG;
                    break;
                }
            if (!gbad[gi]) gbad[gi] = min_shift(w, e0, e1, &gS[gi]);
        }
    });
    int S = 0;
    for (uint32_t gi = 0; gi < G; ++gi) {
        if (gbad[gi]) return gbad[gi];
        S = std::max(S, gS[gi]);
    }
    // One scale serves the batch: exact values and their FP32 roundings do
    // not depend on it.  A graph whose total does not fit at the batch's
    // scale (weights of very different magnitudes across graphs) is reported
    // with LV_ERR_RANGE by lv_batch_run and left out of the run; every other
    // graph is unaffected.
    std::vector<int> status(G, LV_OK);
    host_parallel(G, [&](uint64_t a, uint64_t b, unsigned) {
        for (uint64_t gi = a; gi < b; ++gi)
            if (!total_fits(w, row[graph_off[gi]], row[graph_off[gi + 1]], S))
                status[gi] = LV_ERR_RANGE;
    });
    lv_graph *g = nullptr;
    int rc = graph_new(device, n, ne, true, &g);
    if (rc) return rc;
    rc = graph_upload(g, n, row, col, w, S);
    if (rc) {
        lv_graph_destroy(g);
        return rc;
    }
    g->G = G;
    int e = 0;
    e |= dmalloc(&g->gid, (size_t)n + 1) | dmalloc(&g->gm_d, G) | dmalloc(&g->gstate_d, G);
    e |= dmalloc(&g->voff_d, (size_t)G + 1) | dmalloc(&g->vgather, (size_t)G + 1);
    e |= dmalloc(&g->gcnt, 6ull * G) | dmalloc(&g->gst_d, G) | dmalloc(&g->bsum_d, 1);
    if (!e && (hipHostMalloc((void **)&g->bsum_h, sizeof(lv::BatchSummary),
                             hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess ||
               hipHostGetDevicePointer((void **)&g->bsum_hd, g->bsum_h, 0) != hipSuccess)) {
        (void)hipGetLastError();
        e = LV_ERR_NOMEM;
    }
    if (e) {
        lv_graph_destroy(g);
        return LV_ERR_NOMEM;
    }
    g->voff0.resize(G + 1);
    for (uint32_t gi = 0; gi <= G; ++gi) g->voff0[gi] = (uint32_t)graph_off[gi];
    g->voff = g->voff0;
    g->gstate_h.assign(G, kGsActive);
    g->gstatus0 = std::move(status);
    if (hipMemcpy(g->voff_d, g->voff.data(), (G + 1) * sizeof(uint32_t),
                  hipMemcpyHostToDevice) != hipSuccess) {
        (void)hipGetLastError();
        lv_graph_destroy(g);
        return LV_ERR_HIP;
    }
    lv::launch_batch_gid(g->stream, g->voff_d, G, g->gid, n);
    if (hipStreamSynchronize(g->stream) != hipSuccess) {
        (void)hipGetLastError();
        lv_graph_destroy(g);
        return LV_ERR_HIP;
    }
    *out = g;
    return LV_OK;
}

int lv_batch_run(lv_graph *g, float min_gain, lv_batch_graph_result *per_graph,
                 lv_result *total) {
    if (!g || !g->G) return LV_ERR_ARG;
    LV_HIP(hipSetDevice(g->device));
    const uint32_t G = g->G;
    hipEvent_t t0 = nullptr, t1 = nullptr;
    LV_HIP(hipEventCreate(&t0));
    LV_HIP(hipEventCreate(&t1));
    LV_HIP(hipEventRecord(t0, g->stream));
    // graphs without vertices or with weights out of range do not run
    for (uint32_t gi = 0; gi < G; ++gi)
        g->gstate_h[gi] = (g->voff[gi + 1] > g->voff[gi] && !g->gstatus0[gi]) ? kGsActive : kGsDone;
    lv::launch_fill_u32(g->stream, reinterpret_cast<uint32_t *>(g->gst_d), 0u,
                        (size_t)G * sizeof(lv::GraphState) / 4);
    int rc = push_gstate(g);
    bool more = true;
    while (!rc && more) rc = batch_pass(g, min_gain, &more);
    (void)hipEventRecord(t1, g->stream);
    (void)hipEventSynchronize(t1);
    float ms = 0.0f;
    (void)hipEventElapsedTime(&ms, t0, t1);
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    std::vector<lv::GraphState> st(G);
    if (hipMemcpy(st.data(), g->gst_d, G * sizeof(lv::GraphState), hipMemcpyDeviceToHost) !=
        hipSuccess) {
        (void)hipGetLastError();
        if (!rc) rc = LV_ERR_HIP;
    }
    if (total) {
        memset(total, 0, sizeof(*total));
        total->alg_ms = ms;
        total->status = rc;
        total->n_final = g->n;
    }
    for (uint32_t gi = 0; gi < G; ++gi) {
        const lv::GraphState &b = st[gi];
        if (per_graph) {
            per_graph[gi].q_final = b.q_final;
            per_graph[gi].passes = b.passes;
            per_graph[gi].n_final = g->voff[gi + 1] - g->voff[gi];
            per_graph[gi].total_iters = b.iters;
            per_graph[gi].sum_edges = b.sum_edges;
            per_graph[gi].status = g->gstatus0[gi] ? g->gstatus0[gi] : (int)b.status;
        }
        if (total) {
            total->passes += b.passes;
            total->total_iters += b.iters;
            total->sum_edges += b.sum_edges;
            total->sum_edge_visits += b.sum_visits;
        }
    }
    return rc;
}

int lv_batch_get_assignment(lv_graph *g, uint32_t *local_assignment) {
    if (!g || !g->G || !local_assignment) return LV_ERR_ARG;
    int rc = lv_get_assignment(g, local_assignment);
    if (rc) return rc;
    for (uint32_t gi = 0; gi < g->G; ++gi)
        for (uint32_t v = g->voff0[gi]; v < g->voff0[gi + 1]; ++v) local_assignment[v] -= g->voff[gi];
    return LV_OK;
}

int lv_profile_enable(lv_graph *g, int on) {
    if (!g) return LV_ERR_ARG;
    g->prof_on = on != 0;
    return LV_OK;
}

int lv_profile_select(lv_graph *g, uint32_t family_mask) {
    if (!g) return LV_ERR_ARG;
    g->prof_mask = family_mask;
    return LV_OK;
}

int lv_profile_reset(lv_graph *g) {
    if (!g) return LV_ERR_ARG;
    LV_HIP(hipStreamSynchronize(g->stream));
    for (auto &r : g->prof_pending) {
        g->ev_pool.push_back(r.a);
        g->ev_pool.push_back(r.b);
    }
    g->prof_pending.clear();
    memset(&g->prof, 0, sizeof(g->prof));
    return LV_OK;
}

int lv_profile_get(lv_graph *g, lv_kernel_stats *out) {
    if (!g || !out) return LV_ERR_ARG;
    LV_HIP(hipSetDevice(g->device));
    LV_HIP(hipStreamSynchronize(g->stream));
    for (auto &r : g->prof_pending) {
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, r.a, r.b) == hipSuccess) {
            g->prof.ms[r.fam] += ms;
            g->prof.launches[r.fam] += 1;
            g->prof.bytes[r.fam] += r.bytes;
        } else {
            (void)hipGetLastError();
        }
        g->ev_pool.push_back(r.a);
        g->ev_pool.push_back(r.b);
    }
    g->prof_pending.clear();
    *out = g->prof;
    return LV_OK;
}

}  // extern "C"
