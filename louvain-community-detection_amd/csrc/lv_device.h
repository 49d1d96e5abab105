// lv_device.h -- device-side constants, state views and launcher declarations
// shared by the kernel translation units and the host engine.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lv {

// Reference bucket cut points (modularity.cu:439): bucket b holds
// cut[b-1] < deg <= cut[b].  These are semantic (they define the synchronous
// step sets) and are never tuned.
constexpr uint32_t kCut[8] = {0u, 4u, 8u, 16u, 32u, 84u, 319u, 0x7fffffffu};
// Virtual lanes per vertex = blockDim.x of the reference kernel serving each
// bucket (modularity.cu:443-450).  Semantic: they fix each lane's visiting
// order (per-lane first maximum, bucket 5 round prefixes, last-edge rule).
constexpr uint32_t kLanes[7] = {4u, 8u, 16u, 32u, 32u, 128u, 128u};

// Kernel classes of the move step (performance choice, not semantic).
//   0: deg 0 (never moves)      1-4: buckets 1-4, sub-wave groups of 4/8/16/32
//   5: bucket 5, 32-lane groups 6: bucket 6, one wave64 per vertex
//   7-11: bucket 7 with deg <= 512 / 1024 / 2048 / 4096 / 8192: one workgroup
//        per vertex, LDS table of 1024 / 1280 / 3072 / 5800 / 10240 slots
//   12: bucket 7 rows longer than the tuning's heavy threshold (at most
//       kBlockMax): rows split over many workgroups
constexpr int kNumClasses = 13;
constexpr int kClassHeavy = 12;
constexpr uint32_t kBlockMax = 8192;
constexpr uint32_t kEmpty = 0xffffffffu;
// Exact integer totals are accumulated into kSpread slots (block b adds to
// slot b % kSpread) and summed on the host: same-address atomics from
// thousands of blocks serialise in L2.
constexpr uint32_t kSpread = 256;

__host__ __device__ inline int vertex_class(uint32_t d) {
    if (d == 0) return 0;
    if (d <= 4) return 1;
    if (d <= 8) return 2;
    if (d <= 16) return 3;
    if (d <= 32) return 4;
    if (d <= 84) return 5;
    if (d <= 319) return 6;
    if (d <= 512) return 7;
    if (d <= 1024) return 8;
    if (d <= 2048) return 9;
    if (d <= 4096) return 10;
    if (d <= kBlockMax) return 11;
    return kClassHeavy;
}

// Aggregation community classes (by total member degree).
//   0: no entries (empty row, no kernel)                   (comm_deg == 0)
//   1: LDS merge, one wave per community, 256-slot table   (comm_deg <= 128)
//   2-4: LDS merge, one workgroup, 1024/2048/4096-slot table
//                                          (comm_deg <= 512 / 1024 / 2048)
//   5: global-hash merge + workgroup radix sort            (comm_deg >  2048)
constexpr int kNumAggClasses = 6;
constexpr int kAggGlobalClass = 5;
constexpr uint32_t kAggSmallMax = 128;
constexpr uint32_t kAggMidMax = 2048;

// View of one graph's device state, passed by value to kernels.
struct SweepTail;
struct DevView {
    const uint32_t *row;    // n+1
    const uint32_t *col;    // ne
    const float *w;         // ne
    uint32_t *comm;         // n
    uint32_t *comm_sz;      // n
    uint32_t *new_comm;     // n: a step's decisions (one copy of the state)
    unsigned long long *A;  // n: a_c * 2^S (exact)
    float *k;               // n
    uint32_t n;
    int S;          // fixed-point scale bits
    float scale;    // 2^S
    float m;        // all_w (single-graph mode)
    // batch mode (gid != nullptr): many graphs in one segmented CSR
    const uint32_t *gid;     // vertex -> graph
    const float *gm;         // all_w per graph
    const uint32_t *gstate;  // per graph; only kGraphActive graphs move
    uint32_t G;              // graphs (stride of the per-graph accumulators)
    // Locality order (single-graph mode): when non-null, vertex and community
    // ids on the device are positions in a degree-class order and lid[x] is
    // the id x stands for in the input numbering.  Every comparison of ids
    // that the algorithm makes (tie-breaks, isLowerIdx, renumbering) uses
    // lid, so results do not depend on the order.
    const uint32_t *lid;
    // Sweeps queued ahead of the host's read of the loop state (several per
    // wait, or the next group behind the one waited for): non-null once the
    // pass loop may have stopped on the device; when *halt is set the
    // kernels of the remaining queued sweeps return at entry.
    const uint32_t *halt;
    // Two copies of the community state (small single graphs, where a
    // separate apply launch per step costs more than its work; comm2 == null:
    // one copy, new_comm and k_apply).  A bucket step reads its snapshot
    // from (comm, comm_sz, A) and applies each of its moves at once to
    // (comm2, comm_sz2, A2), which no kernel of the step reads; the next step
    // reads that copy.  The copy a step writes still lacks the previous
    // step's moves: the first cu_nblk workgroups of the step's first kernel
    // apply them (the previous step's vertices cu_verts[0, cu_count), where
    // the two copies can differ) and count them into cu_moved.  Pass init
    // writes both copies (comm2 non-null); other kernels read one copy.
    uint32_t *comm2;
    uint32_t *comm_sz2;
    unsigned long long *A2;
    const uint32_t *cu_verts;
    uint32_t cu_count, cu_nblk;
    unsigned long long *cu_moved;
    // >= 0: the launch's vertex list is the identity run starting there (the
    // class order of a graph in its locality order), so kernels compute
    // vertex ids instead of loading them; -1: load them
    int32_t vbase;
    // Non-null (a sweep's first move kernel, the pass loop decided on the
    // device): workgroup 0 applies the previous sweep's stop rule first
    // (k_collapse_decide's work, one launch fewer per sweep); the kernel's
    // other workgroups run the step ahead of that decision, and the host
    // drops their moves when it stops the pass.
    const SweepTail *tail;
};
constexpr uint32_t kGraphActive = 1;
constexpr uint32_t kGraphWait = 2;  // improved in this pass, waits for the contraction
constexpr uint32_t kGraphDone = 3;

// The loop variables of modularityOptimization (modularity.cu:464-508) for
// one graph, kept on the device so the stop rule is applied there after
// every sweep.
struct PassState {
    float act, old, q_print;
    uint32_t it;     // sweeps of the pass
    uint32_t first;  // the next sweep is the pass's first
    uint32_t ret;    // the pass improved (modularityOptimization's result)
    uint32_t done;   // the pass loop has stopped
    uint32_t err;    // a capacity overflow was flagged during a sweep
};
// A batch graph: its pass loop variables, its m, and its run totals.
struct GraphState {
    float act, old, q_final, m;
    uint32_t pass_iters, iters, first, ret;
    uint32_t passes, status, pad0, pad1;
    unsigned long long ne, sum_edges, sum_visits, pad2;
};
// What the host reads after a batch's group of sweeps.  The first three
// fields are recounted after every sweep (zeroed before it); halt and err
// stay set for the rest of the batch pass.
struct BatchSummary {
    uint32_t active;          // graphs still sweeping
    uint32_t waiting;         // graphs that improved and wait for the contraction
    unsigned long long live;  // vertices of the active graphs
    uint32_t halt;            // no graph is active: later queued sweeps return at entry
    uint32_t err;             // a capacity overflow was flagged
};
constexpr int kBatchSummaryCounted = 4;  // 32-bit words recounted per sweep

// Heavy-vertex state for the chunked kernels (lv_heavy.hip).
constexpr uint32_t kHeavyChunk = 2048;  // entries per workgroup; multiple of 256
constexpr uint32_t kMaxBuckets = 1024;    // buckets per heavy vertex (max)
constexpr uint32_t kHpGroup = 8;          // buckets per workgroup (min per vertex)
struct HeavySplit {
    const uint32_t *verts;               // heavy-class vertices
    uint32_t nh;                         // their count
    uint32_t nchunks;                    // total chunks = coff[nh]
    const uint32_t *coff;                // nh+1 chunk offsets
    const uint32_t *cvtx;                // nchunks: heavy index owning each chunk
    const unsigned long long *slot_off;  // nh+1 table offsets in the arena
    unsigned long long *tab;             // arena of 16-byte slots {key, pad, sum}
    uint32_t *dlist;                     // ne: distinct slots at [row[v], row[v+1])
    uint32_t *eslot;                     // ne: slot of each entry
    unsigned long long *hcnt;            // nh distinct counts, followed by
    unsigned long long *hcew;            // nh comm_e_w sums and
    unsigned long long *hlane;           // nh*128 packed lane maxima (one allocation)
    // partitioned path
    const uint32_t *nbuck;               // nh: buckets per vertex (pow2)
    const uint32_t *bstart;              // nh+1: first bucket of each vertex
    uint32_t nbuckets;                   // total buckets = bstart[nh]
    uint32_t bucket_entries;             // target entries per bucket (tuning)
    const uint32_t *bvtx;                // nbuckets: heavy index owning each bucket
    uint32_t *bcnt, *bpos, *bcur;        // nbuckets / nbuckets+1 / nbuckets
    uint32_t *pc, *pj;                   // heavy-entry space: community, entry index
    float *pw;                           //                    weight
    uint32_t *hflag;                     // nh: LDS overflow, redo on global tables
    uint32_t *flist;                     // 1 + nchunks: count, then the chunks of flagged vertices
    int force_global;                    // test hook: 1 global tables only, 2 mixed
    uint32_t dmax;                       // distinct communities a bucket may hold (test hook)
    uint32_t *scan_tmp;
    uint32_t *err;
};

// ---- device-side pass loop (lv_kernels.hip) ---------------------------------
// After a sweep of one graph: the spread counters (6 groups of kSpread
// slots from `spread`) collapsed into tot (tot[0] = S_in, tot[1..4] = sum a^2
// limbs, tot[5] = moves), Q from them and the stop rule into *ps; *err is
// the sweep's capacity-overflow flag.
struct SweepTail {
    unsigned long long *spread;
    unsigned long long *tot;
    const uint32_t *err;
    float m, min_gain;
    uint32_t max_iters;
    PassState *ps;
    PassState *ps_host;  // device address of the host's mapped copy
};
void launch_collapse_decide(hipStream_t s, const SweepTail &t, int S);
// Batch pass init, after the k_v sums: each active graph's m (a graph
// without weight finishes with LV_ERR_RANGE) and its entries.  acc: per-graph
// exact k totals; voff: vertex offsets, row: the pass's row pointers.
void launch_batch_m(hipStream_t s, uint32_t G, const unsigned long long *acc, int S,
                    const uint32_t *voff, const uint32_t *row, uint32_t *gstate, float *gm,
                    GraphState *gst, int range_status);
// After the pass-init modularity: act = Q0, first sweep next.
void launch_batch_init(hipStream_t s, uint32_t G, const unsigned long long *gcnt, int S,
                       const uint32_t *gstate, GraphState *gst);
// After a batch sweep: every active graph's Q and stop rule; graphs that stop
// wait (improved) or finish.  Counts into the summary (its recounted fields
// zero before the sweep) and leaves every graph's counters zero.
void launch_batch_decide(hipStream_t s, uint32_t G, unsigned long long *gcnt,
                         const uint32_t *err_flag, int S, float min_gain, uint32_t max_iters,
                         const uint32_t *voff, uint32_t *gstate, GraphState *gst,
                         BatchSummary *sum);
// After launch_batch_decide: halt the remaining queued sweeps when no graph
// is active; the summary to the host's mapped copy `host`, its recounted
// fields cleared for the next sweep.
void launch_batch_halt(hipStream_t s, BatchSummary *sum, BatchSummary *host);
// After a batch contraction: waiting graphs become active for the next pass.
void launch_batch_reactivate(hipStream_t s, uint32_t G, uint32_t *gstate, GraphState *gst);

// ---- launchers (lv_kernels.hip) ---------------------------------------------
void launch_fill_u32(hipStream_t s, uint32_t *p, uint32_t v, size_t n);
void launch_fill_u64(hipStream_t s, unsigned long long *p, unsigned long long v, size_t n);
void launch_zero3_u32(hipStream_t s, uint32_t *p0, size_t n0, uint32_t *p1, size_t n1,
                      uint32_t *p2, size_t n2);
// whole contraction of a pass graph of at most kTinyN vertices and kTinyE
// entries (single graph, input numbering) in one workgroup; sizes[0..1] <-
// new vertex and entry counts, new_id <- the new ids (n entries)
constexpr uint32_t kTinyN = 1024, kTinyE = 4096, kTinyThreads = 1024;
void launch_agg_tiny(hipStream_t s, const DevView &g, uint32_t n0, uint32_t *init_comm,
                     uint32_t *orow, uint32_t *ocol, float *ow, uint32_t *sizes,
                     uint32_t *new_id);
// contraction of a pass graph of at most kFewN vertices into at most kFewNN
// communities (single graph, input numbering) through a dense matrix of
// exact sums (mat: kFewNN^2 sums and a kFewNN^2-bit presence map behind them,
// new_id: n + 1); sizes[0] <- communities, and
// when they are at most kFewNN, sizes[1] <- entries and the contracted CSR
// and init_comm written; otherwise nothing else is written
constexpr uint32_t kFewN = 65536, kFewNN = 64, kFewRows = 32;
void launch_agg_few(hipStream_t s, const DevView &g, uint32_t *new_id, unsigned long long *mat,
                    uint32_t n0, uint32_t *init_comm, uint32_t *orow, uint32_t *ocol, float *ow,
                    uint32_t *sizes);
void launch_iota_u32(hipStream_t s, uint32_t *p, size_t n);
void launch_collapse(hipStream_t s, const unsigned long long *spread, int groups,
                     unsigned long long *out);

// batch mode
void launch_batch_gid(hipStream_t s, const uint32_t *voff, uint32_t G, uint32_t *gid,
                      uint32_t n);
void launch_batch_identity(hipStream_t s, const uint32_t *gid, const uint32_t *gstate,
                           uint32_t keep, uint32_t *comm, uint32_t *comm_sz, uint32_t n);
void launch_gather_u32(hipStream_t s, const uint32_t *src, const uint32_t *idx, uint32_t *out,
                       uint32_t count);

// degree classification of the current graph into `order`
// locality order (lv_kernels.hip / lv_heavy.hip)
struct RelabelView {
    const uint32_t *row, *col;  // input-numbered CSR of the pass
    const float *w;
    const uint32_t *row2;       // new row pointers (by position)
    uint32_t *col2;
    float *w2;
    const uint32_t *lid, *pmap;
};
void launch_relabel_map(hipStream_t s, uint32_t *order, uint32_t n, uint32_t *lid,
                        uint32_t *pmap, const uint32_t *row, uint32_t *deg);
void launch_relabel_rows(hipStream_t s, const RelabelView &r, uint32_t x0, uint32_t x1,
                         int lanes);
void launch_relabel_heavy(hipStream_t s, const RelabelView &r, const uint32_t *cvtx,
                          const uint32_t *coff, uint32_t nchunks, uint32_t x0);
// out[b] = i for b in [off[i], off[i+1]), i < count
void launch_owner_fill(hipStream_t s, const uint32_t *off, uint32_t count, uint32_t *out);
void launch_relabel_new_id(hipStream_t s, const uint32_t *comm_sz, const uint32_t *lid, uint32_t n,
                           uint32_t *flags, uint32_t *rank, uint32_t *new_id, uint32_t *scan_tmp);
// rows longer than heavy_min (319 <= heavy_min <= kBlockMax) go to the heavy
// class; with gid non-null (batch), vertices of graphs that are not active
// are left out of the classes
void launch_classify_count(hipStream_t s, const uint32_t *row, uint32_t n, uint32_t heavy_min,
                           const uint32_t *gid, const uint32_t *gstate,
                           uint32_t *counts /*kNumClasses*/,
                           unsigned long long *degsum /*kNumClasses*/);
// run: consecutive vertices per thread (a multiple of 16; longer runs keep
// more neighbouring ids together in each class, fewer workgroups)
void launch_classify_scatter(hipStream_t s, const uint32_t *row, uint32_t n, uint32_t heavy_min,
                             const uint32_t *gid, const uint32_t *gstate, uint32_t run,
                             uint32_t *cursor /*kNumClasses*/, uint32_t *order);

// Pinned-order sums: mode (masked) 0 = k_v (+ optional state init),
// 1 = s_v, 2 = s_v while every community is a singleton (pass init: no
// community gathers); `acc` receives the exact fixed-point total (atomic).
// The six class groups of the pinned-order sums (rows of at most 4, 8, 16,
// 32 and 84 entries, and longer rows) in one launch, longest rows first:
// group i covers count[i] vertices of verts[i] with workgroups
// [bend[i+1], bend[i]) (bend[6] = 0; an empty group repeats the bound;
// bend[0] is the grid size).  Vertices per workgroup: 256, 128, 64, 32, 32,
// 8 (kR1GroupVertices).
constexpr int kR1Groups = 6;
constexpr uint32_t kR1GroupVertices[kR1Groups] = {256, 128, 64, 32, 32, 8};
struct R1Groups {
    const uint32_t *verts[kR1Groups];
    int32_t vbase[kR1Groups];  // >= 0: the group's list is the identity run from there
    uint32_t count[kR1Groups];
    uint32_t bend[kR1Groups];
    // masked sums: sq_blocks workgroups behind the groups add the sum of
    // round32(a_c)^2 over the communities sq_verts[0, sq_count) into limbs
    // (4 arrays of kSpread slots, or G in a batch)
    const uint32_t *sq_verts;
    int32_t sq_vbase;
    uint32_t sq_count, sq_blocks;
    unsigned long long *limbs;
};
void launch_r1_groups(hipStream_t s, const DevView &g, const R1Groups &L, int masked,
                      int init_state, unsigned long long *acc);
// sum of round32(a_c)^2 over the communities in `verts`; limbs: 4 arrays of
// kSpread (single graph) or G (batch) slots
uint32_t sumsq_blocks(uint32_t count);  // workgroups of the a_c^2 sum over count communities

// classes 1-10 (the heavy class goes through launch_heavy_move)
// Launch shapes the tuning selects among compiled variants of the move
// kernels (no result depends on them).
struct MoveShape {
    uint32_t wave_vertices;  // bucket-6 kernel: vertices (one wave each) per workgroup: 1, 2, 4
    uint32_t block_threads;  // rows of 2049..8192 entries: 512 or 1024 threads per vertex
    uint32_t class8_slots;   // rows of 513..1024 entries: LDS table of 1280 or 2048 slots
    // buckets 1-5: a class of at least this many vertices takes the narrow
    // kernel (fewer lanes per vertex, more vertices per wave: faster once the
    // class saturates the device), a smaller one the reference's lane count
    // (shorter chain per vertex: faster while the device is not full)
    uint32_t narrow_min;
};
// default of MoveShape::narrow_min
constexpr uint32_t kNarrowMinVertices = 1u << 19;
// A class's move kernel; with g.cu_count set, catch-up workgroups for the
// previous step's moves go in front of the class's workgroups.
void launch_move_class(hipStream_t s, const DevView &g, int cls, const uint32_t *verts,
                       uint32_t count, const MoveShape &shape);
// The moves of g.cu_verts[0, g.cu_count) applied to the other copy on their own.
void launch_catchup(hipStream_t s, const DevView &g);
// One copy: the step's decisions (new_comm) of verts[0, count) applied in place.
void launch_apply(hipStream_t s, const DevView &g, const uint32_t *verts, uint32_t count,
                  unsigned long long *moved);
void launch_heavy_caps(hipStream_t s, const uint32_t *row, const uint32_t *verts,
                       uint32_t count, unsigned long long *caps);

// ---- launchers (lv_heavy.hip) -----------------------------------------------
void launch_heavy_chunks(hipStream_t s, const uint32_t *row, const uint32_t *verts,
                         uint32_t count, uint32_t *nchunks);
void launch_fill_slots(hipStream_t s, unsigned long long *tab, size_t cap);
void launch_hp_nbuck(hipStream_t s, const uint32_t *row, const uint32_t *verts, uint32_t count,
                     uint32_t target, uint32_t *nbuck);
void launch_heavy_move(hipStream_t s, const DevView &g, const HeavySplit &h);
void launch_heavy_r1(hipStream_t s, const DevView &g, const HeavySplit &h, float *stage,
                     int masked, int init_state, unsigned long long *acc);

// exclusive scan with total at out[n]; input functor chosen by `kind`:
//   0: in[i]   1: in[i] != 0
void launch_scan_u32(hipStream_t s, const uint32_t *in, int kind, uint32_t n, uint32_t *out,
                     uint32_t *tmp /* >= n/4096 + 2 */);
void launch_scan_u64(hipStream_t s, const unsigned long long *in, uint32_t n,
                     unsigned long long *out, unsigned long long *tmp);
size_t scan_tmp_elems(uint32_t n);

// aggregation
struct AggView {
    uint32_t *new_id;       // n+1 (exclusive scan of comm_sz != 0)
    uint32_t *vnid;         // n: new id of each vertex's community, new_id[comm[u]]
    uint32_t *comm_deg;     // n
    uint32_t *epos;         // n+1 (exclusive scan of comm_deg)
    uint32_t *mstart;       // n+1 (exclusive scan of comm_sz)
    uint32_t *mcursor;      // n
    uint32_t *members;      // n
    uint32_t *cnt;          // n'+1 distinct counts by new id, scanned into new row
    uint32_t *tkey;         // ne
    unsigned long long *tval;   // ne
    uint32_t *skey;         // ne (radix scratch)
    unsigned long long *sval;   // ne
    unsigned long long *gtab;   // global-table arena (16-byte slots, shared with heavy rows)
    const unsigned long long *goff;  // per large community: table offsets (count+1)
    uint32_t *mdeg;         // n: degree of members[i]
    uint32_t *moff;         // n+1: exclusive scan of mdeg
    uint32_t *cnk;          // per large community: chunk count
    uint32_t *coff;         // per large community: chunk offsets (count+1)
    uint32_t *dcnt;         // per large community: distinct neighbour communities
    uint32_t nchunks;       // total chunks of the large communities
    uint32_t nn;            // new vertex count (new ids are < nn)
    // global-table merge: the flagged large communities (clist) and their
    // chunks (flist), each list led by its length (nfl, ncl on the host)
    const uint32_t *flist, *clist;
    uint32_t nfl, ncl;
    // batch: graph of each old vertex and first new id of each graph (a
    // community's keys lie in its graph's new-id range); nullptr: one graph
    const uint32_t *kgid, *kbase;
    uint32_t *err;
};
// Partitioned merge of the large communities (per large community i < count:
// nbk/bstart/cflag; per bucket b < NB: the rest).
struct AggPart {
    const uint32_t *nbk;     // buckets of each large community (pow2)
    const uint32_t *bstart;  // count+1: first bucket of each
    const uint32_t *bowner;  // NB: large community of each bucket
    uint32_t *bcnt, *bpos, *bcur;  // NB entry counts / NB+1 ranges / NB cursors
    uint32_t *bnd;           // NB: distinct keys of each bucket
    const uint32_t *boff;    // NB+1: exclusive scan of bnd
    uint32_t *cflag;         // count: bucket overflow -> global-table merge
    uint32_t *nflag;         // number of flagged communities
    uint32_t *fkey;          // ne: new id of each entry, entry order
    float *fw;               //     weight of each entry
    uint32_t *pkey;          // ne: the same, bucket order
    float *pw;
    uint32_t nn;             // new vertex count (keys are < nn)
    uint32_t NB;             // total buckets
    uint32_t dmax;           // distinct keys a bucket may hold (<= kAggBucketDistinct)
};
constexpr uint32_t kAggBucketDistinct = 512;  // per bucket; its LDS table has more slots
// target: entries per key-range bucket of a large community (256..512)
void launch_aggp_nb(hipStream_t s, AggView a, const uint32_t *comms, uint32_t count, uint32_t nn,
                    uint32_t target,
                    uint32_t *nbk);
// phase 0 count, 1 scatter, 2 bucket merge, 3 drop counts of flagged communities
void launch_aggp(hipStream_t s, const DevView &g, AggView a, const AggPart &p,
                 const uint32_t *comms, uint32_t count, int phase);
void launch_aggp_caps(hipStream_t s, AggView a, const AggPart &p, const uint32_t *comms,
                      uint32_t count, unsigned long long *caps);
// Where a contraction writes the new CSR: the row of new vertex u (its
// community's new id) starts at rows[map ? map[u] : u] and holds
// rank_rows[u + 1] - rank_rows[u] entries (rank_rows: row pointers by new id),
// its columns are map[new id] (or the new ids without a map).  With a map the
// graph comes out in the next pass's locality order.
struct EmitTarget {
    const uint32_t *rows;
    const uint32_t *rank_rows;
    const uint32_t *map;
    uint32_t *col;
    float *w;
};
void launch_aggp_emit(hipStream_t s, const DevView &g, AggView a, const AggPart &p,
                      const uint32_t *comms, const EmitTarget &t);
void launch_agg_prep(hipStream_t s, const DevView &g, AggView a);
void launch_agg_members(hipStream_t s, const DevView &g, AggView a);
void launch_agg_classify_count(hipStream_t s, const DevView &g, AggView a, uint32_t *counts);
void launch_agg_classify_scatter(hipStream_t s, const DevView &g, AggView a, uint32_t *cursor,
                                 uint32_t *order);
// lists of the flagged large communities and of their chunks (counts at
// flist[0] and clist[0], cleared before the launch)
void launch_agg_flag_list(hipStream_t s, const uint32_t *flag, const uint32_t *coff,
                          uint32_t count, uint32_t *flist, uint32_t *clist);
void launch_agg_merge(hipStream_t s, const DevView &g, AggView a, int cls, const uint32_t *comms,
                      uint32_t count, int radix_passes);
// degree of every member in member order (mdeg) -- scanned into moff
void launch_agg_mdeg(hipStream_t s, const DevView &g, AggView a);
void launch_agg_large_prep(hipStream_t s, const DevView &g, AggView a, const uint32_t *comms,
                           uint32_t count);
void launch_agg_emit(hipStream_t s, const DevView &g, AggView a, const uint32_t *comms,
                     uint32_t count, const uint32_t *only, const EmitTarget &t);
// vnid[u] = new_id[comm[u]] for u < n: one gather per entry in the merges
void launch_agg_vnid(hipStream_t s, const uint32_t *comm, const uint32_t *new_id, uint32_t n,
                     uint32_t *vnid);
// init_comm <- new_id[comm[init_comm]], through map (new id -> vertex of the
// next pass) when non-null
void launch_agg_compose(hipStream_t s, const uint32_t *comm, const uint32_t *new_id,
                        const uint32_t *map, uint32_t *init_comm, uint32_t n0);

}  // namespace lv
