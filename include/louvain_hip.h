/*
 * louvain_hip.h -- C ABI of the MI355X Louvain engine (liblouvain_hip.so).
 *
 * Plain C: pointers and sizes only, integer status codes, no stdout, no
 * exit().  One lv_graph per device-resident graph; each graph owns its own
 * HIP stream and device buffers.  Calls on different graphs may run from
 * different host threads; calls on one graph must not overlap.
 *
 * The entry points replace the reference's operator API and its I/O:
 *   lv_mtx_read                -> parseGraph               (utils.cu:168-229)
 *   lv_graph_create            -> initDevGraph + copyToDevice + initCommAggr
 *                                 + setInitComm            (utils.cu:70-98, :66-68)
 *   lv_modularity_optimization -> modularityOptimization   (modularity.cuh:6,
 *                                                           modularity.cu:438-520)
 *   lv_aggregation             -> aggregation              (aggregation.cuh:6,
 *                                                           aggregation.cu:387-462)
 *   lv_run                     -> the pass loop of main    (main.cu:41-52)
 *   lv_get_assignment          -> the init_comm download of printCommunities
 *                                 (utils.cu:270-275)
 *   lv_graph_destroy           -> deleteDevGraph + deleteAggrData (utils.cu:26-51)
 * The finer-grained lv_pass_init / lv_step / lv_sweep / lv_modularity expose the
 * pieces of modularityOptimization (init :455-463, one bucket step :468-499,
 * one sweep :467-500, computeModularity :428-436) for step-by-step parity
 * checks.
 */
#ifndef LOUVAIN_HIP_H
#define LOUVAIN_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
enum {
    LV_OK = 0,
    LV_ERR_IO = 1,         /* file cannot be opened (reference: "Error opening file") */
    LV_ERR_FORMAT = 2,     /* malformed .mtx */
    LV_ERR_RANGE = 3,      /* weights not representable in the exact fixed-point range */
    LV_ERR_ARG = 4,        /* bad argument (null pointer, id out of range, size limits) */
    LV_ERR_NOMEM = 5,      /* host or device allocation failed */
    LV_ERR_STATE = 6,      /* call out of order (e.g. lv_step before lv_pass_init) */
    LV_ERR_HIP = 7,        /* HIP runtime error */
    LV_ERR_NO_DEVICE = 8,  /* no usable GPU */
    LV_ERR_OVERFLOW = 9,   /* internal capacity exceeded (reported, never silent) */
    LV_ERR_LIMIT = 10      /* graph larger than a fixed buffer allows */
};

typedef struct lv_graph lv_graph;

/* Host CSR as produced by lv_mtx_read: symmetric rows in file order, each
 * self-loop stored once, |weight|.  row has n+1 entries. */
typedef struct lv_host_csr {
    uint32_t n;
    uint64_t nnz;  /* edge lines in the file */
    uint64_t ne;   /* CSR entries = row[n] */
    uint32_t *row;
    uint32_t *col;
    float *w;
} lv_host_csr;

typedef struct lv_result {
    float q_final;           /* the modularity the reference prints (modularity.cu:504) */
    uint32_t passes;         /* passes that improved and were aggregated */
    uint32_t n_final;        /* vertices of the final graph (= community count) */
    uint32_t total_iters;    /* sweeps over all passes */
    uint64_t sum_edges;      /* sum over passes of the pass graph's CSR entries */
    uint64_t sum_edge_visits;/* sum over passes of (iters+1) * entries */
    double alg_ms;           /* device time of the whole run (hipEvents) */
    double bytes_alg;        /* algorithmic HBM bytes of the run (see DESIGN) */
    int status;
} lv_result;

/* Per-kernel-family device time, collected with hipEvents on the graph's
 * stream while profiling is enabled. */
enum {
    LV_K_CLASSIFY = 0,
    LV_K_R1 = 1,          /* k_v / s_v pinned-order sums */
    LV_K_SUMSQ = 2,       /* sum of a_c^2 (runs inside the LV_K_R1 masked-sum launch) */
    LV_K_MOVE_SMALL = 3,  /* buckets 1-5 */
    LV_K_MOVE_WAVE = 4,   /* bucket 6 */
    LV_K_MOVE_BLOCK = 5,  /* bucket 7, LDS hash */
    LV_K_MOVE_GLOBAL = 6, /* bucket 7, global hash */
    LV_K_APPLY = 7,
    LV_K_AGGREGATE = 8,
    LV_K_RELABEL = 9,     /* locality order of a pass graph */
    LV_K_COUNT = 10
};

typedef struct lv_kernel_stats {
    double ms[LV_K_COUNT];       /* summed launch durations */
    uint64_t launches[LV_K_COUNT];
    double bytes[LV_K_COUNT];    /* algorithmic bytes attributed to the family */
} lv_kernel_stats;

const char *lv_status_string(int status);
int lv_device_count(int *count);

/* ---- input ---------------------------------------------------------------- */
/* Reads a Matrix Market file (or a binary CSR written by lv_csr_write,
 * recognised by its magic) into a host CSR.  Multi-threaded; same result as
 * the reference parser.  Errors: LV_ERR_IO, LV_ERR_FORMAT. */
int lv_mtx_read(const char *path, lv_host_csr *out);
void lv_host_csr_free(lv_host_csr *csr);
/* Binary CSR cache of a parsed graph (no counterpart in the reference, which
 * re-parses the text every run); lv_mtx_read loads it. */
int lv_csr_write(const char *path, const lv_host_csr *csr);

/* ---- graph lifetime -------------------------------------------------------- */
/* Uploads the CSR (row: n+1 offsets, col/w: row[n] entries) to `device`.
 * Limits: n < 2^31, row[n] < 2^32.  Weights must be finite and >= 0. */
int lv_graph_create(int device, uint32_t n, const uint32_t *row, const uint32_t *col,
                    const float *w, lv_graph **out);
int lv_graph_destroy(lv_graph *g);
/* Restore the graph as uploaded (from the handle's pinned host copy of the
 * input) and the identity assignment, so the same input can be processed
 * again.  No counterpart in the reference, which runs one graph per
 * process. */
int lv_graph_reset(lv_graph *g);
/* Locality order of each pass graph on the device: vertices grouped by degree
 * class so the state of high-degree vertices, which most entries point at,
 * stays cache-resident.  mode -1: automatic (pass graphs of at least 2^20
 * vertices), 0: never, 1: every pass.  Results are identical in every mode;
 * ids are compared in the input numbering.  No counterpart in the reference
 * (a performance knob); has no effect on batches. */
int lv_graph_set_locality(lv_graph *g, int mode);

/* Launch-shape parameters.  None of them changes a result: the seven semantic
 * step sets, lane orders and sums are fixed; these only choose which kernel
 * serves a row and how work is cut.  No counterpart in the reference. */
typedef struct lv_tuning {
    uint32_t heavy_min_degree;      /* rows longer than this take the split heavy path
                                       (319..8192, default 4096; shorter bucket-7 rows: one
                                       workgroup with an LDS table) */
    uint32_t bucket_entries;        /* target entries per LDS bucket of a heavy row
                                       (power of two, 64..1024) */
    int32_t locality;               /* -1 automatic, 0 never, 1 every pass */
    uint32_t locality_min_vertices; /* automatic mode: pass graphs from this many vertices */
    uint32_t wave_vertices;         /* bucket-6 rows (85..319 entries): vertices, one wave
                                       each, per workgroup: 1, 2 (default) or 4 */
    uint32_t block_threads;         /* bucket-7 rows of 2049..8192 entries: threads of the
                                       vertex's workgroup, 512 or 1024 (default) */
    uint32_t class8_slots;          /* bucket-7 rows of 513..1024 entries: LDS table slots,
                                       1280 (default) or 2048 */
    uint32_t agg_bucket_entries;    /* contraction of large communities: target entries per
                                       key-range bucket, 256..512 (default 384) */
    uint32_t sweeps_per_launch;     /* sweeps launched per wait of the host in a pass loop
                                       decided on the device: 1 (default), 2 or 4 */
} lv_tuning;
int lv_tuning_default(lv_tuning *t);
int lv_graph_set_tuning(lv_graph *g, const lv_tuning *t);  /* LV_ERR_ARG if out of range */
int lv_graph_get_tuning(lv_graph *g, lv_tuning *t);
/* Autotuner: runs the resident input (lv_graph_reset before each run) under
 * each candidate and keeps the fastest (best of two runs each) on g,
 * returning it with its time.  With candidates NULL it searches every field
 * above one at a time (each over its listed values, the others held at the
 * best found so far; locality only for graphs of at least 2^18 vertices).
 * g is left reset. */
int lv_autotune(lv_graph *g, float min_gain, const lv_tuning *candidates, uint32_t count,
                lv_tuning *best, double *best_ms);

/* Stream: by default every call on a graph runs on the graph's own HIP
 * stream.  lv_graph_set_stream makes later calls enqueue on the caller's
 * stream instead (a hipStream_t passed as void*; NULL returns to the graph's
 * own), so the engine's work is ordered after the caller's earlier work on
 * that stream and before its later work, as the reference's operators are on
 * the default stream (main.cu:43-48).  Kernel chains that run beside the main
 * one fork from and join back into that stream by events.  The engine still
 * synchronises the stream where the host takes a decision (the stop rule
 * after each sweep, the sizes of each aggregation), so the stream must not be
 * under graph capture.  Work queued on the previous stream completes before
 * the call returns.  No counterpart in the reference. */
int lv_graph_set_stream(lv_graph *g, void *stream);

/* ---- the reference operator API ------------------------------------------ */
/* One Louvain pass of local moves (modularityOptimization, modularity.cuh:6).
 * *improved is the reference's return value: false when the first sweep
 * gains less than min_gain.  *q_before is the modularity of the pass graph's
 * singleton partition, computed before the first sweep (when the pass did
 * not improve, the value the reference prints, modularity.cu:504); *q_after
 * the modularity after the last sweep.  Any output may be NULL. */
int lv_modularity_optimization(lv_graph *g, float min_gain, int *improved, float *q_before,
                               float *q_after);
/* Sweeps, q_before and q_after of the last lv_modularity_optimization (or of
 * the last pass of lv_run).  Any output may be NULL. */
int lv_pass_stats(lv_graph *g, uint32_t *sweeps, float *q_before, float *q_after);
/* Contract communities into the next pass's graph; composes the assignment. */
int lv_aggregation(lv_graph *g);
/* The full pass loop: modularity optimisation then aggregation until a pass
 * fails its first sweep. */
int lv_run(lv_graph *g, float min_gain, lv_result *res);

/* The whole run (lv_graph_create + lv_run + lv_get_assignment) with no device
 * allocation.  No counterpart in the reference: for callers that cannot
 * allocate device memory (code under HIP graph capture of the caller's own
 * work, tight serving loops, harnesses that forbid hipMalloc).  Every device
 * buffer is carved from a fixed 128 MiB array that the library holds on each
 * GPU it runs on until the process ends; graphs whose buffers do not fit it
 * (about 1.4M entries with 64k vertices, depending on shape) return
 * LV_ERR_LIMIT.  The graph is the host CSR of lv_graph_create (checked the
 * same way); assignment receives n original-vertex -> community ids; res the
 * figures of lv_run (alg_ms from the host clock).  Runs on `stream` (a
 * hipStream_t; NULL: the null stream), which it synchronises before it
 * returns.  Calls on one GPU share the array and run one at a time (they are
 * serialised inside the library); the array keeps the last call's contents,
 * and each call clears what it needs empty. */
int lv_run_noalloc(int device, uint32_t n, const uint32_t *row, const uint32_t *col,
                   const float *w, float min_gain, void *stream, uint32_t *assignment,
                   lv_result *res);

/* ---- results and state --------------------------------------------------- */
/* original vertex -> vertex of the current graph (length n0) */
int lv_get_assignment(lv_graph *g, uint32_t *assignment);
int lv_graph_size(lv_graph *g, uint32_t *n, uint64_t *ne, uint32_t *n0);
int lv_get_csr(lv_graph *g, uint32_t *row, uint32_t *col, float *w);
/* k (f32), comm, comm_sz, a (exact fixed point a_c * 2^S); any may be NULL */
int lv_get_state(lv_graph *g, float *k, uint32_t *comm, uint32_t *comm_sz, int64_t *a);
int lv_get_scale(lv_graph *g, int *scale_bits, float *m);

/* ---- stepwise pieces of modularityOptimization (parity hooks) ------------- */
int lv_pass_init(lv_graph *g, float *q0);
int lv_step(lv_graph *g, int bucket, uint64_t *moved);
int lv_sweep(lv_graph *g, float *q, uint64_t *moved);
int lv_modularity(lv_graph *g, float *q);

/* ---- batches of independent graphs --------------------------------------- */
/* Many graphs in one segmented CSR: graph g owns vertices
 * [graph_off[g], graph_off[g+1]) (64-bit segment offsets), row/col/w index
 * the concatenation with global vertex ids, and every entry stays inside its
 * graph.  Each graph's result is bit-identical to an lv_run of that graph
 * alone.  No counterpart in the reference, which processes one graph per
 * process (main.cu). */
typedef struct lv_batch_graph_result {
    float q_final;
    uint32_t passes;
    uint32_t n_final;
    uint32_t total_iters;
    uint64_t sum_edges;
    int status;           /* LV_OK, or LV_ERR_RANGE for a graph without weight */
} lv_batch_graph_result;

int lv_batch_create(int device, uint32_t ngraphs, const uint64_t *graph_off, const uint32_t *row,
                    const uint32_t *col, const float *w, lv_graph **out);
int lv_batch_run(lv_graph *batch, float min_gain, lv_batch_graph_result *per_graph,
                 lv_result *total);
/* per original vertex: its community as a 0-based id local to its graph */
int lv_batch_get_assignment(lv_graph *batch, uint32_t *local_assignment);

/* ---- serving --------------------------------------------------------------- */
/* A request loop on one GPU: graphs are admitted as worker slots and HBM
 * budget free up, run with device-resident state on their own streams (so
 * they overlap on the device), and their results are published as they
 * finish.  Each result equals lv_run of that graph alone.  No counterpart in
 * the reference (one graph per process, main.cu). */
typedef struct lv_server lv_server;

typedef struct lv_request_result {
    uint64_t id;
    int status;
    float q_final;
    uint32_t passes;
    uint32_t n_final;
    uint32_t total_iters;
    uint64_t sum_edges;
    double t_submit_ms;  /* server clock: submission, start on a slot, finish */
    double t_start_ms;
    double t_end_ms;
} lv_request_result;

/* device bytes a graph of n vertices / ne entries may occupy while running
 * (every buffer of its handle, the on-demand arenas at their worst case:
 * 84.5 B per entry, 176 B per vertex, 4 MiB) */
uint64_t lv_graph_footprint(uint32_t n, uint64_t ne);
int lv_server_create(int device, uint32_t slots, uint64_t hbm_budget_bytes, float min_gain,
                     int keep_assignments, lv_server **out);
/* copies the CSR; LV_ERR_NOMEM if the graph can never fit the budget */
int lv_server_submit(lv_server *server, uint64_t id, uint32_t n, const uint32_t *row,
                     const uint32_t *col, const float *w);
/* non-blocking: moves up to max finished results into out, *count set */
int lv_server_poll(lv_server *server, lv_request_result *out, uint32_t max, uint32_t *count);
/* with keep_assignments: the finished request's assignment (once) */
int lv_server_get_assignment(lv_server *server, uint64_t id, uint32_t *assignment, uint32_t n);
int lv_server_drain(lv_server *server);  /* waits until every submitted request finished */
int lv_server_destroy(lv_server *server);

/* ---- profiling ------------------------------------------------------------ */
int lv_profile_enable(lv_graph *g, int on);
/* Families that get event pairs while profiling is on (bit f = family f;
 * default all).  An event record costs a few microseconds, so a timed run
 * profiles only the family it reports. */
int lv_profile_select(lv_graph *g, uint32_t family_mask);
int lv_profile_get(lv_graph *g, lv_kernel_stats *out); /* synchronises */
int lv_profile_reset(lv_graph *g);

/* ---- synthetic inputs (seeded, deterministic) ----------------------------- */
/* FEM-band analogue of bcsstk16: nodes of 6 degrees of freedom, dense 6x6
 * blocks, node bandwidth `band`, ~`couple` coupled nodes per node,
 * w = |N(0,1)| + 0.01, self-loops kept. */
int lv_gen_fem(uint32_t nodes, uint32_t band, uint32_t couple, uint64_t seed, lv_host_csr *out);
/* Chung-Lu power-law degrees over a planted partition (unit weights). */
int lv_gen_chung_lu(uint32_t n, double mean_deg, double exponent, uint32_t max_deg,
                    uint32_t groups, double p_ratio, uint64_t seed, lv_host_csr *out);
/* R-MAT (Graph500 style): 2^scale vertices, edge_factor*2^scale edges,
 * symmetrised, deduplicated, vertex ids permuted, w ~ U(wlo, whi) (wlo==whi:
 * unit), self-loops kept once. */
int lv_gen_rmat(uint32_t scale, uint32_t edge_factor, double a, double b, double c,
                float wlo, float whi, uint64_t seed, lv_host_csr *out);
/* Stochastic block model: n vertices in `blocks` blocks, expected degree
 * mean_deg, p_in/p_out = p_ratio, unit weights. */
int lv_gen_sbm(uint32_t n, uint32_t blocks, double mean_deg, double p_ratio, uint64_t seed,
               lv_host_csr *out);

#ifdef __cplusplus
}
#endif
#endif
