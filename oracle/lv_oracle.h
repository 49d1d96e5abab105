/*
 * lv_oracle.h -- CPU restatement of the reference Louvain path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (the HIP library, the
 * CLI) includes, links or calls this code.  It is the checker that the GPU
 * engine is compared against, and the single-core CPU baseline timed beside
 * the GPU in bench.py.
 *
 * The semantics implemented here are the canonical form of the reference's
 * algorithm (modularity.cu, aggregation.cu, main.cu):
 *   - synchronous-snapshot bucket steps with the reference's 7 degree cuts;
 *   - per-lane first-maximum with the reference's lane strides, the bucket 5
 *     round-prefix sums, and the buckets 6-7 "last visited edge" destination;
 *   - every order-dependent weight sum computed exactly in fixed point at a
 *     per-graph power-of-two scale 2^S and rounded once to FP32;
 *   - k_v and s_v in the reference's pinned 32-lane strided + shuffle-tree
 *     FP32 order;
 *   - contracted rows in ascending new-neighbour-id order.
 */
#ifndef LV_ORACLE_H
#define LV_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    OR_OK = 0,
    OR_ERR_IO = 1,        /* file cannot be opened ("Error opening file") */
    OR_ERR_FORMAT = 2,    /* malformed .mtx */
    OR_ERR_RANGE = 3,     /* weights outside the exact fixed-point range */
    OR_ERR_ARG = 4,
    OR_ERR_NOMEM = 5,
    OR_ERR_STATE = 6
};

typedef struct or_graph {
    uint32_t n;        /* vertices */
    uint64_t nnz;      /* lines in the file */
    uint64_t *row;     /* n+1 */
    uint32_t *col;     /* row[n] */
    float *w;          /* row[n] */
} or_graph;

typedef struct or_result {
    float q_final;        /* the value the reference prints (modularity.cu:504) */
    uint32_t passes;      /* passes that improved and were aggregated */
    uint32_t n_final;     /* vertices of the final graph = community count */
    uint32_t total_iters; /* sweeps over all passes */
    uint64_t sum_edges;   /* sum over passes of CSR entries of that pass's graph */
    uint64_t sum_edge_visits; /* sum over passes of (iters+1)*E */
    int status;
} or_result;

typedef struct or_state or_state;

/* .mtx reader with the reference parser's semantics (utils.cu:132-229). */
int or_parse_mtx(const char *path, or_graph *out);
int or_parse_mtx_string(const char *text, or_graph *out);
void or_graph_free(or_graph *g);

/* fixed-point scale: smallest S >= 0 with w*2^S integral for all w. */
int or_scale_bits(uint64_t ne, const float *w, int *s_out);

/* reference pinned FP32 32-lane sums (modularity.cu:27-49, 400-426) */
float or_r1_sum(const float *w, uint32_t deg);
int or_bucket_of(uint32_t deg);

or_state *or_create(uint32_t n, const uint64_t *row, const uint32_t *col,
                    const float *w, int *err);
void or_destroy(or_state *st);

/* Host threads for pass init, bucket steps and modularity (default 1).  The
 * results are the same for every thread count (decisions read only the step
 * snapshot, every total is an exact integer sum); used to time the CPU
 * baseline on all cores. */
int or_set_threads(or_state *st, int nthreads);

int or_pass_init(or_state *st, float *q0);
int or_step(or_state *st, int bucket, uint64_t *moved);
float or_modularity(or_state *st);
int or_sweep(or_state *st, float *q, uint64_t *moved);
int or_modularity_optimization(or_state *st, float min_gain, int *improved,
                               float *q_print, uint32_t *iters);
int or_aggregation(or_state *st);
int or_run(or_state *st, float min_gain, or_result *res);

/* accessors (copy out) */
uint32_t or_n(const or_state *st);
uint64_t or_ne(const or_state *st);
uint32_t or_n0(const or_state *st);
int or_scale(const or_state *st);
float or_m(const or_state *st);
/* modularity at the last pass's init and after its last sweep */
void or_last_pass_q(const or_state *st, float *q_before, float *q_after);
void or_get_csr(const or_state *st, uint64_t *row, uint32_t *col, float *w);
void or_get_k(const or_state *st, float *k);
void or_get_comm(const or_state *st, uint32_t *comm);
void or_get_comm_sz(const or_state *st, uint32_t *comm_sz);
void or_get_a(const or_state *st, int64_t *a);
void or_get_init_comm(const or_state *st, uint32_t *init_comm);

/* FP64 modularity of an assignment on a graph (R6 check). */
double or_q64(uint32_t n, const uint64_t *row, const uint32_t *col,
              const float *w, const uint32_t *assign);

#ifdef __cplusplus
}
#endif
#endif
