/*
 * lv_oracle.c -- CPU restatement of the reference Louvain path.
 *
 * TEST INFRASTRUCTURE ONLY: used by tests/, by the smoke check and by the
 * CPU-baseline leg of bench.py.  The product never links it.
 *
 * Build: gcc -O2 -std=c11 -ffp-contract=off -fno-fast-math (see Makefile).
 * FP32 arithmetic is plain IEEE single precision (SSE, FLT_EVAL_METHOD 0),
 * no FMA contraction, correctly rounded division.
 *
 * Each function cites the reference code it follows.  Where the reference is
 * nondeterministic (FP32 atomics, racy orders) the canonical resolution is
 * stated at the function.
 */
#define _POSIX_C_SOURCE 200809L
#include "lv_oracle.h"

#include <ctype.h>
#include <errno.h>
#include <math.h>
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

typedef __int128 i128;

/* Reference bucket cut points (modularity.cu:439) and virtual lane counts
 * = blockDim.x of the kernel that serves each bucket (modularity.cu:443-450). */
static const uint32_t k_cuts[8] = {0, 4, 8, 16, 32, 84, 319, 0x7fffffffu};
static const uint32_t k_lanes[7] = {4, 8, 16, 32, 32, 128, 128};

#define OR_FIXED_LIMIT ((uint64_t)1 << 61)

/* per-vertex community accumulators of one decider (a generation stamp
 * marks which entries belong to the vertex being decided) */
typedef struct {
    int64_t *acc;
    uint32_t *stamp;
    uint32_t gen;
} or_scratch;

struct or_state {
    /* current graph */
    uint32_t n;
    uint64_t *row;
    uint32_t *col;
    float *w;
    /* original graph size and composed assignment (Graph.init_comm) */
    uint32_t n0;
    uint32_t *init_comm;
    /* fixed-point scale 2^S */
    int S;
    float scale;     /* 2^S */
    float inv_scale; /* 2^-S */
    /* pass state (Graph.neigh_w, all_w, comm, comm_sz, comm_w, new_comm) */
    int pass_ready;
    float *k;
    int64_t *K; /* k_v * 2^S */
    float m;
    uint32_t *comm;
    uint32_t *comm_sz;
    uint32_t *new_comm;
    int64_t *A; /* a_c * 2^S, exact */
    /* per-vertex scratch for the community accumulators */
    int64_t *acc;
    uint32_t *stamp;
    uint32_t stamp_gen;
    /* modularity before the first and after the last sweep of the last pass */
    float q_before, q_after;
    /* host threads of pass init, bucket steps and modularity (1: the plain
     * sequential loops); per-thread accumulator scratch beyond the first */
    int nthreads;
    or_scratch *tscr;
};

/* ------------------------------------------------------------------------ */
/* .mtx parsing: utils.cu:132-229                                            */
/* ------------------------------------------------------------------------ */

typedef struct {
    const char *p, *end;
} or_text;

static int next_line(or_text *t, const char **ls, const char **le) {
    if (t->p >= t->end) return 0;
    const char *s = t->p;
    const char *e = memchr(s, '\n', (size_t)(t->end - s));
    if (!e) e = t->end;
    *ls = s;
    *le = e;
    t->p = (e < t->end) ? e + 1 : e;
    return 1;
}

/* Parse one unsigned decimal token like `istream >> unsigned` does for the
 * well-formed inputs the reference reads.  Returns pointer past the token or
 * NULL. */
static const char *parse_uint(const char *s, const char *e, uint64_t *v) {
    while (s < e && (*s == ' ' || *s == '\t' || *s == '\r')) s++;
    if (s >= e || !isdigit((unsigned char)*s)) return NULL;
    uint64_t x = 0;
    while (s < e && isdigit((unsigned char)*s)) {
        x = x * 10 + (uint64_t)(*s - '0');
        if (x > 0xffffffffull) return NULL;
        s++;
    }
    *v = x;
    return s;
}

/* Optional weight token: `ss >> weight` (utils.cu:192).  If absent the
 * variable keeps its previous value; an unparseable token gives 0 (C++11
 * num_get).  strtof is correctly rounded like libstdc++'s float extraction. */
static void parse_weight(const char *s, const char *e, float *w) {
    while (s < e && (*s == ' ' || *s == '\t' || *s == '\r')) s++;
    if (s >= e) return; /* token absent: carry over */
    char buf[128];
    size_t len = 0;
    while (s < e && !isspace((unsigned char)*s) && len < sizeof(buf) - 1) buf[len++] = *s++;
    buf[len] = 0;
    char *endp = NULL;
    float x = strtof(buf, &endp);
    if (endp == buf) x = 0.0f;
    *w = x;
}

static int parse_edge_line(const char *ls, const char *le, uint64_t *a, uint64_t *b, float *w) {
    const char *s = parse_uint(ls, le, a);
    if (!s) return 0;
    s = parse_uint(s, le, b);
    if (!s) return 0;
    if (s < le && !(*s == ' ' || *s == '\t' || *s == '\r')) return 0;
    parse_weight(s, le, w);
    return 1;
}

typedef struct {
    uint32_t src, dst;
    float w;
} or_trip;

int or_parse_mtx_string(const char *text, or_graph *out) {
    memset(out, 0, sizeof(*out));
    or_text t = {text, text + strlen(text)};
    const char *ls, *le;
    /* header: skip lines beginning with '%' (utils.cu:135-137), then
     * "rows cols nnz"; the reference keeps cols as the vertex count (:140) */
    for (;;) {
        if (!next_line(&t, &ls, &le)) return OR_ERR_FORMAT;
        if (ls < le && *ls == '%') continue;
        break;
    }
    uint64_t rows, cols, nnz;
    const char *s = parse_uint(ls, le, &rows);
    if (!s) return OR_ERR_FORMAT;
    s = parse_uint(s, le, &cols);
    if (!s) return OR_ERR_FORMAT;
    s = parse_uint(s, le, &nnz);
    if (!s) return OR_ERR_FORMAT;
    (void)rows;
    uint32_t n = (uint32_t)cols;

    or_trip *tr = (or_trip *)malloc((nnz ? nnz : 1) * sizeof(or_trip));
    if (!tr) return OR_ERR_NOMEM;
    float wcur = 1.0f; /* utils.cu:187 */
    int zero_based = 0;
    for (uint64_t i = 0; i < nnz; i++) {
        uint64_t a, b;
        if (!next_line(&t, &ls, &le) || !parse_edge_line(ls, le, &a, &b, &wcur)) {
            free(tr);
            return OR_ERR_FORMAT;
        }
        if (a == 0 || b == 0) zero_based = 1; /* isZeroStartingIndex, utils.cu:143-165 */
        tr[i].src = (uint32_t)a;
        tr[i].dst = (uint32_t)b;
        tr[i].w = fabsf(wcur); /* utils.cu:193 */
    }
    /* degree count; self-loop stored once (utils.cu:200-211) */
    uint64_t *row = (uint64_t *)calloc((size_t)n + 1, sizeof(uint64_t));
    if (!row) {
        free(tr);
        return OR_ERR_NOMEM;
    }
    for (uint64_t i = 0; i < nnz; i++) {
        if (!zero_based) {
            if (tr[i].src == 0 || tr[i].dst == 0) { /* unreachable */
                free(tr);
                free(row);
                return OR_ERR_FORMAT;
            }
            tr[i].src--;
            tr[i].dst--;
        }
        if (tr[i].src >= n || tr[i].dst >= n) {
            free(tr);
            free(row);
            return OR_ERR_FORMAT;
        }
        row[tr[i].src + 1]++;
        if (tr[i].src != tr[i].dst) row[tr[i].dst + 1]++;
    }
    for (uint32_t v = 0; v < n; v++) row[v + 1] += row[v];
    uint64_t ne = row[n];
    uint32_t *col = (uint32_t *)malloc((ne ? ne : 1) * sizeof(uint32_t));
    float *w = (float *)malloc((ne ? ne : 1) * sizeof(float));
    uint64_t *fill = (uint64_t *)malloc(((size_t)n + 1) * sizeof(uint64_t));
    if (!col || !w || !fill) {
        free(tr);
        free(row);
        free(col);
        free(w);
        free(fill);
        return OR_ERR_NOMEM;
    }
    memcpy(fill, row, ((size_t)n + 1) * sizeof(uint64_t));
    /* file order within each row (utils.cu:200-224) */
    for (uint64_t i = 0; i < nnz; i++) {
        uint32_t a = tr[i].src, b = tr[i].dst;
        if (a != b) {
            col[fill[a]] = b;
            w[fill[a]++] = tr[i].w;
            col[fill[b]] = a;
            w[fill[b]++] = tr[i].w;
        } else {
            col[fill[a]] = a;
            w[fill[a]++] = tr[i].w;
        }
    }
    free(fill);
    free(tr);
    out->n = n;
    out->nnz = nnz;
    out->row = row;
    out->col = col;
    out->w = w;
    return OR_OK;
}

int or_parse_mtx(const char *path, or_graph *out) {
    memset(out, 0, sizeof(*out));
    FILE *f = fopen(path, "rb");
    if (!f) return OR_ERR_IO;
    if (fseek(f, 0, SEEK_END) != 0) {
        fclose(f);
        return OR_ERR_IO;
    }
    long sz = ftell(f);
    if (sz < 0) {
        fclose(f);
        return OR_ERR_IO;
    }
    rewind(f);
    char *buf = (char *)malloc((size_t)sz + 1);
    if (!buf) {
        fclose(f);
        return OR_ERR_NOMEM;
    }
    size_t got = fread(buf, 1, (size_t)sz, f);
    fclose(f);
    buf[got] = 0;
    int rc = or_parse_mtx_string(buf, out);
    free(buf);
    return rc;
}

void or_graph_free(or_graph *g) {
    if (!g) return;
    free(g->row);
    free(g->col);
    free(g->w);
    memset(g, 0, sizeof(*g));
}

/* ------------------------------------------------------------------------ */
/* Arithmetic helpers                                                        */
/* ------------------------------------------------------------------------ */

/* Smallest S >= 0 such that every finite weight w satisfies w*2^S integral.
 * Rejects non-finite and negative weights (the reference takes abs() at parse,
 * utils.cu:193, so negative weights never reach the algorithm). */
int or_scale_bits(uint64_t ne, const float *w, int *s_out) {
    int S = 0;
    for (uint64_t j = 0; j < ne; j++) {
        float x = w[j];
        if (!(x >= 0.0f) || isinf(x)) return OR_ERR_RANGE;
        if (x == 0.0f) continue;
        uint32_t bits;
        memcpy(&bits, &x, 4);
        int ef = (int)((bits >> 23) & 0xff);
        uint32_t man = bits & 0x7fffffu;
        int e;
        if (ef == 0) {
            e = -149;
        } else {
            man |= 0x800000u;
            e = ef - 150;
        }
        while ((man & 1u) == 0) {
            man >>= 1;
            e++;
        }
        if (-e > S) S = -e;
    }
    *s_out = S;
    return OR_OK;
}

static float pow2f(int e) { return ldexpf(1.0f, e); }

/* Exact value of a non-negative finite float times 2^bits as a 128-bit
 * integer; the caller guarantees x*2^bits is integral and < 2^127. */
static i128 f32_to_fixed128(float x, int bits) {
    if (x == 0.0f) return 0;
    uint32_t u;
    memcpy(&u, &x, 4);
    int ef = (int)((u >> 23) & 0xff);
    uint32_t man = u & 0x7fffffu;
    int e;
    if (ef == 0) {
        e = -149;
    } else {
        man |= 0x800000u;
        e = ef - 150;
    }
    int sh = e + bits;
    if (sh >= 0) return (i128)man << sh;
    return (i128)(man >> (-sh));
}

/* Round an exact fixed-point value X * 2^-S to FP32 (round to nearest even).
 * int64 -> float conversion is correctly rounded; scaling by a power of two is
 * exact in the normal range. */
static float fix2f(const or_state *st, int64_t x) { return (float)x * st->inv_scale; }

/* k_v in the reference's order (modularity.cu:27-49): lane l < 32 sums
 * entries l, l+32, ... sequentially in FP32, then shuffle-down tree with
 * offsets 16,8,4,2,1; the result is lane 0. */
float or_r1_sum(const float *w, uint32_t deg) {
    float p[32];
    for (uint32_t l = 0; l < 32; l++) {
        float s = 0.0f;
        for (uint32_t j = l; j < deg; j += 32) s += w[j];
        p[l] = s;
    }
    for (uint32_t o = 16; o > 0; o >>= 1)
        for (uint32_t l = 0; l < o; l++) p[l] = p[l] + p[l + o];
    return p[0];
}

/* s_v in the reference's order (modularity.cu:400-426): same lanes and tree,
 * only entries whose neighbour is in v's community are added. */
static float r1_sum_masked(const or_state *st, uint32_t v) {
    float p[32];
    uint64_t b = st->row[v];
    uint32_t deg = (uint32_t)(st->row[v + 1] - b);
    uint32_t cv = st->comm[v];
    for (uint32_t l = 0; l < 32; l++) {
        float s = 0.0f;
        for (uint32_t j = l; j < deg; j += 32)
            if (st->comm[st->col[b + j]] == cv) s += st->w[b + j];
        p[l] = s;
    }
    for (uint32_t o = 16; o > 0; o >>= 1)
        for (uint32_t l = 0; l < o; l++) p[l] = p[l] + p[l + o];
    return p[0];
}

int or_bucket_of(uint32_t deg) {
    for (int b = 1; b <= 7; b++)
        if (k_cuts[b - 1] < deg && deg <= k_cuts[b]) return b;
    return 0; /* deg 0: in no bucket, never moves */
}

/* ------------------------------------------------------------------------ */
/* State                                                                     */
/* ------------------------------------------------------------------------ */

static int set_graph(or_state *st, uint32_t n, uint64_t *row, uint32_t *col, float *w) {
    free(st->row);
    free(st->col);
    free(st->w);
    st->n = n;
    st->row = row;
    st->col = col;
    st->w = w;
    st->pass_ready = 0;
    free(st->k);
    free(st->K);
    free(st->comm);
    free(st->comm_sz);
    free(st->new_comm);
    free(st->A);
    free(st->acc);
    free(st->stamp);
    size_t nn = n ? n : 1;
    st->k = (float *)calloc(nn, sizeof(float));
    st->K = (int64_t *)calloc(nn, sizeof(int64_t));
    st->comm = (uint32_t *)calloc(nn, sizeof(uint32_t));
    st->comm_sz = (uint32_t *)calloc(nn, sizeof(uint32_t));
    st->new_comm = (uint32_t *)calloc(nn, sizeof(uint32_t));
    st->A = (int64_t *)calloc(nn, sizeof(int64_t));
    st->acc = (int64_t *)calloc(nn, sizeof(int64_t));
    st->stamp = (uint32_t *)calloc(nn, sizeof(uint32_t));
    st->stamp_gen = 0;
    if (!st->tscr) {
        st->tscr = (or_scratch *)calloc(1, sizeof(or_scratch));
        st->nthreads = 1;
    }
    if (!st->k || !st->K || !st->comm || !st->comm_sz || !st->new_comm || !st->A || !st->acc ||
        !st->stamp || !st->tscr)
        return OR_ERR_NOMEM;
    return OR_OK;
}

or_state *or_create(uint32_t n, const uint64_t *row, const uint32_t *col, const float *w,
                    int *err) {
    *err = OR_OK;
    or_state *st = (or_state *)calloc(1, sizeof(or_state));
    if (!st) {
        *err = OR_ERR_NOMEM;
        return NULL;
    }
    uint64_t ne = row[n];
    int S = 0;
    int rc = or_scale_bits(ne, w, &S);
    if (rc) {
        free(st);
        *err = rc;
        return NULL;
    }
    /* exact total of w*2^S must stay below 2^61 so that every int64
     * accumulator (hw, comm_e_w, a_c, merged W) is exact with headroom */
    i128 tot = 0;
    for (uint64_t j = 0; j < ne; j++) {
        if (w[j] == 0.0f) continue;
        /* a single weight at or above the limit fails before its cast, which
         * would not be representable in int64 for w*2^S >= 2^63 */
        const float x = ldexpf(w[j], S);
        if (S > 62 || !(x < ldexpf(1.0f, 61))) {
            free(st);
            *err = OR_ERR_RANGE;
            return NULL;
        }
        tot += (i128)(int64_t)x;
        if (tot >= (i128)OR_FIXED_LIMIT) {
            free(st);
            *err = OR_ERR_RANGE;
            return NULL;
        }
    }
    st->S = S;
    st->scale = pow2f(S);
    st->inv_scale = pow2f(-S);
    uint64_t *r = (uint64_t *)malloc(((size_t)n + 1) * sizeof(uint64_t));
    uint32_t *c = (uint32_t *)malloc((ne ? ne : 1) * sizeof(uint32_t));
    float *ww = (float *)malloc((ne ? ne : 1) * sizeof(float));
    st->init_comm = (uint32_t *)malloc((n ? n : 1) * sizeof(uint32_t));
    if (!r || !c || !ww || !st->init_comm) {
        free(r);
        free(c);
        free(ww);
        or_destroy(st);
        *err = OR_ERR_NOMEM;
        return NULL;
    }
    memcpy(r, row, ((size_t)n + 1) * sizeof(uint64_t));
    memcpy(c, col, ne * sizeof(uint32_t));
    memcpy(ww, w, ne * sizeof(float));
    for (uint64_t j = 0; j < ne; j++) {
        if (c[j] >= n) {
            free(r);
            free(c);
            free(ww);
            or_destroy(st);
            *err = OR_ERR_ARG;
            return NULL;
        }
    }
    st->n0 = n;
    for (uint32_t v = 0; v < n; v++) st->init_comm[v] = v; /* setInitComm, utils.cu:66-68 */
    rc = set_graph(st, n, r, c, ww);
    if (rc) {
        or_destroy(st);
        *err = rc;
        return NULL;
    }
    return st;
}

/* thread scratch beyond the state's own (entry 0) */
static void free_threads(or_state *st) {
    if (st->tscr)
        for (int t = 1; t < st->nthreads; t++) {
            free(st->tscr[t].acc);
            free(st->tscr[t].stamp);
        }
    free(st->tscr);
    st->tscr = NULL;
    st->nthreads = 1;
}

void or_destroy(or_state *st) {
    if (!st) return;
    free_threads(st);
    free(st->row);
    free(st->col);
    free(st->w);
    free(st->init_comm);
    free(st->k);
    free(st->K);
    free(st->comm);
    free(st->comm_sz);
    free(st->new_comm);
    free(st->A);
    free(st->acc);
    free(st->stamp);
    free(st);
}

/* ------------------------------------------------------------------------ */
/* Host threads (CPU baseline only): vertex ranges are handed out in chunks  */
/* from a shared counter; every per-vertex result depends only on the step   */
/* snapshot and every total is an exact integer sum, so the results do not   */
/* depend on the number of threads or the order of the chunks.               */
/* ------------------------------------------------------------------------ */

typedef void (*or_range_fn)(void *ctx, int t, uint32_t a, uint32_t b);

typedef struct {
    or_range_fn fn;
    void *ctx;
    uint32_t n, chunk;
    uint32_t next; /* shared */
} or_pool;

typedef struct {
    or_pool *pool;
    int t;
} or_worker;

static void *or_work(void *arg) {
    or_worker *wk = (or_worker *)arg;
    or_pool *p = wk->pool;
    for (;;) {
        uint32_t a = __atomic_fetch_add(&p->next, p->chunk, __ATOMIC_RELAXED);
        if (a >= p->n) break;
        uint32_t b = p->n - a < p->chunk ? p->n : a + p->chunk;
        p->fn(p->ctx, wk->t, a, b);
    }
    return NULL;
}

#define OR_MAX_THREADS 256

static void parallel_for(int T, uint32_t n, or_range_fn fn, void *ctx) {
    if (T <= 1 || n < 8192) {
        fn(ctx, 0, 0, n);
        return;
    }
    or_pool pool = {fn, ctx, n, 1024, 0};
    pthread_t th[OR_MAX_THREADS];
    or_worker wk[OR_MAX_THREADS];
    int started = 1;
    for (int t = 0; t < T; t++) wk[t] = (or_worker){&pool, t};
    for (int t = 1; t < T; t++) {
        if (pthread_create(&th[t], NULL, or_work, &wk[t]) != 0) break;
        started++;
    }
    or_work(&wk[0]);
    for (int t = 1; t < started; t++) pthread_join(th[t], NULL);
}

/* ------------------------------------------------------------------------ */
/* Modularity (modularity.cu:400-436)                                        */
/* ------------------------------------------------------------------------ */

/* Q = S_in/(2m) - sum_c a_c^2 / ((4m) m) in FP32 (modularity.cu:435).
 * s_v: pinned order.  S_in: exact sum of the FP32 s_v (R3).  sum a_c^2: FP32
 * squares (square<float>, modularity.cu:20-25) summed exactly at scale 2^2S
 * and rounded once (R4); the reference's thrust order is unpinned. */
typedef struct {
    const or_state *st;
    int64_t sin[OR_MAX_THREADS];
    i128 sq[OR_MAX_THREADS];
} or_qsum;

static void q_range(void *ctx, int t, uint32_t a, uint32_t b) {
    or_qsum *q = (or_qsum *)ctx;
    const or_state *st = q->st;
    int64_t sin_fixed = 0;
    i128 sq = 0;
    for (uint32_t v = a; v < b; v++) {
        float s = r1_sum_masked(st, v);
        sin_fixed += (int64_t)(s * st->scale);
        float av = fix2f(st, st->A[v]);
        sq += f32_to_fixed128(av * av, 2 * st->S);
    }
    q->sin[t] += sin_fixed;
    q->sq[t] += sq;
}

float or_modularity(or_state *st) {
    or_qsum q;
    memset(&q, 0, sizeof(q));
    q.st = st;
    parallel_for(st->nthreads, st->n, q_range, &q);
    int64_t sin_fixed = 0;
    i128 sq = 0;
    for (int t = 0; t < OR_MAX_THREADS; t++) {
        sin_fixed += q.sin[t];
        sq += q.sq[t];
    }
    float S_in = fix2f(st, sin_fixed);
    /* sq * 2^-2S, correctly rounded once */
    float A2 = ldexpf((float)sq, -2 * st->S);
    float m = st->m;
    return (S_in / (2.0f * m)) - (A2 / ((4.0f * m) * m));
}

/* ------------------------------------------------------------------------ */
/* Pass init (modularity.cu:455-463)                                         */
/* ------------------------------------------------------------------------ */

typedef struct {
    or_state *st;
    int64_t tot[OR_MAX_THREADS];
} or_ksum;

static void k_range(void *ctx, int t, uint32_t a, uint32_t b) {
    or_ksum *q = (or_ksum *)ctx;
    or_state *st = q->st;
    int64_t tot = 0;
    for (uint32_t v = a; v < b; v++) {
        uint64_t r = st->row[v];
        float k = or_r1_sum(st->w + r, (uint32_t)(st->row[v + 1] - r));
        st->k[v] = k;
        st->K[v] = (int64_t)(k * st->scale);
        tot += st->K[v];
    }
    q->tot[t] += tot;
}

int or_pass_init(or_state *st, float *q0) {
    or_ksum ks;
    memset(&ks, 0, sizeof(ks));
    ks.st = st;
    parallel_for(st->nthreads, st->n, k_range, &ks);
    int64_t tot = 0;
    for (int t = 0; t < OR_MAX_THREADS; t++) tot += ks.tot[t];
    /* all_w = reduce(neigh_w)/2 (modularity.cu:457-458): exact sum, one
     * rounding, exact halving */
    st->m = fix2f(st, tot) / 2.0f;
    /* a graph without weight would divide by zero in every gain and in Q;
     * it is reported as out of range instead */
    if (!(st->m > 0.0f)) return OR_ERR_RANGE;
    for (uint32_t v = 0; v < st->n; v++) {
        st->comm[v] = v;
        st->comm_sz[v] = 1;
        st->A[v] = st->K[v];
    }
    st->pass_ready = 1;
    float q = or_modularity(st);
    if (q0) *q0 = q;
    return OR_OK;
}

/* ------------------------------------------------------------------------ */
/* One bucket step: computeMove{Warp,Block,Global} + assignCommunity +       */
/* computeNewCommWeights (modularity.cu:95-398, driver :467-499)             */
/* ------------------------------------------------------------------------ */

/* isLowerIdx (modularity.cu:85-93) on the step snapshot */
static int is_lower_idx(const or_state *st, uint32_t cv, uint32_t ce) {
    if (st->comm_sz[cv] == 1 && st->comm_sz[ce] == 1) return ce < cv;
    return 1;
}

/* gain expression as written (modularity.cu:149-152), FP32, no contraction */
static float gain_of(float hw, float m, float kv, float acv, float ac) {
    float t1 = hw / m;
    float t3 = (acv - kv) - ac;
    float t4 = kv * t3;
    float t6 = (2.0f * m) * m;
    return t1 + t4 / t6;
}

static int better(float ga, uint32_t ca, float gb, uint32_t cb) {
    return ga > gb || (ga == gb && ca < cb);
}

/* Decide one vertex: returns its new community. */
static uint32_t decide(const or_state *st, or_scratch *sc, uint32_t v, int bucket) {
    const uint32_t T = k_lanes[bucket - 1];
    const uint64_t base = st->row[v];
    const uint32_t d = (uint32_t)(st->row[v + 1] - base);
    const uint32_t cv = st->comm[v];
    const float kv = st->k[v];
    const float m = st->m;
    const float acv = fix2f(st, st->A[cv]);
    float lane_g[128];
    uint32_t lane_c[128];
    for (uint32_t l = 0; l < T; l++) {
        lane_g[l] = -INFINITY;
        lane_c[l] = 0xffffffffu;
    }
    /* fresh accumulator generation for this vertex */
    sc->gen++;
    if (sc->gen == 0) {
        memset(sc->stamp, 0, (size_t)st->n * sizeof(uint32_t));
        sc->gen = 1;
    }
    const uint32_t gen = sc->gen;
    int64_t *acc = sc->acc;
    uint32_t *stamp = sc->stamp;
    int64_t cew = 0;
    const uint32_t rounds = (d + T - 1) / T;
    /* buckets 1-5: lock-step rounds, the gain of an entry in round r sees the
     * hash sums of all entries of rounds <= r (modularity.cu:312-351).
     * buckets 6-7: complete sums (cross-warp race of :218-258 canonicalised). */
    const int prefix = bucket <= 5;
    for (uint32_t r = 0; r < rounds; r++) {
        uint32_t j0 = r * T, j1 = j0 + T < d ? j0 + T : d;
        for (uint32_t j = j0; j < j1; j++) {
            uint32_t e = st->col[base + j];
            uint32_t ce = st->comm[e];
            int64_t wf = (int64_t)(st->w[base + j] * st->scale);
            if (stamp[ce] != gen) {
                stamp[ce] = gen;
                acc[ce] = 0;
            }
            acc[ce] += wf;
            if (ce == cv && e != v) cew += wf; /* :143-145 */
        }
        if (!prefix) continue;
        for (uint32_t j = j0; j < j1; j++) {
            uint32_t e = st->col[base + j];
            uint32_t ce = st->comm[e];
            if (!is_lower_idx(st, cv, ce)) continue;
            float g = gain_of(fix2f(st, acc[ce]), m, kv, acv, fix2f(st, st->A[ce]));
            uint32_t l = j % T;
            if (g > lane_g[l]) { /* strict: lane keeps its first max (:344) */
                lane_g[l] = g;
                lane_c[l] = ce;
            }
        }
    }
    if (!prefix) {
        for (uint32_t j = 0; j < d; j++) {
            uint32_t e = st->col[base + j];
            uint32_t ce = st->comm[e];
            if (!is_lower_idx(st, cv, ce)) continue;
            float g = gain_of(fix2f(st, acc[ce]), m, kv, acv, fix2f(st, st->A[ce]));
            uint32_t l = j % T;
            if (g > lane_g[l]) {
                lane_g[l] = g;
                lane_c[l] = ce;
            }
        }
    }
    float best_g;
    uint32_t dest;
    if (bucket <= 5) {
        /* shuffle-down reduction (:359-368): the surviving (gain, comm) is
         * the lexicographic best; destination comm[max_e] == that comm (:375) */
        best_g = lane_g[0];
        uint32_t best_c = lane_c[0];
        for (uint32_t l = 1; l < T; l++)
            if (better(lane_g[l], lane_c[l], best_g, best_c)) {
                best_g = lane_g[l];
                best_c = lane_c[l];
            }
        dest = best_c;
    } else {
        /* shared-memory tree (:260-275), lane i takes lane i+o when strictly
         * better; destination is comm of the winner lane's LAST visited entry
         * (graph.comm[e], :279 and :183) */
        float g[128];
        uint32_t c[128], pos[128];
        for (uint32_t l = 0; l < T; l++) {
            g[l] = lane_g[l];
            c[l] = lane_c[l];
            pos[l] = l;
        }
        for (uint32_t o = T / 2; o > 0; o >>= 1)
            for (uint32_t i = 0; i < o; i++)
                if (better(g[i + o], c[i + o], g[i], c[i])) {
                    g[i] = g[i + o];
                    c[i] = c[i + o];
                    pos[i] = pos[i + o];
                }
        best_g = g[0];
        uint32_t wl = pos[0];
        if (best_g == -INFINITY || wl >= d) return cv;
        uint32_t jl = wl + T * ((d - 1 - wl) / T);
        dest = st->comm[st->col[base + jl]];
    }
    if (best_g == -INFINITY) return cv; /* no candidate: stay (D5) */
    float cewf = fix2f(st, cew);
    if (best_g - (cewf / m) > 0.0f) return dest; /* :182, :278, :374 */
    return cv;
}

typedef struct {
    or_state *st;
    int bucket;
    uint64_t moved[OR_MAX_THREADS];
} or_stepctx;

static void decide_range(void *ctx, int t, uint32_t a, uint32_t b) {
    or_stepctx *c = (or_stepctx *)ctx;
    or_state *st = c->st;
    or_scratch *sc = &st->tscr[t];  /* entry 0 views the state's own scratch */
    for (uint32_t v = a; v < b; v++) {
        uint32_t d = (uint32_t)(st->row[v + 1] - st->row[v]);
        if (or_bucket_of(d) != c->bucket) continue;
        st->new_comm[v] = decide(st, sc, v, c->bucket);
    }
}

/* assignCommunity (:382-391) and the community weights; the exact
 * incremental update equals the recompute of :393-398 */
static void apply_range(void *ctx, int t, uint32_t a, uint32_t b) {
    or_stepctx *c = (or_stepctx *)ctx;
    or_state *st = c->st;
    uint64_t mv = 0;
    for (uint32_t v = a; v < b; v++) {
        uint32_t d = (uint32_t)(st->row[v + 1] - st->row[v]);
        if (or_bucket_of(d) != c->bucket) continue;
        uint32_t o = st->comm[v], nc = st->new_comm[v];
        if (o == nc) continue;
        __atomic_fetch_sub(&st->comm_sz[o], 1u, __ATOMIC_RELAXED);
        __atomic_fetch_add(&st->comm_sz[nc], 1u, __ATOMIC_RELAXED);
        __atomic_fetch_sub(&st->A[o], st->K[v], __ATOMIC_RELAXED);
        __atomic_fetch_add(&st->A[nc], st->K[v], __ATOMIC_RELAXED);
        st->comm[v] = nc;
        mv++;
    }
    c->moved[t] += mv;
}

int or_step(or_state *st, int bucket, uint64_t *moved) {
    if (!st->pass_ready) return OR_ERR_STATE;
    if (bucket < 1 || bucket > 7) return OR_ERR_ARG;
    or_stepctx c;
    memset(&c, 0, sizeof(c));
    c.st = st;
    c.bucket = bucket;
    /* decide every vertex of the step from the snapshot ... */
    st->tscr[0] = (or_scratch){st->acc, st->stamp, st->stamp_gen};
    parallel_for(st->nthreads, st->n, decide_range, &c);
    st->stamp_gen = st->tscr[0].gen;
    /* ... then apply them together */
    parallel_for(st->nthreads, st->n, apply_range, &c);
    uint64_t mv = 0;
    for (int t = 0; t < OR_MAX_THREADS; t++) mv += c.moved[t];
    if (moved) *moved = mv;
    return OR_OK;
}

int or_set_threads(or_state *st, int nthreads) {
    if (nthreads < 1 || nthreads > OR_MAX_THREADS) return OR_ERR_ARG;
    const size_t nn = st->n0 ? st->n0 : 1;
    or_scratch *ts = (or_scratch *)calloc((size_t)nthreads, sizeof(or_scratch));
    if (!ts) return OR_ERR_NOMEM;
    for (int t = 1; t < nthreads; t++) {
        ts[t].acc = (int64_t *)calloc(nn, sizeof(int64_t));
        ts[t].stamp = (uint32_t *)calloc(nn, sizeof(uint32_t));
        if (!ts[t].acc || !ts[t].stamp) {
            for (int u = 1; u <= t; u++) {
                free(ts[u].acc);
                free(ts[u].stamp);
            }
            free(ts);
            return OR_ERR_NOMEM;
        }
    }
    free_threads(st);
    st->tscr = ts;
    st->nthreads = nthreads;
    return OR_OK;
}

int or_sweep(or_state *st, float *q, uint64_t *moved) {
    if (!st->pass_ready) return OR_ERR_STATE;
    uint64_t tot = 0;
    for (int b = 1; b <= 7; b++) {
        uint64_t mv = 0;
        int rc = or_step(st, b, &mv);
        if (rc) return rc;
        tot += mv;
    }
    if (q) *q = or_modularity(st);
    if (moved) *moved = tot;
    return OR_OK;
}

#define OR_MAX_ITERS 100000u

/* modularityOptimization (modularity.cu:438-520).  Returns OR_OK; *improved
 * is the reference's return value.  *q_print is the value the reference
 * prints when the first sweep fails (old_modularity, :503-505).  A sweep that
 * moves no vertex ends the loop (the state is a fixed point; with g > 0 the
 * reference stops there too, with g <= 0 it would loop forever). */
int or_modularity_optimization(or_state *st, float min_gain, int *improved, float *q_print,
                               uint32_t *iters) {
    float act = 0.0f, old;
    int rc = or_pass_init(st, &act);
    if (rc) return rc;
    st->q_before = act;
    int first = 1, ret = 1;
    uint32_t it = 0;
    for (;;) {
        uint64_t moved = 0;
        old = act;
        rc = or_sweep(st, &act, &moved);
        if (rc) return rc;
        it++;
        int stop = !(act - old >= min_gain) || moved == 0 || it >= OR_MAX_ITERS;
        if (first && (act - old < min_gain || moved == 0)) {
            if (q_print) *q_print = old;
            ret = 0;
        }
        first = 0;
        if (stop) break;
    }
    st->q_after = act;
    if (ret) {
        /* graphVertToComm (:51-56, :515-517) */
        for (uint32_t i = 0; i < st->n0; i++) st->init_comm[i] = st->comm[st->init_comm[i]];
    }
    if (improved) *improved = ret;
    if (iters) *iters = it;
    return OR_OK;
}

/* ------------------------------------------------------------------------ */
/* Aggregation (aggregation.cu:387-462)                                      */
/* ------------------------------------------------------------------------ */

static int cmp_u32(const void *a, const void *b) {
    uint32_t x = *(const uint32_t *)a, y = *(const uint32_t *)b;
    return x < y ? -1 : x > y;
}

/* new_id(c) = number of non-empty c' < c (flags + exclusive scan, :21-36,
 * :390-393).  Row of new_id(c): distinct neighbour communities d of c's
 * members with W_cd = exact sum of entry weights rounded once to FP32, in
 * ascending new_id(d) (the reference's emit order is racy, :100-102).
 * Communities without entries get an empty row (compressEdges skips them,
 * :362, leaving row_ptr stale).  init_comm <- new_id[init_comm] (:380-385). */
int or_aggregation(or_state *st) {
    if (!st->pass_ready) return OR_ERR_STATE;
    const uint32_t n = st->n;
    uint32_t *new_id = (uint32_t *)malloc((size_t)(n ? n : 1) * sizeof(uint32_t));
    uint64_t *mstart = (uint64_t *)calloc((size_t)n + 1, sizeof(uint64_t));
    uint32_t *members = (uint32_t *)malloc((size_t)(n ? n : 1) * sizeof(uint32_t));
    if (!new_id || !mstart || !members) {
        free(new_id);
        free(mstart);
        free(members);
        return OR_ERR_NOMEM;
    }
    uint32_t nn = 0;
    for (uint32_t c = 0; c < n; c++) {
        new_id[c] = nn;
        if (st->comm_sz[c] != 0) nn++;
    }
    for (uint32_t v = 0; v < n; v++) mstart[st->comm[v] + 1]++;
    for (uint32_t c = 0; c < n; c++) mstart[c + 1] += mstart[c];
    {
        uint64_t *f = (uint64_t *)malloc(((size_t)n + 1) * sizeof(uint64_t));
        if (!f) {
            free(new_id);
            free(mstart);
            free(members);
            return OR_ERR_NOMEM;
        }
        memcpy(f, mstart, ((size_t)n + 1) * sizeof(uint64_t));
        for (uint32_t v = 0; v < n; v++) members[f[st->comm[v]]++] = v;
        free(f);
    }
    uint64_t cap = st->row[n];
    uint64_t *nrow = (uint64_t *)calloc((size_t)nn + 1, sizeof(uint64_t));
    uint32_t *ncol = (uint32_t *)malloc((cap ? cap : 1) * sizeof(uint32_t));
    float *nw = (float *)malloc((cap ? cap : 1) * sizeof(float));
    uint32_t *list = (uint32_t *)malloc((size_t)(n ? n : 1) * sizeof(uint32_t));
    if (!nrow || !ncol || !nw || !list) {
        free(new_id);
        free(mstart);
        free(members);
        free(nrow);
        free(ncol);
        free(nw);
        free(list);
        return OR_ERR_NOMEM;
    }
    uint64_t out = 0;
    for (uint32_t c = 0; c < n; c++) {
        if (st->comm_sz[c] == 0) continue;
        st->stamp_gen++;
        if (st->stamp_gen == 0) {
            memset(st->stamp, 0, (size_t)n * sizeof(uint32_t));
            st->stamp_gen = 1;
        }
        uint32_t gen = st->stamp_gen, cnt = 0;
        for (uint64_t mi = mstart[c]; mi < mstart[c + 1]; mi++) {
            uint32_t v = members[mi];
            for (uint64_t j = st->row[v]; j < st->row[v + 1]; j++) {
                uint32_t d = st->comm[st->col[j]];
                if (st->stamp[d] != gen) {
                    st->stamp[d] = gen;
                    st->acc[d] = 0;
                    list[cnt++] = d;
                }
                st->acc[d] += (int64_t)(st->w[j] * st->scale);
            }
        }
        qsort(list, cnt, sizeof(uint32_t), cmp_u32);
        for (uint32_t i = 0; i < cnt; i++) {
            ncol[out] = new_id[list[i]];
            nw[out] = fix2f(st, st->acc[list[i]]);
            out++;
        }
        nrow[new_id[c] + 1] = out;
    }
    for (uint32_t c = 0; c < nn; c++)
        if (nrow[c + 1] < nrow[c]) nrow[c + 1] = nrow[c];
    for (uint32_t i = 0; i < st->n0; i++) st->init_comm[i] = new_id[st->init_comm[i]];
    free(new_id);
    free(mstart);
    free(members);
    free(list);
    return set_graph(st, nn, nrow, ncol, nw);
}

/* main.cu:43-48 pass loop */
int or_run(or_state *st, float min_gain, or_result *res) {
    memset(res, 0, sizeof(*res));
    for (;;) {
        int improved = 0;
        uint32_t it = 0;
        float qp = 0.0f;
        uint64_t ne = st->row[st->n];
        int rc = or_modularity_optimization(st, min_gain, &improved, &qp, &it);
        if (rc) {
            res->status = rc;
            return rc;
        }
        res->total_iters += it;
        res->sum_edges += ne;
        res->sum_edge_visits += (uint64_t)(it + 1) * ne;
        if (!improved) {
            res->q_final = qp;
            break;
        }
        rc = or_aggregation(st);
        if (rc) {
            res->status = rc;
            return rc;
        }
        res->passes++;
    }
    res->n_final = st->n;
    return OR_OK;
}

/* ------------------------------------------------------------------------ */
/* Accessors                                                                 */
/* ------------------------------------------------------------------------ */

uint32_t or_n(const or_state *st) { return st->n; }
uint64_t or_ne(const or_state *st) { return st->row[st->n]; }
uint32_t or_n0(const or_state *st) { return st->n0; }
int or_scale(const or_state *st) { return st->S; }
float or_m(const or_state *st) { return st->m; }
void or_last_pass_q(const or_state *st, float *q_before, float *q_after) {
    if (q_before) *q_before = st->q_before;
    if (q_after) *q_after = st->q_after;
}
void or_get_csr(const or_state *st, uint64_t *row, uint32_t *col, float *w) {
    memcpy(row, st->row, ((size_t)st->n + 1) * sizeof(uint64_t));
    memcpy(col, st->col, st->row[st->n] * sizeof(uint32_t));
    memcpy(w, st->w, st->row[st->n] * sizeof(float));
}
void or_get_k(const or_state *st, float *k) { memcpy(k, st->k, (size_t)st->n * sizeof(float)); }
void or_get_comm(const or_state *st, uint32_t *c) {
    memcpy(c, st->comm, (size_t)st->n * sizeof(uint32_t));
}
void or_get_comm_sz(const or_state *st, uint32_t *c) {
    memcpy(c, st->comm_sz, (size_t)st->n * sizeof(uint32_t));
}
void or_get_a(const or_state *st, int64_t *a) { memcpy(a, st->A, (size_t)st->n * sizeof(int64_t)); }
void or_get_init_comm(const or_state *st, uint32_t *ic) {
    memcpy(ic, st->init_comm, (size_t)st->n0 * sizeof(uint32_t));
}

/* FP64 modularity of `assign` on the given graph, with the reference's
 * conventions: k_v = row sum (self-loop counted once, utils.cu:207-210),
 * 2m = sum of all entries. */
double or_q64(uint32_t n, const uint64_t *row, const uint32_t *col, const float *w,
              const uint32_t *assign) {
    double two_m = 0.0, in = 0.0;
    double *a = (double *)calloc((size_t)(n ? n : 1), sizeof(double));
    if (!a) return NAN;
    for (uint32_t v = 0; v < n; v++) {
        double k = 0.0;
        for (uint64_t j = row[v]; j < row[v + 1]; j++) {
            k += (double)w[j];
            if (assign[col[j]] == assign[v]) in += (double)w[j];
        }
        two_m += k;
        if (assign[v] < n) a[assign[v]] += k;
    }
    double sq = 0.0;
    for (uint32_t c = 0; c < n; c++) sq += a[c] * a[c];
    free(a);
    return in / two_m - sq / (two_m * two_m);
}
