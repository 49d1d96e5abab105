// lv_io.cpp -- host input path: the .mtx reader with the reference parser's
// semantics, and seeded synthetic graph generators for the benchmark and
// parity workloads.  Pure host C++ (no device code).
#include <algorithm>
#include <atomic>
#include <cerrno>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fcntl.h>
#include <new>
#include <sys/mman.h>
#include <sys/stat.h>
#include <thread>
#include <unistd.h>
#include <vector>

#include "../../include/louvain_hip.h"

namespace {

// ---------------------------------------------------------------------------
// parallel helper
// ---------------------------------------------------------------------------
unsigned n_workers() {
    unsigned h = std::thread::hardware_concurrency();
    if (h == 0) h = 1;
    if (h > 16) h = 16;
    const char *e = getenv("LV_GEN_THREADS");
    if (e && atoi(e) > 0) h = (unsigned)atoi(e);
    return h;
}

template <typename F>
void parallel_for(uint64_t n, F f) {
    unsigned T = n_workers();
    if (n < 4096 || T == 1) {
        f(0, n, 0u);
        return;
    }
    std::vector<std::thread> th;
    for (unsigned t = 0; t < T; ++t) {
        uint64_t a = n * t / T, b = n * (t + 1) / T;
        th.emplace_back([=] { f(a, b, t); });
    }
    for (auto &x : th) x.join();
}

// counter-based PRNG (splitmix64 finaliser over a mixed key)
inline uint64_t mix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}
inline uint64_t rng(uint64_t seed, uint64_t a, uint64_t b = 0) {
    return mix64(seed ^ mix64(a ^ mix64(b + 0x632BE59BD9B4E019ull)));
}
inline double u01(uint64_t r) { return (double)(r >> 11) * (1.0 / 9007199254740992.0); }

int alloc_csr(lv_host_csr *out, uint32_t n, uint64_t ne) {
    memset(out, 0, sizeof(*out));
    if (ne >= 0xffffffffull) return LV_ERR_ARG;
    out->row = (uint32_t *)malloc(((size_t)n + 1) * sizeof(uint32_t));
    out->col = (uint32_t *)malloc((ne ? ne : 1) * sizeof(uint32_t));
    out->w = (float *)malloc((ne ? ne : 1) * sizeof(float));
    if (!out->row || !out->col || !out->w) {
        free(out->row);
        free(out->col);
        free(out->w);
        memset(out, 0, sizeof(*out));
        return LV_ERR_NOMEM;
    }
    out->n = n;
    out->ne = ne;
    return LV_OK;
}

// Build a symmetric CSR from an undirected edge generator: gen(k, &u, &v)
// for k < m.  Rows are sorted ascending and deduplicated; a self-loop is
// stored once.  Weight of {u,v} from wfun(min, max).  Two generation passes
// (count, scatter) so the edge list is never materialised.
template <typename Gen, typename WFun>
int build_sym(uint32_t n, uint64_t m, Gen gen, WFun wfun, lv_host_csr *out) {
    std::vector<std::atomic<uint32_t>> deg(n);
    for (auto &d : deg) d.store(0, std::memory_order_relaxed);
    parallel_for(m, [&](uint64_t a, uint64_t b, unsigned) {
        for (uint64_t k = a; k < b; ++k) {
            uint32_t u, v;
            if (!gen(k, &u, &v)) continue;
            deg[u].fetch_add(1, std::memory_order_relaxed);
            if (u != v) deg[v].fetch_add(1, std::memory_order_relaxed);
        }
    });
    std::vector<uint64_t> off((size_t)n + 1, 0);
    for (uint32_t i = 0; i < n; ++i) off[i + 1] = off[i] + deg[i].load(std::memory_order_relaxed);
    const uint64_t tot = off[n];
    uint32_t *tmp = (uint32_t *)malloc((tot ? tot : 1) * sizeof(uint32_t));
    if (!tmp) return LV_ERR_NOMEM;
    for (uint32_t i = 0; i < n; ++i) deg[i].store(0, std::memory_order_relaxed);
    parallel_for(m, [&](uint64_t a, uint64_t b, unsigned) {
        for (uint64_t k = a; k < b; ++k) {
            uint32_t u, v;
            if (!gen(k, &u, &v)) continue;
            tmp[off[u] + deg[u].fetch_add(1, std::memory_order_relaxed)] = v;
            if (u != v) tmp[off[v] + deg[v].fetch_add(1, std::memory_order_relaxed)] = u;
        }
    });
    // sort + unique each row
    std::vector<uint32_t> uniq(n);
    parallel_for(n, [&](uint64_t a, uint64_t b, unsigned) {
        for (uint64_t i = a; i < b; ++i) {
            uint32_t *p = tmp + off[i], *e = tmp + off[i + 1];
            std::sort(p, e);
            uniq[i] = (uint32_t)(std::unique(p, e) - p);
        }
    });
    uint64_t ne = 0;
    for (uint32_t i = 0; i < n; ++i) ne += uniq[i];
    int rc = alloc_csr(out, n, ne);
    if (rc) {
        free(tmp);
        return rc;
    }
    out->row[0] = 0;
    for (uint32_t i = 0; i < n; ++i) out->row[i + 1] = out->row[i] + uniq[i];
    parallel_for(n, [&](uint64_t a, uint64_t b, unsigned) {
        for (uint64_t i = a; i < b; ++i) {
            const uint32_t *p = tmp + off[i];
            uint32_t o = out->row[i];
            for (uint32_t j = 0; j < uniq[i]; ++j) {
                uint32_t v = p[j];
                out->col[o + j] = v;
                out->w[o + j] = wfun(std::min<uint32_t>((uint32_t)i, v), std::max<uint32_t>((uint32_t)i, v));
            }
        }
    });
    free(tmp);
    out->nnz = 0;
    for (uint32_t i = 0; i < n; ++i)
        for (uint32_t j = out->row[i]; j < out->row[i + 1]; ++j)
            if (out->col[j] >= i) out->nnz++;
    return LV_OK;
}

// ---------------------------------------------------------------------------
// .mtx reader (utils.cu:132-229)
// ---------------------------------------------------------------------------
struct Cursor {
    const char *p, *end;
    bool line(const char **ls, const char **le) {
        if (p >= end) return false;
        const char *s = p;
        const char *e = (const char *)memchr(s, '\n', (size_t)(end - s));
        if (!e) e = end;
        *ls = s;
        *le = e;
        p = e < end ? e + 1 : e;
        return true;
    }
};

inline bool is_blank(char c) { return c == ' ' || c == '\t' || c == '\r' || c == '\v' || c == '\f'; }

const char *read_u32(const char *s, const char *e, uint64_t *v) {
    while (s < e && is_blank(*s)) ++s;
    if (s >= e || *s < '0' || *s > '9') return nullptr;
    uint64_t x = 0;
    while (s < e && *s >= '0' && *s <= '9') {
        x = x * 10 + (uint64_t)(*s - '0');
        if (x > 0xffffffffull) return nullptr;
        ++s;
    }
    *v = x;
    return s;
}

// `ss >> weight` (utils.cu:192): absent -> previous value kept, unparseable
// -> 0 (C++11 num_get), otherwise correctly rounded like strtof.
void read_weight(const char *s, const char *e, float *w) {
    while (s < e && is_blank(*s)) ++s;
    if (s >= e) return;
    char buf[128];
    size_t n = 0;
    while (s < e && !is_blank(*s) && n + 1 < sizeof(buf)) buf[n++] = *s++;
    buf[n] = 0;
    char *endp = nullptr;
    float x = strtof(buf, &endp);
    if (endp == buf) x = 0.0f;
    *w = x;
}

// One parsed edge line.  has_w: the line carried a weight token (a token
// that does not parse still counts: it sets the weight to 0).
struct EdgeLine {
    uint32_t a, b;
    float w;
    bool has_w;
};

// Parse body lines [s, e) whose first line is body line `first` (0-based);
// lines at index >= nnz are ignored like the reference ignores them.
// Returns LV_OK or LV_ERR_FORMAT; sets *zero if an id 0 was seen.
int parse_lines(const char *s, const char *e, uint64_t first, uint64_t nnz, EdgeLine *out,
                bool *zero) {
    Cursor cur{s, e};
    const char *ls, *le;
    for (uint64_t i = first; i < nnz && cur.line(&ls, &le); ++i) {
        uint64_t x, y;
        const char *t = read_u32(ls, le, &x);
        if (t) t = read_u32(t, le, &y);
        if (!t || (t < le && !is_blank(*t))) return LV_ERR_FORMAT;
        EdgeLine &L = out[i];
        L.a = (uint32_t)x;
        L.b = (uint32_t)y;
        L.has_w = false;
        float w = 0.0f;
        const char *u = t;
        while (u < le && is_blank(*u)) ++u;
        if (u < le) {
            read_weight(u, le, &w);
            L.has_w = true;
        }
        L.w = w;
        if (x == 0 || y == 0) *zero = true;  // isZeroStartingIndex (utils.cu:143-165)
    }
    return LV_OK;
}

// .mtx text -> symmetric CSR with the reference semantics (utils.cu:132-229):
// '%' lines before the header skipped, the column count is V, ids 1-based
// unless an id 0 appears, |weight|, a line without a weight keeps the
// previous one (1.0 at the start), rows in file order, a self-loop once.
// The body is split at line boundaries over worker threads; the weight
// carried across a split is resolved afterwards, and each worker then
// scatters the entries of its own row range in file order.
int parse_text(const char *text, size_t len, lv_host_csr *out) {
    memset(out, 0, sizeof(*out));
    Cursor cur{text, text + len};
    const char *ls, *le;
    for (;;) {  // header, skipping '%' lines (utils.cu:135-137)
        if (!cur.line(&ls, &le)) return LV_ERR_FORMAT;
        if (ls < le && *ls == '%') continue;
        break;
    }
    uint64_t rows, cols, nnz;
    const char *s = read_u32(ls, le, &rows);
    if (s) s = read_u32(s, le, &cols);
    if (s) s = read_u32(s, le, &nnz);
    if (!s) return LV_ERR_FORMAT;
    (void)rows;  // the reference keeps the column count (utils.cu:140)
    const uint32_t n = (uint32_t)cols;
    if (n >= 0x80000000u) return LV_ERR_ARG;
    const char *body = cur.p, *end = text + len;
    // split the body into T pieces at line starts; count each piece's lines
    unsigned T = n_workers();
    if ((size_t)(end - body) < (1u << 20)) T = 1;
    std::vector<const char *> cut(T + 1);
    cut[0] = body;
    cut[T] = end;
    for (unsigned t = 1; t < T; ++t) {  // first line start at or after the even split
        const char *q = body + (size_t)(end - body) * t / T;
        if (q <= cut[t - 1]) {
            cut[t] = cut[t - 1];
            continue;
        }
        const char *nl = (const char *)memchr(q - 1, '\n', (size_t)(end - (q - 1)));
        cut[t] = nl ? nl + 1 : end;
    }
    std::vector<uint64_t> lines(T + 1, 0);
    {
        std::vector<std::thread> th;
        for (unsigned t = 0; t < T; ++t)
            th.emplace_back([&, t] {
                uint64_t c = 0;
                for (const char *q = cut[t]; q < cut[t + 1];) {
                    const char *nl = (const char *)memchr(q, '\n', (size_t)(cut[t + 1] - q));
                    ++c;
                    if (!nl) break;
                    q = nl + 1;
                }
                lines[t + 1] = c;
            });
        for (auto &x : th) x.join();
    }
    for (unsigned t = 0; t < T; ++t) lines[t + 1] += lines[t];
    if (lines[T] < nnz) return LV_ERR_FORMAT;
    std::vector<EdgeLine> el(nnz);
    std::vector<int> prc(T, LV_OK);
    std::vector<char> pzero(T, 0);
    {
        std::vector<std::thread> th;
        for (unsigned t = 0; t < T; ++t)
            th.emplace_back([&, t] {
                bool z = false;
                prc[t] = parse_lines(cut[t], cut[t + 1], lines[t], nnz, el.data(), &z);
                pzero[t] = z;
            });
        for (auto &x : th) x.join();
    }
    bool zero_based = false;
    for (unsigned t = 0; t < T; ++t) {
        if (prc[t]) return prc[t];
        zero_based |= pzero[t] != 0;
    }
    // carried weight (`ss >> weight` leaves it unchanged, utils.cu:187-193)
    float wcur = 1.0f;
    for (uint64_t i = 0; i < nnz; ++i) {
        if (el[i].has_w) wcur = el[i].w;
        el[i].w = std::fabs(wcur);
        if (!zero_based) {
            el[i].a--;
            el[i].b--;
        }
        if (el[i].a >= n || el[i].b >= n) return LV_ERR_FORMAT;
    }
    std::vector<uint64_t> deg((size_t)n + 1, 0);
    for (uint64_t i = 0; i < nnz; ++i) {
        deg[el[i].a + 1]++;
        if (el[i].a != el[i].b) deg[el[i].b + 1]++;
    }
    for (uint32_t v = 0; v < n; ++v) deg[v + 1] += deg[v];
    const uint64_t ne = deg[n];
    int rc = alloc_csr(out, n, ne);
    if (rc) return rc;
    for (uint32_t v = 0; v <= n; ++v) out->row[v] = (uint32_t)deg[v];
    // rows in file order; self-loop once (utils.cu:200-224); worker t fills
    // rows [n*t/T, n*(t+1)/T) scanning every line in order
    const unsigned W = (nnz < (1u << 16)) ? 1u : n_workers();
    std::vector<std::thread> th;
    for (unsigned t = 0; t < W; ++t)
        th.emplace_back([&, t] {
            const uint32_t r0 = (uint32_t)((uint64_t)n * t / W);
            const uint32_t r1 = (uint32_t)((uint64_t)n * (t + 1) / W);
            for (uint64_t i = 0; i < nnz; ++i) {
                const uint32_t x = el[i].a, y = el[i].b;
                const float w = el[i].w;
                if (x >= r0 && x < r1) {
                    out->col[deg[x]] = y;
                    out->w[deg[x]++] = w;
                }
                if (x != y && y >= r0 && y < r1) {
                    out->col[deg[y]] = x;
                    out->w[deg[y]++] = w;
                }
            }
        });
    for (auto &x : th) x.join();
    out->nnz = nnz;
    return LV_OK;
}

// Binary CSR cache: "LVCSR001", n (u32), pad, nnz (u64), ne (u64), then
// row[n+1] (u32), col[ne] (u32), w[ne] (f32), little-endian as in memory.
constexpr char kCsrMagic[8] = {'L', 'V', 'C', 'S', 'R', '0', '0', '1'};

int load_binary(const char *p, size_t len, lv_host_csr *out) {
    memset(out, 0, sizeof(*out));
    if (len < 32) return LV_ERR_FORMAT;
    uint32_t n;
    uint64_t nnz, ne;
    memcpy(&n, p + 8, 4);
    memcpy(&nnz, p + 16, 8);
    memcpy(&ne, p + 24, 8);
    if (n >= 0x80000000u || ne >= 0xffffffffull) return LV_ERR_FORMAT;
    const size_t need = 32 + ((size_t)n + 1) * 4 + (size_t)ne * 8;
    if (len != need) return LV_ERR_FORMAT;
    int rc = alloc_csr(out, n, ne);
    if (rc) return rc;
    const char *q = p + 32;
    memcpy(out->row, q, ((size_t)n + 1) * 4);
    q += ((size_t)n + 1) * 4;
    memcpy(out->col, q, (size_t)ne * 4);
    q += (size_t)ne * 4;
    memcpy(out->w, q, (size_t)ne * 4);
    out->nnz = nnz;
    // the same invariants a parsed file has
    bool ok = out->row[0] == 0 && out->row[n] == ne;
    for (uint32_t v = 0; ok && v < n; ++v) ok = out->row[v] <= out->row[v + 1];
    for (uint64_t j = 0; ok && j < ne; ++j) ok = out->col[j] < n;
    if (!ok) {
        lv_host_csr_free(out);
        return LV_ERR_FORMAT;
    }
    return LV_OK;
}

}  // namespace

extern "C" {

int lv_mtx_read(const char *path, lv_host_csr *out) {
    if (!path || !out) return LV_ERR_ARG;
    memset(out, 0, sizeof(*out));
    int fd = open(path, O_RDONLY);
    if (fd < 0) return LV_ERR_IO;
    struct stat st;
    if (fstat(fd, &st) != 0) {
        close(fd);
        return LV_ERR_IO;
    }
    size_t len = (size_t)st.st_size;
    if (len == 0) {
        close(fd);
        return LV_ERR_FORMAT;
    }
    void *p = mmap(nullptr, len, PROT_READ, MAP_PRIVATE, fd, 0);
    close(fd);
    if (p == MAP_FAILED) return LV_ERR_IO;
    const char *t = (const char *)p;
    const int rc = (len >= 8 && memcmp(t, kCsrMagic, 8) == 0) ? load_binary(t, len, out)
                                                               : parse_text(t, len, out);
    munmap(p, len);
    return rc;
}

int lv_csr_write(const char *path, const lv_host_csr *csr) {
    if (!path || !csr || !csr->row || (csr->ne && (!csr->col || !csr->w))) return LV_ERR_ARG;
    FILE *f = fopen(path, "wb");
    if (!f) return LV_ERR_IO;
    char head[32] = {0};
    memcpy(head, kCsrMagic, 8);
    memcpy(head + 8, &csr->n, 4);
    memcpy(head + 16, &csr->nnz, 8);
    memcpy(head + 24, &csr->ne, 8);
    bool ok = fwrite(head, 1, 32, f) == 32;
    ok = ok && fwrite(csr->row, 4, (size_t)csr->n + 1, f) == (size_t)csr->n + 1;
    ok = ok && (!csr->ne || fwrite(csr->col, 4, csr->ne, f) == csr->ne);
    ok = ok && (!csr->ne || fwrite(csr->w, 4, csr->ne, f) == csr->ne);
    ok = (fclose(f) == 0) && ok;
    return ok ? LV_OK : LV_ERR_IO;
}

void lv_host_csr_free(lv_host_csr *csr) {
    if (!csr) return;
    free(csr->row);
    free(csr->col);
    free(csr->w);
    memset(csr, 0, sizeof(*csr));
}

// FEM-band analogue of bcsstk16 (6 DoF per node, dense 6x6 blocks).
int lv_gen_fem(uint32_t nodes, uint32_t band, uint32_t couple, uint64_t seed, lv_host_csr *out) {
    if (!out || nodes == 0 || band == 0) return LV_ERR_ARG;
    memset(out, 0, sizeof(*out));
    const double p = std::min(1.0, (couple / 2.0) / (double)band);
    // coupled node lists
    std::vector<std::vector<uint32_t>> nb(nodes);
    for (uint32_t a = 0; a < nodes; ++a) {
        nb[a].push_back(a);
        for (uint32_t o = 1; o <= band && a + o < nodes; ++o)
            if (u01(rng(seed, a, o)) < p) {
                nb[a].push_back(a + o);
                nb[a + o].push_back(a);
            }
    }
    const uint32_t n = nodes * 6;
    uint64_t ne = 0;
    for (uint32_t a = 0; a < nodes; ++a) ne += 36ull * nb[a].size();
    int rc = alloc_csr(out, n, ne);
    if (rc) return rc;
    auto wfun = [&](uint32_t r, uint32_t c) {
        uint32_t lo = std::min(r, c), hi = std::max(r, c);
        uint64_t x = rng(seed ^ 0x5EEDull, lo, hi);
        double u1 = (double)((x >> 11) + 1) * (1.0 / 9007199254740993.0);
        double u2 = u01(mix64(x));
        float z = (float)(std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2));
        return std::fabs(z) + 0.01f;
    };
    uint32_t o = 0;
    out->row[0] = 0;
    for (uint32_t a = 0; a < nodes; ++a) {
        std::vector<uint32_t> nodes_sorted = nb[a];
        std::sort(nodes_sorted.begin(), nodes_sorted.end());
        for (int i = 0; i < 6; ++i) {
            uint32_t r = a * 6 + i;
            for (uint32_t b : nodes_sorted)
                for (int j = 0; j < 6; ++j) {
                    uint32_t c = b * 6 + j;
                    out->col[o] = c;
                    out->w[o] = wfun(r, c);
                    ++o;
                }
            out->row[r + 1] = o;
        }
    }
    out->nnz = 0;
    for (uint32_t r = 0; r < n; ++r)
        for (uint32_t j = out->row[r]; j < out->row[r + 1]; ++j)
            if (out->col[j] >= r) out->nnz++;
    return LV_OK;
}

// Chung-Lu power-law degrees over a planted partition (unit weights).
int lv_gen_chung_lu(uint32_t n, double mean_deg, double exponent, uint32_t max_deg,
                    uint32_t groups, double p_ratio, uint64_t seed, lv_host_csr *out) {
    if (!out || n < 2 || groups == 0 || exponent <= 1.0) return LV_ERR_ARG;
    memset(out, 0, sizeof(*out));
    std::vector<double> wt(n);
    const double ex = -1.0 / (exponent - 1.0);
    double sum = 0;
    for (uint32_t i = 0; i < n; ++i) sum += (wt[i] = std::pow((double)i + 1.0, ex));
    for (int it = 0; it < 8; ++it) {  // scale to the mean, then cap
        double sc = mean_deg * n / sum;
        sum = 0;
        for (uint32_t i = 0; i < n; ++i) {
            wt[i] = std::min((double)max_deg, wt[i] * sc);
            sum += wt[i];
        }
    }
    // random group per vertex; cumulative weights globally and per group
    std::vector<uint32_t> grp(n), gcount(groups, 0);
    for (uint32_t i = 0; i < n; ++i) gcount[grp[i] = (uint32_t)(rng(seed, 7, i) % groups)]++;
    std::vector<uint64_t> goff(groups + 1, 0);
    for (uint32_t g = 0; g < groups; ++g) goff[g + 1] = goff[g] + gcount[g];
    std::vector<uint32_t> gmem(n), gfill(groups, 0);
    for (uint32_t i = 0; i < n; ++i) gmem[goff[grp[i]] + gfill[grp[i]]++] = i;
    std::vector<double> cdf(n), gcdf(n);
    double acc = 0;
    for (uint32_t i = 0; i < n; ++i) cdf[i] = (acc += wt[i]);
    for (uint32_t g = 0; g < groups; ++g) {
        double a2 = 0;
        for (uint64_t k = goff[g]; k < goff[g + 1]; ++k) gcdf[k] = (a2 += wt[gmem[k]]);
    }
    const double s = (double)n / groups;
    const double f_in = p_ratio * (s - 1) / (p_ratio * (s - 1) + (n - s));
    const uint64_t m = (uint64_t)std::llround(n * mean_deg / 2.0);
    auto gen = [&](uint64_t k, uint32_t *u, uint32_t *v) {
        uint64_t r0 = rng(seed, 1, k), r1 = rng(seed, 2, k), r2 = rng(seed, 3, k);
        uint32_t a = (uint32_t)(std::upper_bound(cdf.begin(), cdf.end(), u01(r0) * acc) - cdf.begin());
        if (a >= n) a = n - 1;
        uint32_t b;
        if (u01(r1) < f_in) {
            uint32_t g = grp[a];
            double tot = gcdf[goff[g + 1] - 1];
            uint64_t idx = (uint64_t)(std::upper_bound(gcdf.begin() + goff[g], gcdf.begin() + goff[g + 1],
                                                       u01(r2) * tot) -
                                      gcdf.begin());
            if (idx >= goff[g + 1]) idx = goff[g + 1] - 1;
            b = gmem[idx];
        } else {
            b = (uint32_t)(std::upper_bound(cdf.begin(), cdf.end(), u01(r2) * acc) - cdf.begin());
            if (b >= n) b = n - 1;
        }
        if (a == b) return false;  // no self-loops in a co-purchase graph
        *u = a;
        *v = b;
        return true;
    };
    return build_sym(n, m, gen, [](uint32_t, uint32_t) { return 1.0f; }, out);
}

// R-MAT, Graph500 style.
int lv_gen_rmat(uint32_t scale, uint32_t edge_factor, double a, double b, double c, float wlo,
                float whi, uint64_t seed, lv_host_csr *out) {
    if (!out || scale == 0 || scale > 30) return LV_ERR_ARG;
    memset(out, 0, sizeof(*out));
    const uint32_t n = 1u << scale;
    const uint32_t mask = n - 1;
    const uint64_t m = (uint64_t)edge_factor << scale;
    const double ab = a + b, abc = a + b + c;
    const uint32_t half = scale / 2 > 0 ? scale / 2 : 1;
    auto perm = [=](uint32_t x) {  // bijection on [0, 2^scale)
        for (int r = 0; r < 3; ++r) {
            x ^= x >> half;
            x = (uint32_t)((x * 0x9E3779B1ull + (0x7F4A7C15u + 0x1234567u * r)) & mask);
        }
        return x & mask;
    };
    // quadrant thresholds on 16-bit uniforms: one 64-bit hash serves 4 levels
    const uint32_t ta = (uint32_t)std::lround(a * 65536.0), tab = (uint32_t)std::lround(ab * 65536.0),
                   tabc = (uint32_t)std::lround(abc * 65536.0);
    auto gen = [=](uint64_t k, uint32_t *u, uint32_t *v) {
        uint32_t x = 0, y = 0;
        uint64_t r = 0;
        for (uint32_t l = 0; l < scale; ++l) {
            if ((l & 3) == 0) r = mix64(seed ^ (k * 0x9E3779B97F4A7C15ull + (l >> 2)));
            const uint32_t q = (uint32_t)(r & 0xffffu);
            r >>= 16;
            uint32_t bx = 0, by = 0;
            if (q < ta) {
            } else if (q < tab) {
                by = 1;
            } else if (q < tabc) {
                bx = 1;
            } else {
                bx = 1;
                by = 1;
            }
            x = (x << 1) | bx;
            y = (y << 1) | by;
        }
        *u = perm(x);
        *v = perm(y);
        return true;
    };
    auto wfun = [=](uint32_t lo, uint32_t hi) {
        if (wlo == whi) return wlo;
        double r = u01(rng(seed ^ 0xABCDEFull, lo, hi));
        return wlo + (whi - wlo) * (float)r;
    };
    return build_sym(n, m, gen, wfun, out);
}

// Stochastic block model with uniform expected degree.
int lv_gen_sbm(uint32_t n, uint32_t blocks, double mean_deg, double p_ratio, uint64_t seed,
               lv_host_csr *out) {
    if (!out || n < 2 || blocks == 0 || blocks > n) return LV_ERR_ARG;
    memset(out, 0, sizeof(*out));
    const double s = (double)n / blocks;
    const double f_in = p_ratio * (s - 1) / (p_ratio * (s - 1) + (n - s));
    const uint64_t m = (uint64_t)std::llround(n * mean_deg / 2.0);
    auto gen = [=](uint64_t k, uint32_t *u, uint32_t *v) {
        uint32_t a = (uint32_t)(rng(seed, 1, k) % n);
        uint32_t b;
        if (u01(rng(seed, 2, k)) < f_in) {
            uint32_t blk = (uint32_t)((uint64_t)a * blocks / n);
            uint32_t lo = (uint32_t)(((uint64_t)blk * n + blocks - 1) / blocks);
            uint32_t hi = (uint32_t)(((uint64_t)(blk + 1) * n + blocks - 1) / blocks);
            if (hi <= lo) hi = lo + 1;
            b = lo + (uint32_t)(rng(seed, 3, k) % (hi - lo));
        } else {
            b = (uint32_t)(rng(seed, 3, k) % n);
        }
        if (a == b) return false;
        *u = a;
        *v = b;
        return true;
    };
    return build_sym(n, m, gen, [](uint32_t, uint32_t) { return 1.0f; }, out);
}

}  // extern "C"
