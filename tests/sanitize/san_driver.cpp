// san_driver.cpp -- host code of the engine and the CPU oracle under
// AddressSanitizer and UndefinedBehaviorSanitizer (CPU only, no GPU needed).
//
// Exercises: the .mtx reader and binary CSR cache (lv_io.cpp), the
// generators, the input checks of lv_graph_create / lv_batch_create /
// lv_run_noalloc (lv_engine.cpp, rejected inputs return before any device
// work; valid ones reach the device query and report LV_ERR_NO_DEVICE on a
// machine without a GPU), and the oracle's parser and full runs, single- and
// multi-threaded.  Exit status 0 when every expectation holds.
//
//   san_driver <bcsstk16.mtx> <scratch dir>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/louvain_hip.h"
#include "../../oracle/lv_oracle.h"

static int failures = 0;
#define EXPECT(cond)                                                    \
    do {                                                                \
        if (!(cond)) {                                                  \
            std::fprintf(stderr, "FAILED %s:%d: %s\n", __FILE__, __LINE__, #cond); \
            ++failures;                                                 \
        }                                                               \
    } while (0)

static std::string write_file(const std::string &dir, const char *name, const char *text) {
    const std::string p = dir + "/" + name;
    FILE *f = std::fopen(p.c_str(), "wb");
    if (!f) return p;
    std::fputs(text, f);
    std::fclose(f);
    return p;
}

// n, row, col, w of an lv_host_csr as the oracle takes them
static or_state *oracle_of(const lv_host_csr &c, int *err) {
    std::vector<uint64_t> row(c.row, c.row + c.n + 1);
    return or_create(c.n, row.data(), c.col, c.w, err);
}

static void check_reader(const std::string &mtx, const std::string &dir) {
    lv_host_csr c;
    EXPECT(lv_mtx_read(mtx.c_str(), &c) == LV_OK);
    EXPECT(c.n == 4884 && c.ne == 290378);
    // binary cache round trip
    const std::string bin = dir + "/b16.lvcsr";
    EXPECT(lv_csr_write(bin.c_str(), &c) == LV_OK);
    lv_host_csr d;
    EXPECT(lv_mtx_read(bin.c_str(), &d) == LV_OK);
    EXPECT(d.n == c.n && d.ne == c.ne);
    EXPECT(std::memcmp(d.row, c.row, (c.n + 1) * sizeof(uint32_t)) == 0);
    EXPECT(std::memcmp(d.col, c.col, c.ne * sizeof(uint32_t)) == 0);
    EXPECT(std::memcmp(d.w, c.w, c.ne * sizeof(float)) == 0);
    lv_host_csr_free(&d);
    // the oracle's parser gives the same CSR
    or_graph og;
    EXPECT(or_parse_mtx(mtx.c_str(), &og) == OR_OK);
    EXPECT(og.n == c.n && og.row[og.n] == c.ne);
    for (uint64_t j = 0; j < c.ne && j < og.row[og.n]; ++j)
        if (og.col[j] != c.col[j] || og.w[j] != c.w[j]) {
            EXPECT(false);
            break;
        }
    or_graph_free(&og);
    lv_host_csr_free(&c);
    // malformed inputs and a truncated cache
    const char *bad[] = {"2 2 2\n1 2 1\n", "2 2 1\n1 3 1\n", "2 2 1\nx y\n", "", "%only\n"};
    for (int i = 0; i < 5; ++i) {
        const std::string p = write_file(dir, ("bad" + std::to_string(i) + ".mtx").c_str(), bad[i]);
        lv_host_csr e;
        EXPECT(lv_mtx_read(p.c_str(), &e) == LV_ERR_FORMAT);
        or_graph o;
        EXPECT(or_parse_mtx(p.c_str(), &o) != OR_OK);
    }
    const std::string trunc = write_file(dir, "trunc.lvcsr", "LVCSR001\x05");
    lv_host_csr t;
    EXPECT(lv_mtx_read(trunc.c_str(), &t) != LV_OK);
    EXPECT(lv_mtx_read((dir + "/missing.mtx").c_str(), &t) == LV_ERR_IO);
    // a carried weight, CRLF, a negative weight, a line past the count
    const std::string odd = write_file(dir, "odd.mtx",
                                       "%%MatrixMarket\r\n3 3 3\r\n2 1 0.5\r\n3 1\r\n3 2 -2\r\n9 9 9\r\n");
    lv_host_csr o2 = {};
    EXPECT(lv_mtx_read(odd.c_str(), &o2) == LV_OK);
    if (o2.row) lv_host_csr_free(&o2);
}

static void check_graph_inputs() {
    lv_graph *g = nullptr;
    const uint32_t row[] = {0, 1, 2};
    const uint32_t col_ok[] = {1, 0}, col_bad[] = {1, 5};
    const float w_ok[] = {1.0f, 1.0f};
    EXPECT(lv_graph_create(0, 2, row, col_bad, w_ok, &g) == LV_ERR_ARG);
    const uint32_t row_bad[] = {0, 2, 1};
    EXPECT(lv_graph_create(0, 2, row_bad, col_ok, w_ok, &g) == LV_ERR_ARG);
    const float bad_w[] = {INFINITY, NAN, -1.0f, 1e-30f, 1e30f, 1.1805916e21f /* 2^70 */};
    for (float x : bad_w) {
        const float w[] = {x, x};
        EXPECT(lv_graph_create(0, 2, row, col_ok, w, &g) == LV_ERR_RANGE);
    }
    // valid input: the checks pass and the device query decides
    const int rc = lv_graph_create(0, 2, row, col_ok, w_ok, &g);
    EXPECT(rc == LV_ERR_NO_DEVICE || rc == LV_OK);
    if (rc == LV_OK) lv_graph_destroy(g);
    // batches: an entry leaving its graph, a negative weight, no column array
    const uint64_t off[] = {0, 2, 4};
    const uint32_t brow[] = {0, 1, 2, 3, 4};
    const uint32_t bcol_out[] = {1, 2, 3, 2}, bcol[] = {1, 0, 3, 2};
    const float bw[] = {1, 1, 1, 1}, bw_neg[] = {1, 1, -1, -1};
    EXPECT(lv_batch_create(0, 2, off, brow, bcol_out, bw, &g) == LV_ERR_ARG);
    EXPECT(lv_batch_create(0, 2, off, brow, bcol, bw_neg, &g) == LV_ERR_RANGE);
    EXPECT(lv_batch_create(0, 2, off, brow, nullptr, nullptr, &g) == LV_ERR_ARG);
    // the allocation-free run validates the same way
    lv_result res;
    uint32_t assign[2];
    EXPECT(lv_run_noalloc(0, 2, row, col_bad, w_ok, 1e-6f, nullptr, assign, &res) != LV_OK);
    EXPECT(lv_run_noalloc(0, 2, row, col_ok, w_ok, 1e-6f, nullptr, nullptr, &res) == LV_ERR_ARG);
    lv_tuning t;
    EXPECT(lv_tuning_default(&t) == LV_OK);
}

static void check_oracle_runs(const std::string &mtx) {
    lv_host_csr graphs[4];
    EXPECT(lv_mtx_read(mtx.c_str(), &graphs[0]) == LV_OK);
    EXPECT(lv_gen_fem(200, 25, 9, 3, &graphs[1]) == LV_OK);
    EXPECT(lv_gen_rmat(12, 16, 0.57, 0.19, 0.19, 0.5f, 1.5f, 5, &graphs[2]) == LV_OK);
    EXPECT(lv_gen_sbm(3000, 16, 12.0, 10.0, 7, &graphs[3]) == LV_OK);
    lv_host_csr cl;
    EXPECT(lv_gen_chung_lu(5000, 9.4, 2.5, 200, 40, 20.0, 9, &cl) == LV_OK);
    lv_host_csr_free(&cl);
    for (auto &c : graphs) {
        or_result r1, r4;
        int err = 0;
        or_state *a = oracle_of(c, &err);
        EXPECT(a != nullptr);
        if (!a) continue;
        EXPECT(or_run(a, 1e-6f, &r1) == OR_OK);
        or_state *b = oracle_of(c, &err);
        EXPECT(b && or_set_threads(b, 4) == OR_OK);
        EXPECT(b && or_run(b, 1e-6f, &r4) == OR_OK);
        EXPECT(r1.q_final == r4.q_final && r1.total_iters == r4.total_iters &&
               r1.n_final == r4.n_final && r1.passes == r4.passes);
        if (b) {
            std::vector<uint32_t> ia(or_n0(a)), ib(or_n0(b));
            or_get_init_comm(a, ia.data());
            or_get_init_comm(b, ib.data());
            EXPECT(ia == ib);
            or_destroy(b);
        }
        or_destroy(a);
        lv_host_csr_free(&c);
    }
}

int main(int argc, char **argv) {
    if (argc < 3) {
        std::fprintf(stderr, "usage: san_driver <bcsstk16.mtx> <scratch dir>\n");
        return 2;
    }
    check_reader(argv[1], argv[2]);
    check_graph_inputs();
    check_oracle_runs(argv[1]);
    std::printf("%s: %d failures\n", failures ? "FAILED" : "ok", failures);
    return failures ? 1 : 0;
}
