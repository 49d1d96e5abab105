// louvain_main.cpp -- command-line drop-in for the reference binary
// (main.cu, utils.cu:231-293):
//
//   louvain -f <file.mtx> -g <min-gain> [-v] [-t]
//
// Output contract (what a caller of the reference binary sees):
//   stdout line 1  final modularity, default ostream formatting (modularity.cu:504)
//   stdout line 2  "alg_ms total_ms" (main.cu:56)
//   with -v        the number of communities, then one line per non-empty
//                  community in increasing id: its 1-based rank, then its
//                  members as 1-based vertex ids in increasing order, each
//                  followed by one space (utils.cu:270-293)
//   errors         usage line on stdout and exit status 1 for a missing -f/-g
//                  or an unknown option; "-g option invalid argument: stof" /
//                  "-g option error out of range: stof" for a gain std::stof
//                  rejects; "Error opening file" for an unreadable input
//                  (utils.cu:171-174, :245-264).
// -t (not in the reference) prints per-run figures on stderr.
#include <getopt.h>

#include <cerrno>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <string>
#include <vector>

#include "../../include/louvain_hip.h"

namespace {

const char kUsage[] = "Usage: ./gpulouvain -f mtx-matrix-file -g min-gain [-v]";

struct CliOptions {
    const char *path = nullptr;
    float min_gain = 0.0f;
    bool have_path = false, have_gain = false;
    bool list_communities = false;  // -v
    bool timings = false;           // -t
};

enum class GainParse { ok, invalid, range };

// std::stof's acceptance rule without exceptions: leading white space
// skipped, the longest prefix strtof converts is used, nothing converted ->
// invalid, ERANGE -> out of range.
GainParse parse_gain(const char *text, float *out) {
    char *end = nullptr;
    errno = 0;
    const float x = strtof(text, &end);
    if (end == text) return GainParse::invalid;
    if (errno == ERANGE) return GainParse::range;
    *out = x;
    return GainParse::ok;
}

// Returns -1 when the options are complete, else the exit status after the
// message has been printed.
int read_options(int argc, char **argv, CliOptions *o) {
    for (int opt; (opt = getopt(argc, argv, "f:g:vt")) != -1;) {
        if (opt == 'f') {
            o->path = optarg;
            o->have_path = true;
        } else if (opt == 'g') {
            const GainParse r = parse_gain(optarg, &o->min_gain);
            if (r != GainParse::ok) {
                std::cout << (r == GainParse::invalid ? "-g option invalid argument: "
                                                      : "-g option error out of range: ")
                          << "stof" << std::endl;
                return 1;
            }
            o->have_gain = true;
        } else if (opt == 'v') {
            o->list_communities = true;
        } else if (opt == 't') {
            o->timings = true;
        } else {  // unknown option or missing argument
            std::cout << kUsage << std::endl;
            return 1;
        }
    }
    if (!o->have_path || !o->have_gain) {
        std::cout << kUsage << std::endl;
        return 1;
    }
    return -1;
}

// Members of every community in increasing vertex order by a counting sort,
// written as one buffer.
void list_communities(uint32_t n_comm, const std::vector<uint32_t> &assign) {
    std::vector<uint64_t> start(n_comm + 1, 0);
    for (uint32_t c : assign) start[c + 1]++;
    for (uint32_t c = 0; c < n_comm; ++c) start[c + 1] += start[c];
    std::vector<uint32_t> members(assign.size());
    std::vector<uint64_t> fill(start.begin(), start.end() - 1);
    for (size_t v = 0; v < assign.size(); ++v) members[fill[assign[v]]++] = (uint32_t)v;
    std::string out = std::to_string(n_comm) + "\n";
    uint64_t rank = 0;
    char num[24];
    for (uint32_t c = 0; c < n_comm; ++c) {
        if (start[c] == start[c + 1]) continue;
        out += std::to_string(++rank);
        out += ' ';
        for (uint64_t i = start[c]; i < start[c + 1]; ++i) {
            const int len = snprintf(num, sizeof(num), "%u ", members[i] + 1u);
            out.append(num, (size_t)len);
        }
        out += '\n';
        if (out.size() > (1u << 20)) {
            std::cout.write(out.data(), (std::streamsize)out.size());
            out.clear();
        }
    }
    std::cout.write(out.data(), (std::streamsize)out.size());
    std::cout.flush();
}

int report(int rc) {
    std::cerr << "louvain: " << lv_status_string(rc) << std::endl;
    return 1;
}

}  // namespace

int main(int argc, char **argv) {
    CliOptions opt;
    const int early = read_options(argc, argv, &opt);
    if (early >= 0) return early;

    lv_host_csr csr;
    int rc = lv_mtx_read(opt.path, &csr);
    if (rc == LV_ERR_IO) {
        std::cout << "Error opening file" << std::endl;
        return 1;
    }
    if (rc) return report(rc);

    // the reference's "copy" time: upload and state set-up (main.cu:19-36)
    const auto t0 = std::chrono::steady_clock::now();
    lv_graph *g = nullptr;
    rc = lv_graph_create(0, csr.n, csr.row, csr.col, csr.w, &g);
    const auto t1 = std::chrono::steady_clock::now();
    lv_host_csr_free(&csr);
    if (rc) return report(rc);
    const float upload_ms = std::chrono::duration<float, std::milli>(t1 - t0).count();

    lv_result res;
    rc = lv_run(g, opt.min_gain, &res);
    if (rc) {
        lv_graph_destroy(g);
        return report(rc);
    }
    const float run_ms = (float)res.alg_ms;
    std::cout << res.q_final << std::endl;
    std::cout << run_ms << " " << run_ms + upload_ms << std::endl;
    if (opt.timings)
        std::cerr << "passes " << res.passes << " iters " << res.total_iters << " communities "
                  << res.n_final << " edges/s " << (double)res.sum_edges / (res.alg_ms * 1e-3)
                  << std::endl;
    if (opt.list_communities) {
        uint32_t n = 0, n0 = 0;
        uint64_t ne = 0;
        lv_graph_size(g, &n, &ne, &n0);
        std::vector<uint32_t> assign(n0);
        rc = lv_get_assignment(g, assign.data());
        if (rc) {
            lv_graph_destroy(g);
            return report(rc);
        }
        list_communities(n, assign);
    }
    lv_graph_destroy(g);
    return 0;
}
