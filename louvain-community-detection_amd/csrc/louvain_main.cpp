// louvain_main.cpp -- command-line drop-in for the reference binary
// (main.cu, utils.cu:231-293):
//
//   louvain -f <file.mtx> -g <min-gain> [-v]
//
// stdout carries exactly what the reference prints: the final modularity
// (modularity.cu:504), then "alg_ms total_ms" (main.cu:56), then with -v the
// community listing of printCommunities (utils.cu:270-293).  Extra per-pass
// figures go to stderr with -t.
#include <getopt.h>
#include <hip/hip_runtime.h>

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <iterator>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/louvain_hip.h"

static void usage_exit() {
    std::cout << "Usage: ./gpulouvain -f mtx-matrix-file -g min-gain [-v]" << std::endl;
    exit(1);
}

// parseArgs (utils.cu:231-267)
static void parse_args(int argc, char **argv, float &threshold, bool &verbose, bool &timing,
                       char *&filename) {
    int c;
    bool fopt = false, gopt = false;
    while ((c = getopt(argc, argv, "f:g:vt")) != -1) {
        switch (c) {
            case 'f':
                filename = optarg;
                fopt = true;
                break;
            case 'g':
                try {
                    threshold = std::stof(std::string(optarg));
                } catch (const std::invalid_argument &ia) {
                    std::cout << "-g option invalid argument: " << ia.what() << std::endl;
                    exit(1);
                } catch (const std::out_of_range &oor) {
                    std::cout << "-g option error out of range: " << oor.what() << std::endl;
                    exit(1);
                }
                gopt = true;
                break;
            case 'v':
                verbose = true;
                break;
            case 't':
                timing = true;
                break;
            case '?':
                usage_exit();
                break;
            default:
                std::cout << "?? arg_parser error ??" << std::endl;
        }
    }
    if (!fopt || !gopt) usage_exit();
}

static int fail(int rc) {
    std::cerr << "louvain: " << lv_status_string(rc) << std::endl;
    return 1;
}

int main(int argc, char **argv) {
    char *filename = nullptr;
    float threshold = 0.0f;
    bool verbose = false, timing = false;
    parse_args(argc, argv, threshold, verbose, timing, filename);

    lv_host_csr csr;
    int rc = lv_mtx_read(filename, &csr);
    if (rc == LV_ERR_IO) {
        std::cout << "Error opening file" << std::endl;  // utils.cu:171-174
        return 1;
    }
    if (rc) return fail(rc);

    // copy phase: upload + state init (main.cu:19-36)
    auto c0 = std::chrono::steady_clock::now();
    lv_graph *g = nullptr;
    rc = lv_graph_create(0, csr.n, csr.row, csr.col, csr.w, &g);
    auto c1 = std::chrono::steady_clock::now();
    if (rc) {
        lv_host_csr_free(&csr);
        return fail(rc);
    }
    const float copy_ms = std::chrono::duration<float, std::milli>(c1 - c0).count();

    lv_result res;
    rc = lv_run(g, threshold, &res);
    if (rc) {
        lv_graph_destroy(g);
        lv_host_csr_free(&csr);
        return fail(rc);
    }
    const float alg_ms = (float)res.alg_ms;
    std::cout << res.q_final << std::endl;                                 // modularity.cu:504
    std::cout << alg_ms << " " << alg_ms + copy_ms << std::endl;           // main.cu:56
    if (timing) {
        std::cerr << "passes " << res.passes << " iters " << res.total_iters << " communities "
                  << res.n_final << " edges/s " << (double)res.sum_edges / (res.alg_ms * 1e-3)
                  << std::endl;
    }
    if (verbose) {  // printCommunities (utils.cu:270-293)
        uint32_t n = 0, n0 = 0;
        uint64_t ne = 0;
        lv_graph_size(g, &n, &ne, &n0);
        std::vector<uint32_t> comm(n0);
        rc = lv_get_assignment(g, comm.data());
        if (rc) {
            lv_graph_destroy(g);
            lv_host_csr_free(&csr);
            return fail(rc);
        }
        std::vector<std::vector<uint32_t>> comm_vert(n);
        for (uint32_t i = 0; i < n0; ++i) comm_vert[comm[i]].push_back(i + 1);
        unsigned int comm_n = 1;
        std::cout << n << std::endl;
        for (uint32_t i = 0; i < n; ++i) {
            if (!comm_vert[i].empty()) {
                std::cout << comm_n << " ";
                std::copy(comm_vert[i].begin(), comm_vert[i].end(),
                          std::ostream_iterator<unsigned int>(std::cout, " "));
                std::cout << std::endl;
                comm_n++;
            }
        }
    }
    lv_graph_destroy(g);
    lv_host_csr_free(&csr);
    return 0;
}
