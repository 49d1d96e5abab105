// lv_host.h -- host-side engine internals shared by the C ABI units
// (lv_engine.cpp, lv_server.cpp); not part of the public interface.
#pragma once

#include <stddef.h>
#include <stdint.h>

#include "../../include/louvain_hip.h"

namespace lv_host {

// lv_graph_create's input checks; *S = fixed-point scale bits of the weights
int check_input(uint32_t n, const uint32_t *row, const uint32_t *col, const float *w, int *S);
// device bytes of a handle for up to n vertices and ne entries, every arena
// that grows on demand counted at its worst case
uint64_t footprint(uint32_t n, uint64_t ne);
// a handle with device buffers for up to n_cap vertices and e_cap entries;
// with `pristine`, a pinned host copy of the input for lv_graph_reset
int graph_new(int device, uint32_t n_cap, uint64_t e_cap, bool pristine, lv_graph **out);
// upload a checked input that fits the handle's capacity (LV_ERR_NOMEM if not)
int graph_upload(lv_graph *g, uint32_t n, const uint32_t *row, const uint32_t *col,
                 const float *w, int S);

// host memory the per-sweep counter readback of run_fixed needs
constexpr size_t kHostCountersBytes = 16384;
// The whole pass loop (lv_run) on a graph whose device buffers are all
// carved from the `cap` bytes at device address `base`, on `stream` (a
// hipStream_t; null: the null stream), with no device or pinned-host
// allocation and no HIP graph.  host_counters: kHostCountersBytes of host
// memory, 8-byte aligned.  LV_ERR_LIMIT if the graph's buffers exceed cap.
int run_fixed(int device, char *base, size_t cap, void *host_counters, uint32_t n,
              const uint32_t *row, const uint32_t *col, const float *w, float min_gain,
              void *stream, uint32_t *assignment, lv_result *res);

}  // namespace lv_host
