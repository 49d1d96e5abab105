// lv_host.h -- host-side engine internals shared by the C ABI units
// (lv_engine.cpp, lv_server.cpp); not part of the public interface.
#pragma once

#include <stdint.h>

#include "../../include/louvain_hip.h"

namespace lv_host {

// lv_graph_create's input checks; *S = fixed-point scale bits of the weights
int check_input(uint32_t n, const uint32_t *row, const uint32_t *col, const float *w, int *S);
// a handle with device buffers for up to n_cap vertices and e_cap entries
int graph_new(int device, uint32_t n_cap, uint64_t e_cap, lv_graph **out);
// upload a checked input that fits the handle's capacity (LV_ERR_NOMEM if not)
int graph_upload(lv_graph *g, uint32_t n, const uint32_t *row, const uint32_t *col,
                 const float *w, int S);

}  // namespace lv_host
