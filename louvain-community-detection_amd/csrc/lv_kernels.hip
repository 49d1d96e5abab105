// lv_kernels.hip -- hand-written CDNA4 (gfx950) kernels of the Louvain engine.
//
// Everything here is integer/byte work plus scalar FP32 gain arithmetic:
// there is no dense contraction, so no MFMA.  Kernels are written for
// 64-wide wavefronts; sub-wave groups of 4/8/16/32 lanes serve low-degree
// vertices, one wave or one workgroup serves a heavier vertex.
//
// Arithmetic contract (shared bit-for-bit with the CPU oracle):
//   * weights, k_v, hash sums, a_c and merged weights are exact integers at
//     a per-graph scale 2^S (int64), rounded once to FP32 when used;
//   * k_v and s_v use the reference's pinned FP32 order: 32 strided
//     sequential partials, then the shuffle-down tree 16,8,4,2,1
//     (modularity.cu:27-49, 400-426);
//   * the gain, move test and modularity are evaluated in FP32 exactly as
//     written in the reference (modularity.cu:149-152, :182, :435), compiled
//     with -ffp-contract=off and IEEE division.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "lv_device.h"
#include "lv_devmath.h"

namespace lv {

// ---------------------------------------------------------------------------
// fills
// ---------------------------------------------------------------------------
__global__ void k_fill_u32(uint32_t *p, uint32_t v, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = v;
}
__global__ void k_fill_u64(unsigned long long *p, unsigned long long v, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = v;
}
// three u32 arrays zeroed by one launch (a pass boundary of a small graph
// is bound by the host's launches)
__global__ void k_zero3_u32(uint32_t *p0, size_t n0, uint32_t *p1, size_t n1, uint32_t *p2,
                            size_t n2) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n0 + n1 + n2; i += stride) {
        if (i < n0) p0[i] = 0u;
        else if (i < n0 + n1) p1[i - n0] = 0u;
        else p2[i - n0 - n1] = 0u;
    }
}
static inline unsigned grid_for(size_t n, unsigned threads, unsigned cap = 65536) {
    size_t g = (n + threads - 1) / threads;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return (unsigned)g;
}
// sums `groups` arrays of kSpread slots into out[0..groups)
// out[gi] = sum of the kSpread slots of group gi (one workgroup of 256)
// (and with `clear`, leaves the slots zero for the next sweep)
__device__ __forceinline__ void collapse_body(unsigned long long *spread, int groups,
                                              unsigned long long *out, bool clear = false) {
    for (int gi = 0; gi < groups; ++gi) {
        unsigned long long x = 0;
        for (uint32_t i = threadIdx.x; i < kSpread; i += 256) {
            x += spread[gi * kSpread + i];
            if (clear) spread[gi * kSpread + i] = 0ull;
        }
        x = wave_sum_u64(x);
        __shared__ unsigned long long part[4];
        if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = x;
        __syncthreads();
        if (threadIdx.x == 0) out[gi] = part[0] + part[1] + part[2] + part[3];
        __syncthreads();
    }
}
__global__ __launch_bounds__(256) void k_collapse(const unsigned long long *spread, int groups,
                                                  unsigned long long *out) {
    collapse_body(const_cast<unsigned long long *>(spread), groups, out);
}
void launch_collapse(hipStream_t s, const unsigned long long *spread, int groups,
                     unsigned long long *out) {
    hipLaunchKernelGGL(k_collapse, dim3(1), dim3(256), 0, s, spread, groups, out);
}

void launch_fill_u32(hipStream_t s, uint32_t *p, uint32_t v, size_t n) {
    if (!n) return;
    hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(n, 256)), dim3(256), 0, s, p, v, n);
}
void launch_zero3_u32(hipStream_t s, uint32_t *p0, size_t n0, uint32_t *p1, size_t n1,
                      uint32_t *p2, size_t n2) {
    const size_t n = n0 + n1 + n2;
    if (!n) return;
    hipLaunchKernelGGL(k_zero3_u32, dim3(grid_for(n, 256)), dim3(256), 0, s, p0, n0, p1, n1, p2,
                       n2);
}
void launch_fill_u64(hipStream_t s, unsigned long long *p, unsigned long long v, size_t n) {
    if (!n) return;
    hipLaunchKernelGGL(k_fill_u64, dim3(grid_for(n, 256)), dim3(256), 0, s, p, v, n);
}

__global__ void k_iota_u32(uint32_t *p, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = (uint32_t)i;
}
void launch_iota_u32(hipStream_t s, uint32_t *p, size_t n) {
    if (!n) return;
    hipLaunchKernelGGL(k_iota_u32, dim3(grid_for(n, 256)), dim3(256), 0, s, p, n);
}

// ---------------------------------------------------------------------------
// batch mode helpers
// ---------------------------------------------------------------------------
// gid[v] = graph whose vertex range [voff[g], voff[g+1]) holds v
__global__ void k_batch_gid(const uint32_t *voff, uint32_t G, uint32_t *gid, uint32_t n) {
    const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n) return;
    uint32_t lo = 0, hi = G;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (voff[mid] <= v) lo = mid;
        else hi = mid;
    }
    gid[v] = lo;
}
void launch_batch_gid(hipStream_t s, const uint32_t *voff, uint32_t G, uint32_t *gid,
                      uint32_t n) {
    if (!n || !G) return;
    hipLaunchKernelGGL(k_batch_gid, dim3((n + 255) / 256), dim3(256), 0, s, voff, G, gid, n);
}
// vertices of graphs that do not aggregate now become their own community, so
// the union contraction leaves those graphs unchanged up to renumbering
__global__ void k_batch_identity(const uint32_t *gid, const uint32_t *gstate, uint32_t keep,
                                 uint32_t *comm, uint32_t *comm_sz, uint32_t n) {
    const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v < n && gstate[gid[v]] != keep) {
        comm[v] = v;
        comm_sz[v] = 1u;
    }
}
void launch_batch_identity(hipStream_t s, const uint32_t *gid, const uint32_t *gstate,
                           uint32_t keep, uint32_t *comm, uint32_t *comm_sz, uint32_t n) {
    if (!n) return;
    hipLaunchKernelGGL(k_batch_identity, dim3((n + 255) / 256), dim3(256), 0, s, gid, gstate, keep,
                       comm, comm_sz, n);
}
__global__ void k_gather_u32(const uint32_t *src, const uint32_t *idx, uint32_t *out,
                             uint32_t count) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) out[i] = src[idx[i]];
}
void launch_gather_u32(hipStream_t s, const uint32_t *src, const uint32_t *idx, uint32_t *out,
                       uint32_t count) {
    if (!count) return;
    hipLaunchKernelGGL(k_gather_u32, dim3((count + 255) / 256), dim3(256), 0, s, src, idx, out,
                       count);
}

// ---------------------------------------------------------------------------
// locality order (single-graph mode, once per pass)
//
// A sweep gathers comm/A/comm_sz at entry endpoints, and endpoints favour
// high-degree vertices.  Renumbering the pass graph so vertices are grouped by
// kernel class puts the hubs' state in one compact range that stays in L2,
// where input-numbered hubs are scattered over the whole array.  Position x
// stands for input id lid[x]; pmap is the inverse.  Rows keep their entry
// order, so every pinned sum and lane assignment is unchanged.
// ---------------------------------------------------------------------------
// lid = order (vertices grouped by class), pmap = its inverse, order -> iota
__global__ void k_relabel_map(uint32_t *order, uint32_t n, uint32_t *lid, uint32_t *pmap) {
    const uint32_t x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x < n) {
        const uint32_t v = order[x];
        lid[x] = v;
        pmap[v] = x;
        order[x] = x;
    }
}
__global__ void k_relabel_deg(const uint32_t *row, const uint32_t *lid, uint32_t n,
                              uint32_t *deg) {
    const uint32_t x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x < n) {
        const uint32_t v = lid[x];
        deg[x] = row[v + 1] - row[v];
    }
}
void launch_relabel_map(hipStream_t s, uint32_t *order, uint32_t n, uint32_t *lid,
                        uint32_t *pmap, const uint32_t *row, uint32_t *deg) {
    if (!n) return;
    hipLaunchKernelGGL(k_relabel_map, dim3((n + 255) / 256), dim3(256), 0, s, order, n, lid, pmap);
    hipLaunchKernelGGL(k_relabel_deg, dim3((n + 255) / 256), dim3(256), 0, s, row, lid, n, deg);
}

// copy the rows of positions [x0, x1) into the new CSR, G lanes per row,
// column ids mapped through pmap; B rounds of loads, then of gathers, in
// flight at once (rounds past the row end reload entry 0: rows here are
// non-empty)
template <int G, int B>
__global__ __launch_bounds__(256) void k_relabel_rows(RelabelView r, uint32_t x0, uint32_t x1) {
    const uint32_t t = blockIdx.x * 256 + threadIdx.x;
    const uint32_t x = x0 + t / G, lane = t % G;
    if (x >= x1) return;
    const uint32_t v = r.lid[x];
    const uint32_t b = r.row[v], d = r.row[v + 1] - b, o = r.row2[x];
    for (uint32_t j0 = lane; j0 < d; j0 += G * B) {
        uint32_t c[B];
        float ww[B];
#pragma unroll
        for (int u = 0; u < B; ++u) {
            const uint32_t j = j0 + (uint32_t)(G * u);
            const uint32_t q = b + (j < d ? j : 0u);
            c[u] = r.col[q];
            ww[u] = r.w[q];
        }
#pragma unroll
        for (int u = 0; u < B; ++u) c[u] = r.pmap[c[u]];
#pragma unroll
        for (int u = 0; u < B; ++u) {
            const uint32_t j = j0 + (uint32_t)(G * u);
            if (j < d) {
                r.col2[o + j] = c[u];
                r.w2[o + j] = ww[u];
            }
        }
    }
}
void launch_relabel_rows(hipStream_t s, const RelabelView &r, uint32_t x0, uint32_t x1,
                         int lanes) {
    if (x1 <= x0) return;
    const uint64_t threads = (uint64_t)(x1 - x0) * (uint64_t)lanes;
    const dim3 grid((unsigned)((threads + 255) / 256));
    if (lanes == 8)
        hipLaunchKernelGGL((k_relabel_rows<8, 4>), grid, dim3(256), 0, s, r, x0, x1);
    else
        hipLaunchKernelGGL((k_relabel_rows<64, 8>), grid, dim3(256), 0, s, r, x0, x1);
}

// new_id under a locality order: rank of lid[c] among the non-empty
// communities in the input numbering (rank = scan over the input numbering)
__global__ void k_scatter_by(const uint32_t *src, const uint32_t *idx, uint32_t n,
                             uint32_t *out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[idx[i]] = src[i];
}
__global__ void k_relabel_newid(const uint32_t *rank, const uint32_t *lid, uint32_t n,
                                uint32_t *new_id) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) new_id[i] = rank[lid[i]];
    else if (i == n) new_id[n] = rank[n];
}
void launch_relabel_new_id(hipStream_t s, const uint32_t *comm_sz, const uint32_t *lid, uint32_t n,
                           uint32_t *flags, uint32_t *rank, uint32_t *new_id, uint32_t *scan_tmp) {
    if (!n) return;
    hipLaunchKernelGGL(k_scatter_by, dim3((n + 255) / 256), dim3(256), 0, s, comm_sz, lid, n,
                       flags);
    launch_scan_u32(s, flags, 1, n, rank, scan_tmp);
    hipLaunchKernelGGL(k_relabel_newid, dim3((n + 256) / 256), dim3(256), 0, s, rank, lid, n,
                       new_id);
}

// ---------------------------------------------------------------------------
// degree classification: once per pass (the reference re-partitions the
// whole vertex set 7 times per sweep, modularity.cu:468-469)
// ---------------------------------------------------------------------------
// class of a row; rows longer than heavy_min (>= 319, so always bucket 7)
// take the split heavy path
__device__ __forceinline__ int row_class(uint32_t d, uint32_t heavy_min) {
    return d > heavy_min ? kClassHeavy : vertex_class(d);
}

__device__ __forceinline__ bool listed(const uint32_t *gid, const uint32_t *gstate, uint32_t v) {
    return !gid || gstate[gid[v]] == kGraphActive;
}
__global__ __launch_bounds__(256) void k_classify_count(const uint32_t *row, uint32_t n,
                                                        uint32_t heavy_min, const uint32_t *gid,
                                                        const uint32_t *gstate, uint32_t *counts,
                                                        unsigned long long *degsum) {
    __shared__ uint32_t hc[kNumClasses];
    __shared__ unsigned long long hd[kNumClasses];
    if (threadIdx.x < kNumClasses) {
        hc[threadIdx.x] = 0;
        hd[threadIdx.x] = 0;
    }
    __syncthreads();
    for (uint32_t v = blockIdx.x * 256 + threadIdx.x; v < n; v += gridDim.x * 256) {
        if (!listed(gid, gstate, v)) continue;
        uint32_t d = row[v + 1] - row[v];
        int c = row_class(d, heavy_min);
        atomicAdd(&hc[c], 1u);
        atomicAdd(&hd[c], (unsigned long long)d);
    }
    __syncthreads();
    if (threadIdx.x < kNumClasses) {
        if (hc[threadIdx.x]) atomicAdd(&counts[threadIdx.x], hc[threadIdx.x]);
        if (hd[threadIdx.x]) atomicAdd(&degsum[threadIdx.x], hd[threadIdx.x]);
    }
}

// cursor[c] starts at the class offset; order within a class is not
// semantic (every vertex of a step decides from the same snapshot)
constexpr uint32_t kScatterPer = 16;
// Vertices into class order. A workgroup takes 256 * run consecutive vertices
// (run a multiple of 16): it counts them per class, reserves each class's
// range with one global cursor add, then places them (a second pass over the
// same rows), so every class holds runs of nearby ids — the order the
// locality renumbering inherits.
__global__ __launch_bounds__(256) void k_classify_scatter(const uint32_t *row, uint32_t n,
                                                          uint32_t heavy_min, const uint32_t *gid,
                                                          const uint32_t *gstate, uint32_t run,
                                                          uint32_t *cursor, uint32_t *order) {
    __shared__ uint32_t hc[kNumClasses];
    __shared__ uint32_t base[kNumClasses];
    if (threadIdx.x < kNumClasses) hc[threadIdx.x] = 0;
    __syncthreads();
    const uint32_t v0 = blockIdx.x * run * 256u + threadIdx.x;
    // 16 rows' lengths per round are loaded before their class is used
    auto lengths = [&](uint32_t q0, int (&c)[16]) {
        uint32_t a[16], b[16];
#pragma unroll
        for (int u = 0; u < 16; ++u) {
            const uint32_t v = v0 + 256u * (q0 + u), vv = v < n ? v : n - 1u;
            a[u] = row[vv];
            b[u] = row[vv + 1];
        }
#pragma unroll
        for (int u = 0; u < 16; ++u) {
            const uint32_t v = v0 + 256u * (q0 + u);
            c[u] = v < n && listed(gid, gstate, v) ? row_class(b[u] - a[u], heavy_min) : -1;
        }
    };
    for (uint32_t q0 = 0; q0 < run; q0 += 16) {
        int c[16];
        lengths(q0, c);
#pragma unroll
        for (int u = 0; u < 16; ++u)
            if (c[u] >= 0) atomicAdd(&hc[c[u]], 1u);
    }
    __syncthreads();
    if (threadIdx.x < kNumClasses) {
        const uint32_t h = hc[threadIdx.x];
        base[threadIdx.x] = h ? atomicAdd(&cursor[threadIdx.x], h) : 0u;
        hc[threadIdx.x] = 0;
    }
    __syncthreads();
    for (uint32_t q0 = 0; q0 < run; q0 += 16) {
        int c[16];
        lengths(q0, c);
#pragma unroll
        for (int u = 0; u < 16; ++u)
            if (c[u] >= 0) order[base[c[u]] + atomicAdd(&hc[c[u]], 1u)] = v0 + 256u * (q0 + u);
    }
}

void launch_classify_count(hipStream_t s, const uint32_t *row, uint32_t n, uint32_t heavy_min,
                           const uint32_t *gid, const uint32_t *gstate, uint32_t *counts,
                           unsigned long long *degsum) {
    if (!n) return;
    hipLaunchKernelGGL(k_classify_count, dim3(grid_for(n, 256, 4096)), dim3(256), 0, s, row, n,
                       heavy_min, gid, gstate, counts, degsum);
}
void launch_classify_scatter(hipStream_t s, const uint32_t *row, uint32_t n, uint32_t heavy_min,
                             const uint32_t *gid, const uint32_t *gstate, uint32_t run,
                             uint32_t *cursor, uint32_t *order) {
    if (!n) return;
    run = run < 16 ? 16u : run & ~15u;
    const uint32_t per = 256u * run;
    hipLaunchKernelGGL(k_classify_scatter, dim3((uint32_t)(((uint64_t)n + per - 1) / per)),
                       dim3(256), 0, s, row, n, heavy_min, gid, gstate, run, cursor, order);
}

// ---------------------------------------------------------------------------
// pinned-order sums k_v / s_v (modularity.cu:27-49 and 400-426)
//
// G lanes per vertex.  For G < 32 the class guarantees deg <= G, and the
// reference's 32-lane tree restricted to G lanes gives the same bits (the
// missing partials are +0).  Lane l sums entries l, l+32, ... sequentially;
// loads are batched B rounds ahead (B = 1 for G < 32), adds stay in order.
// Every load of a batch is issued before the first gather and every gather
// before the first add: rounds past the row end load the row's first entry
// (always in bounds) and are skipped by the adds.  At pass init every
// community is a singleton (comm[u] = u), so the masked sum compares columns
// directly (`ident`) and gathers nothing.
// ---------------------------------------------------------------------------
// One workgroup of 256 threads: vertices [bx * 256/G, (bx + 1) * 256/G) of
// verts.
template <int G, bool MASKED, int B>
__device__ __forceinline__ void r1_group(const DevView &g, const uint32_t *verts, int32_t vb,
                                         uint32_t count, int init_state, int ident,
                                         unsigned long long *acc, uint32_t bx) {
    constexpr int GPB = 256 / G;
    const int lane = threadIdx.x % G;
    const uint32_t vi = bx * GPB + threadIdx.x / G;
    unsigned long long fixed = 0;
    uint32_t slot = 0;
    bool valid = false;
    const uint32_t v = vi < count ? (vb >= 0 ? (uint32_t)vb + vi : verts[vi]) : 0u;
    if (vi < count && v_live(g, v)) {
        const uint32_t base = g.row[v];
        const uint32_t d = g.row[v + 1] - base;
        const uint32_t cv = MASKED ? g.comm[v] : 0u;
        const float *wr = g.w + base;
        const uint32_t *cr = g.col + base;
        float p = 0.0f;
        for (uint32_t j = lane; j < d; j += B * 32) {
            float x[B];
            if (MASKED) {
                // columns, then communities, then the weights of the entries
                // inside v's community only (the other lanes issue no load)
                uint32_t c[B];
                bool keep[B];
#pragma unroll
                for (int u = 0; u < B; ++u) c[u] = cr[j + 32u * u < d ? j + 32u * u : 0u];
                if (!ident) {  // ident: every community is a singleton, comm[u] = u
#pragma unroll
                    for (int u = 0; u < B; ++u) c[u] = g.comm[c[u]];
                }
#pragma unroll
                for (int u = 0; u < B; ++u) {
                    keep[u] = j + 32u * u < d && c[u] == cv;
                    x[u] = 0.0f;
                    if (keep[u]) x[u] = wr[j + 32u * u];
                }
#pragma unroll
                for (int u = 0; u < B; ++u)
                    if (keep[u]) p += x[u];
            } else {
#pragma unroll
                for (int u = 0; u < B; ++u) x[u] = wr[j + 32u * u < d ? j + 32u * u : 0u];
#pragma unroll
                for (int u = 0; u < B; ++u)
                    if (j + 32u * u < d) p += x[u];
            }
        }
#pragma unroll
        for (int o = G / 2; o > 0; o >>= 1) p += __shfl_down(p, o, G);
        if (lane == 0) {
            fixed = f2fix(p, g.scale);
            slot = acc_slot(g, v);
            valid = true;
            if (!MASKED) {
                g.k[v] = p;
                if (init_state) {
                    init_state_of(g, v, fixed);
                }
            }
        }
    }
    block_add(g, acc, slot, fixed, valid);  // exact, order-free
}

// Pinned-order sums of rows of at most L entries (R = 1) or at most 32 R
// entries (L = 32), with Q lanes per vertex: lane q holds virtual lanes q,
// q + Q, ..., q + L - Q of the reference's 32-lane order.  Virtual lane l
// adds positions l, l + 32, ... of the row sequentially; tree levels
// L/2 .. Q pair two virtual lanes of one lane, the levels below Q are
// shuffles among the Q lanes, and the levels at or above L add +0 (the row
// has no entry there) and are skipped.  The additions and their association
// are those of the 32-lane tree, so are the bits, with 32/Q times as many
// vertices per wave in flight: these sums are bound by the latency of their
// dependent loads (row, column, community, weight), not by their bytes.
template <int L, int Q, int R, bool MASKED>
__device__ __forceinline__ void r1_narrow(const DevView &g, const uint32_t *verts, int32_t vb,
                                          uint32_t count, int init_state, int ident,
                                          unsigned long long *acc, uint32_t bx) {
    static_assert(L % Q == 0 && L <= 32 && (L & (L - 1)) == 0 && (R == 1 || L == 32),
                  "lane shape");
    constexpr int V = L / Q, VPB = 256 / Q;
    const int q = threadIdx.x % Q;
    const uint32_t vi = bx * VPB + threadIdx.x / Q;
    unsigned long long fixed = 0;
    uint32_t slot = 0;
    bool valid = false;
    const uint32_t v = vi < count ? (vb >= 0 ? (uint32_t)vb + vi : verts[vi]) : 0u;
    if (vi < count && v_live(g, v)) {
        const uint32_t base = g.row[v];
        const uint32_t d = g.row[v + 1] - base;
        const uint32_t cv = MASKED ? g.comm[v] : 0u;
        // every load before the first add: columns (past the row end the
        // row's first entry stands in), communities, then the weights of the
        // kept entries only
        bool keep[R][V];
        uint32_t c[R][V];
        float x[R][V];
#pragma unroll
        for (int r = 0; r < R; ++r)
#pragma unroll
            for (int i = 0; i < V; ++i) {
                const uint32_t pos = (uint32_t)(q + Q * i + 32 * r);
                keep[r][i] = pos < d;
                c[r][i] = MASKED ? g.col[base + (keep[r][i] ? pos : 0u)] : 0u;
            }
        if (MASKED && !ident) {  // ident: every community is a singleton, comm[u] = u
#pragma unroll
            for (int r = 0; r < R; ++r)
#pragma unroll
                for (int i = 0; i < V; ++i) c[r][i] = g.comm[c[r][i]];
        }
#pragma unroll
        for (int r = 0; r < R; ++r)
#pragma unroll
            for (int i = 0; i < V; ++i) {
                if (MASKED) keep[r][i] = keep[r][i] && c[r][i] == cv;
                x[r][i] = 0.0f;
                if (keep[r][i]) x[r][i] = g.w[base + (uint32_t)(q + Q * i + 32 * r)];
            }
        float part[V];
#pragma unroll
        for (int i = 0; i < V; ++i) {
            part[i] = 0.0f;
#pragma unroll
            for (int r = 0; r < R; ++r)
                if (keep[r][i]) part[i] += x[r][i];
        }
#pragma unroll
        for (int o = L / 2; o >= Q; o >>= 1)
#pragma unroll
            for (int i = 0; i < o / Q; ++i) part[i] += part[i + o / Q];
        float p = part[0];
#pragma unroll
        for (int o = Q / 2; o > 0; o >>= 1) p += __shfl_down(p, o, Q);
        if (q == 0) {
            fixed = f2fix(p, g.scale);
            slot = acc_slot(g, v);
            valid = true;
            if (!MASKED) {
                g.k[v] = p;
                if (init_state) {
                    init_state_of(g, v, fixed);
                }
            }
        }
    }
    block_add(g, acc, slot, fixed, valid);  // exact, order-free
}

// Masked sum of rows of more than 32 entries, 32 lanes per vertex: the
// columns are read four per lane with 16-byte loads from the 4-aligned
// position at or below the row start (BK blocks of 128 positions per round),
// the loading lane gathers the communities and loads the kept weights, and
// the per-entry values are exchanged through LDS so that lane p adds entries
// p, p+32, p+64, ... of the row in order — the same pinned sequence as
// r1_group<32, true, B>. Positions outside the row contribute nothing (their
// column stands in as v itself, always in bounds); non-kept entries add +0,
// which leaves the partial unchanged. The column arrays carry 4 entries of
// padding, so a 16-byte load that starts inside the last row stays inside
// the allocation.
// One workgroup of 256 threads: vertices [8 bx, 8 bx + 8) of verts.
template <int BK>
__device__ __forceinline__ void r1_group32(const DevView &g, const uint32_t *verts, int32_t vb,
                                           uint32_t count, int ident, unsigned long long *acc,
                                           uint32_t bx) {
    __shared__ float xs[8][BK * 128];
    const int lane = threadIdx.x & 31, grp = threadIdx.x >> 5;
    float *xg = xs[grp];
    const uint32_t vi = bx * 8 + grp;
    unsigned long long fixed = 0;
    uint32_t slot = 0;
    bool valid = false;
    const uint32_t v = vi < count ? (vb >= 0 ? (uint32_t)vb + vi : verts[vi]) : 0u;
    if (vi < count && v_live(g, v)) {
        const uint32_t base = g.row[v], end = g.row[v + 1];
        const uint32_t cv = g.comm[v];
        const uint32_t a0 = base & ~3u, rot = (lane + (base & 3u)) & 31u;
        float p = 0.0f;
        for (uint32_t a = a0; a < end; a += BK * 128u) {
            uint4 c4[BK];
#pragma unroll
            for (int k = 0; k < BK; ++k) {
                const uint32_t q = a + 128u * k + 4u * lane;
                c4[k] = *reinterpret_cast<const uint4 *>(g.col + (q < end ? q : a0));
            }
            uint32_t c[BK][4];
#pragma unroll
            for (int k = 0; k < BK; ++k) {
                const uint32_t q = a + 128u * k + 4u * lane;
                c[k][0] = q + 0u >= base && q + 0u < end ? c4[k].x : v;
                c[k][1] = q + 1u >= base && q + 1u < end ? c4[k].y : v;
                c[k][2] = q + 2u >= base && q + 2u < end ? c4[k].z : v;
                c[k][3] = q + 3u >= base && q + 3u < end ? c4[k].w : v;
            }
            if (!ident) {  // ident: every community is a singleton, comm[u] = u
#pragma unroll
                for (int k = 0; k < BK; ++k)
#pragma unroll
                    for (int i = 0; i < 4; ++i) c[k][i] = g.comm[c[k][i]];
            }
            float x[BK][4];
#pragma unroll
            for (int k = 0; k < BK; ++k)
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const uint32_t q = a + 128u * k + 4u * lane + i;
                    // v's own stand-in never passes: the position test fails
                    const bool keep = q >= base && q < end && c[k][i] == cv;
                    x[k][i] = 0.0f;
                    if (keep) x[k][i] = g.w[q];
                }
#pragma unroll
            for (int k = 0; k < BK; ++k)
                *reinterpret_cast<float4 *>(xg + 128 * k + 4 * lane) =
                    make_float4(x[k][0], x[k][1], x[k][2], x[k][3]);
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
            for (int k = 0; k < BK; ++k)
#pragma unroll
                for (int m = 0; m < 4; ++m) p += xg[128 * k + 32 * m + rot];
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }
#pragma unroll
        for (int o = 16; o > 0; o >>= 1) p += __shfl_down(p, o, 32);
        if (lane == 0) {
            fixed = f2fix(p, g.scale);
            slot = acc_slot(g, v);
            valid = true;
        }
    }
    block_add(g, acc, slot, fixed, valid);  // exact, order-free
}

// ---------------------------------------------------------------------------
// sum over communities of round32(a_c)^2, exact at scale 2^2S, as four
// 32-bit limbs accumulated in u64 (modularity.cu:431-432, square<float>)
// ---------------------------------------------------------------------------
// workgroup blk of nblk (256 threads)
__device__ __forceinline__ void sumsq_body(const DevView &g, const uint32_t *verts,
                                           uint32_t count, unsigned long long *limbs,
                                           uint32_t blk, uint32_t nblk) {
    const int twoS = 2 * g.S;
    const uint32_t stride = acc_stride(g);
    for (uint32_t base_i = blk * 256; base_i < count; base_i += nblk * 256) {
        const uint32_t i = base_i + threadIdx.x;
        unsigned long long l0 = 0, l1 = 0, l2 = 0, l3 = 0;
        uint32_t slot = 0;
        bool valid = false;
        const uint32_t c = i < count ? vert_at(g, verts, i) : 0u;
        const unsigned long long a = (i < count && v_live(g, c)) ? g.A[c] : 0ull;
        if (a) {
        valid = true;
        slot = acc_slot(g, c);
        float af = fix2f(a, g.S);
        float a2 = af * af;
        uint32_t u = __float_as_uint(a2);
        int ef = (int)((u >> 23) & 0xff);
        uint32_t man = u & 0x7fffffu;
        int e;
        if (ef == 0) {
            e = -149;
        } else {
            man |= 0x800000u;
            e = ef - 150;
        }
        int sh = e + twoS;  // >= 0: a2 is a multiple of 2^-2S
        unsigned long long lo, hi;
        if (sh <= 0) {  // sh < 0 cannot occur: a2 is a multiple of 2^-2S
            lo = (unsigned long long)man >> (-sh);
            hi = 0ull;
        } else if (sh < 64) {
            lo = (unsigned long long)man << sh;
            hi = (unsigned long long)man >> (64 - sh);
        } else {
            lo = 0ull;
            hi = (unsigned long long)man << (sh - 64);
        }
        l0 = lo & 0xffffffffull;
        l1 = lo >> 32;
        l2 = hi & 0xffffffffull;
        l3 = hi >> 32;
        }
        block_add(g, limbs, slot, l0, valid);
        block_add(g, limbs + stride, slot, l1, valid);
        block_add(g, limbs + 2 * stride, slot, l2, valid);
        block_add(g, limbs + 3 * stride, slot, l3, valid);
    }
}

uint32_t sumsq_blocks(uint32_t count) { return count ? grid_for(count, 256, 8192) : 0u; }

// Every class group's pinned sums in one launch, longest rows first:
// workgroups [0, bend[5]) take group 5 (rows of more than 84 entries, 32
// lanes), [bend[5], bend[4]) group 4 (33-84 entries), then groups 3 (17-32),
// 2 (9-16), 1 (5-8) and 0 (at most 4 entries), so the longest workgroups
// are dispatched first and the short ones fill in behind them.  Groups 0-4
// take 1, 2, 4, 8 and 8 lanes per vertex (r1_narrow).  The groups read the
// same snapshot and add into order-free exact counters, so running them
// side by side in one grid changes no result; on small graphs it replaces
// several dependent launches of a few microseconds each by one.
// CATCHUP: workgroups in front apply the last bucket step's moves to the
// copy of the state that the sums do not read (two copies); a separate
// instantiation, so the sums' code without them is what it was (the catch-up
// path changes their register allocation: R-MAT-24 ran 1.5 % slower).
template <bool MASKED, bool CATCHUP>
__global__ __launch_bounds__(256) void k_r1_groups(DevView g, R1Groups L, int init_state,
                                                   int ident, unsigned long long *acc) {
    if (halted(g)) return;  // the pass loop stopped earlier in this launch
    if (CATCHUP && blockIdx.x < g.cu_nblk) return catch_up(g, blockIdx.x);
    const uint32_t b = CATCHUP ? blockIdx.x - g.cu_nblk : blockIdx.x;
    if (b >= L.bend[0]) {  // behind the groups: the sum of a_c^2 of the same snapshot
        DevView gs = g;
        gs.vbase = L.sq_vbase;
        sumsq_body(gs, L.sq_verts, L.sq_count, L.limbs, b - L.bend[0], L.sq_blocks);
        return;
    }
    if (b < L.bend[5]) {
        if (MASKED)
            r1_group32<2>(g, L.verts[5], L.vbase[5], L.count[5], ident, acc, b);
        else
            r1_group<32, false, 8>(g, L.verts[5], L.vbase[5], L.count[5], init_state, 0, acc,
                                   b);
    } else if (b < L.bend[4]) {
        r1_narrow<32, 8, 3, MASKED>(g, L.verts[4], L.vbase[4], L.count[4], init_state, ident,
                                    acc, b - L.bend[5]);
    } else if (b < L.bend[3]) {
        r1_narrow<32, 8, 1, MASKED>(g, L.verts[3], L.vbase[3], L.count[3], init_state, ident,
                                    acc, b - L.bend[4]);
    } else if (b < L.bend[2]) {
        r1_narrow<16, 4, 1, MASKED>(g, L.verts[2], L.vbase[2], L.count[2], init_state, ident,
                                    acc, b - L.bend[3]);
    } else if (b < L.bend[1]) {
        r1_narrow<8, 2, 1, MASKED>(g, L.verts[1], L.vbase[1], L.count[1], init_state, ident,
                                   acc, b - L.bend[2]);
    } else {
        r1_narrow<4, 1, 1, MASKED>(g, L.verts[0], L.vbase[0], L.count[0], init_state, ident,
                                   acc, b - L.bend[1]);
    }
}

uint32_t catchup_blocks(uint32_t count, uint32_t threads);
void launch_r1_groups(hipStream_t s, const DevView &g, const R1Groups &L, int masked,
                      int init_state, unsigned long long *acc) {
    DevView gv = g;
    gv.cu_nblk = g.cu_count ? catchup_blocks(g.cu_count, 256) : 0u;
    const uint32_t nb = gv.cu_nblk + L.bend[0] + (masked ? L.sq_blocks : 0u);
    if (!nb) return;
    if (masked && gv.cu_nblk)
        hipLaunchKernelGGL((k_r1_groups<true, true>), dim3(nb), dim3(256), 0, s, gv, L,
                           init_state, masked == 2, acc);
    else if (masked)
        hipLaunchKernelGGL((k_r1_groups<true, false>), dim3(nb), dim3(256), 0, s, gv, L,
                           init_state, masked == 2, acc);
    else  // pass init: no catch-up (cu_count is 0)
        hipLaunchKernelGGL((k_r1_groups<false, false>), dim3(nb), dim3(256), 0, s, gv, L,
                           init_state, 0, acc);
}


// ---------------------------------------------------------------------------
// move kernels (computeMoveWarp / computeMoveBlock / computeMoveGlobal,
// modularity.cu:95-380), canonical semantics:
//   * each vertex of the step decides from the step-start snapshot;
//   * virtual lane l of T visits entries l, l+T, ... ; a lane keeps its FIRST
//     maximum (strict >);
//   * buckets 1-5: the gain of an entry in round r sees the sums of entries
//     of rounds <= r; destination = best (gain desc, comm asc) lane's comm;
//   * buckets 6-7: complete sums; the 128-lane tree of modularity.cu:265-275
//     picks the winner lane; destination = comm of that lane's LAST visited
//     entry (graph.comm[e], modularity.cu:183/279);
//   * move iff G - comm_e_w / m > 0 (FP32), else stay; no candidate: stay.
// ---------------------------------------------------------------------------

// The previous sweep's stop rule in workgroup 0 of a sweep's first move
// kernel (any workgroup size up to 1024); defined with k_collapse_decide.
__device__ void decide_front(const SweepTail &t, int S);

// Workgroups in front of a move kernel's own: the previous sweep's stop rule
// (g.tail), then the previous step's catch-up (cu_nblk).  True for those;
// otherwise blk is the kernel's own workgroup index.
__device__ __forceinline__ bool front_work(const DevView &g, uint32_t &blk) {
    blk = blockIdx.x;
    if (g.tail) {
        if (blk == 0) {
            decide_front(*g.tail, g.S);
            return true;
        }
        --blk;
    }
    if (blk < g.cu_nblk) {
        catch_up(g, blk);
        return true;
    }
    blk -= g.cu_nblk;
    return false;
}

// buckets 1-5: T lanes per vertex (T = reference lanes), ROUNDS rounds,
// per-group LDS open-addressing table of CAP slots with exact u64 sums.
template <int T, int ROUNDS, int CAP>
__global__ __launch_bounds__(256) void k_move_small(DevView g, const uint32_t *verts,
                                                    uint32_t count) {
    if (halted(g)) return;  // the pass loop stopped earlier in this launch
    uint32_t bid;
    if (front_work(g, bid)) return;
    constexpr int GPB = 256 / T;
    constexpr int LC = Log2<CAP>::value;
    __shared__ uint32_t skey[GPB * CAP];
    __shared__ unsigned long long sval[GPB * CAP];
    const int grp = threadIdx.x / T, lane = threadIdx.x % T;
    uint32_t *key = skey + grp * CAP;
    unsigned long long *val = sval + grp * CAP;
    for (int i = lane; i < CAP; i += T) {
        key[i] = kEmpty;
        val[i] = 0ull;
    }
    // groups past the end read the last vertex (count >= 1) and write nothing;
    // the vertex state loads are unconditional so none of them waits on a branch
    const uint32_t vi = bid * GPB + grp;
    const uint32_t v = vert_at(g, verts, vi < count ? vi : count - 1u);
    const bool active = vi < count && v_live(g, v);
    // a workgroup with no live vertex (past the end, or only frozen graphs of
    // a batch) skips the loads below
    if (!__syncthreads_or(active)) return;
    const float m = v_m(g, v);
    const uint32_t base = g.row[v], rend = g.row[v + 1];
    const uint32_t d = active ? rend - base : 0u;
    const uint32_t cv = g.comm[v];
    const uint32_t szv = g.comm_sz[cv];
    const float kv = g.k[v];
    const float acv = fix2f(g.A[cv], g.S);
    float best_g = -INFINITY;
    uint32_t best_c = kEmpty;
    unsigned long long cew = 0;
    // every round's loads are issued before the first insert, all of them
    // unconditionally (positions past the row end, or of an inactive group,
    // load entry `base`, which exists: this class has entries), so they are
    // all in flight at once
    uint32_t er[ROUNDS], cr[ROUNDS];
    float wr[ROUNDS];
    bool hs[ROUNDS];
#pragma unroll
    for (int r = 0; r < ROUNDS; ++r) {
        const uint32_t j = (uint32_t)(r * T + lane);
        hs[r] = active && j < d;
        const uint32_t e = base + (hs[r] ? j : 0u);
        er[r] = g.col[e];
        wr[r] = g.w[e];
    }
#pragma unroll
    for (int r = 0; r < ROUNDS; ++r) cr[r] = g.comm[er[r]];
#pragma unroll
    for (int r = 0; r < ROUNDS; ++r) {
        const bool has = hs[r];
        uint32_t e = 0, ce = 0, slot = 0;
        if (has) {
            e = er[r];
            ce = cr[r];
            unsigned long long wf = f2fix(wr[r], g.scale);
            slot = hslot<LC>(ce);
            for (int probe = 0; probe < CAP; ++probe) {
                uint32_t prev = atomicCAS(&key[slot], kEmpty, ce);
                if (prev == kEmpty || prev == ce) break;
                slot = (slot + 1) & (CAP - 1);
            }
            atomicAdd(&val[slot], wf);
            if (ce == cv && e != v) cew += wf;
        }
        __syncthreads();
        if (has) {
            // isLowerIdx (modularity.cu:85-93)
            const bool cand = lower_idx_ok(g, szv, ce, cv);
            if (cand) {
                float gg = gain_of(fix2f(val[slot], g.S), m, kv, acv, fix2f(g.A[ce], g.S));
                if (gg > best_g) {
                    best_g = gg;
                    best_c = ce;
                }
            }
        }
        if (ROUNDS > 1) __syncthreads();
    }
#pragma unroll
    for (int o = T / 2; o > 0; o >>= 1) {
        float og = __shfl_xor(best_g, o, T);
        uint32_t oc = __shfl_xor(best_c, o, T);
        unsigned long long ow = __shfl_xor(cew, o, T);
        if (better(og, oc, best_g, best_c, g.lid)) {
            best_g = og;
            best_c = oc;
        }
        cew += ow;
    }
    if (active && lane == 0) {
        uint32_t nc = cv;
        if (best_g != -INFINITY) {
            float cewf = fix2f(cew, g.S);
            if (best_g - (cewf / m) > 0.0f) nc = best_c;
        }
        commit_move(g, v, kv, cv, nc);
    }
}

// Buckets 1-5 with Q lanes per vertex instead of the reference's T: lane q
// holds virtual lanes q, q + Q, ..., q + T - Q, and a wave holds 64/Q
// vertices.  Round r inserts the entries at positions r T + l of every
// virtual lane l into the vertex's LDS table, then (after a wave-level
// barrier: a vertex's lanes share one wave) every virtual lane scores its
// round-r entry on the sums of rounds <= r and keeps its first maximum —
// exactly the T-lane kernel's per-lane state.  The winner over the virtual
// lanes (gain desc, community asc) is a total order, so folding a lane's
// virtual lanes before the shuffles changes nothing.  These kernels are bound
// by the latency of their dependent loads, so more vertices per wave is
// what makes them faster.
template <int T, int Q, int ROUNDS, int CAP>
__global__ __launch_bounds__(256) void k_move_narrow(DevView g, const uint32_t *verts,
                                                     uint32_t count) {
    if (halted(g)) return;  // the pass loop stopped earlier in this launch
    uint32_t bid;
    if (front_work(g, bid)) return;
    static_assert(T % Q == 0 && 64 % Q == 0 && CAP > T * ROUNDS / 2, "lane shape");
    constexpr int V = T / Q, GPB = 256 / Q, LC = Log2<CAP>::value;
    __shared__ uint32_t skey[GPB * CAP];
    __shared__ unsigned long long sval[GPB * CAP];
    const int grp = threadIdx.x / Q, q = threadIdx.x % Q;
    uint32_t *key = skey + grp * CAP;
    unsigned long long *val = sval + grp * CAP;
    for (int i = q; i < CAP; i += Q) {
        key[i] = kEmpty;
        val[i] = 0ull;
    }
    // groups past the end read the last vertex (count >= 1) and write nothing;
    // the vertex state loads are unconditional so none of them waits on a branch
    const uint32_t vi = bid * GPB + grp;
    const uint32_t v = vert_at(g, verts, vi < count ? vi : count - 1u);
    const bool active = vi < count && v_live(g, v);
    // a workgroup with no live vertex (past the end, or only frozen graphs of
    // a batch) skips the loads below; the barrier also orders the table
    // clearing before the inserts
    if (!__syncthreads_or(active)) return;
    const float m = v_m(g, v);
    const uint32_t base = g.row[v], rend = g.row[v + 1];
    const uint32_t d = active ? rend - base : 0u;
    const uint32_t cv = g.comm[v];
    const uint32_t szv = g.comm_sz[cv];
    const float kv = g.k[v];
    const float acv = fix2f(g.A[cv], g.S);
    float bg[V];
    uint32_t bc[V];
#pragma unroll
    for (int i = 0; i < V; ++i) {
        bg[i] = -INFINITY;
        bc[i] = kEmpty;
    }
    unsigned long long cew = 0;
    // every round's loads before the first insert, unconditionally
    // (positions past the row end, or of an inactive vertex, load entry
    // `base`, which exists: this class has entries)
    uint32_t er[ROUNDS][V], cr[ROUNDS][V];
    float wr[ROUNDS][V];
    bool hs[ROUNDS][V];
#pragma unroll
    for (int r = 0; r < ROUNDS; ++r)
#pragma unroll
        for (int i = 0; i < V; ++i) {
            const uint32_t j = (uint32_t)(r * T + q + Q * i);
            hs[r][i] = active && j < d;
            const uint32_t e = base + (hs[r][i] ? j : 0u);
            er[r][i] = g.col[e];
            wr[r][i] = g.w[e];
        }
#pragma unroll
    for (int r = 0; r < ROUNDS; ++r)
#pragma unroll
        for (int i = 0; i < V; ++i) cr[r][i] = g.comm[er[r][i]];
#pragma unroll
    for (int r = 0; r < ROUNDS; ++r) {
        uint32_t sl[V];
#pragma unroll
        for (int i = 0; i < V; ++i) {
            sl[i] = 0;
            if (!hs[r][i]) continue;
            const uint32_t ce = cr[r][i];
            const unsigned long long wf = f2fix(wr[r][i], g.scale);
            uint32_t slot = hslot<LC>(ce);
            for (int probe = 0; probe < CAP; ++probe) {
                const uint32_t prev = atomicCAS(&key[slot], kEmpty, ce);
                if (prev == kEmpty || prev == ce) break;
                slot = (slot + 1) & (CAP - 1);
            }
            atomicAdd(&val[slot], wf);
            if (ce == cv && er[r][i] != v) cew += wf;
            sl[i] = slot;
        }
        wave_sync_lds();  // round r's inserts before its scores
#pragma unroll
        for (int i = 0; i < V; ++i) {
            if (!hs[r][i]) continue;
            const uint32_t ce = cr[r][i];
            // isLowerIdx (modularity.cu:85-93)
            if (lower_idx_ok(g, szv, ce, cv)) {
                const float gg =
                    gain_of(fix2f(val[sl[i]], g.S), m, kv, acv, fix2f(g.A[ce], g.S));
                if (gg > bg[i]) {  // the virtual lane's first maximum
                    bg[i] = gg;
                    bc[i] = ce;
                }
            }
        }
        if (ROUNDS > 1) wave_sync_lds();  // round r's scores before round r+1's inserts
    }
    float best_g = bg[0];
    uint32_t best_c = bc[0];
#pragma unroll
    for (int i = 1; i < V; ++i)
        if (better(bg[i], bc[i], best_g, best_c, g.lid)) {
            best_g = bg[i];
            best_c = bc[i];
        }
#pragma unroll
    for (int o = Q / 2; o > 0; o >>= 1) {
        const float og = __shfl_xor(best_g, o, Q);
        const uint32_t oc = __shfl_xor(best_c, o, Q);
        const unsigned long long ow = __shfl_xor(cew, o, Q);
        if (better(og, oc, best_g, best_c, g.lid)) {
            best_g = og;
            best_c = oc;
        }
        cew += ow;
    }
    if (active && q == 0) {
        uint32_t nc = cv;
        if (best_g != -INFINITY) {
            const float cewf = fix2f(cew, g.S);
            if (best_g - (cewf / m) > 0.0f) nc = best_c;
        }
        commit_move(g, v, kv, cv, nc);
    }
}

// bucket 6 (84 < deg <= 319): one wave64 per vertex, LDS table per wave.
// Physical lane p visits entries p, p+64, ...: even rounds belong to virtual
// lane p, odd rounds to virtual lane p+64.
#ifndef LV_WAVECAP
#define LV_WAVECAP 400
#endif
constexpr int kWaveCap = LV_WAVECAP;  // slots per wave (at most 319 distinct keys)
constexpr int kWaveMaxPer = 5;        // ceil(319 / 64)
// kWaveWpb waves (vertices) per workgroup: 1, 2 or 4 (a tuning choice)
template <int kWaveWpb>
__global__ __launch_bounds__(64 * kWaveWpb) void k_move_wave(DevView g, const uint32_t *verts,
                                                             uint32_t count) {
    if (halted(g)) return;  // the pass loop stopped earlier in this launch
    if (blockIdx.x < g.cu_nblk) return catch_up(g, blockIdx.x);
    const uint32_t bid = blockIdx.x - g.cu_nblk;
    __shared__ uint32_t skey[kWaveWpb * kWaveCap];
    __shared__ unsigned long long sval[kWaveWpb * kWaveCap];
    __shared__ uint16_t sdist[kWaveWpb][320];
    __shared__ uint32_t sndist[kWaveWpb];
    __shared__ unsigned long long scew[kWaveWpb];
    const int wv = threadIdx.x >> 6, p = threadIdx.x & 63;
    uint32_t *key = skey + wv * kWaveCap;
    unsigned long long *val = sval + wv * kWaveCap;
    for (int i = p; i < kWaveCap; i += 64) {
        key[i] = kEmpty;
        val[i] = 0ull;
    }
    if (p == 0) {
        sndist[wv] = 0;
        scew[wv] = 0;
    }
    // waves past the end read the last vertex (count >= 1) and write nothing
    const uint32_t vi = bid * kWaveWpb + wv;
    const uint32_t v = vert_at(g, verts, vi < count ? vi : count - 1u);
    const bool active = vi < count && v_live(g, v);
    // a workgroup with no live vertex (past the end, or only frozen graphs of
    // a batch) skips the loads below
    if (!__syncthreads_or(active)) return;
    const float m = v_m(g, v);
    const uint32_t base = g.row[v], rend = g.row[v + 1];
    const uint32_t d = active ? rend - base : 0u;
    const uint32_t cv = g.comm[v];
    const uint32_t szv = g.comm_sz[cv];
    const float kv = g.k[v];
    const float acv = fix2f(g.A[cv], g.S);
    uint32_t eslot[kWaveMaxPer];
    unsigned long long cew = 0;
    // all loads of the lane's entries first, then their community gathers
    uint32_t ei[kWaveMaxPer], ci[kWaveMaxPer];
    float wi[kWaveMaxPer];
    // unconditional loads (past the row end, or of an inactive wave: entry
    // `base`, which exists), so every load and then every gather is in flight
    // at once
#pragma unroll
    for (int i = 0; i < kWaveMaxPer; ++i) {
        const uint32_t j = p + 64u * i;
        const bool has = active && j < d;
        const uint32_t e = base + (has ? j : 0u);
        ei[i] = g.col[e];
        wi[i] = g.w[e];
        ci[i] = has ? 0u : kEmpty;
    }
#pragma unroll
    for (int i = 0; i < kWaveMaxPer; ++i) {
        const uint32_t c = g.comm[ei[i]];
        ei[i] = ci[i] == kEmpty ? kEmpty : ei[i];
        ci[i] = c;
    }
#pragma unroll
    for (int i = 0; i < kWaveMaxPer; ++i) {
        eslot[i] = kEmpty;
        if (ei[i] != kEmpty) {
            const uint32_t e = ei[i];
            const uint32_t ce = ci[i];
            unsigned long long wf = f2fix(wi[i], g.scale);
            uint32_t slot = hslot_n<kWaveCap>(ce);
            for (int probe = 0; probe < kWaveCap; ++probe) {
                uint32_t prev = atomicCAS(&key[slot], kEmpty, ce);
                if (prev == kEmpty) {
                    uint32_t idx = atomicAdd(&sndist[wv], 1u);
                    sdist[wv][idx] = (uint16_t)slot;
                    break;
                }
                if (prev == ce) break;
                slot = slot + 1 == (uint32_t)kWaveCap ? 0u : slot + 1;
            }
            atomicAdd(&val[slot], wf);
            if (ce == cv && e != v) cew += wf;
            eslot[i] = slot;
        }
    }
    cew = wave_sum_u64(cew);
    __syncthreads();
    // one gain per distinct community (complete sums), stored over the sum
    const uint32_t nd = active ? sndist[wv] : 0u;
    distinct_gains(g, sdist[wv], nd, p, 64u, key, val, m, kv, acv, szv, cv);
    __syncthreads();
    float g0 = -INFINITY, g1 = -INFINITY;
    uint32_t c0 = kEmpty, c1 = kEmpty;
#pragma unroll
    for (int i = 0; i < kWaveMaxPer; ++i) {
        if (eslot[i] == kEmpty) continue;
        float gg = __uint_as_float((uint32_t)val[eslot[i]]);
        uint32_t c = key[eslot[i]];
        if (i & 1) {
            if (gg > g1) {
                g1 = gg;
                c1 = c;
            }
        } else {
            if (gg > g0) {
                g0 = gg;
                c0 = c;
            }
        }
    }
    float bg;
    uint32_t bc, bp;
    tree128(g0, c0, g1, c1, bg, bc, bp, g.lid);
    // the destination is the community of the winning virtual lane's last
    // entry (decide_lastedge); lane last % 64 gathered it already, as its
    // entry last / 64, so a shuffle replaces the two dependent loads
    bg = __shfl(bg, 0, 64);
    bp = __shfl(bp, 0, 64);
    const bool win = bg != -INFINITY && bp < d;
    const uint32_t last = win ? last_entry(bp, d) : 0u;
    uint32_t mine = ci[0];
#pragma unroll
    for (int i = 1; i < kWaveMaxPer; ++i)
        if (last / 64u == (uint32_t)i) mine = ci[i];
    const uint32_t dest = __shfl(mine, (int)(last & 63u), 64);
    if (active && p == 0) {
        uint32_t nc = cv;
        if (win && bg - (fix2f(cew, g.S) / m) > 0.0f) nc = dest;
        commit_move(g, v, kv, cv, nc);
    }
}

// Combine per-thread lane states into the 128 virtual lanes and run the
// tree.  Thread t of P serves virtual lane t % 128 on entries t, t+P, ...;
// a virtual lane's first maximum is (max gain, then smallest entry index).
template <int P>
__device__ __forceinline__ void block_argmax(float tg, uint32_t tc, uint32_t tj, float *lg,
                                             uint32_t *lc, uint32_t *lj, float &bg, uint32_t &bc,
                                             uint32_t &bp, const uint32_t *lid) {
    lg[threadIdx.x] = tg;
    lc[threadIdx.x] = tc;
    lj[threadIdx.x] = tj;
    __syncthreads();
    if (threadIdx.x < 64) {
        float gv[2];
        uint32_t cvv[2];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const uint32_t vl = threadIdx.x + 64u * h;
            float g = lg[vl];
            uint32_t c = lc[vl], j = lj[vl];
#pragma unroll
            for (int q = 1; q < P / 128; ++q) {
                const uint32_t t2 = vl + 128u * q;
                float g2 = lg[t2];
                uint32_t j2 = lj[t2];
                if (g2 > g || (g2 == g && j2 < j)) {
                    g = g2;
                    c = lc[t2];
                    j = j2;
                }
            }
            gv[h] = g;
            cvv[h] = c;
        }
        tree128(gv[0], cvv[0], gv[1], cvv[1], bg, bc, bp, lid);
    }
}

// bucket 7 with deg <= kBlockMax: one workgroup of P threads per vertex,
// LDS table of CAP slots (any size above MAXDEG: at most MAXDEG distinct
// keys) with linear probing.  The distinct-slot list is 16-bit and the
// per-thread lane states of the final tree reuse the sum array once the
// lane phase is done, so the LDS footprint is about 12 B per slot.
template <int P, int MAXDEG, int CAP, int MINB = 1>
__global__ __launch_bounds__(P, MINB) void k_move_block(DevView g, const uint32_t *verts,
                                                        uint32_t count) {
    if (halted(g)) return;  // the pass loop stopped earlier in this launch
    if (blockIdx.x < g.cu_nblk) return catch_up(g, blockIdx.x);
    const uint32_t bid = blockIdx.x - g.cu_nblk;
    static_assert(CAP > MAXDEG && CAP <= 65536 && 3 * P * 4 <= CAP * 8, "table shape");
    constexpr int MAXPER = (MAXDEG + P - 1) / P;
    __shared__ unsigned long long val[CAP];
    __shared__ uint32_t key[CAP];
    __shared__ uint16_t dist[MAXDEG];
    __shared__ uint32_t ndist;
    __shared__ unsigned long long scew;
    float *lg = reinterpret_cast<float *>(val);  // lane states, after the lane phase
    uint32_t *lc = reinterpret_cast<uint32_t *>(val) + P;
    uint32_t *lj = reinterpret_cast<uint32_t *>(val) + 2 * P;
    const uint32_t t = threadIdx.x;
    for (uint32_t i = t; i < CAP; i += P) {
        key[i] = kEmpty;
        val[i] = 0ull;
    }
    if (t == 0) {
        ndist = 0;
        scew = 0;
    }
    const uint32_t v = vert_at(g, verts, bid);
    if (!v_live(g, v)) return;  // uniform over the workgroup
    const float m = v_m(g, v);
    const uint32_t base = g.row[v];
    const uint32_t d = g.row[v + 1] - base;
    const uint32_t cv = g.comm[v];
    const uint32_t szv = g.comm_sz[cv];
    const float kv = g.k[v];
    const float acv = fix2f(g.A[cv], g.S);
    uint32_t eslot[MAXPER];
    unsigned long long cew = 0;
    // all loads of the thread's entries first, then their community gathers
    uint32_t ei[MAXPER], ci[MAXPER];
    float wi[MAXPER];
    // unconditional (past the row end: entry `base`), so every load and then
    // every gather is in flight at once
#pragma unroll
    for (int i = 0; i < MAXPER; ++i) {
        const uint32_t j = t + (uint32_t)P * i;
        const uint32_t e = base + (j < d ? j : 0u);
        ei[i] = g.col[e];
        wi[i] = g.w[e];
    }
#pragma unroll
    for (int i = 0; i < MAXPER; ++i) {
        ci[i] = g.comm[ei[i]];
        if (t + (uint32_t)P * i >= d) ei[i] = kEmpty;
    }
    __syncthreads();  // table cleared (the loads above do not touch LDS)
#pragma unroll
    for (int i = 0; i < MAXPER; ++i) {
        eslot[i] = kEmpty;
        if (ei[i] != kEmpty) {
            const uint32_t e = ei[i];
            const uint32_t ce = ci[i];
            unsigned long long wf = f2fix(wi[i], g.scale);
            uint32_t slot = hslot_n<CAP>(ce);
            for (int probe = 0; probe < CAP; ++probe) {
                uint32_t prev = atomicCAS(&key[slot], kEmpty, ce);
                if (prev == kEmpty) {
                    dist[atomicAdd(&ndist, 1u)] = (uint16_t)slot;
                    break;
                }
                if (prev == ce) break;
                slot = slot + 1 == (uint32_t)CAP ? 0u : slot + 1;
            }
            atomicAdd(&val[slot], wf);
            if (ce == cv && e != v) cew += wf;
            eslot[i] = slot;
        }
    }
    cew = wave_sum_u64(cew);
    if ((t & 63) == 0 && cew) atomicAdd(&scew, cew);
    __syncthreads();
    const uint32_t nd = ndist;
    distinct_gains(g, dist, nd, t, (uint32_t)P, key, val, m, kv, acv, szv, cv);
    __syncthreads();
    float tg = -INFINITY;
    uint32_t tc = kEmpty, tj = kEmpty;
#pragma unroll
    for (int i = 0; i < MAXPER; ++i) {
        if (eslot[i] == kEmpty) continue;
        float gg = __uint_as_float((uint32_t)val[eslot[i]]);
        if (gg > tg) {
            tg = gg;
            tc = key[eslot[i]];
            tj = t + (uint32_t)P * i;
        }
    }
    float bg = -INFINITY;
    uint32_t bc = kEmpty, bp = kEmpty;
    __syncthreads();  // the lane states overwrite the sums read above
    block_argmax<P>(tg, tc, tj, lg, lc, lj, bg, bc, bp, g.lid);
    if (t == 0) commit_move(g, v, kv, cv, decide_lastedge(g, m, base, d, cv, bg, bp, scew));
}

#ifndef LV_CAP7
#define LV_CAP7 1024
#endif
#ifndef LV_CAP8
#define LV_CAP8 1280
#endif
#ifndef LV_MINB8
#define LV_MINB8 8
#endif
#ifndef LV_CAP9
#define LV_CAP9 3072
#endif
#ifndef LV_CAP10
#define LV_CAP10 5800
#endif
#ifndef LV_P10
#define LV_P10 1024
#endif
#ifndef LV_MINB10
#define LV_MINB10 2
#endif
#ifndef LV_CAP11
#define LV_CAP11 10240
#endif
#ifndef LV_SMALLCAP5
#define LV_SMALLCAP5 256
#endif
uint32_t catchup_blocks(uint32_t count, uint32_t threads);
__global__ __launch_bounds__(256) void k_decide_from(const SweepTail *t, int S);
void launch_move_class(hipStream_t s, const DevView &g, int cls, const uint32_t *verts,
                       uint32_t count, const MoveShape &shape) {
    if (!count) return;
    // catch-up workgroups in front of the class's own, and the previous
    // sweep's stop rule in workgroup 0 of the kernels of buckets 1-5 (a
    // launch of its own before the others: their entry stays as it was)
    auto go = [&](auto kern, uint32_t grid, uint32_t threads, bool front = false) {
        DevView gv = g;
        if (g.tail && !front) {
            hipLaunchKernelGGL(k_decide_from, dim3(1), dim3(256), 0, s, g.tail, g.S);
            gv.tail = nullptr;
        }
        gv.cu_nblk = g.cu_count ? catchup_blocks(g.cu_count, threads) : 0u;
        hipLaunchKernelGGL(kern, dim3(grid + gv.cu_nblk + (gv.tail ? 1u : 0u)), dim3(threads), 0,
                           s, gv, verts, count);
    };
    switch (cls) {
        // large classes: T virtual lanes (the reference's) on Q lanes per
        // vertex, 256/Q vertices per workgroup, 24.6 KB of LDS tables per
        // workgroup; small classes: T lanes per vertex
        case 1:
            if (count >= shape.narrow_min)
                go(k_move_narrow<4, 1, 1, 8>, (count + 255) / 256, 256, true);
            else
                go(k_move_small<4, 1, 8>, (count + 63) / 64, 256, true);
            break;
        case 2:
            if (count >= shape.narrow_min)
                go(k_move_narrow<8, 2, 1, 16>, (count + 127) / 128, 256, true);
            else
                go(k_move_small<8, 1, 16>, (count + 31) / 32, 256, true);
            break;
        case 3:
            if (count >= shape.narrow_min)
                go(k_move_narrow<16, 4, 1, 32>, (count + 63) / 64, 256, true);
            else
                go(k_move_small<16, 1, 32>, (count + 15) / 16, 256, true);
            break;
        case 4:
            if (count >= shape.narrow_min)
                go(k_move_narrow<32, 8, 1, 64>, (count + 31) / 32, 256, true);
            else
                go(k_move_small<32, 1, 64>, (count + 7) / 8, 256, true);
            break;
        case 5:
            if (count >= shape.narrow_min)
                go(k_move_narrow<32, 16, 3, 128>, (count + 15) / 16, 256, true);
            else
                go(k_move_small<32, 3, LV_SMALLCAP5>, (count + 7) / 8, 256, true);
            break;
        case 6:
            if (shape.wave_vertices == 1)
                go(k_move_wave<1>, count, 64);
            else if (shape.wave_vertices == 4)
                go(k_move_wave<4>, (count + 3) / 4, 256);
            else
                go(k_move_wave<2>, (count + 1) / 2, 128);
            break;
        case 7:
            go(k_move_block<256, 512, LV_CAP7>, count, 256);
            break;
        case 8:
            if (shape.class8_slots == 2048)
                go(k_move_block<256, 1024, 2048, 4>, count, 256);
            else
                go(k_move_block<256, 1024, LV_CAP8, LV_MINB8>, count, 256);
            break;
        case 9:
            go(k_move_block<256, 2048, LV_CAP9>, count, 256);
            break;
        case 10:
            if (shape.block_threads == 512)
                go(k_move_block<512, 4096, LV_CAP10, LV_MINB10>, count, 512);
            else
                go(k_move_block<LV_P10, 4096, LV_CAP10, LV_MINB10>, count, LV_P10);
            break;
        case 11:
            if (shape.block_threads == 512)
                go(k_move_block<512, 8192, LV_CAP11, 1>, count, 512);
            else
                go(k_move_block<1024, 8192, LV_CAP11, 1>, count, 1024);
            break;
        default:
            break;
    }
}

// heavy-class table capacities: 2*deg slots
__global__ void k_heavy_caps(const uint32_t *row, const uint32_t *verts, uint32_t count,
                             unsigned long long *caps) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) {
        uint32_t v = verts[i];
        caps[i] = 2ull * (row[v + 1] - row[v]);
    }
}
void launch_heavy_caps(hipStream_t s, const uint32_t *row, const uint32_t *verts, uint32_t count,
                       unsigned long long *caps) {
    if (!count) return;
    hipLaunchKernelGGL(k_heavy_caps, dim3((count + 255) / 256), dim3(256), 0, s, row, verts,
                       count, caps);
}

// ---------------------------------------------------------------------------
// apply one step's moves together in place (assignCommunity,
// modularity.cu:382-391) and update the community weights incrementally and
// exactly (equal to the from-scratch recompute of computeNewCommWeights,
// modularity.cu:393-398): the step's kernels wrote their decisions to new_comm
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_apply(DevView g, const uint32_t *verts, uint32_t count,
                                               unsigned long long *moved) {
    if (halted(g)) return;  // the pass loop stopped earlier in this launch
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    unsigned long long mv = 0;
    const uint32_t v = i < count ? vert_at(g, verts, i) : 0u;
    const bool valid = i < count && v_live(g, v);
    if (valid) {
        const uint32_t o = g.comm[v], nc = g.new_comm[v];
        if (o != nc) {
            const unsigned long long K = f2fix(g.k[v], g.scale);
            atomicSub(&g.comm_sz[o], 1u);
            atomicAdd(&g.comm_sz[nc], 1u);
            atomicAdd(&g.A[o], (unsigned long long)(-(long long)K));
            atomicAdd(&g.A[nc], K);
            g.comm[v] = nc;
            mv = 1;
        }
    }
    block_add(g, moved, valid ? acc_slot(g, v) : 0u, mv, valid);
}
void launch_apply(hipStream_t s, const DevView &g, const uint32_t *verts, uint32_t count,
                  unsigned long long *moved) {
    if (!count) return;
    hipLaunchKernelGGL(k_apply, dim3((count + 255) / 256), dim3(256), 0, s, g, verts, count,
                       moved);
}

// ---------------------------------------------------------------------------
// a step's moves applied to the other copy of the community state on their
// own (after a sweep's last step, or a single step): every workgroup catches
// up (assignCommunity, modularity.cu:382-391, with the community weights
// updated incrementally and exactly, equal to computeNewCommWeights'
// recompute, :393-398)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_catchup(DevView g) {
    if (halted(g)) return;  // the pass loop stopped earlier in this launch
    catch_up(g, blockIdx.x);
}
uint32_t catchup_blocks(uint32_t count, uint32_t threads) {
    const uint64_t per = (uint64_t)threads * kCatchUpPer;
    return (uint32_t)std::min<uint64_t>((count + per - 1) / per, 2048);
}
void launch_catchup(hipStream_t s, const DevView &g) {
    if (!g.cu_count) return;
    DevView gv = g;
    // a kernel of its own: one round per thread, as many workgroups as that takes
    gv.cu_nblk = (g.cu_count + 256 * kCatchUpPer - 1) / (256 * kCatchUpPer);
    hipLaunchKernelGGL(k_catchup, dim3(gv.cu_nblk), dim3(256), 0, s, gv);
}

// ---------------------------------------------------------------------------
// exclusive scan (3 phases: tile sums, scan of tile sums, tile scan)
// ---------------------------------------------------------------------------
constexpr int kScanThreads = 256;
constexpr int kScanItems = 16;
constexpr int kScanTile = kScanThreads * kScanItems;

size_t scan_tmp_elems(uint32_t n) { return (size_t)n / kScanTile + 2; }

template <typename T>
__device__ __forceinline__ T block_excl_scan(T x, T &total) {
    __shared__ T wsum[kScanThreads / 64];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    T v = x;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        T y = __shfl_up(v, o, 64);
        if (lane >= o) v += y;
    }
    if (lane == 63) wsum[wv] = v;
    __syncthreads();
    T off = 0, tot = 0;
#pragma unroll
    for (int i = 0; i < kScanThreads / 64; ++i) {
        T s = wsum[i];
        if (i < wv) off += s;
        tot += s;
    }
    __syncthreads();
    total = tot;
    return off + v - x;
}

template <typename T>
__device__ __forceinline__ T scan_in(const T *in, int kind, size_t i) {
    T x = in[i];
    return kind == 1 ? (T)(x != 0) : x;
}

template <typename T>
__global__ __launch_bounds__(kScanThreads) void k_scan_tiles(const T *in, int kind, uint32_t n,
                                                             T *tsum) {
    const size_t b = (size_t)blockIdx.x * kScanTile + (size_t)threadIdx.x * kScanItems;
    T s = 0;
#pragma unroll
    for (int i = 0; i < kScanItems; ++i)
        if (b + i < n) s += scan_in(in, kind, b + i);
    T tot;
    block_excl_scan<T>(s, tot);
    if (threadIdx.x == 0) tsum[blockIdx.x] = tot;
}

template <typename T>
__global__ __launch_bounds__(kScanThreads) void k_scan_sums(T *tsum, uint32_t nt, T *out,
                                                            uint32_t n) {
    T carry = 0;
    for (uint32_t b = 0; b < nt; b += kScanThreads) {
        uint32_t i = b + threadIdx.x;
        T x = i < nt ? tsum[i] : (T)0;
        T tot;
        T e = block_excl_scan<T>(x, tot);
        if (i < nt) tsum[i] = carry + e;
        carry += tot;
    }
    if (threadIdx.x == 0) out[n] = carry;
}

template <typename T>
__global__ __launch_bounds__(kScanThreads) void k_scan_final(const T *in, int kind, uint32_t n,
                                                             const T *tsum, T *out) {
    const size_t b = (size_t)blockIdx.x * kScanTile + (size_t)threadIdx.x * kScanItems;
    T x[kScanItems];
    T s = 0;
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) {
        x[i] = (b + i < n) ? scan_in(in, kind, b + i) : (T)0;
        s += x[i];
    }
    T tot;
    T e = block_excl_scan<T>(s, tot) + tsum[blockIdx.x];
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) {
        if (b + i < n) out[b + i] = e;
        e += x[i];
    }
}

// Short arrays (at most kScanOneTiles tiles): one workgroup scans them tile
// by tile with a running carry, one launch instead of three -- the passes of
// small graphs are bound by the host's launches, not by the scans' work.
constexpr uint32_t kScanOneTiles = 4;
template <typename T>
__global__ __launch_bounds__(kScanThreads) void k_scan_one(const T *in, int kind, uint32_t n,
                                                           T *out) {
    T carry = 0;
    for (size_t t0 = 0; t0 < n; t0 += kScanTile) {
        const size_t b = t0 + (size_t)threadIdx.x * kScanItems;
        T x[kScanItems];
        T sum = 0;
#pragma unroll
        for (int i = 0; i < kScanItems; ++i) {
            x[i] = (b + i < n) ? scan_in(in, kind, b + i) : (T)0;
            sum += x[i];
        }
        T tot;
        T e = block_excl_scan<T>(sum, tot) + carry;
#pragma unroll
        for (int i = 0; i < kScanItems; ++i) {
            if (b + i < n) out[b + i] = e;
            e += x[i];
        }
        carry += tot;
    }
    if (threadIdx.x == 0) out[n] = carry;
}

template <typename T>
static void scan_impl(hipStream_t s, const T *in, int kind, uint32_t n, T *out, T *tmp) {
    const uint32_t nt = (uint32_t)(((size_t)n + kScanTile - 1) / kScanTile);
    if (nt <= kScanOneTiles) {  // n == 0 included: out[0] = 0
        hipLaunchKernelGGL((k_scan_one<T>), dim3(1), dim3(kScanThreads), 0, s, in, kind, n, out);
        return;
    }
    hipLaunchKernelGGL((k_scan_tiles<T>), dim3(nt), dim3(kScanThreads), 0, s, in, kind, n, tmp);
    hipLaunchKernelGGL((k_scan_sums<T>), dim3(1), dim3(kScanThreads), 0, s, tmp, nt, out, n);
    hipLaunchKernelGGL((k_scan_final<T>), dim3(nt), dim3(kScanThreads), 0, s, in, kind, n, tmp,
                       out);
}
void launch_scan_u32(hipStream_t s, const uint32_t *in, int kind, uint32_t n, uint32_t *out,
                     uint32_t *tmp) {
    scan_impl<uint32_t>(s, in, kind, n, out, tmp);
}
void launch_scan_u64(hipStream_t s, const unsigned long long *in, uint32_t n,
                     unsigned long long *out, unsigned long long *tmp) {
    scan_impl<unsigned long long>(s, in, 0, n, out, tmp);
}

// ---------------------------------------------------------------------------
// aggregation (aggregation.cu:387-462), canonical form:
//   new_id(c) = number of non-empty c' < c in the input numbering (scan of
//   comm_sz != 0, through lid under a locality order);
//   row of new_id(c) = new ids of the distinct neighbour communities d of
//   c's members, ascending, W_cd exact then rounded once;
//   communities without entries get an empty row.
// ---------------------------------------------------------------------------

// comm_deg[c] = sum of member degrees (setCommData, aggregation.cu:21-36)
__global__ void k_agg_prep(DevView g, AggView a) {
    uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v < g.n) {
        uint32_t d = g.row[v + 1] - g.row[v];
        if (d) atomicAdd(&a.comm_deg[g.comm[v]], d);
    }
}
void launch_agg_prep(hipStream_t s, const DevView &g, AggView a) {
    if (!g.n) return;
    hipLaunchKernelGGL(k_agg_prep, dim3((g.n + 255) / 256), dim3(256), 0, s, g, a);
}

// members grouped by community (vertByCommOrder, aggregation.cu:38-45); the
// order inside a community is not observable in the canonical output
__global__ void k_agg_members(DevView g, AggView a) {
    uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v < g.n) {
        uint32_t c = g.comm[v];
        uint32_t r = atomicAdd(&a.mcursor[c], 1u);
        a.members[a.mstart[c] + r] = v;
    }
}
void launch_agg_members(hipStream_t s, const DevView &g, AggView a) {
    if (!g.n) return;
    hipLaunchKernelGGL(k_agg_members, dim3((g.n + 255) / 256), dim3(256), 0, s, g, a);
}

__device__ __forceinline__ int agg_class(uint32_t cdeg) {
    if (cdeg == 0) return 0;
    if (cdeg <= kAggSmallMax) return 1;
    if (cdeg <= 512) return 2;
    if (cdeg <= 1024) return 3;
    if (cdeg <= kAggMidMax) return 4;
    return kAggGlobalClass;
}

__global__ __launch_bounds__(256) void k_agg_class_count(DevView g, AggView a, uint32_t *counts) {
    __shared__ uint32_t hc[kNumAggClasses];
    if (threadIdx.x < kNumAggClasses) hc[threadIdx.x] = 0;
    __syncthreads();
    for (uint32_t c = blockIdx.x * 256 + threadIdx.x; c < g.n; c += gridDim.x * 256)
        if (g.comm_sz[c]) atomicAdd(&hc[agg_class(a.comm_deg[c])], 1u);
    __syncthreads();
    if (threadIdx.x < kNumAggClasses && hc[threadIdx.x])
        atomicAdd(&counts[threadIdx.x], hc[threadIdx.x]);
}
__global__ __launch_bounds__(256) void k_agg_class_scatter(DevView g, AggView a, uint32_t *cursor,
                                                           uint32_t *order) {
    __shared__ uint32_t hc[kNumAggClasses], base[kNumAggClasses];
    if (threadIdx.x < kNumAggClasses) hc[threadIdx.x] = 0;
    __syncthreads();
    int k[kScatterPer];
    uint32_t r[kScatterPer];
#pragma unroll
    for (int q = 0; q < kScatterPer; ++q) {
        const uint32_t c = (blockIdx.x * kScatterPer + q) * 256 + threadIdx.x;
        k[q] = -1;
        r[q] = 0;
        if (c < g.n && g.comm_sz[c]) {
            k[q] = agg_class(a.comm_deg[c]);
            r[q] = atomicAdd(&hc[k[q]], 1u);
        }
    }
    __syncthreads();
    if (threadIdx.x < kNumAggClasses && hc[threadIdx.x])
        base[threadIdx.x] = atomicAdd(&cursor[threadIdx.x], hc[threadIdx.x]);
    __syncthreads();
#pragma unroll
    for (int q = 0; q < kScatterPer; ++q)
        if (k[q] >= 0) order[base[k[q]] + r[q]] = (blockIdx.x * kScatterPer + q) * 256 + threadIdx.x;
}
void launch_agg_classify_count(hipStream_t s, const DevView &g, AggView a, uint32_t *counts) {
    if (!g.n) return;
    hipLaunchKernelGGL(k_agg_class_count, dim3(grid_for(g.n, 256, 4096)), dim3(256), 0, s, g, a,
                       counts);
}
void launch_agg_classify_scatter(hipStream_t s, const DevView &g, AggView a, uint32_t *cursor,
                                 uint32_t *order) {
    if (!g.n) return;
    hipLaunchKernelGGL(k_agg_class_scatter,
                       dim3((g.n + 256 * kScatterPer - 1) / (256 * kScatterPer)), dim3(256), 0, s,
                       g, a, cursor, order);
}

// Bitonic sort of n2 (pow2) (key, slot) pairs in LDS, ascending by key.
template <int P>
__device__ void lds_bitonic(uint32_t *k, uint32_t *sl, uint32_t n2) {
    for (uint32_t kk = 2; kk <= n2; kk <<= 1) {
        for (uint32_t j = kk >> 1; j > 0; j >>= 1) {
            for (uint32_t i = threadIdx.x; i < n2; i += P) {
                uint32_t ixj = i ^ j;
                if (ixj > i) {
                    bool up = (i & kk) == 0;
                    uint32_t a = k[i], b = k[ixj];
                    if ((a > b) == up) {
                        k[i] = b;
                        k[ixj] = a;
                        uint32_t t = sl[i];
                        sl[i] = sl[ixj];
                        sl[ixj] = t;
                    }
                }
            }
            __syncthreads();
        }
    }
}

// Ascending order of nd distinct keys (keyof(i), i < nd) lying in
// [lo, lo + W), without a comparison sort: the keys are binned by
// (k - lo) >> s into at most NB bins, a prefix sum gives every bin its start,
// and a key's rank is its bin's start plus the keys of its bin below it;
// emit(rank, i) is called once per key.  Returns false, with nothing emitted,
// when a bin holds more than kBinMax keys (clustered keys): the caller then
// sorts.  cur/boffs/oidx: LDS u32[NB]; P threads, block-uniform.
constexpr uint32_t kBinMax = 32;
template <uint32_t NB, int P, typename KeyOf, typename Emit>
__device__ bool lds_bin_rank(uint32_t nd, uint32_t lo, uint32_t W, KeyOf keyof, uint32_t *cur,
                             uint32_t *boffs, uint32_t *oidx, uint32_t *wsum, uint32_t *maxbin,
                             Emit emit) {
    static_assert(NB % P == 0 && P % 64 == 0, "bins per thread, whole waves");
    constexpr uint32_t PER = NB / P;
    const uint32_t t = threadIdx.x, lane = t & 63, wv = t >> 6;
    uint32_t sh = 0;
    while (((W - 1u) >> sh) >= NB) ++sh;
    for (uint32_t i = t; i < NB; i += P) cur[i] = 0u;
    if (t == 0) *maxbin = 0u;
    __syncthreads();
    uint32_t mx = 0;  // largest bin this thread saw; one LDS max per wave
    for (uint32_t i = t; i < nd; i += P) {
        const uint32_t c = atomicAdd(&cur[(keyof(i) - lo) >> sh], 1u);
        mx = max(mx, c + 1u);
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) mx = max(mx, (uint32_t)__shfl_xor(mx, o, 64));
    if (lane == 0 && mx) atomicMax(maxbin, mx);
    __syncthreads();
    if (*maxbin > kBinMax) return false;
    uint32_t cnt[PER], own = 0;
#pragma unroll
    for (uint32_t k = 0; k < PER; ++k) {
        cnt[k] = cur[PER * t + k];
        own += cnt[k];
    }
    uint32_t inc = own;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(inc, o, 64);
        if (lane >= (uint32_t)o) inc += y;
    }
    if (lane == 63) wsum[wv] = inc;
    __syncthreads();
    uint32_t run = inc - own;
    for (uint32_t k = 0; k < wv; ++k) run += wsum[k];
#pragma unroll
    for (uint32_t k = 0; k < PER; ++k) {
        boffs[PER * t + k] = run;
        cur[PER * t + k] = run;
        run += cnt[k];
    }
    __syncthreads();
    for (uint32_t i = t; i < nd; i += P) oidx[atomicAdd(&cur[(keyof(i) - lo) >> sh], 1u)] = i;
    __syncthreads();
    for (uint32_t i = t; i < nd; i += P) {
        const uint32_t k = keyof(i), b = (k - lo) >> sh;
        const uint32_t st = boffs[b], en = cur[b];
        uint32_t r = 0;
        for (uint32_t j = st; j < en; ++j) r += keyof(oidx[j]) < k ? 1u : 0u;
        emit(st + r, i);
    }
    return true;
}

// LDS merge: one workgroup of P threads per community; waves walk members,
// lanes walk a member's row.  Output: (new_id(d), W) sorted at
// tkey/tval[epos[c]..],
// distinct count at cnt[new_id[c]].
template <int P, int MAXE, int CAP>
__global__ __launch_bounds__(P) void k_agg_merge_lds(DevView g, AggView a, const uint32_t *comms,
                                                     uint32_t count) {
    // MAXE: member entries of the class (comm_deg <= MAXE, a power of two);
    // CAP: table slots (any size above MAXE)
    static_assert(CAP > MAXE && (MAXE & (MAXE - 1)) == 0, "table shape");
    constexpr int NW = P / 64;
    constexpr uint32_t PER = (MAXE + P - 1) / P;
    __shared__ uint32_t key[CAP];
    __shared__ unsigned long long val[CAP];
    constexpr uint32_t NBIN = MAXE / 2;  // bins for the ranked order
    __shared__ uint32_t skeys[MAXE];
    __shared__ uint32_t sslot[MAXE];
    __shared__ uint32_t bcur[NBIN], boffs[NBIN], oidx[MAXE + 1];
    __shared__ uint32_t ndist, wsum[NW], maxbin, kmin, kmax;
    uint32_t *mb = oidx;  // member entry offsets while merging, ranks afterwards
    const uint32_t c = comms[blockIdx.x];
    for (uint32_t i = threadIdx.x; i < CAP; i += P) {
        key[i] = kEmpty;
        val[i] = 0ull;
    }
    if (threadIdx.x == 0) {
        ndist = 0;
        kmin = 0xffffffffu;
        kmax = 0u;
    }
    // the community's member entries, flat: members have degree >= 1, so at
    // most MAXE of them; offsets relative to the first
    const uint32_t m0 = a.mstart[c], msz = g.comm_sz[c], ne = a.comm_deg[c];
    const uint32_t f0 = a.moff[m0];
    for (uint32_t q = threadIdx.x; q <= msz; q += P) mb[q] = a.moff[m0 + q] - f0;
    __syncthreads();
    // every entry's loads before any insert: member, row start, column and
    // weight, new id of the neighbour's community.  Positions past the end
    // repeat entry 0 (ne >= 1), so no load waits on a branch.
    uint32_t vv[PER], jj[PER], dd[PER], nid[PER];
    float ww[PER];
#pragma unroll
    for (uint32_t u = 0; u < PER; ++u) {
        const uint32_t ru = threadIdx.x + (uint32_t)P * u;
        const uint32_t r = ru < ne ? ru : 0u;
        uint32_t l = 0, h = msz;  // member of flat entry r
        while (h - l > 1) {
            const uint32_t mid = (l + h) >> 1;
            if (mb[mid] <= r) l = mid;
            else h = mid;
        }
        vv[u] = a.members[m0 + l];
        jj[u] = r - mb[l];
    }
#pragma unroll
    for (uint32_t u = 0; u < PER; ++u) jj[u] += g.row[vv[u]];
#pragma unroll
    for (uint32_t u = 0; u < PER; ++u) {
        dd[u] = g.col[jj[u]];
        ww[u] = g.w[jj[u]];
    }
#pragma unroll
    for (uint32_t u = 0; u < PER; ++u) dd[u] = a.vnid[dd[u]];  // keyed by new id
#pragma unroll
    for (uint32_t u = 0; u < PER; ++u) nid[u] = dd[u];
#pragma unroll
    for (uint32_t u = 0; u < PER; ++u) {
        if (threadIdx.x + (uint32_t)P * u >= ne) continue;
        const uint32_t d2 = dd[u];
        uint32_t slot = hslot_n<CAP>(d2);
        for (int probe = 0; probe < CAP; ++probe) {
            const uint32_t prev = atomicCAS(&key[slot], kEmpty, d2);
            if (prev == kEmpty) {
                const uint32_t idx = atomicAdd(&ndist, 1u);
                skeys[idx] = nid[u];  // sorted by new id
                sslot[idx] = slot;
                break;
            }
            if (prev == d2) break;
            slot = slot + 1 == (uint32_t)CAP ? 0u : slot + 1;
        }
        atomicAdd(&val[slot], f2fix(ww[u], g.scale));
    }
    __syncthreads();
    const uint32_t nd = ndist;
    const uint32_t out = a.epos[c];
    auto keyof = [&](uint32_t i) { return skeys[i]; };
    auto emit = [&](uint32_t pos, uint32_t i) {
        a.tkey[out + pos] = skeys[i];
        a.tval[out + pos] = val[sslot[i]];
    };
    // the bins span the community's own key range (in a batch the keys of
    // one graph's communities are a narrow slice of [0, nn))
    {
        uint32_t lo = 0xffffffffu, hi = 0u;
        for (uint32_t i = threadIdx.x; i < nd; i += P) {
            lo = min(lo, skeys[i]);
            hi = max(hi, skeys[i]);
        }
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            lo = min(lo, (uint32_t)__shfl_xor(lo, o, 64));
            hi = max(hi, (uint32_t)__shfl_xor(hi, o, 64));
        }
        if ((threadIdx.x & 63) == 0 && lo <= hi) {
            atomicMin(&kmin, lo);
            atomicMax(&kmax, hi);
        }
        __syncthreads();
    }
    // oidx holds up to MAXE distinct keys (nd >= 1: the community has entries)
    if (!lds_bin_rank<NBIN, P>(nd, kmin, kmax - kmin + 1u, keyof, bcur, boffs, oidx, wsum,
                               &maxbin, emit)) {
        uint32_t n2 = 1;
        while (n2 < nd) n2 <<= 1;
        for (uint32_t i = nd + threadIdx.x; i < n2; i += P) {
            skeys[i] = kEmpty;
            sslot[i] = 0;
        }
        __syncthreads();
        lds_bitonic<P>(skeys, sslot, n2);
        for (uint32_t i = threadIdx.x; i < nd; i += P) {
            a.tkey[out + i] = skeys[i];
            a.tval[out + i] = val[sslot[i]];
        }
    }
    if (threadIdx.x == 0) a.cnt[a.new_id[c]] = nd;
}

// Small communities (at most kAggSmallMax member entries): one wave each,
// four per workgroup, with wave-private LDS tables, so the waves never wait
// for each other.  The member entries are walked flat (two per lane, every
// load before any insert), distinct neighbour communities (by new id) are
// merged in an open-addressing table with exact sums, and each distinct key
// lands at its rank among the community's keys (the count of smaller ones),
// which is the ascending order without a sort.
__global__ __launch_bounds__(256) void k_agg_merge_small(DevView g, AggView a,
                                                         const uint32_t *comms, uint32_t count) {
    constexpr int CAP = 256, LC = 8, N = (int)kAggSmallMax;
    static_assert(N == 128, "two flat entries per lane of a wave");
    __shared__ uint32_t skey_all[4][CAP];
    __shared__ unsigned long long sval_all[4][CAP];
    __shared__ uint32_t sk_all[4][N];
    __shared__ uint32_t ss_all[4][N];
    __shared__ uint32_t mb_all[4][N + 1];
    __shared__ uint32_t nd_all[4];
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const uint32_t ci = blockIdx.x * 4 + wv;
    if (ci >= count) return;  // wave-uniform; no workgroup barrier below
    uint32_t *key = skey_all[wv];
    unsigned long long *val = sval_all[wv];
    uint32_t *sk = sk_all[wv], *ss = ss_all[wv], *mb = mb_all[wv];
    for (int i = lane; i < CAP; i += 64) {
        key[i] = kEmpty;
        val[i] = 0ull;
    }
    if (lane == 0) nd_all[wv] = 0;
    const uint32_t c = comms[ci];
    // the community's member entries, flat (members have degree >= 1, so at
    // most N of them): offsets of the members' rows relative to the first
    const uint32_t m0 = a.mstart[c], msz = g.comm_sz[c], ne = a.comm_deg[c];
    const uint32_t f0 = a.moff[m0];
    for (uint32_t q = lane; q <= msz; q += 64) mb[q] = a.moff[m0 + q] - f0;
    wave_sync_lds();
    // flat entries lane and lane + 64: every load of both before any insert
    // (member, row start, column and weight, new id of the neighbour's
    // community); positions past the end repeat entry 0 (ne >= 1)
    uint32_t vv[2], jj[2], dd[2];
    float ww[2];
#pragma unroll
    for (uint32_t u = 0; u < 2; ++u) {
        const uint32_t ru = lane + 64u * u;
        const uint32_t r = ru < ne ? ru : 0u;
        uint32_t l = 0, h = msz;  // member of flat entry r
        while (h - l > 1) {
            const uint32_t mid = (l + h) >> 1;
            if (mb[mid] <= r) l = mid;
            else h = mid;
        }
        vv[u] = a.members[m0 + l];
        jj[u] = r - mb[l];
    }
#pragma unroll
    for (uint32_t u = 0; u < 2; ++u) jj[u] += g.row[vv[u]];
#pragma unroll
    for (uint32_t u = 0; u < 2; ++u) {
        dd[u] = g.col[jj[u]];
        ww[u] = g.w[jj[u]];
    }
#pragma unroll
    for (uint32_t u = 0; u < 2; ++u) dd[u] = a.vnid[dd[u]];  // keyed by new id
#pragma unroll
    for (uint32_t u = 0; u < 2; ++u) {
        if (lane + 64u * u >= ne) continue;
        const uint32_t d2 = dd[u];
        uint32_t slot = hslot<LC>(d2);
        for (int probe = 0; probe < CAP; ++probe) {
            const uint32_t prev = atomicCAS(&key[slot], kEmpty, d2);
            if (prev == kEmpty) {
                const uint32_t idx = atomicAdd(&nd_all[wv], 1u);
                sk[idx] = d2;
                ss[idx] = slot;
                break;
            }
            if (prev == d2) break;
            slot = (slot + 1) & (CAP - 1);
        }
        atomicAdd(&val[slot], f2fix(ww[u], g.scale));
    }
    wave_sync_lds();
    const uint32_t nd = nd_all[wv];
    const uint32_t out = a.epos[c];
    // rank of each distinct key among the nd keys (all distinct): its place
    // in ascending new-id order
    for (uint32_t i = lane; i < nd; i += 64) {
        const uint32_t k = sk[i];
        uint32_t rank = 0;
        for (uint32_t t = 0; t < nd; ++t) rank += sk[t] < k ? 1u : 0u;
        a.tkey[out + rank] = k;
        a.tval[out + rank] = val[ss[i]];
    }
    if (lane == 0) a.cnt[a.new_id[c]] = nd;
}

// Large communities (comm_deg > 2048): their member entries form the flat
// range [epos[c], epos[c] + comm_deg[c]) of the member-ordered entry sequence
// (moff = scan of member degrees).  One workgroup per chunk of kAggChunk flat
// entries merges its chunk in LDS, then adds each distinct neighbour
// community once to the community's global table and lists newly claimed
// slots; k_agg_merge_global then sorts the list.
constexpr uint32_t kAggChunk = 2048;

// Chunk `blk` of the large communities `comms`: the community (ci, c), its
// flat entry range [f0, f1), and in LDS the member-entry offsets mb[0..nm] of
// the members i0.. covering it (members of a large community have degree
// >= 1, so at most kAggChunk of them cover a chunk).  Block-uniform.
struct AggChunk {
    uint32_t ci, c, f0, f1, i0, nm;
};
__device__ AggChunk agg_chunk(const DevView &g, const AggView &a, const uint32_t *comms,
                              uint32_t count, uint32_t blk, uint32_t *mb, uint32_t *sh) {
    AggChunk x;
    uint32_t lo = 0, hi = count;  // largest ci with coff[ci] <= blk
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (a.coff[mid] <= blk) lo = mid;
        else hi = mid;
    }
    x.ci = lo;
    x.c = comms[x.ci];
    const uint32_t k = blk - a.coff[x.ci];
    const uint32_t cb = a.epos[x.c], ce = cb + a.comm_deg[x.c];
    x.f0 = cb + k * kAggChunk;
    x.f1 = min(ce, x.f0 + kAggChunk);
    const uint32_t m0 = a.mstart[x.c], m1 = m0 + g.comm_sz[x.c];
    if (threadIdx.x == 0) {
        uint32_t l = m0, h = m1;  // last member with moff <= f0
        while (h - l > 1) {
            const uint32_t mid = (l + h) >> 1;
            if (a.moff[mid] <= x.f0) l = mid;
            else h = mid;
        }
        sh[0] = l;
        sh[1] = min(m1 - l, (uint32_t)kAggChunk);
    }
    __syncthreads();
    x.i0 = sh[0];
    x.nm = sh[1];
    for (uint32_t q = threadIdx.x; q <= x.nm; q += blockDim.x) mb[q] = a.moff[x.i0 + q];
    __syncthreads();
    return x;
}
// CSR entry of flat entry f of the chunk
__device__ __forceinline__ uint32_t agg_entry(const DevView &g, const AggView &a,
                                              const AggChunk &x, const uint32_t *mb, uint32_t f) {
    uint32_t l = 0, h = x.nm;  // member of flat entry f
    while (h - l > 1) {
        const uint32_t mid = (l + h) >> 1;
        if (mb[mid] <= f) l = mid;
        else h = mid;
    }
    const uint32_t v = a.members[x.i0 + l];
    return g.row[v] + (f - mb[l]);
}

__global__ void k_agg_mdeg(DevView g, AggView a) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < g.n) {
        const uint32_t v = a.members[i];
        a.mdeg[i] = g.row[v + 1] - g.row[v];
    }
}
__global__ void k_agg_chunk_counts(AggView a, const uint32_t *comms, uint32_t count) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) a.cnk[i] = (a.comm_deg[comms[i]] + kAggChunk - 1) / kAggChunk;
}

__global__ __launch_bounds__(256) void k_agg_insert_chunks(DevView g, AggView a,
                                                           const uint32_t *comms, uint32_t count) {
    constexpr int LCAP = 2 * kAggChunk, LC = Log2<LCAP>::value, PER = kAggChunk / 256;
    __shared__ uint32_t lkey[LCAP];
    __shared__ unsigned long long lval[LCAP];
    __shared__ uint32_t ldist[kAggChunk];
    __shared__ uint32_t mb[kAggChunk + 1];
    __shared__ uint32_t nld, nnew, newbase, sh[2];
    for (int i = threadIdx.x; i < LCAP; i += 256) {
        lkey[i] = kEmpty;
        lval[i] = 0ull;
    }
    if (threadIdx.x == 0) {
        nld = 0;
        nnew = 0;
    }
    // one workgroup per listed chunk (chunks of flagged communities only)
    const AggChunk x = agg_chunk(g, a, comms, count, a.flist[1 + blockIdx.x], mb, sh);
    const uint32_t ci = x.ci, f0 = x.f0, f1 = x.f1;
    const uint32_t cb = a.epos[x.c];
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const uint32_t f = f0 + threadIdx.x + 256u * q;
        if (f < f1) {
            const uint32_t j = agg_entry(g, a, x, mb, f);
            const uint32_t dd = a.vnid[g.col[j]];  // keyed by new id
            const unsigned long long wf = f2fix(g.w[j], g.scale);
            uint32_t slot = hslot<LC>(dd);
            for (int probe = 0; probe < LCAP; ++probe) {
                const uint32_t prev = atomicCAS(&lkey[slot], kEmpty, dd);
                if (prev == kEmpty) {
                    ldist[atomicAdd(&nld, 1u)] = slot;
                    break;
                }
                if (prev == dd) break;
                slot = (slot + 1) & (LCAP - 1);
            }
            atomicAdd(&lval[slot], wf);
        }
    }
    __syncthreads();
    const unsigned long long off = a.goff[ci];
    const uint32_t cap = (uint32_t)(a.goff[ci + 1] - off);
    unsigned long long *tab = a.gtab + 2ull * off;
    const uint32_t nd = nld;
    uint32_t my_new[PER], my_rank[PER], my_key[PER];
    int nmine = 0;
    for (uint32_t i = threadIdx.x; i < nd; i += 256) {
        const uint32_t ls = ldist[i];
        const uint32_t dd = lkey[ls];
        uint32_t slot = hslot_cap(dd, cap);
        uint32_t probe = 0;
        bool claimed = false;
        for (; probe < cap; ++probe) {
            const uint32_t prev = atomicCAS(slot_key(tab, slot), kEmpty, dd);
            if (prev == kEmpty) {
                claimed = true;
                break;
            }
            if (prev == dd) break;
            slot = next_slot(slot, cap);
        }
        if (probe == cap) {  // cannot happen: cap = 2*comm_deg
            atomicOr(a.err, 1u);
            continue;
        }
        atomicAdd(slot_sum(tab, slot), lval[ls]);
        if (claimed) {
            my_new[nmine] = slot;
            my_key[nmine] = dd;  // sorted by new id
            my_rank[nmine] = atomicAdd(&nnew, 1u);
            ++nmine;
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) newbase = nnew ? atomicAdd(&a.dcnt[ci], nnew) : 0u;
    __syncthreads();
    uint32_t *ok = a.tkey + cb;
    unsigned long long *ov = a.tval + cb;
    for (int q = 0; q < nmine; ++q) {
        ok[newbase + my_rank[q]] = my_key[q];
        ov[newbase + my_rank[q]] = my_new[q];
    }
}

// global-hash merge for large communities, then an in-workgroup stable LSD
// radix sort of the distinct keys (8-bit digits) with scratch skey/sval.
constexpr int kRadixThreads = 1024;
__global__ __launch_bounds__(kRadixThreads) void k_agg_merge_global(DevView g, AggView a,
                                                                    const uint32_t *comms,
                                                                    uint32_t count, int passes) {
    constexpr int P = kRadixThreads;
    constexpr int NW = P / 64;
    __shared__ uint32_t hist[256];
    __shared__ uint32_t binbase[256];
    __shared__ uint32_t ctot[256];
    __shared__ uint32_t wcnt[NW][256];
    const uint32_t ci = a.clist[1 + blockIdx.x];  // one workgroup per flagged community
    const uint32_t c = comms[ci];
    unsigned long long *tab = a.gtab + 2ull * a.goff[ci];
    const uint32_t out = a.epos[c];
    uint32_t *ok = a.tkey + out;
    unsigned long long *ov = a.tval + out;
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    // entries were inserted by k_agg_insert_chunks; the nd distinct keys are
    // listed at ok[0..nd) with their table slots in ov[0..nd)
    const uint32_t nd = a.dcnt[ci];
    // gather sums and clean the arena
    for (uint32_t i = threadIdx.x; i < nd; i += P) {
        uint32_t slot = (uint32_t)ov[i];
        ov[i] = *slot_sum(tab, slot);
        tab[2ull * slot] = (unsigned long long)kEmpty;
        *slot_sum(tab, slot) = 0ull;
    }
    __syncthreads();
    uint32_t *kin = ok, *kout = a.skey + out;
    unsigned long long *vin = ov, *vout = a.sval + out;
    for (int ps = 0; ps < passes; ++ps) {
        const int shift = 8 * ps;
        for (int i = threadIdx.x; i < 256; i += P) hist[i] = 0;
        for (int i = threadIdx.x; i < NW * 256; i += P) (&wcnt[0][0])[i] = 0;
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < nd; i += P) atomicAdd(&hist[(kin[i] >> shift) & 255u], 1u);
        __syncthreads();
        if (threadIdx.x == 0) {
            uint32_t run = 0;
            for (int b = 0; b < 256; ++b) {
                binbase[b] = run;
                run += hist[b];
            }
        }
        __syncthreads();
        for (uint32_t ch = 0; ch < nd; ch += P) {
            const uint32_t i = ch + threadIdx.x;
            const bool has = i < nd;
            const uint32_t kk = has ? kin[i] : 0u;
            const uint32_t dig = (kk >> shift) & 255u;
            unsigned long long eq = __ballot(has);
#pragma unroll
            for (int bit = 0; bit < 8; ++bit) {
                unsigned long long bb = __ballot(has && ((dig >> bit) & 1u));
                eq &= ((dig >> bit) & 1u) ? bb : ~bb;
            }
            const uint32_t rank = __popcll(eq & ((1ull << lane) - 1ull));
            if (has && rank == 0) wcnt[wv][dig] = __popcll(eq);
            __syncthreads();
            if (threadIdx.x < 256) {
                uint32_t run = 0;
                for (int w2 = 0; w2 < NW; ++w2) {
                    uint32_t x = wcnt[w2][threadIdx.x];
                    wcnt[w2][threadIdx.x] = run;
                    run += x;
                }
                ctot[threadIdx.x] = run;
            }
            __syncthreads();
            if (has) {
                uint32_t pos = binbase[dig] + wcnt[wv][dig] + rank;
                kout[pos] = kk;
                vout[pos] = vin[i];
            }
            __syncthreads();
            if (threadIdx.x < 256) {
                binbase[threadIdx.x] += ctot[threadIdx.x];
                for (int w2 = 0; w2 < NW; ++w2) wcnt[w2][threadIdx.x] = 0;
            }
            __syncthreads();
        }
        uint32_t *tk = kin;
        kin = kout;
        kout = tk;
        unsigned long long *tv = vin;
        vin = vout;
        vout = tv;
    }
    if (kin != ok) {
        for (uint32_t i = threadIdx.x; i < nd; i += P) {
            ok[i] = kin[i];
            ov[i] = vin[i];
        }
    }
    if (threadIdx.x == 0) a.cnt[a.new_id[c]] = nd;
}

#ifndef LV_AGGCAP2
#define LV_AGGCAP2 768
#endif
#ifndef LV_AGGCAP3
#define LV_AGGCAP3 1536
#endif
#ifndef LV_AGGCAP4
#define LV_AGGCAP4 3072
#endif
void launch_agg_merge(hipStream_t s, const DevView &g, AggView a, int cls, const uint32_t *comms,
                      uint32_t count, int radix_passes) {
    if (!count || cls == 0) return;  // class 0: no entries, cnt stays 0
    if (cls == 1)
        hipLaunchKernelGGL(k_agg_merge_small, dim3((count + 3) / 4), dim3(256), 0, s, g, a, comms,
                           count);
    else if (cls == 2)
        hipLaunchKernelGGL((k_agg_merge_lds<256, 512, LV_AGGCAP2>), dim3(count), dim3(256), 0, s, g, a, comms,
                           count);
    else if (cls == 3)
        hipLaunchKernelGGL((k_agg_merge_lds<256, 1024, LV_AGGCAP3>), dim3(count), dim3(256), 0, s, g, a, comms,
                           count);
    else if (cls == 4)
        hipLaunchKernelGGL((k_agg_merge_lds<256, 2048, LV_AGGCAP4>), dim3(count), dim3(256), 0, s, g, a, comms,
                           count);
    else if (a.ncl) {  // the listed (flagged) communities and chunks only
        hipLaunchKernelGGL(k_agg_insert_chunks, dim3(a.nfl), dim3(256), 0, s, g, a, comms,
                           count);
        hipLaunchKernelGGL(k_agg_merge_global, dim3(a.ncl), dim3(kRadixThreads), 0, s, g, a,
                           comms, count, radix_passes);
    }
}

__global__ void k_agg_flag_list(const uint32_t *flag, const uint32_t *coff, uint32_t count,
                                uint32_t *flist, uint32_t *clist) {
    const uint32_t ci = blockIdx.x * blockDim.x + threadIdx.x;
    if (ci >= count || !flag[ci]) return;
    clist[1 + atomicAdd(clist, 1u)] = ci;
    const uint32_t c0 = coff[ci], nc = coff[ci + 1] - c0;
    const uint32_t pos = atomicAdd(flist, nc);
    for (uint32_t t = 0; t < nc; ++t) flist[1 + pos + t] = c0 + t;
}
void launch_agg_flag_list(hipStream_t s, const uint32_t *flag, const uint32_t *coff,
                          uint32_t count, uint32_t *flist, uint32_t *clist) {
    if (!count) return;
    hipLaunchKernelGGL(k_agg_flag_list, dim3((count + 255) / 256), dim3(256), 0, s, flag, coff,
                       count, flist, clist);
}

void launch_agg_mdeg(hipStream_t s, const DevView &g, AggView a) {
    if (g.n)
        hipLaunchKernelGGL(k_agg_mdeg, dim3((g.n + 255) / 256), dim3(256), 0, s, g, a);
}
void launch_agg_large_prep(hipStream_t s, const DevView &g, AggView a, const uint32_t *comms,
                           uint32_t count) {
    if (count)
        hipLaunchKernelGGL(k_agg_chunk_counts, dim3((count + 255) / 256), dim3(256), 0, s, a,
                           comms, count);
}

// ---------------------------------------------------------------------------
// Large communities, partitioned (the default for comm_deg > kAggMidMax).
// Each community's entries are split by the new id of their neighbour
// community into key ranges (buckets of about the tuning's target entries); the
// buckets are merged and sorted in LDS, and since bucket b's keys all precede
// bucket b+1's, the bucket outputs in order are the community's sorted row.
//   P0 count    new id of every entry (staged in entry order), bucket
//               histogram per chunk
//   (scan of the bucket counts -> bucket ranges)
//   P1 scatter  (key, weight) into the bucket ranges, runs written contiguously
//   P2 bucket   LDS table of exact sums, distinct keys sorted, written at the
//               bucket's place in the community's output range
// A bucket with more distinct keys than its table holds flags its community,
// which is then merged by the global-table kernels above instead.
// ---------------------------------------------------------------------------
constexpr uint32_t kAggMaxBuckets = 4096;
#ifndef LV_AGGBCAP
#define LV_AGGBCAP 1024
#endif
constexpr int kAggBucketCap = LV_AGGBCAP;  // LDS table slots of one bucket (> kAggBucketDistinct)

__device__ __forceinline__ uint32_t agg_bucket(uint32_t key, uint32_t nb, uint32_t nn) {
    return (uint32_t)(((unsigned long long)key * nb) / nn);
}
// The key range [kb, kb + span) of community c's neighbours: the graph's
// new ids in a batch (edges never leave a graph), all of [0, nn) otherwise.
// Buckets split this range evenly.
__device__ __forceinline__ void agg_keys(const AggView &a, uint32_t nn, uint32_t c, uint32_t &kb,
                                         uint32_t &span) {
    if (a.kgid) {
        const uint32_t gi = a.kgid[c];
        kb = a.kbase[gi];
        span = a.kbase[gi + 1] - kb;
    } else {
        kb = 0;
        span = nn;
    }
}

// buckets of a large community: pow2 >= its entries / target, but no more
// than its key range needs (span / 64 keys per bucket at least, so a small
// graph's community with many entries and few distinct neighbours does not
// dispatch hundreds of empty buckets)
__global__ void k_aggp_nb(AggView a, const uint32_t *comms, uint32_t count, uint32_t nn,
                          uint32_t target, uint32_t *nbk) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) {
        const uint32_t c = comms[i];
        const uint32_t d = a.comm_deg[c];
        uint32_t kb, span;
        agg_keys(a, nn, c, kb, span);
        uint32_t b = 1;
        while (b < kAggMaxBuckets && (unsigned long long)b * target < d &&
               (unsigned long long)b * 64u < span)
            b <<= 1;
        nbk[i] = b;
    }
}
void launch_aggp_nb(hipStream_t s, AggView a, const uint32_t *comms, uint32_t count, uint32_t nn,
                    uint32_t target, uint32_t *nbk) {
    if (!count) return;
    hipLaunchKernelGGL(k_aggp_nb, dim3((count + 255) / 256), dim3(256), 0, s, a, comms, count,
                       nn, target, nbk);
}

// P0: new id of every entry of the chunk, staged in entry order, counted per
// bucket
__global__ __launch_bounds__(256) void k_aggp_count(DevView g, AggView a, AggPart p,
                                                    const uint32_t *comms, uint32_t count) {
    __shared__ uint32_t mb[kAggChunk + 1];
    __shared__ uint32_t hist[kAggMaxBuckets];
    __shared__ uint32_t sh[2];
    for (uint32_t b = threadIdx.x; b < kAggMaxBuckets; b += 256) hist[b] = 0;
    const AggChunk x = agg_chunk(g, a, comms, count, blockIdx.x, mb, sh);
    const uint32_t nb = p.nbk[x.ci];
    uint32_t kb, span;
    agg_keys(a, p.nn, x.c, kb, span);
    constexpr uint32_t PER = kAggChunk / 256;
    // each dependent load level for all of the thread's entries at once:
    // member, row start, column + weight, new id of the neighbour's
    // community.  Positions past the chunk end repeat its first entry, so no
    // load waits on a branch.
    uint32_t vv[PER], jj[PER], kk[PER];
    float ww[PER];
#pragma unroll
    for (uint32_t u = 0; u < PER; ++u) {
        const uint32_t fu = x.f0 + threadIdx.x + 256u * u;
        const uint32_t f = fu < x.f1 ? fu : x.f0;
        uint32_t l = 0, h = x.nm;  // member of flat entry f
        while (h - l > 1) {
            const uint32_t mid = (l + h) >> 1;
            if (mb[mid] <= f) l = mid;
            else h = mid;
        }
        vv[u] = a.members[x.i0 + l];
        jj[u] = f - mb[l];
    }
#pragma unroll
    for (uint32_t u = 0; u < PER; ++u) jj[u] += g.row[vv[u]];
#pragma unroll
    for (uint32_t u = 0; u < PER; ++u) {
        kk[u] = g.col[jj[u]];
        ww[u] = g.w[jj[u]];
    }
#pragma unroll
    for (uint32_t u = 0; u < PER; ++u) kk[u] = a.vnid[kk[u]];
#pragma unroll
    for (uint32_t u = 0; u < PER; ++u) {
        const uint32_t f = x.f0 + threadIdx.x + 256u * u;
        if (f >= x.f1) continue;
        p.fkey[f] = kk[u];
        p.fw[f] = ww[u];
        atomicAdd(&hist[agg_bucket(kk[u] - kb, nb, span)], 1u);
    }
    __syncthreads();
    const uint32_t b0 = p.bstart[x.ci];
    for (uint32_t b = threadIdx.x; b < nb; b += 256)
        if (hist[b]) atomicAdd(&p.bcnt[b0 + b], hist[b]);
}

// P1: the chunk's entries ordered by bucket in LDS, then each bucket's run
// written to consecutive addresses of its range
__global__ __launch_bounds__(256) void k_aggp_scatter(AggView a, AggPart p, const uint32_t *comms,
                                                      uint32_t count) {
    constexpr uint32_t PER = kAggMaxBuckets / 256;
    static_assert(PER % 2 == 0 && kAggChunk < 65536, "16-bit counts, in pairs");
    // per bucket: entries of the chunk (16-bit halves of 32-bit words, so a
    // 32-bit LDS atomic counts two buckets), their offset in the
    // bucket-ordered chunk, their position in the bucket's range
    __shared__ uint32_t lcnt2[kAggMaxBuckets / 2];
    __shared__ uint16_t loff[kAggMaxBuckets];
    __shared__ uint32_t lbase[kAggMaxBuckets];
    __shared__ uint32_t sk[kAggChunk];
    __shared__ float sw[kAggChunk];
    uint32_t *wsum = sk;  // scan partials; read before the staging below starts
    auto count_of = [&](uint32_t b) { return (lcnt2[b >> 1] >> ((b & 1u) << 4)) & 0xffffu; };
    uint32_t lo = 0, hi = count;  // chunk -> large community
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (a.coff[mid] <= blockIdx.x) lo = mid;
        else hi = mid;
    }
    const uint32_t ci = lo, c = comms[ci];
    const uint32_t cb = a.epos[c], ce = cb + a.comm_deg[c];
    const uint32_t f0 = cb + (blockIdx.x - a.coff[ci]) * kAggChunk;
    const uint32_t cnt = min(ce, f0 + kAggChunk) - f0;
    const uint32_t nb = p.nbk[ci];
    uint32_t kb, span;
    agg_keys(a, p.nn, c, kb, span);
    for (uint32_t b = threadIdx.x; b < kAggMaxBuckets / 2; b += 256) lcnt2[b] = 0;
    __syncthreads();
    uint32_t key[kAggChunk / 256], rank[kAggChunk / 256], bk[kAggChunk / 256];
    float wq[kAggChunk / 256];
    // all loads first (past the chunk end: its first entry), then the ranks
#pragma unroll
    for (uint32_t q = 0; q < kAggChunk / 256; ++q) {
        const uint32_t i = threadIdx.x + 256u * q;
        const uint32_t f = f0 + (i < cnt ? i : 0u);
        key[q] = p.fkey[f];
        wq[q] = p.fw[f];
    }
#pragma unroll
    for (uint32_t q = 0; q < kAggChunk / 256; ++q) {
        const uint32_t i = threadIdx.x + 256u * q;
        bk[q] = kEmpty;
        if (i < cnt) {
            bk[q] = agg_bucket(key[q] - kb, nb, span);
            const uint32_t sh = (bk[q] & 1u) << 4;
            rank[q] = (atomicAdd(&lcnt2[bk[q] >> 1], 1u << sh) >> sh) & 0xffffu;
        }
    }
    __syncthreads();
    // exclusive scan of the counts: thread t owns buckets PER*t .. PER*t+PER-1
    const uint32_t t = threadIdx.x, lane = t & 63, wv = t >> 6;
    uint32_t own = 0;
    for (uint32_t k = 0; k < PER; ++k) own += count_of(PER * t + k);
    uint32_t inc = own;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(inc, o, 64);
        if (lane >= (uint32_t)o) inc += y;
    }
    if (lane == 63) wsum[wv] = inc;
    const uint32_t b0 = p.bstart[ci];
    for (uint32_t b = t; b < nb; b += 256) {
        const uint32_t cb2 = count_of(b);
        if (cb2) lbase[b] = atomicAdd(&p.bcur[b0 + b], cb2);
    }
    __syncthreads();
    uint32_t run = inc - own;
    for (uint32_t k = 0; k < wv; ++k) run += wsum[k];
    for (uint32_t k = 0; k < PER; ++k) {
        loff[PER * t + k] = (uint16_t)run;
        run += count_of(PER * t + k);
    }
    __syncthreads();
#pragma unroll
    for (uint32_t q = 0; q < kAggChunk / 256; ++q) {
        if (bk[q] == kEmpty) continue;
        const uint32_t sp = loff[bk[q]] + rank[q];
        sk[sp] = key[q];
        sw[sp] = wq[q];
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < cnt; i += 256) {
        const uint32_t k = sk[i];
        const uint32_t b = agg_bucket(k - kb, nb, span);
        const uint32_t pos = lbase[b] + (i - loff[b]);
        p.pkey[pos] = k;
        p.pw[pos] = sw[i];
    }
}

__global__ __launch_bounds__(256) void k_aggp_bucket(DevView g, AggView a, AggPart p,
                                                     const uint32_t *comms) {
    static_assert(kAggBucketCap > (int)kAggBucketDistinct, "table shape");
    constexpr uint32_t DMAX = kAggBucketDistinct;
    __shared__ uint32_t key[kAggBucketCap];
    __shared__ unsigned long long val[kAggBucketCap];
    __shared__ uint16_t dslot[DMAX];
    __shared__ uint32_t skeys[DMAX], sslot[DMAX], oslot[DMAX];  // bins, or the sort
    __shared__ uint32_t wsum[4], maxbin;
    __shared__ uint32_t nd_s, over;
    const uint32_t b = blockIdx.x;
    const uint32_t ci = p.bowner[b], c = comms[ci];
    const uint32_t q0 = p.bpos[b], q1 = p.bpos[b + 1];
    if (threadIdx.x == 0) p.bnd[b] = 0;
    if (q0 == q1) return;  // uniform
    // where the bucket's output goes and its key range, loaded now so their
    // latency overlaps the inserts
    const uint32_t bs = p.bstart[ci];
    const uint32_t out = a.epos[c] + (q0 - p.bpos[bs]);
    const unsigned long long nb = p.nbk[ci], bb = b - bs;
    for (uint32_t i = threadIdx.x; i < kAggBucketCap; i += 256) {
        key[i] = kEmpty;
        val[i] = 0ull;
    }
    if (threadIdx.x == 0) {
        nd_s = 0;
        over = 0;
    }
    __syncthreads();
    // entries in batches of 8 per thread, loaded before any is inserted
    for (uint32_t base = q0; base < q1; base += 256u * 8u) {
        uint32_t kk[8];
        float ww[8];
#pragma unroll
        for (uint32_t u = 0; u < 8; ++u) {
            const uint32_t q = base + threadIdx.x + 256u * u;
            kk[u] = kEmpty;
            if (q < q1) {
                kk[u] = p.pkey[q];
                ww[u] = p.pw[q];
            }
        }
#pragma unroll
        for (uint32_t u = 0; u < 8; ++u) {
            // the wave's entries of its first key are summed before the table
            // sees them: a key that dominates a bucket (the community's own
            // id, in graphs with few communities) costs one LDS atomic per
            // wave instead of 64 on one slot (every lane takes part)
            uint32_t k = kk[u];
            unsigned long long wf = k == kEmpty ? 0ull : f2fix(ww[u], g.scale);
            const unsigned long long act = __ballot(k != kEmpty);
            if (act) {  // uniform
                const int lead = __ffsll((long long)act) - 1;
                const uint32_t k0 = __shfl(k, lead, 64);
                const bool same = k != kEmpty && k == k0;
                const unsigned long long s = wave_sum_u64(same ? wf : 0ull);
                if ((int)(threadIdx.x & 63u) == lead) wf = s;
                else if (same) k = kEmpty;  // folded into the leader's entry
            }
            if (k == kEmpty) continue;
            uint32_t slot = hslot_n<kAggBucketCap>(k);
            int probe = 0;
            for (; probe < kAggBucketCap; ++probe) {
                const uint32_t prev = atomicCAS(&key[slot], kEmpty, k);
                if (prev == kEmpty) {
                    const uint32_t idx = atomicAdd(&nd_s, 1u);
                    if (idx < p.dmax) dslot[idx] = (uint16_t)slot;
                    else over = 1u;  // more distinct keys than a bucket holds
                    break;
                }
                if (prev == k) break;
                slot = slot + 1 == (uint32_t)kAggBucketCap ? 0u : slot + 1;
            }
            if (probe == kAggBucketCap) {
                over = 1u;
                continue;
            }
            atomicAdd(&val[slot], wf);
        }
    }
    __syncthreads();
    if (over) {  // uniform: the community goes to the global-table merge
        if (threadIdx.x == 0 && atomicOr(&p.cflag[ci], 1u) == 0u) atomicAdd(p.nflag, 1u);
        return;
    }
    const uint32_t nd = nd_s;
    // the bucket's key range: keys k with floor((k - kb) * nb / span) == bb
    uint32_t kb, span;
    agg_keys(a, p.nn, c, kb, span);
    const uint32_t lo = kb + (uint32_t)((bb * span + nb - 1) / nb);
    const uint32_t hi = kb + (uint32_t)(((bb + 1) * span + nb - 1) / nb);
    auto keyof = [&](uint32_t i) { return key[dslot[i]]; };
    auto emit = [&](uint32_t pos, uint32_t i) {
        const uint32_t slot = dslot[i];
        a.tkey[out + pos] = key[slot];
        a.tval[out + pos] = val[slot];
    };
    if (!lds_bin_rank<DMAX, 256>(nd, lo, hi - lo, keyof, skeys, sslot, oslot, wsum, &maxbin,
                                 emit)) {
        uint32_t n2 = 1;
        while (n2 < nd) n2 <<= 1;
        for (uint32_t i = threadIdx.x; i < n2; i += 256) {
            skeys[i] = i < nd ? key[dslot[i]] : kEmpty;
            sslot[i] = i < nd ? dslot[i] : 0u;
        }
        __syncthreads();
        lds_bitonic<256>(skeys, sslot, n2);
        for (uint32_t i = threadIdx.x; i < nd; i += 256) {
            a.tkey[out + i] = skeys[i];
            a.tval[out + i] = val[sslot[i]];
        }
    }
    if (threadIdx.x == 0) {
        p.bnd[b] = nd;
        atomicAdd(&a.cnt[a.new_id[c]], nd);
    }
}

// communities that overflowed restart on the global-table merge: their
// partial distinct counts are dropped
__global__ void k_aggp_reset_flagged(AggView a, AggPart p, const uint32_t *comms,
                                     uint32_t count) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count && p.cflag[i]) a.cnt[a.new_id[comms[i]]] = 0;
}
// table capacities for flagged communities only (others get none)
__global__ void k_aggp_caps(AggView a, AggPart p, const uint32_t *comms, uint32_t count,
                            unsigned long long *caps) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) {
        // 2*deg slots for a flagged community (large: deg > 2048), none else
        caps[i] = p.cflag[i] ? 2ull * a.comm_deg[comms[i]] : 0ull;
    }
}

// sorted distinct lists of the bucket outputs -> the new CSR (one workgroup
// per bucket; buckets of flagged communities are emitted per community)
__global__ __launch_bounds__(256) void k_aggp_emit(DevView g, AggView a, AggPart p,
                                                   const uint32_t *comms, EmitTarget t) {
    const uint32_t b = blockIdx.x;
    const uint32_t ci = p.bowner[b];
    if (p.cflag[ci]) return;
    const uint32_t nd = p.bnd[b];
    if (!nd) return;
    const uint32_t c = comms[ci], first = p.bstart[ci];
    const uint32_t src = a.epos[c] + (p.bpos[b] - p.bpos[first]);
    const uint32_t nid = a.new_id[c];
    const uint32_t dst = t.rows[t.map ? t.map[nid] : nid] + (p.boff[b] - p.boff[first]);
    for (uint32_t i = threadIdx.x; i < nd; i += 256) {
        const uint32_t k = a.tkey[src + i];
        t.col[dst + i] = t.map ? t.map[k] : k;
        t.w[dst + i] = fix2f(a.tval[src + i], g.S);
    }
}

void launch_aggp(hipStream_t s, const DevView &g, AggView a, const AggPart &p,
                 const uint32_t *comms, uint32_t count, int phase) {
    if (!count) return;
    switch (phase) {
        case 0:
            hipLaunchKernelGGL(k_aggp_count, dim3(a.nchunks), dim3(256), 0, s, g, a, p, comms,
                               count);
            break;
        case 1:
            hipLaunchKernelGGL(k_aggp_scatter, dim3(a.nchunks), dim3(256), 0, s, a, p, comms,
                               count);
            break;
        case 2:
            hipLaunchKernelGGL(k_aggp_bucket, dim3(p.NB), dim3(256), 0, s, g, a, p, comms);
            break;
        case 3:
            hipLaunchKernelGGL(k_aggp_reset_flagged, dim3((count + 255) / 256), dim3(256), 0, s, a,
                               p, comms, count);
            break;
        default:
            break;
    }
}
void launch_aggp_caps(hipStream_t s, AggView a, const AggPart &p, const uint32_t *comms,
                      uint32_t count, unsigned long long *caps) {
    if (!count) return;
    hipLaunchKernelGGL(k_aggp_caps, dim3((count + 255) / 256), dim3(256), 0, s, a, p, comms,
                       count, caps);
}
void launch_aggp_emit(hipStream_t s, const DevView &g, AggView a, const AggPart &p,
                      const uint32_t *comms, const EmitTarget &t) {
    if (!p.NB) return;
    hipLaunchKernelGGL(k_aggp_emit, dim3(p.NB), dim3(256), 0, s, g, a, p, comms, t);
}

// ---------------------------------------------------------------------------
// The whole contraction of a tiny pass graph (at most kTinyN vertices and
// kTinyE entries) in one workgroup: new ids (non-empty communities in id
// order, aggregation.cu:21-36,390-393), every entry keyed by (new id of its
// row's community, new id of its column's community) and sorted in LDS,
// exact sums per key, rows in ascending neighbour id (the order the merge
// kernels give), and init_comm composed through comm
// (aggregation.cu:380-385).  One launch instead of about thirty: the
// boundaries between the passes of a small graph are bound by the host's
// launches.
// ---------------------------------------------------------------------------
template <int NT>
__device__ __forceinline__ uint32_t wg_excl_scan(uint32_t x, uint32_t &total, uint32_t *wsum) {
    const uint32_t lane = threadIdx.x & 63u, wv = threadIdx.x >> 6;
    uint32_t v = x;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(v, o, 64);
        if (lane >= (uint32_t)o) v += y;
    }
    if (lane == 63u) wsum[wv] = v;
    __syncthreads();
    uint32_t off = 0, tot = 0;
#pragma unroll
    for (int i = 0; i < NT / 64; ++i) {
        const uint32_t s = wsum[i];
        if ((uint32_t)i < wv) off += s;
        tot += s;
    }
    __syncthreads();
    total = tot;
    return off + v - x;
}

__global__ __launch_bounds__(kTinyThreads) void k_agg_tiny(DevView g, uint32_t n0,
                                                           uint32_t *init_comm, uint32_t *orow,
                                                           uint32_t *ocol, float *ow,
                                                           uint32_t *sizes, uint32_t *new_id) {
    constexpr uint32_t PER = kTinyE / kTinyThreads;
    __shared__ uint32_t nid[kTinyN];
    __shared__ uint32_t cnt[kTinyN];
    __shared__ uint32_t key[kTinyE];
    __shared__ uint16_t idx[kTinyE];
    __shared__ unsigned long long val[kTinyE];
    __shared__ uint32_t wsum[kTinyThreads / 64];
    const uint32_t t = threadIdx.x, n = g.n;
    // new ids: one community per thread
    uint32_t nn;
    {
        const uint32_t f = (t < n && g.comm_sz[t] != 0u) ? 1u : 0u;
        const uint32_t e = wg_excl_scan<kTinyThreads>(f, nn, wsum);
        if (t < n) nid[t] = new_id[t] = e;
        cnt[t] = 0u;
    }
    __syncthreads();
    // (row community, column community) of every entry: one vertex per thread
    const uint32_t ne = g.row[n];
    if (t < n) {
        const uint32_t a = nid[g.comm[t]] * nn;
        for (uint32_t j = g.row[t], je = g.row[t + 1]; j < je; ++j) {
            key[j] = a + nid[g.comm[g.col[j]]];
            val[j] = f2fix(g.w[j], g.scale);
        }
    }
    uint32_t np2 = 1;
    while (np2 < ne) np2 <<= 1;
    for (uint32_t i = t; i < np2; i += kTinyThreads) {
        idx[i] = (uint16_t)i;
        if (i >= ne) key[i] = kEmpty;  // above every key (a key is below kTinyN^2)
    }
    __syncthreads();
    // bitonic sort of the keys (with their entry index) in LDS
    for (uint32_t k = 2; k <= np2; k <<= 1)
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            for (uint32_t i = t; i < np2; i += kTinyThreads) {
                const uint32_t l = i ^ j;
                if (l > i) {
                    const uint32_t ki = key[i], kl = key[l];
                    if ((ki > kl) == ((i & k) == 0u)) {
                        key[i] = kl;
                        key[l] = ki;
                        const uint16_t x = idx[i];
                        idx[i] = idx[l];
                        idx[l] = x;
                    }
                }
            }
            __syncthreads();
        }
    // one output entry per run of equal keys, in key order: rows by new id,
    // columns ascending within a row
    uint32_t h[PER], s = 0;
#pragma unroll
    for (uint32_t q = 0; q < PER; ++q) {
        const uint32_t i = t * PER + q;
        h[q] = (i < ne && (i == 0u || key[i] != key[i - 1u])) ? 1u : 0u;
        s += h[q];
    }
    uint32_t ne2;
    uint32_t pos = wg_excl_scan<kTinyThreads>(s, ne2, wsum);
#pragma unroll
    for (uint32_t q = 0; q < PER; ++q) {
        if (!h[q]) continue;
        const uint32_t i = t * PER + q, kk = key[i];
        unsigned long long sum = 0;
        for (uint32_t r = i; r < ne && key[r] == kk; ++r) sum += val[idx[r]];
        const uint32_t a = kk / nn;
        ocol[pos] = kk - a * nn;
        ow[pos] = fix2f(sum, g.S);
        atomicAdd(&cnt[a], 1u);
        ++pos;
    }
    __syncthreads();
    // row pointers; communities whose members have no entries get empty rows
    {
        uint32_t tot;
        const uint32_t rp = wg_excl_scan<kTinyThreads>(t < nn ? cnt[t] : 0u, tot, wsum);
        if (t < nn) orow[t] = rp;
        if (t == 0) {
            orow[nn] = tot;
            sizes[0] = nn;
            sizes[1] = tot;
        }
    }
    // init_comm holds vertices of this pass: their community's new id (n0
    // is 0 when the input graph is large: a launch of its own composes it)
    for (uint32_t i = t; i < n0; i += kTinyThreads) init_comm[i] = nid[g.comm[init_comm[i]]];
}
// an input graph larger than this composes init_comm in a launch of its own
// (one workgroup would walk it alone)
constexpr uint32_t kComposeInKernel = 8192;
__global__ void k_agg_compose(const uint32_t *comm, const uint32_t *new_id, const uint32_t *map,
                              uint32_t *init_comm, uint32_t n0);
void launch_agg_tiny(hipStream_t s, const DevView &g, uint32_t n0, uint32_t *init_comm,
                     uint32_t *orow, uint32_t *ocol, float *ow, uint32_t *sizes,
                     uint32_t *new_id) {
    const bool in_kernel = n0 <= kComposeInKernel;
    hipLaunchKernelGGL(k_agg_tiny, dim3(1), dim3(kTinyThreads), 0, s, g, in_kernel ? n0 : 0u,
                       init_comm, orow, ocol, ow, sizes, new_id);
    if (!in_kernel)
        hipLaunchKernelGGL(k_agg_compose, dim3((n0 + 255) / 256), dim3(256), 0, s, g.comm, new_id,
                           nullptr, init_comm, n0);
}

// ---------------------------------------------------------------------------
// Contraction into few communities (at most kFewNN non-empty ones, a pass
// graph of at most kFewN vertices, single graph, input numbering): the
// contracted graph is a dense kFewNN x kFewNN matrix of exact sums.  Three
// launches instead of about thirty: new ids (one workgroup), the entries
// added into per-workgroup LDS matrices and then into one global matrix,
// and the CSR written from the matrix in ascending columns (one workgroup).
// A pair of communities joined only by entries of weight zero still gets its
// entry (as in the general path), so a presence bitmap follows the matrix.
// More communities than kFewNN: every kernel returns at entry and the host
// takes the general path.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(1024) void k_few_ids(DevView g, uint32_t *new_id,
                                                  unsigned long long *mat, uint32_t *sizes) {
    __shared__ uint32_t wsum[16];
    const uint32_t n = g.n;
    uint32_t carry = 0;
    for (uint32_t c0 = 0; c0 < n; c0 += 1024) {
        const uint32_t c = c0 + threadIdx.x;
        const uint32_t f = (c < n && g.comm_sz[c] != 0u) ? 1u : 0u;
        uint32_t tot;
        const uint32_t e = wg_excl_scan<1024>(f, tot, wsum);
        if (c < n) new_id[c] = carry + e;
        carry += tot;
    }
    // the matrix and, behind it, its presence bitmap
    for (uint32_t i = threadIdx.x; i < kFewNN * kFewNN + kFewNN * kFewNN / 64; i += 1024)
        mat[i] = 0ull;
    if (threadIdx.x == 0) sizes[0] = carry;
}

__global__ __launch_bounds__(256) void k_few_sums(DevView g, const uint32_t *new_id,
                                                  unsigned long long *mat,
                                                  const uint32_t *sizes) {
    __shared__ unsigned long long lm[kFewNN * kFewNN];
    __shared__ uint32_t lp[kFewNN * kFewNN / 32];
    __shared__ uint32_t lrow[kFewRows + 1];
    const uint32_t nn = sizes[0];
    if (nn > kFewNN) return;  // uniform: the general path takes over
    uint32_t *pres = reinterpret_cast<uint32_t *>(mat + kFewNN * kFewNN);
    // the workgroup's rows [r0, r1): their offsets in LDS, their entries
    // [lrow[0], lrow[r1 - r0]) shared out over the threads
    const uint32_t r0 = blockIdx.x * kFewRows, r1 = min(g.n, r0 + kFewRows), nr = r1 - r0;
    for (uint32_t i = threadIdx.x; i <= nr; i += 256) lrow[i] = g.row[r0 + i];
    for (uint32_t i = threadIdx.x; i < nn * nn; i += 256) lm[i] = 0ull;
    for (uint32_t i = threadIdx.x; i < kFewNN * kFewNN / 32; i += 256) lp[i] = 0u;
    __syncthreads();
    for (uint32_t j = lrow[0] + threadIdx.x; j < lrow[nr]; j += 256) {
        uint32_t lo = 0, hi = nr;  // lrow[lo] <= j < lrow[hi]: the last such lo owns j
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (lrow[mid] <= j) lo = mid;
            else hi = mid;
        }
        const uint32_t a = new_id[g.comm[r0 + lo]], b = new_id[g.comm[g.col[j]]];
        const uint32_t ab = a * nn + b;
        atomicAdd(&lm[ab], f2fix(g.w[j], g.scale));
        atomicOr(&lp[ab >> 5], 1u << (ab & 31u));
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < nn * nn; i += 256)
        if (lm[i]) atomicAdd(&mat[i], lm[i]);
    for (uint32_t i = threadIdx.x; i < kFewNN * kFewNN / 32; i += 256)
        if (lp[i]) atomicOr(&pres[i], lp[i]);
}

__global__ __launch_bounds__(1024) void k_few_emit(DevView g, const uint32_t *new_id,
                                                   const unsigned long long *mat, uint32_t n0,
                                                   uint32_t *init_comm, uint32_t *orow,
                                                   uint32_t *ocol, float *ow, uint32_t *sizes) {
    __shared__ uint32_t cnt[kFewNN];
    __shared__ uint32_t start[kFewNN + 1];
    const uint32_t nn = sizes[0];
    if (nn > kFewNN) return;
    const uint32_t *pres = reinterpret_cast<const uint32_t *>(mat + kFewNN * kFewNN);
    auto present = [&](uint32_t ab) { return (pres[ab >> 5] >> (ab & 31u)) & 1u; };
    const uint32_t t = threadIdx.x;
    if (t < nn) {  // one row per thread: its columns with entries
        uint32_t c = 0;
        for (uint32_t b = 0; b < nn; ++b) c += present(t * nn + b);
        cnt[t] = c;
    }
    __syncthreads();
    if (t == 0) {
        uint32_t r = 0;
        for (uint32_t a = 0; a < nn; ++a) {
            start[a] = r;
            r += cnt[a];
        }
        start[nn] = r;
        sizes[1] = r;
    }
    __syncthreads();
    if (t <= nn) orow[t] = start[t];
    if (t < nn) {
        uint32_t o = start[t];
        for (uint32_t b = 0; b < nn; ++b) {
            if (present(t * nn + b)) {
                ocol[o] = b;
                ow[o] = fix2f(mat[t * nn + b], g.S);
                ++o;
            }
        }
    }
    for (uint32_t i = t; i < n0; i += 1024) init_comm[i] = new_id[g.comm[init_comm[i]]];
}

// init_comm of a large input graph, when the pass left at most kFewNN
// communities
__global__ void k_few_compose(const uint32_t *comm, const uint32_t *new_id, uint32_t *init_comm,
                              uint32_t n0, const uint32_t *sizes) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (sizes[0] <= kFewNN && i < n0) init_comm[i] = new_id[comm[init_comm[i]]];
}
void launch_agg_few(hipStream_t s, const DevView &g, uint32_t *new_id, unsigned long long *mat,
                    uint32_t n0, uint32_t *init_comm, uint32_t *orow, uint32_t *ocol, float *ow,
                    uint32_t *sizes) {
    const bool in_kernel = n0 <= kComposeInKernel;
    hipLaunchKernelGGL(k_few_ids, dim3(1), dim3(1024), 0, s, g, new_id, mat, sizes);
    hipLaunchKernelGGL(k_few_sums, dim3((g.n + kFewRows - 1) / kFewRows), dim3(256), 0, s, g,
                       new_id, mat, sizes);
    hipLaunchKernelGGL(k_few_emit, dim3(1), dim3(1024), 0, s, g, new_id, mat,
                       in_kernel ? n0 : 0u, init_comm, orow, ocol, ow, sizes);
    if (!in_kernel)
        hipLaunchKernelGGL(k_few_compose, dim3((n0 + 255) / 256), dim3(256), 0, s, g.comm, new_id,
                           init_comm, n0, sizes);
}

// compressEdges (aggregation.cu:359-378): one wave per community copies its
// sorted distinct list into the new CSR with renumbered ids and FP32 weights
__global__ __launch_bounds__(256) void k_agg_emit(DevView g, AggView a, const uint32_t *comms,
                                                  uint32_t count, const uint32_t *only,
                                                  EmitTarget t) {
    const uint32_t ci = blockIdx.x * 4 + (threadIdx.x >> 6);
    const uint32_t lane = threadIdx.x & 63;
    if (ci >= count || (only && !only[ci])) return;
    const uint32_t c = comms[ci];
    const uint32_t nid = a.new_id[c];
    const uint32_t o = t.rows[t.map ? t.map[nid] : nid];
    const uint32_t nd = t.rank_rows[nid + 1] - t.rank_rows[nid];
    const uint32_t src = a.epos[c];
    for (uint32_t i = lane; i < nd; i += 64) {
        const uint32_t k = a.tkey[src + i];  // already renumbered
        t.col[o + i] = t.map ? t.map[k] : k;
        t.w[o + i] = fix2f(a.tval[src + i], g.S);
    }
}
void launch_agg_emit(hipStream_t s, const DevView &g, AggView a, const uint32_t *comms,
                     uint32_t count, const uint32_t *only, const EmitTarget &t) {
    if (!count) return;
    hipLaunchKernelGGL(k_agg_emit, dim3((count + 3) / 4), dim3(256), 0, s, g, a, comms, count,
                       only, t);
}

// graphVertToComm + graphVertToNewComm (modularity.cu:51-56,
// aggregation.cu:380-385) fused: init_comm <- new_id[comm[init_comm]]
__global__ void k_agg_vnid(const uint32_t *comm, const uint32_t *new_id, uint32_t n,
                           uint32_t *vnid) {
    const uint32_t u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u < n) vnid[u] = new_id[comm[u]];
}
void launch_agg_vnid(hipStream_t s, const uint32_t *comm, const uint32_t *new_id, uint32_t n,
                     uint32_t *vnid) {
    if (!n) return;
    hipLaunchKernelGGL(k_agg_vnid, dim3((n + 255) / 256), dim3(256), 0, s, comm, new_id, n, vnid);
}
__global__ void k_agg_compose(const uint32_t *comm, const uint32_t *new_id, const uint32_t *map,
                              uint32_t *init_comm, uint32_t n0) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n0) {
        const uint32_t u = new_id[comm[init_comm[i]]];
        init_comm[i] = map ? map[u] : u;
    }
}
void launch_agg_compose(hipStream_t s, const uint32_t *comm, const uint32_t *new_id,
                        const uint32_t *map, uint32_t *init_comm, uint32_t n0) {
    if (!n0) return;
    hipLaunchKernelGGL(k_agg_compose, dim3((n0 + 255) / 256), dim3(256), 0, s, comm, new_id, map,
                       init_comm, n0);
}

// ---------------------------------------------------------------------------
// The pass loop's decisions on the device (modularity.cu:501-508 per graph):
// the FP32 modularity from the exact counters, in the reference's
// expression (modularity.cu:428-436), then the stop rule.
// ---------------------------------------------------------------------------

// Correctly rounded (nearest-even) u128 -> f32.
__device__ __forceinline__ float u128_to_f32_rn(unsigned __int128 x) {
    const unsigned long long hi = (unsigned long long)(x >> 64);
    if (!hi) return u64_to_f32_rn((unsigned long long)x);
    const int lz = __clzll(hi);  // the top bit moves to bit 127
    const unsigned __int128 y = x << lz;
    const unsigned long long h = (unsigned long long)(y >> 64), l = (unsigned long long)y;
    unsigned long long mant = h >> 40;  // 24 bits
    const unsigned long long rest = h & ((1ull << 40) - 1ull), half = 1ull << 39;
    if (rest > half || (rest == half && (l != 0ull || (mant & 1ull)))) mant++;
    return ldexpf((float)(uint32_t)mant, (127 - lz) - 23);
}

// Q = S_in/(2m) - sum a^2/((4m) m): S_in and sum a^2 (at scale 2^2S, four
// 32-bit limbs each summed in a u64) exact, each rounded once
__device__ __forceinline__ float q_exact(unsigned long long sin, const unsigned long long *limb,
                                         size_t stride, int S, float m) {
    unsigned __int128 sq = 0;
#pragma unroll
    for (int l = 0; l < 4; ++l) sq += (unsigned __int128)limb[l * stride] << (32 * l);
    const float s_in = fix2f(sin, S);
    const float a2 = ldexpf(u128_to_f32_rn(sq), -2 * S);
    return (s_in / (2.0f * m)) - (a2 / ((4.0f * m) * m));
}

// one sweep's end: old/act, the first sweep's verdict, the stop rule (a
// sweep without moves is a fixed point and also ends the loop)
__device__ __forceinline__ bool stop_rule(float q, unsigned long long moved, float min_gain,
                                          uint32_t max_iters, float &act, float &old,
                                          uint32_t &it, uint32_t &first, uint32_t &ret,
                                          float &q_print) {
    old = act;
    act = q;
    ++it;
    const bool stop = !(act - old >= min_gain) || moved == 0ull || it >= max_iters;
    if (first && (act - old < min_gain || moved == 0ull)) {
        q_print = old;
        ret = 0u;
    }
    first = 0u;
    return stop;
}

// The stop rule of the pass loop (modularity.cu:503-508) after a sweep,
// from the collapsed exact counters tot = {S_in, sum a^2 limbs[4], moved}.
__device__ __forceinline__ void sweep_decide(const unsigned long long *tot,
                                             const uint32_t *err_flag, int S, float m,
                                             float min_gain, uint32_t max_iters, PassState *ps) {
    if (ps->done) return;
    if (*err_flag) {
        ps->err = 1u;
        ps->done = 1u;
        return;
    }
    PassState p = *ps;
    const float q = q_exact(tot[0], tot + 1, 1, S, m);
    if (stop_rule(q, tot[5], min_gain, max_iters, p.act, p.old, p.it, p.first, p.ret, p.q_print))
        p.done = 1u;
    *ps = p;
}

// After a sweep's sums: the spread counters (6 groups) collapsed into tot,
// then the stop rule, in one workgroup (one launch instead of two).
// The spread slots are left zero for the next sweep, and the loop state is
// also written to the host's mapped copy, which the host reads after the
// launch (no copy launch).
__device__ void decide_front(const SweepTail &t, int S) {
    __shared__ unsigned long long part[16];
    const uint32_t nw = (blockDim.x + 63u) >> 6;
    for (int gi = 0; gi < 6; ++gi) {
        unsigned long long x = 0;
        for (uint32_t i = threadIdx.x; i < kSpread; i += blockDim.x) {
            x += t.spread[gi * kSpread + i];
            t.spread[gi * kSpread + i] = 0ull;
        }
        x = wave_sum_u64(x);
        if ((threadIdx.x & 63u) == 0) part[threadIdx.x >> 6] = x;
        __syncthreads();
        if (threadIdx.x == 0) {
            unsigned long long s = 0;
            for (uint32_t w = 0; w < nw; ++w) s += part[w];
            t.tot[gi] = s;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        sweep_decide(t.tot, t.err, S, t.m, t.min_gain, t.max_iters, t.ps);
        *t.ps_host = *t.ps;
        __threadfence_system();
    }
}
__global__ __launch_bounds__(256) void k_collapse_decide(SweepTail t, int S) {
    decide_front(t, S);
}
__global__ __launch_bounds__(256) void k_decide_from(const SweepTail *t, int S) {
    decide_front(*t, S);
}
void launch_collapse_decide(hipStream_t s, const SweepTail &t, int S) {
    hipLaunchKernelGGL(k_collapse_decide, dim3(1), dim3(256), 0, s, t, S);
}

__global__ __launch_bounds__(256) void k_batch_m(uint32_t G, const unsigned long long *acc, int S,
                                                 const uint32_t *voff, const uint32_t *row,
                                                 uint32_t *gstate, float *gm, GraphState *gst,
                                                 int range_status) {
    const uint32_t gi = blockIdx.x * 256 + threadIdx.x;
    if (gi >= G || gstate[gi] != kGraphActive) return;
    GraphState &b = gst[gi];
    b.ne = row[voff[gi + 1]] - row[voff[gi]];
    // all_w = reduce(neigh_w) / 2 (modularity.cu:457-458): exact sum, one
    // rounding, exact halving
    const float m = fix2f(acc[gi], S) / 2.0f;
    gm[gi] = m;
    b.m = m;
    if (!(m > 0.0f)) {  // no weight: the reference would divide by zero
        b.status = (uint32_t)range_status;
        gstate[gi] = kGraphDone;
    }
}
void launch_batch_m(hipStream_t s, uint32_t G, const unsigned long long *acc, int S,
                    const uint32_t *voff, const uint32_t *row, uint32_t *gstate, float *gm,
                    GraphState *gst, int range_status) {
    hipLaunchKernelGGL(k_batch_m, dim3((G + 255) / 256), dim3(256), 0, s, G, acc, S, voff, row,
                       gstate, gm, gst, range_status);
}

__global__ __launch_bounds__(256) void k_batch_init(uint32_t G, const unsigned long long *gcnt,
                                                    int S, const uint32_t *gstate,
                                                    GraphState *gst) {
    const uint32_t gi = blockIdx.x * 256 + threadIdx.x;
    if (gi >= G || gstate[gi] != kGraphActive) return;
    GraphState &b = gst[gi];
    b.act = q_exact(gcnt[gi], gcnt + G + gi, G, S, b.m);
    b.first = 1u;
    b.ret = 1u;
    b.pass_iters = 0u;
}
void launch_batch_init(hipStream_t s, uint32_t G, const unsigned long long *gcnt, int S,
                       const uint32_t *gstate, GraphState *gst) {
    hipLaunchKernelGGL(k_batch_init, dim3((G + 255) / 256), dim3(256), 0, s, G, gcnt, S, gstate,
                       gst);
}

__global__ __launch_bounds__(256) void k_batch_decide(uint32_t G, unsigned long long *gcnt,
                                                      const uint32_t *err_flag, int S,
                                                      float min_gain, uint32_t max_iters,
                                                      const uint32_t *voff, uint32_t *gstate,
                                                      GraphState *gst, BatchSummary *sum) {
    const uint32_t gi = blockIdx.x * 256 + threadIdx.x;
    uint32_t state = 0;
    if (gi < G) {
        state = gstate[gi];
        if (state == kGraphActive) {
            GraphState b = gst[gi];
            const float q = q_exact(gcnt[gi], gcnt + G + gi, G, S, b.m);
            float qp = b.q_final;
            if (stop_rule(q, gcnt[5ull * G + gi], min_gain, max_iters, b.act, b.old,
                          b.pass_iters, b.first, b.ret, qp)) {
                b.iters += b.pass_iters;
                b.sum_edges += b.ne;
                b.sum_visits += (unsigned long long)(b.pass_iters + 1u) * b.ne;
                state = b.ret ? kGraphWait : kGraphDone;
                gstate[gi] = state;
            }
            b.q_final = qp;
            gst[gi] = b;
        }
        // the graph's exact counters start the next sweep at zero
        for (int f = 0; f < 6; ++f) gcnt[(size_t)f * G + gi] = 0ull;
    }
    const unsigned long long live =
        state == kGraphActive ? (unsigned long long)(voff[gi + 1] - voff[gi]) : 0ull;
    const unsigned long long a = wave_sum_u64(state == kGraphActive ? 1ull : 0ull);
    const unsigned long long w = wave_sum_u64(state == kGraphWait ? 1ull : 0ull);
    const unsigned long long l = wave_sum_u64(live);
    if ((threadIdx.x & 63) == 0) {
        if (a) atomicAdd(&sum->active, (uint32_t)a);
        if (w) atomicAdd(&sum->waiting, (uint32_t)w);
        if (l) atomicAdd(&sum->live, l);
    }
    if (gi == 0 && *err_flag) sum->err = 1u;
}
// No graph active: the remaining queued sweeps halt.  The summary goes to
// the host's mapped copy, which the host reads after the group, and its
// recounted fields are cleared for the next sweep's decide.
__global__ void k_batch_halt(BatchSummary *sum, BatchSummary *host) {
    if (sum->active == 0u) sum->halt = 1u;
    *host = *sum;
    __threadfence_system();
    sum->active = 0u;
    sum->waiting = 0u;
    sum->live = 0ull;
}
void launch_batch_halt(hipStream_t s, BatchSummary *sum, BatchSummary *host) {
    hipLaunchKernelGGL(k_batch_halt, dim3(1), dim3(1), 0, s, sum, host);
}
void launch_batch_decide(hipStream_t s, uint32_t G, unsigned long long *gcnt,
                         const uint32_t *err_flag, int S, float min_gain, uint32_t max_iters,
                         const uint32_t *voff, uint32_t *gstate, GraphState *gst,
                         BatchSummary *sum) {
    hipLaunchKernelGGL(k_batch_decide, dim3((G + 255) / 256), dim3(256), 0, s, G, gcnt, err_flag,
                       S, min_gain, max_iters, voff, gstate, gst, sum);
}

__global__ __launch_bounds__(256) void k_batch_reactivate(uint32_t G, uint32_t *gstate,
                                                          GraphState *gst) {
    const uint32_t gi = blockIdx.x * 256 + threadIdx.x;
    if (gi < G && gstate[gi] == kGraphWait) {
        gstate[gi] = kGraphActive;
        gst[gi].passes++;
    }
}
void launch_batch_reactivate(hipStream_t s, uint32_t G, uint32_t *gstate, GraphState *gst) {
    hipLaunchKernelGGL(k_batch_reactivate, dim3((G + 255) / 256), dim3(256), 0, s, G, gstate,
                       gst);
}

}  // namespace lv
