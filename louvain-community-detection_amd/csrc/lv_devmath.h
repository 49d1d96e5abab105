// lv_devmath.h -- device arithmetic shared by the kernel translation units:
// exact fixed point <-> FP32, the reference's gain expression and arg-max
// order, hashing, wave reductions and the 128-lane winner tree.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "lv_device.h"

namespace lv {

// ---------------------------------------------------------------------------
// arithmetic helpers
// ---------------------------------------------------------------------------

// Correctly rounded (nearest-even) u64 -> f32.
__device__ __forceinline__ float u64_to_f32_rn(unsigned long long u) {
    if (u < (1ull << 24)) return (float)(uint32_t)u;
    int lz = __clzll(u);
    int sh = 40 - lz;  // keep the top 24 bits
    unsigned long long mant = u >> sh;
    unsigned long long rem = u & ((1ull << sh) - 1ull);
    unsigned long long half = 1ull << (sh - 1);
    if (rem > half || (rem == half && (mant & 1ull))) mant++;
    return ldexpf((float)(uint32_t)mant, sh);
}

// exact fixed point (x * 2^-S) rounded once to f32
__device__ __forceinline__ float fix2f(unsigned long long x, int S) {
    return ldexpf(u64_to_f32_rn(x), -S);
}

// f32 (a multiple of 2^-S, >= 0) -> exact fixed point
__device__ __forceinline__ unsigned long long f2fix(float x, float scale) {
    return (unsigned long long)(x * scale);
}

// Gain of moving to community C (modularity.cu:149-152): FP32, this
// association, no contraction.
__device__ __forceinline__ float gain_of(float hw, float m, float kv, float acv, float ac) {
    float t1 = hw / m;
    float t3 = (acv - kv) - ac;
    float t4 = kv * t3;
    float t6 = (2.0f * m) * m;
    return t1 + t4 / t6;
}

// input-numbering id of a device id (identity without a locality order)
__device__ __forceinline__ uint32_t lid_of(const uint32_t *lid, uint32_t c) {
    return (lid && c != kEmpty) ? lid[c] : c;
}

// reduction order of the reference's arg-max (modularity.cu:169-179, 363);
// ids are compared in the input numbering, fetched only on a gain tie
__device__ __forceinline__ bool better(float ga, uint32_t ca, float gb, uint32_t cb,
                                       const uint32_t *lid) {
    return ga > gb || (ga == gb && lid_of(lid, ca) < lid_of(lid, cb));
}

// isLowerIdx (modularity.cu:85-93): between two singleton communities only a
// move to the lower id is allowed
__device__ __forceinline__ bool lower_idx_ok(const DevView &g, uint32_t szv, uint32_t c,
                                             uint32_t cv) {
    if (szv == 1u && g.comm_sz[c] == 1u) return lid_of(g.lid, c) < lid_of(g.lid, cv);
    return true;
}

template <int LOG2CAP>
__device__ __forceinline__ uint32_t hslot(uint32_t key) {
    return (key * 0x9E3779B1u) >> (32 - LOG2CAP);
}
// home slot in a table of any size CAP (multiply-high of the same hash; for a
// power of two it is hslot)
template <int CAP>
__device__ __forceinline__ uint32_t hslot_n(uint32_t key) {
    return (uint32_t)(((unsigned long long)(key * 0x9E3779B1u) * (unsigned)CAP) >> 32);
}

// Global hash tables (heavy rows in a sweep, large communities in the
// contraction; never live at the same time) share one arena of 16-byte slots
// {u32 key, u32 zero, u64 sum}, so a probe and its sum share a cache line.
// A table of d entries has exactly 2d slots (load factor <= 1/2 however many
// keys are distinct); the home slot is hslot_n's multiply-high with the
// capacity known at run time, and probing wraps at the capacity.  An empty
// slot is {kEmpty, 0}: every user returns its slots to that state.
__device__ __forceinline__ uint32_t *slot_key(unsigned long long *tab, uint32_t s) {
    return reinterpret_cast<uint32_t *>(tab + 2ull * s);
}
__device__ __forceinline__ unsigned long long *slot_sum(unsigned long long *tab, uint32_t s) {
    return tab + 2ull * s + 1ull;
}
__device__ __forceinline__ uint32_t hslot_cap(uint32_t key, uint32_t cap) {
    return (uint32_t)(((unsigned long long)(key * 0x9E3779B1u) * cap) >> 32);
}
__device__ __forceinline__ uint32_t next_slot(uint32_t s, uint32_t cap) {
    return s + 1u == cap ? 0u : s + 1u;
}

template <int N>
struct Log2 {
    static constexpr int value = 1 + Log2<N / 2>::value;
};
template <>
struct Log2<1> {
    static constexpr int value = 0;
};

// LDS writes of a wave's lanes visible to the wave's other lanes (a
// wave-level barrier: for tables private to one wave, no workgroup barrier)
__device__ __forceinline__ void wave_sync_lds() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// wave-level sum of u64
__device__ __forceinline__ unsigned long long wave_sum_u64(unsigned long long x) {
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
    return x;
}

// The 128-virtual-lane arg-max tree of computeMoveBlock/Global
// (modularity.cu:260-275) on one wave: lane p holds virtual lanes p and
// p+64.  Returns the winner (gain, comm, virtual lane) in lane 0.
__device__ __forceinline__ void tree128(float g0, uint32_t c0, float g1, uint32_t c1, float &bg,
                                        uint32_t &bc, uint32_t &bp, const uint32_t *lid) {
    const uint32_t p = __lane_id();
    bg = g0;
    bc = c0;
    bp = p;
    if (better(g1, c1, g0, c0, lid)) {
        bg = g1;
        bc = c1;
        bp = p + 64;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        float og = __shfl_down(bg, o, 64);
        uint32_t oc = __shfl_down(bc, o, 64);
        uint32_t op = __shfl_down(bp, o, 64);
        if (better(og, oc, bg, bc, lid)) {
            bg = og;
            bc = oc;
            bp = op;
        }
    }
}

// One FP32 gain per distinct candidate community (complete sums), written
// over its exact sum: val[slot] <- bits of the gain, -inf for a candidate
// isLowerIdx rejects.  Distinct slots dist[i], i = first, first + stride, ...
// < nd.
template <typename D>
__device__ __forceinline__ void distinct_gains(const DevView &g, const D *dist, uint32_t nd,
                                               uint32_t first, uint32_t stride,
                                               const uint32_t *key, unsigned long long *val,
                                               float m, float kv, float acv, uint32_t szv,
                                               uint32_t cv) {
    // B distinct communities per thread per round; gathers are unconditional
    // (empty positions gather the vertex's own community).  B = 1 measured
    // fastest (2 and 4 were 0.3 % and 1.1 % slower end to end on R-MAT-24).
    constexpr int B = 1;
    for (uint32_t i0 = first; i0 < nd; i0 += stride * B) {
        uint32_t sl[B], c[B], sz[B];
        unsigned long long a[B];
#pragma unroll
        for (int b = 0; b < B; ++b) {
            const uint32_t i = i0 + stride * b;
            sl[b] = i < nd ? dist[i] : kEmpty;
            c[b] = sl[b] != kEmpty ? key[sl[b]] : cv;
        }
#pragma unroll
        for (int b = 0; b < B; ++b) {
            a[b] = g.A[c[b]];
            // only isLowerIdx of a singleton needs the size: otherwise every
            // lane reads the vertex's own entry (one line per vertex)
            sz[b] = g.comm_sz[szv == 1u ? c[b] : cv];
        }
#pragma unroll
        for (int b = 0; b < B; ++b) {
            if (sl[b] == kEmpty) continue;
            // isLowerIdx (modularity.cu:85-93)
            const bool cand = !(szv == 1u && sz[b] == 1u) || lid_of(g.lid, c[b]) < lid_of(g.lid, cv);
            float gg = -INFINITY;
            if (cand) gg = gain_of(fix2f(val[sl[b]], g.S), m, kv, acv, fix2f(a[b], g.S));
            val[sl[b]] = (unsigned long long)__float_as_uint(gg);
        }
    }
}

// last entry visited by virtual lane wl of 128 on a row of d entries
__device__ __forceinline__ uint32_t last_entry(uint32_t wl, uint32_t d) {
    return wl + 128u * ((d - 1u - wl) / 128u);
}

// final decision for buckets 6-7
__device__ __forceinline__ uint32_t decide_lastedge(const DevView &g, float m, uint32_t base,
                                                    uint32_t d, uint32_t cv, float bg,
                                                    uint32_t bp, unsigned long long cew) {
    if (bg == -INFINITY || bp >= d) return cv;
    uint32_t dest = g.comm[g.col[base + last_entry(bp, d)]];
    float cewf = fix2f(cew, g.S);
    if (bg - (cewf / m) > 0.0f) return dest;
    return cv;
}

// the pass loop has stopped on the device: kernels of the remaining
// queued sweeps return at entry (uniform over the grid)
__device__ __forceinline__ bool halted(const DevView &g) { return g.halt && *g.halt; }

// ---- single-graph / batch mode --------------------------------------------
// a vertex takes part in the current sweep (batch: its graph is active)
__device__ __forceinline__ bool v_live(const DevView &g, uint32_t v) {
    return !g.gid || g.gstate[g.gid[v]] == kGraphActive;
}
// the modularity normaliser m of the vertex's graph
__device__ __forceinline__ float v_m(const DevView &g, uint32_t v) {
    return g.gid ? g.gm[g.gid[v]] : g.m;
}
// accumulator slot for an exact total: the vertex's graph in batch mode, a
// spread slot (block % kSpread) in single-graph mode
__device__ __forceinline__ uint32_t acc_slot(const DevView &g, uint32_t v) {
    return g.gid ? g.gid[v] : (blockIdx.x % kSpread);
}
__device__ __forceinline__ uint32_t acc_stride(const DevView &g) {
    return g.gid ? g.G : kSpread;
}
// Exact total contribution of a workgroup.  Single-graph mode: the block is
// reduced in LDS and adds once to spread slot block % kSpread.  Batch mode:
// wave_add to the per-graph slot.  Every thread of the block must call it.
__device__ void wave_add(unsigned long long *base, uint32_t key, unsigned long long x,
                         bool valid);
__device__ __forceinline__ void block_add(const DevView &g, unsigned long long *base,
                                          uint32_t key, unsigned long long x, bool valid) {
    if (g.gid) {
        wave_add(base, key, x, valid);
        return;
    }
    __shared__ unsigned long long part[16];
    const unsigned long long s = wave_sum_u64(valid ? x : 0ull);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long tot = 0;
        const uint32_t nw = (blockDim.x + 63) / 64;
        for (uint32_t i = 0; i < nw; ++i) tot += part[i];
        if (tot) atomicAdd(base + blockIdx.x % kSpread, tot);
    }
    __syncthreads();
}

// base[key] += x over the valid lanes of a wave: one atomic when the valid
// lanes share the key, else one per lane.  Call from every lane of the wave.
__device__ __forceinline__ void wave_add(unsigned long long *base, uint32_t key,
                                         unsigned long long x, bool valid) {
    const unsigned long long vm = __ballot(valid);
    if (!vm) return;
    const int first = __ffsll((long long)vm) - 1;
    const uint32_t k0 = __shfl(key, first, 64);
    if (__ballot(valid && key != k0) == 0ull) {
        const unsigned long long s = wave_sum_u64(valid ? x : 0ull);
        if ((int)__lane_id() == first && s) atomicAdd(base + k0, s);
    } else if (valid && x) {
        atomicAdd(base + key, x);
    }
}

// pass init (modularity.cu:459-463): v alone in its community, in both
// copies of the community state when the view carries the second
__device__ __forceinline__ void init_state_of(const DevView &g, uint32_t v,
                                              unsigned long long fixed) {
    g.comm[v] = v;
    g.comm_sz[v] = 1u;
    g.A[v] = fixed;
    if (g.comm2) {
        g.comm2[v] = v;
        g.comm_sz2[v] = 1u;
        g.A2[v] = fixed;
    }
}

// vertex i of a launch's list (computed when the list is an identity run)
__device__ __forceinline__ uint32_t vert_at(const DevView &g, const uint32_t *verts, uint32_t i) {
    return g.vbase >= 0 ? (uint32_t)g.vbase + i : verts[i];
}

// ---- applying a step's moves (assignCommunity, modularity.cu:382-391, and
// computeNewCommWeights, :393-398, as exact incremental updates) -------------

// moves v (weighted degree kv) from community o to n in the copy (comm2,
// comm_sz2, A2); integer atomics, so any order gives the same state
__device__ __forceinline__ void apply_move2(const DevView &g, uint32_t v, float kv, uint32_t o,
                                            uint32_t n) {
    const unsigned long long K = f2fix(kv, g.scale);
    atomicSub(&g.comm_sz2[o], 1u);
    atomicAdd(&g.comm_sz2[n], 1u);
    atomicAdd(&g.A2[o], (unsigned long long)(-(long long)K));
    atomicAdd(&g.A2[n], K);
    g.comm2[v] = n;
}

// A step's decision for v (community cv in the snapshot).  With two copies
// (comm2 non-null) a move goes at once to the copy the step writes, which the
// step's other vertices never read; otherwise the decision goes to new_comm
// and k_apply applies the step's moves after it.
__device__ __forceinline__ void commit_move(const DevView &g, uint32_t v, float kv, uint32_t cv,
                                            uint32_t nc) {
    if (!g.comm2)
        g.new_comm[v] = nc;
    else if (nc != cv)
        apply_move2(g, v, kv, cv, nc);
}

// The previous step's moves, applied to the copy this step writes: that
// copy holds the state before the previous step, the snapshot the state
// after it, and only the previous step's vertices can differ between them.
// Workgroup blk of the cu_nblk catch-up workgroups; every thread of the
// workgroup calls it.  Counts the moves into cu_moved (the sweep's total).
constexpr uint32_t kCatchUpPer = 4;  // vertices per thread and round
__device__ inline void catch_up(const DevView &g, uint32_t blk) {
    const uint32_t bd = blockDim.x, t = threadIdx.x, n = g.cu_count;
    for (uint32_t b0 = blk * bd * kCatchUpPer; b0 < n; b0 += g.cu_nblk * bd * kCatchUpPer) {
        // every load of the round is issued before the first use, none under
        // a branch (positions past the end load the last vertex)
        uint32_t vv[kCatchUpPer], o[kCatchUpPer], c[kCatchUpPer], gi[kCatchUpPer];
        float kv[kCatchUpPer];
#pragma unroll
        for (uint32_t u = 0; u < kCatchUpPer; ++u) {
            const uint32_t i = b0 + u * bd + t;
            vv[u] = g.cu_verts[i < n ? i : n - 1u];
        }
#pragma unroll
        for (uint32_t u = 0; u < kCatchUpPer; ++u) {
            o[u] = g.comm2[vv[u]];
            c[u] = g.comm[vv[u]];
            kv[u] = g.k[vv[u]];
            gi[u] = g.gid ? g.gid[vv[u]] : 0u;
        }
        unsigned long long moves = 0;
#pragma unroll
        for (uint32_t u = 0; u < kCatchUpPer; ++u) {
            const bool mv = b0 + u * bd + t < n && o[u] != c[u];
            if (mv) apply_move2(g, vv[u], kv[u], o[u], c[u]);
            if (g.gid)
                wave_add(g.cu_moved, gi[u], 1ull, mv);
            else
                moves += mv ? 1ull : 0ull;
        }
        if (!g.gid) block_add(g, g.cu_moved, 0u, moves, true);
    }
}

}  // namespace lv
