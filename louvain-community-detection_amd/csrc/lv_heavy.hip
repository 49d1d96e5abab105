// lv_heavy.hip -- heavy-tail vertices (bucket 7, deg > kBlockMax) spread over
// many workgroups.
//
// A vertex with hundreds of thousands of entries (R-MAT hubs) would make one
// workgroup the critical path of its step.  Its row is cut into chunks of
// kHeavyChunk entries; every kernel below runs one workgroup per chunk, and
// the vertex's state lives in global memory:
//   H1 insert   exact u64 hash sums into the vertex's global table, distinct
//               slot list, per-entry slot, comm_e_w
//   H2 gain     one FP32 gain per distinct community (complete sums)
//   H3 lanes    per virtual lane (128 lanes, entries l, l+128, ...) the first
//               maximum (max gain, then smallest entry index) as one packed
//               u64 combined with atomicMax
//   H4 decide   one wave per vertex: 128-lane tree of computeMoveBlock/Global
//               (modularity.cu:165-187), last-edge destination
//   H5 clean    reset the touched table slots
// The result is the same as one workgroup walking the whole row,
// because every combination step is exact or order-free.
//
// The pinned FP32 sums (k_v, s_v) of these rows are sequential per lane by
// definition (modularity.cu:34-43): one workgroup per vertex streams the row
// through LDS in a lane-transposed layout so 32 summing lanes read LDS while
// three loader waves fetch the next chunk.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "lv_device.h"
#include "lv_devmath.h"

namespace lv {

__device__ __forceinline__ uint32_t f32_ord(float x) {
    uint32_t u = __float_as_uint(x);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float f32_unord(uint32_t o) {
    uint32_t u = (o & 0x80000000u) ? (o & 0x7fffffffu) : ~o;
    return __uint_as_float(u);
}

struct HeavyCtx {
    uint32_t i, v, base, d, c;
    unsigned long long off;
    uint32_t cap;
    bool live;
};

// context of heavy chunk b (the workgroup's own chunk unless given)
__device__ __forceinline__ HeavyCtx heavy_ctx(const DevView &g, const HeavySplit &h,
                                              uint32_t b = 0xffffffffu) {
    if (b == 0xffffffffu) b = blockIdx.x;
    HeavyCtx x;
    x.i = h.cvtx[b];
    x.c = b - h.coff[x.i];
    x.v = h.verts[x.i];
    x.base = g.row[x.v];
    x.d = g.row[x.v + 1] - x.base;
    x.off = h.slot_off[x.i];
    x.cap = (uint32_t)(h.slot_off[x.i + 1] - x.off);
    x.live = v_live(g, x.v);
    return x;
}

// H1: the chunk's entries are first merged in an LDS table (exact u64 sums),
// then each distinct community of the chunk is added to the vertex's global
// table once: global atomics per chunk = distinct communities, not entries.
constexpr int kLocalCap = 2 * kHeavyChunk;
constexpr int kPerThread = kHeavyChunk / 256;

__global__ __launch_bounds__(256) void k_heavy_insert(DevView g, HeavySplit h) {
    constexpr int LC = Log2<kLocalCap>::value;
    __shared__ uint32_t lkey[kLocalCap];
    __shared__ unsigned long long lval[kLocalCap];
    __shared__ uint32_t lglob[kLocalCap];
    __shared__ uint32_t ldist[kHeavyChunk];
    __shared__ uint32_t nld, nnew, newbase;
    __shared__ unsigned long long part[4];
    const uint32_t nfl = h.flist[0];
    for (uint32_t fb = blockIdx.x; fb < nfl; fb += gridDim.x) {
    __syncthreads();  // the previous chunk's LDS reads are done
    const HeavyCtx x = heavy_ctx(g, h, h.flist[1 + fb]);
    if (!x.live) continue;  // uniform over the workgroup
    unsigned long long *tab = h.tab + 2ull * x.off;
    for (int i = threadIdx.x; i < kLocalCap; i += 256) {
        lkey[i] = kEmpty;
        lval[i] = 0ull;
    }
    if (threadIdx.x == 0) {
        nld = 0;
        nnew = 0;
    }
    const uint32_t cv = g.comm[x.v];
    const uint32_t j0 = x.c * kHeavyChunk;
    const uint32_t j1 = min(x.d, j0 + kHeavyChunk);
    __syncthreads();
    uint32_t lslot[kPerThread];
    unsigned long long cew = 0;
#pragma unroll
    for (int q = 0; q < kPerThread; ++q) {
        const uint32_t j = j0 + threadIdx.x + 256u * q;
        lslot[q] = kEmpty;
        if (j < j1) {
            const uint32_t e = g.col[x.base + j];
            const uint32_t ce = g.comm[e];
            const unsigned long long wf = f2fix(g.w[x.base + j], g.scale);
            uint32_t slot = hslot<LC>(ce);
            for (int probe = 0; probe < kLocalCap; ++probe) {
                const uint32_t prev = atomicCAS(&lkey[slot], kEmpty, ce);
                if (prev == kEmpty) {
                    ldist[atomicAdd(&nld, 1u)] = slot;
                    break;
                }
                if (prev == ce) break;
                slot = (slot + 1) & (kLocalCap - 1);
            }
            atomicAdd(&lval[slot], wf);
            if (ce == cv && e != x.v) cew += wf;
            lslot[q] = slot;
        }
    }
    cew = wave_sum_u64(cew);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = cew;
    __syncthreads();
    // flush the chunk's distinct communities into the vertex's global table
    const uint32_t nd = nld;
    uint32_t my_new[kPerThread * 2];
    uint32_t my_rank[kPerThread * 2];
    int nmine = 0;
    for (uint32_t i = threadIdx.x; i < nd; i += 256) {
        const uint32_t ls = ldist[i];
        const uint32_t ce = lkey[ls];
        uint32_t slot = hslot_cap(ce, x.cap);
        uint32_t probe = 0;
        bool claimed = false;
        for (; probe < x.cap; ++probe) {
            const uint32_t prev = atomicCAS(slot_key(tab, slot), kEmpty, ce);
            if (prev == kEmpty) {
                claimed = true;
                break;
            }
            if (prev == ce) break;
            slot = next_slot(slot, x.cap);
        }
        if (probe == x.cap) {  // cannot happen: cap = 2*deg
            atomicOr(h.err, 1u);
            slot = 0;
        }
        atomicAdd(slot_sum(tab, slot), lval[ls]);
        lglob[ls] = slot;
        if (claimed) {
            my_new[nmine] = slot;
            my_rank[nmine] = atomicAdd(&nnew, 1u);
            ++nmine;
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const unsigned long long s = part[0] + part[1] + part[2] + part[3];
        if (s) atomicAdd(&h.hcew[x.i], s);
        newbase = nnew ? (uint32_t)atomicAdd(&h.hcnt[x.i], (unsigned long long)nnew) : 0u;
    }
    __syncthreads();
    for (int q = 0; q < nmine; ++q) h.dlist[x.base + newbase + my_rank[q]] = my_new[q];
#pragma unroll
    for (int q = 0; q < kPerThread; ++q) {
        const uint32_t j = j0 + threadIdx.x + 256u * q;
        if (lslot[q] != kEmpty) h.eslot[x.base + j] = lglob[lslot[q]];
    }
    }
}

__global__ __launch_bounds__(256) void k_heavy_gain(DevView g, HeavySplit h) {
    const uint32_t nfl = h.flist[0];
    for (uint32_t fb = blockIdx.x; fb < nfl; fb += gridDim.x) {
    const HeavyCtx x = heavy_ctx(g, h, h.flist[1 + fb]);
    if (!x.live) continue;  // uniform over the workgroup
    unsigned long long *tab = h.tab + 2ull * x.off;
    const uint32_t nd = (uint32_t)h.hcnt[x.i];
    const uint32_t i0 = x.c * kHeavyChunk;
    if (i0 >= nd) continue;
    const uint32_t i1 = min(nd, i0 + kHeavyChunk);
    const uint32_t cv = g.comm[x.v];
    const uint32_t szv = g.comm_sz[cv];
    const float kv = g.k[x.v];
    const float acv = fix2f(g.A[cv], g.S);
    for (uint32_t i = i0 + threadIdx.x; i < i1; i += 256) {
        const uint32_t slot = h.dlist[x.base + i];
        const uint32_t c = *slot_key(tab, slot);
        float gg = -INFINITY;
        // isLowerIdx (modularity.cu:85-93)
        const bool cand = lower_idx_ok(g, szv, c, cv);
        if (cand)
            gg = gain_of(fix2f(*slot_sum(tab, slot), g.S), v_m(g, x.v), kv, acv, fix2f(g.A[c], g.S));
        *slot_sum(tab, slot) = (unsigned long long)__float_as_uint(gg);
    }
    }
}

__global__ __launch_bounds__(256) void k_heavy_lanes(DevView g, HeavySplit h) {
    __shared__ float sg[256];
    __shared__ uint32_t sj[256];
    const uint32_t nfl = h.flist[0];
    for (uint32_t fb = blockIdx.x; fb < nfl; fb += gridDim.x) {
    __syncthreads();  // the previous chunk's LDS reads are done
    const HeavyCtx x = heavy_ctx(g, h, h.flist[1 + fb]);
    if (!x.live) continue;  // uniform over the workgroup
    unsigned long long *tab = h.tab + 2ull * x.off;
    const uint32_t j0 = x.c * kHeavyChunk;  // multiple of 128: thread t serves lane t % 128
    const uint32_t j1 = min(x.d, j0 + kHeavyChunk);
    float tg = -INFINITY;
    uint32_t tj = 0xffffffffu;
    for (uint32_t j = j0 + threadIdx.x; j < j1; j += 256) {
        const float gg = __uint_as_float((uint32_t)*slot_sum(tab, h.eslot[x.base + j]));
        if (gg > tg) {  // entries visited in increasing order: first maximum
            tg = gg;
            tj = j;
        }
    }
    sg[threadIdx.x] = tg;
    sj[threadIdx.x] = tj;
    __syncthreads();
    if (threadIdx.x < 128) {
        const float g2 = sg[threadIdx.x + 128];
        const uint32_t j2 = sj[threadIdx.x + 128];
        if (g2 > tg || (g2 == tg && j2 < tj)) {
            tg = g2;
            tj = j2;
        }
        if (tg != -INFINITY) {
            const unsigned long long packed =
                ((unsigned long long)f32_ord(tg) << 32) | (unsigned long long)(0xffffffffu - tj);
            atomicMax(&h.hlane[(size_t)x.i * 128 + threadIdx.x], packed);
        }
    }
    }
}

// one wave per heavy vertex
__global__ __launch_bounds__(256) void k_heavy_decide(DevView g, HeavySplit h) {
    if (halted(g)) return;  // the pass loop stopped earlier in this launch
    const uint32_t i = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (i >= h.nh) return;
    const uint32_t p = threadIdx.x & 63;
    const uint32_t v = h.verts[i];
    if (!v_live(g, v)) return;  // uniform over the wave
    const uint32_t base = g.row[v];
    const uint32_t d = g.row[v + 1] - base;
    float gl[2];
    uint32_t cl[2];
#pragma unroll
    for (int q = 0; q < 2; ++q) {
        const unsigned long long k = h.hlane[(size_t)i * 128 + p + 64 * q];
        if (k == 0ull) {
            gl[q] = -INFINITY;
            cl[q] = kEmpty;
        } else {
            gl[q] = f32_unord((uint32_t)(k >> 32));
            const uint32_t j = 0xffffffffu - (uint32_t)(k & 0xffffffffull);
            cl[q] = g.comm[g.col[base + j]];
        }
    }
    float bg;
    uint32_t bc, bp;
    tree128(gl[0], cl[0], gl[1], cl[1], bg, bc, bp, g.lid);
    if (p == 0) {
        const uint32_t cv = g.comm[v];
        commit_move(g, v, g.k[v], cv,
                    decide_lastedge(g, v_m(g, v), base, d, cv, bg, bp, h.hcew[i]));
    }
}

__global__ __launch_bounds__(256) void k_heavy_clean(DevView g, HeavySplit h) {
    const uint32_t nfl = h.flist[0];
    for (uint32_t fb = blockIdx.x; fb < nfl; fb += gridDim.x) {
    const HeavyCtx x = heavy_ctx(g, h, h.flist[1 + fb]);
    if (!x.live) continue;  // uniform over the workgroup
    unsigned long long *tab = h.tab + 2ull * x.off;
    const uint32_t nd = (uint32_t)h.hcnt[x.i];
    const uint32_t i0 = x.c * kHeavyChunk;
    if (i0 >= nd) continue;
    const uint32_t i1 = min(nd, i0 + kHeavyChunk);
    for (uint32_t i = i0 + threadIdx.x; i < i1; i += 256) {
        const uint32_t slot = h.dlist[x.base + i];
        tab[2ull * slot] = (unsigned long long)kEmpty;
        tab[2ull * slot + 1ull] = 0ull;
    }
    }
}

__global__ void k_heavy_chunks(const uint32_t *row, const uint32_t *verts, uint32_t count,
                               uint32_t *nchunks) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) {
        const uint32_t v = verts[i];
        nchunks[i] = (row[v + 1] - row[v] + kHeavyChunk - 1) / kHeavyChunk;
    }
}

// ---------------------------------------------------------------------------
// Partitioned heavy path (the default): a heavy row is split into hash
// buckets of its neighbour communities, sized so a bucket's distinct
// communities fit an LDS table; every bucket is then accumulated, scored and
// reduced to per-virtual-lane maxima entirely in LDS.
//   P0 count    community of every entry (kept in eslot), bucket histogram,
//               comm_e_w
//   (scan of the bucket counts -> bucket positions in heavy-entry space)
//   P1 scatter  (community, weight, entry index) into the bucket's range
//   P2 bucket   LDS table of exact sums, gains per distinct community, per
//               virtual lane first maximum via an LDS atomicMax on
//               (gain, -entry), merged into the vertex's lanes with one
//               global atomicMax per lane
// A bucket whose distinct communities overflow the LDS table flags its
// vertex; flagged vertices are redone by the global-table kernels above.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t bucket_of(uint32_t c, int lb) {
    return lb ? ((c * 0x9E3779B1u) >> (32 - lb)) : 0u;
}

__global__ __launch_bounds__(256) void k_hp_count(DevView g, HeavySplit h) {
    if (halted(g)) return;  // the pass loop stopped earlier in this launch
    __shared__ uint32_t hist[kMaxBuckets];
    __shared__ unsigned long long part[4];
    const HeavyCtx x = heavy_ctx(g, h);
    if (!x.live) return;  // uniform over the workgroup
    const uint32_t nb = h.nbuck[x.i];
    const int lb = 31 - __clz(nb);
    for (uint32_t b = threadIdx.x; b < nb; b += 256) hist[b] = 0;
    __syncthreads();
    const uint32_t cv = g.comm[x.v];
    const uint32_t j0 = x.c * kHeavyChunk;
    const uint32_t j1 = min(x.d, j0 + kHeavyChunk);
    unsigned long long cew = 0;
    // columns first (past the chunk end: its first entry), then every
    // community, then the LDS work
    uint32_t es[kPerThread], cs[kPerThread];
    float ws[kPerThread];
#pragma unroll
    for (int q = 0; q < kPerThread; ++q) {
        const uint32_t j = j0 + threadIdx.x + 256u * q;
        es[q] = g.col[x.base + (j < j1 ? j : j0)];
    }
#pragma unroll
    for (int q = 0; q < kPerThread; ++q) cs[q] = g.comm[es[q]];
    // weights only where comm_e_w needs them (entries in v's community that
    // are not self-loops); the other lanes issue no load
#pragma unroll
    for (int q = 0; q < kPerThread; ++q) {
        const uint32_t j = j0 + threadIdx.x + 256u * q;
        ws[q] = 0.0f;
        if (j < j1 && cs[q] == cv && es[q] != x.v) ws[q] = g.w[x.base + j];
    }
#pragma unroll
    for (int q = 0; q < kPerThread; ++q) {
        const uint32_t j = j0 + threadIdx.x + 256u * q;
        if (j >= j1) continue;
        h.eslot[x.base + j] = cs[q];
        atomicAdd(&hist[bucket_of(cs[q], lb)], 1u);
        if (cs[q] == cv && es[q] != x.v) cew += f2fix(ws[q], g.scale);
    }
    cew = wave_sum_u64(cew);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = cew;
    __syncthreads();
    const uint32_t b0 = h.bstart[x.i];
    for (uint32_t b = threadIdx.x; b < nb; b += 256)
        if (hist[b]) atomicAdd(&h.bcnt[b0 + b], hist[b]);
    if (threadIdx.x == 0) {
        const unsigned long long s = part[0] + part[1] + part[2] + part[3];
        if (s) atomicAdd(&h.hcew[x.i], s);
    }
}

// The chunk's entries are first ordered by bucket in LDS, so each bucket's
// run of entries is written to its range with consecutive addresses.
__global__ __launch_bounds__(256) void k_hp_scatter(DevView g, HeavySplit h) {
    if (halted(g)) return;  // the pass loop stopped earlier in this launch
    static_assert(kMaxBuckets == 4 * 256, "four buckets per thread in the offset scan");
    static_assert(kHeavyChunk < 65536, "16-bit chunk counts and offsets");
    // per bucket: entries of the chunk (16-bit halves of 32-bit words, so a
    // 32-bit LDS atomic counts two buckets), their offset in the
    // bucket-ordered chunk, their position in the bucket's range
    __shared__ uint32_t lcnt2[kMaxBuckets / 2];
    __shared__ uint16_t loff[kMaxBuckets];
    __shared__ uint32_t lbase[kMaxBuckets];
    __shared__ uint32_t sc[kHeavyChunk], sj[kHeavyChunk];
    __shared__ float sw[kHeavyChunk];
    uint32_t *wsum = sc;  // scan partials; read before the staging below starts
    const HeavyCtx x = heavy_ctx(g, h);
    if (!x.live) return;
    const uint32_t nb = h.nbuck[x.i];
    const int lb = 31 - __clz(nb);
    for (uint32_t b = threadIdx.x; b < kMaxBuckets / 2; b += 256) lcnt2[b] = 0;
    __syncthreads();
    const uint32_t j0 = x.c * kHeavyChunk;
    const uint32_t j1 = min(x.d, j0 + kHeavyChunk);
    const uint32_t cnt = j1 - j0;
    uint32_t rank[kPerThread], bk[kPerThread], cc[kPerThread];
    float wq[kPerThread];
    // all loads first (past the chunk end: its first entry), then the ranks
#pragma unroll
    for (int q = 0; q < kPerThread; ++q) {
        const uint32_t j = j0 + threadIdx.x + 256u * q;
        const uint32_t e = x.base + (j < j1 ? j : j0);
        cc[q] = h.eslot[e];
        wq[q] = g.w[e];
    }
#pragma unroll
    for (int q = 0; q < kPerThread; ++q) {
        const uint32_t j = j0 + threadIdx.x + 256u * q;
        bk[q] = kEmpty;
        if (j < j1) {
            bk[q] = bucket_of(cc[q], lb);
            const uint32_t sh = (bk[q] & 1u) << 4;
            rank[q] = (atomicAdd(&lcnt2[bk[q] >> 1], 1u << sh) >> sh) & 0xffffu;
        }
    }
    __syncthreads();
    // exclusive scan of the counts: thread t owns buckets 4t..4t+3
    const uint32_t t = threadIdx.x, lane = t & 63, wv = t >> 6;
    const uint32_t p01 = lcnt2[2 * t], p23 = lcnt2[2 * t + 1];
    const uint32_t c0 = p01 & 0xffffu, c1 = p01 >> 16, c2 = p23 & 0xffffu, c3 = p23 >> 16;
    const uint32_t own = c0 + c1 + c2 + c3;
    uint32_t inc = own;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(inc, o, 64);
        if (lane >= (uint32_t)o) inc += y;
    }
    if (lane == 63) wsum[wv] = inc;
    const uint32_t b0 = h.bstart[x.i];
    for (uint32_t b = t; b < nb; b += 256) {
        const uint32_t cb = (lcnt2[b >> 1] >> ((b & 1u) << 4)) & 0xffffu;
        if (cb) lbase[b] = atomicAdd(&h.bcur[b0 + b], cb);
    }
    __syncthreads();
    uint32_t base = inc - own;
    for (uint32_t k = 0; k < wv; ++k) base += wsum[k];
    loff[4 * t] = (uint16_t)base;
    loff[4 * t + 1] = (uint16_t)(base + c0);
    loff[4 * t + 2] = (uint16_t)(base + c0 + c1);
    loff[4 * t + 3] = (uint16_t)(base + c0 + c1 + c2);
    __syncthreads();
#pragma unroll
    for (int q = 0; q < kPerThread; ++q) {
        if (bk[q] == kEmpty) continue;
        const uint32_t j = j0 + threadIdx.x + 256u * q;
        const uint32_t sp = loff[bk[q]] + rank[q];
        sc[sp] = cc[q];
        sj[sp] = j;
        sw[sp] = wq[q];
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < cnt; i += 256) {
        const uint32_t c = sc[i];
        const uint32_t b = bucket_of(c, lb);
        const uint32_t pos = lbase[b] + (i - loff[b]);
        h.pc[pos] = c;
        h.pw[pos] = sw[i];
        h.pj[pos] = sj[i];
    }
}

// CAP: LDS table slots of one bucket (any size), DMAX: distinct keys it may
// hold before the bucket overflows to the global tables.  A bucket of at
// most 4 x 256 entries keeps each thread's entries and their table slots in
// registers between the insert and the lane phase; longer buckets (skewed
// hashes) read their entries again.
template <int kBucketCap>
__device__ __forceinline__ uint32_t pslot_n(uint32_t c) {
    return (uint32_t)(((unsigned long long)(c * 0x85EBCA6Bu) * (unsigned)kBucketCap) >> 32);
}
template <int kBucketCap>
__device__ __forceinline__ uint32_t hp_insert(uint32_t *key, unsigned long long *val,
                                              uint16_t *dist, uint32_t *nd_s, uint32_t *over,
                                              uint32_t c, unsigned long long wf, uint32_t dmax) {
    uint32_t slot = pslot_n<kBucketCap>(c);
    for (int probe = 0; probe < kBucketCap; ++probe) {
        const uint32_t prev = atomicCAS(&key[slot], kEmpty, c);
        if (prev == kEmpty) {
            const uint32_t idx = atomicAdd(nd_s, 1u);
            if (idx < dmax) dist[idx] = (uint16_t)slot;
            else *over = 1u;  // more distinct communities than the table allows
            atomicAdd(&val[slot], wf);
            return slot;
        }
        if (prev == c) {
            atomicAdd(&val[slot], wf);
            return slot;
        }
        slot = slot + 1 == (uint32_t)kBucketCap ? 0u : slot + 1;
    }
    *over = 1u;
    return kEmpty;
}

// One workgroup per kHpGroup consecutive buckets of one vertex (every
// heavy vertex has a multiple of kHpGroup buckets): the buckets are merged
// one after the other through the same LDS table, their per-lane maxima
// gather in LDS, and the vertex's lanes get one global atomicMax per lane
// per group instead of per bucket.
template <int kBucketCap, int kDmax>
__global__ __launch_bounds__(256) void k_hp_bucket(DevView g, HeavySplit h) {
    if (halted(g)) return;  // the pass loop stopped earlier in this launch
    static_assert(kBucketCap > kDmax && kBucketCap <= 65536, "table shape");
    __shared__ uint32_t key[kBucketCap];
    __shared__ unsigned long long val[kBucketCap];
    __shared__ uint16_t dist[kDmax];
    __shared__ unsigned long long lanes[128];
    __shared__ uint32_t nd_s, over;
    const uint32_t bg0 = blockIdx.x * kHpGroup;
    const uint32_t i = h.bvtx[bg0];
    const uint32_t v = h.verts[i];
    if (!v_live(g, v)) return;
    // the vertex's own state, loaded now so its latency overlaps the inserts
    const float m = v_m(g, v);
    const uint32_t cv = g.comm[v];
    const uint32_t szv = g.comm_sz[cv];
    const float kv = g.k[v];
    const unsigned long long acvx = g.A[cv];
    if (threadIdx.x < 128) lanes[threadIdx.x] = 0ull;
    const uint32_t dmax = min(h.dmax, (uint32_t)kDmax);
    for (uint32_t kb = 0; kb < kHpGroup; ++kb) {
        const uint32_t p0 = h.bpos[bg0 + kb], p1 = h.bpos[bg0 + kb + 1];
        if (p0 == p1) continue;  // uniform
        for (int s2 = threadIdx.x; s2 < kBucketCap; s2 += 256) {
            key[s2] = kEmpty;
            val[s2] = 0ull;
        }
        if (threadIdx.x == 0) {
            nd_s = 0;
            over = 0;
        }
        __syncthreads();
        constexpr uint32_t R = 4;
        const bool cached = p1 - p0 <= 256u * R;  // uniform
        uint32_t cs[R], js[R], ss[R];
        if (cached) {
            float ws[R];
#pragma unroll
            for (uint32_t u = 0; u < R; ++u) {  // all loads first, unconditional
                const uint32_t qu = p0 + threadIdx.x + 256u * u;
                const uint32_t q = qu < p1 ? qu : p0;
                cs[u] = h.pc[q];
                ws[u] = h.pw[q];
                js[u] = h.pj[q];
            }
#pragma unroll
            for (uint32_t u = 0; u < R; ++u)
                if (p0 + threadIdx.x + 256u * u >= p1) cs[u] = kEmpty;
#pragma unroll
            for (uint32_t u = 0; u < R; ++u)
                ss[u] = cs[u] == kEmpty
                            ? kEmpty
                            : hp_insert<kBucketCap>(key, val, dist, &nd_s, &over, cs[u],
                                                    f2fix(ws[u], g.scale), dmax);
        } else {
            for (uint32_t q = p0 + threadIdx.x; q < p1; q += 256)
                hp_insert<kBucketCap>(key, val, dist, &nd_s, &over, h.pc[q],
                                      f2fix(h.pw[q], g.scale), dmax);
        }
        __syncthreads();
        if (over) {  // uniform: redo this vertex with the global-table kernels
            if (threadIdx.x == 0) atomicOr(&h.hflag[i], 1u);
            return;
        }
        const uint32_t nd = nd_s;
        const float acv = fix2f(acvx, g.S);
        distinct_gains(g, dist, nd, threadIdx.x, 256u, key, val, m, kv, acv, szv, cv);
        __syncthreads();
        auto lane_max = [&](uint32_t slot, uint32_t j) {
            const float gg = __uint_as_float((uint32_t)val[slot]);
            if (gg != -INFINITY) {
                const unsigned long long packed =
                    ((unsigned long long)f32_ord(gg) << 32) | (unsigned long long)(0xffffffffu - j);
                atomicMax(&lanes[j & 127u], packed);
            }
        };
        if (cached) {
#pragma unroll
            for (uint32_t u = 0; u < R; ++u)
                if (ss[u] != kEmpty) lane_max(ss[u], js[u]);
        } else {
            for (uint32_t q = p0 + threadIdx.x; q < p1; q += 256) {
                const uint32_t c = h.pc[q];
                uint32_t slot = pslot_n<kBucketCap>(c);
                while (key[slot] != c)
                    slot = slot + 1 == (uint32_t)kBucketCap ? 0u : slot + 1;  // present
                lane_max(slot, h.pj[q]);
            }
        }
        __syncthreads();  // the lane phase read the table the next bucket clears
    }
    if (threadIdx.x < 128 && lanes[threadIdx.x])
        atomicMax(&h.hlane[(size_t)i * 128 + threadIdx.x], lanes[threadIdx.x]);
}

// per heavy vertex: bucket count pow2 >= ceil(deg / target), at least
// kHpGroup, at most kMaxBuckets
__global__ void k_hp_nbuck(const uint32_t *row, const uint32_t *verts, uint32_t count,
                           uint32_t target, uint32_t *nbuck) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) {
        const uint32_t v = verts[i];
        const uint32_t d = row[v + 1] - row[v];
        uint32_t b = kHpGroup;
        while (b < kMaxBuckets && (uint64_t)b * target < d) b <<= 1;
        nbuck[i] = b;
    }
}
void launch_hp_nbuck(hipStream_t s, const uint32_t *row, const uint32_t *verts, uint32_t count,
                     uint32_t target, uint32_t *nbuck) {
    if (!count) return;
    hipLaunchKernelGGL(k_hp_nbuck, dim3((count + 255) / 256), dim3(256), 0, s, row, verts, count,
                       target, nbuck);
}

// test hook: every odd heavy vertex also takes the global-table path, so
// both paths serve one step
__global__ void k_hp_flag_odd(HeavySplit h) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < h.nh && (i & 1u)) h.hflag[i] = 1u;
}

// flagged vertices restart on the global-table kernels: their state is
// cleared and their chunks are appended to the list those kernels walk
// (flist[0] counts the listed chunks, cleared before this launch)
__global__ void k_hp_reset_flagged(HeavySplit h) {
    __shared__ uint32_t pos;
    const uint32_t i = blockIdx.x;
    if (i >= h.nh || !h.hflag[i]) return;  // uniform over the workgroup
    const uint32_t c0 = h.coff[i], nc = h.coff[i + 1] - c0;
    if (threadIdx.x == 0) {
        h.hcnt[i] = 0ull;
        h.hcew[i] = 0ull;
        pos = atomicAdd(h.flist, nc);
    }
    if (threadIdx.x < 128) h.hlane[(size_t)i * 128 + threadIdx.x] = 0ull;
    __syncthreads();
    for (uint32_t t = threadIdx.x; t < nc; t += blockDim.x) h.flist[1 + pos + t] = c0 + t;
}

__global__ void k_fill_slots(unsigned long long *tab, size_t cap) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < cap; i += stride) {
        tab[2 * i] = (unsigned long long)kEmpty;
        tab[2 * i + 1] = 0ull;
    }
}
void launch_fill_slots(hipStream_t s, unsigned long long *tab, size_t cap) {
    if (!cap) return;
    size_t gsz = (cap + 255) / 256;
    if (gsz > 65536) gsz = 65536;
    hipLaunchKernelGGL(k_fill_slots, dim3((unsigned)gsz), dim3(256), 0, s, tab, cap);
}

void launch_heavy_chunks(hipStream_t s, const uint32_t *row, const uint32_t *verts,
                         uint32_t count, uint32_t *nchunks) {
    if (!count) return;
    hipLaunchKernelGGL(k_heavy_chunks, dim3((count + 255) / 256), dim3(256), 0, s, row, verts,
                       count, nchunks);
}

#ifndef LV_HPCAP
#define LV_HPCAP 1280
#endif
#ifndef LV_HPDMAX
#define LV_HPDMAX 640
#endif
constexpr uint32_t kFallbackGrid = 1024;  // workgroups of the global-table kernels

void launch_heavy_move(hipStream_t s, const DevView &g, const HeavySplit &hin) {
    if (!hin.nh || !hin.nchunks) return;
    const HeavySplit &h = hin;
    const dim3 grid(h.nchunks), block(256);
    launch_fill_u64(s, h.hcnt, 0ull, (size_t)h.nh * 130);  // hcnt | hcew | hlane
    if (h.force_global == 1) {  // every heavy vertex on the global tables
        launch_fill_u32(s, h.hflag, 1u, h.nh);
    } else {
        launch_fill_u32(s, h.hflag, 0u, h.nh);
        launch_fill_u32(s, h.bcnt, 0u, h.nbuckets);
        hipLaunchKernelGGL(k_hp_count, grid, block, 0, s, g, h);
        launch_scan_u32(s, h.bcnt, 0, h.nbuckets, h.bpos, h.scan_tmp);
        (void)hipMemcpyAsync(h.bcur, h.bpos, (size_t)h.nbuckets * sizeof(uint32_t),
                             hipMemcpyDeviceToDevice, s);
        hipLaunchKernelGGL(k_hp_scatter, grid, block, 0, s, g, h);
        // buckets of up to 512 target entries: 2048-slot tables; larger: 4096
        const dim3 bgrid(h.nbuckets / kHpGroup);
        if (h.bucket_entries <= 512)
            hipLaunchKernelGGL((k_hp_bucket<LV_HPCAP, LV_HPDMAX>), bgrid, block, 0, s, g, h);
        else
            hipLaunchKernelGGL((k_hp_bucket<4096, 2048>), bgrid, block, 0, s, g, h);
        if (h.force_global == 2)
            hipLaunchKernelGGL(k_hp_flag_odd, dim3((h.nh + 255) / 256), dim3(256), 0, s, h);
    }
    // overflowing vertices (if any) redo the step on the global tables; those
    // kernels walk the list of their chunks with a fixed grid, so a step
    // without overflow dispatches a few idle workgroups, not one per chunk
    launch_fill_u32(s, h.flist, 0u, 1);
    hipLaunchKernelGGL(k_hp_reset_flagged, dim3(h.nh), dim3(128), 0, s, h);
    const dim3 fgrid(h.nchunks < kFallbackGrid ? h.nchunks : kFallbackGrid);
    hipLaunchKernelGGL(k_heavy_insert, fgrid, block, 0, s, g, h);
    hipLaunchKernelGGL(k_heavy_gain, fgrid, block, 0, s, g, h);
    hipLaunchKernelGGL(k_heavy_lanes, fgrid, block, 0, s, g, h);
    hipLaunchKernelGGL(k_heavy_decide, dim3((h.nh + 3) / 4), block, 0, s, g, h);
    hipLaunchKernelGGL(k_heavy_clean, fgrid, block, 0, s, g, h);
}

// locality order: heavy rows are copied one chunk of kHeavyChunk entries per
// workgroup (positions x0 + i, i < nh, chunk offsets coff)
__global__ __launch_bounds__(256) void k_relabel_heavy(RelabelView r, const uint32_t *cvtx,
                                                       const uint32_t *coff, uint32_t x0) {
    const uint32_t i = cvtx[blockIdx.x];
    const uint32_t c = blockIdx.x - coff[i];
    const uint32_t x = x0 + i, v = r.lid[x];
    const uint32_t b = r.row[v], d = r.row[v + 1] - b, o = r.row2[x];
    const uint32_t j0 = c * kHeavyChunk, j1 = min(d, j0 + kHeavyChunk);
    // all loads, then all gathers, then the stores (past the chunk end: its
    // first entry again)
    uint32_t cs[kPerThread];
    float ws[kPerThread];
#pragma unroll
    for (int q = 0; q < kPerThread; ++q) {
        const uint32_t j = j0 + threadIdx.x + 256u * q;
        const uint32_t e = b + (j < j1 ? j : j0);
        cs[q] = r.col[e];
        ws[q] = r.w[e];
    }
#pragma unroll
    for (int q = 0; q < kPerThread; ++q) cs[q] = r.pmap[cs[q]];
#pragma unroll
    for (int q = 0; q < kPerThread; ++q) {
        const uint32_t j = j0 + threadIdx.x + 256u * q;
        if (j < j1) {
            r.col2[o + j] = cs[q];
            r.w2[o + j] = ws[q];
        }
    }
}
void launch_relabel_heavy(hipStream_t s, const RelabelView &r, const uint32_t *cvtx,
                          const uint32_t *coff, uint32_t nchunks, uint32_t x0) {
    if (!nchunks) return;
    hipLaunchKernelGGL(k_relabel_heavy, dim3(nchunks), dim3(256), 0, s, r, cvtx, coff, x0);
}

// one wave per owner: a vertex with a thousand buckets is 16 rounds of its
// wave, not a thousand stores of one thread
__global__ __launch_bounds__(256) void k_owner_fill(const uint32_t *off, uint32_t count,
                                                    uint32_t *out) {
    const uint32_t i = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (i >= count) return;
    const uint32_t b1 = off[i + 1];
    for (uint32_t b = off[i] + (threadIdx.x & 63); b < b1; b += 64) out[b] = i;
}
void launch_owner_fill(hipStream_t s, const uint32_t *off, uint32_t count, uint32_t *out) {
    if (!count) return;
    hipLaunchKernelGGL(k_owner_fill, dim3((count + 3) / 4), dim3(256), 0, s, off, count, out);
}

// ---------------------------------------------------------------------------
// pinned-order sums of heavy rows
// ---------------------------------------------------------------------------

// masked weights of the heavy rows, staged contiguously: w if the neighbour
// is in v's community, else +0 (adding +0 to a partial >= +0 changes nothing)
__global__ __launch_bounds__(256) void k_heavy_stage(DevView g, HeavySplit h, float *stage,
                                                     int ident) {
    if (halted(g)) return;  // the pass loop stopped earlier in this launch
    const HeavyCtx x = heavy_ctx(g, h);
    if (!x.live) return;  // uniform over the workgroup
    const uint32_t cv = g.comm[x.v];
    const uint32_t j0 = x.c * kHeavyChunk;
    const uint32_t j1 = min(x.d, j0 + kHeavyChunk);
    // columns (past the chunk end: its first entry), then communities, then
    // the weights of the entries inside v's community only (the other lanes
    // issue no load)
    uint32_t cs[kPerThread];
    float ws[kPerThread];
#pragma unroll
    for (int q = 0; q < kPerThread; ++q) {
        const uint32_t j = j0 + threadIdx.x + 256u * q;
        cs[q] = g.col[x.base + (j < j1 ? j : j0)];
    }
    if (!ident) {  // ident: every community is a singleton, comm[u] = u
#pragma unroll
        for (int q = 0; q < kPerThread; ++q) cs[q] = g.comm[cs[q]];
    }
#pragma unroll
    for (int q = 0; q < kPerThread; ++q) {
        const uint32_t j = j0 + threadIdx.x + 256u * q;
        ws[q] = 0.0f;
        if (j < j1 && cs[q] == cv) ws[q] = g.w[x.base + j];
    }
#pragma unroll
    for (int q = 0; q < kPerThread; ++q) {
        const uint32_t j = j0 + threadIdx.x + 256u * q;
        if (j < j1) stage[x.base + j] = ws[q];
    }
}

constexpr int kR1K = 192;                 // entries per lane per chunk
constexpr int kR1Chunk = 32 * kR1K;       // 6144 entries
constexpr int kR1Per = kR1Chunk / 192;    // loads per loader thread and chunk

__global__ __launch_bounds__(256) void k_heavy_r1(DevView g, HeavySplit h, const float *stage,
                                                  int masked, int init_state,
                                                  unsigned long long *acc) {
    if (halted(g)) return;  // the pass loop stopped earlier in this launch
    __shared__ float buf[2][32][kR1K + 1];
    const uint32_t v = h.verts[blockIdx.x];
    if (!v_live(g, v)) return;  // uniform over the workgroup
    const uint32_t base = g.row[v];
    const uint32_t d = g.row[v + 1] - base;  // > 0: a heavy row
    const float *src = masked ? stage + base : g.w + base;
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const uint32_t nch = (d + kR1Chunk - 1) / kR1Chunk;
    // loader waves 1-3: all of a thread's loads in flight at once, then the
    // transposed LDS stores (entry 32 i + l of the chunk -> buf[l][i])
    auto load = [&](uint32_t c, int b) {
        const int t = (int)threadIdx.x - 64;
        float x[kR1Per];
#pragma unroll
        for (int i = 0; i < kR1Per; ++i) {
            const uint32_t j = c * kR1Chunk + (uint32_t)(t + 192 * i);
            x[i] = src[j < d ? j : d - 1];
        }
#pragma unroll
        for (int i = 0; i < kR1Per; ++i) {
            const int idx = t + 192 * i;
            const uint32_t j = c * kR1Chunk + (uint32_t)idx;
            buf[b][idx & 31][idx >> 5] = j < d ? x[i] : 0.0f;
        }
    };
    if (wv != 0) load(0, 0);
    __syncthreads();
    float p = 0.0f;
    for (uint32_t c = 0; c < nch; ++c) {
        if (wv != 0) {
            if (c + 1 < nch) load(c + 1, (c + 1) & 1);
        } else if (lane < 32) {
            // rounds holding entries of the row; the zero padding of the last
            // round leaves p unchanged (p is never -0)
            const uint32_t rem = d - c * kR1Chunk;
            const int ni = rem >= (uint32_t)kR1Chunk ? kR1K : (int)((rem + 31) / 32);
            const float *r = &buf[c & 1][lane][0];
            for (int i = 0; i < ni; ++i) p += r[i];
        }
        __syncthreads();
    }
    if (wv == 0) {
#pragma unroll
        for (int o = 16; o > 0; o >>= 1) p += __shfl_down(p, o, 32);
        if (lane == 0) {
            const unsigned long long fixed = f2fix(p, g.scale);
            if (!masked) {
                g.k[v] = p;
                if (init_state) {
                    init_state_of(g, v, fixed);
                }
            }
            if (fixed) atomicAdd(&acc[acc_slot(g, v)], fixed);
        }
    }
}

void launch_heavy_r1(hipStream_t s, const DevView &g, const HeavySplit &h, float *stage,
                     int masked, int init_state, unsigned long long *acc) {
    if (!h.nh) return;
    if (masked)
        hipLaunchKernelGGL(k_heavy_stage, dim3(h.nchunks), dim3(256), 0, s, g, h, stage,
                           masked == 2 ? 1 : 0);
    hipLaunchKernelGGL(k_heavy_r1, dim3(h.nh), dim3(256), 0, s, g, h, stage, masked, init_state,
                       acc);
}

}  // namespace lv
