"""Lane-level simulation of the reference move kernels, for pinning the oracle.

Test helper only.  It follows the CUDA source line by line rather than the
oracle's formulation: every vertex is simulated as the reference's thread
group (computeMoveWarp, modularity.cu:286-380; computeMoveBlock /
computeMoveGlobal, modularity.cu:95-284), lane by lane, with lock-step rounds,
the shuffle-down reduction including reads past the group, the shared-memory
tree of the block kernels, and the `e` / `max_e` registers the winner lane
uses for its write.  The only departures are the canonical resolutions the
oracle also documents: hash sums are exact integers (scaled by 2^S) rounded
once, the block kernels see complete sums, and a group whose winner lane
writes nothing keeps the vertex in place.

Pure Python, for graphs of a few hundred entries.
"""
import numpy as np

F32 = np.float32
CUTS = [0, 4, 8, 16, 32, 84, 319, 0x7FFFFFFF]
BLOCKS = [(4, 32), (8, 16), (16, 8), (32, 4), (32, 4), (128, 1), (128, 1)]
NEG_INF = F32(-np.inf)
GARBAGE = 0xDEADBEEF  # uninitialised max_comm register of a lane without candidates


def int_to_f32(x):
    """non-negative integer -> float32, round half to even, exactly once"""
    assert x >= 0
    nb = x.bit_length()
    if nb <= 24:
        return F32(x)
    sh = nb - 24
    mant, rem = x >> sh, x & ((1 << sh) - 1)
    half = 1 << (sh - 1)
    if rem > half or (rem == half and (mant & 1)):
        mant += 1
    return F32(float(mant) * 2.0 ** sh)


def f32_of_fixed(x, S):
    """exact integer x * 2^-S rounded once to float32"""
    return F32(int_to_f32(x) * F32(2.0 ** -S))


def fixed_of_f32(x, S):
    return int(float(x) * (2 ** S))


def pinned_sum(ws):
    """sumNeighWeights / computeCommNeighSum lanes + shuffle tree (32 lanes)"""
    part = [F32(0.0)] * 32
    for lane in range(32):
        s = F32(0.0)
        for j in range(lane, len(ws), 32):
            s = F32(s + ws[j])
        part[lane] = s
    for off in (16, 8, 4, 2, 1):
        part = [F32(part[i] + part[i + off]) if i + off < 32 else F32(part[i] + part[i])
                for i in range(32)]
    return part[0]


class Sim:
    def __init__(self, row, col, w, S):
        self.row = [int(x) for x in row]
        self.col = [int(x) for x in col]
        self.w = [F32(x) for x in w]
        self.S = S
        self.n = len(self.row) - 1
        self.k = [pinned_sum(self.w[self.row[v]:self.row[v + 1]]) for v in range(self.n)]
        self.kfix = [fixed_of_f32(x, S) for x in self.k]
        self.m = F32(f32_of_fixed(sum(self.kfix), S) / F32(2.0))
        self.comm = list(range(self.n))
        self.comm_sz = [1] * self.n
        self.A = list(self.kfix)

    def deg(self, v):
        return self.row[v + 1] - self.row[v]

    def is_lower_idx(self, v, e):
        vc, nc = self.comm[v], self.comm[e]
        if self.comm_sz[vc] == 1 and self.comm_sz[nc] == 1:
            return nc < vc
        return True

    def gain(self, hw_fixed, v, e):
        m = self.m
        vw = self.k[v]
        a_v = f32_of_fixed(self.A[self.comm[v]], self.S)
        a_e = f32_of_fixed(self.A[self.comm[e]], self.S)
        t = F32(F32(hw_fixed_to_f(hw_fixed, self.S)) / m)
        num = F32(vw * F32(F32(a_v - vw) - a_e))
        den = F32(F32(F32(2.0) * m) * m)
        return F32(t + F32(num / den))

    # computeMoveWarp: T lanes, lock-step rounds
    def warp_vertex(self, v, T):
        b, deg = self.row[v], self.deg(v)
        w_fix = [fixed_of_f32(x, self.S) for x in self.w]
        hashw = {}
        comm_e_w = 0
        lanes = [{"max": NEG_INF, "max_e": None, "max_comm": GARBAGE} for _ in range(T)]
        rounds = (deg + T - 1) // T
        for r in range(rounds):
            active = [ln for ln in range(T) if ln + r * T < deg]
            # every active lane inserts before any lane reads (lock-step)
            for ln in active:
                j = ln + r * T
                e = self.col[b + j]
                hashw[self.comm[e]] = hashw.get(self.comm[e], 0) + w_fix[b + j]
                if self.comm[e] == self.comm[v] and e != v:
                    comm_e_w += w_fix[b + j]
            for ln in active:
                j = ln + r * T
                e = self.col[b + j]
                if self.is_lower_idx(v, e):
                    g = self.gain(hashw[self.comm[e]], v, e)
                    if g > lanes[ln]["max"]:
                        lanes[ln]["max"] = g
                        lanes[ln]["max_e"] = e
                        lanes[ln]["max_comm"] = self.comm[e]
        # shuffle-down reduction; lanes past the group read foreign values
        reg = [(lanes[ln]["max"], ln, lanes[ln]["max_comm"]) for ln in range(T)]
        off = T // 2
        while off > 0:
            nxt = []
            for ln in range(T):
                src = ln + off
                tmp = reg[src] if src < T else (F32(np.inf), 999, 0)  # foreign lane
                cur = reg[ln]
                if tmp[0] > cur[0] or (tmp[0] == cur[0] and tmp[2] < cur[2]):
                    cur = tmp
                nxt.append(cur)
            reg = nxt
            off //= 2
        pos = reg[0][1]
        if pos >= deg or pos >= T:
            return self.comm[v]  # no lane writes new_comm: canonical stay
        lane = lanes[pos]
        cew = f32_of_fixed(comm_e_w, self.S)
        if lane["max"] == NEG_INF:
            return self.comm[v]
        if F32(lane["max"] - F32(cew / self.m)) > F32(0.0):
            return self.comm[lane["max_e"]]
        return self.comm[v]

    # computeMoveBlock / computeMoveGlobal with complete sums
    def block_vertex(self, v, T):
        b, deg = self.row[v], self.deg(v)
        w_fix = [fixed_of_f32(x, self.S) for x in self.w]
        hashw = {}
        comm_e_w = 0
        for j in range(deg):
            e = self.col[b + j]
            hashw[self.comm[e]] = hashw.get(self.comm[e], 0) + w_fix[b + j]
            if self.comm[e] == self.comm[v] and e != v:
                comm_e_w += w_fix[b + j]
        lanes = []
        for ln in range(T):
            mx, mc, e_reg = NEG_INF, GARBAGE, None
            j = ln
            while j < deg:
                e_reg = self.col[b + j]
                if self.is_lower_idx(v, e_reg):
                    g = self.gain(hashw[self.comm[e_reg]], v, e_reg)
                    if g > mx:
                        mx, mc = g, self.comm[e_reg]
                j += T
            lanes.append({"max": mx, "comm": mc, "e": e_reg})
        cache_max = [ln["max"] for ln in lanes]
        cache_pos = list(range(T))
        cache_comm = [ln["comm"] for ln in lanes]
        off = T // 2
        while off > 0:
            for i in range(off):
                if cache_max[i + off] > cache_max[i] or (
                        cache_max[i + off] == cache_max[i] and cache_comm[i + off] < cache_comm[i]):
                    cache_max[i] = cache_max[i + off]
                    cache_pos[i] = cache_pos[i + off]
                    cache_comm[i] = cache_comm[i + off]
            off //= 2
        pos = cache_pos[0]
        if pos >= deg:
            return self.comm[v]
        lane = lanes[pos]
        if lane["max"] == NEG_INF:
            return self.comm[v]
        cew = f32_of_fixed(comm_e_w, self.S)
        if F32(lane["max"] - F32(cew / self.m)) > F32(0.0):
            return self.comm[lane["e"]]  # graph.comm[e]: the last visited entry
        return self.comm[v]

    def step(self, bucket):
        lb, ub = CUTS[bucket - 1], CUTS[bucket]
        T = BLOCKS[bucket - 1][0]
        verts = [v for v in range(self.n) if lb < self.deg(v) <= ub]
        new = {}
        for v in verts:
            new[v] = self.warp_vertex(v, T) if T <= 32 else self.block_vertex(v, T)
        for v in verts:  # assignCommunity + exact community weights
            o, nc = self.comm[v], new[v]
            self.comm_sz[o] -= 1
            self.comm_sz[nc] += 1
            self.A[o] -= self.kfix[v]
            self.A[nc] += self.kfix[v]
            self.comm[v] = nc

    def modularity(self):
        s_in = 0
        for v in range(self.n):
            b, e = self.row[v], self.row[v + 1]
            ws = [self.w[j] if self.comm[self.col[j]] == self.comm[v] else F32(0.0)
                  for j in range(b, e)]
            s_in += fixed_of_f32(pinned_sum(ws), self.S)
        sq = 0
        for c in range(self.n):
            a = f32_of_fixed(self.A[c], self.S)
            sq += int(float(F32(a * a)) * (2 ** (2 * self.S)))
        S_in = f32_of_fixed(s_in, self.S)
        a2 = F32(float(int_to_f32(sq)) * 2.0 ** (-2 * self.S))
        m = self.m
        return F32(F32(S_in / F32(F32(2.0) * m)) - F32(a2 / F32(F32(F32(4.0) * m) * m)))


def hw_fixed_to_f(x, S):
    return f32_of_fixed(x, S)
