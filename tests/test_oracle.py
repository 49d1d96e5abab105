"""Pins the CPU oracle (CPU only).

Every check here is independent of the oracle's own formulation: values
derived by hand from the reference source, closed forms, FP64 recomputes,
invariants, and a lane-level simulation of the reference kernels
(tests/refsim.py) on random small graphs.
"""
import numpy as np
import pytest

import refsim

F32 = np.float32


def mtx(n, edges):
    lines = [f"{n} {n} {len(edges)}"]
    for e in edges:
        lines.append(" ".join(str(x) for x in e))
    return "\n".join(lines) + "\n"


# ---------------------------------------------------------------- parser ----

def test_parse_semantics(orc):
    # '%' comments, 1-based ids, self-loop stored once, abs(), 2-token lines
    # carry the previous weight, file order inside rows (utils.cu:132-229)
    text = "%%MatrixMarket matrix coordinate real symmetric\n% c\n4 4 5\n1 1 -2.5\n2 1 0.5\n3 2\n4 4 3\n1 3 -1\n"
    row, col, w = orc.parse_mtx(text=text)
    assert row.tolist() == [0, 3, 5, 7, 8]
    assert col.tolist() == [0, 1, 2, 0, 2, 1, 0, 3]
    assert w.tolist() == [2.5, 0.5, 1.0, 0.5, 0.5, 0.5, 1.0, 3.0]


def test_parse_zero_based(orc):
    row, col, w = orc.parse_mtx(text="3 3 2\n0 1 1\n1 2 2\n")
    assert row.tolist() == [0, 1, 3, 4]
    assert col.tolist() == [1, 0, 2, 1]


def test_parse_default_weight_and_errors(orc):
    row, col, w = orc.parse_mtx(text="2 2 1\n1 2\n")
    assert w.tolist() == [1.0, 1.0]
    for bad in ["", "2 2 2\n1 2 1\n", "2 2 1\n1 3 1\n", "2 2 1\nx y\n"]:
        with pytest.raises(orc.OracleError):
            orc.parse_mtx(text=bad)


def test_bcsstk16_facts(orc, bcsstk16_text):
    row, col, w = orc.parse_mtx(text=bcsstk16_text)
    deg = np.diff(row.astype(np.int64))
    assert row.size - 1 == 4884 and col.size == 290378
    assert deg.max() == 81
    assert int((col == np.repeat(np.arange(4884), deg)).sum()) == 4884  # self-loops
    # zero-weight entries counted straight from the text: an off-diagonal line
    # of weight 0 gives two CSR entries, a diagonal one gives one
    zeros = 0
    for line in bcsstk16_text.splitlines()[1:]:
        if line.startswith("%"):
            continue
        p = line.split()
        if len(p) >= 3 and float(p[2]) == 0.0:
            zeros += 1 if p[0] == p[1] else 2
    assert zeros == 1492
    assert int((w == 0).sum()) == zeros
    assert orc.scale_bits(w) == 29


def test_weights_beyond_the_exact_range(orc):
    """A weight whose fixed-point value reaches 2^61 is refused before its
    conversion to an integer (it would not be representable in int64)."""
    row = np.array([0, 1, 2], np.uint64)
    col = np.array([1, 0], np.uint32)
    for big in (1e30, 2.0 ** 70, 2.0 ** 61):
        with pytest.raises(orc.OracleError):
            orc.Oracle(row, col, np.array([big, big], np.float32))
    # just below the limit: accepted, m is the weight itself
    x = np.float32(2.0 ** 59)
    o = orc.Oracle(row, col, np.array([x, x], np.float32))
    o.pass_init()
    assert o.m == x


# ------------------------------------------------------- arithmetic pins ----

def test_r1_sum_order(orc):
    # lane 0 holds entries 0 and 32: 1 + 2^-24 rounds to 1 before lane 1's
    # 2^-24 arrives; a pairwise or sequential order would give 1 + 2^-23
    w = np.zeros(33, np.float32)
    w[0] = 1.0
    w[32] = 2.0 ** -24
    w[1] = 2.0 ** -24
    assert orc.r1_sum(w) == F32(1.0)
    assert F32(np.float64(w.astype(np.float64).sum())) == F32(1.0 + 2.0 ** -23)
    # dyadic weights: any order is exact
    w = np.arange(1, 101, dtype=np.float32) / 8
    assert orc.r1_sum(w) == F32(w.astype(np.float64).sum())
    assert orc.r1_sum(np.zeros(0, np.float32)) == 0


def test_scale_bits(orc):
    assert orc.scale_bits(np.array([1.0, 3.0], np.float32)) == 0
    assert orc.scale_bits(np.array([0.5], np.float32)) == 1
    assert orc.scale_bits(np.array([0.75, 2.0], np.float32)) == 2
    assert orc.scale_bits(np.array([0.1], np.float32)) == 27  # 13421773 * 2^-27
    assert orc.scale_bits(np.array([0.0, 4.0], np.float32)) == 0
    with pytest.raises(orc.OracleError):
        orc.scale_bits(np.array([np.inf], np.float32))
    # tiny weights leave the exact fixed-point range
    row = np.array([0, 1, 2], np.uint64)
    with pytest.raises(orc.OracleError):
        orc.Oracle(row, np.array([1, 0], np.uint32), np.array([1e-30, 1e-30], np.float32))


# ------------------------------------------------------ closed forms --------

TWO_TRI = mtx(6, [(1, 2, 1), (2, 3, 1), (1, 3, 1), (4, 5, 1), (5, 6, 1), (4, 6, 1), (3, 4, 1)])


def test_two_triangles(orc):
    row, col, w = orc.parse_mtx(text=TWO_TRI)
    o = orc.Oracle(row, col, w)
    # singleton Q0 = -sum k^2 / (4 m^2) = -(4+4+9+9+4+4)/196
    assert abs(float(o.pass_init()) - (-34 / 196)) < 1e-7
    assert o.m == F32(7.0)
    res = o.run(1e-6)
    assert abs(res.q_final - 5 / 14) < 1e-7
    assert o.init_comm().tolist() == [0, 0, 0, 1, 1, 1]
    assert res.n_final == 2


def test_two_triangles_first_step_and_contraction(orc):
    row, col, w = orc.parse_mtx(text=TWO_TRI)
    o = orc.Oracle(row, col, w)
    o.pass_init()
    # every vertex is in bucket 1, all decide from the singleton snapshot:
    #  0: neighbours 1,2 have higher ids -> stays
    #  1: candidate 0 -> 0;  2: candidates 1,0, equal gains -> lower id 0
    #  3: candidate 2, gain 1/7 - 3*3/98 = 5/98 > 0 -> 2
    #  4: candidate 3, gain 1/7 - 2*3/98 = 8/98 -> 3
    #  5: candidates 3 (8/98) and 4 (1/7 - 2*2/98 = 10/98) -> 4
    o.step(1)
    assert o.comm().tolist() == [0, 0, 0, 2, 3, 4]
    res = o.run(1e-6)
    assert res.n_final == 2
    r, c, ww = o.csr()
    # final graph = contraction into the two triangles: internal edges counted
    # from both ends (3*2), the bridge once per side
    assert r.tolist() == [0, 2, 4]
    assert c.tolist() == [0, 1, 0, 1]
    assert ww.tolist() == [6.0, 1.0, 1.0, 6.0]
    # its singleton Q equals the partition's Q on the original graph
    q_next = o.pass_init()
    assert abs(float(q_next) - orc.q64(row, col, w, o.init_comm())) < 1e-7


def test_path_first_step(orc):
    # P4 0-1-2-3, unit weights, bucket 1 (deg <= 4), all singletons:
    #  0: only neighbour 1 has a higher id -> no candidate, stays
    #  1: candidate 0, gain 1/3 - 2*1/18 = 2/9 > 0 -> 0
    #  2: candidate 1, gain 1/3 - 2*2/18 = 1/9 > 0 -> 1
    #  3: candidate 2, gain 1/3 - 1*2/18 = 2/9 > 0 -> 2
    row, col, w = orc.parse_mtx(text=mtx(4, [(1, 2, 1), (2, 3, 1), (3, 4, 1)]))
    o = orc.Oracle(row, col, w)
    o.pass_init()
    assert o.step(1) == 3
    assert o.comm().tolist() == [0, 0, 1, 2]
    assert o.comm_sz().tolist() == [2, 1, 1, 0]


def test_tie_lowest_comm(orc):
    # vertex 3 (0-based) joined to singletons 1 and 2 with equal gains: the
    # shuffle reduction breaks the tie toward the lower community id
    for order in ([(4, 2, 1), (4, 3, 1)], [(4, 3, 1), (4, 2, 1)]):
        row, col, w = orc.parse_mtx(text=mtx(4, order))
        o = orc.Oracle(row, col, w)
        o.pass_init()
        o.step(1)
        assert o.comm().tolist() == [0, 1, 2, 1]


def _lane_first_max_graph():
    # pool 0..84 (deg 31), B=85, A=86 (deg 1), v=87 (deg 33, bucket 5),
    # others 88..118 (deg 86).  v's row: A, 31 others, B.  In step 5 only A
    # (lane 0, round 0) and B (lane 0, round 1) are candidates, with equal
    # gains; lane 0 keeps its first maximum, so v joins A although B < A.
    edges = [(88, 87, 1)]
    edges += [(88, 89 + i, 1) for i in range(31)]
    edges += [(88, 86, 1)]
    for o_ in range(89, 120):
        for p in range(1, 86):
            edges.append((o_, p, 1))
    return mtx(119, edges)


def test_lane_keeps_first_max(orc):
    row, col, w = orc.parse_mtx(text=_lane_first_max_graph())
    o = orc.Oracle(row, col, w)
    o.pass_init()
    for b in range(1, 6):
        o.step(b)
    assert o.comm()[87] == 86


def test_block_last_edge_destination(orc):
    # A=0, v=1 (deg 130, bucket 6), Y=2 (deg 129).  v's row: A then 129 x Y.
    # Only A is a candidate for v (lane 0, entry 0); lane 0's last visited
    # entry is 128 (-> Y), so v moves to comm[Y] = 2 (computeMoveBlock writes
    # graph.comm[e]).  Y's only candidate is v; it moves to comm[v] = 1.
    edges = [(2, 1, 1)] + [(2, 3, 1)] * 129
    row, col, w = orc.parse_mtx(text=mtx(3, edges))
    o = orc.Oracle(row, col, w)
    o.pass_init()
    for b in range(1, 7):
        o.step(b)
    assert o.comm().tolist() == [0, 2, 1]


# --------------------------------------------- lane-level simulation --------

def _random_graph(rng, n, kind):
    edges = []
    if kind == "sparse":
        m = int(rng.integers(n, 4 * n))
        for _ in range(m):
            edges.append((int(rng.integers(1, n + 1)), int(rng.integers(1, n + 1))))
    else:  # hubs made of multi-edges: buckets 5-7
        for _ in range(int(rng.integers(n, 2 * n))):
            edges.append((int(rng.integers(1, n + 1)), int(rng.integers(1, n + 1))))
        for hub in rng.choice(np.arange(1, n + 1), size=3, replace=False):
            for _ in range(int(rng.choice([40, 100, 200, 400]))):
                edges.append((int(hub), int(rng.integers(1, n + 1))))
    wkind = rng.integers(0, 3)
    out = []
    for a, b in edges:
        if wkind == 0:
            wt = 1
        elif wkind == 1:
            wt = float(rng.integers(1, 8)) / 4
        else:
            wt = float(np.float32(rng.random() + 0.01))
        out.append((a, b, repr(wt)))
    return mtx(n, out)


@pytest.mark.parametrize("seed", range(40))
def test_oracle_matches_lane_simulation(orc, seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(6, 40))
    text = _random_graph(rng, n, "sparse" if seed % 2 == 0 else "hubs")
    row, col, w = orc.parse_mtx(text=text)
    o = orc.Oracle(row, col, w)
    q0 = o.pass_init()
    sim = refsim.Sim(row, col, w, o.scale)
    assert np.array_equal(np.array(sim.k, np.float32), o.k())
    assert sim.m == o.m
    assert sim.modularity() == q0
    for it in range(3):
        for b in range(1, 8):
            o.step(b)
            sim.step(b)
            assert o.comm().tolist() == sim.comm, (seed, it, b)
            assert o.a_fixed().tolist() == sim.A
        assert sim.modularity() == o.modularity()


# ------------------------------------------------------------- invariants ----

def test_bcsstk16_run_invariants(orc, bcsstk16_text):
    row, col, w = orc.parse_mtx(text=bcsstk16_text)
    o = orc.Oracle(row, col, w)
    res = o.run(1e-6)
    assign = o.init_comm()
    q64 = orc.q64(row, col, w, assign)
    assert abs(res.q_final - q64) <= 1e-6
    assert res.n_final == len(np.unique(assign)) == assign.max() + 1
    assert res.q_final > 0.7


def test_contraction_invariants(orc, lv):
    row, col, w = lv.gen_chung_lu(3000, 8.0, 2.5, 200, 30, 20.0, 3)
    o = orc.Oracle(row.astype(np.uint64), col, w)
    total = np.float64(w.astype(np.float64).sum())
    for _ in range(3):
        imp, _, _ = o.modularity_optimization(1e-6)
        if not imp:
            break
        comm = o.comm()
        o.aggregation()
        r, c, ww = o.csr()
        # unit weights: exact conservation and symmetry
        assert ww.astype(np.float64).sum() == total
        dense = {}
        for v in range(r.size - 1):
            for j in range(int(r[v]), int(r[v + 1])):
                dense[(v, int(c[j]))] = ww[j]
            assert np.all(np.diff(c[int(r[v]):int(r[v + 1])].astype(np.int64)) > 0)
        for (a, b), x in dense.items():
            assert dense[(b, a)] == x
        # renumbering is the rank of each non-empty community
        ids = np.unique(comm)
        assert r.size - 1 == ids.size
        assert abs(float(o.pass_init()) - orc.q64(row.astype(np.uint64), col, w,
                                                  o.init_comm())) < 1e-6


def test_threads_do_not_change_results(lv, orc):
    """The oracle on several host threads (the all-core CPU baseline) gives
    the single-thread result: decisions read only the step snapshot and every
    total is an exact integer sum."""
    for r, c, w in [lv.gen_rmat(15, 16, 0.57, 0.19, 0.19, 0.5, 1.5, 21),
                    lv.gen_sbm(40000, 30, 12.0, 10.0, 22)]:
        out = []
        for threads in (1, 4):
            o = orc.Oracle(r.astype(np.uint64), c, w)
            o.set_threads(threads)
            res = o.run(1e-6)
            out.append((F32(res.q_final).view(np.uint32), res.passes, res.total_iters,
                         res.n_final, res.sum_edges, o.init_comm().tobytes()))
        assert out[0] == out[1]

