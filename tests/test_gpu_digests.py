"""GPU results on the full-size workloads against the oracle's committed
digests (tests/golden/digests_*.json, written on the CPU by
tests/golden/make_digests.py).

* R-MAT scale 24 (16.8M vertices, 520.7M entries): per pass the sweep count,
  the FP32 bits of Q before and after, the size and SHA-256 of the contracted
  CSR and of the composed assignment; then the final printed Q, passes,
  sweeps, community count, summed entries and assignment of a full lv_run,
  with the locality order on and off.
* The 4,096-graph SBM batch (33.9M vertices, 665M entries) in one
  lv_batch_run: every graph's Q bits, passes, sweeps, community count, summed
  entries and assignment.
* The serving pool (12 SBM graphs of 1k-200k vertices, R-MAT scale 14-21)
  streamed through the server with every graph submitted twice and admitted
  concurrently: every served result against the oracle.

Bit-exact throughout (modularity.cu:438-520, aggregation.cu:387-462,
main.cu:43-48 under the canonical arithmetic of the oracle).
"""
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
MIN_GAIN = 1e-6


def _digest(name):
    with open(os.path.join(GOLDEN, name)) as f:
        return json.load(f)


@pytest.fixture(scope="module")
def W(lv):
    if lv.device_count() < 1:
        pytest.fail("no GPU visible: GPU tests must run on a GPU box")
    from louvain_amd import workloads
    return workloads


def _qb(q):
    return int(np.float32(q).view(np.uint32))


def _short(W, a):
    return W.sha256_arrays(np.asarray(a, dtype=np.uint32))[:16]


def test_rmat24_matches_oracle_digests(lv, orc, W):
    d = _digest("digests_rmat-24-ef16.json")
    row, col, w = W.make_graph(lv, W.CONFIGS["c3"])
    print(f"[case] rmat24 n={row.size - 1} ne={col.size}", flush=True)
    assert W.csr_digest(row, col, w) == d["input"], "generator output changed: regenerate digests"
    with lv.Graph(row, col, w) as g:
        # pass by pass through the reference operator API
        for p, exp in enumerate(d["passes_detail"]):
            imp, qb, qa, it = g.modularity_optimization(MIN_GAIN)
            print(f"[case] rmat24 pass {p}: sweeps={it} improved={imp}", flush=True)
            if p == 0:
                S, m = g.scale()
                assert (S, _qb(m)) == (d["scale_bits"], d["m_bits"])
            assert (imp, it) == (exp["improved"], exp["sweeps"]), p
            assert (_qb(qb), _qb(qa)) == (exp["q_before_bits"], exp["q_after_bits"]), p
            if not imp:
                assert _qb(qb) == d["q_bits"]  # what the reference prints
                break
            g.aggregation()
            n, ne, _ = g.size()
            assert (n, ne) == (exp["next_n"], exp["next_ne"]), p
            r2, c2, w2 = g.csr()
            assert W.csr_digest(r2, c2, w2) == exp["next_csr"], f"contracted CSR of pass {p}"
            del r2, c2, w2
            a = g.assignment()
            assert W.sha256_arrays(a) == exp["assign_after"], f"assignment after pass {p}"
        assert g.size()[0] == d["n_final"]
        # the whole run, with the locality order on (automatic) and off
        for loc in (-1, 0):
            g.reset()
            g.set_locality(loc)
            r = g.run(MIN_GAIN)
            a = g.assignment()
            print(f"[case] rmat24 lv_run loc={loc}: {r.alg_ms:.1f} ms", flush=True)
            assert _qb(r.q_final) == d["q_bits"], loc
            assert (r.passes, r.total_iters, r.n_final, r.sum_edges) == \
                (d["passes"], d["sweeps"], d["n_final"], d["sum_edges"]), loc
            assert W.sha256_arrays(a) == d["assign"], loc
    q64 = orc.q64(row.astype(np.uint64), col, w, a)
    assert abs(float(np.float32(d["q_final"])) - q64) <= 1e-6, q64


def test_batch4096_matches_oracle_digests(lv, W):
    d = _digest("digests_batch-4096-sbm.json")
    graphs = W.make_graph(lv, W.CONFIGS["c4"])
    G = len(graphs)
    print(f"[case] batch of {G} SBM graphs, {sum(c.size for _, c, _ in graphs)} entries",
          flush=True)
    assert G == d["graphs"]
    assert [W.csr_digest(*x)[:16] for x in graphs] == d["input"], "generator output changed"
    with lv.Batch(graphs) as b:
        per, tot = b.run(MIN_GAIN)
        assigns = b.assignments()
    print(f"[case] batch run {tot.alg_ms:.1f} ms", flush=True)
    bad = []
    for i in range(G):
        got = (per[i].status, _qb(per[i].q_final), per[i].passes, per[i].total_iters,
               per[i].n_final, per[i].sum_edges, _short(W, assigns[i]))
        exp = (0, d["q_bits"][i], d["passes"][i], d["sweeps"][i], d["n_final"][i],
               d["sum_edges"][i], d["assign"][i])
        if got != exp:
            bad.append((i, got, exp))
    assert not bad, f"{len(bad)} graphs differ, first: {bad[:3]}"


def test_served_pool_matches_oracle_digests(lv, W):
    d = _digest("digests_serve-pool.json")["graphs"]
    seed = W.CONFIGS["c5"]["args"][2]
    sbm, rmat = W.serve_pool(lv, seed)
    pool = sbm + rmat
    assert [W.csr_digest(*x)[:16] for x in pool] == [e["input"] for e in d]
    # every pool graph twice, large and small interleaved, all submitted at
    # once: the server admits them as its slots and budget free up
    idx = sorted(range(len(pool)), key=lambda i: -pool[i][1].size)
    order = []
    for k in range(len(idx) // 2):
        order += [idx[k], idx[-1 - k]]
    order += order[::-1]
    print(f"[case] server: {len(order)} requests over {len(pool)} graphs", flush=True)
    with lv.Server(slots=8, hbm_budget=200 << 30, keep_assignments=True) as s:
        for rid, gi in enumerate(order):
            s.submit(rid, *pool[gi])
        s.drain()
        out = s.poll(4096)
        assert len(out) == len(order)
        for res in out:
            gi = order[res.id]
            e = d[gi]
            got = (res.status, _qb(res.q_final), res.passes, res.total_iters, res.n_final,
                   res.sum_edges)
            assert got == (0, e["q_bits"], e["passes"], e["sweeps"], e["n_final"],
                           e["sum_edges"]), (res.id, gi)
            a = s.assignment(res.id, pool[gi][0].size - 1)
            assert _short(W, a) == e["assign"], (res.id, gi)
    # the requests did run side by side: some ran while others were running
    spans = sorted((r.t_start_ms, r.t_end_ms) for r in out)
    overlaps = sum(1 for (s0, _), (_, e1) in zip(spans[1:], spans[:-1]) if s0 < e1)
    assert overlaps > 0
