"""GPU parity: the HIP engine (through the C ABI) against the CPU oracle.

Bit-exact for every integer and FP32 quantity: k_v, m, community assignment,
community sizes and exact community weights after every bucket step, the
FP32 modularity after every sweep, renumbering and the contracted CSR after
every aggregation, and the final printed modularity and assignment.  The
final modularity is also checked against an FP64 recompute on the original
graph (|dQ| <= 1e-6).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

Q64_TOL = 1e-6


def _graphs(lv, orc, text):
    out = {}
    out["bcsstk16"] = orc.parse_mtx(text=text)
    out["fem"] = lv.gen_fem(814, 25, 9, 1)
    out["chung_lu_20k"] = lv.gen_chung_lu(20000, 9.4, 2.5, 420, 100, 20.0, 5)
    out["rmat12"] = lv.gen_rmat(12, 16, 0.57, 0.19, 0.19, 0.5, 1.5, 7)
    out["rmat14_unit"] = lv.gen_rmat(14, 16, 0.57, 0.19, 0.19, 1.0, 1.0, 9)
    out["sbm_5k"] = lv.gen_sbm(5000, 20, 12.0, 10.0, 11)
    out["rmat16"] = lv.gen_rmat(16, 16, 0.57, 0.19, 0.19, 0.5, 1.5, 13)
    return out


@pytest.fixture(scope="module")
def graphs(lv, orc, bcsstk16_text):
    if lv.device_count() < 1:
        pytest.fail("no GPU visible: GPU tests must run on a GPU box")
    return _graphs(lv, orc, bcsstk16_text)


def _u32(a):
    return np.asarray(a).astype(np.uint32)


def _cmp_state(g, o, where):
    k, comm, sz, a = g.state()
    assert np.array_equal(comm, o.comm()), f"comm differs {where}"
    assert np.array_equal(sz, o.comm_sz()), f"comm_sz differs {where}"
    assert np.array_equal(a, o.a_fixed()), f"a_c differs {where}"
    assert np.array_equal(k.view(np.uint32), o.k().view(np.uint32)), f"k differs {where}"


def _cmp_csr(g, o, where):
    row, col, w = g.csr()
    orow, ocol, ow = o.csr()
    assert np.array_equal(row.astype(np.uint64), orow), f"row_ptr differs {where}"
    assert np.array_equal(col, ocol), f"col differs {where}"
    assert np.array_equal(w.view(np.uint32), ow.view(np.uint32)), f"weights differ {where}"


# locality 0: input numbering on the device; 1: every pass graph in locality
# order (ids compared through lid) -- results must be identical either way
@pytest.mark.parametrize("loc", [0, 1])
@pytest.mark.parametrize("name", ["bcsstk16", "fem", "chung_lu_20k", "rmat12", "rmat14_unit",
                                  "sbm_5k", "rmat16"])
def test_stepwise_first_pass(lv, orc, graphs, name, loc):
    row, col, w = graphs[name]
    print(f"[case] stepwise {name} loc={loc} n={row.size - 1} ne={col.size}", flush=True)
    o = orc.Oracle(row.astype(np.uint64), col, w)
    with lv.Graph(_u32(row), col, w) as g:
        g.set_locality(loc)
        s, m = g.scale()
        assert s == o.scale
        q0 = g.pass_init()
        oq0 = o.pass_init()
        assert g.scale()[1] == o.m
        assert q0.view(np.uint32) == oq0.view(np.uint32), (q0, oq0)
        _cmp_state(g, o, "after pass init")
        for it in range(3):
            for b in range(1, 8):
                mv = g.step(b)
                omv = o.step(b)
                assert mv == omv, f"moved count differs it={it} b={b}"
                _cmp_state(g, o, f"it={it} bucket={b}")
            q = g.modularity()
            oq = o.modularity()
            assert q.view(np.uint32) == oq.view(np.uint32), (it, q, oq)


def _with_zero_links(row, col, w, pairs):
    """The CSR with undirected links of weight zero added (rows kept sorted)."""
    row = np.asarray(row).astype(np.int64)
    n = row.size - 1
    rows = [list(zip(col[row[v]:row[v + 1]].tolist(), w[row[v]:row[v + 1]].tolist()))
            for v in range(n)]
    for a, b in pairs:
        rows[a].append((b, 0.0))
        rows[b].append((a, 0.0))
    for v in {x for p in pairs for x in p}:
        rows[v].sort(key=lambda t: t[0])
    nrow = np.concatenate([[0], np.cumsum([len(x) for x in rows])]).astype(np.uint32)
    ncol = np.array([t[0] for x in rows for t in x], np.uint32)
    nw = np.array([t[1] for x in rows for t in x], np.float32)
    return nrow, ncol, nw


def test_small_contraction_paths(lv, orc, graphs):
    """Contractions of small pass graphs take one of three paths: one
    workgroup for a tiny graph (at most 1024 vertices and 4096 entries), a
    dense matrix of exact sums when at most 64 communities remain of at most
    65,536 vertices, the general path otherwise.  Every contraction gives the
    oracle's CSR, weights bit for bit, and assignment.  The FEM graph gets
    three long links of weight zero: its first contraction (23 communities,
    the matrix path) keeps them as entries of weight zero, as the general
    path does; bcsstk16 carries 1,492 entries of weight zero through the
    general and the tiny path."""
    fem = _with_zero_links(*graphs["fem"], [(0, 4883), (100, 4000), (2000, 4500)])
    cases = {"fem_zero_links": fem, "bcsstk16": graphs["bcsstk16"], "sbm_5k": graphs["sbm_5k"]}
    gain = np.float32(1e-6)
    seen, zero_kept = set(), None
    for name, (row, col, w) in cases.items():
        print(f"[case] small contractions {name}", flush=True)
        o = orc.Oracle(np.asarray(row).astype(np.uint64), col, w)
        with lv.Graph(_u32(row), col, w) as g:
            g.set_locality(0)
            for p in range(20):
                imp = g.modularity_optimization(gain)[0]
                assert imp == o.modularity_optimization(gain)[0], f"{name} pass {p}"
                if not imp:
                    break
                n, ne, _ = g.size()
                g.aggregation()
                o.aggregation()
                nn = g.size()[0]
                seen.add("tiny" if n <= 1024 and ne <= 4096 else
                         "few" if n <= 65536 and nn <= 64 else "general")
                _cmp_csr(g, o, f"{name} after aggregation {p}")
                assert np.array_equal(g.assignment(), o.init_comm()), f"{name} pass {p}"
                if name == "fem_zero_links" and p == 0:
                    zero_kept = int((g.csr()[2] == 0).sum())
    assert seen == {"tiny", "few", "general"}, seen
    assert zero_kept == 6, zero_kept


# locality 0: input numbering on the device; 1: every pass graph in locality
# order (ids compared through lid) -- results must be identical either way
@pytest.mark.parametrize("loc", [0, 1])
@pytest.mark.parametrize("name", ["bcsstk16", "fem", "chung_lu_20k", "rmat12", "rmat14_unit",
                                  "sbm_5k", "rmat16"])
def test_passes_and_aggregation(lv, orc, graphs, name, loc):
    row, col, w = graphs[name]
    print(f"[case] passes {name} loc={loc}", flush=True)
    o = orc.Oracle(row.astype(np.uint64), col, w)
    gain = np.float32(1e-6)
    with lv.Graph(_u32(row), col, w) as g:
        g.set_locality(loc)
        for p in range(20):
            imp, qb, qa, it = g.modularity_optimization(gain)
            oimp, oqp, oit = o.modularity_optimization(gain)
            oqb, oqa = o.last_pass_q()
            assert imp == oimp, f"pass {p}: improved differs"
            assert it == oit, f"pass {p}: sweeps differ"
            assert qb.view(np.uint32) == oqb.view(np.uint32), f"pass {p}: q_before"
            assert qa.view(np.uint32) == oqa.view(np.uint32), f"pass {p}: q_after"
            _cmp_state(g, o, f"end of pass {p}")
            if not imp:
                # the value the reference prints is the pass's starting Q
                assert qb.view(np.uint32) == oqp.view(np.uint32)
                break
            g.aggregation()
            o.aggregation()
            _cmp_csr(g, o, f"after aggregation {p}")
            assert np.array_equal(g.assignment(), o.init_comm()), f"assignment pass {p}"


# locality 0: input numbering on the device; 1: every pass graph in locality
# order (ids compared through lid) -- results must be identical either way
@pytest.mark.parametrize("loc", [0, 1])
@pytest.mark.parametrize("name", ["bcsstk16", "fem", "chung_lu_20k", "rmat12", "rmat14_unit",
                                  "sbm_5k", "rmat16"])
def test_full_run(lv, orc, graphs, name, loc):
    row, col, w = graphs[name]
    print(f"[case] run {name} loc={loc}", flush=True)
    o = orc.Oracle(row.astype(np.uint64), col, w)
    ores = o.run(1e-6)
    with lv.Graph(_u32(row), col, w) as g:
        g.set_locality(loc)
        res = g.run(1e-6)
        assign = g.assignment()
        n_final = g.size()[0]
    assert np.float32(res.q_final).view(np.uint32) == np.float32(ores.q_final).view(np.uint32)
    assert res.passes == ores.passes and res.n_final == ores.n_final == n_final
    assert res.total_iters == ores.total_iters
    assert res.sum_edges == ores.sum_edges
    assert np.array_equal(assign, o.init_comm())
    q64 = orc.q64(row.astype(np.uint64), col, w, assign)
    assert abs(float(res.q_final) - q64) <= Q64_TOL, (res.q_final, q64)


def test_determinism(lv, graphs):
    row, col, w = graphs["rmat12"]
    outs = []
    for _ in range(3):
        with lv.Graph(_u32(row), col, w) as g:
            r = g.run(1e-6)
            outs.append((np.float32(r.q_final).view(np.uint32), g.assignment().tobytes()))
    assert all(x == outs[0] for x in outs)


def test_edge_cases(lv, orc):
    # two triangles joined by an edge; isolated vertices; self-loops; star
    cases = {
        "two_triangles": "6 6 7\n1 2 1\n2 3 1\n1 3 1\n4 5 1\n5 6 1\n4 6 1\n3 4 1\n",
        "isolated_and_loops": "8 8 6\n1 1 2\n1 2 1\n2 3 1\n3 1 1\n5 5 3\n6 7 0.5\n",
        "star": "40 40 39\n" + "".join(f"1 {i} 1\n" for i in range(2, 41)),
    }
    # high-degree vertices through duplicate edges (buckets 6 and 7, heavy class)
    rng = np.random.default_rng(5)
    lines = []
    for hub, deg in [(1, 300), (2, 1500), (3, 5000), (4, 21000)]:
        for _ in range(deg):
            lines.append(f"{hub} {int(rng.integers(4, 60))} {1 + int(rng.integers(0, 3))}\n")
    cases["dup_hubs"] = f"60 60 {len(lines)}\n" + "".join(lines)
    for name, text in cases.items():
        row, col, w = orc.parse_mtx(text=text)
        o = orc.Oracle(row, col, w)
        ores = o.run(1e-6)
        for loc in (0, 1):
            print(f"[case] edge {name} loc={loc}", flush=True)
            with lv.Graph(_u32(row), col, w) as g:
                g.set_locality(loc)
                res = g.run(1e-6)
                assign = g.assignment()
            assert np.float32(res.q_final).view(np.uint32) == \
                np.float32(ores.q_final).view(np.uint32), (name, loc)
            assert np.array_equal(assign, o.init_comm()), (name, loc)


def test_copurchase_full_size(lv, orc):
    """Full-size co-purchase workload (262 111 vertices) against the oracle."""
    row, col, w = lv.gen_chung_lu(262111, 9.4, 2.5, 420, 2000, 20.0, 2)
    print(f"[case] copurchase full n={row.size - 1} ne={col.size}", flush=True)
    o = orc.Oracle(row.astype(np.uint64), col, w)
    ores = o.run(1e-6)
    for loc in (0, 1):
        print(f"[case] copurchase full loc={loc}", flush=True)
        with lv.Graph(row, col, w) as g:
            g.set_locality(loc)
            res = g.run(1e-6)
            assign = g.assignment()
        assert np.float32(res.q_final).view(np.uint32) == \
            np.float32(ores.q_final).view(np.uint32), loc
        assert np.array_equal(assign, o.init_comm()), loc
        assert res.total_iters == ores.total_iters and res.n_final == ores.n_final, loc


def test_cli_output_matches_reference_format(lv, orc, bcsstk16_path):
    """`louvain -f bcsstk16.mtx -g 1e-6 -v`: line 1 the final modularity as the
    reference's default ostream prints it (modularity.cu:504), line 2 two
    times (main.cu:56), then the community listing of printCommunities
    (utils.cu:270-293), all from the oracle's result."""
    import subprocess
    row, col, w = orc.parse_mtx(path=bcsstk16_path)
    o = orc.Oracle(row, col, w)
    res = o.run(1e-6)
    assign = o.init_comm()
    r = subprocess.run([lv.CLI_PATH, "-f", bcsstk16_path, "-g", "0.000001", "-v"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    lines = r.stdout.split("\n")
    assert lines[0] == "%g" % float(np.float32(res.q_final))
    t = lines[1].split(" ")
    assert len(t) == 2 and float(t[0]) > 0 and float(t[1]) >= float(t[0])
    assert lines[2] == str(res.n_final)
    expect = []
    for c in range(res.n_final):
        members = np.nonzero(assign == c)[0] + 1
        expect.append(f"{c + 1} " + "".join(f"{m} " for m in members))
    assert lines[3:3 + res.n_final] == expect
    assert lines[3 + res.n_final:] == [""]


def _mixed_batch(lv, orc, bcsstk16_text):
    rng = np.random.default_rng(21)
    graphs = [orc.parse_mtx(text=bcsstk16_text), lv.gen_fem(200, 25, 9, 3),
              lv.gen_chung_lu(3000, 9.4, 2.5, 200, 30, 20.0, 4),
              lv.gen_rmat(11, 16, 0.57, 0.19, 0.19, 0.5, 1.5, 5),
              orc.parse_mtx(text="5 5 0\n"),                       # no edges
              orc.parse_mtx(text="6 6 7\n1 2 1\n2 3 1\n1 3 1\n4 5 1\n5 6 1\n4 6 1\n3 4 1\n")]
    for i in range(24):
        n = int(np.exp(rng.uniform(np.log(100), np.log(4000))))
        blocks = int(rng.integers(8, 33))
        deg = float(rng.uniform(8, 32))
        graphs.append(lv.gen_sbm(n, min(blocks, n), deg, 10.0, 100 + i))
    return [(r.astype(np.uint32), c, w) for r, c, w in graphs]


def test_batch_matches_single_runs(lv, orc, bcsstk16_text):
    """Every graph of a mixed batch gets the result of its own single run."""
    graphs = _mixed_batch(lv, orc, bcsstk16_text)
    print(f"[case] batch of {len(graphs)} graphs", flush=True)
    with lv.Batch(graphs) as b:
        per, tot = b.run(1e-6)
        assigns = b.assignments()
    for i, (r, c, w) in enumerate(graphs):
        o = orc.Oracle(r.astype(np.uint64), c, w)
        try:
            ores = o.run(1e-6)
        except orc.OracleError:
            assert per[i].status != 0, i  # graph without weight
            continue
        assert per[i].status == 0
        assert np.float32(per[i].q_final).view(np.uint32) == \
            np.float32(ores.q_final).view(np.uint32), (i, per[i].q_final, ores.q_final)
        assert per[i].passes == ores.passes and per[i].n_final == ores.n_final, i
        assert per[i].total_iters == ores.total_iters, i
        assert per[i].sum_edges == ores.sum_edges, i
        assert np.array_equal(assigns[i], o.init_comm()), i


def test_batch_compaction(lv, orc):
    """A batch pass rebuilds its degree classes from the active graphs once
    these hold at most half of the listed vertices (LV_COMPACT_MIN=1 lets a
    small batch do so, in a child process so the setting does not leak);
    every graph still gets the result of its own run."""
    import os
    import subprocess
    import sys
    code = (
        "import sys, numpy as np; sys.path.insert(0, %r); sys.path.insert(0, %r)\n"
        "from loader import load_package; import oracle\n"
        "lv = load_package(); rng = np.random.default_rng(7); graphs = []\n"
        "for i in range(40):\n"
        "    n = int(np.exp(rng.uniform(np.log(60), np.log(3000))))\n"
        "    r, c, w = lv.gen_sbm(n, min(int(rng.integers(4, 33)), n), float(rng.uniform(4, 24)), 10.0, 300 + i)\n"
        "    graphs.append((r.astype(np.uint32), c, w))\n"
        "with lv.Batch(graphs) as b:\n"
        "    per, tot = b.run(1e-6); assigns = b.assignments()\n"
        "for i, (r, c, w) in enumerate(graphs):\n"
        "    o = oracle.Oracle(r.astype(np.uint64), c, w); ores = o.run(1e-6)\n"
        "    assert np.float32(per[i].q_final).view(np.uint32) == np.float32(ores.q_final).view(np.uint32), i\n"
        "    assert per[i].total_iters == ores.total_iters and per[i].passes == ores.passes, i\n"
        "    assert np.array_equal(assigns[i], o.init_comm()), i\n"
        "print('compaction ok')\n"
    ) % (os.path.join(lv.HERE), os.path.dirname(lv.HERE))
    env = dict(os.environ, LV_COMPACT_MIN="1")
    print("[case] batch of 40 SBM graphs with class compaction", flush=True)
    r = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True,
                       timeout=240)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "compaction ok" in r.stdout


def test_batch_reset_repeatable(lv, orc, bcsstk16_text):
    graphs = _mixed_batch(lv, orc, bcsstk16_text)[:10]
    with lv.Batch(graphs) as b:
        p1, _ = b.run(1e-6)
        a1 = b.assignments()
        b.reset()
        p2, _ = b.run(1e-6)
        a2 = b.assignments()
    for x, y in zip(p1, p2):
        assert np.float32(x.q_final).view(np.uint32) == np.float32(y.q_final).view(np.uint32)
    assert all(np.array_equal(x, y) for x, y in zip(a1, a2))


@pytest.mark.parametrize("mode", ["1", "2", "dmax", "hostloop", "graph", "onestream"])
def test_heavy_global_table_path(lv, orc, mode):
    """The global-table kernels that take over heavy vertices whose hash
    buckets overflow LDS: forced for every heavy vertex (LV_HEAVY_GLOBAL=1),
    for every other one after the partitioned path ran (=2), and through
    buckets that really overflow a reduced capacity (LV_HEAVY_DMAX) so both
    paths serve one step; the stop rule on the host after every sweep
    (LV_HOST_LOOP, as carved handles run); the sweeps captured once per pass
    and replayed as a HIP graph (LV_GRAPH, the heavy chain and the block
    classes as graph branches instead of two live streams); and every chain
    on one stream (LV_ONE_STREAM, as profiler runs use). Each in a child
    process so the setting does not leak."""
    import os
    import subprocess
    import sys
    code = (
        "import sys, numpy as np; sys.path.insert(0, %r); sys.path.insert(0, %r)\n"
        "from loader import load_package; import oracle\n"
        "lv = load_package()\n"
        "for sc, tune in [(16, {}), (18, {}), (18, {'bucket_entries': 64})]:\n"
        "    r, c, w = lv.gen_rmat(sc, 16, 0.57, 0.19, 0.19, 0.5, 1.5, 13)\n"
        "    o = oracle.Oracle(r.astype(np.uint64), c, w); ores = o.run(1e-6)\n"
        "    g = lv.Graph(r, c, w); g.set_tuning(**tune); res = g.run(1e-6)\n"
        "    assert np.float32(res.q_final).view(np.uint32) == np.float32(ores.q_final).view(np.uint32), (sc, tune)\n"
        "    assert np.array_equal(g.assignment(), o.init_comm()), (sc, tune)\n"
        "print('global-table path ok')\n"
    ) % (os.path.join(lv.HERE), os.path.dirname(lv.HERE))
    env = dict(os.environ, **({"LV_HEAVY_DMAX": "48"} if mode == "dmax" else
                              {"LV_HOST_LOOP": "1"} if mode == "hostloop" else
                              {"LV_GRAPH": "1"} if mode == "graph" else
                              {"LV_ONE_STREAM": "1"} if mode == "onestream" else
                              {"LV_HEAVY_GLOBAL": mode}))
    print(f"[case] heavy global tables mode {mode}", flush=True)
    r = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True,
                       timeout=240)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "global-table path ok" in r.stdout


@pytest.mark.parametrize("env", [{"LV_AGG_GLOBAL": "1"}, {"LV_AGG_DMAX": "24"}])
def test_large_community_merge_paths(lv, orc, env):
    """Large communities are merged by key-range buckets in LDS; a community
    whose bucket overflows goes to the global-table merge.  Every community
    on the global tables (LV_AGG_GLOBAL), and a tiny bucket capacity so both
    paths serve communities of the same aggregation (LV_AGG_DMAX), each give
    the oracle's contracted graphs -- in a child process so the setting does
    not leak."""
    import os
    import subprocess
    import sys
    code = (
        "import sys, numpy as np; sys.path.insert(0, %r); sys.path.insert(0, %r)\n"
        "from loader import load_package; import oracle\n"
        "lv = load_package()\n"
        "for r, c, w in [lv.gen_rmat(16, 16, 0.57, 0.19, 0.19, 0.5, 1.5, 13),\n"
        "                lv.gen_chung_lu(60000, 9.4, 2.5, 420, 300, 20.0, 5)]:\n"
        "    o = oracle.Oracle(r.astype(np.uint64), c, w)\n"
        "    with lv.Graph(r, c, w) as g:\n"
        "        for p in range(20):\n"
        "            imp = g.modularity_optimization(1e-6)[0]\n"
        "            oimp, oqp, _ = o.modularity_optimization(1e-6)\n"
        "            assert imp == oimp\n"
        "            if not imp: break\n"
        "            g.aggregation(); o.aggregation()\n"
        "            gr, gc, gw = g.csr(); orow, ocol, ow = o.csr()\n"
        "            assert np.array_equal(gr.astype(np.uint64), orow) and np.array_equal(gc, ocol)\n"
        "            assert np.array_equal(gw.view(np.uint32), ow.view(np.uint32))\n"
        "print('merge paths ok')\n"
    ) % (os.path.join(lv.HERE), os.path.dirname(lv.HERE))
    print(f"[case] large-community merge {env}", flush=True)
    r = subprocess.run([sys.executable, "-c", code], env=dict(os.environ, **env),
                       capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "merge paths ok" in r.stdout


def test_narrow_move_kernels_everywhere(lv):
    """Buckets 1-5 take the narrow move kernels (fewer lanes per vertex than
    the reference's T) for classes of at least 2^19 vertices; LV_NARROW_MIN=0
    sends every class there, so graphs with rows in every bucket, and a
    batch, give the oracle's results through them -- in a child process so
    the setting does not leak."""
    import os
    import subprocess
    import sys
    code = (
        "import sys, numpy as np; sys.path.insert(0, %r); sys.path.insert(0, %r)\n"
        "from loader import load_package; import oracle\n"
        "lv = load_package()\n"
        "qb = lambda q: int(np.float32(q).view(np.uint32))\n"
        "graphs = [lv.gen_sbm(20000, 40, 20.0, 10.0, 61), lv.gen_fem(814, 25, 9, 62),\n"
        "          lv.gen_chung_lu(60000, 9.4, 2.5, 420, 300, 20.0, 63),\n"
        "          lv.gen_rmat(14, 16, 0.57, 0.19, 0.19, 0.5, 1.5, 64)]\n"
        "for r, c, w in graphs:\n"
        "    o = oracle.Oracle(r.astype(np.uint64), c, w); ores = o.run(1e-6)\n"
        "    with lv.Graph(r, c, w) as g:\n"
        "        res = g.run(1e-6); a = g.assignment()\n"
        "    assert (qb(res.q_final), res.passes, res.total_iters) == \\\n"
        "        (qb(ores.q_final), ores.passes, ores.total_iters)\n"
        "    assert np.array_equal(a, o.init_comm())\n"
        "small = [lv.gen_sbm(n, 8, 18.0, 6.0, 70 + i) for i, n in enumerate([300, 2000, 700, 5000])]\n"
        "with lv.Batch(small) as b:\n"
        "    per, tot = b.run(1e-6); assigns = b.assignments()\n"
        "for i, (r, c, w) in enumerate(small):\n"
        "    o = oracle.Oracle(r.astype(np.uint64), c, w); ores = o.run(1e-6)\n"
        "    assert (qb(per[i].q_final), per[i].passes) == (qb(ores.q_final), ores.passes)\n"
        "    assert np.array_equal(assigns[i], o.init_comm())\n"
        "print('narrow kernels ok')\n"
    ) % (os.path.join(lv.HERE), os.path.dirname(lv.HERE))
    print("[case] narrow move kernels for every class", flush=True)
    r = subprocess.run([sys.executable, "-c", code], env=dict(os.environ, LV_NARROW_MIN="0"),
                       capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "narrow kernels ok" in r.stdout


@pytest.mark.parametrize("fused,pipe", [("0", "1"), ("1", "0")])
def test_state_copies_either_way(lv, fused, pipe):
    """Passes of small single graphs sweep with two copies of the community
    state (a step's moves applied inside its kernels, the next step catching
    up on them); larger graphs and batches keep one copy and an apply kernel
    per step.  LV_FUSED forces one way for everything: the state after every
    bucket step, full runs with heavy rows (R-MAT-18, heavy threshold 512)
    and a batch give the oracle's results either way -- in a child process
    so the setting does not leak.  LV_PIPELINE crosses it: small graphs'
    pass loops keep the next captured launch queued by default, large ones
    do not; each copy mode runs here with the other choice."""
    import os
    import subprocess
    import sys
    code = (
        "import sys, numpy as np; sys.path.insert(0, %r); sys.path.insert(0, %r)\n"
        "from loader import load_package; import oracle\n"
        "lv = load_package()\n"
        "qb = lambda q: int(np.float32(q).view(np.uint32))\n"
        "r, c, w = lv.gen_chung_lu(20000, 9.4, 2.5, 420, 100, 20.0, 81)\n"
        "o = oracle.Oracle(r.astype(np.uint64), c, w)\n"
        "with lv.Graph(r, c, w) as g:\n"
        "    assert qb(g.pass_init()) == qb(o.pass_init())\n"
        "    for it in range(3):\n"
        "        for b in range(1, 8):\n"
        "            assert g.step(b) == o.step(b), (it, b)\n"
        "            k, comm, sz, a = g.state()\n"
        "            assert np.array_equal(comm, o.comm()) and np.array_equal(sz, o.comm_sz()), (it, b)\n"
        "            assert np.array_equal(a, o.a_fixed()), (it, b)\n"
        "        assert qb(g.modularity()) == qb(o.modularity()), it\n"
        "graphs = [(lv.gen_rmat(18, 16, 0.57, 0.19, 0.19, 0.5, 1.5, 82), {'heavy_min_degree': 512}),\n"
        "          (lv.gen_fem(814, 25, 9, 83), {}), (lv.gen_sbm(20000, 40, 20.0, 10.0, 84), {})]\n"
        "for (r, c, w), tune in graphs:\n"
        "    o = oracle.Oracle(r.astype(np.uint64), c, w); ores = o.run(1e-6)\n"
        "    with lv.Graph(r, c, w) as g:\n"
        "        g.set_tuning(**tune); res = g.run(1e-6); a = g.assignment()\n"
        "    assert (qb(res.q_final), res.passes, res.total_iters) == \\\n"
        "        (qb(ores.q_final), ores.passes, ores.total_iters)\n"
        "    assert np.array_equal(a, o.init_comm())\n"
        "small = [lv.gen_sbm(n, 8, 18.0, 6.0, 90 + i) for i, n in enumerate([300, 2000, 700, 5000])]\n"
        "with lv.Batch(small) as b:\n"
        "    per, tot = b.run(1e-6); assigns = b.assignments()\n"
        "for i, (r, c, w) in enumerate(small):\n"
        "    o = oracle.Oracle(r.astype(np.uint64), c, w); ores = o.run(1e-6)\n"
        "    assert (qb(per[i].q_final), per[i].passes) == (qb(ores.q_final), ores.passes)\n"
        "    assert np.array_equal(assigns[i], o.init_comm())\n"
        "print('state copies ok')\n"
    ) % (os.path.join(lv.HERE), os.path.dirname(lv.HERE))
    print(f"[case] state copies LV_FUSED={fused} LV_PIPELINE={pipe}", flush=True)
    r = subprocess.run([sys.executable, "-c", code],
                       env=dict(os.environ, LV_FUSED=fused, LV_PIPELINE=pipe),
                       capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "state copies ok" in r.stdout


def _server_pool(lv):
    return [lv.gen_sbm(3000, 12, 14.0, 10.0, 31), lv.gen_sbm(20000, 40, 20.0, 10.0, 32),
            lv.gen_rmat(12, 16, 0.57, 0.19, 0.19, 0.5, 1.5, 33),
            lv.gen_rmat(14, 16, 0.57, 0.19, 0.19, 0.5, 1.5, 34), lv.gen_fem(300, 25, 9, 35),
            lv.gen_chung_lu(8000, 9.4, 2.5, 300, 50, 20.0, 36)]


@pytest.mark.parametrize("slots,budget", [(4, 200 << 30), (3, None)])
def test_server_results_match_oracle(lv, orc, slots, budget):
    """Requests admitted concurrently (and, with a tight budget, one at a
    time) give exactly the oracle's result for their graph."""
    pool = _server_pool(lv)
    expect = []
    for r, c, w in pool:
        o = orc.Oracle(r.astype(np.uint64), c, w)
        ores = o.run(1e-6)
        expect.append((np.float32(ores.q_final).view(np.uint32), ores.passes, ores.n_final,
                       ores.total_iters, ores.sum_edges, o.init_comm()))
    if budget is None:  # room for the largest graph only
        budget = max(lv.lib().lv_graph_footprint(r.size - 1, c.size) for r, c, _ in pool)
    print(f"[case] server slots={slots} budget={budget}", flush=True)
    order = [i % len(pool) for i in range(3 * len(pool))]
    with lv.Server(slots=slots, hbm_budget=budget, keep_assignments=True) as s:
        for rid, gi in enumerate(order):
            s.submit(rid, *pool[gi])
        s.drain()
        out = s.poll(4096)
        assert len(out) == len(order)
        for res in out:
            gi = order[res.id]
            assert res.status == 0
            q, passes, nf, iters, edges, assign = expect[gi]
            assert np.float32(res.q_final).view(np.uint32) == q
            assert (res.passes, res.n_final, res.total_iters, res.sum_edges) == \
                (passes, nf, iters, edges)
            assert res.t_submit_ms <= res.t_start_ms <= res.t_end_ms
            got = s.assignment(res.id, pool[gi][0].size - 1)
            assert np.array_equal(got, assign)


@pytest.mark.parametrize("tune", [
    dict(heavy_min_degree=319, bucket_entries=64),
    dict(heavy_min_degree=1024, bucket_entries=1024),
    dict(heavy_min_degree=2048, bucket_entries=256, locality=1),
    dict(heavy_min_degree=4096, bucket_entries=512, locality=0),
    dict(heavy_min_degree=8192, bucket_entries=512, locality=1),
    dict(wave_vertices=1, block_threads=512, class8_slots=2048, agg_bucket_entries=256),
    dict(wave_vertices=4, heavy_min_degree=8192, block_threads=512, agg_bucket_entries=512,
         sweeps_per_launch=2),
    dict(sweeps_per_launch=4, locality=1),
])
def test_tuning_never_changes_results(lv, orc, graphs, tune):
    """Launch-shape parameters (which kernel serves a bucket-7 row, heavy
    bucket sizes, locality order) leave every result bit-identical."""
    row, col, w = graphs["rmat16"]
    print(f"[case] tuning {tune}", flush=True)
    o = orc.Oracle(row.astype(np.uint64), col, w)
    ores = o.run(1e-6)
    with lv.Graph(_u32(row), col, w) as g:
        g.set_tuning(**tune)
        assert all(g.get_tuning()[k] == v for k, v in tune.items())
        res = g.run(1e-6)
        assign = g.assignment()
    assert np.float32(res.q_final).view(np.uint32) == np.float32(ores.q_final).view(np.uint32)
    assert res.total_iters == ores.total_iters
    assert np.array_equal(assign, o.init_comm())


def test_tuning_change_between_passes(lv, orc, graphs):
    """A contraction that writes the next pass in its locality order also
    classifies it for the current heavy threshold; a threshold changed
    before the next pass init makes that pass classify the ordered graph
    again (without ordering it twice).  Passes and contractions stay equal
    to the oracle's, the contracted graph is reported in the input
    numbering, and a second pass init in a row changes nothing."""
    row, col, w = graphs["rmat16"]
    print("[case] tuning change between passes", flush=True)
    o = orc.Oracle(row.astype(np.uint64), col, w)
    gain = np.float32(1e-6)
    with lv.Graph(_u32(row), col, w) as g:
        g.set_locality(1)
        for p, heavy in enumerate([4096, 512, 8192, 1024, 4096, 512, 2048]):
            if p == 1:  # a second pass init in a row: same state, same Q
                assert g.pass_init().view(np.uint32) == g.pass_init().view(np.uint32)
            g.set_tuning(heavy_min_degree=heavy)
            imp, qb, qa, it = g.modularity_optimization(gain)
            oimp, oqp, oit = o.modularity_optimization(gain)
            oqb, oqa = o.last_pass_q()
            assert (imp, it) == (oimp, oit), f"pass {p}"
            assert qa.view(np.uint32) == oqa.view(np.uint32), f"pass {p}"
            _cmp_state(g, o, f"end of pass {p}")
            if not imp:
                break
            g.aggregation()
            o.aggregation()
            _cmp_csr(g, o, f"after aggregation {p}")
            assert np.array_equal(g.assignment(), o.init_comm()), f"assignment pass {p}"


def test_autotune(lv, orc, graphs):
    """The autotuner times candidate launch shapes on the resident graph and
    keeps the fastest; a run afterwards still gives the oracle's result, and
    out-of-range tunings are refused."""
    row, col, w = graphs["rmat16"]
    o = orc.Oracle(row.astype(np.uint64), col, w)
    ores = o.run(1e-6)
    with lv.Graph(_u32(row), col, w) as g:
        with pytest.raises(lv.LvError):
            g.set_tuning(heavy_min_degree=100)  # would pull bucket-6 rows into step 7
        with pytest.raises(lv.LvError):
            g.set_tuning(bucket_entries=300)
        best, ms = g.autotune(1e-6)
        print(f"[case] autotune best={best} ms={ms:.2f}", flush=True)
        assert ms > 0 and g.get_tuning() == best
        res = g.run(1e-6)
        assign = g.assignment()
    assert np.float32(res.q_final).view(np.uint32) == np.float32(ores.q_final).view(np.uint32)
    assert np.array_equal(assign, o.init_comm())


def _hip():
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipStreamCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
    hip.hipStreamDestroy.argtypes = [ctypes.c_void_p]
    return hip


def test_caller_stream(lv, orc, graphs):
    """Calls on a graph enqueue on a stream the caller owns
    (lv_graph_set_stream), pass by pass and as a whole run, with the
    oracle's results; the graph's own stream is restored with NULL."""
    import ctypes
    hip = _hip()
    st = ctypes.c_void_p()
    assert hip.hipStreamCreate(ctypes.byref(st)) == 0
    try:
        row, col, w = graphs["rmat14_unit"]
        o = orc.Oracle(row.astype(np.uint64), col, w)
        print("[case] caller stream rmat14_unit", flush=True)
        with lv.Graph(_u32(row), col, w) as g:
            g.set_stream(st.value)
            for p in range(20):
                imp, qb, qa, it = g.modularity_optimization(1e-6)
                oimp, _, oit = o.modularity_optimization(1e-6)
                assert (imp, it) == (oimp, oit), p
                assert (qb, qa) == o.last_pass_q(), p
                if not imp:
                    break
                g.aggregation()
                o.aggregation()
            assert np.array_equal(g.assignment(), o.init_comm())
            g.reset()
            r = g.run(1e-6)
            a = g.assignment()
            g.set_stream(None)
            g.reset()
            r2 = g.run(1e-6)
            assert np.array_equal(g.assignment(), a)
        assert np.array_equal(a, o.init_comm())
        assert np.float32(r.q_final).view(np.uint32) == np.float32(r2.q_final).view(np.uint32)
    finally:
        hip.hipStreamDestroy(st)


@pytest.mark.parametrize("k", [2, 4])
def test_batch_sweeps_per_launch(lv, orc, bcsstk16_text, k):
    """Several sweeps captured to one launch: each graph's stop rule is
    applied on the device after every sweep, and the sweeps captured after
    the last graph stopped return at entry; every graph still gets the
    oracle's result."""
    graphs = _mixed_batch(lv, orc, bcsstk16_text)
    print(f"[case] batch of {len(graphs)} graphs, {k} sweeps per launch", flush=True)
    with lv.Batch(graphs) as b:
        b.set_tuning(sweeps_per_launch=k)
        per, _ = b.run(1e-6)
        assigns = b.assignments()
    for i, (r, c, w) in enumerate(graphs):
        o = orc.Oracle(r.astype(np.uint64), c, w)
        try:
            ores = o.run(1e-6)
        except orc.OracleError:
            assert per[i].status != 0, i
            continue
        assert np.float32(per[i].q_final).view(np.uint32) == \
            np.float32(ores.q_final).view(np.uint32), i
        assert (per[i].passes, per[i].total_iters) == (ores.passes, ores.total_iters), i
        assert np.array_equal(assigns[i], o.init_comm()), i


def test_batch_reports_graph_out_of_range(lv, orc):
    """One fixed-point scale serves a batch.  A graph whose total does not fit
    at the batch's scale (weights of 2^22 beside a graph that needs 2^-40) is
    reported with LV_ERR_RANGE and left out; every other graph still gets the
    oracle's result."""
    r0, c0, w0 = lv.gen_sbm(2000, 12, 10.0, 10.0, 41)
    w0 = (w0 * np.float32(2.0 ** -40)).astype(np.float32)  # needs S = 40
    r1, c1, w1 = lv.gen_sbm(1500, 10, 10.0, 10.0, 42)
    w1 = np.full_like(w1, 2.0 ** 22)                       # 2^62 each at S = 40
    r2, c2, w2 = lv.gen_fem(200, 25, 9, 43)
    graphs = [(r0, c0, w0), (r1, c1, w1), (r2, c2, w2)]
    print("[case] batch with an out-of-range graph", flush=True)
    with lv.Batch(graphs) as b:
        per, _ = b.run(1e-6)
        assigns = b.assignments()
    assert per[1].status == 3
    for i in (0, 2):
        r, c, w = graphs[i]
        o = orc.Oracle(r.astype(np.uint64), c, w)
        ores = o.run(1e-6)
        assert per[i].status == 0
        assert np.float32(per[i].q_final).view(np.uint32) == \
            np.float32(ores.q_final).view(np.uint32), i
        assert np.array_equal(assigns[i], o.init_comm()), i
    # alone, the refused graph runs (its own scale is 0)
    orc.Oracle(r1.astype(np.uint64), c1, w1).run(1e-6)
