"""lv_run_noalloc (every device buffer carved from the library's fixed
scratch array) against lv_run (the same run on a handle with its own
buffers) and the oracle, through the C ABI: two calls in a row, the second
on a smaller graph, then a graph too large for the array (LV_ERR_LIMIT), a
call on a caller-owned stream, and a call after the refusal."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _qb(q):
    return int(np.float32(q).view(np.uint32))


def _check(lv, orc, graph, tag, stream=None):
    row, col, w = graph
    print(f"[case] noalloc {tag}: n={row.size - 1} ne={col.size}", flush=True)
    r, a = lv.run_noalloc(row, col, w, 1e-6, stream=stream)
    with lv.Graph(row, col, w) as g:
        ref = g.run(1e-6)
        ref_a = g.assignment()
    assert (r.status, _qb(r.q_final), r.passes, r.total_iters, r.n_final, r.sum_edges) == \
        (0, _qb(ref.q_final), ref.passes, ref.total_iters, ref.n_final, ref.sum_edges), tag
    assert np.array_equal(a, ref_a), tag
    o = orc.Oracle(row.astype(np.uint64), col, w)
    ores = o.run(1e-6)
    assert _qb(r.q_final) == _qb(ores.q_final), tag
    assert np.array_equal(a, o.init_comm()), tag


def _with_hub(row, col, w, leaves, seed):
    """The graph plus edges of weight 1 from vertex 0 to `leaves` others."""
    n = row.size - 1
    src = np.repeat(np.arange(n, dtype=np.int64), np.diff(row.astype(np.int64)))
    dst = np.random.default_rng(seed).choice(np.arange(1, n), leaves, replace=False)
    s2 = np.concatenate([src, np.zeros(leaves, np.int64), dst])
    d2 = np.concatenate([col.astype(np.int64), dst, np.zeros(leaves, np.int64)])
    w2 = np.concatenate([w, np.ones(2 * leaves, np.float32)])
    order = np.lexsort((d2, s2))
    s2, d2, w2 = s2[order], d2[order], w2[order]
    r = np.zeros(n + 1, np.int64)
    np.add.at(r, s2 + 1, 1)
    return np.cumsum(r).astype(np.uint32), d2.astype(np.uint32), w2.astype(np.float32)


def test_noalloc_matches_lv_run(lv, orc):
    if lv.device_count() < 1:
        pytest.fail("no GPU visible: GPU tests must run on a GPU box")
    big = lv.gen_rmat(14, 16, 0.57, 0.19, 0.19, 0.5, 1.5, 51)
    # a hub of 6000 entries takes the split heavy-row path (rows over 4096)
    small = _with_hub(*lv.gen_sbm(20000, 20, 10.0, 10.0, 52), 6000, 5)
    assert np.diff(small[0].astype(np.int64)).max() > 4096
    assert small[1].size < big[1].size
    _check(lv, orc, big, "rmat14")
    _check(lv, orc, small, "sbm20k with a hub")
    # far beyond the array: refused before any device work
    huge = lv.gen_rmat(18, 16, 0.57, 0.19, 0.19, 0.5, 1.5, 53)
    with pytest.raises(lv.LvError) as e:
        lv.run_noalloc(*huge)
    assert e.value.rc == 10
    # a caller-owned stream, after the refusal
    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipStreamCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
    hip.hipStreamDestroy.argtypes = [ctypes.c_void_p]
    st = ctypes.c_void_p()
    assert hip.hipStreamCreate(ctypes.byref(st)) == 0
    try:
        _check(lv, orc, lv.gen_fem(814, 25, 9, 54), "fem on a caller stream", stream=st.value)
    finally:
        hip.hipStreamDestroy(st)
