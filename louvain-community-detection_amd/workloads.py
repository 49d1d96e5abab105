"""Seeded synthetic workloads: the five configurations the engine is measured
and checked on.

Every graph here is produced by the engine's own host generators
(``lv_gen_*`` in liblouvain_hip.so, counter-based PRNGs with fixed seeds), so
the same call gives byte-identical CSRs on any host.  ``bench.py`` times these
workloads, the committed result digests under ``tests/golden/`` are computed
from them, and the GPU tests check the engine against those digests.

    c1  FEM-band analogue of bcsstk16 (4,884 vertices, ~284k entries)
    c2  co-purchase-like Chung-Lu graph over a planted partition (262,111 vertices)
    c3  R-MAT scale 24, edge factor 16, w ~ U(0.5, 1.5) (16.8M vertices, 520.7M entries)
    c4  4,096 stochastic-block-model graphs in one segmented CSR
    c5  Poisson request stream over a pool of SBM and R-MAT graphs
"""
import hashlib

import numpy as np

MIN_GAIN = 1e-6

CONFIGS = {
    "c1": dict(workload="fem-band-4884", gen="fem", args=(814, 25, 9, 1)),
    "c2": dict(workload="copurchase-chung-lu-262111", gen="chung_lu",
               args=(262111, 9.4, 2.5, 420, 2000, 20.0, 2)),
    "c3": dict(workload="rmat-24-ef16", gen="rmat", args=(24, 16, 0.57, 0.19, 0.19, 0.5, 1.5, 3)),
    # (graphs, V min, V max, blocks min, blocks max, mean degree min, max, p_in/p_out, seed)
    "c4": dict(workload="batch-4096-sbm", gen="batch",
               args=(4096, 100, 50000, 8, 64, 8.0, 32.0, 10.0, 4)),
    # (requests, mean arrival rate per second, seed)
    "c5": dict(workload="serve-poisson-200rps", gen="serve", args=(400, 200.0, 5)),
}


def batch_specs(G, vmin, vmax, bmin, bmax, dmin, dmax, p_ratio, seed):
    """Per-graph SBM parameters: V log-uniform, blocks and mean degree uniform."""
    rng = np.random.default_rng(seed)
    n = np.exp(rng.uniform(np.log(vmin), np.log(vmax), G)).astype(np.int64)
    blocks = rng.integers(bmin, bmax + 1, G)
    deg = rng.uniform(dmin, dmax, G)
    return [(int(n[i]), int(min(blocks[i], n[i])), float(deg[i]), p_ratio, seed * 100003 + i)
            for i in range(G)]


def make_batch(lv, args, workers=16):
    from concurrent.futures import ThreadPoolExecutor
    specs = batch_specs(*args)
    with ThreadPoolExecutor(max_workers=workers) as ex:  # generators release the GIL
        return list(ex.map(lambda sp: lv.gen_sbm(*sp), specs))


def make_graph(lv, cfg):
    """The graph (row, col, w) of c1-c3, or the list of graphs of c4."""
    if cfg["gen"] == "batch":
        return make_batch(lv, cfg["args"])
    return getattr(lv, "gen_" + cfg["gen"])(*cfg["args"])


def sym_weights(row, col, lo, hi, seed):
    """w ~ U(lo, hi) per undirected pair, equal on both directions."""
    n = row.size - 1
    src = np.repeat(np.arange(n, dtype=np.uint64), np.diff(row.astype(np.int64)))
    a = np.minimum(src, col.astype(np.uint64))
    b = np.maximum(src, col.astype(np.uint64))
    x = (a * np.uint64(0x9E3779B97F4A7C15) + b * np.uint64(0xBF58476D1CE4E5B9) +
         np.uint64(seed)) & np.uint64(0xFFFFFFFFFFFFFFFF)
    x ^= x >> np.uint64(31)
    x = (x * np.uint64(0x94D049BB133111EB)) & np.uint64(0xFFFFFFFFFFFFFFFF)
    x ^= x >> np.uint64(29)
    u = (x >> np.uint64(11)).astype(np.float64) / float(1 << 53)
    return (np.float32(lo) + np.float32(hi - lo) * u.astype(np.float32)).astype(np.float32)


def serve_pool_specs(seed):
    """Generator calls of the serving pool: 12 SBM graphs with 1k-200k
    vertices, then R-MAT scale 14-21 (edge factor 16), all with
    w ~ U(0.5, 1.5).  Each item is (kind, generator args, weight seed)."""
    rng = np.random.default_rng(seed)
    items = []
    for i in range(12):
        n = int(np.exp(rng.uniform(np.log(1000), np.log(200000))))
        items.append(("sbm", (n, int(rng.integers(8, 65)), float(rng.uniform(8, 32)), 10.0,
                              seed * 1000 + i), 7 + i))
    for sc in range(14, 22):
        items.append(("rmat", (sc, 16, 0.57, 0.19, 0.19, 0.5, 1.5, seed * 1000 + 100 + sc), None))
    return items


def serve_graph(lv, item):
    kind, args, wseed = item
    if kind == "rmat":
        return lv.gen_rmat(*args)
    r, c, _ = lv.gen_sbm(*args)
    return r, c, sym_weights(r, c, 0.5, 1.5, wseed)


def serve_pool(lv, seed):
    """The serving pool as (sbm, rmat), lists of (row, col, w)."""
    from concurrent.futures import ThreadPoolExecutor
    items = serve_pool_specs(seed)
    with ThreadPoolExecutor(max_workers=8) as ex:  # generators release the GIL
        graphs = list(ex.map(lambda it: serve_graph(lv, it), items))
    n_sbm = sum(1 for it in items if it[0] == "sbm")
    return graphs[:n_sbm], graphs[n_sbm:]


def serve_stream(nreq, rate, seed, n_sbm, n_rmat):
    """Arrival times (s) and pool indices of the Poisson request stream."""
    rng = np.random.default_rng(seed + 1)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, nreq))
    kinds = rng.integers(0, 2, nreq)
    picks = [int(rng.integers(0, n_sbm)) if k == 0 else n_sbm + int(rng.integers(0, n_rmat))
             for k in kinds]
    return arrivals, picks


def sha256_arrays(*arrays):
    """Hex SHA-256 over the raw little-endian bytes of the arrays, in order."""
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def csr_digest(row, col, w):
    """Digest of a CSR as the engine reports it: u32 row offsets, u32 columns,
    f32 weights."""
    return sha256_arrays(np.asarray(row, dtype=np.uint32), np.asarray(col, dtype=np.uint32),
                         np.asarray(w, dtype=np.float32))


def q_bits(q):
    return int(np.float32(q).view(np.uint32))
