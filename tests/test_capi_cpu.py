"""CPU-side checks of the product library and CLI (no GPU compute).

The C ABI library loads and exports every symbol include/louvain_hip.h
declares; the product .mtx reader agrees with the oracle's; argument
validation and the reference CLI's error behaviour (utils.cu:231-267,
:171-174) hold without touching a GPU.
"""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "louvain_hip.h")


def header_symbols():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(lv_[a-z0-9_]+)\s*\(", text)))


def test_exports_every_declared_symbol(lv):
    L = lv.lib()
    syms = header_symbols()
    assert len(syms) >= 20
    missing = [s for s in syms if not hasattr(L, s)]
    assert not missing, missing
    assert sorted(lv.EXPORTS) == syms


def test_status_strings(lv):
    L = lv.lib()
    assert L.lv_status_string(1) == b"Error opening file"
    assert L.lv_status_string(0) == b"ok"


def test_mtx_reader_matches_oracle(lv, orc, bcsstk16_path, tmp_path):
    row, col, w = lv.read_mtx(bcsstk16_path)
    orow, ocol, ow = orc.parse_mtx(path=bcsstk16_path)
    assert np.array_equal(row.astype(np.uint64), orow)
    assert np.array_equal(col, ocol)
    assert np.array_equal(w.view(np.uint32), ow.view(np.uint32))
    cases = ["%%MatrixMarket x\n% c\n4 4 5\n1 1 -2.5\n2 1 0.5\n3 2\n4 4 3\n1 3 -1\n",
             "3 3 2\n0 1 1\n1 2 2\n", "2 2 1\n1 2\n", "5 5 3\n1 2 1e-3\n2 3 0.1\n5 5 7\n"]
    for i, text in enumerate(cases):
        p = tmp_path / f"c{i}.mtx"
        p.write_text(text)
        a = lv.read_mtx(str(p))
        b = orc.parse_mtx(text=text)
        assert np.array_equal(a[0].astype(np.uint64), b[0])
        assert np.array_equal(a[1], b[1])
        assert np.array_equal(a[2].view(np.uint32), b[2].view(np.uint32))


def _big_mtx_text(n, lines, seed, zero_line=None):
    """A multi-megabyte .mtx body exercising every parser rule across the
    worker splits: lines without a weight (carried), unparseable weights
    (-> 0), negative weights (abs), self-loops, CRLF endings, blank padding,
    and trailing lines beyond nnz."""
    rng = np.random.default_rng(seed)
    out = ["%%MatrixMarket matrix coordinate real general\n", "% comment\n",
           f"{n} {n} {lines}\n"]
    for i in range(lines):
        a = int(rng.integers(1, n))  # ids valid both 1-based and 0-based
        b = a if rng.random() < 0.02 else int(rng.integers(1, n))
        if zero_line is not None and i == zero_line:
            a = 0
        kind = rng.random()
        if kind < 0.25:
            wtxt = ""  # carried from the previous line
        elif kind < 0.27:
            wtxt = " nan?" if rng.random() < 0.5 else " x"  # does not parse -> 0
        else:
            wtxt = f" {rng.normal() * 3:.7g}"
        end = "\r\n" if rng.random() < 0.05 else "\n"
        out.append(f"{a}\t{b}{wtxt}{end}" if rng.random() < 0.1 else f"{a} {b}{wtxt}{end}")
    out.append("1 1 5\n")  # beyond nnz: ignored
    out.append("garbage beyond nnz\n")
    return "".join(out)


@pytest.mark.parametrize("zero_line", [None, 150000])
def test_mtx_reader_multithreaded_matches_oracle(lv, orc, tmp_path, zero_line):
    """Bodies of several MB are split over worker threads; the carried weight
    and the 0/1-based decision cross the splits exactly like one sequential
    pass of the reference parser."""
    text = _big_mtx_text(50000, 200000, 7, zero_line)
    assert len(text) > (1 << 21)
    p = tmp_path / "big.mtx"
    p.write_bytes(text.encode())
    a = lv.read_mtx(str(p))
    b = orc.parse_mtx(text=text)
    assert np.array_equal(a[0].astype(np.uint64), b[0])
    assert np.array_equal(a[1], b[1])
    assert np.array_equal(a[2].view(np.uint32), b[2].view(np.uint32))


def test_binary_csr_roundtrip(lv, bcsstk16_path, tmp_path):
    row, col, w = lv.read_mtx(bcsstk16_path)
    p = tmp_path / "g.lvcsr"
    lv.write_csr(str(p), row, col, w)
    r2, c2, w2 = lv.read_mtx(str(p))
    assert np.array_equal(row, r2) and np.array_equal(col, c2)
    assert np.array_equal(w.view(np.uint32), w2.view(np.uint32))
    raw = bytearray(p.read_bytes())
    raw[32 + 4 * (row.size - 1) + 4 * 5] = 0xFF  # a column id far out of range
    raw[32 + 4 * (row.size - 1) + 4 * 5 + 3] = 0x7F
    bad = tmp_path / "bad.lvcsr"
    bad.write_bytes(bytes(raw[:-4]))  # truncated
    with pytest.raises(lv.LvError) as e:
        lv.read_mtx(str(bad))
    assert e.value.rc == 2
    bad.write_bytes(bytes(raw))
    with pytest.raises(lv.LvError) as e:
        lv.read_mtx(str(bad))
    assert e.value.rc == 2


def test_mtx_reader_errors(lv, tmp_path):
    with pytest.raises(lv.LvError) as e:
        lv.read_mtx(str(tmp_path / "missing.mtx"))
    assert e.value.rc == 1
    for i, bad in enumerate(["2 2 2\n1 2 1\n", "2 2 1\n1 3 1\n", "2 2 1\nx y\n"]):
        p = tmp_path / f"b{i}.mtx"
        p.write_text(bad)
        with pytest.raises(lv.LvError) as e:
            lv.read_mtx(str(p))
        assert e.value.rc == 2


def test_graph_create_validates_before_device(lv):
    L = lv.lib()
    h = ctypes.c_void_p()
    # column out of range
    rc = L.lv_graph_create(0, 2, np.array([0, 1, 2], np.uint32), np.array([1, 5], np.uint32),
                           np.array([1, 1], np.float32), ctypes.byref(h))
    assert rc == 4
    # non-monotone row pointer
    rc = L.lv_graph_create(0, 2, np.array([0, 2, 1], np.uint32), np.array([1, 0], np.uint32),
                           np.array([1, 1], np.float32), ctypes.byref(h))
    assert rc == 4
    # non-finite and out-of-range weights: too fine for a shift of 62 bits,
    # or single weights beyond the 2^61 limit of the exact sums (rejected
    # before their conversion to integers)
    for bad in (np.inf, np.nan, 1e-30, 1e30, 2.0 ** 70, 2.0 ** 61):
        rc = L.lv_graph_create(0, 2, np.array([0, 1, 2], np.uint32),
                               np.array([1, 0], np.uint32),
                               np.array([bad, bad], np.float32), ctypes.byref(h))
        assert rc == 3


def test_batch_create_validates_before_device(lv):
    L = lv.lib()
    h = ctypes.c_void_p()
    off = np.array([0, 2, 4], np.uint64)
    row = np.array([0, 1, 2, 3, 4], np.uint32)
    w = np.ones(4, np.float32)
    # entries present but no column or weight array (a second handle on the
    # library whose argument types admit NULL)
    L2 = ctypes.CDLL(lv.LIB_PATH)
    L2.lv_batch_create.argtypes = [ctypes.c_int, ctypes.c_uint32, ctypes.c_void_p,
                                   ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                   ctypes.POINTER(ctypes.c_void_p)]
    assert L2.lv_batch_create(0, 2, off.ctypes.data, row.ctypes.data, None, None,
                              ctypes.byref(h)) == 4
    # an entry of graph 0 pointing into graph 1
    col = np.array([1, 2, 3, 2], np.uint32)
    assert L.lv_batch_create(0, 2, off, row, col, w, ctypes.byref(h)) == 4
    # a negative weight in one graph
    col = np.array([1, 0, 3, 2], np.uint32)
    assert L.lv_batch_create(0, 2, off, row, col, np.array([1, 1, -1, -1], np.float32),
                             ctypes.byref(h)) == 3


def test_no_device_is_an_error(lv):
    if lv.device_count() > 0:
        pytest.skip("a GPU is visible")
    with pytest.raises(lv.LvError) as e:
        lv.Graph(np.array([0, 1, 2], np.uint32), np.array([1, 0], np.uint32),
                 np.array([1, 1], np.float32))
    assert e.value.rc == 8


def test_footprint_bound(lv):
    """lv_graph_footprint: every device buffer of a handle with the on-demand
    arenas at their worst case.  Per CSR entry: two working CSRs (8 B each),
    the heavy-row lists (8), the contraction's sort keys and sums (24), the
    shared global tables at 2 slots of 16 B (32), bucket and lane scratch; the
    pristine input stays on the host."""
    f = lv.lib().lv_graph_footprint
    n, ne = 1 << 20, 1 << 24
    per_entry = (f(n, 2 * ne) - f(n, ne)) / ne
    per_vertex = (f(2 * n, ne) - f(n, ne)) / n
    assert 84 <= per_entry < 85, per_entry
    assert per_vertex < 200, per_vertex
    # the R-MAT scale-24 handle (16.8M vertices, 520.7M entries): under 48 GB
    assert f(1 << 24, 520_749_004) < 48e9
    assert f(100, 0) < 8 << 20  # fixed part: small arrays and slack


def test_generators_deterministic(lv):
    a = lv.gen_fem(814, 25, 9, 1)
    b = lv.gen_fem(814, 25, 9, 1)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
    row, col, w = a
    assert row.size - 1 == 4884
    assert 250_000 < col.size < 320_000
    assert (w > 0.01 - 1e-7).all()
    deg = np.diff(row.astype(np.int64))
    # symmetric, sorted rows, one self-loop per row
    assert (col == np.repeat(np.arange(4884), deg)).sum() == 4884
    r1 = lv.gen_rmat(10, 16, 0.57, 0.19, 0.19, 0.5, 1.5, 3)
    r2 = lv.gen_rmat(10, 16, 0.57, 0.19, 0.19, 0.5, 1.5, 3)
    assert all(np.array_equal(x, y) for x, y in zip(r1, r2))
    row, col, w = r1
    dense = {}
    for v in range(row.size - 1):
        seg = col[row[v]:row[v + 1]]
        assert np.all(np.diff(seg.astype(np.int64)) > 0)
        for j in range(row[v], row[v + 1]):
            dense[(v, int(col[j]))] = w[j]
    assert all(dense[(b_, a_)] == x for (a_, b_), x in dense.items())
    assert ((w >= 0.5) & (w <= 1.5)).all()


def test_copurchase_generator_shape(lv):
    row, col, w = lv.gen_chung_lu(262111, 9.4, 2.5, 420, 2000, 20.0, 2)
    assert row.size - 1 == 262111
    assert 2_200_000 < col.size < 2_600_000   # ~1.23M undirected edges
    assert (w == 1.0).all()


# ------------------------------------------------------------------ CLI ----

def _cli(lv, *args):
    return subprocess.run([lv.CLI_PATH, *args], capture_output=True, text=True, timeout=60)


def test_cli_usage(lv):
    for args in ([], ["-f", "x.mtx"], ["-g", "0.1"], ["-q"]):
        r = _cli(lv, *args)
        assert r.returncode == 1
        assert "Usage: ./gpulouvain -f mtx-matrix-file -g min-gain [-v]" in r.stdout


def test_cli_bad_gain(lv):
    r = _cli(lv, "-f", "x.mtx", "-g", "abc")
    assert r.returncode == 1
    assert r.stdout.startswith("-g option invalid argument: ")
    r = _cli(lv, "-f", "x.mtx", "-g", "1e99")
    assert r.returncode == 1
    assert r.stdout.startswith("-g option error out of range: ")


def test_cli_missing_file(lv, tmp_path):
    r = _cli(lv, "-f", str(tmp_path / "nope.mtx"), "-g", "0.01")
    assert r.returncode == 1
    assert r.stdout == "Error opening file\n"


def test_tuning_defaults(lv):
    t = lv.Tuning()
    assert lv.lib().lv_tuning_default(ctypes.byref(t)) == 0
    assert t.as_dict() == {"heavy_min_degree": 4096, "bucket_entries": 512, "locality": -1,
                           "locality_min_vertices": 1 << 20, "wave_vertices": 2,
                           "block_threads": 1024, "class8_slots": 1280,
                           "agg_bucket_entries": 384, "sweeps_per_launch": 1}
