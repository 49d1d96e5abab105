"""bench.py's host-side parts on the CPU: the multi-GPU refusal, the rocprofv3
summaries (kernel trace and the copy-scaled HBM counters) on small databases
of the profiler's schema, and the roofline traffic read back from a summary."""
import json
import os
import sqlite3
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BENCH = os.path.join(ROOT, "bench.py")


def _bench(*args):
    return subprocess.run([sys.executable, BENCH, *args], capture_output=True, text=True,
                          timeout=120)


def test_more_than_one_gpu_is_refused():
    r = _bench("--gpus", "2")
    assert r.returncode == 2
    assert "one GPU" in json.loads(r.stdout)["error"]


def _kernel_db(path, rows):
    con = sqlite3.connect(path)
    con.execute("create table kernels (id integer, name text, start integer, end integer)")
    con.executemany("insert into kernels values (?, ?, ?, ?)",
                    [(i, n, s, e) for i, (n, s, e) in enumerate(rows)])
    con.commit()
    con.close()


def _pmc_db(path, counter, rows):
    con = sqlite3.connect(path)
    con.execute("create table pmc_events (name text, counter_name text, counter_value real, "
                "dispatch_id integer)")
    con.executemany("insert into pmc_events values (?, ?, ?, ?)",
                    [(n, counter, v, d) for d, (n, v) in enumerate(rows)])
    con.commit()
    con.close()


def test_kernel_summary(tmp_path):
    db = str(tmp_path / "k.db")
    # two launches of one block kernel variant (10 us, 30 us), one of another
    # variant (20 us), one scan (40 us): the family averages 20 us per launch
    _kernel_db(db, [("void lv::k_move_block<128, 1024>(lv::DevView)", 0, 10_000),
                    ("void lv::k_move_block<128, 1024>(lv::DevView)", 50_000, 80_000),
                    ("void lv::k_move_block<1280, 1024>(lv::DevView)", 100_000, 120_000),
                    ("void lv::k_scan_block(unsigned int const*)", 200_000, 240_000)])
    r = _bench("--summarize-kernels", db)
    assert r.returncode == 0, r.stderr
    fam = [ln for ln in r.stdout.splitlines() if ln.startswith("| move_block |")]
    assert fam == ["| move_block | 3 | 0.060 | 20.00 | 60.00 |"]
    assert "| (shared scans/fills/copies) | 1 | 0.040 | 40.00 | 40.00 |" in r.stdout
    assert "kernel time 0.10 ms over 4 dispatches" in r.stdout


def test_pmc_summary_scales_by_the_copy_kernel(tmp_path):
    f, w, cf, cw = (str(tmp_path / f"{x}.db") for x in ("f", "w", "cf", "cw"))
    # counters in KiB per dispatch; the copy kernel of 1 MiB each way reads
    # as 512 KiB fetched (scale 2) and 1024 KiB written (scale 1)
    _pmc_db(f, "FETCH_SIZE", [("k_move_block<128, 1024>(x)", 100.0),
                              ("k_move_block<1280, 1024>(x)", 50.0),
                              ("k_r1_groups<true>(x)", 10.0), ("k_scan_block(x)", 7.0)])
    _pmc_db(w, "WRITE_SIZE", [("k_move_block<128, 1024>(x)", 1.0), ("k_r1_groups<true>(x)", 2.0)])
    _pmc_db(cf, "FETCH_SIZE", [("small(x)", 1.0), ("copy(x)", 512.0)])
    _pmc_db(cw, "WRITE_SIZE", [("small(x)", 1.0), ("copy(x)", 1024.0)])
    r = _bench("--summarize-pmc", f, w, cf, cw, str(1 << 20))
    assert r.returncode == 0, r.stderr
    assert "FETCH x2.0000, WRITE x1.0000" in r.stdout
    block = r.stdout.split("```json\n")[1].split("```")[0]
    d = json.loads(block)["per_run_bytes"]
    assert d["move_block"] == {"fetch": 150 * 1024 * 2.0, "write": 1024.0}
    assert d["r1"] == {"fetch": 10 * 1024 * 2.0, "write": 2 * 1024.0}
    assert set(d) == {"move_block", "r1"}  # the scan is attributed to no family
    # bench reads the same block back for roofline.traffic (bytes per launch)
    sys.path.insert(0, ROOT)
    try:
        import bench
    finally:
        sys.path.pop(0)
    prof = tmp_path / "profiles"
    prof.mkdir()
    (prof / "pmc_toy.md").write_text("# toy\n\n" + r.stdout)
    old = bench.ROOT
    bench.ROOT = str(tmp_path)
    try:
        assert bench.measured_traffic("toy", "move_block", 3) == pytest.approx(
            (150 * 1024 * 2.0 + 1024.0) / 3)
        assert bench.measured_traffic("toy", "apply", 3) is None
        assert bench.measured_traffic("other", "move_block", 3) is None
    finally:
        bench.ROOT = old
