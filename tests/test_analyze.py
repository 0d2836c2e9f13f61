"""BenchDB result analysis (skinny_matmul_amd.analyze) on the host.

The grids, efficiencies and speed-ups are checked on hand-made `benchmarks`
tables whose expected values follow from the definitions in
benchmark_analyzer.py (makeGrid :38-53, rooflinePlot :84-88, speedupPlot
:160-169): a planted wrong axis order, a dropped time filter or a ratio
taken the wrong way round each fails one assertion.  The perf driver's
real table layout comes from `skinny_benchdb_check` (BenchDB itself).
"""
import json
import math
import os
import sqlite3
import subprocess

import numpy as np
import pytest

from skinny_matmul_amd import analyze as an

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "skinny_matmul_amd", "bin")

SCHEMA = ("CREATE TABLE benchmarks (multype TEXT, device TEXT, types TEXT, M INTEGER, "
          "N INTEGER, name TEXT, inplace INTEGER, zerobeta INTEGER, branch TEXT, "
          "usr1_name TEXT, usr1_val TEXT, blockCount INTEGER, K INTEGER, time REAL, "
          "flops REAL, bw REAL, l2bw REAL, sharedbw REAL, occupancy REAL, drambw REAL)")


def _db(path, rows):
    con = sqlite3.connect(path)
    con.execute(SCHEMA)
    for r in rows:
        d = dict(multype="TSMM", device="gfx950", types="DR", name="SKINNY", inplace=0,
                 zerobeta=1, branch="main", usr1_name="", usr1_val="", blockCount=0,
                 K=1 << 20, time=0.01, flops=0.0, bw=0.0, l2bw=0, sharedbw=0, occupancy=0,
                 drambw=0)
        d.update(r)
        con.execute("INSERT INTO benchmarks VALUES (" + ",".join("?" * 20) + ")",
                    [d[k] for k in ("multype", "device", "types", "M", "N", "name", "inplace",
                                    "zerobeta", "branch", "usr1_name", "usr1_val",
                                    "blockCount", "K", "time", "flops", "bw", "l2bw",
                                    "sharedbw", "occupancy", "drambw")])
    con.commit()
    con.close()


def test_grid_axes_and_filters(tmp_path):
    p = str(tmp_path / "a.db")
    _db(p, [dict(M=3, N=1, bw=100.0), dict(M=1, N=2, bw=200.0),
            dict(M=2, N=2, bw=300.0, time=5e-5),          # too short: 0
            dict(M=2, N=1, bw=float("nan")),               # not finite: 0
            dict(M=1, N=1, bw=None)])                      # missing: 0
    conns = an.connect([p])
    g = an.fetch_grid(conns, "TSMM", "gfx950", "DR", "SKINNY", 0, 1, "bw")
    assert g.shape == (3, 4)              # (max N + 1, max M + 1)
    assert g[1][3] == 100.0 and g[2][1] == 200.0   # grid[N][M]
    assert g[2][2] == 0.0 and g[1][2] == 0.0 and g[1][1] == 0.0
    assert np.count_nonzero(g) == 2
    # another configuration key selects nothing
    assert an.fetch_grid(conns, "TSMM", "gfx950", "DR", "SKINNY", 1, 1, "bw").size == 0


def test_first_database_with_rows_wins(tmp_path):
    p1, p2 = str(tmp_path / "1.db"), str(tmp_path / "2.db")
    _db(p1, [dict(M=1, N=1, bw=1.0, name="OTHER")])
    _db(p2, [dict(M=1, N=1, bw=2.0), dict(M=2, N=2, bw=3.0)])
    conns = an.connect([p1, p2])
    g = an.fetch_grid(conns, "TSMM", "gfx950", "DR", "SKINNY", 0, 1, "bw")
    assert g[1][1] == 2.0 and g[2][2] == 3.0


def test_roofline_fraction_closed_form():
    # TSMTTSM at M = N, fp64: intensity M/8 flop/B (2 M N K flop over
    # (M + N) K 8 bytes); at stream 5000 GB/s and peak 40000 GFlop/s the ridge
    # is at 8 flop/B, M = 64: below it the bandwidth term wins, above the flops
    bw_c, fp_c = 5000.0, 40000.0
    bw = np.zeros((129, 129))
    fl = np.zeros((129, 129))
    for m in (8, 64, 128):
        t = 1.0  # a kernel at exactly the roofline
        intensity = m / 8.0
        rate = min(bw_c * intensity, fp_c)
        fl[m][m] = rate
        bw[m][m] = rate / intensity
    eff = an.roofline_grid(bw, fl, bw_c, fp_c)
    for m in (8, 64, 128):
        assert math.isclose(eff[m][m], 1.0)
    assert math.isclose(bw[128][128] / bw_c, 0.5)      # compute-bound: half the stream
    assert math.isclose(fl[8][8] / fp_c, 0.125)        # bandwidth-bound: 1/8 of peak


def test_speedup_log2_and_mask():
    g1 = np.array([[0.0, 4.0], [8.0, 0.1]])
    g2 = np.array([[1.0, 2.0], [16.0, 5.0]])
    s = an.speedup_grid(g1, g2)
    assert s[0][1] == 1.0          # 2x faster
    assert s[1][0] == -1.0         # 2x slower
    assert s[0][0] == 0.0 and s[1][1] == 0.0   # unmeasured entries count as equal


def test_ceilings_by_operation():
    c = {"read_GBs": 7000.0, "copy_GBs": 5500.0, "fp64_fma": 59.0, "fp64_mfma": 47.0}
    assert an.ceilings_for("TSMTTSM", "DR", c) == (7000.0, 59000.0)
    assert an.ceilings_for("TSMM", "DC", c) == (5500.0, 59000.0)
    assert an.ceilings_for("TSMM", "FR", c) == (5500.0, 157300.0)  # datasheet FP32
    # the 4x4x4 FP64 form and the measured FP32 matrix-core peaks take part
    c.update({"fp64_mfma_4x4x4": 67.0, "fp32_mfma": 126.0, "fp32_mfma_4x4x1": 131.0})
    assert an.ceilings_for("TSMTTSM", "HR", c) == (7000.0, 67000.0)
    assert an.ceilings_for("TSMTTSM", "FC", c) == (7000.0, 131000.0)


def test_cli_writes_maps_and_index(tmp_path):
    p = str(tmp_path / "b.db")
    rows = []
    for m in range(1, 9):
        for n in range(1, 9):
            rows.append(dict(multype="TSMTTSM", M=m, N=n, bw=1000.0 * (m + n),
                             flops=10.0 * m * n))
            rows.append(dict(multype="TSMTTSM", M=m, N=n, name="OZAKI3",
                             bw=500.0 * (m + n), flops=5.0 * m * n))
    _db(p, rows)
    ceil = tmp_path / "c.json"
    ceil.write_text(json.dumps({"read_GBs": 7000.0, "copy_GBs": 5500.0, "fp64_fma": 59.0,
                                "fp64_mfma": 47.0}))
    out = tmp_path / "plots"
    assert an.main([p, "--out", str(out), "--ceilings", str(ceil), "--format", "png",
                    "--vs", "SKINNY", "OZAKI3"]) == 0
    idx = (out / "index.md").read_text()
    assert "| TSMTTSM | gfx950 | DR | SKINNY | 0 | 1 | 7000 | 59000 |" in idx
    # best efficiency: bw 16000 at M = N = 8 over 7000 GB/s
    assert f"{16000.0 / 7000.0:.3f}" in idx
    assert "geometric-mean speed-up 2.000x" in idx
    pngs = sorted(f for f in os.listdir(out) if f.endswith(".png"))
    assert any(f.startswith("roofline_TSMTTSM") for f in pngs)
    assert any(f.startswith("speedup_") for f in pngs)
    for f in pngs:
        assert (out / f).read_bytes()[:8] == b"\x89PNG\r\n\x1a\n"


def test_reads_the_perf_driver_table(tmp_path):
    exe = os.path.join(BIN, "skinny_benchdb_check")
    if not os.access(exe, os.X_OK):
        pytest.fail("skinny_benchdb_check is not built")
    db = str(tmp_path / "bench.db")
    r = subprocess.run([exe, db], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    conns = an.connect([db])
    cfgs = an.configurations(conns)
    assert cfgs
    g = an.fetch_grid(conns, *cfgs[0], "bw")
    assert g.size > 0 and np.max(g) > 0
