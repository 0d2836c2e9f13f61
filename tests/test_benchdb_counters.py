"""The BenchDB counter merge on the CPU: synthetic rocprofv3 counter files
whose values make every expected column follow from the definitions."""
import csv
import os
import sqlite3

import pytest

CU = 256
KEYS = ("multype", "device", "types", "M", "N", "name", "inplace", "zerobeta", "branch",
        "usr1_name", "usr1_val")


def _write_counters(path, dispatches):
    """dispatches: [(kernel name, grid size, {counter: value})]"""
    with open(path, "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(["Dispatch_Id", "Kernel_Name", "Grid_Size", "Counter_Name", "Counter_Value"])
        for i, (name, grid, ctr) in enumerate(dispatches, 1):
            for k, v in ctr.items():
                # two rows per counter (per-XCD pieces): the tool sums them
                w.writerow([i, name, grid, k, v / 2])
                w.writerow([i, name, grid, k, v / 2])


def _run(counter, copy_val, seg_vals, occ=False):
    fill0 = ("void skinny::fill_kernel<0>(void*, long)", 256, {counter: 1.0})
    fill2 = ("void skinny::fill_kernel<2>(void*, long)", 256, {counter: 1.0})
    copy1 = ("void skinny::(anonymous namespace)::stream_kernel<1, 4, true>(...)", 256 * CU,
             copy_val)
    copy8 = ("void skinny::(anonymous namespace)::stream_kernel<1, 4, false>(...)", 2048 * CU,
             copy_val if not occ else {"SQ_WAVE_CYCLES": 9.0, "SQ_BUSY_CYCLES": 1.0})
    d = [fill0, copy1, copy8]            # data fills and the calibration copies
    for vals in seg_vals:                 # one marked segment per row
        d.append(fill0)
        for v in vals:
            d.append(("void skinny::tsmttsm_kernel<...>", 512 * 256, v))
        d.append(fill2)
    return d


def test_merge_fills_drambw_and_occupancy(tmp_path):
    from skinny_matmul_amd import benchdb_counters as bc
    db = str(tmp_path / "b.db")
    con = sqlite3.connect(db)
    con.execute("CREATE TABLE benchmarks (multype TEXT, device TEXT, types TEXT, M INTEGER, "
                "N INTEGER, name TEXT, inplace INTEGER, zerobeta INTEGER, branch TEXT, "
                "usr1_name TEXT, usr1_val TEXT, blockCount INTEGER, K INTEGER, time REAL, "
                "flops REAL, bw REAL, l2bw REAL, sharedbw REAL, occupancy REAL, drambw REAL)")
    rows = [("TSMTTSM", "gpu x", "DR", 8, 8, "SKINNY", 0, 0, "main", "", "", 0.002),
            ("TSMTTSM", "gpu x", "DR", 8, 8, "ROCBLAS", 0, 0, "main", "", "", 0.004)]
    for r in rows:
        con.execute("INSERT INTO benchmarks (multype, device, types, M, N, name, inplace, "
                    "zerobeta, branch, usr1_name, usr1_val, time, drambw, occupancy) "
                    "VALUES (?,?,?,?,?,?,?,?,?,?,?,?,0,0)", r)
    con.commit()
    con.close()
    marks = tmp_path / "m.tsv"
    marks.write_text("".join("\t".join(str(x) for x in r) + "\n" for r in rows))
    # the copy counts half its bytes in FETCH_SIZE (factor 2) and all of them
    # in WRITE_SIZE (factor 1); sizes in KiB
    gib_kib = (1 << 30) / 1024
    f = tmp_path / "f.csv"
    _write_counters(f, _run("FETCH_SIZE", {"FETCH_SIZE": gib_kib / 2},
                            [[{"FETCH_SIZE": 1000.0}], [{"FETCH_SIZE": 600.0},
                                                        {"FETCH_SIZE": 400.0}]]))
    w = tmp_path / "w.csv"
    _write_counters(w, _run("WRITE_SIZE", {"WRITE_SIZE": gib_kib},
                            [[{"WRITE_SIZE": 10.0}], [{"WRITE_SIZE": 30.0}]]))
    o = tmp_path / "o.csv"
    # copy at one workgroup per CU: ratio 0.5 -> factor 0.25 (0.125 / 0.5)
    _write_counters(o, _run("SQ_WAVE_CYCLES", {"SQ_WAVE_CYCLES": 1.0, "SQ_BUSY_CYCLES": 2.0},
                            [[{"SQ_WAVE_CYCLES": 4.0, "SQ_BUSY_CYCLES": 2.0}],
                             [{"SQ_WAVE_CYCLES": 1.0, "SQ_BUSY_CYCLES": 1.0},
                              {"SQ_WAVE_CYCLES": 1.0, "SQ_BUSY_CYCLES": 3.0}]], occ=True))
    r = bc.merge(db, str(marks), str(f), str(w), str(o), cu_count=CU)
    assert r["fetch_factor"] == pytest.approx(2.0)
    assert r["write_factor"] == pytest.approx(1.0)
    con = sqlite3.connect(db)
    got = {row[0]: row[1:] for row in
           con.execute("SELECT name, drambw, occupancy FROM benchmarks ORDER BY name")}
    con.close()
    # SKINNY: (2 * 1000 + 10) KiB in 2 ms; occupancy 4 / 2 * 0.25
    assert got["SKINNY"][0] == pytest.approx((2 * 1000 + 10) * 1024 / 0.002 / 1e9)
    assert got["SKINNY"][1] == pytest.approx(0.5)
    # ROCBLAS: two dispatches summed: (2 * 1000 + 30) KiB in 4 ms; 2 / 4 * 0.25
    assert got["ROCBLAS"][0] == pytest.approx((2 * 1000 + 30) * 1024 / 0.004 / 1e9)
    assert got["ROCBLAS"][1] == pytest.approx(0.125)


def test_merge_refuses_misaligned_runs(tmp_path):
    from skinny_matmul_amd import benchdb_counters as bc
    marks = tmp_path / "m.tsv"
    marks.write_text("TSMTTSM\tg\tDR\t1\t1\tSKINNY\t0\t0\tmain\t\t\t0.001\n")
    gib_kib = (1 << 30) / 1024
    f = tmp_path / "f.csv"
    _write_counters(f, _run("FETCH_SIZE", {"FETCH_SIZE": gib_kib}, [[{"FETCH_SIZE": 1.0}]] * 2))
    w = tmp_path / "w.csv"
    _write_counters(w, _run("WRITE_SIZE", {"WRITE_SIZE": gib_kib}, [[{"WRITE_SIZE": 1.0}]]))
    with pytest.raises(SystemExit, match="segment counts differ"):
        bc.merge(str(tmp_path / "x.db"), str(marks), str(f), str(w), dry_run=True)


def test_profile_traffic_per_marked_call(tmp_path):
    """profile_traffic: calibrated bytes per marked call, kernels listed,
    entries merged into an existing traffic file."""
    import json
    from skinny_matmul_amd import profile_traffic as pt
    marks = tmp_path / "m.tsv"
    marks.write_text("tsmttsm_f64_M8_K2^20\ttsmttsm\t8\t134217728\n"
                     "tsmm_f64_M8_K2^20\ttsmm\t8\t134218240\n")
    gib_kib = (1 << 30) / 1024
    f = tmp_path / "f.csv"
    _write_counters(f, _run("FETCH_SIZE", {"FETCH_SIZE": gib_kib / 2},
                            [[{"FETCH_SIZE": 65536.0}, {"FETCH_SIZE": 8.0}],
                             [{"FETCH_SIZE": 32768.0}]]))
    w = tmp_path / "w.csv"
    _write_counters(w, _run("WRITE_SIZE", {"WRITE_SIZE": gib_kib},
                            [[{"WRITE_SIZE": 4.0}, {"WRITE_SIZE": 1.0}],
                             [{"WRITE_SIZE": 65536.0}]]))
    ff, wf, rows = pt.summarize(str(marks), str(f), str(w))
    assert ff == pytest.approx(2.0) and wf == pytest.approx(1.0)
    assert rows[0]["hbm"] == pytest.approx((2 * (65536 + 8) + 5) * 1024)
    assert rows[1]["hbm"] == pytest.approx((2 * 32768 + 65536) * 1024)
    assert rows[0]["kernels"] == ["tsmttsm_kernel<...>"] * 2
    js = tmp_path / "t.json"
    js.write_text(json.dumps({"other": {"hbm_bytes_per_launch": 1.0}}))
    import sys
    argv = sys.argv
    sys.argv = ["x", "--marks", str(marks), "--fetch", str(f), "--write", str(w),
                "--json", str(js)]
    try:
        assert pt.main() == 0
    finally:
        sys.argv = argv
    d = json.loads(js.read_text())
    assert set(d) == {"other", "tsmttsm_f64_M8_K2^20", "tsmm_f64_M8_K2^20"}
    assert d["tsmm_f64_M8_K2^20"]["hbm_bytes_per_launch"] == pytest.approx(rows[1]["hbm"])
