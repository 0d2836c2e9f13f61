"""The C++ drivers: built, usage errors handled, and BenchDB on the host.

`skinny_benchdb_check` writes rows through BenchDB (the perf driver's
sqlite store) and replaces one by key; the file is read back here with
Python's sqlite3.  The GPU runs of the drivers are in test_gpu_drivers.py.
"""
import os
import sqlite3
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "skinny_matmul_amd", "bin")
DRIVERS = ("skinny_test", "skinny_perf", "skinny_numeric_tests", "skinny_svqb",
           "skinny_benchdb_check")


@pytest.mark.parametrize("name", DRIVERS)
def test_driver_built(name):
    assert os.access(os.path.join(BIN, name), os.X_OK), f"{name} is not built"


def test_benchdb_roundtrip(tmp_path):
    db = str(tmp_path / "bench.db")
    r = subprocess.run([os.path.join(BIN, "skinny_benchdb_check"), db],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    con = sqlite3.connect(db)
    rows = con.execute("SELECT M, name, blockCount, K, time, bw FROM benchmarks "
                       "ORDER BY M").fetchall()
    assert rows == [(1, "SKINNY", 512, 1 << 26, 0.001, 100.0),
                    (2, "SKINNY", 1024, 1 << 26, 0.5, 4321.0),
                    (3, "SKINNY", 512, 1 << 26, 0.003, 300.0)]
    cols = [c[1] for c in con.execute("PRAGMA table_info(benchmarks)")]
    for c in ("multype", "device", "types", "M", "N", "name", "inplace", "zerobeta",
              "branch", "usr1_name", "usr1_val", "blockCount", "K", "time", "flops",
              "bw", "l2bw", "sharedbw", "occupancy", "drambw"):
        assert c in cols
    assert con.execute("SELECT table_text FROM tuning").fetchall() == [("0 0 8 8 26 1 512\n",)]


@pytest.mark.parametrize("name", ("skinny_test", "skinny_perf"))
def test_usage_errors_without_gpu(name):
    # argument errors are reported before any device call
    r = subprocess.run([os.path.join(BIN, name), "bogus"], capture_output=True,
                       text=True, timeout=60)
    assert r.returncode == 2 and "usage" in r.stderr


@pytest.mark.parametrize("types", ("FR", "DR", "FC", "DC", "HR"))
def test_cpp_header_type_macros(tmp_path, types):
    """The drop-in header compiles and links with each type macro defined,
    as the reference's drivers are built (-DDR etc., types.hpp:9-59), and
    the macro-selected registries are the per-type ones."""
    src = tmp_path / "use.cpp"
    src.write_text(
        '#include "skinny/skinny.hpp"\n'
        "#include <cstdio>\n"
        "int main() {\n"
        "  MatmulFunctionType f = getEnabledTSMTTSMVersions()[0].first;\n"
        "  auto a = getEnabledTSMTTSMVersions();\n"
        "  auto b = getEnabledTSMMVersions();\n"
        "  std::printf(\"%s %d %zu %zu %s\\n\", types.c_str(), flopsPerCell, a.size(),\n"
        "              b.size(), a[0].second.c_str());\n"
        "  return f ? 0 : 1;\n"
        "}\n")
    exe = str(tmp_path / "use")
    r = subprocess.run(["g++", "-std=c++17", f"-D{types}", "-I", os.path.join(ROOT, "include"),
                        str(src), "-L", os.path.join(ROOT, "skinny_matmul_amd", "lib"),
                        "-lskinny", "-o", exe], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
