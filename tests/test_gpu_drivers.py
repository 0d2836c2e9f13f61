"""The C++ drivers on the GPU (the reference's test, perf, numeric_tests and
svqb programs): exact parity against rocBLAS for every type and both ops,
the cancellation known-answer test, a perf sweep into BenchDB and the SVQB
precision table.  Each driver is one child process using the GPU at a time.
"""
import os
import sqlite3
import subprocess

import pytest

from gpu_util import note

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "skinny_matmul_amd", "bin")


def _run(args, timeout=240):
    note("driver: " + " ".join(args))
    r = subprocess.run([os.path.join(BIN, args[0]), *args[1:]], capture_output=True,
                       text=True, timeout=timeout)
    return r


@pytest.mark.parametrize("op,types,rng", [
    ("tsmttsm", "DR", ["1", "9"]), ("tsmttsm", "DR", ["15", "17", "30", "33"]),
    ("tsmttsm", "FR", ["1", "6"]), ("tsmttsm", "DC", ["1", "5"]),
    ("tsmttsm", "FC", ["2", "3"]), ("tsmttsm", "HR", ["1", "4"]),
    ("tsmm", "DR", ["1", "9"]), ("tsmm", "DR", ["60", "66"]),
    ("tsmm", "FR", ["1", "5"]), ("tsmm", "DC", ["1", "4"]), ("tsmm", "FC", ["7", "8"]),
])
def test_skinny_test_exact(gpu, op, types, rng):
    r = _run(["skinny_test", op, types, *rng, "--samples", "3", "--mib", "32"])
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-2000:]
    assert "Failed" not in r.stdout and "Passed" in r.stdout


def test_numeric_kat(gpu):
    r = _run(["skinny_numeric_tests"], timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "PseudoQuad result is within 4 ulp" in r.stdout


def test_perf_into_benchdb(gpu, tmp_path):
    db = str(tmp_path / "b.db")
    r = _run(["skinny_perf", "tsmttsm", "DR", "4", "5", "--mib", "256", "--db", db,
              "--tune", "--usr1", "check=1"], timeout=400)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    con = sqlite3.connect(db)
    rows = con.execute("SELECT M, N, name, inplace, zerobeta, bw, time FROM benchmarks "
                       "WHERE usr1_name = 'check'").fetchall()
    names = {(m, n, name, inpl) for m, n, name, inpl, zb, bw, t in rows}
    # both variants (library FP64 and Ozaki), A*B and Gram, both widths
    for M in (4, 5):
        for name in ("SKINNY", "OZAKI3"):
            for inpl in (0, 1):
                assert (M, M, name, inpl) in names
    assert all(bw > 0 and t > 0 for *_, bw, t in rows)
    assert con.execute("SELECT COUNT(*) FROM tuning").fetchone()[0] == 1


def test_svqb_precision_table(gpu):
    r = _run(["skinny_svqb", "--K", "2000000", "--reps", "1", "--min-pert", "0.01"],
             timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    rows = [l for l in r.stdout.splitlines() if l.startswith("| ") and "pert" not in l]
    assert len(rows) >= 8
    pert, fp32, fp64, dd, oz3, dev = [float(x) for x in rows[0].strip("| ").split(" | ")]
    # pert = 4: FP64-or-better Gram matrices orthogonalise to near double
    # precision.  The +-1e5 spikes dominate two Gram entries and make columns
    # 0 and 1 nearly collinear after scaling, which amplifies the 3-limb
    # Ozaki Gram's error (about 2^-21 of sum |a||b|, plus the coarse scale of
    # the spikes' own 64-row chunk) to 1e-4: it sits between FP32 and FP64
    assert max(fp64, dd, dev) < 1e-8 and oz3 < 1e-3 and oz3 < fp32, rows[0]
