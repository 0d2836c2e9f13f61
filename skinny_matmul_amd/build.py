"""Build libskinny.so (gfx950) and the C++ drivers in-tree with hipcc.

    python -m skinny_matmul_amd.build [--jobs N] [--verbose] [--no-drivers]

Each HIP translation unit under csrc/ is compiled to an object with hipcc for
gfx950 and the objects are linked into skinny_matmul_amd/lib/libskinny.so.
Packed-FP32 instruction selection is disabled for all device code.

The drivers under drivers/ (host C++ over the C ABI) are linked against it
into skinny_matmul_amd/bin/: skinny_test (exact parity against rocBLAS),
skinny_perf (throughput sweeps into BenchDB), skinny_numeric_tests (the
cancellation known-answer test), skinny_svqb (Gram precision vs SVQB
orthogonality) and skinny_benchdb_check (BenchDB on the host only).
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import subprocess
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIBDIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIBDIR, "libskinny.so")
DRIVERS = os.path.join(HERE, "drivers")
BINDIR = os.path.join(HERE, "bin")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("SKINNY_ARCH", "gfx950")

SOURCES = [
    "runtime.hip",
    "tsmttsm.hip",
    "htsmttsm.hip",
    "tsmm.hip",
    "tsmm_lds.hip",
    "tsmm_q4.hip",
    "tsmm_wide.hip",
    "capi.hip",
    "microbench.hip",
    "svqb.hip",
    "svqb_fused.hip",
    "batch.hip",
    "tune.hip",
    "ozaki.hip",
    "noalloc.hip",
]

CFLAGS = [
    "-O3",
    "-std=c++17",
    "-fPIC",
    f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
    "-Wno-unused-variable",
    "-I", os.path.join(ROOT, "include"),
    "-I", CSRC,
]


def _compile(src: str, verbose: bool) -> str:
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC, *CFLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
    if verbose:
        print(" ".join(cmd), flush=True)
    t0 = time.time()
    r = subprocess.run(cmd, capture_output=True, text=True)
    # hipcc's host pass does not know the packed-fp32 feature: harmless
    err = "\n".join(l for l in r.stderr.splitlines()
                    if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{err}")
    if verbose and err.strip():
        print(err, file=sys.stderr)
    if verbose:
        print(f"  {src}: {time.time() - t0:.1f}s", flush=True)
    return obj


def build(jobs: int = 8, verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        objs = list(ex.map(lambda s: _compile(s, verbose), SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, LIB)
    return LIB


# driver -> (source, extra link libraries); all link libskinny
DRIVER_TARGETS = {
    "skinny_test": ("test.cpp", ["-lrocblas"]),
    "skinny_perf": ("perf.cpp", ["-l:libsqlite3.so.0", "-lrocblas"]),
    "skinny_numeric_tests": ("numeric_tests.cpp", []),
    "skinny_svqb": ("svqb.cpp", []),
}


def _git_branch() -> str:
    try:
        r = subprocess.run(["git", "-C", ROOT, "rev-parse", "--abbrev-ref", "HEAD"],
                           capture_output=True, text=True, timeout=10)
        b = r.stdout.strip()
        return b if r.returncode == 0 and b and b.isprintable() and '"' not in b else "main"
    except (OSError, subprocess.SubprocessError):
        return "main"


def build_drivers(jobs: int = 8, verbose: bool = False) -> list:
    """Compile the host drivers (needs libskinny.so built first)."""
    os.makedirs(BINDIR, exist_ok=True)
    branch = _git_branch()

    def one(item):
        name, (src, libs) = item
        out = os.path.join(BINDIR, name)
        cmd = [HIPCC, "-O2", "-std=c++17", f"-DSKINNY_BRANCH=\"{branch}\"",
               "-I", os.path.join(ROOT, "include"), "-I", DRIVERS,
               os.path.join(DRIVERS, src), "-L", LIBDIR, "-lskinny", *libs,
               "-Wl,-rpath,$ORIGIN/../lib", "-o", out + ".tmp"]
        if verbose:
            print(" ".join(cmd), flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"driver {name} failed:\n{r.stderr}")
        os.replace(out + ".tmp", out)
        return out

    outs = []
    with cf.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        outs = list(ex.map(one, DRIVER_TARGETS.items()))
    # BenchDB check: host only, plain g++
    out = os.path.join(BINDIR, "skinny_benchdb_check")
    cmd = ["g++", "-O2", "-std=c++17", os.path.join(DRIVERS, "benchdb_check.cpp"),
           "-l:libsqlite3.so.0", "-o", out + ".tmp"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"benchdb_check failed:\n{r.stderr}")
    os.replace(out + ".tmp", out)
    outs.append(out)
    return outs


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=8)
    ap.add_argument("--verbose", action="store_true")
    ap.add_argument("--no-drivers", action="store_true")
    a = ap.parse_args()
    t0 = time.time()
    out = build(a.jobs, a.verbose)
    print(f"built {out} in {time.time() - t0:.1f}s")
    if not a.no_drivers:
        t1 = time.time()
        bins = build_drivers(a.jobs, a.verbose)
        print(f"built {len(bins)} drivers in {time.time() - t1:.1f}s")


if __name__ == "__main__":
    main()
