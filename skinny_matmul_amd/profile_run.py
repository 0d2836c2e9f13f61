"""Run one tall-skinny product repeatedly, for profilers (rocprofv3).

    python -m skinny_matmul_amd.profile_run --op tsmm --M 64 --logK 26 --reps 3
    python -m skinny_matmul_amd.profile_run --op copy --logK 26 --reps 3
    python -m skinny_matmul_amd.profile_run --op ozaki --M 128 --logK 24 --S 3 [--gram]
    python -m skinny_matmul_amd.profile_run --op hr --M 64 --logK 22
    python -m skinny_matmul_amd.profile_run --op headline --logK 26 --tune-only --tune-out T
    python -m skinny_matmul_amd.profile_run --op headline --logK 26 --tune-in T --marks m.tsv
    python -m skinny_matmul_amd.profile_run --op dc --logK 26 --marks m.tsv

``--op copy`` runs the HBM copy microbenchmark over arrays of the same byte
size as A (the calibration kernel for FETCH_SIZE / WRITE_SIZE counters).

``--op headline`` runs the fp64 sweep of bench.py once (TSMTTSM then TSMM at
M = N = 1, 2, ..., 64, same inputs and calls) and ``--op dc`` the complex-
double TSMM of the same widths (out of place, bench.py's C2). Both first run
every call once unmarked (so one-time workspace growth stays out of the
marked calls), then the copy microbenchmark over 1 GiB arrays (counter
calibration), then each call again between two one-element fills that mark
it in a counter trace: fill_uniform (fill_kernel<0>) before, fill_int
(fill_kernel<2>) after. ``--marks`` writes one line per marked call: its
key, the op, M and its algorithmic bytes (profile_traffic turns the counter
runs into HBM bytes per call). The autotuner's choices can be saved in a run
without a profiler (``--tune-only --tune-out``) and loaded under one
(``--tune-in``), so the counted kernels are the tuned ones and tuning is not
profiled.
Inputs come from the device counter RNG; nothing is checked here.
"""
from __future__ import annotations

import argparse
import sys


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--op", choices=("tsmm", "tsmttsm", "ozaki", "hr", "copy", "headline", "dc"),
                    required=True)
    ap.add_argument("--S", type=int, default=4, help="Ozaki limb count")
    ap.add_argument("--M", type=int, default=64)
    ap.add_argument("--N", type=int, help="TSMM: width of C (default M)")
    ap.add_argument("--logK", type=int, default=26)
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--gram", action="store_true", help="TSMTTSM/Ozaki/HR: B is A")
    ap.add_argument("--marks", help="headline/dc: write the marked calls here")
    ap.add_argument("--tune-in", help="headline: load the autotuner table from this file")
    ap.add_argument("--tune-out", help="headline: autotune each shape, save the table here")
    ap.add_argument("--tune-only", action="store_true", help="headline: stop after tuning")
    a = ap.parse_args()

    import torch

    import skinny_matmul_amd as sk

    K = 1 << a.logK
    if a.op in ("headline", "dc"):
        return marked_sweep(sk, torch, a, K)
    M = a.M
    if a.op == "copy":
        gbs = sk.mb_stream("copy", K * M * 8, a.reps)
        print(f"copy {K * M * 8} B per array: {gbs:.1f} GB/s")
        return 0
    A = torch.empty(K * M, dtype=torch.float64, device="cuda")
    B = torch.empty(K * M, dtype=torch.float64, device="cuda")
    sk.fill_uniform(A, A.numel(), 1, 0)
    sk.fill_uniform(B, B.numel(), 1, 1)
    Bs = torch.empty(M * M, dtype=torch.float64, device="cuda")
    sk.fill_uniform(Bs, M * M, 1, 2)
    Cs = torch.zeros(M * M, dtype=torch.float64, device="cuda")
    N = a.N or M
    if a.op == "tsmm" and N != M:
        Bs = torch.empty(M * N, dtype=torch.float64, device="cuda")
        sk.fill_uniform(Bs, M * N, 1, 2)
        B = torch.empty(K * N, dtype=torch.float64, device="cuda")
    Cw = B  # TSMM output (K x N)
    torch.cuda.synchronize()
    if a.gram:
        B = A
    for _ in range(a.reps):
        if a.op == "tsmm":
            ok = sk.dtsmm(0, M, N, K, A, M, 1.0, Bs, M, 0.0, Cw, N)
        elif a.op == "ozaki":
            ok = sk.dtsmttsm_ozaki(0, M, M, K, A, M, 1.0, B, M, 0.0, Cs, M, slices=a.S)
        elif a.op == "hr":
            ok = sk.htsmttsm(0, M, M, K, A, M, 1.0, B, M, 0.0, Cs, M)
        else:
            ok = sk.dtsmttsm(0, M, M, K, A, M, 1.0, B, M, 0.0, Cs, M)
        assert ok
    torch.cuda.synchronize()
    print(f"{a.op} M={M} N={N} K=2^{a.logK}: {a.reps} calls done")
    return 0


WIDTHS = (1, 2, 4, 8, 16, 32, 64)


def marked_sweep(sk, torch, a, K: int) -> int:
    """bench.py's headline sweep (--op headline) or its C2 complex-double
    TSMM (--op dc), one call per shape, each between two marker fills."""
    dev = torch.device("cuda:0")
    Mmax = max(WIDTHS)
    calls = []
    if a.op == "headline":
        A = torch.empty(K * Mmax, dtype=torch.float64, device=dev)
        B = torch.empty(K * Mmax, dtype=torch.float64, device=dev)
        Cs = torch.zeros(Mmax * Mmax, dtype=torch.float64, device=dev)
        sk.fill_uniform(A, A.numel(), 1, 0)
        sk.fill_uniform(B, B.numel(), 1, 1)
        Bsm = {}
        for M in WIDTHS:
            Bsm[M] = torch.empty(M * M, dtype=torch.float64, device=dev)
            sk.fill_uniform(Bsm[M], M * M, 1, 2 + M)
        for M in WIDTHS:
            calls.append((f"tsmttsm_f64_M{M}_K2^{a.logK}", "tsmttsm", M, 2 * M * K * 8,
                          lambda M=M: sk.dtsmttsm(0, M, M, K, A, M, 1.0, B, M, 0.0, Cs, M)))
            calls.append((f"tsmm_f64_M{M}_K2^{a.logK}", "tsmm", M, (2 * M * K + M * M) * 8,
                          lambda M=M: sk.dtsmm(0, M, M, K, A, M, 1.0, Bsm[M], M, 0.0, B, M)))
        if a.tune_in:
            with open(a.tune_in) as f:
                sk.tune_import(f.read())
        elif a.tune_out:
            for _, op, M, _, _ in calls:
                sk.autotune(op, M, M, K, reps=3)
            with open(a.tune_out, "w") as f:
                f.write(sk.tune_export())
            if a.tune_only:
                print(f"tuned {len(calls)} shapes into {a.tune_out}")
                return 0
    else:
        A = torch.empty(2 * K * Mmax, dtype=torch.float64, device=dev)
        Cc = torch.empty(2 * K * Mmax, dtype=torch.float64, device=dev)
        sk.fill_uniform(A, A.numel(), 3, 0)
        Bc = {}
        for M in WIDTHS:
            Bc[M] = torch.empty(2 * M * M, dtype=torch.float64, device=dev)
            sk.fill_uniform(Bc[M], Bc[M].numel(), 3, 1)
            calls.append((f"tsmm_c128_M{M}_K2^{a.logK}", "tsmm", M, 2 * (2 * M * K + M * M) * 8,
                          lambda M=M: sk.ztsmm(0, M, M, K, A, M, 1.0, Bc[M], M, 0.0, Cc, M)))
    mark = torch.empty(1, dtype=torch.float64, device=dev)
    for _, _, _, _, fn in calls:
        assert fn()  # unmarked: the library's workspace reaches its size first
    torch.cuda.synchronize()
    gbs = sk.mb_stream("copy", 1 << 30, 1)
    print(f"calibration copy over 1 GiB arrays: {gbs:.1f} GB/s")
    lines = []
    for key, op, M, nbytes, fn in calls:
        sk.fill_uniform(mark, 1, 7, 0)
        assert fn(), key
        sk.fill_int(mark, 1, 7, 0)
        lines.append(f"{key}\t{op}\t{M}\t{nbytes}\n")
    torch.cuda.synchronize()
    if a.marks:
        with open(a.marks, "w") as f:
            f.writelines(lines)
    print(f"{a.op} K=2^{a.logK}: {len(calls)} marked calls done")
    return 0


if __name__ == "__main__":
    sys.exit(main())
