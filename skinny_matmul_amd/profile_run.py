"""Run one tall-skinny product repeatedly, for profilers (rocprofv3).

    python -m skinny_matmul_amd.profile_run --op tsmm --M 64 --logK 26 --reps 3
    python -m skinny_matmul_amd.profile_run --op copy --logK 26 --reps 3
    python -m skinny_matmul_amd.profile_run --op ozaki --M 128 --logK 24 --S 3 [--gram]
    python -m skinny_matmul_amd.profile_run --op hr --M 64 --logK 22

``--op copy`` runs the HBM copy microbenchmark over arrays of the same byte
size as A (the calibration kernel for FETCH_SIZE / WRITE_SIZE counters).
Inputs come from the device counter RNG; nothing is checked here.
"""
from __future__ import annotations

import argparse
import sys


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--op", choices=("tsmm", "tsmttsm", "ozaki", "hr", "copy"), required=True)
    ap.add_argument("--S", type=int, default=4, help="Ozaki limb count")
    ap.add_argument("--M", type=int, default=64)
    ap.add_argument("--logK", type=int, default=26)
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--gram", action="store_true", help="TSMTTSM/Ozaki/HR: B is A")
    a = ap.parse_args()

    import torch

    import skinny_matmul_amd as sk

    K = 1 << a.logK
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
    Cw = B  # TSMM output (K x M)
    torch.cuda.synchronize()
    if a.gram:
        B = A
    for _ in range(a.reps):
        if a.op == "tsmm":
            ok = sk.dtsmm(0, M, M, K, A, M, 1.0, Bs, M, 0.0, Cw, M)
        elif a.op == "ozaki":
            ok = sk.dtsmttsm_ozaki(0, M, M, K, A, M, 1.0, B, M, 0.0, Cs, M, slices=a.S)
        elif a.op == "hr":
            ok = sk.htsmttsm(0, M, M, K, A, M, 1.0, B, M, 0.0, Cs, M)
        else:
            ok = sk.dtsmttsm(0, M, M, K, A, M, 1.0, B, M, 0.0, Cs, M)
        assert ok
    torch.cuda.synchronize()
    print(f"{a.op} M=N={M} K=2^{a.logK}: {a.reps} calls done")
    return 0


if __name__ == "__main__":
    sys.exit(main())
