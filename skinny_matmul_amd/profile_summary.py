"""Summarise a rocprofv3 kernel trace of the headline benchmark per shape.

    python -m skinny_matmul_amd.profile_summary <prefix>_kernel_trace.csv \
        [--steps 3] [--warmup 1] [--logK 26]

The trace comes from

    rocprofv3 --kernel-trace --stats --output-format csv -d <dir> -o bench -- \
        python3 bench.py --steps 3 --warmup 1 --no-cpu --headline-only

bench.py autotunes first (many candidate launches), then runs `warmup +
steps` sweeps of the 14 calls TSMTTSM M=N=1, TSMM M=N=1, TSMTTSM M=N=2, ...
TSMM M=N=64.  The last sweeps are taken from the end of the trace: every
product kernel dispatch is one call, and a TSMTTSM call may be followed by
its partial-sum kernel (tsmttsm_reduce_kernel).  Prints a markdown table of
average durations per call and GB/s of algorithmic bytes (perf.cu:209-231
accounting: (M + N) K 8 bytes for either product).
"""
from __future__ import annotations

import argparse
import csv
import sys

WIDTHS = (1, 2, 4, 8, 16, 32, 64)
REDUCE = "tsmttsm_reduce_kernel"


def short(name: str) -> str:
    name = name.replace("void skinny::", "")
    return name.split("(")[0]


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("trace")
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--logK", type=int, default=26)
    a = ap.parse_args()

    with open(a.trace, newline="") as f:
        rows = [r for r in csv.DictReader(f) if r.get("Kind", "KERNEL_DISPATCH") == "KERNEL_DISPATCH"]
    rows.sort(key=lambda r: int(r["Start_Timestamp"]))
    disp = [(short(r["Kernel_Name"]), int(r["End_Timestamp"]) - int(r["Start_Timestamp"]))
            for r in rows
            if "tsmttsm" in r["Kernel_Name"] or "tsmm" in r["Kernel_Name"]]

    ops = [(op, M) for M in WIDTHS for op in ("tsmttsm", "tsmm")]
    ncalls = (a.steps + a.warmup) * len(ops)
    # walk back from the end: product dispatches are calls, reduce dispatches
    # belong to the call before them
    calls = []
    pending = 0
    for name, ns in reversed(disp):
        if len(calls) == ncalls:
            break
        if name.startswith(REDUCE):
            pending += ns
            continue
        calls.append((name, ns, pending))
        pending = 0
    if len(calls) < ncalls:
        print(f"only {len(calls)} product dispatches, need {ncalls}", file=sys.stderr)
        return 1
    calls.reverse()
    calls = calls[a.warmup * len(ops):]  # the timed sweeps

    K = 1 << a.logK
    print("| call | kernel | calls | avg ms (product) | avg ms (partial sums) | "
          "avg ms (both) | algorithmic bytes | GB/s |")
    print("|---|---|---|---|---|---|---|---|")
    total_ns = 0
    total_bytes = 0
    for i, (op, M) in enumerate(ops):
        mine = calls[i::len(ops)]
        names = sorted({c[0] for c in mine})
        prod = sum(c[1] for c in mine) / len(mine)
        red = sum(c[2] for c in mine) / len(mine)
        nbytes = (M + M) * K * 8
        total_ns += (prod + red) * len(mine)
        total_bytes += nbytes * len(mine)
        print(f"| {op} M=N={M} | `{'`, `'.join(names)}` | {len(mine)} | {prod / 1e6:.3f} | "
              f"{red / 1e6:.4f} | {(prod + red) / 1e6:.3f} | {nbytes} | "
              f"{nbytes / (prod + red):.0f} |")
    print(f"\nAll timed calls: {total_bytes} B in {total_ns / 1e6:.3f} ms of kernel time, "
          f"{total_bytes / total_ns:.0f} GB/s.")
    return 0


if __name__ == "__main__":
    sys.exit(main())
