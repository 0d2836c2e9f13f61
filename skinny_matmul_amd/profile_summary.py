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
accounting: (M + N) K 8 bytes for either product, plus M N 8 for TSMM).

With --svqb the summary covers the SVQB measurement instead: the hipGraph
replays of two sweeps over 64 blocks (K = 2^22, M = 8, 16, 32, 64 in turn),
per replay (one eigensolve launch per sweep).  The trace is of

    python3 bench.py --svqb-only --no-fused

or of the whole benchmark, where the library's kernels between the batched
launch and the first fused launch are taken.
"""
from __future__ import annotations

import argparse
import csv
import sys

WIDTHS = (1, 2, 4, 8, 16, 32, 64)
REDUCE = "tsmttsm_reduce_kernel"


def short(name: str) -> str:
    for pre in ("void skinny::", "skinny::"):
        if name.startswith(pre):
            name = name[len(pre):]
    return name.split("(")[0]


def load_trace(path: str):
    """Kernel dispatches in start order, from a rocprofv3 kernel-trace CSV or
    from its default SQLite output (`<prefix>_results.db`, `kernels` view).
    A run with a cooperative launch can end the profiler before it writes the
    CSV files; the database is written as the run goes."""
    if path.endswith(".db"):
        import sqlite3
        con = sqlite3.connect(path)
        rows = [{"Kernel_Name": n, "Start_Timestamp": s0, "End_Timestamp": e0}
                for n, s0, e0 in con.execute("select name, start, end from kernels")]
        con.close()
    else:
        with open(path, newline="") as f:
            rows = [r for r in csv.DictReader(f)
                    if r.get("Kind", "KERNEL_DISPATCH") == "KERNEL_DISPATCH"]
    rows.sort(key=lambda r: int(r["Start_Timestamp"]))
    return rows


SVQB_KERNELS = ("tsmttsm_kernel", "tsmttsm_reduce_kernel", "tsmm_", "svqb_")


def svqb_summary(rows) -> int:
    names = [short(r["Kernel_Name"]) for r in rows]
    last_batch = max((i for i, n in enumerate(names) if n.startswith("batch_kernel")),
                     default=-1)
    first_fused = next((i for i, n in enumerate(names)
                        if n.startswith("svqb_fused_kernel")), len(names))
    # bench.py marks the end of the input construction (whose Gram and TSMM
    # calls would otherwise count as replays) with a one-element
    # fill_kernel<2> just before the replays
    first_eig = next((i for i, n in enumerate(names) if n.startswith("svqb_eig_kernel")),
                     len(names))
    start = max([last_batch] + [i for i in range(first_eig)
                                if names[i].startswith("fill_kernel<2>")])
    win = [(names[i], int(rows[i]["End_Timestamp"]) - int(rows[i]["Start_Timestamp"]))
           for i in range(start + 1, first_fused)
           if names[i].startswith(SVQB_KERNELS) and "htsmttsm" not in names[i]]
    neig = sum(1 for n, _ in win if n.startswith("svqb_eig_kernel"))
    if neig < 2 or neig % 2:
        print(f"found {neig} eigensolve launches, expected two per replay", file=sys.stderr)
        return 1
    reps = neig // 2
    agg = {}
    for n, ns in win:
        c, t = agg.get(n, (0, 0))
        agg[n] = (c + 1, t + ns)
    print(f"{reps} replays of the two-sweep graph; per replay:\n")
    print("| kernel | launches | ms | share |")
    print("|---|---|---|---|")
    total = sum(t for _, t in agg.values()) / reps
    for n, (c, t) in sorted(agg.items(), key=lambda kv: -kv[1][1]):
        print(f"| `{n}` | {c // reps} | {t / reps / 1e6:.3f} | {t / reps / total:.1%} |")
    nbytes = sum(2 * 3 * M * (1 << 22) * 8 for M in (8, 16, 32, 64) for _ in range(16))
    print(f"\nKernel time per replay: {total / 1e6:.3f} ms; {nbytes} algorithmic bytes "
          f"(Gram K M 8 + update 2 K M 8 per block and sweep): {nbytes / total:.0f} GB/s.")
    return 0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("trace")
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--logK", type=int, default=26)
    ap.add_argument("--svqb", action="store_true")
    a = ap.parse_args()

    rows = load_trace(a.trace)
    if a.svqb:
        return svqb_summary(rows)
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
        # perf.cu:209-231: TSMM also writes C once per call (M N 8 bytes)
        nbytes = (M + M) * K * 8 + (M * M * 8 if op == "tsmm" else 0)
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
