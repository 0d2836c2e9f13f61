"""Result analysis over BenchDB: roofline-efficiency, absolute and speed-up
heat maps of the `benchmarks` table that `skinny_perf` fills.

    python -m skinny_matmul_amd.analyze DB [DB ...] --out DIR
        [--ceilings ceilings.json | --stream-bw GBs --peak-fp GFLOPs]
        [--format pdf|png|svg] [--branch NAME] [--vs NAME1 NAME2]

The counterpart of benchmark_analyzer.py (its makeGrid, fetchGrid,
rooflinePlot, absolutePlot and speedupOver/speedupPlot, lines 38-200), over
the same table and columns (perf.cu:286-307):

* a grid holds one value per (N, M): `grid[N][M]`, rows taken from the first
  database that has any for the query; entries with time < 1e-4 s or a
  non-finite or missing value are 0 (makeGrid, benchmark_analyzer.py:38-53);
* roofline efficiency = max(bw / stream_bw, flops / peak_fp) per entry
  (rooflinePlot, :84-88), colour range [0, 1];
* speed-up of variant 1 over variant 2 = log2(grid1 / grid2), with entries
  where either grid is below 0.2 counted as equal (speedupPlot, :160-169);
* absolute maps of bw (GB/s) or flops (GFlop/s).

The ceilings are this chip's measured ones (`python -m
skinny_matmul_amd.ceilings --json FILE`): a TSMTTSM row is a read stream
(read_GBs), a TSMM row reads and writes (copy_GBs); FP64 types use the
best of the FP64 vector and matrix-core (16x16x4 and 4x4x4) peaks, FP32
types the best measured FP32 matrix-core peak (the 157.3 TFLOP/s datasheet
figure when none was measured).  Both can be overridden.

Each configuration (multype, device, types, name, inplace, zerobeta) found
in the databases gets its maps plus a line in DIR/index.md with the
median and best efficiency and the M = N diagonal.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sqlite3
import sys
from typing import Iterable, Sequence

import numpy as np

KEY_COLS = ("multype", "device", "types", "name", "inplace", "zerobeta")
VALUE_NAMES = {"flops": "GFlop/s", "bw": "GB/s"}


def connect(paths: Sequence[str]) -> list:
    conns = []
    for p in paths:
        c = sqlite3.connect(p)
        c.row_factory = sqlite3.Row
        conns.append(c)
    return conns


def _query(conns, where: str, args: tuple) -> list:
    """Rows of the first database with any match (fetchGrid's rule)."""
    for c in conns:
        res = c.execute("SELECT * FROM benchmarks WHERE " + where, args).fetchall()
        if res:
            return res
    return []


def make_grid(rows: Iterable, value: str, max_n: int = -1, max_m: int = -1) -> np.ndarray:
    """grid[N][M] = row[value]; 0 for missing, non-finite, or time < 1e-4 s."""
    rows = list(rows)
    if not rows:
        return np.zeros((0, 0))
    if max_n < 0:
        max_n = max(int(r["N"]) for r in rows)
    if max_m < 0:
        max_m = max(int(r["M"]) for r in rows)
    g = np.zeros((max_n + 1, max_m + 1))
    for r in rows:
        n, m = int(r["N"]), int(r["M"])
        if n > max_n or m > max_m:
            continue
        v = r[value]
        t = r["time"]
        if v is None or t is None or t < 1e-4 or not math.isfinite(float(v)):
            g[n][m] = 0.0
        else:
            g[n][m] = float(v)
    return g


def fetch_grid(conns, multype, device, types, name, inplace, zerobeta, value,
               branch=None, max_n=-1, max_m=-1) -> np.ndarray:
    where = "multype=? AND device=? AND types=? AND name=? AND inplace=? AND zerobeta=?"
    args = (multype, device, types, name, int(inplace), int(zerobeta))
    if branch is not None:
        where += " AND branch=?"
        args += (branch,)
    return make_grid(_query(conns, where, args), value, max_n, max_m)


def roofline_grid(bw: np.ndarray, flops: np.ndarray, stream_bw: float,
                  peak_fp: float) -> np.ndarray:
    """Fraction of the roofline reached: the larger of the bandwidth and the
    flop-rate fractions (bw in GB/s against stream_bw, flops in GFlop/s
    against peak_fp)."""
    return np.maximum(bw / stream_bw, flops / peak_fp)


def speedup_grid(g1: np.ndarray, g2: np.ndarray) -> np.ndarray:
    """log2(g1 / g2) over the common shape; entries where either value is
    below 0.2 (unmeasured) count as equal (0)."""
    n = min(g1.shape[0], g2.shape[0])
    m = min(g1.shape[1], g2.shape[1])
    a = g1[:n, :m].astype(float).copy()
    b = g2[:n, :m].astype(float).copy()
    mask = (a < 0.2) | (b < 0.2)
    a[mask] = 1.0
    b[mask] = 1.0
    return np.log2(a / b)


def configurations(conns, branch=None) -> list:
    """Every distinct (multype, device, types, name, inplace, zerobeta)."""
    seen = []
    for c in conns:
        q = "SELECT DISTINCT " + ", ".join(KEY_COLS) + " FROM benchmarks"
        args = ()
        if branch is not None:
            q += " WHERE branch=?"
            args = (branch,)
        for r in c.execute(q, args).fetchall():
            t = tuple(r)
            if t not in seen:
                seen.append(t)
    return sorted(seen, key=lambda t: tuple(str(x) for x in t))


# the ceilings.py keys whose best value is the compute peak of a type set
FP64_PEAKS = ("fp64_fma", "fp64_mfma", "fp64_mfma_4x4x4")
FP32_PEAKS = ("fp32_mfma", "fp32_mfma_32x32x2", "fp32_mfma_4x4x1")


def ceilings_for(multype: str, types: str, ceil: dict) -> tuple:
    """(stream GB/s, peak GFlop/s) for one configuration from measured
    ceilings (ceilings.py's JSON)."""
    bw = ceil.get("read_GBs" if multype.upper() == "TSMTTSM" else "copy_GBs")
    if types in ("DR", "DC", "HR"):
        tf = max(ceil.get(k, 0.0) for k in FP64_PEAKS)
    else:
        # the measured FP32 matrix-core peaks, else the datasheet figure
        tf = max((ceil.get(k, 0.0) for k in FP32_PEAKS), default=0.0) or 157.3
    return float(bw), float(tf) * 1000.0


# ---------------------------------------------------------------------------
# output
# ---------------------------------------------------------------------------
def _slug(*parts) -> str:
    return "_".join(str(p).replace(" ", "_").replace("/", "-") for p in parts)


def plot_grid(grid: np.ndarray, path: str, label: str, vmin=None, vmax=None,
              cmap="jet", fmt_speedup=False) -> None:
    """A heat map with M on the x axis and N on the y axis (origin lower),
    the layout of the reference's plots."""
    import matplotlib
    matplotlib.use("Agg")
    import matplotlib.pyplot as plt
    from matplotlib import ticker

    fig = plt.figure()
    plt.imshow(grid, interpolation="nearest", origin="lower", vmin=vmin, vmax=vmax,
               cmap=plt.get_cmap(cmap))
    plt.xlabel("M")
    plt.ylabel("N")
    plt.xlim(0.5, max(1.5, grid.shape[1] - 0.5))
    plt.ylim(0.5, max(1.5, grid.shape[0] - 0.5))
    if fmt_speedup:
        def f(x, pos):
            if x == 0:
                return "    = "
            return "{:5.1f}".format(round(float(np.power(2.0, abs(x))), 1)) + r"$\times$"
        plt.colorbar(format=ticker.FuncFormatter(f))
    else:
        plt.colorbar()
    plt.xticks([1] + list(range(1, grid.shape[1]))[7::8])
    plt.yticks([1] + list(range(1, grid.shape[0]))[7::8])
    plt.grid(linestyle="-", alpha=0.1, color="black")
    plt.title(label, fontsize=9)
    fig.savefig(path, bbox_inches="tight")
    plt.close(fig)


def _diag(grid: np.ndarray, widths=(1, 2, 4, 8, 16, 32, 64)) -> str:
    out = []
    for w in widths:
        if w < grid.shape[0] and w < grid.shape[1] and grid[w][w] > 0:
            out.append(f"{w}: {grid[w][w]:.3g}")
    return ", ".join(out)


def analyse(db_paths: Sequence[str], out_dir: str, ceilings: dict | None = None,
            stream_bw: float | None = None, peak_fp: float | None = None,
            fmt: str = "pdf", branch: str | None = None, vs: tuple | None = None,
            plots: bool = True) -> list:
    """Writes the maps and DIR/index.md; returns the index lines."""
    os.makedirs(out_dir, exist_ok=True)
    conns = connect(db_paths)
    lines = ["# BenchDB analysis", "",
             f"databases: {', '.join(os.path.basename(p) for p in db_paths)}", "",
             "| multype | device | types | name | inplace | zerobeta | stream GB/s | "
             "peak GFlop/s | median eff. | best eff. | eff. at M = N | GB/s at M = N | files |",
             "|---|---|---|---|---|---|---|---|---|---|---|---|---|"]
    for cfg in configurations(conns, branch):
        multype, device, types, name, inplace, zerobeta = cfg
        bwg = fetch_grid(conns, *cfg, "bw", branch)
        flg = fetch_grid(conns, *cfg, "flops", branch)
        if bwg.size == 0:
            continue
        sbw, pfp = ceilings_for(multype, types, ceilings or {}) if ceilings else (0.0, 0.0)
        if stream_bw:
            sbw = stream_bw
        if peak_fp:
            pfp = peak_fp
        files = []
        eff = None
        if sbw > 0 and pfp > 0:
            eff = roofline_grid(bwg, flg, sbw, pfp)
        slug = _slug(multype, device, types, name, f"ip{inplace}", f"zb{zerobeta}")
        if plots:
            if eff is not None:
                p = os.path.join(out_dir, f"roofline_{slug}.{fmt}")
                plot_grid(eff, p, f"{multype} {name} {types} roofline fraction", 0, 1)
                files.append(os.path.basename(p))
            for value in ("bw", "flops"):
                g = bwg if value == "bw" else flg
                p = os.path.join(out_dir, f"{value}_{slug}.{fmt}")
                plot_grid(g, p, f"{multype} {name} {types} {VALUE_NAMES[value]}")
                files.append(os.path.basename(p))
        measured = eff[eff > 0] if eff is not None else np.zeros(0)
        med = f"{np.median(measured):.3f}" if measured.size else "-"
        best = f"{np.max(measured):.3f}" if measured.size else "-"
        lines.append(f"| {multype} | {device} | {types} | {name} | {inplace} | {zerobeta} | "
                     f"{sbw:.0f} | {pfp:.0f} | {med} | {best} | "
                     f"{_diag(eff) if eff is not None else '-'} | {_diag(bwg)} | "
                     f"{' '.join(files)} |")
    if vs:
        n1, n2 = vs
        lines += ["", f"## speed-up {n1} over {n2} (bw)", ""]
        for cfg in configurations(conns, branch):
            multype, device, types, name, inplace, zerobeta = cfg
            if name != n1:
                continue
            g1 = fetch_grid(conns, multype, device, types, n1, inplace, zerobeta, "bw", branch)
            g2 = fetch_grid(conns, multype, device, types, n2, inplace, zerobeta, "bw", branch)
            if g1.size == 0 or g2.size == 0:
                continue
            sp = speedup_grid(g1, g2)
            f = os.path.join(out_dir, "speedup_" + _slug(multype, device, types, n1, "vs", n2,
                                                         f"ip{inplace}", f"zb{zerobeta}")
                             + f".{fmt}")
            if plots:
                spread = float(np.max(np.abs(sp))) or 1.0
                plot_grid(sp, f, f"{multype} {types} {n1} vs {n2}", -spread, spread, "bwr",
                          fmt_speedup=True)
            n = min(g1.shape[0], g2.shape[0])
            m = min(g1.shape[1], g2.shape[1])
            both = (g1[:n, :m] >= 0.2) & (g2[:n, :m] >= 0.2)
            gm = 2.0 ** float(np.mean(sp[both])) if both.any() else float("nan")
            lines.append(f"* {multype} {types} inplace={inplace} zerobeta={zerobeta}: "
                         f"geometric-mean speed-up {gm:.3f}x over {int(both.sum())} shapes "
                         f"({os.path.basename(f)})")
    with open(os.path.join(out_dir, "index.md"), "w") as fh:
        fh.write("\n".join(lines) + "\n")
    return lines


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("db", nargs="+")
    ap.add_argument("--out", default="plots")
    ap.add_argument("--ceilings", default=None, help="JSON written by skinny_matmul_amd.ceilings")
    ap.add_argument("--stream-bw", type=float, default=None, help="GB/s")
    ap.add_argument("--peak-fp", type=float, default=None, help="GFlop/s")
    ap.add_argument("--format", default="pdf", choices=("pdf", "png", "svg"))
    ap.add_argument("--branch", default=None)
    ap.add_argument("--vs", nargs=2, default=None, metavar=("NAME1", "NAME2"))
    ap.add_argument("--no-plots", action="store_true")
    a = ap.parse_args(argv)
    ceil = None
    if a.ceilings:
        with open(a.ceilings) as f:
            ceil = json.load(f)
    if ceil is None and not (a.stream_bw and a.peak_fp):
        print("no ceilings: give --ceilings FILE or both --stream-bw and --peak-fp; "
              "efficiency maps are skipped", file=sys.stderr)
    lines = analyse(a.db, a.out, ceil, a.stream_bw, a.peak_fp, a.format, a.branch,
                    tuple(a.vs) if a.vs else None, not a.no_plots)
    print("\n".join(lines))
    return 0


if __name__ == "__main__":
    sys.exit(main())
