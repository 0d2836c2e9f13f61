"""HBM bytes per call from two rocprofv3 counter runs of a marked sweep.

    python -m skinny_matmul_amd.profile_traffic --marks m.tsv \
        --fetch F/run_counter_collection.csv --write W/run_counter_collection.csv \
        [--json profiles/traffic.json]

The runs are of `profile_run --op headline` or `--op dc` (one with
`--pmc FETCH_SIZE`, one with `--pmc WRITE_SIZE`, same arguments).  Each run
holds the copy microbenchmark over 1 GiB arrays, then one marked segment per
line of the marks file (the dispatches between a fill_kernel<0> and the next
fill_kernel<2>).  FETCH_SIZE and WRITE_SIZE (KiB per dispatch) are scaled by
the factors that make the copy's dispatches read and write exactly 2^30
bytes; the HBM bytes of a call are the scaled sums over its dispatches (the
product kernel and, for TSMTTSM, its partial-sum kernel).  Prints a markdown
table; with --json, stores each call under its key (the `roofline.traffic`
source of bench.py) and keeps the other entries of the file.
"""
from __future__ import annotations

import argparse
import json
import sys

from .benchdb_counters import byte_factor, load_dispatches, segments


def short(name: str) -> str:
    for pre in ("void skinny::", "skinny::"):
        if name.startswith(pre):
            name = name[len(pre):]
    return name.split("(")[0]


def read_marks(path: str):
    out = []
    with open(path) as f:
        for line in f:
            p = line.rstrip("\n").split("\t")
            if len(p) >= 4:
                out.append({"key": p[0], "op": p[1], "M": int(p[2]), "bytes": int(p[3])})
    return out


def summarize(marks: str, fetch: str, write: str):
    calls = read_marks(marks)
    fd, wd = load_dispatches(fetch), load_dispatches(write)
    fs, ws = segments(fd), segments(wd)
    if not (len(fs) == len(ws) == len(calls)):
        raise SystemExit(f"segment counts differ: marks {len(calls)}, fetch run {len(fs)}, "
                         f"write run {len(ws)}")
    ff, wf = byte_factor(fd, "FETCH_SIZE"), byte_factor(wd, "WRITE_SIZE")
    rows = []
    for c, f, w in zip(calls, fs, ws):
        kf = [short(d[1]) for d in f]
        if kf != [short(d[1]) for d in w]:
            raise SystemExit(f"{c['key']}: the two runs launched different kernels")
        fk = sum(d[3].get("FETCH_SIZE", 0.0) for d in f)
        wk = sum(d[3].get("WRITE_SIZE", 0.0) for d in w)
        hbm = ff * fk * 1024.0 + wf * wk * 1024.0
        rows.append(dict(c, kernels=kf, fetch_kib=fk, write_kib=wk, hbm=hbm))
    return ff, wf, rows


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--marks", required=True)
    ap.add_argument("--fetch", required=True)
    ap.add_argument("--write", required=True)
    ap.add_argument("--json", help="traffic file to update (e.g. profiles/traffic.json)")
    a = ap.parse_args()
    ff, wf, rows = summarize(a.marks, a.fetch, a.write)
    print(f"FETCH_SIZE x {ff:.4f}, WRITE_SIZE x {wf:.4f} (calibrated on the copy kernel's "
          f"dispatches, 2^30 bytes read and written each)\n")
    print("| call | kernels | FETCH_SIZE KiB | WRITE_SIZE KiB | HBM bytes (calibrated) | "
          "algorithmic bytes | ratio |")
    print("|---|---|---|---|---|---|---|")
    for r in rows:
        print(f"| {r['key']} | `{'`, `'.join(r['kernels'])}` | {r['fetch_kib']:.0f} | "
              f"{r['write_kib']:.0f} | {r['hbm']:.4e} | {r['bytes']} | "
              f"{r['hbm'] / r['bytes']:.4f} |")
    if a.json:
        try:
            with open(a.json) as f:
                d = json.load(f)
        except OSError:
            d = {}
        for r in rows:
            d[r["key"]] = {
                "hbm_bytes_per_launch": r["hbm"],
                "algorithmic_bytes": r["bytes"],
                "fetch_kib": r["fetch_kib"],
                "write_kib": r["write_kib"],
                "fetch_factor": ff,
                "write_factor": wf,
                "kernels": r["kernels"],
            }
        with open(a.json, "w") as f:
            json.dump(d, f, indent=1, sort_keys=True)
            f.write("\n")
    return 0


if __name__ == "__main__":
    sys.exit(main())
