"""Fill BenchDB's drambw and occupancy columns from rocprofv3 counters.

    python -m skinny_matmul_amd.benchdb_counters --db benchmarks.db \
        --marks marks.tsv --fetch F.csv --write W.csv [--occupancy O.csv]

The counter files are the `*_counter_collection.csv` of three rocprofv3 runs
of the same replay (`skinny_perf <op> <type> [range] --counters marks.tsv`),
one per counter set, since one run cannot hold them all:

    rocprofv3 --pmc FETCH_SIZE --output-format csv -d F -o r -- skinny_perf ...
    rocprofv3 --pmc WRITE_SIZE ...
    rocprofv3 --pmc SQ_WAVE_CYCLES SQ_BUSY_CYCLES ...

The replay first runs the HBM copy microbenchmark over 1 GiB arrays, then
per stored row a one-element fill_uniform (start marker), the row's call and
a one-element fill_int (end marker).  This tool

* splits each run's dispatches into the marked segments (one per line of
  the marks file, in order);
* calibrates the byte counters on the copy kernel's dispatches, each of
  which reads and writes exactly one array: bytes = FETCH_SIZE (KiB) x 1024
  x f_read and WRITE_SIZE x 1024 x f_write, with f = 2^30 / the counted
  bytes of a copy dispatch (the median over the copies);
* calibrates occupancy on the copy dispatches launched with one 256-thread
  workgroup per CU (4 of the CU's 32 wave slots held for the whole run):
  occupancy = SQ_WAVE_CYCLES / SQ_BUSY_CYCLES x (0.125 / that ratio of the
  copy), the mean fraction of wave slots in use while the CUs are busy;
* writes drambw = calibrated bytes of the row's dispatches / the row's
  stored time (GB/s, the traffic the hardware moved) and occupancy into the
  row with the same key columns.

The counter-derived columns take the place of the reference's CUPTI metrics
(perf.cu:239-269, metrics.cuh:257-383).
"""
from __future__ import annotations

import argparse
import csv
import sqlite3
import statistics
import sys

COPY_BYTES = 1 << 30
KEYS = ("multype", "device", "types", "M", "N", "name", "inplace", "zerobeta", "branch",
        "usr1_name", "usr1_val")


def load_dispatches(path: str):
    """[(dispatch id, kernel name, grid size, {counter: summed value})] in
    dispatch order."""
    by_id = {}
    with open(path, newline="") as f:
        for r in csv.DictReader(f):
            d = int(r["Dispatch_Id"])
            e = by_id.setdefault(d, [r["Kernel_Name"], int(r.get("Grid_Size") or 0), {}])
            name = r["Counter_Name"]
            e[2][name] = e[2].get(name, 0.0) + float(r["Counter_Value"])
    return [(d, *by_id[d]) for d in sorted(by_id)]


def segments(disp):
    """Dispatches between each start marker (fill_kernel<0>) and the next end
    marker (fill_kernel<2>), in order."""
    out, cur = [], None
    for d in disp:
        name = d[1]
        if "fill_kernel<0>" in name:
            cur = []
        elif "fill_kernel<2>" in name:
            if cur is not None:
                out.append(cur)
            cur = None
        elif cur is not None:
            cur.append(d)
    return out


def copy_dispatches(disp):
    return [d for d in disp if "stream_kernel<1," in d[1]]


def byte_factor(disp, counter: str) -> float:
    vals = [d[3][counter] * 1024.0 for d in copy_dispatches(disp) if d[3].get(counter)]
    if not vals:
        raise SystemExit(f"no copy dispatches with {counter} to calibrate on")
    return COPY_BYTES / statistics.median(vals)


def occupancy_factor(disp, cu_count: int) -> float:
    one = [d for d in copy_dispatches(disp) if d[2] == 256 * cu_count]
    ratios = [d[3]["SQ_WAVE_CYCLES"] / d[3]["SQ_BUSY_CYCLES"] for d in one
              if d[3].get("SQ_BUSY_CYCLES")]
    if not ratios:
        raise SystemExit("no one-workgroup-per-CU copy dispatch to calibrate occupancy on")
    return 0.125 / statistics.median(ratios)


def read_marks(path: str):
    rows = []
    with open(path) as f:
        for line in f:
            p = line.rstrip("\n").split("\t")
            if len(p) < 12:
                continue
            key = dict(zip(KEYS, p[:11]))
            for k in ("M", "N", "inplace", "zerobeta"):
                key[k] = int(key[k])
            rows.append((key, float(p[11])))
    return rows


def merge(db: str, marks: str, fetch: str, write: str, occupancy: str | None = None,
          cu_count: int = 256, dry_run: bool = False):
    rows = read_marks(marks)
    fd, wd = load_dispatches(fetch), load_dispatches(write)
    fs, ws = segments(fd), segments(wd)
    if not (len(fs) == len(ws) == len(rows)):
        raise SystemExit(f"segment counts differ: marks {len(rows)}, fetch run {len(fs)}, "
                         f"write run {len(ws)}")
    ff, wf = byte_factor(fd, "FETCH_SIZE"), byte_factor(wd, "WRITE_SIZE")
    occ = None
    if occupancy:
        od = load_dispatches(occupancy)
        os_ = segments(od)
        if len(os_) != len(rows):
            raise SystemExit(f"occupancy run has {len(os_)} segments, marks {len(rows)}")
        of = occupancy_factor(od, cu_count)
        occ = []
        for seg in os_:
            wc = sum(d[3].get("SQ_WAVE_CYCLES", 0.0) for d in seg)
            bc = sum(d[3].get("SQ_BUSY_CYCLES", 0.0) for d in seg)
            occ.append(wc / bc * of if bc else 0.0)
    out = []
    for i, (key, t) in enumerate(rows):
        nbytes = (ff * sum(d[3].get("FETCH_SIZE", 0.0) for d in fs[i]) * 1024.0 +
                  wf * sum(d[3].get("WRITE_SIZE", 0.0) for d in ws[i]) * 1024.0)
        drambw = nbytes / t / 1e9 if t > 0 else 0.0
        out.append((key, drambw, occ[i] if occ else None, nbytes))
    if not dry_run:
        con = sqlite3.connect(db)
        for key, drambw, o, _ in out:
            where = " AND ".join(f"{k} = ?" for k in KEYS)
            sets = "drambw = ?" + (", occupancy = ?" if o is not None else "")
            vals = [drambw] + ([o] if o is not None else []) + [key[k] for k in KEYS]
            con.execute(f"UPDATE benchmarks SET {sets} WHERE {where}", vals)
        con.commit()
        con.close()
    return {"fetch_factor": ff, "write_factor": wf, "rows": out}


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--db", required=True)
    ap.add_argument("--marks", required=True)
    ap.add_argument("--fetch", required=True)
    ap.add_argument("--write", required=True)
    ap.add_argument("--occupancy")
    ap.add_argument("--cu-count", type=int, default=256)
    ap.add_argument("--dry-run", action="store_true")
    a = ap.parse_args()
    r = merge(a.db, a.marks, a.fetch, a.write, a.occupancy, a.cu_count, a.dry_run)
    print(f"FETCH_SIZE x {r['fetch_factor']:.4f}, WRITE_SIZE x {r['write_factor']:.4f} "
          f"(calibrated on the copy kernel)")
    for key, drambw, o, nbytes in r["rows"]:
        occ = "" if o is None else f" occupancy {o:.3f}"
        print(f"{key['multype']} {key['types']} M={key['M']} N={key['N']} {key['name']} "
              f"inplace={key['inplace']} zerobeta={key['zerobeta']}: "
              f"{nbytes / 1e6:.1f} MB, drambw {drambw:.1f} GB/s{occ}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
