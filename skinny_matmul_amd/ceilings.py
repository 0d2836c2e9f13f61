"""Measure this MI355X's own ceilings: HBM streams and compute peaks.

    python -m skinny_matmul_amd.ceilings [--gib 4] [--reps 5] [--json out.json]

The analogue of the reference's membw and roofline microbenchmarks
(microbenches/membw.cu:19-329, roofline.cu:19-66), run through the library's
`skinny_mb_stream` / `skinny_mb_peak`:

* HBM streams over arrays of --gib GiB each (far above the 256 MB Infinity
  Cache): read, copy, triad, write and update (a += s b), GB/s of array
  traffic, best of --reps;
* FP64 vector FMA, FP64 matrix-core (16x16x4) and INT8 matrix-core
  (32x32x32) peaks in TFLOP/s (TOP/s).

These are the measured denominators of the roofline fractions; the datasheet
figures (8 TB/s, 78.6 TF FP64, 5 POP/s INT8) are upper bounds only.
"""
from __future__ import annotations

import argparse
import json
import sys

from . import mb_peak, mb_stream

STREAMS = ("read", "copy", "triad", "write", "update")
PEAKS = ("fp64_fma", "fp64_mfma", "int8_mfma")


def measure(gib: float = 4.0, reps: int = 5) -> dict:
    out = {}
    nbytes = int(gib * (1 << 30))
    for k in STREAMS:
        out[k + "_GBs"] = round(mb_stream(k, nbytes, reps), 1)
        print(f"{k:8s} {out[k + '_GBs']:9.1f} GB/s", flush=True)
    for k in PEAKS:
        unit = "TOP/s" if k.startswith("int8") else "TFLOP/s"
        out[k] = round(mb_peak(k, max(1, reps // 2)), 2)
        print(f"{k:9s} {out[k]:9.2f} {unit}", flush=True)
    return out


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--gib", type=float, default=4.0)
    ap.add_argument("--reps", type=int, default=5)
    ap.add_argument("--json", default=None)
    args = ap.parse_args(argv)
    res = measure(args.gib, args.reps)
    if args.json:
        with open(args.json, "w") as f:
            json.dump(res, f, indent=1)
    return 0


if __name__ == "__main__":
    sys.exit(main())
