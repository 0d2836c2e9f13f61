"""Measure this MI355X's own ceilings: HBM streams and compute peaks.

    python -m skinny_matmul_amd.ceilings [--gib 4] [--reps 5] [--json out.json]
                                         [--hierarchy]

The analogue of the reference's membw and roofline microbenchmarks
(microbenches/membw.cu:19-329, roofline.cu:19-66), run through the library's
`skinny_mb_stream` / `skinny_mb_peak`:

* HBM streams over arrays of --gib GiB each (far above the 256 MB Infinity
  Cache): read, copy, triad, write and update (a += s b), GB/s of array
  traffic, best of --reps; and the best of 40 other copy patterns
  (nontemporal stores, per-wave bursts of 16 KiB to 1 MiB);
* FP64 vector FMA, FP64 matrix-core (16x16x4 and four-block 4x4x4), INT8
  matrix-core (32x32x32) and FP32 matrix-core (16x16x4, 32x32x2 and the
  16-block 4x4x1) peaks in TFLOP/s (TOP/s), each with the shader clock
  measured while its loop runs.

* with --hierarchy: ns per dependent load over working sets from 4 KiB to
  1 GiB, GB/s re-reading buffers from 4 MiB to 8 GiB, LDS GB/s and latency
  (the role of microbenches/instruction_rates.cu), and the cost of a
  grid-wide barrier at 1-8 workgroups per CU (global_sync/sync.cu).

These are the measured denominators of the roofline fractions; the datasheet
figures (8 TB/s, 78.6 TF FP64, 5 POP/s INT8) are upper bounds only.
"""
from __future__ import annotations

import argparse
import json
import sys

from . import mb_clock, mb_peak, mb_stream

STREAMS = ("read", "copy", "copy_variants", "triad", "write", "update")
PEAKS = ("fp64_fma", "fp64_mfma", "fp64_mfma_4x4x4", "int8_mfma", "fp32_mfma",
         "fp32_mfma_32x32x2", "fp32_mfma_4x4x1")


def measure(gib: float = 4.0, reps: int = 5) -> dict:
    out = {}
    nbytes = int(gib * (1 << 30))
    for k in STREAMS:
        out[k + "_GBs"] = round(mb_stream(k, nbytes, reps), 1)
        print(f"{k:13s} {out[k + '_GBs']:9.1f} GB/s", flush=True)
    for k in PEAKS:
        unit = "TOP/s" if k.startswith("int8") else "TFLOP/s"
        out[k] = round(mb_peak(k, max(1, reps // 2)), 2)
        out[k + "_clock_MHz"] = round(mb_clock(k), 0)
        print(f"{k:15s} {out[k]:9.2f} {unit} at {out[k + '_clock_MHz']:.0f} MHz", flush=True)
    out["idle_clock_MHz"] = round(mb_clock("idle"), 0)
    return out


LATENCY_SIZES = tuple(1 << k for k in range(12, 31))          # 4 KiB .. 1 GiB
BANDWIDTH_SIZES = tuple(1 << k for k in range(22, 34))        # 4 MiB .. 8 GiB


def hierarchy(reps: int = 3) -> dict:
    """Latency and re-read bandwidth by working-set size, and the LDS
    figures: where L1 (32 KiB per CU), L2 (4 MiB per XCD), the 256 MiB
    Infinity Cache and HBM take over."""
    from . import mb_cache
    out = {"latency_ns": {}, "bandwidth_GBs": {}}
    for b in LATENCY_SIZES:
        v = round(mb_cache("latency", b, reps), 1)
        out["latency_ns"][str(b)] = v
        print(f"latency   {b / 1024:12.0f} KiB {v:9.1f} ns", flush=True)
    for b in BANDWIDTH_SIZES:
        v = round(mb_cache("bandwidth", b, reps), 1)
        out["bandwidth_GBs"][str(b)] = v
        print(f"bandwidth {b / 2**20:12.0f} MiB {v:9.1f} GB/s", flush=True)
    from . import mb_barrier
    out["barrier_us"] = {}
    for per in (1, 2, 4, 8):
        v = mb_barrier(per, reps)
        if v is None:
            break
        out["barrier_us"][str(per)] = round(v, 2)
        print(f"grid barrier, {per} workgroup(s) per CU: {v:.2f} us", flush=True)
    out["lds_GBs"] = round(mb_cache("lds_bandwidth", 0, reps), 1)
    out["lds_latency_ns"] = round(mb_cache("lds_latency", 0, reps), 2)
    print(f"LDS {out['lds_GBs']:.1f} GB/s, {out['lds_latency_ns']:.2f} ns per dependent read",
          flush=True)
    return out


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--gib", type=float, default=4.0)
    ap.add_argument("--reps", type=int, default=5)
    ap.add_argument("--json", default=None)
    ap.add_argument("--hierarchy", action="store_true",
                    help="also sweep latency and bandwidth over working-set sizes")
    args = ap.parse_args(argv)
    res = measure(args.gib, args.reps)
    if args.hierarchy:
        res["hierarchy"] = hierarchy(max(1, args.reps // 2))
    if args.json:
        with open(args.json, "w") as f:
            json.dump(res, f, indent=1)
    return 0


if __name__ == "__main__":
    sys.exit(main())
