#!/usr/bin/env python3
"""Headline benchmark: FP64 TSMTTSM and TSMM throughput on one MI355X.

One step = the metric's sweep: for M = N in {1, 2, 4, 8, 16, 32, 64},
TSMTTSM C = A^T B (A, B row-major K x M, K = 2^26) followed by TSMM
C = A B (A row-major K x M, B M x M), alpha = 1, beta = 0, out of place,
14 calls on inputs already resident in HBM (generated on the device from the
counter RNG, U[-1,1), seed 1).  Bytes are the reference's algorithmic counts
(perf.cu:209-231):  TSMTTSM (M + N) K 8,  TSMM (M K + N K + M N) 8.

Prints ONE JSON line:  value = total algorithmic bytes of the timed steps /
wall time (GB/s); per-width GB/s from HIP events on the launch stream;
roofline of the dominant (longest) kernel against the MI355X HBM peak; and a
CPU baseline (the oracle's OpenMP FP64 loops, timed on this host on a bounded
sample of the same sweep).

    python bench.py [--gpus 1] [--steps 10] [--warmup 2] [--no-cpu]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

WIDTHS = (1, 2, 4, 8, 16, 32, 64)
HBM_PEAK_GBS = 8000.0


def log(msg: str) -> None:
    sys.stderr.write(msg + "\n")
    sys.stderr.flush()


def op_bytes(op: str, M: int, N: int, K: int) -> int:
    if op == "tsmttsm":
        return (M + N) * K * 8
    return (M * K + N * K + M * N) * 8


def traffic_from_profiles(kernel_key: str):
    """HBM bytes per launch measured with rocprofv3 PMC counters, if a
    committed summary for this kernel exists (profiles/traffic.json)."""
    p = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(p) as f:
            d = json.load(f)
        v = d.get(kernel_key)
        return None if v is None else float(v["hbm_bytes_per_launch"])
    except (OSError, ValueError, KeyError):
        return None


def _time_ms(torch, fn, reps=3):
    """Median milliseconds of fn() over reps calls (one untimed call first),
    HIP events on the current stream."""
    fn()
    ts = []
    for _ in range(reps):
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        fn()
        e1.record()
        torch.cuda.synchronize()
        ts.append(e0.elapsed_time(e1))
    ts.sort()
    return ts[len(ts) // 2]


def extra_configs(sk, torch, logK: int) -> dict:
    """The other benchmark configurations (not part of `value`):
    C2 complex-double TSMM (out of place and in place) and in-place FP64
    TSMM over the same widths at K = 2^logK; C3 Ozaki-INT8 against FP64 TSMTTSM at M = N = 128, K = 2^24,
    for the Gram product and a general one, errors against the
    double-double (HR) result; C4 one persistent batched
    launch over 512 uneven products; C5 two SVQB sweeps over 64 blocks
    (K = 2^22, M cycling 8..64) replayed from a hipGraph."""
    import random

    dev = torch.device("cuda:0")
    out = {}
    K = 1 << logK

    # ---- C2: DC TSMM (out of place and in place) and in-place DR TSMM ----
    dc, dci, ip = {}, {}, {}
    for M in WIDTHS:
        log(f"C2 width {M}")
        A = torch.empty(2 * K * M, dtype=torch.float64, device=dev)
        Cc = torch.empty(2 * K * M, dtype=torch.float64, device=dev)
        Bc = torch.empty(2 * M * M, dtype=torch.float64, device=dev)
        sk.fill_uniform(A, A.numel(), 3, 0)
        sk.fill_uniform(Bc, Bc.numel(), 3, 1)
        nb = op_bytes("tsmm", M, M, K) * 2
        ms = _time_ms(torch, lambda: sk.ztsmm(0, M, M, K, A, M, 1.0, Bc, M, 0.0, Cc, M))
        dc[str(M)] = round(nb / (ms * 1e-3) / 1e9, 1)
        ms = _time_ms(torch, lambda: sk.ztsmm(0, M, M, K, A, M, 1.0, Bc, M, 0.0, A, M))
        dci[str(M)] = round(nb / (ms * 1e-3) / 1e9, 1)
        del Cc
        X = A[: K * M]
        Br = Bc[: M * M]
        ms = _time_ms(torch, lambda: sk.dtsmm(0, M, M, K, X, M, 1.0, Br, M, 0.0, X, M))
        ip[str(M)] = round(op_bytes("tsmm", M, M, K) / (ms * 1e-3) / 1e9, 1)
        del A, Bc, X, Br
        torch.cuda.empty_cache()
    out["c2_dc_tsmm_GBs"] = dc
    out["c2_dc_inplace_tsmm_GBs"] = dci
    out["c2_inplace_tsmm_GBs"] = ip

    # ---- C3: Ozaki vs FP64 vs dd, M = N = 128, K = 2^24 ------------------
    # C3 data: column j = (v + 1e-8 u_j) 2^e_j with v, u_j ~ U[-1,1) and
    # integer e_j ~ U[-20, 20] (exact scaling).  Two cases: the Gram product
    # A^T A (B is A: read once, tiles below the diagonal mirrored) and a
    # general A^T B with B built the same way from other noise.
    M = 128
    K3 = 1 << 24
    erng = random.Random(5)
    v = torch.empty(K3, dtype=torch.float64, device=dev)
    sk.fill_uniform(v, K3, 5, 0)

    def c3_matrix(stream_id):
        X = torch.empty(K3 * M, dtype=torch.float64, device=dev)
        sk.fill_uniform(X, X.numel(), 5, stream_id)
        scale = torch.tensor([2.0 ** erng.randint(-20, 20) for _ in range(M)],
                             dtype=torch.float64, device=dev)
        X.view(K3, M).mul_(1e-8).add_(v[:, None]).mul_(scale)
        return X

    A = c3_matrix(1)
    c3 = {"workload": "TSMTTSM A^T B, M=N=128, K=2^24, columns (v + 1e-8 u_j) 2^e_j",
          "error": "relative Frobenius against the double-double (HR) result",
          "tolerance_8Keps": 8 * K3 * 2.0 ** -53}
    Cd = torch.zeros(M * M, dtype=torch.float64, device=dev)
    Cq = torch.zeros(2 * M * M, dtype=torch.float64, device=dev)
    for case in ("gram", "general"):
        log(f"C3 ozaki {case}")
        B = A if case == "gram" else c3_matrix(2)
        assert sk.htsmttsm_dd(0, M, M, K3, A, M, 1.0, B, M, 0.0, Cq, M)
        torch.cuda.synchronize()
        q = Cq.view(M * M, 2)
        ref_hi, ref_lo = q[:, 0].clone(), q[:, 1].clone()

        def rel_err(C):
            return float(((C - ref_hi) - ref_lo).norm() / ref_hi.norm())

        r = {}
        ms64 = _time_ms(torch, lambda: sk.dtsmttsm(0, M, M, K3, A, M, 1.0, B, M, 0.0, Cd, M), 5)
        r["fp64_ms"] = round(ms64, 3)
        r["fp64_rel_err"] = rel_err(Cd)
        for S in (3, 4, 5):
            ms = _time_ms(torch, lambda: sk.dtsmttsm_ozaki(0, M, M, K3, A, M, 1.0, B, M, 0.0,
                                                          Cd, M, slices=S), 5)
            r[f"ozaki{S}_ms"] = round(ms, 3)
            r[f"ozaki{S}_speedup"] = round(ms64 / ms, 3)
            r[f"ozaki{S}_rel_err"] = rel_err(Cd)
        c3[case] = r
        del B, q, ref_hi, ref_lo
    out["c3"] = c3
    del A, v, Cd, Cq
    torch.cuda.empty_cache()

    # ---- C4: 512 uneven products in one persistent launch ----------------
    log("C4 batch")
    rng = random.Random(4)
    choices = (1, 2, 3, 4, 6, 8, 12, 16, 24, 32, 48, 64)
    shapes = []
    total = 0
    for i in range(512):
        Kb = int(2 ** rng.uniform(12, 22))
        Mb, Nb = rng.choice(choices), rng.choice(choices)
        op = "tsmttsm" if i % 2 == 0 else "tsmm"
        sizes = (Kb * Mb, Kb * Nb, Mb * Nb) if op == "tsmttsm" else (Kb * Mb, Mb * Nb, Kb * Nb)
        shapes.append((op, Mb, Nb, Kb, sizes))
        total += sum((n + 31) // 32 * 32 for n in sizes)
    pool = torch.empty(total, dtype=torch.float64, device=dev)
    sk.fill_uniform(pool, pool.numel(), 7, 0)
    base, off, ops, nbytes = pool.data_ptr(), 0, [], 0
    for op, Mb, Nb, Kb, sizes in shapes:
        ptrs = []
        for n in sizes:
            ptrs.append(base + 8 * off)
            off += (n + 31) // 32 * 32
        a, b, c = ptrs
        if op == "tsmttsm":
            ops.append((op, Mb, Nb, Kb, a, Mb, 1.0, b, Nb, 0.0, c, Mb))
        else:
            ops.append((op, Mb, Nb, Kb, a, Mb, 1.0, b, Mb, 0.0, c, Nb))
        nbytes += op_bytes(op, Mb, Nb, Kb) if op == "tsmm" else (Mb + Nb) * Kb * 8
    batch = sk.Batch(ops)
    ms = _time_ms(torch, batch.run, 5)
    out["c4"] = {"ops": len(ops), "items": batch.items, "ms": round(ms, 3),
                 "GBs": round(nbytes / (ms * 1e-3) / 1e9, 1), "bytes": nbytes}
    batch.close()
    del pool
    torch.cuda.empty_cache()

    out["c5"] = svqb_config(sk, torch)
    return out


def svqb_config(sk, torch, fused: bool = True) -> dict:
    """C5: two SVQB sweeps over 64 blocks (K = 2^22, M cycling 8, 16, 32,
    64) replayed from a hipGraph, and (fused) the same sweeps as one
    cooperative launch with grid barriers."""
    dev = torch.device("cuda:0")
    # ---- C5: SVQB, 64 blocks, K = 2^22, two sweeps from a hipGraph -------
    log("C5 svqb")
    K5 = 1 << 22
    blocks, bufs = [], []
    for i in range(64):
        M5 = (8, 16, 32, 64)[i % 4]
        X = torch.empty(K5 * M5, dtype=torch.float64, device=dev)
        sk.fill_uniform(X, X.numel(), 9, i)
        sigma = torch.logspace(0, -6, M5, dtype=torch.float64, device=dev)
        X.view(K5, M5).mul_(sigma)
        bufs.append((X, X.clone()))
        blocks.append((M5, K5, X, M5))
    graph = sk.Graph(blocks, sweeps=2)
    torch.cuda.synchronize()

    def replay():
        for X, X0 in bufs:
            X.copy_(X0)
        graph.launch()

    # restore + replay; the restore copies are timed first and removed, and
    # the errors are taken after the last replay
    ms_copy = _time_ms(torch, lambda: [X.copy_(X0) for X, X0 in bufs], 3)
    ms_all = _time_ms(torch, replay, 3)
    errs = [sk.orth_error(M5, K5, X, M5) for (M5, K5, X, _) in blocks]
    svqb_bytes = sum(2 * 3 * M5 * K5 * 8 for (M5, K5, _, _) in blocks)
    ms_graph = max(ms_all - ms_copy, 1e-6)
    c5 = {"blocks": 64, "K": K5, "sweeps": 2, "ms": round(ms_graph, 3),
          "ms_per_block": round(ms_graph / 64, 4),
          "GBs": round(svqb_bytes / (ms_graph * 1e-3) / 1e9, 1),
          "max_orth_err_fro": max(errs)}
    if fused:
        # the same two sweeps as one cooperative launch with grid barriers
        # (synchronous call: its per-call buffers are included in the time)
        def fused_run():
            for X, X0 in bufs:
                X.copy_(X0)
            sk.dsvqb_fused(blocks, sweeps=2)
        ms_fused = max(_time_ms(torch, fused_run, 3) - ms_copy, 1e-6)
        errs_fused = [sk.orth_error(M5, K5, X, M5) for (M5, K5, X, _) in blocks]
        c5["fused_grid_barrier_ms"] = round(ms_fused, 3)
        c5["fused_max_orth_err_fro"] = max(errs_fused)
    graph.close()
    del bufs, blocks
    torch.cuda.empty_cache()
    return c5


def cpu_baseline(K_cpu: int, min_seconds: float, threads: int):
    import numpy as np

    import oracle

    oracle.build()
    Mmax = max(WIDTHS)
    A = oracle.uniform(K_cpu * Mmax, 1, 0)
    B = oracle.uniform(K_cpu * Mmax, 1, 1)
    C = np.zeros(Mmax * Mmax)
    T = np.zeros(K_cpu * Mmax)
    Bs = {M: oracle.uniform(M * M, 1, 2 + M) for M in WIDTHS}
    total_bytes = 0
    t0 = time.perf_counter()
    sweeps = 0
    while True:
        for M in WIDTHS:
            oracle.tsmttsm_par(threads, M, M, K_cpu, A, M, 1.0, B, M, 0.0, C, M)
            total_bytes += op_bytes("tsmttsm", M, M, K_cpu)
            oracle.tsmm_par(threads, M, M, K_cpu, A, M, 1.0, Bs[M], M, 0.0, T, M)
            total_bytes += op_bytes("tsmm", M, M, K_cpu)
        sweeps += 1
        el = time.perf_counter() - t0
        if el >= min_seconds:
            break
    return {
        "value": round(total_bytes / el / 1e9, 3),
        "unit": "GB/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"{sweeps} sweep(s) of the same 14 FP64 calls at K=2^{K_cpu.bit_length() - 1} "
                   f"(OpenMP oracle loops, {threads} threads), {el:.1f} s"),
    }


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--logK", type=int, default=26)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-tune", action="store_true",
                    help="skip the untimed autotuning of each shape")
    ap.add_argument("--headline-only", action="store_true",
                    help="skip the C2-C5 measurements reported beside the headline")
    ap.add_argument("--svqb-only", action="store_true",
                    help="run only the C5 SVQB measurement and print its JSON")
    ap.add_argument("--no-fused", action="store_true",
                    help="with --svqb-only: skip the cooperative (fused) SVQB launch")
    args = ap.parse_args()

    if args.gpus != 1:
        log("bench.py runs on one GPU only: the tall-skinny products here are "
            "single-device kernels (K-sharding across GPUs is described but not "
            "built); rerun with --gpus 1")
        return 2

    import torch

    import skinny_matmul_amd as sk

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    sk.lib()
    if args.svqb_only:
        print(json.dumps({"c5": svqb_config(sk, torch, fused=not args.no_fused)}), flush=True)
        return 0
    dev = torch.device("cuda:0")
    K = 1 << args.logK
    Mmax = max(WIDTHS)
    log(f"allocating 2 x {K * Mmax * 8 / 2**30:.0f} GiB")
    A = torch.empty(K * Mmax, dtype=torch.float64, device=dev)
    B = torch.empty(K * Mmax, dtype=torch.float64, device=dev)
    Cs = torch.zeros(Mmax * Mmax, dtype=torch.float64, device=dev)
    sk.fill_uniform(A, A.numel(), 1, 0)
    sk.fill_uniform(B, B.numel(), 1, 1)
    Bsm = {}
    for M in WIDTHS:
        Bsm[M] = torch.empty(M * M, dtype=torch.float64, device=dev)
        sk.fill_uniform(Bsm[M], M * M, 1, 2 + M)
    torch.cuda.synchronize()

    ops = []
    for M in WIDTHS:
        ops.append(("tsmttsm", M))
        ops.append(("tsmm", M))

    def run(op, M):
        if op == "tsmttsm":
            ok = sk.dtsmttsm(0, M, M, K, A, M, 1.0, B, M, 0.0, Cs, M)
        else:
            ok = sk.dtsmm(0, M, M, K, A, M, 1.0, Bsm[M], M, 0.0, B, M)
        assert ok, (op, M)

    step_bytes = sum(op_bytes(op, M, M, K) for op, M in ops)
    tuned = {}
    if not args.no_tune:
        # untimed: the library's autotuner picks unroll and grid per shape
        for op, M in ops:
            tuned[f"{op}_{M}"] = round(sk.autotune(op, M, M, K, reps=3), 3)
        log(f"autotuned: {sk.tune_export().strip()}")
    for _ in range(args.warmup):
        for op, M in ops:
            run(op, M)
    torch.cuda.synchronize()
    log("warmup done")

    evs = {o: [] for o in ops}
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        for o in ops:
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record()
            run(*o)
            e1.record()
            evs[o].append((e0, e1))
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    wall = t1 - t0

    per = {}
    for o in ops:
        ms = sum(a.elapsed_time(b) for a, b in evs[o]) / len(evs[o])
        per[o] = ms
    per_width = {
        op: {str(M): round(op_bytes(op, M, M, K) / (per[(op, M)] * 1e-3) / 1e9, 1)
             for (o2, M) in ops if o2 == op}
        for op in ("tsmttsm", "tsmm")
    }
    dom = max(ops, key=lambda o: per[o])
    dom_bytes = op_bytes(dom[0], dom[1], dom[1], K)
    achieved = dom_bytes / (per[dom] * 1e-3) / 1e9
    kernel_key = f"{dom[0]}_f64_M{dom[1]}_K2^{args.logK}"
    traffic = traffic_from_profiles(kernel_key)
    value = step_bytes * args.steps / wall / 1e9
    log(f"per-width GB/s: {per_width}")
    del A, B, Cs, Bsm
    torch.cuda.empty_cache()
    extras = None if args.headline_only else extra_configs(sk, torch, args.logK)
    if extras:
        log(f"extras: {json.dumps(extras)}")

    cpu = None
    if not args.no_cpu:
        log("cpu baseline ...")
        cpu = cpu_baseline(1 << 22, args.cpu_seconds, min(16, os.cpu_count() or 1))

    out = {
        "metric": "TSMTTSM/TSMM GB/s (% HBM roofline) fp64 M=N 1\u201364; Ozaki vs fp64 at M=N=128",
        "value": round(value, 1),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic U[-1,1) from a counter RNG (seed 1), generated in HBM",
        "config": {
            "workload": (f"TSMTTSM C=A^T B then TSMM C=A B, fp64, M=N in {list(WIDTHS)}, "
                         f"K=2^{args.logK}, alpha=1, beta=0, out of place, "
                         f"{'default' if args.no_tune else 'autotuned'} blockCount"),
            "K": K,
            "widths": list(WIDTHS),
            "calls_per_step": len(ops),
            "bytes_per_step": step_bytes,
        },
        "per_width_GBs": per_width,
        "roofline": {
            "kernel": kernel_key,
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
        },
        "cpu_baseline": cpu,
        "configs": extras,
    }
    if extras and "c3" in extras:
        # the metric's second half: Ozaki (3 limbs) against the FP64 kernel
        # at M = N = 128, K = 2^24 on the cancelling C3 columns
        c3 = extras["c3"]
        out["ozaki_vs_fp64"] = {
            case: {"speedup": c3[case]["ozaki3_speedup"], "ozaki3_ms": c3[case]["ozaki3_ms"],
                   "fp64_ms": c3[case]["fp64_ms"], "rel_err": c3[case]["ozaki3_rel_err"],
                   "tolerance_8Keps": c3["tolerance_8Keps"]}
            for case in ("gram", "general")}
    print(json.dumps(out))
    sys.stdout.flush()
    return 0


if __name__ == "__main__":
    sys.exit(main())
