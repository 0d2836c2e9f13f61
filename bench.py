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
C3_WIDTHS = (16, 64, 128, 256)
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
    TSMM over the same widths at K = 2^logK; C3 FP64, double-double (HR)
    and Ozaki-INT8 TSMTTSM at M = N in {16, 64, 128, 256}, K = 2^24, for the
    Gram product and a general one, errors against the double-double
    result; C4 one persistent batched
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

    # ---- C3: precision sweep, M = N in {16, 64, 128, 256}, K = 2^24 -------
    # C3 data: column j = (v + 1e-8 u_j) 2^e_j with v ~ U[-1,1) (stream 0),
    # u_j ~ U[-1,1) from stream 1 + j (column-wise, so any column can be
    # regenerated on its own) and integer e_j ~ U[-20, 20] (exact scaling).
    # Two cases: the Gram product A^T A (B is A) and a general A^T B with B
    # built from the same v and other noise.  Every path is timed (FP64,
    # double-double HR, Ozaki with 3-5 limbs) and its error is the relative
    # Frobenius distance to the double-double result, which the GPU tests
    # hold within 1e-28 of the exact product on the same data.
    K3 = 1 << 24
    v = torch.empty(K3, dtype=torch.float64, device=dev)
    sk.fill_uniform(v, K3, 5, 0)
    u = torch.empty(K3, dtype=torch.float64, device=dev)

    def c3_matrix(M, stream_base, erng):
        X = torch.empty((K3, M), dtype=torch.float64, device=dev)
        for j in range(M):
            sk.fill_uniform(u, K3, 5, stream_base + j)
            col = u * 1e-8
            col += v
            col *= 2.0 ** erng.randint(-20, 20)
            X[:, j] = col
        return X.view(-1)

    c3 = {"workload": "TSMTTSM A^T B, M=N in {16,64,128,256}, K=2^24, columns (v + 1e-8 u_j) 2^e_j",
          "error": "relative Frobenius distance to the double-double (HR) result",
          "tolerance_8Keps": 8 * K3 * 2.0 ** -53}
    for M in C3_WIDTHS:
        erng = random.Random(5 + M)
        A = c3_matrix(M, 1, erng)
        Cd = torch.zeros(M * M, dtype=torch.float64, device=dev)
        Cq = torch.zeros(2 * M * M, dtype=torch.float64, device=dev)
        for case in ("gram", "general"):
            log(f"C3 M={M} {case}")
            B = A if case == "gram" else c3_matrix(M, 1001, erng)
            ms_dd = _time_ms(torch, lambda: sk.htsmttsm_dd(0, M, M, K3, A, M, 1.0, B, M, 0.0,
                                                          Cq, M), 3)
            torch.cuda.synchronize()
            q = Cq.view(M * M, 2)
            ref_hi, ref_lo = q[:, 0].clone(), q[:, 1].clone()

            def rel_err(C):
                return float(((C - ref_hi) - ref_lo).norm() / ref_hi.norm())

            r = {"dd_ms": round(ms_dd, 3)}
            ms64 = _time_ms(torch, lambda: sk.dtsmttsm(0, M, M, K3, A, M, 1.0, B, M, 0.0,
                                                       Cd, M), 5)
            r["fp64_ms"] = round(ms64, 3)
            r["fp64_rel_err"] = rel_err(Cd)
            ms_hr = _time_ms(torch, lambda: sk.htsmttsm(0, M, M, K3, A, M, 1.0, B, M, 0.0,
                                                        Cd, M), 3)
            r["hr_ms"] = round(ms_hr, 3)
            r["hr_rel_err"] = rel_err(Cd)
            for S in (3, 4, 5):
                ms = _time_ms(torch, lambda: sk.dtsmttsm_ozaki(0, M, M, K3, A, M, 1.0, B, M, 0.0,
                                                              Cd, M, slices=S), 5)
                r[f"ozaki{S}_ms"] = round(ms, 3)
                r[f"ozaki{S}_speedup"] = round(ms64 / ms, 3)
                r[f"ozaki{S}_rel_err"] = rel_err(Cd)
            c3.setdefault(str(M), {})[case] = r
            del B, q, ref_hi, ref_lo
        del A, Cd, Cq
        torch.cuda.empty_cache()
    out["c3"] = c3
    del v, u
    torch.cuda.empty_cache()

    # ---- C4: 512 uneven TSMTTSM + TSMM pairs in one persistent launch -----
    # pair i: C1 = A^T B (TSMTTSM, M x N) and C2 = A W (TSMM, K x N) on the
    # same A (K x M), B (K x N), W (M x N); K log-uniform in [2^12, 2^22].
    log("C4 batch")
    rng = random.Random(4)
    choices = (1, 2, 3, 4, 6, 8, 12, 16, 24, 32, 48, 64)
    pairs = []
    total = 0

    def rnd(n):
        return (n + 31) // 32 * 32

    for i in range(512):
        Kb = int(2 ** rng.uniform(12, 22))
        Mb, Nb = rng.choice(choices), rng.choice(choices)
        sizes = (Kb * Mb, Kb * Nb, Mb * Nb, Mb * Nb, Kb * Nb)  # A, B, W, C1, C2
        pairs.append((Mb, Nb, Kb, sizes))
        total += sum(rnd(n) for n in sizes)
    pool = torch.empty(total, dtype=torch.float64, device=dev)
    sk.fill_uniform(pool, pool.numel(), 7, 0)
    off, ops, views, nbytes = 0, [], [], 0
    for Mb, Nb, Kb, sizes in pairs:
        v5 = []
        for n in sizes:
            v5.append(pool[off:off + n])
            off += rnd(n)
        a, b, w, c1, c2 = v5
        ops.append(("tsmttsm", Mb, Nb, Kb, a, Mb, 1.0, b, Nb, 0.0, c1, Mb))
        ops.append(("tsmm", Mb, Nb, Kb, a, Mb, 1.0, w, Mb, 0.0, c2, Nb))
        views.append((Mb, Nb, Kb, a, b, w, c1, c2))
        nbytes += (Mb + Nb) * Kb * 8 + op_bytes("tsmm", Mb, Nb, Kb)
    batch = sk.Batch(ops)
    ms_auto = _time_ms(torch, batch.run, 5)
    default_cls = batch.classes()
    # the batch's own tuner picks the tile class of each op shape, then the
    # item size and grid of the plan
    _, tuned_ib, tuned_grid = batch.tune(reps=2)
    tuned_cls = batch.classes()
    ms = _time_ms(torch, batch.run, 5)
    # per-instance error against an FP64 library product (rocBLAS through
    # torch): the whole M x N result of each TSMTTSM, 64 sampled rows of each
    # TSMM output; relative Frobenius, against 8 K eps and 8 M eps
    eps = 2.0 ** -53
    worst_ts, worst_tm, bad = 0.0, 0.0, 0
    g = torch.Generator(device="cpu").manual_seed(4)
    for Mb, Nb, Kb, a, b, w, c1, c2 in views:
        A2 = a.view(Kb, Mb)
        ref1 = A2.t().contiguous().mm(b.view(Kb, Nb))
        got1 = c1.view(Nb, Mb).t()
        e1 = float((got1 - ref1).norm() / ref1.norm().clamp_min(1e-300))
        rows = torch.randint(0, Kb, (64,), generator=g).to(dev)
        ref2 = A2.index_select(0, rows).mm(w.view(Nb, Mb).t().contiguous())
        got2 = c2.view(Kb, Nb).index_select(0, rows)
        e2 = float((got2 - ref2).norm() / ref2.norm().clamp_min(1e-300))
        worst_ts, worst_tm = max(worst_ts, e1 / (Kb * eps)), max(worst_tm, e2 / (Mb * eps))
        bad += (e1 > 8 * Kb * eps) + (e2 > 8 * Mb * eps)
    out["c4"] = {"pairs": len(views), "ops": len(ops), "items": batch.items, "ms": round(ms, 3),
                 "GBs": round(nbytes / (ms * 1e-3) / 1e9, 1), "bytes": nbytes,
                 "untuned_plan_ms": round(ms_auto, 3), "tuned_item_bytes": tuned_ib,
                 "tuned_grid": tuned_grid,
                 "ops_tuned_to_another_tile_class": sum(
                     1 for c0, c1 in zip(default_cls, tuned_cls) if c0 != c1),
                 "max_tsmttsm_err_over_K_eps": round(worst_ts, 3),
                 "max_tsmm_err_over_M_eps": round(worst_tm, 3),
                 "instances_over_8_eps_bound": int(bad)}
    batch.close()
    del pool, views, ops
    torch.cuda.empty_cache()

    out["c5"] = svqb_config(sk, torch)
    return out


def _colmajor(T):
    """An M x N torch matrix as the column-major buffer the TSMM operand B
    and the TSMTTSM result C use (entry (m, n) at n M + m)."""
    return T.t().contiguous().view(-1)


def svqb_input(sk, torch, M: int, K: int, seed: int, stream: int):
    """X = Q diag(logspace(0, -6, M)) W^T (K x M row-major) on the device:
    Q from U[-1,1) columns made orthonormal by two Cholesky QR passes (the
    library's Gram product, a Cholesky factor, the library's TSMM by its
    inverse), W a random orthogonal M x M from a QR on the host.  X has the
    singular values logspace(0, -6, M) up to rounding."""
    dev = torch.device("cuda:0")
    X = torch.empty(K * M, dtype=torch.float64, device=dev)
    sk.fill_uniform(X, X.numel(), seed, stream)
    G = torch.empty(M * M, dtype=torch.float64, device=dev)
    for _ in range(2):
        assert sk.dtsmttsm(0, M, M, K, X, M, 1.0, X, M, 0.0, G, M)
        L = torch.linalg.cholesky(G.view(M, M).t().cpu())  # G = L L^T
        Rinv = torch.linalg.inv(L.t())                       # R = L^T, Q = X R^-1
        assert sk.dtsmm(0, M, M, K, X, M, 1.0, _colmajor(Rinv).to(dev), M, 0.0, X, M)
    gen = torch.Generator().manual_seed(seed * 1000 + stream)
    W, _ = torch.linalg.qr(torch.randn(M, M, dtype=torch.float64, generator=gen))
    sigma = torch.logspace(0, -6, M, dtype=torch.float64)
    T = torch.diag(sigma) @ W.t()                            # X <- Q T
    assert sk.dtsmm(0, M, M, K, X, M, 1.0, _colmajor(T).to(dev), M, 0.0, X, M)
    return X


def svqb_config(sk, torch, fused: bool = True) -> dict:
    """C5: two SVQB sweeps over 64 blocks (K = 2^22, M cycling 8, 16, 32,
    64; X = Q diag(logspace(0, -6, M)) W^T) replayed from a hipGraph, and
    (fused) the same sweeps as one cooperative launch with grid barriers."""
    dev = torch.device("cuda:0")
    # ---- C5: SVQB, 64 blocks, K = 2^22, two sweeps from a hipGraph -------
    log("C5 svqb")
    K5 = 1 << 22
    blocks, bufs = [], []
    for i in range(64):
        M5 = (8, 16, 32, 64)[i % 4]
        X = svqb_input(sk, torch, M5, K5, 9, i)
        bufs.append((X, X.clone()))
        blocks.append((M5, K5, X, M5))
    graph = sk.Graph(blocks, sweeps=2)
    # a one-element fill (fill_kernel<2>) marks in a kernel trace where the
    # input construction ends and the replays begin
    sk.fill_int(torch.empty(1, dtype=torch.float64, device=dev), 1, 5, 0)
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


def _cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(K_cpu: int, min_seconds: float, threads: int):
    """The oracle's OpenMP FP64 loops over the same 14 calls at K = K_cpu
    (TSMTTSM: K-chunks with private accumulators; TSMM: rows), on `threads`
    threads for at least `min_seconds`, then on one thread for at least
    half as long."""
    import numpy as np

    import oracle

    oracle.build()
    Mmax = max(WIDTHS)
    A = oracle.uniform(K_cpu * Mmax, 1, 0)
    B = oracle.uniform(K_cpu * Mmax, 1, 1)
    C = np.zeros(Mmax * Mmax)
    T = np.zeros(K_cpu * Mmax)
    Bs = {M: oracle.uniform(M * M, 1, 2 + M) for M in WIDTHS}

    def run(nthreads, K, seconds):
        total_bytes = 0
        t0 = time.perf_counter()
        sweeps = 0
        while True:
            for M in WIDTHS:
                oracle.tsmttsm_par(nthreads, M, M, K, A, M, 1.0, B, M, 0.0, C, M)
                total_bytes += op_bytes("tsmttsm", M, M, K)
                oracle.tsmm_par(nthreads, M, M, K, A, M, 1.0, Bs[M], M, 0.0, T, M)
                total_bytes += op_bytes("tsmm", M, M, K)
            sweeps += 1
            el = time.perf_counter() - t0
            if el >= seconds:
                return total_bytes / el / 1e9, sweeps, el

    gbs, sweeps, el = run(threads, K_cpu, min_seconds)
    # one thread over the first quarter of the rows (a sweep at K_cpu takes
    # one thread about as long as the whole multi-thread sample)
    K1 = K_cpu // 4
    gbs1, sweeps1, el1 = run(1, K1, min_seconds / 2)
    try:
        affinity = len(os.sched_getaffinity(0))
    except AttributeError:
        affinity = os.cpu_count() or 1
    return {
        "value": round(gbs, 3),
        "unit": "GB/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"{sweeps} sweep(s) of the same 14 FP64 calls at K=2^{K_cpu.bit_length() - 1} "
                   f"(OpenMP oracle loops, {threads} threads), {el:.1f} s; one thread: "
                   f"{sweeps1} sweep(s) at K=2^{K1.bit_length() - 1}, {el1:.1f} s"),
        "single_core_GBs": round(gbs1, 3),
        "nproc": os.cpu_count(),
        "affinity_cpus": affinity,
        "omp_max_threads": oracle.max_threads(),
        "cpu_model": _cpu_model(),
    }


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--logK", type=int, default=26)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
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
        # every CPU this process may run on, within the 16-core share of one
        # GPU on a shared host
        try:
            ncpu = len(os.sched_getaffinity(0))
        except AttributeError:
            ncpu = os.cpu_count() or 1
        cpu = cpu_baseline(1 << 22, args.cpu_seconds, max(1, min(16, ncpu)))

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
            "traffic_source": (f"profiles/traffic.json[{kernel_key}] (rocprofv3 FETCH_SIZE and "
                               f"WRITE_SIZE, calibrated on a copy)" if traffic is not None else
                               f"none: profiles/traffic.json holds no counter summary for "
                               f"{kernel_key}"),
        },
        "cpu_baseline": cpu,
        "configs": extras,
    }
    if extras and "c3" in extras:
        # the metric's second half: Ozaki (3 limbs) against the FP64 kernel
        # at M = N = 128, K = 2^24 on the cancelling C3 columns
        c3 = extras["c3"]["128"]
        out["ozaki_vs_fp64"] = {
            case: {"speedup": c3[case]["ozaki3_speedup"], "ozaki3_ms": c3[case]["ozaki3_ms"],
                   "fp64_ms": c3[case]["fp64_ms"], "rel_err": c3[case]["ozaki3_rel_err"],
                   "tolerance_8Keps": extras["c3"]["tolerance_8Keps"]}
            for case in ("gram", "general")}
    print(json.dumps(out))
    sys.stdout.flush()
    return 0


if __name__ == "__main__":
    sys.exit(main())
