"""Persistent batched launch (replaces multi_dispatch.cuh) on the GPU.

Uneven (K, M, N) TSMTTSM and TSMM ops with integer data are compared bitwise
with the oracle, op by op; the batch is relaunched (queue reset by the
kernel), its FP64 TSMTTSM results are schedule-independent (bitwise equal
across launches on U[-1,1) data), and it replays correctly from a captured
graph with eager work between replays.
"""
import random

import numpy as np
import pytest

from gpu_util import dev, host, note

pytestmark = pytest.mark.gpu
WIDTHS = (1, 2, 3, 4, 6, 8, 12, 16, 24, 32, 48, 64)


def _make_ops(orc, rng, n, Kmax, integer=True, beta_choices=(0.0, 1.0)):
    ops, refs = [], []
    for i in range(n):
        kind = rng.choice(("tsmttsm", "tsmm"))
        M, N = rng.choice(WIDTHS), rng.choice(WIDTHS)
        K = int(2 ** rng.uniform(4, np.log2(Kmax)))
        seed = rng.randint(0, 1 << 30)
        gen = (lambda n_, s_: orc.integers(n_, seed, s_)) if integer else \
              (lambda n_, s_: orc.uniform(n_, seed, s_))
        alpha = float(rng.choice((1.0, 2.0, -1.0)))
        beta = float(rng.choice(beta_choices))
        if kind == "tsmttsm":
            self_ = M == N and rng.random() < 0.3
            lda = M + rng.randint(0, 3)
            ldb = lda if self_ else N + rng.randint(0, 3)
            ldc = M + rng.randint(0, 2)
            A = gen(K * lda, 0)
            B = A if self_ else gen(K * ldb, 1)
            C = gen(N * ldc, 2)
            want = C.copy()
            orc.tsmttsm(M, N, K, A, lda, alpha, B, ldb, beta, want, ldc)
            dA = dev(A)
            dB = dA if self_ else dev(B)
            dC = dev(C)
        else:
            inplace = rng.random() < 0.3
            ldb = M + rng.randint(0, 3)
            if inplace:
                ldc = max(N, M) + rng.randint(0, 2)
                lda = ldc
            else:
                lda = M + rng.randint(0, 3)
                ldc = N + rng.randint(0, 3)
            B = gen(N * ldb, 1)
            C = gen(K * ldc, 2)
            A = C.copy() if inplace else gen(K * lda, 0)
            want = C.copy()
            orc.tsmm(M, N, K, want if inplace else A, lda, alpha, B, ldb, beta, want, ldc)
            dC = dev(C)
            dA = dC if inplace else dev(A)
            dB = dev(B)
        ops.append((kind, M, N, K, dA, lda, alpha, dB, ldb, beta, dC, ldc))
        refs.append((kind, M, N, K, C, want, dC))
    return ops, refs


def test_batch_integer_bitwise(gpu, orc):
    rng = random.Random(42)
    ops, refs = _make_ops(orc, rng, 60, 60000)
    note(f"batch: {len(ops)} ops")
    b = gpu.Batch(ops, item_bytes=1 << 18)  # small items: many per op
    assert b.items > len(ops)
    b.run()
    for (kind, M, N, K, C0, want, dC) in refs:
        got = host(dC)
        assert np.array_equal(got, want), (kind, M, N, K)
    b.close()


def test_batch_gram_ops(gpu, orc):
    """Gram products (B is A) in a batch take the symmetric item classes:
    integer data, bitwise against the oracle, symmetric results."""
    rng = random.Random(5)
    ops, refs = [], []
    for M in (1, 8, 12, 16, 20, 32, 33, 48, 64):
        K = rng.randint(3000, 40000)
        lda = M + rng.randint(0, 3)
        seed = rng.randint(0, 1 << 30)
        A = orc.integers(K * lda, seed, 0)
        C = orc.integers(M * M, seed, 2)
        beta = float(rng.choice((0.0, 1.0)))
        want = C.copy()
        orc.tsmttsm(M, M, K, A, lda, 1.0, A, lda, beta, want, M)
        dA, dC = dev(A), dev(C)
        ops.append(("tsmttsm", M, M, K, dA, lda, 1.0, dA, lda, beta, dC, M))
        refs.append((M, K, want, dC))
    note(f"batch gram: {len(ops)} ops")
    b = gpu.Batch(ops, item_bytes=1 << 18)
    b.run()
    for (M, K, want, dC) in refs:
        got = host(dC)
        assert np.array_equal(got, want), (M, K)
    b.close()


def test_batch_relaunch_and_reproducible(gpu, orc):
    rng = random.Random(7)
    ops, refs = _make_ops(orc, rng, 40, 40000, integer=False, beta_choices=(0.0,))
    ops = [o for o in ops if o[0] == "tsmttsm"]
    refs = [r for r in refs if r[0] == "tsmttsm"]
    b = gpu.Batch(ops, item_bytes=1 << 17)
    outs = []
    for rep in range(3):
        note(f"batch relaunch {rep}")
        b.run()
        outs.append([host(r[6]).copy() for r in refs])
    for rep in (1, 2):
        for a, c in zip(outs[0], outs[rep]):
            assert np.array_equal(a, c)
    for (kind, M, N, K, C0, want, dC), got in zip(refs, outs[0]):
        scale = max(1.0, np.abs(want).max())
        assert np.max(np.abs(got - want)) <= 8 * K * np.finfo(float).eps * scale
    b.close()


def test_batch_graph_replay(gpu, orc):
    import torch
    rng = random.Random(99)
    ops, refs = _make_ops(orc, rng, 30, 30000, beta_choices=(0.0,))
    # in-place TSMM ops are not idempotent under replay: drop them here
    keep = [i for i, o in enumerate(ops) if not (o[0] == "tsmm" and o[4] is o[10])]
    ops = [ops[i] for i in keep]
    refs = [refs[i] for i in keep]
    b = gpu.Batch(ops, item_bytes=1 << 18)
    s = torch.cuda.Stream()
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        b.run(stream=s)
    for rep in range(3):
        note(f"batch graph replay {rep}")
        for r in refs:  # restore C (padding included; beta == 0 so C is not read)
            r[6].copy_(torch.from_numpy(r[4]).cuda())
        junk = torch.randn(1 << 18, device="cuda").sum()
        g.replay()
        torch.cuda.synchronize()
        for (kind, M, N, K, C0, want, dC) in refs:
            assert np.array_equal(host(dC), want), (rep, kind, M, N, K)
        del junk
    b.close()


def test_batch_tune_replans_and_stays_exact(gpu, orc):
    """skinny_dbatch_tune times plans over item sizes and grids and keeps the
    fastest; the tuned batch (outputs restored after the tuning runs) gives
    the oracle's results bitwise on integer data, and runs again."""
    import torch
    rng = random.Random(7)
    ops, refs = _make_ops(orc, rng, 40, 200000)
    b = gpu.Batch(ops)
    note(f"batch tune: {len(ops)} ops, {b.items} items")
    ms, ib, g = b.tune(reps=1)
    cus = gpu.cu_count()
    assert ms > 0
    assert ib in (0, 1 << 20, 2 << 20, 4 << 20, 8 << 20, 16 << 20)
    assert 1 <= g <= 2 * cus
    for _ in range(2):
        for (kind, M, N, K, C0, want, dC) in refs:
            dC.copy_(torch.from_numpy(C0))
        torch.cuda.synchronize()
        b.run()
        for (kind, M, N, K, C0, want, dC) in refs:
            assert np.array_equal(host(dC), want), (kind, M, N, K, ib, g)
    b.close()


# Which tile classes can compute an op, stated from the kernels' shapes: a
# packed TSMTTSM class puts P rows of both operands into the 16 matrix-core
# lanes, a subtile class covers 16 MT x 16 NT entries, a Gram class one
# square tile of a Gram op; a TSMM class covers 4 S columns of A and 16 NT of
# C, a row class M <= 1 or 2 with N <= 2.
_PACK = (16, 8, 5, 4, 3, 2)
_SUB = (1, 2, 4)


def _ts_valid(cls, M, N, gram):
    if cls < 6:
        return _PACK[cls] * max(M, N) <= 16
    if cls >= 15:
        return gram and M <= (16, 32, 64)[cls - 15]
    return 16 * _SUB[(cls - 6) // 3] >= M and 16 * _SUB[(cls - 6) % 3] >= N


def _tm_valid(cls, M, N):
    if cls >= 15:
        return M <= cls - 14 and N <= 2
    return 4 * (1, 2, 4, 8, 16)[cls // 3] >= M and 16 * _SUB[cls % 3] >= N


def _one_op(orc, rng, kind, M, N, K, gram=False, inplace=False, beta=0.0):
    seed = rng.randint(0, 1 << 30)
    if kind == "tsmttsm":
        lda = M + rng.randint(0, 2)
        ldb = lda if gram else N + rng.randint(0, 2)
        A = orc.integers(K * lda, seed, 0)
        B = A if gram else orc.integers(K * ldb, seed, 1)
        C = orc.integers(N * M, seed, 2)
        want = C.copy()
        orc.tsmttsm(M, N, K, A, lda, 1.0, B, ldb, beta, want, M)
        dA = dev(A)
        dB = dA if gram else dev(B)
        dC = dev(C)
        return (kind, M, N, K, dA, lda, 1.0, dB, ldb, beta, dC, M), C, want, dC
    ldb = M + rng.randint(0, 2)
    if inplace:
        lda = ldc = max(M, N) + rng.randint(0, 1)
    else:
        lda, ldc = M + rng.randint(0, 2), N + rng.randint(0, 2)
    B = orc.integers(N * ldb, seed, 1)
    C = orc.integers(K * ldc, seed, 2)
    A = C.copy() if inplace else orc.integers(K * lda, seed, 0)
    want = C.copy()
    orc.tsmm(M, N, K, want if inplace else A, lda, 1.0, B, ldb, beta, want, ldc)
    dC = dev(C)
    dA = dC if inplace else dev(A)
    return (kind, M, N, K, dA, lda, 1.0, dev(B), ldb, beta, dC, ldc), C, want, dC


TS_SHAPES = [(1, 1, False), (2, 3, False), (4, 4, False), (5, 2, False), (8, 8, False),
             (3, 3, True), (8, 8, True), (12, 3, False), (16, 16, True), (20, 40, False),
             (24, 24, True), (33, 64, False), (64, 64, True)]
TM_SHAPES = [(1, 1, False, 0.0), (1, 2, True, 1.0), (2, 2, False, 1.0), (2, 1, True, 0.0),
             (3, 5, False, 0.0), (8, 8, False, 0.0), (8, 8, True, 1.0), (12, 20, False, 1.0),
             (33, 17, False, 0.0), (64, 64, True, 0.0)]


def test_batch_every_tile_class_exact(gpu, orc):
    """skinny_dbatch_set_class accepts exactly the tile classes that can
    compute an op, and every accepted class gives the oracle's result bitwise
    on integer data (TSMTTSM general and Gram, TSMM out of place and in place,
    beta 0 and 1); a refused class leaves the plan as it was."""
    import torch
    rng = random.Random(11)
    cases = [("tsmttsm", M, N, g, False, 0.0) for (M, N, g) in TS_SHAPES]
    cases += [("tsmm", M, N, False, ip, beta) for (M, N, ip, beta) in TM_SHAPES]
    for kind, M, N, gram, inplace, beta in cases:
        K = rng.randint(3000, 9000)
        op, C0, want, dC = _one_op(orc, rng, kind, M, N, K, gram, inplace, beta)
        b = gpu.Batch([op], item_bytes=1 << 16)
        default = b.classes()[0]
        ncls = 18 if kind == "tsmttsm" else 17
        tried = 0
        for cls in list(range(ncls)) + [-1]:
            valid = cls < 0 or (_ts_valid(cls, M, N, gram) if kind == "tsmttsm"
                                else _tm_valid(cls, M, N))
            note(f"batch class: {kind} M={M} N={N} gram={gram} inplace={inplace} "
                 f"beta={beta} class={cls}")
            before = b.classes()[0]
            ok = b.set_class(0, cls)
            assert ok == valid, (kind, M, N, cls)
            if not ok:
                assert b.classes()[0] == before
                continue
            assert b.classes()[0] == (default if cls < 0 else cls)
            dC.copy_(torch.from_numpy(C0))
            torch.cuda.synchronize()
            b.run()
            assert np.array_equal(host(dC), want), (kind, M, N, gram, inplace, beta, cls)
            tried += 1
        assert tried >= 2
        b.close()


def test_batch_tune_classes_valid_and_kept(gpu, orc):
    """After skinny_dbatch_tune every op runs a class that can compute it, a
    class the caller set is kept, and the results stay exact."""
    import torch
    rng = random.Random(21)
    ops, refs = [], []
    for (M, N, g) in TS_SHAPES[:8]:
        op, C0, want, dC = _one_op(orc, rng, "tsmttsm", M, N, 40000, g)
        ops.append(op)
        refs.append((C0, want, dC))
    for (M, N, ip, beta) in TM_SHAPES[:5]:
        op, C0, want, dC = _one_op(orc, rng, "tsmm", M, N, 40000, False, ip, beta)
        ops.append(op)
        refs.append((C0, want, dC))
    b = gpu.Batch(ops)
    assert b.set_class(4, 6)  # TSMTTSM 8 x 8 on the unpacked 16 x 16 tile
    b.tune(reps=1)
    cls = b.classes()
    assert cls[4] == 6
    for i, o in enumerate(ops):
        kind, M, N = o[0], o[1], o[2]
        gram = kind == "tsmttsm" and o[4] is o[7]
        assert (_ts_valid(cls[i], M, N, gram) if kind == "tsmttsm"
                else _tm_valid(cls[i], M, N)), (i, kind, M, N, cls[i])
    for C0, want, dC in refs:
        dC.copy_(torch.from_numpy(C0))
    torch.cuda.synchronize()
    b.run()
    for (C0, want, dC), o in zip(refs, ops):
        assert np.array_equal(host(dC), want), o[:4]
    b.close()
