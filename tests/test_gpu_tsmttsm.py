"""TSMTTSM parity on the GPU: libskinny (C ABI) against the CPU oracle.

Follows test.cu:245-293: integer data in [-2, 2] (every sum exact, so the
comparison is bitwise whatever the summation order), padded leading
dimensions ld = dim + U[0, 4], blockCount ~ U[1, 200], self (A == B) and
beta in {0, 1}.  Floating-point data is checked against the exact
(superaccumulator) oracle within 8 * K * eps relative Frobenius error.
"""
import json
import os
import random

import numpy as np
import pytest

from gpu_util import EPS32, EPS64, dev, host, note, rel_fro

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _run(sk, blockCount, M, N, K, A, lda, alpha, B, ldb, beta, C, ldc, self_=False):
    dA = dev(A)
    dB = dA if self_ else dev(B)
    dC = dev(C)
    ok = sk.tsmttsm(blockCount, M, N, K, dA, lda, alpha, dB, ldb, beta, dC, ldc)
    return ok, host(dC)


def test_golden_integer(gpu, orc):
    with open(os.path.join(GOLD, "tsmttsm_int.json")) as f:
        cases = json.load(f)
    for c in cases:
        for bc in (0, 1, 7, 200):
            note(f"tsmttsm golden M={c['M']} N={c['N']} K={c['K']} bc={bc}")
            A = np.array(c["A"], dtype=np.float64)
            B = A if c["self"] else np.array(c["B"], dtype=np.float64)
            C = np.array(c["C0"], dtype=np.float64)
            ok, got = _run(gpu, bc, c["M"], c["N"], c["K"], A, c["lda"], c["alpha"], B,
                           c["ldb"], c["beta"], C, c["ldc"], c["self"])
            assert ok
            assert np.array_equal(got, np.array(c["C"], dtype=np.float64)), (c["M"], c["N"], bc)


def _fuzz_one(sk, orc, rng, M, N, dtype, Kmax):
    self_ = rng.random() < 0.5 and M == N
    beta = float(rng.randint(0, 1))
    lda = M + rng.randint(0, 4)
    ldb = lda if self_ else N + rng.randint(0, 4)
    ldc = M + rng.randint(0, 4)
    K = rng.randint(1, Kmax)
    bc = rng.randint(1, 200)
    seed = rng.randint(0, 1 << 30)
    note(f"tsmttsm fuzz {np.dtype(dtype).name} M={M} N={N} K={K} lda={lda} ldb={ldb} "
         f"ldc={ldc} bc={bc} self={self_} beta={beta} seed={seed}")
    A = orc.integers(K * lda, seed, 0).astype(dtype)
    B = A if self_ else orc.integers(K * ldb, seed, 1).astype(dtype)
    C = orc.integers(N * ldc, seed, 2).astype(dtype)
    want = C.copy()
    orc.tsmttsm(M, N, K, A, lda, 1.0, B, ldb, beta, want, ldc)
    ok, got = _run(sk, bc, M, N, K, A, lda, 1.0, B, ldb, beta, C, ldc, self_)
    assert ok
    assert np.array_equal(got, want), (M, N, K, lda, ldb, ldc, bc, self_, beta)


@pytest.mark.parametrize("M", list(range(1, 17)) + [20, 24, 31, 32, 33, 40, 48, 63, 64])
def test_fuzz_square_fp64(gpu, orc, M):
    rng = random.Random(1000 + M)
    for _ in range(6):
        _fuzz_one(gpu, orc, rng, M, M, np.float64, 40000)


def test_fuzz_rectangular_fp64(gpu, orc):
    rng = random.Random(7)
    for _ in range(40):
        M = rng.randint(1, 64)
        N = rng.randint(1, 64)
        _fuzz_one(gpu, orc, rng, M, N, np.float64, 20000)


@pytest.mark.parametrize("M,N", [(65, 65), (96, 96), (128, 128), (100, 7), (3, 130), (256, 256)])
def test_fuzz_wide_fp64(gpu, orc, M, N):
    rng = random.Random(M * 7 + N)
    for _ in range(2):
        _fuzz_one(gpu, orc, rng, M, N, np.float64, 3000)


# 64 x 64 tiles run on the 4x4x4 matrix-core kernel: full and ragged tiles,
# even and odd leading dimensions (16-byte pair loads or element loads), the
# Gram product, one workgroup (in-kernel epilogue) and several grid tiles
@pytest.mark.parametrize("M,N,ld_pad,self_,bc", [
    (64, 64, 0, False, 0), (64, 64, 1, False, 37), (64, 64, 0, True, 0), (64, 64, 1, True, 5),
    (50, 50, 0, False, 1), (57, 61, 3, False, 64), (49, 49, 2, True, 1), (63, 63, 1, True, 200),
    (64, 52, 0, False, 3), (128, 64, 0, False, 0), (128, 128, 1, True, 9), (130, 70, 2, False, 1)])
def test_q4_tiles(gpu, orc, M, N, ld_pad, self_, bc):
    K = 4099
    lda = M + ld_pad
    ldb = lda if self_ else N + ld_pad
    ldc = M + 1
    note(f"tsmttsm q4 M={M} N={N} lda={lda} self={self_} bc={bc}")
    A = orc.integers(K * lda, 77, 0)
    B = A if self_ else orc.integers(K * ldb, 77, 1)
    for beta in (0.0, 1.0):
        C = orc.integers(N * ldc, 77, 2)
        want = C.copy()
        orc.tsmttsm(M, N, K, A, lda, 1.0, B, ldb, beta, want, ldc)
        ok, got = _run(gpu, bc, M, N, K, A, lda, 1.0, B, ldb, beta, C, ldc, self_)
        assert ok
        assert np.array_equal(got, want), (M, N, lda, self_, bc, beta)


@pytest.mark.parametrize("M,N", [(1, 1), (4, 4), (8, 8), (5, 11), (16, 16), (32, 32), (64, 64)])
def test_fuzz_fp32(gpu, orc, M, N):
    rng = random.Random(M * 31 + N)
    for _ in range(4):
        _fuzz_one(gpu, orc, rng, M, N, np.float32, 30000)


@pytest.mark.parametrize("M", [1, 2, 4, 8, 16, 32, 64])
def test_uniform_fp64_vs_exact(gpu, orc, M):
    """U[-1,1) data, K = 2^16: <= 8 K eps relative Frobenius vs the exact result."""
    K = 1 << 16
    note(f"tsmttsm uniform M=N={M} K={K}")
    A = orc.uniform(K * M, 1, 0)
    B = orc.uniform(K * M, 1, 1)
    hi, lo = orc.xtsmttsm(M, M, K, A, M, B, M)
    ok, got = _run(gpu, 0, M, M, K, A, M, 1.0, B, M, 0.0, np.zeros(M * M), M)
    assert ok
    err = rel_fro(got - lo, hi)
    assert err <= 8 * K * EPS64, err
    # fixed blockCount: bitwise reproducible
    ok, again = _run(gpu, 0, M, M, K, A, M, 1.0, B, M, 0.0, np.zeros(M * M), M)
    assert np.array_equal(got, again)


@pytest.mark.parametrize("M", [2, 17, 32, 33, 48, 64])
def test_gram_mirror_matches_general(gpu, orc, M):
    """A Gram product (B is A) accumulates only the subtiles on and above the
    diagonal and mirrors them.  On uniform data the result is exactly
    symmetric and meets the exact oracle's bound; with the unroll and grid
    pinned through the tuning table it equals the general kernel on a copy
    of A bitwise (the same products summed in the same order) where both run
    on the 16x16x4 instruction.  From 49 columns the general product runs on
    the 4x4x4 instruction, whose summation order differs: there the two
    agree within the bound only."""
    import torch
    K = 1 << 16
    A = orc.uniform(K * M, 12, M)
    dA, dA2 = dev(A), dev(A.copy())
    hi, lo = orc.xtsmttsm(M, M, K, A, M, A, M)
    h = gpu.Handle()
    try:
        for uvar, G in ((None, None), (0, 37), (1, 37), (1, 300)):
            if uvar is not None:
                gpu.tune_clear(handle=h)
                gpu.tune_import(f"0 0 {M} {M} 16 {uvar} {G}\n", handle=h)
            note(f"tsmttsm gram mirror M={M} K={K} uvar={uvar} G={G}")
            out = []
            for dB in (dA, dA2):
                dC = dev(np.zeros(M * M))
                assert gpu.dtsmttsm(0, M, M, K, dA, M, 1.0, dB, M, 0.0, dC, M, handle=h)
                torch.cuda.synchronize()
                out.append(host(dC))
            gram, gen = out
            G2 = gram.reshape(M, M)
            assert np.array_equal(G2, G2.T)
            assert rel_fro(gram - lo, hi) <= 8 * K * EPS64
            if uvar is not None and M < 49:
                assert np.array_equal(gram, gen)
            assert rel_fro(gen - lo, hi) <= 8 * K * EPS64
    finally:
        h.close()


@pytest.mark.parametrize("fp", [np.float64, np.float32])
def test_narrow_asymmetric(gpu, orc, fp):
    """One width 1..8 and the other 5..16 (both orientations), where the
    matrix-core tiles are mostly masked lanes; FP64 products with a 9..16
    wide operand of even leading dimension pair rows in 16-byte loads.
    Padded and unpadded leading dimensions (odd ones take the general
    kernel), integer data, bitwise against the oracle, with the default grid
    and random ones, beta in {0, 1}."""
    rng = random.Random(321 + (fp is np.float32))
    shapes = [(lo, hi) for lo in (1, 2, 3, 4) for hi in (5, 8, 9, 12, 13, 16)]
    shapes += [(lo, hi) for lo in (5, 7, 8) for hi in (9, 10, 16)]
    shapes += [(hi, lo) for lo, hi in shapes]
    for M, N in shapes:
        cls = lambda w: 1 if w <= 1 else 2 if w <= 2 else 4 if w <= 4 else 8 if w <= 8 \
            else 12 if w <= 12 else 16
        for lda, ldb in ((M, N), (cls(M) + 4, cls(N) + rng.randint(0, 4)),
                         (cls(M) + 5, N + 1)):
            K = rng.randint(1, 3000) if rng.random() < 0.3 else rng.randint(3000, 70000)
            beta = float(rng.randint(0, 1))
            bc = rng.choice((0, rng.randint(1, 300)))
            seed = rng.randint(0, 1 << 30)
            note(f"tsmttsm narrow {np.dtype(fp).name} M={M} N={N} K={K} lda={lda} ldb={ldb} "
                 f"bc={bc} beta={beta} seed={seed}")
            A = orc.integers(K * lda, seed, 0).astype(fp)
            B = orc.integers(K * ldb, seed, 1).astype(fp)
            C = orc.integers(N * M, seed, 2).astype(fp)
            want = C.copy()
            orc.tsmttsm(M, N, K, A, lda, 1.0, B, ldb, beta, want, M)
            ok, got = _run(gpu, bc, M, N, K, A, lda, 1.0, B, ldb, beta, C, M)
            assert ok
            assert np.array_equal(got, want), (M, N, K, lda, ldb, bc, beta)


@pytest.mark.parametrize("M,N", [(12, 1), (16, 8), (9, 3), (3, 13), (6, 1), (1, 8), (5, 4)])
def test_narrow_asymmetric_alpha_beta(gpu, orc, M, N):
    """Paired-row shapes with alpha = 1.5, beta = -0.5 on U[-1,1) data: within
    8 K eps of the exact oracle (the epilogue is beta C + alpha sum)."""
    K = 1 << 16
    for lda, ldb in ((M, N), (M + 1, N + 2)):
        note(f"tsmttsm narrow alpha/beta M={M} N={N} K={K} lda={lda} ldb={ldb}")
        A = orc.uniform(K * lda, 21, M)
        B = orc.uniform(K * ldb, 22, N)
        C0 = orc.uniform(M * N, 23, 0)
        hi, lo = orc.xtsmttsm(M, N, K, A, lda, B, ldb)
        want = 1.5 * (hi + lo) - 0.5 * C0
        ok, got = _run(gpu, 0, M, N, K, A, lda, 1.5, B, ldb, -0.5, C0.copy(), M)
        assert ok
        assert rel_fro(got, want) <= 8 * K * EPS64


def test_c1_config(gpu, orc):
    """K = 2^16, M = N = 8, alpha = 1, beta = 0, seed 1 (streams 0 and 1)."""
    M = N = 8
    K = 1 << 16
    A = orc.uniform(K * M, 1, 0)
    B = orc.uniform(K * N, 1, 1)
    want = np.zeros(M * N)
    orc.tsmttsm(M, N, K, A, M, 1.0, B, N, 0.0, want, M)
    hi, lo = orc.xtsmttsm(M, N, K, A, M, B, N)
    ok, got = _run(gpu, 0, M, N, K, A, M, 1.0, B, N, 0.0, np.zeros(M * N), M)
    assert ok
    assert rel_fro(got, want) <= 8 * K * EPS64
    assert rel_fro(got - lo, hi) <= 8 * K * EPS64


def test_alpha_beta_and_nan_beta0(gpu, orc):
    M, N, K = 6, 5, 1000
    A = orc.integers(K * M, 4, 0)
    B = orc.integers(K * N, 4, 1)
    C = orc.integers(M * N, 4, 2)
    want = C.copy()
    orc.tsmttsm(M, N, K, A, M, -2.0, B, N, 3.0, want, M)
    ok, got = _run(gpu, 0, M, N, K, A, M, -2.0, B, N, 3.0, C.copy(), M)
    assert ok and np.array_equal(got, want)
    # beta == 0: C is not read (NaN in C must not leak)
    ok, got = _run(gpu, 0, M, N, K, A, M, 1.0, B, N, 0.0, np.full(M * N, np.nan), M)
    want = np.zeros(M * N)
    orc.tsmttsm(M, N, K, A, M, 1.0, B, N, 0.0, want, M)
    assert ok and np.array_equal(got, want)


def test_edge_sizes(gpu, orc):
    # K = 0: C = beta * C;  K = 1;  M = N = 0 is a no-op
    C = np.arange(9, dtype=np.float64)
    ok, got = _run(gpu, 0, 3, 3, 0, np.zeros(3), 3, 1.0, np.zeros(3), 3, 2.0, C.copy(), 3)
    assert ok and np.array_equal(got, 2.0 * C)
    ok, got = _run(gpu, 5, 3, 3, 1, np.array([1.0, 2, 3]), 3, 1.0, np.array([1.0, 0, -1]), 3, 0.0, np.zeros(9), 3)
    assert ok and np.array_equal(got, np.array([1.0, 2, 3, 0, 0, 0, -1, -2, -3]))
    ok, _ = _run(gpu, 0, 0, 3, 10, np.zeros(30), 1, 1.0, np.zeros(30), 3, 0.0, np.zeros(3), 1)
    assert ok


def test_invalid_arguments_raise(gpu):
    import torch
    A = torch.zeros(100, dtype=torch.float64, device="cuda")
    with pytest.raises(gpu.SkinnyError):
        gpu.dtsmttsm(0, 4, 4, 10, A, 3, 1.0, A, 4, 0.0, A, 4)  # lda < M
    with pytest.raises(gpu.SkinnyError):
        gpu.dtsmttsm(0, -1, 4, 10, A, 4, 1.0, A, 4, 0.0, A, 4)


def test_many_blockcounts_reproducible(gpu, orc):
    M = N = 16
    K = 100000
    A = orc.uniform(K * M, 3, 0)
    B = orc.uniform(K * N, 3, 1)
    hi, lo = orc.xtsmttsm(M, N, K, A, M, B, N)
    for bc in (1, 2, 17, 64, 255, 256, 1000, 4096):
        note(f"tsmttsm blockCount={bc}")
        ok, got = _run(gpu, bc, M, N, K, A, M, 1.0, B, N, 0.0, np.zeros(M * N), M)
        assert ok
        assert rel_fro(got - lo, hi) <= 8 * K * EPS64


def test_concurrent_streams_share_a_handle(gpu, orc):
    """Calls on two streams through one handle overlap without sharing
    split-K partials or arrival counters (each stream has its own
    workspace): both results stay bit-exact on integer data."""
    import torch
    M = N = 16
    K = 1 << 22
    note("tsmttsm two streams, one handle")
    outs = []
    streams = [torch.cuda.Stream(), torch.cuda.Stream()]
    data = []
    for i in range(2):
        A = torch.empty(K * M, dtype=torch.float64, device="cuda")
        B = torch.empty(K * N, dtype=torch.float64, device="cuda")
        gpu.fill_int(A, A.numel(), 40 + i, 0)
        gpu.fill_int(B, B.numel(), 40 + i, 1)
        data.append((A, B, torch.zeros(M * N, dtype=torch.float64, device="cuda")))
    torch.cuda.synchronize()
    for rep in range(3):
        for (A, B, C), s in zip(data, streams):
            with torch.cuda.stream(s):
                assert gpu.dtsmttsm(512, M, N, K, A, M, 1.0, B, N, 0.0, C, M)
        torch.cuda.synchronize()
        for A, B, C in data:
            ref = A.view(K, M).t().mm(B.view(K, N))
            assert torch.equal(C.view(N, M).t(), ref), rep
