"""TSMM parity on the GPU: libskinny (C ABI) against the CPU oracle.

Follows test.cu:245-293 for TSMM: integer data (bitwise comparison), padded
leading dimensions, random blockCount, beta in {0, 1}, and the in-place case
C == A with ldc = max(N + U[0,4], M) (test.cu:267) where the oracle reads the
old A as tsmm/cublas.cuh:27-33 does.
"""
import json
import os
import random

import numpy as np
import pytest

from gpu_util import EPS32, EPS64, dev, host, note, rel_fro

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _run(sk, bc, M, N, K, A, lda, alpha, B, ldb, beta, C, ldc, inplace=False):
    dC = dev(C)
    dA = dC if inplace else dev(A)
    dB = dev(B)
    ok = sk.tsmm(bc, M, N, K, dA, lda, alpha, dB, ldb, beta, dC, ldc)
    return ok, host(dC)


def test_golden_integer(gpu, orc):
    with open(os.path.join(GOLD, "tsmm_int.json")) as f:
        cases = json.load(f)
    for c in cases:
        for bc in (0, 1, 3, 200):
            note(f"tsmm golden M={c['M']} N={c['N']} K={c['K']} inplace={c['inplace']} bc={bc}")
            B = np.array(c["B"], dtype=np.float64)
            C = np.array(c["C0"], dtype=np.float64)
            A = C if c["inplace"] else np.array(c["A"], dtype=np.float64)
            ok, got = _run(gpu, bc, c["M"], c["N"], c["K"], A, c["lda"], c["alpha"], B,
                           c["ldb"], c["beta"], C, c["ldc"], c["inplace"])
            assert ok
            assert np.array_equal(got, np.array(c["C"], dtype=np.float64))


def _fuzz_one(sk, orc, rng, M, N, dtype, Kmax, inplace=None):
    if inplace is None:
        inplace = rng.random() < 0.5
    beta = float(rng.randint(0, 1))
    ldb = M + rng.randint(0, 4)
    if inplace:
        ldc = max(N + rng.randint(0, 4), M)
        lda = ldc
    else:
        lda = M + rng.randint(0, 4)
        ldc = N + rng.randint(0, 4)
    K = rng.randint(1, Kmax)
    bc = rng.randint(1, 200)
    seed = rng.randint(0, 1 << 30)
    note(f"tsmm fuzz {np.dtype(dtype).name} M={M} N={N} K={K} lda={lda} ldb={ldb} "
         f"ldc={ldc} bc={bc} inplace={inplace} beta={beta} seed={seed}")
    B = orc.integers(N * ldb, seed, 1).astype(dtype)
    C = orc.integers(K * ldc, seed, 2).astype(dtype)
    A = C.copy() if inplace else orc.integers(K * lda, seed, 0).astype(dtype)
    want = C.copy()
    orc.tsmm(M, N, K, want if inplace else A, lda, 1.0, B, ldb, beta, want, ldc)
    ok, got = _run(sk, bc, M, N, K, A, lda, 1.0, B, ldb, beta, C, ldc, inplace)
    assert ok
    assert np.array_equal(got, want), (M, N, K, lda, ldb, ldc, bc, inplace, beta)


@pytest.mark.parametrize("M", list(range(1, 17)) + [20, 24, 31, 32, 33, 40, 48, 63, 64])
def test_fuzz_square_fp64(gpu, orc, M):
    rng = random.Random(2000 + M)
    for _ in range(6):
        _fuzz_one(gpu, orc, rng, M, M, np.float64, 30000)


def test_fuzz_rectangular_fp64(gpu, orc):
    rng = random.Random(11)
    for _ in range(40):
        _fuzz_one(gpu, orc, rng, rng.randint(1, 64), rng.randint(1, 64), np.float64, 20000)


@pytest.mark.parametrize("inplace", [False, True])
def test_narrow_grid_fp64(gpu, orc, inplace):
    """Every (M, N) in 1..9 x 1..5, both sides of the switch between
    whole-row threads and the matrix cores (which depends on M, N, in place
    and beta), beta in {0, 1}, the default grid and a random one."""
    rng = random.Random(77 + inplace)
    for M in range(1, 10):
        for N in range(1, 6):
            for beta in (0.0, 1.0):
                if inplace:
                    ldc = max(M, N) + rng.randint(0, 2)
                    lda = ldc
                else:
                    lda, ldc = M + rng.randint(0, 2), N + rng.randint(0, 2)
                K = rng.randint(1000, 5000)
                seed = rng.randint(0, 1 << 30)
                for bc in (0, rng.randint(1, 300)):
                    note(f"tsmm narrow M={M} N={N} K={K} lda={lda} ldc={ldc} "
                         f"inplace={inplace} beta={beta} bc={bc} seed={seed}")
                    B = orc.integers(N * M, seed, 1)
                    C = orc.integers(K * ldc, seed, 2)
                    A = C.copy() if inplace else orc.integers(K * lda, seed, 0)
                    want = C.copy()
                    orc.tsmm(M, N, K, want if inplace else A, lda, 1.0, B, M, beta,
                             want, ldc)
                    ok, got = _run(gpu, bc, M, N, K, A, lda, 1.0, B, M, beta, C, ldc,
                                   inplace)
                    assert ok
                    assert np.array_equal(got, want), (M, N, K, lda, ldc, bc, beta)


@pytest.mark.parametrize("inplace", [False, True])
@pytest.mark.parametrize("off", [0, 1])
def test_narrow_grid_fp32(gpu, orc, inplace, off):
    """FP32 around the switch between whole-row threads (16-byte row accesses
    when the leading dimensions are multiples of 4 and the rows 16-byte
    aligned) and the matrix cores: every M in 1..12 with N in {1..6, 8, 9},
    beta in {0, 1}, the default grid and a random one.  off = 0 keeps A and C
    16-byte aligned with leading dimensions rounded up to 4; off = 1 starts
    them one float into their buffers, with other leading dimensions."""
    import torch
    rng = random.Random(91 + 2 * off + inplace)
    for M in range(1, 13):
        for N in (1, 2, 3, 4, 5, 6, 8, 9):
            for beta in (0.0, 1.0):
                if off == 0:
                    r4 = lambda x: (x + 3) // 4 * 4
                    ldc = r4(max(M, N)) if inplace else r4(N)
                    lda = ldc if inplace else r4(M) + 4 * rng.randint(0, 1)
                elif inplace:
                    ldc = max(M, N) + rng.randint(0, 2)
                    lda = ldc
                else:
                    lda, ldc = M + rng.randint(0, 2), N + rng.randint(0, 2)
                K = rng.randint(1000, 5000)
                seed = rng.randint(0, 1 << 30)
                for bc in (0, rng.randint(1, 300)):
                    note(f"tsmm fp32 narrow M={M} N={N} K={K} lda={lda} ldc={ldc} off={off} "
                         f"inplace={inplace} beta={beta} bc={bc} seed={seed}")
                    B = orc.integers(N * M, seed, 1).astype(np.float32)
                    C = orc.integers(K * ldc, seed, 2).astype(np.float32)
                    A = C.copy() if inplace else orc.integers(K * lda, seed, 0).astype(np.float32)
                    want = C.copy()
                    orc.tsmm(M, N, K, want if inplace else A, lda, 1.0, B, M, beta,
                             want, ldc)
                    bufC = torch.zeros(C.size + 4, dtype=torch.float32, device="cuda")
                    bufC[off:off + C.size] = torch.from_numpy(C).cuda()
                    dC = bufC[off:off + C.size]
                    if inplace:
                        dA = dC
                    else:
                        bufA = torch.zeros(A.size + 4, dtype=torch.float32, device="cuda")
                        bufA[off:off + A.size] = torch.from_numpy(A).cuda()
                        dA = bufA[off:off + A.size]
                    assert gpu.tsmm(bc, M, N, K, dA, lda, 1.0, dev(B), M, beta, dC, ldc)
                    assert np.array_equal(host(dC), want), (M, N, K, lda, ldc, bc, beta)


@pytest.mark.parametrize("M,N", [(1, 1), (4, 4), (8, 8), (5, 11), (16, 16), (32, 32), (64, 64)])
def test_fuzz_fp32(gpu, orc, M, N):
    rng = random.Random(M * 17 + N)
    for _ in range(4):
        _fuzz_one(gpu, orc, rng, M, N, np.float32, 30000)


@pytest.mark.parametrize("M", [1, 2, 4, 8, 16, 32, 64])
def test_uniform_fp64(gpu, orc, M):
    K = 1 << 15
    note(f"tsmm uniform M=N={M} K={K}")
    A = orc.uniform(K * M, 5, 0)
    B = orc.uniform(M * M, 5, 1)
    C = orc.uniform(K * M, 5, 2)
    want = C.copy()
    orc.tsmm(M, M, K, A, M, 1.5, B, M, -0.5, want, M)
    ok, got = _run(gpu, 0, M, M, K, A, M, 1.5, B, M, -0.5, C, M)
    assert ok
    assert rel_fro(got, want) <= 8 * M * EPS64


def test_beta0_does_not_read_C(gpu, orc):
    M, N, K = 5, 3, 700
    A = orc.integers(K * M, 8, 0)
    B = orc.integers(N * M, 8, 1)
    want = np.zeros(K * N)
    orc.tsmm(M, N, K, A, M, 1.0, B, M, 0.0, want, N)
    ok, got = _run(gpu, 0, M, N, K, A, M, 1.0, B, M, 0.0, np.full(K * N, np.nan), N)
    assert ok and np.array_equal(got, want)


def test_unsupported_overlap_returns_false(gpu):
    import torch
    buf = torch.zeros(1000, dtype=torch.float64, device="cuda")
    B = torch.zeros(16, dtype=torch.float64, device="cuda")
    # A == C with lda != ldc: the reference variants return false (SKIP)
    assert gpu.dtsmm(0, 4, 4, 10, buf, 8, 1.0, B, 4, 0.0, buf, 4) is False
    # partially overlapping A and C
    assert gpu.dtsmm(0, 4, 4, 10, buf, 4, 1.0, B, 4, 0.0, buf[2:], 4) is False


@pytest.mark.parametrize("M,N", [(7, 1), (12, 4), (5, 5), (10, 8), (40, 3), (64, 2)])
def test_transposed_tile_alpha_beta(gpu, orc, M, N):
    """Narrow C on the matrix cores' transposed tile (N <= 4, or N <= 8 with
    beta != 0): alpha = 1.5, beta in {0, -0.5} on U[-1,1) data, out of
    place, within the FP64 tolerance of the oracle."""
    K = (1 << 15) + 7
    for beta in (0.0, -0.5):
        note(f"tsmm transposed tile M={M} N={N} K={K} beta={beta}")
        A = orc.uniform(K * M, 31, M)
        B = orc.uniform(M * N, 32, N)
        C = orc.uniform(K * N, 33, 0)
        want = C.copy()
        orc.tsmm(M, N, K, A, M, 1.5, B, M, beta, want, N)
        ok, got = _run(gpu, 0, M, N, K, A, M, 1.5, B, M, beta, C, N)
        assert ok
        assert rel_fro(got, want) <= 8 * (M + 1) * EPS64


@pytest.mark.parametrize("M,N", [(4, 4), (8, 3), (16, 16), (33, 20), (64, 64)])
def test_inplace_beta_uniform(gpu, orc, M, N):
    """In place with beta != 0 and N <= M: beta * C joins the operand as
    beta * I (C's old columns are A's own); non-integer alpha, beta within
    the FP64 tolerance of the oracle's beta C + alpha A B."""
    K = 1 << 14
    note(f"tsmm in place beta uniform M={M} N={N} K={K}")
    C = orc.uniform(K * M, 6, 0)
    B = orc.uniform(M * N, 6, 1)
    want = C.copy()
    orc.tsmm(M, N, K, want, M, 0.75, B, M, -1.25, want, M)
    ok, got = _run(gpu, 0, M, N, K, C, M, 0.75, B, M, -1.25, C, M, inplace=True)
    assert ok
    assert rel_fro(got, want) <= 8 * (M + 1) * EPS64


# A 57..64 columns wide with C 49..64 (the widest matrix-core class): full
# width with aligned rows, ragged widths and odd leading dimensions, beta !=
# 0, in place (beta folded into the operand) and ragged K
@pytest.mark.parametrize("M,N,lda_pad,beta,inplace", [
    (64, 64, 0, 0.0, False), (64, 64, 0, 1.0, False), (64, 64, 1, 0.0, False),
    (64, 64, 0, 1.0, True), (64, 64, 0, 0.0, True), (60, 64, 0, 0.0, False),
    (64, 50, 2, 1.0, False), (57, 49, 0, 1.0, True), (63, 63, 3, 0.0, True)])
def test_wide64_tsmm(gpu, orc, M, N, lda_pad, beta, inplace):
    K = 4099
    ld = max(M, N) + lda_pad if inplace else M + lda_pad
    ldc = ld if inplace else N + 1
    note(f"tsmm wide64 M={M} N={N} lda={ld} beta={beta} inplace={inplace}")
    A = orc.integers(K * ld, 91, 0)
    Bm = orc.integers(M * N, 91, 1)
    C0 = A.copy() if inplace else orc.integers(K * ldc, 91, 2)
    want = C0.copy()
    orc.tsmm(M, N, K, A.copy() if inplace else A, ld, 1.0, Bm, M, beta, want, ldc)
    dA = dev(A)
    dC = dA if inplace else dev(C0)
    assert gpu.dtsmm(0, M, N, K, dA, ld, 1.0, dev(Bm), M, beta, dC, ldc)
    assert np.array_equal(host(dC), want), (M, N, ld, beta, inplace)
