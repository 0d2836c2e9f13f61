"""Ozaki-split TSMTTSM on the INT8 matrix cores.

* the INT8 32x32x32 result layout the kernel assumes, read back from the
  device;
* blockCount = 1: bitwise equal to the oracle's restatement of the same
  sequence (limbs, INT32 levels, flush points, FP64 reconstruction order),
  including ragged K, padded shapes and magnitudes that grow along K;
* default blockCount: error against the exact product within the scheme's
  bound and within 8 K eps at K = 2^24 (3e-8) from 4 limbs on.
"""
import numpy as np
import pytest

from gpu_util import dev, host, note, rel_fro

pytestmark = pytest.mark.gpu


def test_i8_layout(gpu):
    rows, cols = gpu.probe_i8_layout()
    for lane in range(64):
        for r in range(16):
            assert cols[lane, r] == lane % 32
            assert rows[lane, r] == 8 * (r // 4) + 4 * (lane // 32) + (r % 4), (lane, r, rows[lane])


@pytest.mark.parametrize("M,N,K,S", [(5, 3, 1000, 4), (32, 32, 3001, 3), (128, 128, 2000, 4),
                                     (130, 70, 999, 5), (1, 1, 100, 4)])
def test_bitwise_blockcount1(gpu, orc, M, N, K, S):
    note(f"ozaki bitwise M={M} N={N} K={K} S={S}")
    A = orc.uniform(K * M, M + N, 0)
    B = orc.uniform(K * N, M + N, 1)
    C0 = orc.uniform(M * N, M + N, 2)
    want = C0.copy()
    orc.ozaki_tsmttsm(M, N, K, A, M, 1.5, B, N, -1.0, want, M, S)
    dC = dev(C0.copy())
    assert gpu.dtsmttsm_ozaki(1, M, N, K, dev(A), M, 1.5, dev(B), N, -1.0, dC, M, slices=S)
    assert np.array_equal(host(dC), want)


def test_bitwise_growth(gpu, orc):
    M, N, K = 16, 9, 40000
    A = orc.uniform(K * M, 9, 0)
    B = orc.uniform(K * N, 9, 1)
    scale = np.ldexp(1.0, np.arange(K) // 3000)
    A = (A.reshape(K, M) * scale[:, None]).ravel()
    want = np.zeros(M * N)
    orc.ozaki_tsmttsm(M, N, K, A, M, 1.0, B, N, 0.0, want, M, 4)
    dC = dev(np.zeros(M * N))
    assert gpu.dtsmttsm_ozaki(1, M, N, K, dev(A), M, 1.0, dev(B), N, 0.0, dC, M, slices=4)
    assert np.array_equal(host(dC), want)


def test_bitwise_outlier(gpu, orc):
    """An outlier chunk followed by ordinary data: the growth flush and the
    scale restart of the next chunk, bitwise against the oracle."""
    M, N, K = 6, 5, 4000
    note(f"ozaki outlier M={M} N={N} K={K}")
    A = orc.uniform(K * M, 21, 0).copy()
    B = orc.uniform(K * N, 21, 1).copy()
    A[100 * M] = 1e5
    B[2500 * N + 3] = -3e7
    for S in (3, 4):
        want = np.zeros(M * N)
        orc.ozaki_tsmttsm(M, N, K, A, M, 1.0, B, N, 0.0, want, M, S)
        dC = dev(np.zeros(M * N))
        assert gpu.dtsmttsm_ozaki(1, M, N, K, dev(A), M, 1.0, dev(B), N, 0.0, dC, M, slices=S)
        assert np.array_equal(host(dC), want), S


@pytest.mark.parametrize("S", [3, 4, 5])
def test_c3_error(gpu, orc, S):
    M = 128
    K = 1 << 17
    note(f"ozaki c3 M={M} K={K} S={S}")
    A = orc.c3_columns(M, K, 5)
    ms = np.array([0, 17, 64, 127, 5, 99])
    ns = np.array([0, 3, 64, 127, 120, 99])
    hi, lo = orc.xtsmttsm_entries(K, A, M, A, M, ms, ns)
    dA = dev(A)
    dC = dev(np.zeros(M * M))
    assert gpu.dtsmttsm_ozaki(0, M, M, K, dA, M, 1.0, dA, M, 0.0, dC, M, slices=S)
    got = host(dC).reshape(M, M).T  # column-major C: got[m, n]
    err = np.sqrt(np.sum((got[ms, ns] - hi - lo) ** 2)) / np.sqrt(np.sum(hi ** 2))
    assert err <= 8 * (S + 1) * 2.0 ** (-7 * S)
    assert err <= 8 * 2 ** 24 * np.finfo(float).eps
