"""Pin the oracle's Ozaki-split restatement on the CPU.

Checks: exactness on data the limbs represent exactly (small integers, so
every product and the reconstruction are exact), error against the exact
superaccumulator decreasing with the limb count and within the scheme's
bound on U[-1,1) and on cancelling columns (within the 8 K eps tolerance at
K = 2^24 from 3 limbs on, thanks to rounded limbs and the level s + t =
S + 2), alpha/beta handling with growing magnitudes, and outliers.
"""
import numpy as np
import pytest


def test_integers_exact(orc):
    M, N, K = 7, 5, 700
    A = orc.integers(K * M, 1, 0, -100, 100)
    B = orc.integers(K * N, 1, 1, -100, 100)
    want = np.zeros(M * N)
    orc.tsmttsm(M, N, K, A, M, 1.0, B, N, 0.0, want, M)
    for S in (3, 4, 5):
        got = np.zeros(M * N)
        orc.ozaki_tsmttsm(M, N, K, A, M, 1.0, B, N, 0.0, got, M, S)
        assert np.array_equal(got, want), S


@pytest.mark.parametrize("data", ["uniform", "c3"])
def test_error_vs_exact(orc, data):
    M = N = 24
    K = 3000
    if data == "uniform":
        A = orc.uniform(K * M, 2, 0)
        B = orc.uniform(K * N, 2, 1)
    else:
        A = orc.c3_columns(M, K, 3)
        B = A
    hi, lo = orc.xtsmttsm(M, N, K, A, M, B, N)
    errs = []
    for S in (3, 4, 5):
        got = np.zeros(M * N)
        orc.ozaki_tsmttsm(M, N, K, A, M, 1.0, B, N, 0.0, got, M, S)
        errs.append(np.linalg.norm((got - hi) - lo) / np.linalg.norm(hi))
    assert errs[0] > errs[1] > errs[2]
    # truncated scheme: limb truncation and the dropped levels s + t > S + 1
    # each contribute about S 2^(-7 S) of sum |a||b|
    for S, err in zip((3, 4, 5), errs):
        assert err <= 8 * (S + 1) * 2.0 ** (-7 * S), (S, err)
    # on C3 data the C3 tolerance, 8 K eps at K = 2^24 (1.5e-8), holds from 3
    # limbs on (U[-1,1) data need 4: its result is small next to sum |a||b|)
    assert errs[0 if data == "c3" else 1] <= 8 * 2 ** 24 * np.finfo(float).eps
    # the level s + t = S + 2 is kept: about 2^-7 below the truncated scheme
    for S, err in zip((3, 4, 5), errs):
        assert err <= 8 * (S + 1) * 2.0 ** (-7 * S - 7) * (32 if data == "uniform" else 1), (S, err)


def test_alpha_beta_and_growth(orc):
    M, N, K = 3, 4, 5000
    A = orc.uniform(K * M, 4, 0)
    B = orc.uniform(K * N, 4, 1)
    # magnitude grows along k: exercises the rescale-and-flush path
    scale = np.ldexp(1.0, (np.arange(K) // 700))
    A = (A.reshape(K, M) * scale[:, None]).ravel()
    C0 = orc.uniform(M * N, 4, 2)
    got = C0.copy()
    orc.ozaki_tsmttsm(M, N, K, A, M, 2.0, B, N, -0.5, got, M, 5)
    want = C0.copy()
    orc.tsmttsm(M, N, K, A, M, 2.0, B, N, -0.5, want, M)
    assert np.linalg.norm(got - want) <= 1e-9 * np.linalg.norm(want)


def test_outlier_coarsens_only_its_chunk(orc):
    """A 1e5 outlier in column 0 raises that column's scale for its own
    32-row chunk only: the next chunk is 9+ bits below the scale and restarts
    it.  Column 0's error stays near one chunk's share (a scale kept for the
    whole run would make it about K / 32 times larger) and the other
    columns keep the scheme's accuracy."""
    M = N = 6
    K = 6000
    A = orc.uniform(K * M, 9, 0).copy()
    B = orc.uniform(K * N, 9, 1)
    A[100 * M] = 1e5
    hi, lo = orc.xtsmttsm(M, N, K, A, M, B, N)
    keep = np.ones(K, bool)
    keep[100] = False
    ref = np.abs(A.reshape(K, M)[keep]).T @ np.abs(B.reshape(K, N)[keep])
    got = np.zeros(M * N)
    orc.ozaki_tsmttsm(M, N, K, A, M, 1.0, B, N, 0.0, got, M, 4)
    err = np.abs((got - hi) - lo).reshape(N, M).T / ref
    assert err[0].max() <= 1e-5, err[0].max()
    assert err[1:].max() <= 8 * 5 * 2.0 ** -28, err[1:].max()


@pytest.mark.parametrize("M,K", [(5, 999), (128, 3001), (200, 2000), (130, 40)])
def test_gram_path(orc, M, K):
    """A Gram product (B is A) takes 64-row chunks on the diagonal tiles, one
    scale per column for both operands: exact on integers, symmetric bit for
    bit (the off-diagonal tiles are transposes of each other), and within the
    scheme's bound of the exact result like the general path on a copy."""
    A = orc.integers(K * M, 3, 0, -100, 100)
    want = np.zeros(M * M)
    orc.tsmttsm(M, M, K, A, M, 1.0, A, M, 0.0, want, M)
    got = np.zeros(M * M)
    orc.ozaki_tsmttsm(M, M, K, A, M, 1.0, A, M, 0.0, got, M, 3)
    assert np.array_equal(got, want)

    A = orc.c3_columns(M, K, 7)
    hi, lo = orc.xtsmttsm(M, M, K, A, M, A, M)
    for S in (3, 4):
        gram = np.zeros(M * M)
        orc.ozaki_tsmttsm(M, M, K, A, M, 1.0, A, M, 0.0, gram, M, S)
        g2 = gram.reshape(M, M)
        assert np.array_equal(g2, g2.T), S
        general = np.zeros(M * M)
        orc.ozaki_tsmttsm(M, M, K, A, M, 1.0, A.copy(), M, 0.0, general, M, S)
        for res in (gram, general):
            err = np.linalg.norm((res - hi) - lo) / np.linalg.norm(hi)
            assert err <= 8 * (S + 1) * 2.0 ** (-7 * S - 7), (S, err)
