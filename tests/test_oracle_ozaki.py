"""Pin the oracle's Ozaki-split restatement on the CPU.

Checks: exactness on data the limbs represent exactly (small integers, so
every product and the reconstruction are exact), error against the exact
superaccumulator decreasing with the limb count and within the scheme's
bound (tests/ozaki_bound.py) on U[-1,1) and on cancelling columns (within the
8 K eps tolerance at K = 2^24 from 3 limbs on), the signed-byte limbs
against a direct digit expansion, alpha/beta handling with growing
magnitudes, and outliers.
"""
import numpy as np
import pytest

from ozaki_bound import ozaki_bound


def _cond(A, B, K, M, N, hi):
    """||(|A|^T |B|)||_F / ||A^T B||_F: the product's conditioning."""
    ref = np.abs(A.reshape(K, M)).T @ np.abs(B.reshape(K, N))
    return np.linalg.norm(ref) / np.linalg.norm(hi)


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
    # the scheme's bound against sum |a||b| (tests/ozaki_bound.py); measured
    # errors sit 40-75 times below it
    cond = _cond(A, B, K, M, N, hi)
    for S, err in zip((3, 4, 5), errs):
        assert err <= ozaki_bound(S) * cond, (S, err, cond)
    # on C3 data the C3 tolerance, 8 K eps at K = 2^24 (1.5e-8), holds from 3
    # limbs on (U[-1,1) data need 4: its result is small next to sum |a||b|)
    assert errs[0 if data == "c3" else 1] <= 8 * 2 ** 24 * np.finfo(float).eps


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
    assert err[1:].max() <= 8 * ozaki_bound(4), err[1:].max()


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
        cond = _cond(A, A, K, M, M, hi)
        for res in (gram, general):
            err = np.linalg.norm((res - hi) - lo) / np.linalg.norm(hi)
            assert err <= ozaki_bound(S) * cond, (S, err)


@pytest.mark.parametrize("S", [3, 4, 5])
def test_signed_byte_limbs(orc, S):
    """One row of a column at a time (K = 1, M = N = 1, x * 1.0 with 1.0 on
    its own scale) returns the limb integer q = rint(x 2^(8S - 2 - E)) times
    2^(E - 8S + 2): the limbs are the balanced base-256 digits of q, which
    this checks against q computed here from the scale.  Values straddle
    rounding ties and both signs."""
    rng = np.random.default_rng(S)
    xs = list(rng.uniform(-1, 1, 200)) + [0.75, -0.75, 0.5 + 2.0 ** -30, -(0.5 + 2.0 ** -30)]
    sh = 8 * S - 2
    for x in xs:
        e = np.frexp(abs(x))[1]
        q = np.rint(np.ldexp(x, sh - e))  # numpy rint: ties to even
        want = np.ldexp(q, e - sh)
        got = np.zeros(1)
        orc.ozaki_tsmttsm(1, 1, 1, np.array([x]), 1, 1.0, np.array([1.0]), 1, 0.0, got, 1, S)
        assert got[0] == want, (x, got[0], want)
