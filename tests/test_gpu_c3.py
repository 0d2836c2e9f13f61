"""The precision configuration on the GPU: cancelling columns at every width.

Inputs follow the precision sweep: column j = (v + 1e-8 u_j) 2^e_j with v and
u_j ~ U[-1,1) from the counter RNG and integer e_j in [-20, 20] (exact
scaling), M = N in {16, 64, 128, 256}.  The Gram product A^T A and a general
A^T B (B from the same v with other noise, so its columns are nearly
parallel to A's) run on the three paths:

* FP64 TSMTTSM: relative Frobenius error <= 8 K eps;
* HR (double-double accumulation, numeric_tests.cu:58-72 compares double
  with dd on cancelling data): the dd result within 1e-28;
* Ozaki-INT8 with S = 3, 4, 5 limbs: within the scheme's bound
  (tests/ozaki_bound.py, against sum |a||b|) and within 8 K eps of the
  configuration's K = 2^24 (1.49e-8; the limb error does not grow with K).

Errors are taken on sampled entries (m, n) against the exact
(superaccumulator) product, which the host computes per entry.
"""
import numpy as np
import pytest

from gpu_util import EPS64, dev, host, note
from ozaki_bound import ozaki_bound

pytestmark = pytest.mark.gpu

K = 1 << 16
TOL_C3 = 8 * (1 << 24) * EPS64


def _c3_pair(orc, M, seed):
    """(A, B): A from orc.c3_columns; B shares v and the exponents with
    fresh noise streams."""
    A = orc.c3_columns(M, K, seed)
    v = orc.uniform(K, seed, 0)
    e = orc.integers(M, seed, 999, -20, 20).astype(int)
    X = np.empty((K, M))
    for j in range(M):
        X[:, j] = np.ldexp(v + 1e-8 * orc.uniform(K, seed, 5000 + j), int(e[j]))
    return A, np.ascontiguousarray(X).ravel()


def _entries(M):
    rng = np.random.default_rng(M)
    ms = list(rng.integers(0, M, 20)) + [0, M - 1, 0, M - 1, M // 2]
    ns = list(rng.integers(0, M, 20)) + [0, M - 1, M - 1, 0, M // 2]
    return np.array(ms), np.array(ns)


def _rel(got, hi, lo):
    return float(np.sqrt(np.sum(((got - hi) - lo) ** 2)) / np.sqrt(np.sum(hi ** 2)))


def _rel_dd(s, t, hi, lo):
    return float(np.sqrt(np.sum(((s - hi) + (t - lo)) ** 2)) / np.sqrt(np.sum(hi ** 2)))


@pytest.mark.parametrize("M", [16, 64, 128, 256])
@pytest.mark.parametrize("case", ["gram", "general"])
def test_c3_all_paths(gpu, orc, M, case):
    A, B = _c3_pair(orc, M, 31)
    if case == "gram":
        B = A
    ms, ns = _entries(M)
    hi, lo = orc.xtsmttsm_entries(K, A, M, B, M, ms, ns)
    dA = dev(A)
    dB = dA if case == "gram" else dev(B)
    dC = dev(np.zeros(M * M))

    def pick(C):  # column-major C: entry (m, n) at n * M + m
        return C[ns * M + ms]

    note(f"c3 {case} M={M} K={K}: fp64")
    assert gpu.dtsmttsm(0, M, M, K, dA, M, 1.0, dB, M, 0.0, dC, M)
    err64 = _rel(pick(host(dC)), hi, lo)
    assert err64 <= 8 * K * EPS64, err64

    note(f"c3 {case} M={M} K={K}: dd")
    Cdd = dev(np.zeros(2 * M * M))
    assert gpu.htsmttsm_dd(0, M, M, K, dA, M, 1.0, dB, M, 0.0, Cdd, M)
    q = host(Cdd).reshape(-1, 2)
    errdd = _rel_dd(pick(q[:, 0]), pick(q[:, 1]), hi, lo)
    assert errdd <= 1e-28, errdd
    # the double result of the HR path is the dd sum rounded once
    assert gpu.htsmttsm(0, M, M, K, dA, M, 1.0, dB, M, 0.0, dC, M)
    got = pick(host(dC))
    assert np.all(np.abs(got - (hi + lo)) <= np.spacing(np.abs(hi)))

    Ak, Bk = np.abs(A.reshape(K, M)), np.abs(B.reshape(K, M))
    absprod = np.einsum("km,km->m", Ak[:, ms], Bk[:, ns])
    cond = float(np.linalg.norm(absprod) / np.linalg.norm(hi))
    for S in (3, 4, 5):
        note(f"c3 {case} M={M} K={K}: ozaki S={S}")
        assert gpu.dtsmttsm_ozaki(0, M, M, K, dA, M, 1.0, dB, M, 0.0, dC, M, slices=S)
        err = _rel(pick(host(dC)), hi, lo)
        assert err <= ozaki_bound(S) * cond, (S, err)
        assert err <= TOL_C3, (S, err)
