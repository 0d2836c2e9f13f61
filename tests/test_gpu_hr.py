"""HR (double inputs, PseudoQuad accumulation) on the GPU.

* numeric_tests.cu's known-answer input (numeric_tests.cu:23-53) with
  blockCount = 1: the device's row-to-thread assignment and combination
  order are then fixed, and a host emulation with the oracle's dd operations
  (the same error-free operation sequence) must match the GPU bitwise.
* U[-1,1) data and cancelling columns: the dd result is within 1e-28
  relative Frobenius error of the exact (superaccumulator) product, and the
  double result within one ulp of it.
"""
import json
import os

import numpy as np
import pytest

from gpu_util import dev, host, note

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _kat():
    with open(os.path.join(GOLD, "kat_numeric.json")) as f:
        d = json.load(f)
    N = d["N"]
    a = np.full(N, 0.1)
    b = np.full(N, 0.1)
    for k, v in d["a_special"].items():
        a[int(k)] = v
    for k, v in d["b_special"].items():
        b[int(k)] = v
    return d, a, b


def _emulate_g1(orc, a, b, alpha):
    """blockCount = 1, M = N = 1: 256 row slots; slot t sums rows t, t+256, ..."""
    K = len(a)
    slots = []
    for t in range(256):
        s = (0.0, 0.0)
        for r in range(t, K, 256):
            s = orc.dd_slow_axpy2(s, a[r], b[r])
        slots.append(s)
    part = (0.0, 0.0, 0.0)
    for s in slots:  # slot combination in LDS, dd plus rounding remainders
        part = orc.dda_add(part, s)
    tot = orc.dda_add((0.0, 0.0, 0.0), part)  # last arrival over the single partial
    return orc.dd_scale2(orc.dd_renormalize(orc.dda_value(tot)), alpha)


def test_kat_bitwise_against_emulation(gpu, orc):
    d, a, b = _kat()
    N = d["N"]
    note(f"htsmttsm KAT K={N} blockCount=1")
    want = _emulate_g1(orc, a, b, 1.0)
    dA, dB = dev(a), dev(b)
    Cdd = dev(np.zeros(2))
    assert gpu.htsmttsm_dd(1, 1, 1, N, dA, 1, 1.0, dB, 1, 0.0, Cdd, 1)
    got = host(Cdd)
    assert got[0] == want[0] and got[1] == want[1], (got, want)
    Cd = dev(np.zeros(1))
    assert gpu.htsmttsm(1, 1, 1, N, dA, 1, 1.0, dB, 1, 0.0, Cd, 1)
    assert host(Cd)[0] == want[0] + want[1]
    # report what numeric_tests.cu prints: double, PseudoQuad, expected
    C64 = dev(np.zeros(1))
    assert gpu.dtsmttsm(1000, 1, 1, N, dA, 1, 1.0, dB, 1, 0.0, C64, 1)
    note(f"KAT: double {host(C64)[0]!r}  PseudoQuad {want[0] + want[1]!r}  "
         f"exact {d['exact_hi']!r}  (numeric_tests.cu prints {d['reference_printed']!r})")


@pytest.mark.parametrize("bc", [1, 2, 16, 17, 100, 435, 1000, 4000])
def test_kat_accurate_any_blockcount(gpu, bc):
    """The cancellation KAT at any blockCount: partials near 2^124, 2^62 and
    1 are combined with the rounding remainders kept, so the dd result stays
    within a few ulp of the exact value whatever the partition (plain dd
    partial sums lost up to 14% of it)."""
    d, a, b = _kat()
    N = d["N"]
    note(f"htsmttsm KAT K={N} blockCount={bc}")
    dA, dB = dev(a), dev(b)
    Cd = dev(np.zeros(1))
    assert gpu.htsmttsm(bc, 1, 1, N, dA, 1, 1.0, dB, 1, 0.0, Cd, 1)
    got = host(Cd)[0]
    ex = d["exact_hi"]
    assert abs(got - ex) <= 4 * np.spacing(ex), (bc, got, ex)


def _rel_dd(hi, lo, ehi, elo):
    num = np.sqrt(np.sum(((hi - ehi) + (lo - elo)) ** 2))
    den = np.sqrt(np.sum(ehi ** 2))
    return num / den


@pytest.mark.parametrize("M,N", [(1, 1), (3, 5), (8, 8), (16, 16), (33, 17), (64, 64), (100, 70)])
def test_uniform_dd_1e28(gpu, orc, M, N):
    K = 20000
    note(f"htsmttsm uniform M={M} N={N} K={K}")
    A = orc.uniform(K * M, 11, 0)
    B = orc.uniform(K * N, 11, 1)
    ehi, elo = orc.xtsmttsm(M, N, K, A, M, B, N)
    dA, dB = dev(A), dev(B)
    Cdd = dev(np.zeros(2 * M * N))
    assert gpu.htsmttsm_dd(0, M, N, K, dA, M, 1.0, dB, N, 0.0, Cdd, M)
    got = host(Cdd).reshape(-1, 2)
    assert _rel_dd(got[:, 0], got[:, 1], ehi, elo) <= 1e-28
    Cd = dev(np.zeros(M * N))
    assert gpu.htsmttsm(0, M, N, K, dA, M, 1.0, dB, N, 0.0, Cd, M)
    d = host(Cd)
    assert np.all(np.abs(d - ehi) <= np.spacing(np.abs(ehi)))


def test_cancelling_columns_dd(gpu, orc):
    """columns = (v + 1e-8 u_j) * 2^e_j with integer e_j in [-20, 20]"""
    M = 16
    K = 1 << 16
    v = orc.uniform(K, 21, 0)
    rng = np.random.default_rng(21)
    e = rng.integers(-20, 21, M)
    X = np.empty((K, M))
    for j in range(M):
        X[:, j] = np.ldexp(v + 1e-8 * orc.uniform(K, 21, 1 + j), int(e[j]))
    A = np.ascontiguousarray(X).ravel()
    ehi, elo = orc.xtsmttsm(M, M, K, A, M, A, M)
    dA = dev(A)
    Cdd = dev(np.zeros(2 * M * M))
    assert gpu.htsmttsm_dd(0, M, M, K, dA, M, 1.0, dA, M, 0.0, Cdd, M)
    got = host(Cdd).reshape(-1, 2)
    assert _rel_dd(got[:, 0], got[:, 1], ehi, elo) <= 1e-28


def test_alpha_beta_dd(gpu, orc):
    M, N, K = 4, 3, 5000
    A = orc.uniform(K * M, 12, 0)
    B = orc.uniform(K * N, 12, 1)
    C0 = orc.uniform(M * N, 12, 2)
    want = C0.copy()
    orc.htsmttsm(M, N, K, A, M, 2.0, B, N, -0.5, want, M)
    Cd = dev(C0.copy())
    assert gpu.htsmttsm(0, M, N, K, dev(A), M, 2.0, dev(B), N, -0.5, Cd, M)
    got = host(Cd)
    assert np.all(np.abs(got - want) <= 2 * np.spacing(np.abs(want)))
