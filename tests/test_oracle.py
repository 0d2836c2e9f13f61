"""Pin the CPU oracle (CPU only).

The oracle is checked against (1) the golden fixtures computed with exact
Python integer arithmetic, (2) numpy's BLAS-backed products (a library
routine), (3) exact rational arithmetic for the error-free transformations and
the superaccumulator, and (4) the numeric_tests.cu known-answer input.  Each
check is chosen so that a plausible slip in the restatement (a transposed
operand, a dropped beta term, conjugation, a wrong leading dimension, an
in-place read-after-write) fails.
"""
import json
import math
import os
import random
from fractions import Fraction

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def load(name):
    with open(os.path.join(GOLD, name)) as f:
        return json.load(f)


# ---------------------------------------------------------------------------
# golden integer fixtures, bitwise
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("case", load("tsmttsm_int.json"), ids=lambda c: f"{c['M']}x{c['N']}x{c['K']}")
def test_oracle_tsmttsm_golden(orc, case):
    A = np.array(case["A"], dtype=np.float64)
    B = A if case["self"] else np.array(case["B"], dtype=np.float64)
    C = np.array(case["C0"], dtype=np.float64)
    orc.tsmttsm(case["M"], case["N"], case["K"], A, case["lda"], case["alpha"], B,
                case["ldb"], case["beta"], C, case["ldc"])
    assert np.array_equal(C, np.array(case["C"], dtype=np.float64))


@pytest.mark.parametrize("case", load("tsmm_int.json"), ids=lambda c: f"{c['M']}x{c['N']}x{c['K']}{'ip' if c['inplace'] else ''}")
def test_oracle_tsmm_golden(orc, case):
    B = np.array(case["B"], dtype=np.float64)
    C = np.array(case["C0"], dtype=np.float64)
    A = C if case["inplace"] else np.array(case["A"], dtype=np.float64)
    orc.tsmm(case["M"], case["N"], case["K"], A, case["lda"], case["alpha"], B,
             case["ldb"], case["beta"], C, case["ldc"])
    assert np.array_equal(C, np.array(case["C"], dtype=np.float64))


@pytest.mark.parametrize("case", load("tsmttsm_int.json"), ids=lambda c: f"{c['M']}x{c['N']}x{c['K']}")
def test_oracle_stsmttsm_golden(orc, case):
    """orc_stsmttsm (FR) on the same fixtures held as float32: small integers
    and their products and sums are exact in float32, so the result must
    equal the stored C bit for bit (the FR type of types.hpp:29-37)."""
    A = np.array(case["A"], dtype=np.float32)
    B = A if case["self"] else np.array(case["B"], dtype=np.float32)
    C = np.array(case["C0"], dtype=np.float32)
    assert np.abs(np.array(case["C"])).max() < 2 ** 24
    orc.tsmttsm(case["M"], case["N"], case["K"], A, case["lda"], case["alpha"], B,
                case["ldb"], case["beta"], C, case["ldc"])
    assert C.dtype == np.float32
    assert np.array_equal(C, np.array(case["C"], dtype=np.float32))


@pytest.mark.parametrize("case", load("tsmm_int.json"), ids=lambda c: f"{c['M']}x{c['N']}x{c['K']}{'ip' if c['inplace'] else ''}")
def test_oracle_stsmm_golden(orc, case):
    """orc_stsmm (FR) on the TSMM fixtures as float32, in place included:
    bitwise equal to the stored C."""
    B = np.array(case["B"], dtype=np.float32)
    C = np.array(case["C0"], dtype=np.float32)
    A = C if case["inplace"] else np.array(case["A"], dtype=np.float32)
    orc.tsmm(case["M"], case["N"], case["K"], A, case["lda"], case["alpha"], B,
             case["ldb"], case["beta"], C, case["ldc"])
    assert C.dtype == np.float32
    assert np.array_equal(C, np.array(case["C"], dtype=np.float32))


def _ulps32(got, want):
    """Distance in float32 units in the last place (same-sign finite values)."""
    g = got.astype(np.float32).view(np.int32).astype(np.int64)
    w = want.astype(np.float32).view(np.int32).astype(np.int64)
    return np.abs(g - w)


def test_oracle_fp32_accumulates_in_double(orc):
    """On non-integer float32 data the FR oracle is the double-precision sum
    of the exact float32 products, rounded once to float32: within one
    float32 ulp of numpy's float64 product (another summation order, error
    far below half a float32 ulp) and equal in almost every entry.  A
    float32 running sum, the plausible slip, misses by many ulps at this K."""
    M, N, K = 6, 5, 4096
    A = orc.uniform_f(K * M, 3, 0)
    B = orc.uniform_f(K * N, 3, 1)
    C0 = orc.uniform_f(M * N, 3, 2)
    C = C0.copy()
    orc.tsmttsm(M, N, K, A, M, 0.75, B, N, -0.5, C, M)
    S = A.reshape(K, M).astype(np.float64).T @ B.reshape(K, N).astype(np.float64)
    want = (0.75 * S - 0.5 * C0.reshape(N, M).T.astype(np.float64)).astype(np.float32)
    got = C.reshape(N, M).T
    d = _ulps32(got, want)
    assert d.max() <= 1 and (d == 0).mean() >= 0.9
    # the same sums accumulated in float32 fail this bound
    f32 = np.zeros((M, N), dtype=np.float32)
    Af, Bf = A.reshape(K, M), B.reshape(K, N)
    for k in range(K):
        f32 += np.outer(Af[k], Bf[k]).astype(np.float32)
    slip = (np.float32(0.75) * f32 - np.float32(0.5) * C0.reshape(N, M).T).astype(np.float32)
    assert _ulps32(slip, want).max() > 4

    # TSMM: rows of A times B (column-major M x N), beta != 0
    Kr = 513
    A = orc.uniform_f(Kr * M, 4, 0)
    Bm = orc.uniform_f(M * N, 4, 1)
    C0 = orc.uniform_f(Kr * N, 4, 2)
    C = C0.copy()
    orc.tsmm(M, N, Kr, A, M, 1.25, Bm, M, 0.5, C, N)
    S = A.reshape(Kr, M).astype(np.float64) @ Bm.reshape(N, M).T.astype(np.float64)
    want = (1.25 * S + 0.5 * C0.reshape(Kr, N).astype(np.float64)).astype(np.float32)
    d = _ulps32(C.reshape(Kr, N), want)
    assert d.max() <= 1 and (d == 0).mean() >= 0.9


def _cz(v):
    return np.array([complex(a, b) for a, b in v], dtype=np.complex128)


@pytest.mark.parametrize("case", load("complex_int.json"), ids=lambda c: c["op"])
def test_oracle_complex_golden(orc, case):
    A, B, C = _cz(case["A"]), _cz(case["B"]), _cz(case["C0"])
    alpha, beta = complex(*case["alpha"]), complex(*case["beta"])
    fn = orc.tsmttsm if case["op"] == "ztsmttsm" else orc.tsmm
    fn(case["M"], case["N"], case["K"], A, case["lda"], alpha, B, case["ldb"], beta, C, case["ldc"])
    assert np.array_equal(C, _cz(case["C"]))
    # the same through complex64 (integers are exact in fp32 too)
    A32, B32, C32 = A.astype(np.complex64), B.astype(np.complex64), _cz(case["C0"]).astype(np.complex64)
    fn(case["M"], case["N"], case["K"], A32, case["lda"], alpha, B32, case["ldb"], beta, C32, case["ldc"])
    assert np.array_equal(C32, _cz(case["C"]).astype(np.complex64))


# ---------------------------------------------------------------------------
# against numpy (library GEMM) with leading-dimension padding
# ---------------------------------------------------------------------------
def _views_tsmttsm(A, B, C, M, N, K, lda, ldb, ldc):
    Av = A[: K * lda].reshape(K, lda)[:, :M]
    Bv = B[: K * ldb].reshape(K, ldb)[:, :N]
    Cv = C[: N * ldc].reshape(N, ldc)[:, :M].T  # column-major M x N
    return Av, Bv, Cv


@pytest.mark.parametrize("M,N,K", [(1, 1, 1000), (4, 9, 777), (16, 16, 500), (33, 65, 301), (64, 1, 129)])
def test_oracle_tsmttsm_vs_numpy(orc, M, N, K):
    rng = np.random.default_rng(M * 1000 + N)
    lda, ldb, ldc = M + 3, N + 1, M + 2
    A = rng.integers(-2, 3, K * lda).astype(np.float64)
    B = rng.integers(-2, 3, K * ldb).astype(np.float64)
    C = rng.integers(-2, 3, N * ldc).astype(np.float64)
    C0 = C.copy()
    Av, Bv, Cv = _views_tsmttsm(A, B, C, M, N, K, lda, ldb, ldc)
    want = 2.0 * (Av.T @ Bv) - 1.0 * Cv
    orc.tsmttsm(M, N, K, A, lda, 2.0, B, ldb, -1.0, C, ldc)
    got = C.reshape(N, ldc)[:, :M].T
    assert np.array_equal(got, want)  # integer data: exact
    # padding rows of the column-major C are untouched
    assert np.array_equal(C.reshape(N, ldc)[:, M:], C0.reshape(N, ldc)[:, M:])


def test_oracle_tsmttsm_float_tolerance(orc):
    M, N, K = 8, 8, 65536
    A = orc.uniform(K * M, 1, 0)
    B = orc.uniform(K * N, 1, 1)
    C = np.zeros(M * N)
    orc.tsmttsm(M, N, K, A, M, 1.0, B, N, 0.0, C, M)
    want = (A.reshape(K, M).T @ B.reshape(K, N))
    got = C.reshape(N, M).T
    err = np.linalg.norm(got - want) / np.linalg.norm(want)
    assert err <= 8 * K * np.finfo(np.float64).eps


def test_oracle_beta_zero_does_not_read_C(orc):
    M, N, K = 3, 2, 10
    A = np.ones(K * M)
    B = np.ones(K * N)
    C = np.full(M * N, np.nan)
    orc.tsmttsm(M, N, K, A, M, 1.0, B, N, 0.0, C, M)
    assert np.all(C == K)
    C2 = np.full(K * N, np.nan)
    orc.tsmm(M, N, K, A, M, 1.0, np.ones(N * M), M, 0.0, C2, N)
    assert np.all(C2 == M)


def test_oracle_tsmm_inplace_matches_copy(orc):
    rng = np.random.default_rng(5)
    M, N, K, ld = 6, 4, 300, 7
    A = rng.standard_normal(K * ld)
    B = rng.standard_normal(N * M)
    ref = A.copy()
    orc.tsmm(M, N, K, A.copy(), ld, 1.5, B, M, 0.5, ref, ld)  # out of place, C preloaded with A
    ip = A.copy()
    orc.tsmm(M, N, K, ip, ld, 1.5, B, M, 0.5, ip, ld)
    assert np.array_equal(ip, ref)
    # and the closed form through numpy
    Av = A.reshape(K, ld)[:, :M]
    Bv = B.reshape(N, M).T
    want = 1.5 * (Av @ Bv) + 0.5 * A.reshape(K, ld)[:, :N]
    assert np.allclose(ip.reshape(K, ld)[:, :N], want, rtol=1e-13, atol=1e-13)


def test_oracle_complex_is_transpose_not_conjugate(orc):
    K = 5
    A = np.full(K, 1j, dtype=np.complex128)
    B = np.full(K, 1j, dtype=np.complex128)
    C = np.zeros(1, dtype=np.complex128)
    orc.tsmttsm(1, 1, K, A, 1, 1.0, B, 1, 0.0, C, 1)
    assert C[0] == -K  # i * i summed; a conjugate transpose would give +K


def test_oracle_parallel_baseline_matches(orc):
    M, N, K = 16, 8, 100003
    A = orc.integers(K * M, 3, 0)
    B = orc.integers(K * N, 3, 1)
    C1 = np.zeros(M * N)
    C2 = np.zeros(M * N)
    orc.tsmttsm(M, N, K, A, M, 1.0, B, N, 0.0, C1, M)
    orc.tsmttsm_par(4, M, N, K, A, M, 1.0, B, N, 0.0, C2, M)
    assert np.array_equal(C1, C2)
    T1 = np.zeros(K * N)
    T2 = np.zeros(K * N)
    Bm = orc.integers(M * N, 3, 2)
    orc.tsmm(M, N, K, A, M, 1.0, Bm, M, 0.0, T1, N)
    orc.tsmm_par(4, M, N, K, A, M, 1.0, Bm, M, 0.0, T2, N)
    assert np.array_equal(T1, T2)


# ---------------------------------------------------------------------------
# counter RNG: grid, range and an independent Python restatement
# ---------------------------------------------------------------------------
def _mix(z):
    m = (1 << 64) - 1
    z = (z + 0x9E3779B97F4A7C15) & m
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & m
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & m
    return z ^ (z >> 31)


def test_rng_matches_python_restatement(orc):
    seed, stream = 12345, 7
    key = _mix(seed ^ _mix((stream + 0x632BE59BD9B4E019) & ((1 << 64) - 1)))
    assert orc.lib().orc_stream_key(seed, stream) == key
    got = orc.uniform(100, seed, stream, 1000)
    for i in range(100):
        b = _mix((key + (1000 + i) * 0xD1B54A32D192ED03) & ((1 << 64) - 1))
        assert got[i] == (b >> 11) * 2.0 ** -52 - 1.0
    big = orc.uniform(1 << 16, 1, 0)
    assert big.min() >= -1.0 and big.max() < 1.0
    assert np.all(np.ldexp(big + 1.0, 52) == np.floor(np.ldexp(big + 1.0, 52)))
    ints = orc.integers(1 << 14, 1, 0)
    assert set(np.unique(ints)) == {-2.0, -1.0, 0.0, 1.0, 2.0}


def test_rng_splitmix64_published_values(orc):
    """orc_rng_bits(key, 0) is SplitMix64's output for state `key`: with key
    = (n - 1) * golden it must give the published first outputs of
    SplitMix64 seeded with 0 (Steele, Lea, Flood 2014; the xoshiro
    reference seeding generator)."""
    golden = 0x9E3779B97F4A7C15
    want = (0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F)
    for n, w in enumerate(want):
        assert orc.lib().orc_rng_bits((n * golden) & ((1 << 64) - 1), 0) == w


# ---------------------------------------------------------------------------
# PseudoQuad error-free transformations, exact rationals
# ---------------------------------------------------------------------------
def _rand_doubles(rng, n):
    out = []
    for _ in range(n):
        out.append(math.ldexp(rng.uniform(-1, 1), rng.randint(-60, 60)))
    return out


def test_eft_two_sum_and_fast2sum_exact(orc):
    rng = random.Random(1)
    for a, b in zip(_rand_doubles(rng, 300), _rand_doubles(rng, 300)):
        s, t = orc.gen2sum(a, b)
        assert Fraction(s) + Fraction(t) == Fraction(a) + Fraction(b)
        assert s == a + b
        x, y = (a, b) if abs(a) >= abs(b) else (b, a)
        s2, t2 = orc.fast2sum(x, y)
        assert Fraction(s2) + Fraction(t2) == Fraction(x) + Fraction(y)


def test_eft_two_prod_exact(orc):
    rng = random.Random(2)
    for a, b in zip(_rand_doubles(rng, 300), _rand_doubles(rng, 300)):
        s, t = orc.fma2mult(a, b)
        assert Fraction(s) + Fraction(t) == Fraction(a) * Fraction(b)


def _dd_rand(rng):
    s = math.ldexp(rng.uniform(-1, 1), rng.randint(-20, 20))
    t = s * 2.0 ** -53 * rng.uniform(-1, 1)
    return orc_renorm(s, t)


def orc_renorm(s, t):
    r = s + t
    return r, float(Fraction(s) + Fraction(t) - Fraction(r))


def test_dd_ops_accuracy(orc):
    rng = random.Random(3)
    u = Fraction(1, 2 ** 104)
    for _ in range(300):
        x = _dd_rand(rng)
        y = _dd_rand(rng)
        ex = Fraction(x[0]) + Fraction(x[1]) + Fraction(y[0]) + Fraction(y[1])
        r = orc.dd_accu(x, y)
        got = Fraction(r[0]) + Fraction(r[1])
        assert abs(got - ex) <= 4 * u * (abs(Fraction(x[0])) + abs(Fraction(y[0]))) + Fraction(1, 2 ** 1000)
        b = math.ldexp(rng.uniform(-1, 1), rng.randint(-10, 10))
        r = orc.dd_scale2(x, b)
        ex = (Fraction(x[0]) + Fraction(x[1])) * Fraction(b)
        assert abs(Fraction(r[0]) + Fraction(r[1]) - ex) <= 4 * u * abs(ex)
        a, c = rng.uniform(-1, 1), rng.uniform(-1, 1)
        ex = Fraction(x[0]) + Fraction(x[1]) + Fraction(a) * Fraction(c)
        scale = abs(Fraction(x[0])) + abs(Fraction(a) * Fraction(c))
        r = orc.dd_slow_axpy2(x, a, c)
        assert abs(Fraction(r[0]) + Fraction(r[1]) - ex) <= 8 * u * scale
        # the cheap axpy2 is only double-accurate under cancellation
        rc = orc.dd_axpy2(x, a, c)
        assert abs(Fraction(rc[0]) + Fraction(rc[1]) - ex) <= Fraction(1, 2 ** 52) * scale
        rn = orc.dd_renormalize(r)
        assert Fraction(rn[0]) + Fraction(rn[1]) == Fraction(r[0]) + Fraction(r[1])
        assert rn[0] == rn[0] + rn[1]


# ---------------------------------------------------------------------------
# superaccumulator against exact rationals
# ---------------------------------------------------------------------------
def test_xdot_exact_vs_fraction(orc):
    rng = random.Random(4)
    for trial in range(20):
        n = rng.randint(1, 300)
        a = [math.ldexp(rng.uniform(-1, 1), rng.randint(-300, 300)) for _ in range(n)]
        b = [math.ldexp(rng.uniform(-1, 1), rng.randint(-300, 300)) for _ in range(n)]
        # plant exact cancellation
        a.append(a[0]); b.append(-b[0])
        ex = sum(Fraction(x) * Fraction(y) for x, y in zip(a, b))
        hi, lo = orc.xdot(np.array(a), np.array(b))
        assert hi == float(ex)  # correctly rounded (ties are measure-zero here)
        if ex != 0:
            assert abs(Fraction(hi) + Fraction(lo) - ex) <= abs(ex) * Fraction(1, 2 ** 105)


def test_cheap_axpy2_loses_bits_under_cancellation(orc):
    """PseudoQuad.cuh:96-108 (axpy2): cp = r.s - abp is inexact when the
    product is not a double; slow_axpy2 (PseudoQuad.cuh:115-118) is not."""
    x = (0.06607522506561458, -4.651654860799798e-18)
    a, c = -0.09021660270633802, 0.6058924087222506
    ex = Fraction(x[0]) + Fraction(x[1]) + Fraction(a) * Fraction(c)
    cheap = orc.dd_axpy2(x, a, c)
    slow = orc.dd_slow_axpy2(x, a, c)
    assert abs(Fraction(cheap[0]) + Fraction(cheap[1]) - ex) == Fraction(1, 2 ** 59)
    assert abs(Fraction(slow[0]) + Fraction(slow[1]) - ex) <= abs(ex) * Fraction(1, 2 ** 104)


def _kat_input():
    d = load("kat_numeric.json")
    N = d["N"]
    a = np.full(N, 0.1)
    b = np.full(N, 0.1)
    for k, v in d["a_special"].items():
        a[int(k)] = v
    for k, v in d["b_special"].items():
        b[int(k)] = v
    return d, a, b


def test_dda_add_keeps_rounding_remainders(orc):
    """dda_add: (s, t) is bitwise dd_accu, and s + t + u equals the exact sum
    up to the rounding of the third word (checked with exact rationals),
    including dd partials of very different magnitudes."""
    rng = random.Random(5)
    cases = [((2.0 ** 124, 2.55, 0.0), (2.0 ** 62 * 0.1, 2.56)),
             ((2.0 ** 124, 4.6e17, 3.5), (-(2.0 ** 124), 1e-3)),
             ((1.0, 2.0 ** -60, 0.0), (-1.0, 2.0 ** -70))]
    for _ in range(300):
        x = _rand_doubles(rng, 2)
        y = _rand_doubles(rng, 2)
        xs, xt = orc.dd_renormalize((max(x, key=abs), min(x, key=abs)))
        ys, yt = orc.dd_renormalize((max(y, key=abs), min(y, key=abs)))
        cases.append(((xs, xt, 0.0), (ys, yt)))
    for x3, y in cases:
        r = orc.dda_add(x3, y)
        assert (r[0], r[1]) == orc.dd_accu(x3[:2], y)
        exact = sum(Fraction(v) for v in x3) + Fraction(y[0]) + Fraction(y[1])
        got = Fraction(r[0]) + Fraction(r[1]) + Fraction(r[2])
        assert abs(got - exact) <= abs(Fraction(r[2])) * Fraction(1, 2 ** 52)
    # the numeric_tests.cu situation: 2^124 and 2^62-sized partials with small
    # terms; plain dd loses the small terms, the triple keeps them exactly
    big, mid = 2.0 ** 124, 2.0 ** 62 * 0.1
    parts = [(big, 2.5), (mid, 1.25), (-mid, 0.75), (-big, 0.5)]
    acc, dd = (0.0, 0.0, 0.0), (0.0, 0.0)
    for p in parts:
        acc = orc.dda_add(acc, p)
        dd = orc.dd_accu(dd, p)
    assert sum(orc.dda_value(acc)) == 5.0
    assert sum(dd) != 5.0


def test_kat_numeric_tests_exact(orc):
    d, a, b = _kat_input()
    hi, lo = orc.xdot(a, b)
    assert hi == d["exact_hi"] and lo == d["exact_lo"]
    # the value numeric_tests.cu:71 prints, (N-2)*0.01, is not the exact result
    assert hi != d["reference_printed"]
    # sequential fp64 loses the small terms entirely
    C = np.zeros(1)
    orc.tsmttsm(1, 1, d["N"], a, 1, 1.0, b, 1, 0.0, C, 1)
    assert abs(C[0] - hi) > 1.0


def test_htsmttsm_dd_close_to_exact(orc):
    M, N, K = 4, 3, 20000
    A = orc.uniform(K * M, 9, 0)
    B = orc.uniform(K * N, 9, 1)
    hi, lo = orc.xtsmttsm(M, N, K, A, M, B, N)
    dh, dl = orc.htsmttsm_dd(M, N, K, A, M, 1.0, B, N, M)
    num = 0.0
    den = 0.0
    for e in range(M * N):
        ex = Fraction(hi[e]) + Fraction(lo[e])
        num += float((Fraction(dh[e]) + Fraction(dl[e]) - ex) ** 2)
        den += float(ex ** 2)
    assert math.sqrt(num / den) <= 1e-28
    # double output: within one ulp of the exact value
    C = np.zeros(M * N)
    orc.htsmttsm(M, N, K, A, M, 1.0, B, N, 0.0, C, M)
    assert np.all(np.abs(C - hi) <= np.spacing(np.abs(hi)))


def test_xtsmttsm_entries_match_full(orc):
    M, N, K = 5, 6, 3000
    A = orc.uniform(K * M, 2, 0)
    B = orc.uniform(K * N, 2, 1)
    hi, lo = orc.xtsmttsm(M, N, K, A, M, B, N)
    ms = np.array([0, 4, 2, 3])
    ns = np.array([5, 0, 2, 1])
    eh, el = orc.xtsmttsm_entries(K, A, M, B, N, ms, ns)
    for i in range(4):
        assert eh[i] == hi[ns[i] * M + ms[i]] and el[i] == lo[ns[i] * M + ms[i]]
