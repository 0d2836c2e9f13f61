"""CPU oracle for the tall-skinny products (test infrastructure only).

Only tests/, ``__graft_entry__.smoke()`` and bench.py's ``cpu_baseline`` leg may
import this package, and only as the checker or as the timed CPU baseline.  The
product path (``skinny_matmul_amd``) never imports it.

The C restatement lives in ``oracle.c`` (reference file:line citations there);
this module only binds it with ctypes and adds numpy conveniences.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

i64 = C.c_int64
u64 = C.c_uint64
dbl = C.c_double
flt = C.c_float
vp = C.c_void_p


def build(force: bool = False) -> str:
    """Compile oracle.c into liboracle.so with the committed Makefile."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s"] + (["-B"] if force else []) + ["-C", _HERE],
                       check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        sig = {
            "orc_stream_key": (u64, [u64, u64]),
            "orc_rng_bits": (u64, [u64, u64]),
            "orc_fill_uniform": (None, [vp, i64, u64, u64, i64]),
            "orc_fill_uniform_f": (None, [vp, i64, u64, u64, i64]),
            "orc_fill_int": (None, [vp, i64, u64, u64, i64, C.c_int, C.c_int]),
            "orc_dtsmttsm": (None, [i64, i64, i64, vp, i64, dbl, vp, i64, dbl, vp, i64]),
            "orc_stsmttsm": (None, [i64, i64, i64, vp, i64, flt, vp, i64, flt, vp, i64]),
            "orc_ztsmttsm": (None, [i64, i64, i64, vp, i64, vp, vp, i64, vp, vp, i64]),
            "orc_ctsmttsm": (None, [i64, i64, i64, vp, i64, vp, vp, i64, vp, vp, i64]),
            "orc_dtsmm": (None, [i64, i64, i64, vp, i64, dbl, vp, i64, dbl, vp, i64]),
            "orc_stsmm": (None, [i64, i64, i64, vp, i64, flt, vp, i64, flt, vp, i64]),
            "orc_ztsmm": (None, [i64, i64, i64, vp, i64, vp, vp, i64, vp, vp, i64]),
            "orc_ctsmm": (None, [i64, i64, i64, vp, i64, vp, vp, i64, vp, vp, i64]),
            "orc_dtsmttsm_par": (None, [C.c_int, i64, i64, i64, vp, i64, dbl, vp, i64, dbl, vp, i64]),
            "orc_dtsmm_par": (None, [C.c_int, i64, i64, i64, vp, i64, dbl, vp, i64, dbl, vp, i64]),
            "orc_max_threads": (C.c_int, []),
            "orc_dd_fast2sum": (None, [dbl, dbl, vp, vp]),
            "orc_dd_gen2sum": (None, [dbl, dbl, vp, vp]),
            "orc_dd_fma2mult": (None, [dbl, dbl, vp, vp]),
            "orc_dd_accu": (None, [vp, vp, vp]),
            "orc_dd_axpy2": (None, [vp, dbl, dbl, vp]),
            "orc_dd_slow_axpy2": (None, [vp, dbl, dbl, vp]),
            "orc_dd_scale2": (None, [vp, dbl, vp]),
            "orc_dd_renormalize": (None, [vp, vp]),
            "orc_dda_add": (None, [vp, vp, vp]),
            "orc_dda_add3": (None, [vp, vp, vp]),
            "orc_dda_value": (None, [vp, vp]),
            "orc_htsmttsm": (None, [i64, i64, i64, vp, i64, dbl, vp, i64, dbl, vp, i64]),
            "orc_htsmttsm_dd": (None, [i64, i64, i64, vp, i64, dbl, vp, i64, vp, vp, i64]),
            "orc_xdot": (None, [i64, vp, i64, vp, i64, vp, vp]),
            "orc_xtsmttsm": (None, [i64, i64, i64, vp, i64, vp, i64, vp, vp, i64]),
            "orc_xtsmttsm_entries": (None, [i64, vp, i64, vp, i64, i64, vp, vp, vp, vp]),
            "orc_ozaki_tsmttsm": (None, [i64, i64, i64, vp, i64, dbl, vp, i64, dbl, vp, i64, C.c_int]),
            "orc_jacobi_eig": (C.c_int, [i64, vp, i64, vp, vp, i64, C.c_int]),
            "orc_svqb_transform": (None, [i64, vp, i64, vp, i64]),
            "orc_svqb_sweep": (None, [i64, i64, vp, i64]),
            "orc_gram_error": (None, [i64, i64, vp, i64, vp, vp]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _p(a: np.ndarray):
    assert a.flags["C_CONTIGUOUS"] or a.flags["F_CONTIGUOUS"] or a.ndim <= 1
    return a.ctypes.data_as(vp)


# --------------------------------------------------------------------------
# data generation (identical to the device generator)
# --------------------------------------------------------------------------
def uniform(n: int, seed: int, stream: int, offset: int = 0) -> np.ndarray:
    out = np.empty(n, dtype=np.float64)
    lib().orc_fill_uniform(_p(out), n, seed, stream, offset)
    return out


def uniform_f(n: int, seed: int, stream: int, offset: int = 0) -> np.ndarray:
    out = np.empty(n, dtype=np.float32)
    lib().orc_fill_uniform_f(_p(out), n, seed, stream, offset)
    return out


def integers(n: int, seed: int, stream: int, lo: int = -2, hi: int = 2,
             offset: int = 0) -> np.ndarray:
    out = np.empty(n, dtype=np.float64)
    lib().orc_fill_int(_p(out), n, seed, stream, offset, lo, hi)
    return out


# --------------------------------------------------------------------------
# products on flat buffers with explicit leading dimensions
# --------------------------------------------------------------------------
def _scalar_pair(v, dt):
    z = complex(v)
    return np.array([z.real, z.imag], dtype=dt)


def tsmttsm(M, N, K, A, lda, alpha, B, ldb, beta, C, ldc):
    """In-place update of the flat buffer C (dtype picks the routine)."""
    L = lib()
    if C.dtype == np.float64:
        L.orc_dtsmttsm(M, N, K, _p(A), lda, float(alpha), _p(B), ldb, float(beta), _p(C), ldc)
    elif C.dtype == np.float32:
        L.orc_stsmttsm(M, N, K, _p(A), lda, float(alpha), _p(B), ldb, float(beta), _p(C), ldc)
    elif C.dtype == np.complex128:
        L.orc_ztsmttsm(M, N, K, _p(A), lda, _p(_scalar_pair(alpha, np.float64)), _p(B), ldb,
                       _p(_scalar_pair(beta, np.float64)), _p(C), ldc)
    elif C.dtype == np.complex64:
        L.orc_ctsmttsm(M, N, K, _p(A), lda, _p(_scalar_pair(alpha, np.float32)), _p(B), ldb,
                       _p(_scalar_pair(beta, np.float32)), _p(C), ldc)
    else:
        raise TypeError(C.dtype)
    return C


def tsmm(M, N, K, A, lda, alpha, B, ldb, beta, C, ldc):
    L = lib()
    if C.dtype == np.float64:
        L.orc_dtsmm(M, N, K, _p(A), lda, float(alpha), _p(B), ldb, float(beta), _p(C), ldc)
    elif C.dtype == np.float32:
        L.orc_stsmm(M, N, K, _p(A), lda, float(alpha), _p(B), ldb, float(beta), _p(C), ldc)
    elif C.dtype == np.complex128:
        L.orc_ztsmm(M, N, K, _p(A), lda, _p(_scalar_pair(alpha, np.float64)), _p(B), ldb,
                    _p(_scalar_pair(beta, np.float64)), _p(C), ldc)
    elif C.dtype == np.complex64:
        L.orc_ctsmm(M, N, K, _p(A), lda, _p(_scalar_pair(alpha, np.float32)), _p(B), ldb,
                    _p(_scalar_pair(beta, np.float32)), _p(C), ldc)
    else:
        raise TypeError(C.dtype)
    return C


def htsmttsm(M, N, K, A, lda, alpha, B, ldb, beta, C, ldc):
    lib().orc_htsmttsm(M, N, K, _p(A), lda, float(alpha), _p(B), ldb, float(beta), _p(C), ldc)
    return C


def htsmttsm_dd(M, N, K, A, lda, alpha, B, ldb, ldc):
    hi = np.zeros(ldc * N, dtype=np.float64)
    lo = np.zeros(ldc * N, dtype=np.float64)
    lib().orc_htsmttsm_dd(M, N, K, _p(A), lda, float(alpha), _p(B), ldb, _p(hi), _p(lo), ldc)
    return hi, lo


def tsmttsm_par(nthreads, M, N, K, A, lda, alpha, B, ldb, beta, C, ldc):
    lib().orc_dtsmttsm_par(nthreads, M, N, K, _p(A), lda, float(alpha), _p(B), ldb,
                           float(beta), _p(C), ldc)
    return C


def tsmm_par(nthreads, M, N, K, A, lda, alpha, B, ldb, beta, C, ldc):
    lib().orc_dtsmm_par(nthreads, M, N, K, _p(A), lda, float(alpha), _p(B), ldb,
                        float(beta), _p(C), ldc)
    return C


def xdot(a: np.ndarray, b: np.ndarray):
    hi = np.zeros(1)
    lo = np.zeros(1)
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    lib().orc_xdot(len(a), _p(a), 1, _p(b), 1, _p(hi), _p(lo))
    return float(hi[0]), float(lo[0])


def xtsmttsm(M, N, K, A, lda, B, ldb):
    """Exact A^T B as (hi, lo) column-major M x N arrays (ldc = M)."""
    hi = np.zeros(M * N)
    lo = np.zeros(M * N)
    lib().orc_xtsmttsm(M, N, K, _p(A), lda, _p(B), ldb, _p(hi), _p(lo), M)
    return hi, lo


def xtsmttsm_entries(K, A, lda, B, ldb, ms, ns):
    ms = np.ascontiguousarray(ms, dtype=np.int64)
    ns = np.ascontiguousarray(ns, dtype=np.int64)
    hi = np.zeros(len(ms))
    lo = np.zeros(len(ms))
    lib().orc_xtsmttsm_entries(K, _p(A), lda, _p(B), ldb, len(ms), _p(ms), _p(ns), _p(hi), _p(lo))
    return hi, lo


# dd primitives -------------------------------------------------------------
def _pair2(fn, a, b):
    s = np.zeros(1)
    t = np.zeros(1)
    fn(float(a), float(b), _p(s), _p(t))
    return float(s[0]), float(t[0])


def fast2sum(a, b):
    return _pair2(lib().orc_dd_fast2sum, a, b)


def gen2sum(a, b):
    return _pair2(lib().orc_dd_gen2sum, a, b)


def fma2mult(a, b):
    return _pair2(lib().orc_dd_fma2mult, a, b)


def dd_accu(x, y):
    r = np.zeros(2)
    lib().orc_dd_accu(_p(np.array(x, dtype=np.float64)), _p(np.array(y, dtype=np.float64)), _p(r))
    return float(r[0]), float(r[1])


def dd_axpy2(c, a, b):
    r = np.zeros(2)
    lib().orc_dd_axpy2(_p(np.array(c, dtype=np.float64)), float(a), float(b), _p(r))
    return float(r[0]), float(r[1])


def dd_slow_axpy2(c, a, b):
    r = np.zeros(2)
    lib().orc_dd_slow_axpy2(_p(np.array(c, dtype=np.float64)), float(a), float(b), _p(r))
    return float(r[0]), float(r[1])


def dd_scale2(x, b):
    r = np.zeros(2)
    lib().orc_dd_scale2(_p(np.array(x, dtype=np.float64)), float(b), _p(r))
    return float(r[0]), float(r[1])


def dda_add(x3, y):
    """(s, t, u) + dd y (or + triple y, whose third word is added to u)."""
    r = np.zeros(3)
    if len(y) == 3:
        lib().orc_dda_add3(_p(np.array(x3, dtype=np.float64)),
                           _p(np.array(y, dtype=np.float64)), _p(r))
    else:
        lib().orc_dda_add(_p(np.array(x3, dtype=np.float64)),
                          _p(np.array(y, dtype=np.float64)), _p(r))
    return float(r[0]), float(r[1]), float(r[2])


def dda_value(x3):
    r = np.zeros(2)
    lib().orc_dda_value(_p(np.array(x3, dtype=np.float64)), _p(r))
    return float(r[0]), float(r[1])


def dd_renormalize(x):
    r = np.zeros(2)
    lib().orc_dd_renormalize(_p(np.array(x, dtype=np.float64)), _p(r))
    return float(r[0]), float(r[1])


def max_threads() -> int:
    return int(lib().orc_max_threads())


# SVQB ----------------------------------------------------------------------
def jacobi_eig(S: np.ndarray, max_sweeps: int = 40):
    """Eigen-decomposition of a symmetric matrix: (w, V), V columns."""
    M = S.shape[0]
    A = np.array(S, dtype=np.float64, order="F").copy(order="F")
    Af = A.ravel(order="F").copy()
    w = np.zeros(M)
    V = np.zeros(M * M)
    sweeps = lib().orc_jacobi_eig(M, _p(Af), M, _p(w), _p(V), M, max_sweeps)
    return w, V.reshape(M, M, order="F"), sweeps


def svqb_transform(S: np.ndarray) -> np.ndarray:
    M = S.shape[0]
    Sf = np.array(S, dtype=np.float64).ravel(order="F").copy()
    T = np.zeros(M * M)
    lib().orc_svqb_transform(M, _p(Sf), M, _p(T), M)
    return T.reshape(M, M, order="F")


def svqb_sweep(M, K, X, ldx):
    lib().orc_svqb_sweep(M, K, _p(X), ldx)
    return X


def gram_error(M, K, X, ldx):
    f = np.zeros(1)
    m = np.zeros(1)
    lib().orc_gram_error(M, K, _p(X), ldx, _p(f), _p(m))
    return float(f[0]), float(m[0])


def svqb_test_block(M, K, seed, cond_exp=6.0):
    """X = Q diag(logspace(0, -cond_exp, M)) W^T, Q (K x M) and W (M x M)
    orthonormal from a seeded QR; row-major K x M."""
    rng = np.random.default_rng(seed)
    Q, _ = np.linalg.qr(rng.standard_normal((K, M)))
    W, _ = np.linalg.qr(rng.standard_normal((M, M)))
    sig = np.logspace(0, -cond_exp, M)
    X = (Q * sig) @ W.T
    return np.ascontiguousarray(X).ravel()


def ozaki_tsmttsm(M, N, K, A, lda, alpha, B, ldb, beta, Cm, ldc, slices):
    """Ozaki-split TSMTTSM as the device runs it with blockCount = 1."""
    lib().orc_ozaki_tsmttsm(M, N, K, _p(A), lda, float(alpha), _p(B), ldb, float(beta),
                            _p(Cm), ldc, slices)
    return Cm


def c3_columns(M, K, seed):
    """C3 inputs: column j = (v + 1e-8 u_j) 2^e_j, v and u_j ~ U[-1,1) from the
    counter RNG, e_j integer in [-20, 20]; row-major K x M."""
    v = uniform(K, seed, 0)
    e = (integers(M, seed, 999, -20, 20)).astype(int)
    X = np.empty((K, M))
    for j in range(M):
        X[:, j] = np.ldexp(v + 1e-8 * uniform(K, seed, 1 + j), int(e[j]))
    return np.ascontiguousarray(X).ravel()
