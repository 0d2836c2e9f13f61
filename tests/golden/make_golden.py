"""Write the golden fixtures under tests/golden/ (run from the repo root).

    python tests/golden/make_golden.py

The reference (te42kyfo/skinny_matmul) ships no stored vectors and its path is
CUDA-only, so nothing here runs it.  The expected outputs are computed with
exact Python integer / rational arithmetic, independently of the C oracle and
of the HIP library, from the closed forms of the reference semantics:

  TSMTTSM  C(m,n) = alpha * sum_k A(k,m) B(k,n) + beta * C(m,n)
           (cublasXgemm(N, T, M, N, K, ...), tsmttsm/gen_cublas.cuh:22-24)
  TSMM     C(k,n) = alpha * sum_m A(k,m) B(m,n) + beta * C(k,n)
           (cublasXgemm(T, N, N, K, M, ...), tsmm/cublas.cuh:37-39)

with integer inputs in [-2, 2] as test.cu:82-95 draws them, so every sum is
exact and the comparison is bitwise (test.cu:174).  The known-answer input of
numeric_tests.cu:23-53 is stored with its exact dot product.
"""
from __future__ import annotations

import json
import os
import random
from fractions import Fraction

HERE = os.path.dirname(os.path.abspath(__file__))


def rint(rng, n):
    return [rng.randint(-2, 2) for _ in range(n)]


def tsmttsm_case(rng, M, N, K, lda, ldb, ldc, alpha, beta, self_=False):
    A = rint(rng, K * lda)
    B = A if self_ else rint(rng, K * ldb)
    C0 = rint(rng, N * ldc)
    C = list(C0)
    for n in range(N):
        for m in range(M):
            s = sum(A[k * lda + m] * B[k * ldb + n] for k in range(K))
            C[n * ldc + m] = alpha * s + (beta * C0[n * ldc + m] if beta != 0 else 0)
    return dict(op="tsmttsm", M=M, N=N, K=K, lda=lda, ldb=ldb, ldc=ldc,
                alpha=alpha, beta=beta, self=self_, A=A, B=B, C0=C0, C=C)


def tsmm_case(rng, M, N, K, lda, ldb, ldc, alpha, beta, inplace=False):
    if inplace:
        assert lda == ldc
    A = rint(rng, K * lda)
    B = rint(rng, N * ldb)
    C0 = list(A) if inplace else rint(rng, K * ldc)
    C = list(C0)
    for k in range(K):
        for n in range(N):
            s = sum(A[k * lda + m] * B[n * ldb + m] for m in range(M))
            C[k * ldc + n] = alpha * s + (beta * C0[k * ldc + n] if beta != 0 else 0)
    return dict(op="tsmm", M=M, N=N, K=K, lda=lda, ldb=ldb, ldc=ldc,
                alpha=alpha, beta=beta, inplace=inplace, A=A, B=B, C0=C0, C=C)


def ztsmttsm_case(rng, M, N, K, alpha, beta):
    """complex, plain transpose (CUBLAS_OP_T, gen_cublas.cuh:30-33)"""
    A = [complex(*rint(rng, 2)) for _ in range(K * M)]
    B = [complex(*rint(rng, 2)) for _ in range(K * N)]
    C0 = [complex(*rint(rng, 2)) for _ in range(M * N)]
    C = list(C0)
    for n in range(N):
        for m in range(M):
            s = sum(A[k * M + m] * B[k * N + n] for k in range(K))
            C[n * M + m] = alpha * s + beta * C0[n * M + m]
    enc = lambda v: [[z.real, z.imag] for z in v]
    return dict(op="ztsmttsm", M=M, N=N, K=K, lda=M, ldb=N, ldc=M,
                alpha=[alpha.real, alpha.imag], beta=[beta.real, beta.imag],
                A=enc(A), B=enc(B), C0=enc(C0), C=enc(C))


def ztsmm_case(rng, M, N, K, alpha, beta):
    A = [complex(*rint(rng, 2)) for _ in range(K * M)]
    B = [complex(*rint(rng, 2)) for _ in range(N * M)]
    C0 = [complex(*rint(rng, 2)) for _ in range(K * N)]
    C = list(C0)
    for k in range(K):
        for n in range(N):
            s = sum(A[k * M + m] * B[n * M + m] for m in range(M))
            C[k * N + n] = alpha * s + beta * C0[k * N + n]
    enc = lambda v: [[z.real, z.imag] for z in v]
    return dict(op="ztsmm", M=M, N=N, K=K, lda=M, ldb=M, ldc=N,
                alpha=[alpha.real, alpha.imag], beta=[beta.real, beta.imag],
                A=enc(A), B=enc(B), C0=enc(C0), C=enc(C))


def kat_numeric():
    """numeric_tests.cu:23-53: N = 111113, every entry 0.1, planted +-2^62."""
    N = 111113
    big = float(1 << 62)
    a_special = {0: big, 1234: big, 41234: -big, N - 1: big}
    b_special = {0: big, N - 1: -big}
    tenth = Fraction(0.1)  # the double nearest 0.1, exactly
    exact = Fraction(0)
    for i in set(a_special) | set(b_special):
        exact += Fraction(a_special.get(i, 0.1)) * Fraction(b_special.get(i, 0.1))
    exact += (N - len(set(a_special) | set(b_special))) * tenth * tenth
    hi = float(exact)
    lo = float(exact - Fraction(hi))
    # the reference prints 0.1 * 0.1 * (N - 2) as "Should be" (numeric_tests.cu:71)
    printed = 0.1 * 0.1 * (N - 2)
    return dict(N=N, big=big, a_special={str(k): v for k, v in a_special.items()},
                b_special={str(k): v for k, v in b_special.items()},
                exact_hi=hi, exact_lo=lo, exact_num=str(exact.numerator),
                exact_den=str(exact.denominator), reference_printed=printed)


def main():
    rng = random.Random(20160421)
    ts = [
        tsmttsm_case(rng, 1, 1, 9, 1, 1, 1, 1, 0),
        tsmttsm_case(rng, 3, 5, 37, 3, 7, 4, 1, 1),
        tsmttsm_case(rng, 8, 8, 64, 8, 8, 8, 1, 0),
        tsmttsm_case(rng, 7, 7, 41, 9, 9, 8, 2, -1, self_=True),
        tsmttsm_case(rng, 17, 12, 23, 19, 12, 17, 1, 1),
        tsmttsm_case(rng, 33, 2, 50, 33, 4, 35, -1, 1),
    ]
    tm = [
        tsmm_case(rng, 1, 1, 13, 1, 1, 1, 1, 0),
        tsmm_case(rng, 4, 3, 29, 6, 5, 3, 1, 1),
        tsmm_case(rng, 8, 8, 33, 8, 8, 8, 2, 0),
        tsmm_case(rng, 5, 5, 31, 7, 5, 7, 1, 1, inplace=True),
        tsmm_case(rng, 3, 6, 17, 6, 3, 6, 1, 0, inplace=True),
        tsmm_case(rng, 20, 9, 21, 21, 22, 10, -1, 1),
    ]
    zs = [
        ztsmttsm_case(rng, 2, 3, 11, complex(1, 0), complex(0, 0)),
        ztsmttsm_case(rng, 4, 4, 19, complex(1, 1), complex(-1, 2)),
        ztsmm_case(rng, 3, 2, 13, complex(1, 0), complex(0, 0)),
        ztsmm_case(rng, 4, 4, 9, complex(2, -1), complex(1, 1)),
    ]
    with open(os.path.join(HERE, "tsmttsm_int.json"), "w") as f:
        json.dump(ts, f)
    with open(os.path.join(HERE, "tsmm_int.json"), "w") as f:
        json.dump(tm, f)
    with open(os.path.join(HERE, "complex_int.json"), "w") as f:
        json.dump(zs, f)
    with open(os.path.join(HERE, "kat_numeric.json"), "w") as f:
        json.dump(kat_numeric(), f, indent=1)
    print("wrote fixtures to", HERE)


if __name__ == "__main__":
    main()
