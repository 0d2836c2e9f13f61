"""TSMM widths above 128 real-view columns (the wide kernel), and the in-place
entry points in the GHOST kernels' argument order.

The reference takes any width out of place (tsmm/var1.cuh) and up to 256
columns in place (tsmm/varip1..3.cuh: 256 / varN threads per row).  The
checks follow test.cu:245-293: integer data, so every product and sum is
exact and results are compared bitwise with the oracle; padded leading
dimensions; beta in {0, 1} (and complex beta); the in-place case C == A with
the oracle reading the old A (tsmm/cublas.cuh:27-33).
"""
import random

import numpy as np
import pytest

from gpu_util import EPS64, dev, host, note, rel_fro

pytestmark = pytest.mark.gpu


def cint(orc, n, seed, stream):
    re = orc.integers(n, seed, 2 * stream)
    im = orc.integers(n, seed, 2 * stream + 1)
    return (re + 1j * im).astype(np.complex128)


def _case(gpu, orc, M, N, K, dt, inplace, beta, alpha=1.0, pad=(1, 2, 3), seed=7, bc=0):
    cplx = np.issubdtype(dt, np.complexfloating)
    ldb = M + pad[1]
    if inplace:
        ldc = max(N, M) + pad[2]
        lda = ldc
    else:
        lda = M + pad[0]
        ldc = N + pad[2]
    note(f"tsmm wide {np.dtype(dt).name} M={M} N={N} K={K} lda={lda} ldb={ldb} ldc={ldc} "
         f"inplace={inplace} alpha={alpha} beta={beta} bc={bc} seed={seed}")
    gen = (lambda n, s: cint(orc, n, seed, s)) if cplx else (lambda n, s: orc.integers(n, seed, s))
    B = gen(N * ldb, 1).astype(dt)
    C = gen(K * ldc, 2).astype(dt)
    A = C.copy() if inplace else gen(K * lda, 0).astype(dt)
    want = C.copy()
    orc.tsmm(M, N, K, want if inplace else A, lda, alpha, B, ldb, beta, want, ldc)
    dC = dev(C)
    dA = dC if inplace else dev(A)
    assert gpu.tsmm(bc, M, N, K, dA, lda, alpha, dev(B), ldb, beta, dC, ldc)
    got = host(dC)
    assert np.array_equal(got, want), (M, N, K, inplace, beta)


WIDE_REAL = [(129, 1), (1, 129), (129, 129), (130, 3), (3, 130), (200, 64), (64, 200),
             (256, 256), (333, 257), (512, 100), (1000, 17), (17, 1000)]


@pytest.mark.parametrize("M,N", WIDE_REAL)
def test_wide_fp64_integer(gpu, orc, M, N):
    rng = random.Random(M * 1009 + N)
    K = max(1, min(4000, 2_000_000 // (M * N)) + rng.randint(0, 40))
    for inplace in (False, True):
        for beta in (0.0, 1.0):
            _case(gpu, orc, M, N, K, np.float64, inplace, beta, seed=rng.randint(0, 1 << 30),
                  bc=rng.choice((0, 1, 3)))


def test_wide_fp64_1024(gpu, orc):
    # the largest width, several bands per workgroup (grid of 7) and a ragged
    # last band
    for inplace in (False, True):
        _case(gpu, orc, 1024, 1024, 531, np.float64, inplace, 1.0, bc=7)


@pytest.mark.parametrize("M,N", [(129, 129), (300, 40), (40, 300), (1024, 64)])
def test_wide_fp32_integer(gpu, orc, M, N):
    for inplace in (False, True):
        _case(gpu, orc, M, N, 777, np.float32, inplace, 1.0)


@pytest.mark.parametrize("M,N", [(65, 65), (100, 40), (40, 100), (65, 1), (1, 65),
                                 (256, 256), (512, 31)])
def test_wide_complex_integer(gpu, orc, M, N):
    rng = random.Random(M * 31 + N)
    K = max(1, min(2000, 500_000 // (M * N)))
    for inplace in (False, True):
        for beta in (0.0, 1.0, complex(1, -1)):
            alpha = complex(rng.randint(-2, 2), rng.randint(-2, 2))
            _case(gpu, orc, M, N, K, np.complex128, inplace, beta, alpha=alpha,
                  seed=rng.randint(0, 1 << 30))


def test_wide_complex64(gpu, orc):
    _case(gpu, orc, 200, 150, 333, np.complex64, True, complex(0, 1), alpha=complex(1, 1))
    _case(gpu, orc, 150, 200, 333, np.complex64, False, 0.0, alpha=complex(2, -1))


def test_wide_uniform_fp64(gpu, orc):
    M = N = 256
    K = 1 << 13
    note(f"tsmm wide uniform M=N={M} K={K}")
    A = orc.uniform(K * M, 5, 0)
    B = orc.uniform(M * N, 5, 1)
    C = orc.uniform(K * N, 5, 2)
    want = C.copy()
    orc.tsmm(M, N, K, A, M, 1.5, B, M, -0.5, want, N)
    dC = dev(C)
    assert gpu.dtsmm(0, M, N, K, dev(A), M, 1.5, dev(B), M, -0.5, dC, N)
    assert rel_fro(host(dC), want) <= 8 * M * EPS64


def test_wide_reproducible_and_limits(gpu, orc):
    import torch
    M, N, K = 300, 300, 5000
    A = dev(orc.uniform(K * M, 3, 0))
    B = dev(orc.uniform(M * N, 3, 1))
    C1 = torch.zeros(K * N, dtype=torch.float64, device="cuda")
    C2 = torch.zeros_like(C1)
    assert gpu.dtsmm(0, M, N, K, A, M, 1.0, B, M, 0.0, C1, N)
    assert gpu.dtsmm(5, M, N, K, A, M, 1.0, B, M, 0.0, C2, N)
    torch.cuda.synchronize()
    assert torch.equal(C1, C2)
    # widths above 1024 (512 complex) are not supported: false, not a crash
    big = torch.zeros(1025 * 4, dtype=torch.float64, device="cuda")
    Bb = torch.zeros(1025 * 1025, dtype=torch.float64, device="cuda")
    Cb = torch.zeros(4, dtype=torch.float64, device="cuda")
    assert gpu.dtsmm(0, 1025, 1, 4, big, 1025, 1.0, Bb, 1025, 0.0, Cb, 1) is False
    assert gpu.dtsmm(0, 1, 1025, 4, Cb, 1, 1.0, Bb, 1, 0.0, big, 1025) is False
    zb = torch.zeros(513 * 4, dtype=torch.complex128, device="cuda")
    zB = torch.zeros(513 * 513, dtype=torch.complex128, device="cuda")
    zC = torch.zeros(4, dtype=torch.complex128, device="cuda")
    assert gpu.ztsmm(0, 513, 1, 4, zb, 513, 1.0, zB, 513, 0.0, zC, 1) is False


@pytest.mark.parametrize("dt", [np.float64, np.float32, np.complex128, np.complex64])
@pytest.mark.parametrize("nin,nout", [(8, 8), (4, 12), (40, 24), (150, 150), (300, 200)])
def test_ghost_inplace_order(gpu, orc, dt, nin, nout):
    """x <- alpha x w + beta x with x row-major (stride stridex) and w column-major
    (column stride stridew): fix_ip_ghost_kernel / var_ip_ghost_kernel
    (tsmm/fix_ip_ghost.cuh:4-21, tsmm/var_ip_ghost.cuh:7-27)."""
    cplx = np.issubdtype(dt, np.complexfloating)
    nrows = 1531
    stridex = max(nin, nout) + 3
    stridew = nin + 1
    seed = nin * 7 + nout
    gen = (lambda n, s: cint(orc, n, seed, s)) if cplx else (lambda n, s: orc.integers(n, seed, s))
    x = gen(nrows * stridex, 0).astype(dt)
    w = gen(nout * stridew, 1).astype(dt)
    alpha = complex(1, -2) if cplx else 2.0
    beta = complex(-1, 1) if cplx else -1.0
    note(f"tsmm ghost order {np.dtype(dt).name} nin={nin} nout={nout} nrows={nrows}")
    want = x.copy()
    orc.tsmm(nin, nout, nrows, want, stridex, alpha, w, stridew, beta, want, stridex)
    dx = dev(x)
    assert gpu.tsmm_inplace(dx, dev(w), alpha, beta, nrows, nin, nout, stridex, stridew)
    assert np.array_equal(host(dx), want)
