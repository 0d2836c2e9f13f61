"""Complex TSMTTSM / TSMM (DC, FC) and real TSMM widths 65..128 on the GPU.

Integer-valued complex data keeps every product and sum exact, so results
are compared bitwise with the oracle (the reference compares bitwise against
cuBLAS Zgemm/Cgemm, test.cu:174, gen_cublas.cuh:142-151, tsmm/cublas.cuh:
204-213).  Complex TSMTTSM is a plain transpose (CUBLAS_OP_T, no conjugate).
"""
import json
import os
import random

import numpy as np
import pytest

from gpu_util import EPS64, dev, host, note, rel_fro

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _cz(v):
    return np.array([complex(a, b) for a, b in v], dtype=np.complex128)


def cint(orc, n, seed, stream):
    re = orc.integers(n, seed, 2 * stream)
    im = orc.integers(n, seed, 2 * stream + 1)
    return (re + 1j * im).astype(np.complex128)


def test_golden_complex(gpu):
    with open(os.path.join(GOLD, "complex_int.json")) as f:
        cases = json.load(f)
    for c in cases:
        for dt in (np.complex128, np.complex64):
            for bc in (0, 3):
                note(f"complex golden {c['op']} {np.dtype(dt).name} M={c['M']} N={c['N']} bc={bc}")
                A, B, C = (_cz(c[k]).astype(dt) for k in ("A", "B", "C0"))
                dA, dB, dC = dev(A), dev(B), dev(C)
                fn = gpu.tsmttsm if c["op"] == "ztsmttsm" else gpu.tsmm
                ok = fn(bc, c["M"], c["N"], c["K"], dA, c["lda"], complex(*c["alpha"]), dB,
                        c["ldb"], complex(*c["beta"]), dC, c["ldc"])
                assert ok
                assert np.array_equal(host(dC), _cz(c["C"]).astype(dt))


def _fuzz_tsmttsm(gpu, orc, rng, M, N, dt, Kmax):
    self_ = M == N and rng.random() < 0.5
    lda = M + rng.randint(0, 4)
    ldb = lda if self_ else N + rng.randint(0, 4)
    ldc = M + rng.randint(0, 4)
    K = rng.randint(1, Kmax)
    bc = rng.randint(1, 200)
    alpha = complex(rng.randint(-2, 2), rng.randint(-2, 2))
    beta = complex(rng.randint(-1, 1), rng.randint(-1, 1))
    seed = rng.randint(0, 1 << 30)
    note(f"ztsmttsm fuzz {np.dtype(dt).name} M={M} N={N} K={K} lda={lda} ldb={ldb} ldc={ldc} "
         f"bc={bc} self={self_} alpha={alpha} beta={beta} seed={seed}")
    A = cint(orc, K * lda, seed, 0).astype(dt)
    B = A if self_ else cint(orc, K * ldb, seed, 1).astype(dt)
    C = cint(orc, N * ldc, seed, 2).astype(dt)
    want = C.copy()
    orc.tsmttsm(M, N, K, A, lda, alpha, B, ldb, beta, want, ldc)
    dA = dev(A)
    dB = dA if self_ else dev(B)
    dC = dev(C)
    assert gpu.tsmttsm(bc, M, N, K, dA, lda, alpha, dB, ldb, beta, dC, ldc)
    assert np.array_equal(host(dC), want), (M, N, K, bc, self_)


def _fuzz_tsmm(gpu, orc, rng, M, N, dt, Kmax):
    inplace = rng.random() < 0.5
    ldb = M + rng.randint(0, 4)
    if inplace:
        ldc = max(N + rng.randint(0, 4), M)
        lda = ldc
    else:
        lda = M + rng.randint(0, 4)
        ldc = N + rng.randint(0, 4)
    K = rng.randint(1, Kmax)
    bc = rng.randint(1, 200)
    alpha = complex(rng.randint(-2, 2), rng.randint(-2, 2))
    beta = complex(rng.randint(-1, 1), rng.randint(-1, 1))
    seed = rng.randint(0, 1 << 30)
    note(f"ztsmm fuzz {np.dtype(dt).name} M={M} N={N} K={K} lda={lda} ldb={ldb} ldc={ldc} "
         f"bc={bc} inplace={inplace} alpha={alpha} beta={beta} seed={seed}")
    B = cint(orc, N * ldb, seed, 1).astype(dt)
    C = cint(orc, K * ldc, seed, 2).astype(dt)
    A = C.copy() if inplace else cint(orc, K * lda, seed, 0).astype(dt)
    want = C.copy()
    orc.tsmm(M, N, K, want if inplace else A, lda, alpha, B, ldb, beta, want, ldc)
    dC = dev(C)
    dA = dC if inplace else dev(A)
    dB = dev(B)
    assert gpu.tsmm(bc, M, N, K, dA, lda, alpha, dB, ldb, beta, dC, ldc)
    assert np.array_equal(host(dC), want), (M, N, K, bc, inplace)


@pytest.mark.parametrize("M", [1, 2, 3, 4, 7, 8, 13, 16, 24, 32, 33, 48, 64])
def test_fuzz_ztsmttsm(gpu, orc, M):
    rng = random.Random(300 + M)
    for _ in range(3):
        _fuzz_tsmttsm(gpu, orc, rng, M, M, np.complex128, 8000)
    _fuzz_tsmttsm(gpu, orc, rng, M, rng.randint(1, 64), np.complex128, 4000)


@pytest.mark.parametrize("M", [1, 2, 3, 4, 5, 8, 12, 16, 20, 31, 32, 40, 64])
def test_fuzz_ztsmm(gpu, orc, M):
    rng = random.Random(400 + M)
    for _ in range(3):
        _fuzz_tsmm(gpu, orc, rng, M, M, np.complex128, 8000)
    _fuzz_tsmm(gpu, orc, rng, M, rng.randint(1, 64), np.complex128, 4000)


@pytest.mark.parametrize("dt", [np.complex128, np.complex64])
def test_tsmm_one_column_real_beta(gpu, orc, dt):
    """M = N = 1 with a real beta runs on the real whole-row kernel over the
    interleaved view; complex alpha, padded strides, in and out of place."""
    rng = random.Random(91)
    for inplace in (False, True):
        for beta in (0.0, 1.0, -1.0):
            for bc in (0, 7, 300):
                lda = 1 + rng.randint(0, 2)
                ldc = lda if inplace else 1 + rng.randint(0, 2)
                K = rng.randint(2000, 9000)
                alpha = complex(rng.randint(-2, 2), rng.randint(-2, 2))
                seed = rng.randint(0, 1 << 30)
                note(f"ztsmm one column {np.dtype(dt).name} K={K} lda={lda} ldc={ldc} "
                     f"inplace={inplace} alpha={alpha} beta={beta} bc={bc} seed={seed}")
                B = cint(orc, 1, seed, 1).astype(dt)
                C = cint(orc, K * ldc, seed, 2).astype(dt)
                A = C.copy() if inplace else cint(orc, K * lda, seed, 0).astype(dt)
                want = C.copy()
                orc.tsmm(1, 1, K, want if inplace else A, lda, alpha, B, 1, beta, want, ldc)
                dC = dev(C)
                dA = dC if inplace else dev(A)
                assert gpu.tsmm(bc, 1, 1, K, dA, lda, alpha, dev(B), 1, beta, dC, ldc)
                assert np.array_equal(host(dC), want), (K, lda, ldc, inplace, beta, bc)


@pytest.mark.parametrize("M,N", [(1, 1), (4, 4), (8, 8), (16, 16), (32, 32), (5, 9),
                                 (17, 17), (33, 33), (40, 17), (17, 40), (48, 64), (64, 64),
                                 (64, 8), (57, 3), (24, 61)])
def test_fuzz_complex64(gpu, orc, M, N):
    rng = random.Random(M * 13 + N)
    for _ in range(2):
        _fuzz_tsmttsm(gpu, orc, rng, M, N, np.complex64, 8000)
        _fuzz_tsmm(gpu, orc, rng, M, N, np.complex64, 8000)


@pytest.mark.parametrize("M", [8, 32, 64])
def test_ztsmm_uniform_tolerance(gpu, orc, M):
    K = 1 << 14
    note(f"ztsmm uniform M=N={M} K={K}")
    A = (orc.uniform(K * M, 6, 0) + 1j * orc.uniform(K * M, 6, 1)).astype(np.complex128)
    B = (orc.uniform(M * M, 6, 2) + 1j * orc.uniform(M * M, 6, 3)).astype(np.complex128)
    want = np.zeros(K * M, dtype=np.complex128)
    orc.tsmm(M, M, K, A, M, 1.0, B, M, 0.0, want, M)
    dC = dev(np.zeros(K * M, dtype=np.complex128))
    assert gpu.tsmm(0, M, M, K, dev(A), M, 1.0, dev(B), M, 0.0, dC, M)
    assert rel_fro(host(dC), want) <= 8 * M * EPS64


@pytest.mark.parametrize("M,N", [(65, 65), (96, 40), (128, 128), (20, 100), (128, 3)])
def test_real_tsmm_wide(gpu, orc, M, N):
    rng = random.Random(M + 7 * N)
    for _ in range(2):
        inplace = rng.random() < 0.5 and True
        ldb = M + rng.randint(0, 4)
        if inplace:
            ldc = max(N + rng.randint(0, 4), M)
            lda = ldc
        else:
            lda = M + rng.randint(0, 4)
            ldc = N + rng.randint(0, 4)
        K = rng.randint(1, 5000)
        beta = float(rng.randint(0, 1))
        seed = rng.randint(0, 1 << 30)
        note(f"dtsmm wide M={M} N={N} K={K} inplace={inplace} beta={beta} seed={seed}")
        B = orc.integers(N * ldb, seed, 1)
        C = orc.integers(K * ldc, seed, 2)
        A = C.copy() if inplace else orc.integers(K * lda, seed, 0)
        want = C.copy()
        orc.tsmm(M, N, K, want if inplace else A, lda, 1.0, B, ldb, beta, want, ldc)
        dC = dev(C)
        dA = dC if inplace else dev(A)
        assert gpu.dtsmm(0, M, N, K, dA, lda, 1.0, dev(B), ldb, beta, dC, ldc)
        assert np.array_equal(host(dC), want)


@pytest.mark.parametrize("dt", [np.complex128, np.complex64])
@pytest.mark.parametrize("half", [(0, 0), (1, 0), (0, 1), (1, 1)])
def test_ztsmm_wide_alignment(gpu, orc, dt, half):
    """Complex TSMM above 32 columns: the three-product kernels need A and C
    aligned to a whole complex element (16 bytes for double, 8 for float);
    with either at half that, the real-view kernel runs.  Both give the
    oracle's result bitwise on integer data, out of place and in place, beta
    0 and complex."""
    import torch
    wide = dt == np.complex128
    real, npreal = (torch.float64, np.float64) if wide else (torch.float32, np.float32)
    fn = gpu.ztsmm if wide else gpu.ctsmm
    rb = np.dtype(dt).itemsize // 2  # bytes per real part
    offA, offC = half[0] * rb, half[1] * rb
    rng = random.Random(900 + offA + 2 * offC + rb)
    for M, N, inplace, beta in ((48, 48, False, 0.0), (64, 33, False, complex(1, -1)),
                                (40, 40, True, complex(-1, 1)), (64, 64, True, 0.0),
                                (17, 64, False, complex(0, 1))):
        K = rng.randint(2000, 9000)
        seed = rng.randint(0, 1 << 30)
        ldb = M
        ldc = max(M, N) if inplace else N
        lda = ldc if inplace else M
        note(f"tsmm alignment {np.dtype(dt).name} offA={offA} offC={offC} M={M} N={N} K={K} "
             f"inplace={inplace}")
        B = cint(orc, N * ldb, seed, 1).astype(dt)
        C = cint(orc, K * ldc, seed, 2).astype(dt)
        A = C.copy() if inplace else cint(orc, K * lda, seed, 0).astype(dt)
        want = C.copy()
        orc.tsmm(M, N, K, want if inplace else A, lda, complex(2, 1), B, ldb, beta, want, ldc)
        # complex buffers at byte offsets inside real device buffers
        bufC = torch.zeros(2 * C.size + 2, dtype=real, device="cuda")
        oc = half[1]
        bufC[oc:oc + 2 * C.size] = torch.from_numpy(C.view(npreal)).cuda()
        if inplace:
            if offA != offC:
                continue
            addrA = addrC = bufC.data_ptr() + offC
        else:
            bufA = torch.zeros(2 * A.size + 2, dtype=real, device="cuda")
            oa = half[0]
            bufA[oa:oa + 2 * A.size] = torch.from_numpy(A.view(npreal)).cuda()
            addrA, addrC = bufA.data_ptr() + offA, bufC.data_ptr() + offC
        dB = dev(B)
        assert fn(0, M, N, K, addrA, lda, complex(2, 1), dB, ldb, beta, addrC, ldc)
        got = host(bufC[oc:oc + 2 * C.size]).view(dt)
        assert np.array_equal(got, want), (M, N, inplace)


@pytest.mark.parametrize("dt", [np.complex128, np.complex64])
@pytest.mark.parametrize("M,lda", [(33, 48), (40, 48), (49, 64), (20, 32), (17, 24)])
def test_ztsmm_reads_only_m_columns(gpu, orc, dt, M, lda):
    """The wide complex TSMM reads only the first M columns of each row of
    A: the columns M .. lda - 1 hold NaN, which any read of them would carry
    into C (the kernels' A slices can reach past M when M is not a multiple
    of their slice width).  Integer data, bitwise against the oracle, beta 0
    and complex."""
    rng = random.Random(77 + M + np.dtype(dt).itemsize)
    fn = gpu.ztsmm if dt == np.complex128 else gpu.ctsmm
    N = 40
    for beta in (0.0, complex(1, -2)):
        K = rng.randint(3000, 7000)
        seed = rng.randint(0, 1 << 30)
        note(f"tsmm NaN padding {np.dtype(dt).name} M={M} lda={lda} N={N} K={K} beta={beta}")
        A = cint(orc, K * lda, seed, 0).astype(dt).reshape(K, lda)
        A[:, M:] = complex(np.nan, np.nan)
        A = A.ravel()
        B = cint(orc, N * M, seed, 1).astype(dt)
        C = cint(orc, K * N, seed, 2).astype(dt)
        want = C.copy()
        orc.tsmm(M, N, K, A, lda, complex(2, 1), B, M, beta, want, N)
        assert not np.isnan(want).any()
        dC = dev(C)
        assert fn(0, M, N, K, dev(A), lda, complex(2, 1), dev(B), M, beta, dC, N)
        assert np.array_equal(host(dC), want), (M, lda, beta)
