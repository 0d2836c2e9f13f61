"""Parity at the benchmark's full size (K = 2^26) and of the device data.

* The device counter RNG (`skinny_fill_*`) equals the oracle's bit for bit,
  so the host can regenerate any slice of the benchmark's inputs.
* Integer data with entries in [-2, 2] keeps every partial sum an exact
  integer (|sum| <= 4 K = 2^28 < 2^53), so TSMTTSM and TSMM results at
  K = 2^26 must equal the product computed by rocBLAS (torch.mm in FP64 or
  complex128, exact on such data, any summation order) bit for bit: out of
  place, in place (C == A) and the complex-double TSMM of the DC sweep.
* U[-1,1) data is compared with rocBLAS within the 8 K eps tolerance.
"""
import numpy as np
import pytest

from gpu_util import EPS64, note

pytestmark = pytest.mark.gpu
K_FULL = 1 << 26


def _free():
    import torch
    torch.cuda.synchronize()
    torch.cuda.empty_cache()


def test_device_fill_matches_oracle(gpu, orc):
    import torch
    n, seed, stream, offset = 1 << 20, 5, 3, 12345
    note("fill parity: uniform, uniform_f, int")
    d = torch.empty(n, dtype=torch.float64, device="cuda")
    gpu.fill_uniform(d, n, seed, stream, offset)
    assert np.array_equal(d.cpu().numpy(), orc.uniform(n, seed, stream, offset))
    f = torch.empty(n, dtype=torch.float32, device="cuda")
    gpu.fill_uniform_f(f, n, seed, stream, offset)
    assert np.array_equal(f.cpu().numpy(), orc.uniform_f(n, seed, stream, offset))
    gpu.fill_int(d, n, seed, stream, -7, 9, offset)
    assert np.array_equal(d.cpu().numpy(), orc.integers(n, seed, stream, -7, 9, offset))


@pytest.mark.parametrize("M", [1, 8, 64])
def test_tsmttsm_full_k_integer_bitwise(gpu, M):
    import torch
    K = K_FULL
    note(f"fullsize tsmttsm int M=N={M} K=2^26")
    A = torch.empty(K * M, dtype=torch.float64, device="cuda")
    B = torch.empty(K * M, dtype=torch.float64, device="cuda")
    gpu.fill_int(A, A.numel(), 11, 0)
    gpu.fill_int(B, B.numel(), 11, 1)
    C = torch.full((M * M,), float("nan"), dtype=torch.float64, device="cuda")
    assert gpu.dtsmttsm(0, M, M, K, A, M, 1.0, B, M, 0.0, C, M)
    ref = A.view(K, M).t().mm(B.view(K, M))  # M x M
    got = C.view(M, M).t()  # C is column-major
    assert torch.equal(got, ref)
    # the Gram (A == B) case reads A once
    assert gpu.dtsmttsm(0, M, M, K, A, M, 1.0, A, M, 0.0, C, M)
    ref = A.view(K, M).t().mm(A.view(K, M))
    assert torch.equal(C.view(M, M).t(), ref)
    del A, B, C, ref, got
    _free()


@pytest.mark.parametrize("M", [1, 8, 64])
def test_tsmm_full_k_integer_bitwise(gpu, M):
    import torch
    K, N = K_FULL, M
    note(f"fullsize tsmm int M=N={M} K=2^26 out of place and in place")
    A = torch.empty(K * M, dtype=torch.float64, device="cuda")
    gpu.fill_int(A, A.numel(), 12, 0)
    Bt = torch.empty(N * M, dtype=torch.float64, device="cuda")  # column-major M x N
    gpu.fill_int(Bt, Bt.numel(), 12, 1)
    ref = A.view(K, M).mm(Bt.view(N, M).t())  # K x N
    C = torch.full((K * N,), float("nan"), dtype=torch.float64, device="cuda")
    assert gpu.dtsmm(0, M, N, K, A, M, 1.0, Bt, M, 0.0, C, N)
    assert torch.equal(C.view(K, N), ref)
    del C
    _free()
    # in place: C aliases A (lda == ldc == M == N)
    assert gpu.dtsmm(0, M, N, K, A, M, 1.0, Bt, M, 0.0, A, N)
    assert torch.equal(A.view(K, N), ref)
    del A, Bt, ref
    _free()


def test_ztsmm_full_k_integer_bitwise(gpu):
    import torch
    K, M = K_FULL, 16
    N = M
    note(f"fullsize ztsmm int M=N={M} K=2^26")
    Ar = torch.empty(2 * K * M, dtype=torch.float64, device="cuda")
    gpu.fill_int(Ar, Ar.numel(), 13, 0)
    Br = torch.empty(2 * N * M, dtype=torch.float64, device="cuda")
    gpu.fill_int(Br, Br.numel(), 13, 1)
    Ac = torch.view_as_complex(Ar.view(K, M, 2))
    Bc = torch.view_as_complex(Br.view(N, M, 2))  # column-major M x N: row n = column n
    ref = Ac.mm(Bc.t().contiguous())  # plain transpose, no conjugation
    Cr = torch.full((2 * K * N,), float("nan"), dtype=torch.float64, device="cuda")
    assert gpu.ztsmm(0, M, N, K, Ar, M, 1.0, Br, M, 0.0, Cr, N)
    assert torch.equal(torch.view_as_complex(Cr.view(K, N, 2)), ref)
    del Cr
    _free()
    assert gpu.ztsmm(0, M, N, K, Ar, M, 1.0, Br, M, 0.0, Ar, N)  # in place
    assert torch.equal(torch.view_as_complex(Ar.view(K, N, 2)), ref)
    del Ar, Br, Ac, Bc, ref
    _free()


def test_uniform_full_k_within_tolerance(gpu):
    import torch
    K, M = K_FULL, 64
    note("fullsize uniform tsmttsm/tsmm M=N=64 K=2^26 vs rocBLAS")
    A = torch.empty(K * M, dtype=torch.float64, device="cuda")
    B = torch.empty(K * M, dtype=torch.float64, device="cuda")
    gpu.fill_uniform(A, A.numel(), 1, 0)
    gpu.fill_uniform(B, B.numel(), 1, 1)
    C = torch.empty(M * M, dtype=torch.float64, device="cuda")
    assert gpu.dtsmttsm(0, M, M, K, A, M, 1.0, B, M, 0.0, C, M)
    ref = A.view(K, M).t().mm(B.view(K, M))
    err = (C.view(M, M).t() - ref).norm() / ref.norm()
    assert float(err) <= 8 * K * EPS64, float(err)
    Bs = B[: M * M]
    T = torch.empty(K * M, dtype=torch.float64, device="cuda")
    assert gpu.dtsmm(0, M, M, K, A, M, 1.0, Bs, M, 0.0, T, M)
    ref = A.view(K, M).mm(Bs.view(M, M).t())
    err = (T.view(K, M) - ref).norm() / ref.norm()
    assert float(err) <= 8 * M * EPS64, float(err)
    del A, B, C, T, ref
    _free()


@pytest.mark.parametrize("M,N", [(8, 8), (64, 64), (5, 1)])
def test_tsmm_full_k_sampled_rows_from_rng(gpu, orc, M, N):
    """U[-1,1) data at K = 2^26: rows of C sampled across the whole range
    (first, last and 4094 random) are recomputed on the host from the
    counter RNG (the oracle regenerates each row of A from its offset) and
    must agree with the device within the dot-product error bound
    2 M eps sum_m |a_m b_mn| per entry (both sides sum M exact products in
    FP64, in different orders).  Out of place with beta = 0, and in place
    with beta != 0 (C == A, old C = the row of A)."""
    import torch
    K = K_FULL
    seed = 17
    rng = np.random.default_rng(M * 100 + N)
    rows = np.unique(np.concatenate([[0, K - 1], rng.integers(0, K, 4094)]))
    Bh = orc.uniform(M * N, seed, 1)  # column-major M x N, ldb = M
    Bm = Bh.reshape(N, M).T  # Bm[m, n]
    Ah = np.concatenate([orc.uniform(M, seed, 0, int(r) * M) for r in rows]).reshape(-1, M)
    bound_base = np.abs(Ah) @ np.abs(Bm)  # sum_m |a_m b_mn| per sampled row
    idx = torch.from_numpy(rows).cuda()

    note(f"fullsize tsmm sampled rows M={M} N={N} K=2^26 out of place")
    A = torch.empty(K * M, dtype=torch.float64, device="cuda")
    gpu.fill_uniform(A, A.numel(), seed, 0)
    dB = torch.from_numpy(Bh).cuda()
    C = torch.full((K * N,), float("nan"), dtype=torch.float64, device="cuda")
    assert gpu.dtsmm(0, M, N, K, A, M, 1.0, dB, M, 0.0, C, N)
    got = C.view(K, N).index_select(0, idx).cpu().numpy()
    want = np.zeros(len(rows) * N)
    orc.tsmm(M, N, len(rows), Ah.ravel(), M, 1.0, Bh, M, 0.0, want, N)
    want = want.reshape(-1, N)
    assert np.all(np.abs(got - want) <= 2 * M * EPS64 * bound_base)
    del C
    _free()

    if N <= M:
        note(f"fullsize tsmm sampled rows M={M} N={N} K=2^26 in place, beta != 0")
        alpha, beta = 0.75, -1.5
        assert gpu.dtsmm(0, M, N, K, A, M, alpha, dB, M, beta, A, M)
        got = A.view(K, M)[:, :N].index_select(0, idx).cpu().numpy()
        want = Ah[:, :N].copy().ravel()
        orc.tsmm(M, N, len(rows), Ah.ravel(), M, alpha, Bh, M, beta, want, N)
        want = want.reshape(-1, N)
        # the device folds beta I into the operand: M + 1 products per entry
        bound = 2 * (M + 2) * EPS64 * (abs(alpha) * bound_base + abs(beta) * np.abs(Ah[:, :N]))
        assert np.all(np.abs(got - want) <= bound)
    del A, dB, idx
    _free()
