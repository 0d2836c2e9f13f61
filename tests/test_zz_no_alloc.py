"""skinny_dtsmttsm_noalloc: the FP64 TSMTTSM with its partials in the
library's fixed per-GPU scratch array, called through the C ABI, against
skinny_dtsmttsm (the entry point whose kernels it shares)."""
import ctypes as C

import numpy as np
import pytest

from gpu_util import EPS64, dev, host, note, rel_fro

pytestmark = pytest.mark.gpu

# (M, N, K, blockCount, same grid as skinny_dtsmttsm): a packed narrow
# product, the paired-row kernel, a 64 x 64 tile whose default split does not
# fit the array (a smaller split is planned), several grid tiles, and one
# workgroup (in-kernel epilogue through the arrival counters)
CASES = [
    (8, 8, 1 << 20, 0, True),
    (12, 1, 1 << 19, 0, True),
    (64, 64, 1 << 18, 0, False),
    (100, 70, 1 << 16, 24, True),
    (5, 3, 1 << 14, 1, True),
]


def _call(L, fn, M, N, K, A, lda, alpha, B, ldb, beta, Cm, ldc, G, handle=0):
    return fn(handle, G, M, N, K, A.data_ptr(), lda, alpha, B.data_ptr(), ldb, beta,
              Cm.data_ptr(), ldc, 0)


def _pair(gpu, orc, M, N, K, G, seed, exact=True):
    """Both entry points on the same inputs (beta != 0 reads the same C);
    returns (noalloc result, workspace result, exact result or None)."""
    import torch
    L = gpu.lib()
    A = orc.uniform(K * M, seed, 0)
    B = orc.uniform(K * N, seed, 1)
    C0 = orc.uniform(M * N, seed, 2)
    dA, dB = dev(A), dev(B)
    c1, c2 = dev(C0), dev(C0)
    note(f"noalloc M={M} N={N} K={K} G={G}")
    st = _call(L, L.skinny_dtsmttsm_noalloc, M, N, K, dA, M, 1.5, dB, N, 0.5, c1, M, G)
    assert st == 0, gpu._last_error()
    st = _call(L, L.skinny_dtsmttsm, M, N, K, dA, M, 1.5, dB, N, 0.5, c2, M, G)
    assert st == 0, gpu._last_error()
    torch.cuda.synchronize()
    want = None
    if exact:
        hi, lo = orc.xtsmttsm(M, N, K, A, M, B, N)
        want = 1.5 * (hi + lo) + 0.5 * C0
    return host(c1), host(c2), want


def test_matches_workspace_entry_point(gpu, orc):
    """Bitwise equal to skinny_dtsmttsm for the same grid, two calls in a row
    on the shared array, the second smaller than the first; and within
    8 K eps of the exact result."""
    for M, N, K, G, same in CASES:
        got, ref, want = _pair(gpu, orc, M, N, K, G, 11 + M)
        if same:
            assert np.array_equal(got, ref), (M, N, K, G)
        assert rel_fro(got, want) <= 8 * K * EPS64, (M, N, K, G)
        assert rel_fro(ref, want) <= 8 * K * EPS64
        # the same shape again, then a smaller one: the array keeps the last
        # call's contents and must not leak them into the next result
        got2, _, _ = _pair(gpu, orc, M, N, K, G, 11 + M, exact=False)
        assert np.array_equal(got, got2)
        Ms, Ns, Ks = max(1, M // 2), max(1, N // 2), K // 4
        small, small_ref, small_want = _pair(gpu, orc, Ms, Ns, Ks, G, 23 + M)
        if same:
            assert np.array_equal(small, small_ref)
        assert rel_fro(small, small_want) <= 8 * Ks * EPS64


def test_default_grid_bitwise_when_it_fits(gpu, orc):
    """blockCount 0 plans the same grid as skinny_dtsmttsm when its partials
    fit the array, so the two results are bitwise equal."""
    M = N = 32
    K = 1 << 20
    got, ref, _ = _pair(gpu, orc, M, N, K, 0, 5, exact=False)
    assert np.array_equal(got, ref)


def test_rejects_grid_beyond_scratch(gpu):
    """An explicit blockCount whose partials exceed the array is refused
    before anything is launched, with a message naming the limit."""
    import torch
    L = gpu.lib()
    M = N = 64
    K = 1 << 16
    dA = torch.zeros(K * M, dtype=torch.float64, device="cuda")
    dC = torch.zeros(M * N, dtype=torch.float64, device="cuda")
    G = gpu.NOALLOC_PARTIAL_BYTES // (M * N * 8) + 1
    note(f"noalloc reject G={G}")
    st = _call(L, L.skinny_dtsmttsm_noalloc, M, N, K, dA, M, 1.0, dA, N, 0.0, dC, M, G)
    assert st == gpu.INVALID_ARG
    assert "bytes the library holds" in gpu._last_error()
    # a grid just inside the limit runs (480 partials of 32 KiB plus the
    # 22 second-level group sums the plan reserves)
    Gmax = 480
    st = _call(L, L.skinny_dtsmttsm_noalloc, M, N, K, dA, M, 1.0, dA, N, 0.0, dC, M, Gmax)
    assert st == 0, gpu._last_error()
    torch.cuda.synchronize()
    assert float(dC.abs().max()) == 0.0


def test_with_handle_and_stream(gpu, orc):
    """A given handle supplies the device and grid; a side stream orders the
    call like any other."""
    import torch
    L = gpu.lib()
    h = gpu.Handle()
    s = torch.cuda.Stream()
    M, N, K = 16, 16, 1 << 18
    A = orc.uniform(K * M, 7, 0)
    B = orc.uniform(K * N, 7, 1)
    dA, dB = dev(A), dev(B)
    c1 = torch.zeros(M * N, dtype=torch.float64, device="cuda")
    c2 = torch.zeros(M * N, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    note("noalloc with handle on a side stream")
    st = L.skinny_dtsmttsm_noalloc(h.value, 0, M, N, K, dA.data_ptr(), M, 1.0, dB.data_ptr(),
                                   N, 0.0, c1.data_ptr(), M, s.cuda_stream)
    assert st == 0, gpu._last_error()
    s.synchronize()
    assert gpu.dtsmttsm(0, M, N, K, dA, M, 1.0, dB, N, 0.0, c2, M, handle=h)
    assert np.array_equal(host(c1), host(c2))
