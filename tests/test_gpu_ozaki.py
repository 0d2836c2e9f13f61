"""Ozaki-split TSMTTSM on the INT8 matrix cores.

* the INT8 32x32x32 result layout the kernel assumes, read back from the
  device;
* blockCount = 1: bitwise equal to the oracle's restatement of the same
  sequence (limbs, INT32 levels, flush points, FP64 reconstruction order),
  including ragged K, padded shapes and magnitudes that grow along K;
* Gram products (B is A): the diagonal tiles' 64-row chunks and the
  mirrored tiles above the diagonal, bitwise against the oracle at
  blockCount = 1 (ragged K, beta != 0, several tiles, an outlier);
* default blockCount: error against the exact product within the scheme's
  bound and within 8 K eps at K = 2^24 (1.5e-8) from 3 limbs on;
* launches of several tiles (placed by chunk set on one XCD): each tile
  bitwise equal to the one-tile product of its column slices;
* operands above 2 GiB (4 GiB each, every limb count, Gram too): exact on
  small-integer data, against the FP64 kernel and host int64 dot products.
"""
import numpy as np
import pytest

from gpu_util import dev, host, note, rel_fro
from ozaki_bound import ozaki_bound

pytestmark = pytest.mark.gpu


def test_i8_layout(gpu):
    rows, cols = gpu.probe_i8_layout()
    for lane in range(64):
        for r in range(16):
            assert cols[lane, r] == lane % 32
            assert rows[lane, r] == 8 * (r // 4) + 4 * (lane // 32) + (r % 4), (lane, r, rows[lane])


@pytest.mark.parametrize("M,N,K,S", [(5, 3, 1000, 4), (32, 32, 3001, 3), (128, 128, 2000, 4),
                                     (130, 70, 999, 5), (1, 1, 100, 4), (64, 64, 5000, 3),
                                     (64, 17, 3001, 5), (33, 64, 777, 4), (65, 64, 2000, 3),
                                     (64, 65, 1500, 4)])
def test_bitwise_blockcount1(gpu, orc, M, N, K, S):
    note(f"ozaki bitwise M={M} N={N} K={K} S={S}")
    A = orc.uniform(K * M, M + N, 0)
    B = orc.uniform(K * N, M + N, 1)
    C0 = orc.uniform(M * N, M + N, 2)
    want = C0.copy()
    orc.ozaki_tsmttsm(M, N, K, A, M, 1.5, B, N, -1.0, want, M, S)
    dC = dev(C0.copy())
    assert gpu.dtsmttsm_ozaki(1, M, N, K, dev(A), M, 1.5, dev(B), N, -1.0, dC, M, slices=S)
    assert np.array_equal(host(dC), want)


@pytest.mark.parametrize("M,N,S,gram", [(48, 40, 3, False), (48, 48, 5, True), (64, 64, 4, True),
                                         (20, 64, 5, False)])
def test_bitwise_strided_small(gpu, orc, M, N, S, gram):
    """The small-width tiles ([A B] and two row blocks of A) with leading
    dimensions beyond the widths, bitwise against the oracle; the padding
    columns hold NaN, which any read of them would carry into C."""
    K = 3001
    lda, ldb = M + 3, (M + 3 if gram else N + 5)
    note(f"ozaki strided M={M} N={N} S={S} gram={gram}")
    A = orc.uniform(K * lda, 51, 0).reshape(K, lda).copy()
    A[:, M:] = np.nan
    if gram:
        B = A
    else:
        B = orc.uniform(K * ldb, 51, 1).reshape(K, ldb).copy()
        B[:, N:] = np.nan
    A, B = A.ravel(), B.ravel()
    C0 = orc.uniform(M * N, 51, 2)
    want = C0.copy()
    orc.ozaki_tsmttsm(M, N, K, A, lda, 1.5, A if gram else B, ldb, -1.0, want, M, S)
    dA = dev(A)
    dB = dA if gram else dev(B)
    dC = dev(C0.copy())
    assert gpu.dtsmttsm_ozaki(1, M, N, K, dA, lda, 1.5, dB, ldb, -1.0, dC, M, slices=S)
    got = host(dC)
    assert np.all(np.isfinite(got))
    assert np.array_equal(got, want)


def test_bitwise_growth(gpu, orc):
    M, N, K = 16, 9, 40000
    A = orc.uniform(K * M, 9, 0)
    B = orc.uniform(K * N, 9, 1)
    scale = np.ldexp(1.0, np.arange(K) // 3000)
    A = (A.reshape(K, M) * scale[:, None]).ravel()
    want = np.zeros(M * N)
    orc.ozaki_tsmttsm(M, N, K, A, M, 1.0, B, N, 0.0, want, M, 4)
    dC = dev(np.zeros(M * N))
    assert gpu.dtsmttsm_ozaki(1, M, N, K, dev(A), M, 1.0, dev(B), N, 0.0, dC, M, slices=4)
    assert np.array_equal(host(dC), want)


def test_bitwise_outlier(gpu, orc):
    """An outlier chunk followed by ordinary data: the growth flush and the
    scale restart of the next chunk, bitwise against the oracle."""
    M, N, K = 6, 5, 4000
    note(f"ozaki outlier M={M} N={N} K={K}")
    A = orc.uniform(K * M, 21, 0).copy()
    B = orc.uniform(K * N, 21, 1).copy()
    A[100 * M] = 1e5
    B[2500 * N + 3] = -3e7
    for S in (3, 4):
        want = np.zeros(M * N)
        orc.ozaki_tsmttsm(M, N, K, A, M, 1.0, B, N, 0.0, want, M, S)
        dC = dev(np.zeros(M * N))
        assert gpu.dtsmttsm_ozaki(1, M, N, K, dev(A), M, 1.0, dev(B), N, 0.0, dC, M, slices=S)
        assert np.array_equal(host(dC), want), S


@pytest.mark.parametrize("M,K,S", [(5, 999, 3), (128, 3001, 3), (128, 40, 4), (128, 20, 3),
                                   (200, 2000, 4), (300, 130, 5), (130, 7000, 3),
                                   (64, 3001, 4), (37, 1000, 5), (64, 40000, 3), (65, 2000, 3),
                                   (1, 300, 3)])
def test_bitwise_gram(gpu, orc, M, K, S):
    note(f"ozaki gram bitwise M={M} K={K} S={S}")
    A = orc.uniform(K * M, 2 * M, 0)
    C0 = orc.uniform(M * M, 2 * M, 2)
    want = C0.copy()
    orc.ozaki_tsmttsm(M, M, K, A, M, 1.5, A, M, -1.0, want, M, S)
    dA = dev(A)
    dC = dev(C0.copy())
    assert gpu.dtsmttsm_ozaki(1, M, M, K, dA, M, 1.5, dA, M, -1.0, dC, M, slices=S)
    assert np.array_equal(host(dC), want)


def test_bitwise_gram_outlier(gpu, orc):
    """Outliers in one 32-row half of a 64-row block: the column's shared
    scale, the growth flush and the restart, bitwise against the oracle
    (M = 40: two 64-row blocks per chunk; row 100 is in the second block of
    chunk 0)."""
    M, K = 40, 6000
    note(f"ozaki gram outlier M={M} K={K}")
    A = orc.uniform(K * M, 23, 0).copy()
    A[100 * M + 3] = 1e5      # row 100: second half of chunk 1
    A[2000 * M + 7] = -3e7    # row 2000: first half of chunk 31
    A[4033 * M + 7] = 1e-6    # tiny rows do not trigger anything
    for S in (3, 4, 5):
        want = np.zeros(M * M)
        orc.ozaki_tsmttsm(M, M, K, A, M, 1.0, A, M, 0.0, want, M, S)
        dA = dev(A)
        dC = dev(np.zeros(M * M))
        assert gpu.dtsmttsm_ozaki(1, M, M, K, dA, M, 1.0, dA, M, 0.0, dC, M, slices=S)
        assert np.array_equal(host(dC), want), S


@pytest.mark.parametrize("M,N,S", [(256, 256, 3), (128, 128, 4), (256, 192, 5)])
def test_shared_tiles_match_single_tiles(gpu, orc, M, N, S):
    """Several tiles of one launch share rows: their workgroups are placed by
    chunk set on one XCD and load without the nontemporal hint.  Placement
    does not change what a tile computes (the same chunk sets, partials and
    reduction order), so each tile of the whole product is bitwise the
    product of its own column slices, computed alone at the same
    blockCount (a one-tile launch, placed as before).  blockCount 16 takes
    the XCD placement; 12 does not."""
    import torch
    K = 40000
    TN = 128 if S == 3 else 64
    A = torch.from_numpy(orc.uniform(K * M, 31, 0)).cuda()
    B = torch.from_numpy(orc.uniform(K * N, 31, 1)).cuda()
    for bc in (16, 12):
        note(f"ozaki shared tiles M={M} N={N} S={S} blockCount={bc}")
        full = torch.zeros(M * N, dtype=torch.float64, device="cuda")
        assert gpu.dtsmttsm_ozaki(bc, M, N, K, A, M, 1.0, B, N, 0.0, full, M, slices=S)
        full_h = host(full).reshape(N, M)  # column-major: [n, m]
        for m0 in range(0, M, 128):
            for n0 in range(0, N, TN):
                mt, nt = min(128, M - m0), min(TN, N - n0)
                part = torch.zeros(mt * nt, dtype=torch.float64, device="cuda")
                assert gpu.dtsmttsm_ozaki(bc, mt, nt, K, A[m0:], M, 1.0, B[n0:], N, 0.0,
                                          part, mt, slices=S)
                got = host(part).reshape(nt, mt)
                assert np.array_equal(full_h[n0:n0 + nt, m0:m0 + mt], got), (bc, m0, n0)


@pytest.mark.parametrize("S", [3, 4, 5])
def test_c3_error(gpu, orc, S):
    M = 128
    K = 1 << 17
    note(f"ozaki c3 M={M} K={K} S={S}")
    A = orc.c3_columns(M, K, 5)
    ms = np.array([0, 17, 64, 127, 5, 99])
    ns = np.array([0, 3, 64, 127, 120, 99])
    hi, lo = orc.xtsmttsm_entries(K, A, M, A, M, ms, ns)
    dA = dev(A)
    dC = dev(np.zeros(M * M))
    assert gpu.dtsmttsm_ozaki(0, M, M, K, dA, M, 1.0, dA, M, 0.0, dC, M, slices=S)
    got = host(dC).reshape(M, M).T  # column-major C: got[m, n]
    err = np.sqrt(np.sum((got[ms, ns] - hi - lo) ** 2)) / np.sqrt(np.sum(hi ** 2))
    Ak = np.abs(A.reshape(K, M))
    cond = np.linalg.norm(np.einsum("km,km->m", Ak[:, ms], Ak[:, ns])) / np.linalg.norm(hi)
    assert err <= ozaki_bound(S) * cond, (S, err, cond)
    assert err <= 8 * 2 ** 24 * np.finfo(float).eps


@pytest.mark.parametrize("M,N,S,gram", [(128, 128, 3, False), (128, 128, 3, True),
                                        (130, 70, 4, False), (64, 96, 5, False)])
def test_operands_above_2GiB_exact(gpu, M, N, S, gram):
    """Operands of more than 2 GiB (K = 2^22 + 37 rows, default blockCount):
    the per-chunk buffer resources address them in full.  On integer data in
    [-2, 2] every limb, INT32 level and flush is exact, so the result is the
    exact product: bitwise equal to the FP64 kernel's (also exact), and three
    entries recomputed on the host from the device's columns in int64."""
    import torch
    K = (1 << 22) + 37
    note(f"ozaki above 2 GiB: M={M} N={N} S={S} gram={gram} K={K}")
    A = torch.empty(K * M, dtype=torch.float64, device="cuda")
    gpu.fill_int(A, A.numel(), 77, 0, -2, 2)
    if gram:
        B, N = A, M
    else:
        B = torch.empty(K * N, dtype=torch.float64, device="cuda")
        gpu.fill_int(B, B.numel(), 77, 1, -2, 2)
    assert A.numel() * 8 > (2 << 30)
    Co = torch.zeros(M * N, dtype=torch.float64, device="cuda")
    Cf = torch.zeros(M * N, dtype=torch.float64, device="cuda")
    assert gpu.dtsmttsm_ozaki(0, M, N, K, A, M, 1.0, B, N, 0.0, Co, M, slices=S)
    assert gpu.dtsmttsm(0, M, N, K, A, M, 1.0, B, N, 0.0, Cf, M)
    torch.cuda.synchronize()
    got, ref = host(Co), host(Cf)
    assert np.array_equal(got, ref)
    A2, B2 = A.view(K, M), B.view(K, N)
    for m, n in ((0, 0), (M - 1, N - 1), (5, 17)):
        a = host(A2[:, m].contiguous()).astype(np.int64)
        b = host(B2[:, n].contiguous()).astype(np.int64)
        assert got[n * M + m] == float(np.dot(a, b)), (m, n)
    del A, B
    torch.cuda.empty_cache()
