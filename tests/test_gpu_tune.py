"""Autotuner on the GPU: tuned choices are remembered, exported, imported,
and calls with blockCount == 0 stay correct under them."""
import numpy as np
import pytest

from gpu_util import EPS64, dev, host, note, rel_fro

pytestmark = pytest.mark.gpu


def test_autotune_roundtrip_and_parity(gpu, orc):
    h = gpu.Handle()
    K = 1 << 18
    for op, M in (("tsmttsm", 8), ("tsmttsm", 48), ("tsmm", 16)):
        note(f"autotune {op} M=N={M} K={K}")
        ms = gpu.autotune(op, M, M, K, reps=2, handle=h)
        assert ms > 0
    text = gpu.tune_export(handle=h)
    lines = [l.split() for l in text.strip().splitlines()]
    assert len(lines) == 3
    assert ["0", "0", "8", "8", "18"] == lines[0][:5] or any(l[:5] == ["0", "0", "8", "8", "18"] for l in lines)
    # tuned calls are correct
    for M in (8, 48):
        A = orc.uniform(K * M, 3, 0)
        B = orc.uniform(K * M, 3, 1)
        hi, lo = orc.xtsmttsm(M, M, K, A, M, B, M)
        dC = dev(np.zeros(M * M))
        assert gpu.dtsmttsm(0, M, M, K, dev(A), M, 1.0, dev(B), M, 0.0, dC, M, handle=h)
        assert rel_fro(host(dC) - lo, hi) <= 8 * K * EPS64
    M = 16
    A = orc.integers(K * M, 4, 0)
    Bm = orc.integers(M * M, 4, 1)
    want = np.zeros(K * M)
    orc.tsmm(M, M, K, A, M, 1.0, Bm, M, 0.0, want, M)
    dC = dev(np.zeros(K * M))
    assert gpu.dtsmm(0, M, M, K, dev(A), M, 1.0, dev(Bm), M, 0.0, dC, M, handle=h)
    assert np.array_equal(host(dC), want)
    # import into a fresh handle reproduces the table
    h2 = gpu.Handle()
    gpu.tune_import(text, handle=h2)
    assert gpu.tune_export(handle=h2) == text
    gpu.tune_clear(handle=h2)
    assert gpu.tune_export(handle=h2) == ""


# (elem_bytes, M, N, beta0, inplace) -> workgroups per CU of the untuned grid.
# Whole-row kernel (narrow C): FP64 four per CU for M = 1 or beta != 0, two
# otherwise; FP32 four, three for C 8 wide, eight for M <= 4, N <= 3 with
# beta != 0 out of place and for M = N = 1.  Matrix-core kernel: eight for A
# at most 4 columns wide, two otherwise.
DEFAULT_GRID_CASES = [
    (8, 5, 1, False, False, 4),   # whole rows, beta != 0 (M = 5, N = 1, beta = 1)
    (8, 5, 1, True, False, 2),    # whole rows, beta == 0, M > 1
    (8, 1, 1, True, False, 4),    # whole rows, M = 1
    (8, 2, 2, True, True, 2),     # whole rows in place
    (8, 2, 4, False, True, 4),    # whole rows in place, beta != 0, C wider than A
    (8, 2, 5, False, True, 8),    # matrix cores in place, beta != 0, A 2 wide
    (8, 4, 4, True, False, 8),    # matrix cores, A 4 wide
    (8, 8, 8, True, False, 2),    # matrix cores, A 8 wide
    (8, 64, 64, True, False, 2),  # matrix cores, A 64 wide
    (8, 8, 3, False, False, 2),   # matrix cores (transposed tile), beta != 0
    (4, 3, 1, True, False, 4),    # FP32 whole rows
    (4, 1, 1, True, False, 8),    # FP32 whole rows, M = N = 1
    (4, 4, 2, False, False, 8),   # FP32 whole rows, beta != 0, narrow A and C
    (4, 4, 4, True, False, 4),    # FP32 whole rows, C 4 wide
    (4, 12, 8, False, False, 3),  # FP32 whole rows, C 8 wide
    (4, 8, 3, True, False, 2),    # FP32 matrix cores, A 8 wide
    (4, 4, 6, True, False, 8),    # FP32 matrix cores, A 4 wide
    (4, 16, 16, True, False, 2),  # FP32 matrix cores
]


@pytest.mark.parametrize("eb,M,N,beta0,inplace,per_cu", DEFAULT_GRID_CASES)
def test_default_tsmm_grid(gpu, eb, M, N, beta0, inplace, per_cu):
    """The untuned TSMM grid follows the measured per-CU rule for each kernel
    class; a fresh handle holds no autotuned entries."""
    h = gpu.Handle()
    cus = gpu.cu_count(handle=h)
    K = 1 << 26
    got = gpu.tuned_blockcount_tsmm(M, N, K, eb, handle=h, beta0=beta0, inplace=inplace)
    assert got == per_cu * cus, (eb, M, N, beta0, inplace, got, cus)
    # a short K caps the grid at one workgroup per 256 rows
    assert gpu.tuned_blockcount_tsmm(M, N, 1000, eb, handle=h, beta0=beta0,
                                     inplace=inplace) == 4


def test_autotune_every_type(gpu, orc):
    """Every type tunes through its own entry point and keeps its own table
    entry; calls with blockCount 0 then run the tuned grid and stay exact on
    integer data (FR, DC, FC) or within the double-double bound (HR)."""
    import torch
    h = gpu.Handle()
    K = 1 << 16
    M, N = 12, 20
    for op, dt in (("tsmttsm", "FR"), ("tsmttsm", "DC"), ("tsmttsm", "FC"), ("tsmttsm", "HR"),
                   ("tsmm", "FR"), ("tsmm", "DC"), ("tsmm", "FC")):
        note(f"autotune {op} {dt} M={M} N={N} K={K}")
        assert gpu.autotune(op, M, N, K, reps=1, handle=h, dtype=dt) > 0
    lines = {tuple(l.split()[:4]) for l in gpu.tune_export(handle=h).strip().splitlines()}
    for key in (("0", "1"), ("0", "2"), ("0", "3"), ("0", "4"), ("1", "1"), ("1", "2"), ("1", "3")):
        assert key + (str(M), str(N)) in lines, (key, lines)
    # tuned calls: complex double TSMTTSM and float TSMM on integer data
    Ar = orc.integers(2 * K * M, 5, 0)
    Br = orc.integers(2 * K * N, 5, 1)
    Ac = Ar.view(np.complex128)
    Bc = Br.view(np.complex128)
    want = np.zeros(M * N, dtype=np.complex128)
    orc.tsmttsm(M, N, K, Ac, M, 1.0, Bc, N, 0.0, want, M)
    dC = dev(np.zeros(2 * M * N))
    assert gpu.ztsmttsm(0, M, N, K, dev(Ar), M, 1.0, dev(Br), N, 0.0, dC, M, handle=h)
    assert np.array_equal(host(dC).view(np.complex128), want)
    A32 = orc.integers(K * M, 6, 0).astype(np.float32)
    W32 = orc.integers(M * N, 6, 1).astype(np.float32)
    want32 = np.zeros(K * N, dtype=np.float32)
    orc.tsmm(M, N, K, A32, M, 1.0, W32, M, 0.0, want32, N)
    dC32 = torch.zeros(K * N, dtype=torch.float32, device="cuda")
    assert gpu.stsmm(0, M, N, K, dev(A32), M, 1.0, dev(W32), M, 0.0, dC32, N, handle=h)
    assert np.array_equal(host(dC32), want32)
    # HR with its tuned grid: within the double-double bound of the exact sum
    A = orc.uniform(K * M, 7, 0)
    B = orc.uniform(K * N, 7, 1)
    hi, lo = orc.xtsmttsm(M, N, K, A, M, B, N)
    Cdd = dev(np.zeros(2 * M * N))
    assert gpu.htsmttsm_dd(0, M, N, K, dev(A), M, 1.0, dev(B), N, 0.0, Cdd, M, handle=h)
    q = host(Cdd).reshape(-1, 2)
    err = np.sqrt(np.sum(((q[:, 0] - hi) + (q[:, 1] - lo)) ** 2)) / np.sqrt(np.sum(hi ** 2))
    assert err <= 1e-28, err
    h.close()


@pytest.mark.parametrize("dt,M,N,gram", [("d", 1, 1, False), ("d", 4, 4, True), ("d", 8, 3, False),
                                         ("d", 48, 48, False), ("d", 64, 64, False),
                                         ("d", 64, 64, True), ("d", 100, 70, False),
                                         ("s", 16, 16, False), ("s", 64, 64, True),
                                         ("z", 8, 8, False)])
def test_one_launch_variant_exact(gpu, orc, dt, M, N, gram):
    """Variant bit 1 (the partials of G > 1 workgroups summed by the last
    arrivals of the same launch: one group, or groups and then the last
    group) gives the oracle's result bitwise on integer data, as the
    two-kernel form does, for several grids."""
    import torch
    K = 123457
    log2K = K.bit_length() - 1
    dtype = {"d": 0, "s": 1, "z": 2}[dt]
    seed = M * 131 + N
    A = orc.integers(K * M * (2 if dt == "z" else 1), seed, 0)
    B = A if gram else orc.integers(K * N * (2 if dt == "z" else 1), seed, 1)
    if dt == "z":
        Ac, Bc = A.view(np.complex128), B.view(np.complex128)
        want = np.zeros(M * N, dtype=np.complex128)
        orc.tsmttsm(M, N, K, Ac, M, 1.0, Bc, N, 0.0, want, M)
    else:
        want = np.zeros(M * N)
        orc.tsmttsm(M, N, K, A, M, 1.0, B, N, 0.0, want, M)
    tdev = torch.float32 if dt == "s" else torch.float64
    dA = torch.from_numpy(A).to(tdev).cuda()
    dB = dA if gram else torch.from_numpy(B).to(tdev).cuda()
    h = gpu.Handle()
    for variant in (0, 2, 3):
        for G in (7, 100, 600):
            note(f"one launch: {dt} M={M} N={N} gram={gram} variant={variant} G={G}")
            gpu.tune_clear(handle=h)
            gpu.tune_import(f"0 {dtype} {M} {N} {log2K} {variant} {G}\n", handle=h)
            dC = torch.zeros(M * N * (2 if dt == "z" else 1), dtype=tdev, device="cuda")
            if dt == "d":
                assert gpu.dtsmttsm(0, M, N, K, dA, M, 1.0, dB, N, 0.0, dC, M, handle=h)
            elif dt == "s":
                assert gpu.stsmttsm(0, M, N, K, dA, M, 1.0, dB, N, 0.0, dC, M, handle=h)
            else:
                assert gpu.ztsmttsm(0, M, N, K, dA, M, 1.0, dB, N, 0.0, dC, M, handle=h)
            got = host(dC).astype(np.float64)
            if dt == "z":
                got = got.view(np.complex128)
            assert np.array_equal(got, want), (variant, G)


@pytest.mark.parametrize("dt", ["d", "s"])
@pytest.mark.parametrize("M,N", [(4, 4), (2, 3), (12, 8), (6, 1), (1, 16), (16, 3), (3, 5)])
def test_tsmm_form_variant_exact(gpu, orc, dt, M, N):
    """Variant 1 (bit 0) of a real TSMM entry runs the other kernel form (whole
    rows per thread where the rule picks the matrix cores, and the reverse);
    both give the oracle's result bitwise on integer data, beta 0 and 1, out
    of place and in place, and the variant survives an export."""
    import torch
    K = 54321
    log2K = K.bit_length() - 1
    dtype = {"d": 0, "s": 1}[dt]
    npdt = np.float64 if dt == "d" else np.float32
    h = gpu.Handle()
    for variant in (0, 1):
        gpu.tune_clear(handle=h)
        gpu.tune_import(f"1 {dtype} {M} {N} {log2K} {variant} 300\n", handle=h)
        assert f"1 {dtype} {M} {N} {log2K} {variant} 300" in gpu.tune_export(handle=h)
        for beta, inplace in ((0.0, False), (1.0, False), (0.0, True), (-1.0, True)):
            seed = M * 7 + N + 100 * variant
            ldc = max(M, N) if inplace else N
            lda = ldc if inplace else M
            note(f"tsmm form variant {dt} M={M} N={N} K={K} variant={variant} beta={beta} "
                 f"inplace={inplace}")
            B = orc.integers(N * M, seed, 1).astype(npdt)
            C = orc.integers(K * ldc, seed, 2).astype(npdt)
            A = C.copy() if inplace else orc.integers(K * lda, seed, 0).astype(npdt)
            want = C.copy()
            orc.tsmm(M, N, K, want if inplace else A, lda, 2.0, B, M, beta, want, ldc)
            dC = torch.from_numpy(C).cuda()
            dA = dC if inplace else torch.from_numpy(A).cuda()
            dB = torch.from_numpy(B).cuda()
            fn = gpu.dtsmm if dt == "d" else gpu.stsmm
            assert fn(0, M, N, K, dA, lda, 2.0, dB, M, beta, dC, ldc, handle=h)
            assert np.array_equal(host(dC), want), (variant, beta, inplace)
