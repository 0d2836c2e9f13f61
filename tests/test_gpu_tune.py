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
