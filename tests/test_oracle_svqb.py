"""Pin the oracle's SVQB restatement (svqb/svqb.cu:72-130) on the CPU.

The Jacobi eigensolver is checked against numpy.linalg.eigh (LAPACK) and
against the defining identities S V = V diag(w), V^T V = I; the SVQB transform
against T^T S T = I; a full sweep against the orthonormality it must produce
and the span it must keep.
"""
import numpy as np
import pytest


@pytest.mark.parametrize("M", [1, 2, 3, 8, 17, 64])
def test_jacobi_vs_eigh(orc, M):
    rng = np.random.default_rng(M)
    A = rng.standard_normal((M, M))
    S = A + A.T
    w, V, sweeps = orc.jacobi_eig(S)
    wref = np.linalg.eigvalsh(S)
    assert np.allclose(np.sort(w), wref, rtol=1e-12, atol=1e-12 * np.abs(wref).max())
    assert np.allclose(V.T @ V, np.eye(M), atol=1e-13)
    assert np.allclose(S @ V, V * w, atol=1e-12 * np.abs(wref).max())
    assert sweeps < 40


@pytest.mark.parametrize("M", [3, 8, 32])
def test_svqb_transform_identity(orc, M):
    K = 4000
    X = orc.svqb_test_block(M, K, 5, cond_exp=3.0).reshape(K, M)
    S = X.T @ X
    T = orc.svqb_transform(S)
    assert np.allclose(T.T @ S @ T, np.eye(M), atol=1e-9)


@pytest.mark.parametrize("M", [4, 16, 64])
def test_svqb_two_sweeps(orc, M):
    K = 6000
    X0 = orc.svqb_test_block(M, K, 7, cond_exp=6.0)
    X = X0.copy()
    orc.svqb_sweep(M, K, X, M)
    orc.svqb_sweep(M, K, X, M)
    fro, mx = orc.gram_error(M, K, X, M)
    assert fro < 1e-12 and mx < 1e-12
    # span kept: X0 = X (X^T X0)
    Xn = X.reshape(K, M)
    R = Xn.T @ X0.reshape(K, M)
    assert np.linalg.norm(X0.reshape(K, M) - Xn @ R) <= 1e-10 * np.linalg.norm(X0)
