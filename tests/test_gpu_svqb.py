"""SVQB on the GPU (svqb/svqb.cu:72-130): eager and hipGraph paths.

Checks: ||X^T X - I||_F after two sweeps (svqb.cu:184-211, computed exactly
by the oracle), the span of X kept, agreement with the oracle's CPU sweep up
to the sign/order freedom of eigenvectors, and bitwise agreement of every
graph replay with the eager result.
"""
import numpy as np
import pytest

from gpu_util import dev, host, note

pytestmark = pytest.mark.gpu


def _match_columns(Xg, Xc):
    """each GPU column equals some CPU column up to sign"""
    G = np.abs(Xg.T @ Xc)
    return np.max(np.abs(np.sort(G, axis=1)[:, -1] - 1.0))


@pytest.mark.parametrize("M", [1, 3, 8, 16, 32, 64])
def test_svqb_eager(gpu, orc, M):
    K = 8192
    note(f"svqb eager M={M} K={K}")
    X0 = orc.svqb_test_block(M, K, 100 + M, cond_exp=6.0)
    dX = dev(X0.copy())
    gpu.dsvqb([(M, K, dX, M)], sweeps=2)
    Xg = host(dX)
    fro, mx = orc.gram_error(M, K, Xg, M)
    assert fro < 1e-11, fro
    assert abs(gpu.orth_error(M, K, dX, M) - fro) < 1e-13
    Xn = Xg.reshape(K, M)
    R = Xn.T @ X0.reshape(K, M)
    assert np.linalg.norm(X0.reshape(K, M) - Xn @ R) <= 1e-9 * np.linalg.norm(X0)


@pytest.mark.parametrize("M", [2, 3, 8, 16, 33, 64])
def test_svqb_one_sweep_matches_oracle(gpu, orc, M):
    """kappa = 1e3: the spectrum of D S D is well separated, so each output
    column is determined up to sign; GPU and oracle agree to ~kappa^2 eps."""
    K = 8192
    note(f"svqb one sweep vs oracle M={M}")
    X0 = orc.svqb_test_block(M, K, 150 + M, cond_exp=3.0)
    dX = dev(X0.copy())
    gpu.dsvqb([(M, K, dX, M)], sweeps=1)
    Xc = X0.copy()
    orc.svqb_sweep(M, K, Xc, M)
    assert _match_columns(host(dX).reshape(K, M), Xc.reshape(K, M)) < 1e-8


def test_svqb_many_blocks_and_ld(gpu, orc):
    blocks = []
    data = []
    for i, M in enumerate((8, 16, 5, 32)):
        K = 3000 + 1000 * i
        ld = M + i
        X0 = orc.svqb_test_block(M, K, 200 + i, cond_exp=5.0).reshape(K, M)
        buf = np.zeros((K, ld))
        buf[:, :M] = X0
        d = dev(buf.ravel())
        blocks.append((M, K, d, ld))
        data.append((M, K, ld, d, buf.copy()))
    note("svqb batched 4 blocks")
    gpu.dsvqb(blocks, sweeps=2)
    for M, K, ld, d, buf0 in data:
        out = host(d).reshape(K, ld)
        assert np.array_equal(out[:, M:], buf0[:, M:])  # padding untouched
        X = np.ascontiguousarray(out[:, :M]).ravel()
        fro, _ = orc.gram_error(M, K, X, M)
        assert fro < 1e-11


def test_svqb_graph_replays_match_eager(gpu, orc):
    import torch
    M, K = 16, 20000
    X0 = orc.svqb_test_block(M, K, 300, cond_exp=6.0)
    eager = dev(X0.copy())
    gpu.dsvqb([(M, K, eager, M)], sweeps=2)
    want = host(eager)
    dX = dev(X0.copy())
    src = dev(X0.copy())
    g = gpu.Graph([(M, K, dX, M)], sweeps=2)
    for rep in range(3):
        note(f"svqb graph replay {rep}")
        dX.copy_(src)
        junk = torch.randn(1 << 20, device="cuda", dtype=torch.float64).sum()  # eager work between
        g.launch()
        torch.cuda.synchronize()
        assert np.array_equal(host(dX), want), rep
        del junk
    g.close()


@pytest.mark.parametrize("Ms", [(8,), (3, 16), (8, 16, 32, 64)])
def test_svqb_fused_grid_barrier(gpu, orc, Ms):
    """One cooperative launch per call: orthogonality as the eager path, the
    span kept, and bitwise reproducible from call to call."""
    K = 6000
    data = []
    for i, M in enumerate(Ms):
        X0 = orc.svqb_test_block(M, K + 100 * i, 400 + 10 * i + M, cond_exp=6.0)
        data.append((M, K + 100 * i, X0))
    note(f"svqb fused Ms={Ms}")
    outs = []
    for rep in range(2):
        devs = [dev(X0.copy()) for (_, _, X0) in data]
        gpu.dsvqb_fused([(M, Kb, d, M) for (M, Kb, _), d in zip(data, devs)], sweeps=2)
        outs.append([host(d) for d in devs])
    for (M, Kb, X0), got, again in zip(data, outs[0], outs[1]):
        assert np.array_equal(got, again)
        fro, _ = orc.gram_error(M, Kb, got, M)
        assert fro < 1e-11, (M, fro)
        Xn = got.reshape(Kb, M)
        R = Xn.T @ X0.reshape(Kb, M)
        assert np.linalg.norm(X0.reshape(Kb, M) - Xn @ R) <= 1e-9 * np.linalg.norm(X0)


def test_svqb_fused_one_sweep_matches_oracle(gpu, orc):
    M, K = 16, 8192
    X0 = orc.svqb_test_block(M, K, 170, cond_exp=3.0)
    dX = dev(X0.copy())
    note("svqb fused one sweep vs oracle")
    gpu.dsvqb_fused([(M, K, dX, M)], sweeps=1)
    Xc = X0.copy()
    orc.svqb_sweep(M, K, Xc, M)
    assert _match_columns(host(dX).reshape(K, M), Xc.reshape(K, M)) < 1e-8
