"""Helpers for the GPU parity tests (device copies, error norms, progress)."""
import sys

import numpy as np


def note(msg: str) -> None:
    """Name the case about to run on the GPU, outside pytest's capture."""
    sys.__stderr__.write(msg + "\n")
    sys.__stderr__.flush()


def dev(a: np.ndarray):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def host(t) -> np.ndarray:
    import torch
    torch.cuda.synchronize()
    return t.cpu().numpy()


def rel_fro(got: np.ndarray, want: np.ndarray) -> float:
    d = np.linalg.norm((got - want).ravel())
    n = np.linalg.norm(np.asarray(want).ravel())
    return float(d / n) if n > 0 else float(d)


EPS64 = float(np.finfo(np.float64).eps)
EPS32 = float(np.finfo(np.float32).eps)
