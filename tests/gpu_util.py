"""Helpers shared by the GPU parity tests (inputs are seeded; the oracle is the checker)."""
import sys

import numpy as np
import torch


def say(*a):
    """Names the case about to run on the GPU (flushed before the GPU step)."""
    print(*a, file=sys.stderr, flush=True)


def dev():
    assert torch.cuda.is_available(), "GPU tests need a ROCm GPU"
    return torch.device("cuda", 0)


def rel_l2(a, b):
    a = np.asarray(a.cpu() if isinstance(a, torch.Tensor) else a)
    b = np.asarray(b.cpu() if isinstance(b, torch.Tensor) else b)
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


def points(rng, N, lim=3 * np.pi - 1e-3):
    return [v.astype(np.float32) for v in rng.uniform(-lim, lim, (3, N))]
