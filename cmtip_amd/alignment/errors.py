"""Angular error between quaternions (cmtip/alignment/errors.py:4-39).

Metric only; numpy or torch inputs, computed with the same steps as the reference.
"""
import numpy as np
import torch


def alignment_error(q1, q2):
    """theta = arccos(round(2 <q1, q2>^2 - 1, 5)) in degrees (errors.py:4-23)."""
    if isinstance(q1, torch.Tensor) or isinstance(q2, torch.Tensor):
        q1 = torch.as_tensor(q1)
        q2 = torch.as_tensor(q2, device=q1.device)
        ip = (q1 * q2).sum(-1)
        c = torch.round(2 * ip * ip - 1, decimals=5)
        return torch.rad2deg(torch.arccos(c))
    inner_product = np.sum(q1 * q2, axis=-1)
    cos_theta = np.around(2 * np.square(inner_product) - 1, decimals=5)
    return np.rad2deg(np.arccos(cos_theta))


def mean_error(errors, symmetry=2):
    """Mean error after folding by the particle's symmetry (errors.py:26-39)."""
    errors = np.asarray(errors.cpu() if isinstance(errors, torch.Tensor) else errors)
    interval = 360 / symmetry
    symm_vals = np.arange(0, 180 + interval, interval)
    symm_vals = np.tile(symm_vals, errors.shape[0]).reshape(errors.shape[0], len(symm_vals))
    symm_diffs = np.abs(errors - symm_vals.T)
    return np.mean(np.min(symm_diffs, axis=0))
