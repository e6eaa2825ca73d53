"""Orientation matching restated in FP64 (TEST INFRASTRUCTURE ONLY).

Follows cmtip/alignment/align.py:11-36 and :98-131 with the model slices from the
FP64 NUFFT oracle and the distances from sklearn's euclidean_distances on float64
inputs (the reference calls it on float32 inputs, which sklearn upcasts to float64
in chunks, sklearn/metrics/pairwise.py:418-421).
"""
import numpy as np
from sklearn.metrics.pairwise import euclidean_distances

from . import geom, nufft


def generate_weights(pixel_position_reciprocal, order=0):
    s_mag = np.linalg.norm(np.asarray(pixel_position_reciprocal, dtype=np.float64), axis=0) * 1e-10
    w = 1.0 / (s_mag ** order)
    return w / np.sum(w)


def compute_slices(quats, pixel_position_reciprocal, reciprocal_extent, ac, eps=1e-12, exact=False):
    """Re(type 2) of ac at the reference's float32 coordinates, FP64 arithmetic."""
    x, y, z = geom.rotated_coords(quats, pixel_position_reciprocal, reciprocal_extent)
    f = np.asarray(ac, dtype=np.float64)
    c = nufft.nudft2(f, x, y, z) if exact else nufft.nufft2(f, x, y, z, eps=eps)
    return c.real


def distances(quats, pixel_position_reciprocal, reciprocal_extent, slices_, ac, nbatches=1, order=0,
              eps=1e-12):
    """Full (n_ref, n_img) distance matrix of align.py:113-128 and the per-batch ratios."""
    n_ref = len(quats)
    quot = n_ref // nbatches
    P = pixel_position_reciprocal.shape[-1]
    data = np.asarray(slices_, dtype=np.float64).reshape(-1, P)
    w = generate_weights(pixel_position_reciprocal, order).reshape(-1)
    D = np.zeros((n_ref, data.shape[0]))
    ratios = []
    for nb in range(nbatches):
        start, end = nb * quot, (nb + 1) * quot
        model = compute_slices(quats[start:end], pixel_position_reciprocal, reciprocal_extent, ac, eps)
        model = model.reshape(end - start, P)
        ratio = data.std() / model.std()
        ratios.append(ratio)
        D[start:end] = euclidean_distances(model * ratio * w, data * w)
    return D, ratios


def match(quats, pixel_position_reciprocal, reciprocal_extent, slices_, ac, nbatches=1, order=0, eps=1e-12):
    """(index of the best reference per image, distance matrix)."""
    D, _ = distances(quats, pixel_position_reciprocal, reciprocal_extent, slices_, ac, nbatches, order, eps)
    return np.argmin(D, axis=0), D


def tie_band_ok(D, c=1.0, rel=None):
    """Images whose best-to-second-best squared-distance gap exceeds the FP32 bound
    c * P-independent relative margin ``rel`` of the best squared distance."""
    s = np.sort(D, axis=0)
    d1, d2 = s[0] ** 2, s[1] ** 2
    rel = 1e-4 if rel is None else rel
    return (d2 - d1) > c * rel * np.maximum(d1, 1e-30)
