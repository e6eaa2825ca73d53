"""Rotations and rotated q-vectors (TEST INFRASTRUCTURE ONLY).

quaternion2rot3d: cmtip/eval/geom.py:8-47 (scalar-first), with the quaternion
products and sums in the quaternion's float32 precision and the 2(.) and 1 - 2(.)
steps in float64.  rotated_coords: cmtip/alignment/align.py:51-57 and
cmtip/autocorrelation/autocorrelation.py:12-29,133-135,210: np.linalg.inv of that
matrix, the einsum "ijk,klm->jilm" in float64, astype(float32), then
/ reciprocal_extent and * pi in float32.
"""
import numpy as np


def quaternion2rot3d(quat):
    q = np.asarray(quat, dtype=np.float32).reshape(-1, 4)
    q0, q1, q2, q3 = q[:, 0], q[:, 1], q[:, 2], q[:, 3]
    q01, q02, q03 = q0 * q1, q0 * q2, q0 * q3
    q11, q12, q13 = q1 * q1, q1 * q2, q1 * q3
    q22, q23, q33 = q2 * q2, q2 * q3, q3 * q3
    d = lambda v: v.astype(np.float64)
    R = np.zeros((q.shape[0], 3, 3))
    R[:, 0, 0] = 1.0 - 2.0 * d(q22 + q33)
    R[:, 0, 1] = 2.0 * d(q12 - q03)
    R[:, 0, 2] = 2.0 * d(q13 + q02)
    R[:, 1, 0] = 2.0 * d(q12 + q03)
    R[:, 1, 1] = 1.0 - 2.0 * d(q11 + q33)
    R[:, 1, 2] = 2.0 * d(q23 - q01)
    R[:, 2, 0] = 2.0 * d(q13 - q02)
    R[:, 2, 1] = 2.0 * d(q23 + q01)
    R[:, 2, 2] = 1.0 - 2.0 * d(q11 + q22)
    return R


def rotated_vectors(quats, pixel_position_reciprocal):
    """Unscaled H, K, L (float32), shape (n_orient,) + pixel shape[1:]."""
    rotmat = np.linalg.inv(quaternion2rot3d(quats))
    s = np.asarray(pixel_position_reciprocal, dtype=np.float64)
    H, K, L = np.einsum("ijk,klm->jilm", rotmat, s)
    return H.astype(np.float32), K.astype(np.float32), L.astype(np.float32)


def rotated_coords(quats, pixel_position_reciprocal, reciprocal_extent):
    """Flattened float32 coordinates x = fl32(H) / fl32(ext) * pi as the reference forms them."""
    H, K, L = rotated_vectors(quats, pixel_position_reciprocal)
    ext = np.float32(reciprocal_extent)
    pi32 = np.float32(np.pi)
    return tuple((v.flatten() / ext) * pi32 for v in (H, K, L))


def alignment_error(q1, q2):
    """cmtip/alignment/errors.py:4-23."""
    ip = np.sum(np.asarray(q1) * np.asarray(q2), axis=-1)
    return np.rad2deg(np.arccos(np.around(2 * ip * ip - 1, decimals=5)))
