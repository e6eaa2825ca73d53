"""Pins the rotation convention and the alignment-error metric: the reference's
known answers (cmtip/tests/test_errors.py:10-27) and closed forms of the rotation
(cmtip/eval/geom.py:37-45)."""
import numpy as np

from oracle import geom


def test_alignment_error_known_answers():
    q1 = np.array([1.0, 0, 0, 0])
    for q in ([0.0, 1, 0, 0], [0.0, 0, 1, 0], [0.0, 0, 0, 1]):
        assert geom.alignment_error(q1, np.array(q)) == 180
    for q in ([1.0, 1, 0, 0], [1.0, 0, 1, 0], [1.0, 0, 0, 1]):
        assert geom.alignment_error(q1, np.array(q) / np.sqrt(2)) == 90


def test_rotation_closed_forms():
    R = geom.quaternion2rot3d(np.array([[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 1, 0], [0, 0, 0, 1]], np.float32))
    np.testing.assert_allclose(R[0], np.eye(3))
    np.testing.assert_allclose(R[1], np.diag([1.0, -1, -1]))
    np.testing.assert_allclose(R[2], np.diag([-1.0, 1, -1]))
    np.testing.assert_allclose(R[3], np.diag([-1.0, -1, 1]))
    # 90 degrees about z maps x to y
    q = np.array([[np.cos(np.pi / 4), 0, 0, np.sin(np.pi / 4)]], np.float32)
    np.testing.assert_allclose(geom.quaternion2rot3d(q)[0] @ [1, 0, 0], [0, 1, 0], atol=1e-7)


def test_rotation_orthogonal_and_sign_invariant(rng):
    q = rng.standard_normal((50, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    R = geom.quaternion2rot3d(q.astype(np.float32))
    np.testing.assert_allclose(np.einsum("nij,nkj->nik", R, R), np.broadcast_to(np.eye(3), R.shape), atol=1e-6)
    np.testing.assert_allclose(np.linalg.det(R), 1.0, atol=1e-6)
    np.testing.assert_allclose(geom.quaternion2rot3d((-q).astype(np.float32)), R, atol=0)


def test_rotated_coords_layout(rng):
    # orientation-major flattening j = o * P + p (align.py:52 einsum "ijk,klm->jilm")
    q = rng.standard_normal((3, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    s = rng.standard_normal((3, 1, 5)).astype(np.float32)
    x, y, z = geom.rotated_coords(q, s, 2.0)
    R = np.linalg.inv(geom.quaternion2rot3d(q))
    for o in range(3):
        v = R[o] @ s[:, 0, :].astype(np.float64)
        np.testing.assert_array_equal(x[o * 5:(o + 1) * 5], v[0].astype(np.float32) / np.float32(2.0) * np.float32(np.pi))
        np.testing.assert_array_equal(z[o * 5:(o + 1) * 5], v[2].astype(np.float32) / np.float32(2.0) * np.float32(np.pi))
