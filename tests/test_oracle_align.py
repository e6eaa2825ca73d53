"""The oracle's matching weights and the alignment-error metric, pinned by closed forms
and hand-computed values (no GPU)."""
import numpy as np
import pytest
import torch

from cmtip_amd import synth
from cmtip_amd.alignment import errors
from oracle import align as OA


def test_generate_weights_closed_forms():
    # align.py:11-24: w_p = (1e-10 |s_p|)^(-order) / sum over p
    s, _ = synth.detector(24)
    mag = np.linalg.norm(s.astype(np.float64), axis=0).ravel()
    P = mag.size
    # order 0: uniform 1 / P
    np.testing.assert_allclose(OA.generate_weights(s, 0).ravel(), np.full(P, 1.0 / P), rtol=1e-15)
    # order -2 (the MPI drivers' default): proportional to |s|^2, summing to one; the
    # 1e-10 scale cancels in the normalisation
    np.testing.assert_allclose(OA.generate_weights(s, -2).ravel(), mag ** 2 / np.sum(mag ** 2), rtol=1e-12)
    # order 1: proportional to 1 / |s|
    np.testing.assert_allclose(OA.generate_weights(s, 1).ravel(), (1 / mag) / np.sum(1 / mag), rtol=1e-12)
    for order in (0, -2, 1, 2):
        w = OA.generate_weights(s, order)
        assert abs(w.sum() - 1.0) < 1e-12 and np.all(w > 0)


def test_generate_weights_planted_sign_error_fails():
    # a weight with the exponent's sign flipped (|s|^order) is not the order -2 form
    s, _ = synth.detector(16)
    mag = np.linalg.norm(s.astype(np.float64), axis=0).ravel()
    wrong = (1 / mag ** 2) / np.sum(1 / mag ** 2)
    assert not np.allclose(OA.generate_weights(s, -2).ravel(), wrong, rtol=1e-3)


@pytest.mark.parametrize("symmetry,want", [(2, (10 + 10 + 85 + 0 + 30) / 5), (4, (10 + 10 + 5 + 0 + 30) / 5),
                                           (1, (10 + 170 + 95 + 0 + 150) / 5)])
def test_mean_error_symmetry_fold(symmetry, want):
    # errors.py:26-39: each error is folded to its distance from the nearest multiple
    # of 360 / symmetry in [0, 180 + 360 / symmetry), then averaged
    e = np.array([10.0, 170.0, 95.0, 0.0, 150.0])
    assert errors.mean_error(e, symmetry) == pytest.approx(want, abs=1e-12)
    assert errors.mean_error(torch.from_numpy(e), symmetry) == pytest.approx(want, abs=1e-12)


def test_mean_error_of_alignment_errors():
    # 180 degree rotations are folded to 0 by a 2-fold symmetry; 90 degrees stay
    q1 = np.array([[1.0, 0, 0, 0]] * 3)
    q2 = np.array([[0.0, 1, 0, 0], [1.0, 0, 0, 1] / np.sqrt(2), [1.0, 0, 0, 0]])
    e = errors.alignment_error(q1, q2)
    np.testing.assert_allclose(e, [180.0, 90.0, 0.0], atol=1e-9)
    assert errors.mean_error(e, 2) == pytest.approx(30.0, abs=1e-9)
    assert errors.mean_error(e, 1) == pytest.approx(90.0, abs=1e-9)
