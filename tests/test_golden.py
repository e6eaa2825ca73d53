"""The oracle reproduces the committed golden fixtures (tests/golden/*.npz, written by
tests/golden/make_fixtures.py): a change to the oracle's arithmetic shows up here."""
import os

import numpy as np
import pytest

from tests.golden import make_fixtures as MF

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _load(name):
    return np.load(os.path.join(GOLDEN, name + ".npz"))


@pytest.mark.parametrize("name", sorted(MF.FIXTURES))
def test_oracle_reproduces_fixture(name):
    want = _load(name)
    got = MF.FIXTURES[name]()
    assert sorted(got) == sorted(want.files)
    for key in want.files:
        a, b = np.asarray(got[key]), want[key]
        if a.dtype.kind in "biu":
            np.testing.assert_array_equal(a, b, err_msg=f"{name}.{key}")
        else:
            np.testing.assert_allclose(a, b, rtol=1e-10, atol=1e-12 * float(np.abs(b).max() + 1e-300),
                                       err_msg=f"{name}.{key}")


def test_fixture_contents():
    # the match fixture plants 12 references among the images and every image lies
    # outside the tie band; the alignment errors follow the package's metric
    from cmtip_amd.alignment import errors
    m = _load("slices_match_m21")
    assert np.array_equal(m["argmin"][:12], np.arange(12))
    assert m["ok"].mean() > 0.9
    g = _load("geom")
    np.testing.assert_allclose(np.asarray(errors.alignment_error(g["q1"], g["q2"])), g["err"], rtol=1e-12)
