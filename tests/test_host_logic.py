"""Host-side logic that needs no GPU: packed matching keys, reference-count rounding,
the pixel order that forms 8 x 8 tiles, the autotuner's candidate bounds and the bench
contract for more than one GPU."""
import json
import os
import subprocess
import sys

import numpy as np
import torch

from cmtip_amd import synth
from cmtip_amd.alignment import align as AL
from cmtip_amd.nufft.plan import spread_candidates

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _pack(d2, r):
    # the kernels' key: (float32 bits of max(d2, 0)) << 32 | r, compared as unsigned
    bits = np.maximum(np.asarray(d2, np.float32), np.float32(0)).view(np.uint32).astype(np.uint64)
    return (bits << np.uint64(32)) | np.asarray(r, np.uint64)


def test_key_order_is_distance_then_index():
    rng = np.random.default_rng(0)
    d2 = np.round(rng.random(5000) * 50).astype(np.float32) / 8  # many ties
    r = rng.permutation(5000).astype(np.uint64)
    keys = _pack(d2, r)
    order = np.argsort(keys, kind="stable")
    want = np.lexsort((r, d2))  # by d2, then by smaller reference index (np.argmin)
    np.testing.assert_array_equal(order, want)
    # negative squared distances clamp to 0 (sklearn's clamp)
    assert _pack([-1.5], [3])[0] == _pack([0.0], [3])[0]


def test_key_helpers_unpack():
    d2 = np.array([0.0, 1.5, 3.25e-7, 8e4], np.float32)
    r = np.array([7, 0, 123456, 4294967294], np.uint64)
    keys = torch.from_numpy(_pack(d2, r).view(np.int64))
    assert torch.equal(AL.key_index(keys), torch.from_numpy(r.astype(np.int64)))
    np.testing.assert_array_equal(AL.key_d2(keys).numpy(), d2)


def test_reference_count_rounding():
    # align.py:98-101: the reference set grows to a multiple of nbatches
    assert AL._round_up_refs(10000, 1) == 10000
    assert AL._round_up_refs(10000, 3) == 10002
    assert AL._round_up_refs(7, 7) == 7
    assert AL._round_up_refs(5, 4) == 8


def test_pixel_order_tiles():
    for side in (128, 96, 64):
        s, _ = synth.detector(side)
        order = AL.pixel_order(torch.from_numpy(s.reshape(3, -1))).numpy()
        assert sorted(order.tolist()) == list(range(side * side))
        # every run of 64 is an 8 x 8 block of the panel
        blocks = order.reshape(-1, 64)
        rows, cols = blocks // side, blocks % side
        assert np.all(rows.max(1) - rows.min(1) == 7) and np.all(cols.max(1) - cols.min(1) == 7)


def test_spread_candidates_fit_lds():
    for w in (4, 7, 10):
        for typ in (1, 2):
            cands = spread_candidates(w, typ)
            assert cands, (w, typ)
            per_cell = 16 if typ == 1 else 4
            for bs, chunk, wpc in cands:
                S = bs + w - 1
                assert bs >= 2 and chunk == 0 and wpc == 0
                assert S * S * (S + 1) * per_cell + (16384 if typ == 1 else 256) <= 160 * 1024
    # the default subgrid (16^3 cells at w = 7) is among the candidates
    assert (10, 0, 0) in spread_candidates(7, 1)


def test_bench_refuses_more_than_one_gpu():
    res = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                         text=True, timeout=120)
    assert res.returncode == 2
    msg = json.loads(res.stdout.strip().splitlines()[-1])
    assert "one GPU" in msg["error"]


def _quotient_by_double_product(h, ext):
    # the kernels' coordinate scaling: fl32(fl64(h * fl64(1 / ext)))
    return (h.astype(np.float64) * (1.0 / np.float64(ext))).astype(np.float32)


def test_scaled_coordinates_equal_ieee_division():
    # rot_coord forms fl32(H) / fl32(ext) as the double product with 1 / ext rounded to
    # double; a quotient of two floats never lies within 2^-49 of a float rounding
    # midpoint, so this rounds like the IEEE float division numpy performs, bit for bit
    rng = np.random.default_rng(7)
    exts = [synth.detector(n)[1] for n in (16, 64, 128)]
    exts += list(rng.uniform(0.5, 2.0, 5).astype(np.float32) * np.float32(2.0) ** rng.integers(-30, 30, 5))
    exts += [np.float32(3.0), np.float32(1.0) + np.float32(2.0 ** -23), np.float32(2.0 ** 20)]
    for ext in exts:
        ext = np.float32(ext)
        # every mantissa pattern class: random bits over the exponents near ext, plus
        # near-integer multiples of ext (quotients close to representable values)
        bits = rng.integers(0, 2 ** 23, 400_000, dtype=np.int64)
        e = rng.integers(-20, 21, bits.size) + int(np.frexp(ext)[1]) + 126
        h = ((e << 23) | bits).astype(np.uint32).view(np.float32)
        h = np.concatenate([h, -h, (np.arange(1, 200_001, dtype=np.float32) * ext),
                            np.nextafter(np.arange(1, 50_001, dtype=np.float32) * ext, np.float32(np.inf))])
        want = h / ext
        got = _quotient_by_double_product(h, ext)
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), float(ext)


def test_float_reciprocal_scaling_is_caught():
    # the check above is sensitive: a float reciprocal (fl32(h * fl32(1 / ext))) differs
    rng = np.random.default_rng(8)
    ext = np.float32(synth.detector(128)[1])
    h = (rng.standard_normal(200_000) * ext).astype(np.float32)
    assert not np.array_equal(h * (np.float32(1.0) / ext), h / ext)
    assert np.array_equal(_quotient_by_double_product(h, ext), h / ext)


def test_rotation_derivative_matches_autograd():
    # the closed-form dR/dq of the quaternion backward (cmtip_amd.autograd) against torch's
    # own derivative of quaternion_rotation (which normalises q first): at unit q they
    # agree after projecting out the radial direction; a wrong coefficient fails
    from cmtip_amd import autograd as AG
    rng = np.random.default_rng(3)
    D = AG.rotation_derivative("cpu", torch.float64)
    for _ in range(5):
        q = torch.tensor(rng.standard_normal(4), dtype=torch.float64)
        q = q / torch.linalg.vector_norm(q)
        J = torch.autograd.functional.jacobian(AG.quaternion_rotation, q)  # (3, 3, 4)
        proj = torch.eye(4, dtype=torch.float64) - torch.outer(q, q)
        JD = torch.einsum("ijml,l->ijm", D, q) @ proj
        assert torch.allclose(J, JD, atol=1e-12)
    bad = D.clone()
    bad[1, 2, 3, 2] = -bad[1, 2, 3, 2]
    JB = torch.einsum("ijml,l->ijm", bad, q) @ proj
    assert not torch.allclose(J, JB, atol=1e-6)
