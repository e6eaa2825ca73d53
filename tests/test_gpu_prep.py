"""The data layer on the GPU (cmtip/prep_data.py:96-262) against the numpy oracle:
binning bit-identical to numpy's float32 means, clipping / trimming / assembly as
exact gathers, index-map binning exact."""
import numpy as np
import pytest
import torch

from cmtip_amd import prep_data as PD
from cmtip_amd import synth
from oracle import prep as OP
from tests.gpu_util import dev, say

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("b", [2, 3, 4, 8, 16])
def test_bin_data_bit_identical(b):
    rng = np.random.default_rng(b)
    # a wide dynamic range so that the order of the float32 additions matters
    a = (rng.random((5, 2, 16 * b, 8 * b)) * np.exp(rng.standard_normal((5, 2, 16 * b, 8 * b)) * 4)).astype(np.float32)
    say(f"bin_data b={b}")
    got = PD.bin_data(a, b)
    want = OP.bin_data(a, b)
    assert got.dtype == np.float32 and got.shape == want.shape
    assert np.array_equal(got, want)
    # flattened layout (load_h5's (n, 1, P)) with det_shape
    flat = a.reshape(5, 1, -1)
    g2 = PD.bin_data(flat, b, det_shape=(2, 16 * b, 8 * b))
    assert np.array_equal(g2, OP.bin_data(flat, b, det_shape=(2, 16 * b, 8 * b)))


def test_bin_data_full_size_stack():
    # 2000 images of 256 x 256 pixels binned by 2 (the 128 x 128 images of the M = 81
    # workload), Poisson counts, on the device
    d = dev()
    g = torch.Generator(device=d).manual_seed(3)
    x = torch.poisson(torch.rand((2000, 1, 256 * 256), device=d, generator=g) * 30, generator=g)
    say("bin_data 2000 x 256^2 by 2")
    got = PD.bin_data(x, 2, det_shape=(1, 256, 256))
    assert isinstance(got, torch.Tensor) and got.shape == (2000, 1, 128 * 128)
    want = OP.bin_data(x.cpu().numpy(), 2, det_shape=(1, 256, 256))
    assert np.array_equal(got.cpu().numpy(), want)


@pytest.mark.parametrize("res_limit", [9.0, 12.0, 20.0, 1e6])
def test_clip_data_and_dataset(res_limit):
    imap, s, det_shape, ext = synth.panel_detector()
    n = 7
    imgs = np.random.default_rng(5).random((n, 1, s.shape[-1])).astype(np.float32)
    say(f"clip res_limit={res_limit}")
    got = PD.clip_data(imgs, s, res_limit)
    want = OP.clip_data(imgs, s, res_limit)
    assert got.shape == want.shape and np.array_equal(got, want)
    gp, gi = PD.clip_dataset(s, imgs, res_limit)
    wp, wi = OP.clip_dataset(s, imgs, res_limit)
    assert np.array_equal(gp, wp) and np.array_equal(gi, wi)
    # several panels: the selection of panel 0 applies to every panel (prep_data.py:167-168)
    s3 = np.concatenate([s[:, :, :4096]] * 3, axis=1)
    arr = np.random.default_rng(6).random((2, 3, 4096)).astype(np.float32)
    assert np.array_equal(PD.clip_data(arr, s3, res_limit), OP.clip_data(arr, s3, res_limit))


def test_clip_dataset_zero_limit_passthrough():
    s, ext = synth.detector(16)
    imgs = np.ones((2, 1, 256), np.float32)
    p, i = PD.clip_dataset(s, imgs, 0)
    assert p is s and i is imgs


def test_assemble_image_stack_batch():
    imap, s, det_shape, ext = synth.panel_detector(px=32, py=24, grid=(2, 3), gap=3)
    stack = np.random.default_rng(7).random((4,) + det_shape).astype(np.float32)
    say("assemble 2 x 3 panels")
    got = PD.assemble_image_stack_batch(stack, imap)
    want = OP.assemble_image_stack_batch(stack, imap)
    assert got.dtype == np.float64 and np.array_equal(got, want)
    # colliding pixels keep the last one written
    im2 = imap.copy()
    im2[1, :5, :5] = im2[0, :5, :5]
    assert np.array_equal(PD.assemble_image_stack_batch(stack, im2), OP.assemble_image_stack_batch(stack, im2))


@pytest.mark.parametrize("res_limit", [12.0, 20.0, 0, 1.0])
def test_trim_dataset(res_limit):
    imap, s, det_shape, ext = synth.panel_detector(px=64, py=64, grid=(2, 2), gap=4)
    n = 6
    imgs = np.random.default_rng(8).random((n, 1, s.shape[-1])).astype(np.float32)
    say(f"trim res_limit={res_limit}")
    gp, gi = PD.trim_dataset(imap, s, imgs, det_shape, res_limit)
    wp, wi = OP.trim_dataset(imap, s, imgs, det_shape, res_limit)
    assert gp.shape == wp.shape and gi.shape == wi.shape
    assert gp.dtype == wp.dtype and gi.dtype == wi.dtype
    assert np.array_equal(gp, wp) and np.array_equal(gi, wi)


@pytest.mark.parametrize("b", [2, 4])
def test_bin_pixel_index_map(b):
    imap, s, det_shape, ext = synth.panel_detector(px=32, py=48, grid=(2, 2), gap=5)
    # a scrambled map as well, so the minimum over each 2 x 2 corner is exercised
    rng = np.random.default_rng(9)
    scr = rng.integers(0, 500, imap.shape).astype(np.int64)
    for m in (imap, scr):
        got = PD.bin_pixel_index_map(m, b)
        want = OP.bin_pixel_index_map(m, b)
        assert got.dtype == want.dtype and np.array_equal(got, want)
