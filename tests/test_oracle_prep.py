"""The data-layer oracle (oracle/prep.py) pinned by explicit loops and closed forms
(no GPU)."""
import numpy as np

from cmtip_amd import synth
from oracle import prep as OP


def test_bin_data_is_the_block_mean():
    rng = np.random.default_rng(0)
    a = rng.random((3, 2, 12, 8)).astype(np.float32)
    got = OP.bin_data(a, 4)
    want = np.zeros((3, 2, 3, 2))
    for i in range(3):
        for j in range(2):
            want[:, :, i, j] = a[:, :, 4 * i:4 * i + 4, 4 * j:4 * j + 4].astype(np.float64).mean(axis=(2, 3))
    assert got.dtype == np.float32 and got.shape == (3, 2, 3, 2)
    np.testing.assert_allclose(got, want, rtol=1e-6)
    # flattened layout with det_shape, exact on integers
    b = np.arange(2 * 4 * 4, dtype=np.float32).reshape(2, 1, 16)
    g = OP.bin_data(b, 2, det_shape=(1, 4, 4))
    assert g.shape == (2, 1, 4)
    np.testing.assert_array_equal(g[0, 0], [2.5, 4.5, 10.5, 12.5])


def test_clip_data_keeps_the_low_resolution_disc():
    # resolution falls with |s|, so the kept pixels are those inside a radius
    s, ext = synth.detector(32)
    res_limit = 12.0
    got = OP.clip_data(s, s, res_limit)
    mag = np.linalg.norm(s.astype(np.float64)[:, 0, :], axis=0)
    want_n = int(np.sum(1e10 / mag > res_limit))
    assert got.shape == (3, 1, want_n)
    np.testing.assert_array_equal(got[:, 0], s[:, 0, 1e10 / mag > res_limit])
    ps, it = OP.clip_dataset(s, np.ones((5, 1, s.shape[-1]), np.float32), res_limit)
    assert it.shape == (5, 1, want_n) and ps.shape == (3, 1, want_n)


def test_assemble_and_trim_single_panel():
    # one panel whose index map is the identity: assembly is the panel itself, and
    # trimming keeps the centred 2 nkeep square with nkeep found from the diagonal
    imap, s, det_shape, ext = synth.panel_detector(px=96, py=96, grid=(1, 1), gap=0)
    n = 3
    imgs = np.random.default_rng(1).random((n, 1, 96 * 96)).astype(np.float32)
    asm = OP.assemble_image_stack_batch(imgs.reshape((n,) + det_shape), imap)
    np.testing.assert_array_equal(asm, imgs.reshape(n, 96, 96))
    assert asm.dtype == np.float64
    res_limit = 14.0
    pp, it = OP.trim_dataset(imap, s, imgs, det_shape, res_limit)
    mag = np.linalg.norm(s.astype(np.float64).reshape(3, 96, 96), axis=0)
    c = 48
    best = min(range(20, c), key=lambda i: abs(1e10 / mag[c - i, c - i] - res_limit))
    want = imgs.reshape(n, 96, 96)[:, c - best:c + best, c - best:c + best].reshape(n, 1, -1)
    np.testing.assert_array_equal(it, want)
    assert pp.shape == (3, 1, (2 * best) ** 2)
    # a limit beyond the data's resolution returns the inputs unchanged
    pp2, it2 = OP.trim_dataset(imap, s, imgs, det_shape, 1.0)
    assert pp2 is s and it2 is imgs


def test_assemble_collisions_keep_the_last_pixel():
    imap = np.zeros((2, 1, 2, 2), np.int64)
    imap[0, 0, :, 1] = [0, 1]
    imap[1, 0, :, 1] = [1, 2]   # panel 1 pixel 0 lands on panel 0 pixel 1
    stack = np.array([[[[1.0, 2.0]], [[3.0, 4.0]]]])
    np.testing.assert_array_equal(OP.assemble_image_stack_batch(stack, imap), [[[1.0, 3.0, 4.0]]])


def test_bin_pixel_index_map_identity():
    imap, s, det_shape, ext = synth.panel_detector(px=8, py=6, grid=(1, 2), gap=2)
    got = OP.bin_pixel_index_map(imap, 2)
    assert got.shape == (2, 4, 3, 2)
    # the identity-like map bins to the binned coordinates
    np.testing.assert_array_equal(got[0, :, :, 0], np.repeat(np.arange(4)[:, None], 3, 1))
    np.testing.assert_array_equal(got[1, :, :, 1], (8 + np.arange(0, 6, 2))[None, :].repeat(4, 0) // 2)
