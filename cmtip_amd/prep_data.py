"""The data layer on the GPU (cmtip/prep_data.py:96-262): detector binning, resolution
clipping and trimming, panel assembly and index-map binning.

Same names, arguments and return layouts as the reference.  The image data move only
through the library's kernels (``cmtip_bin_images``, ``cmtip_resolution_select``,
``cmtip_gather_columns``, ``cmtip_assemble_index``, ``cmtip_bin_index_map``); the few
scalar decisions of trim_dataset (the trimming width) are taken on the host from the
assembled pixel positions, as small arrays.  numpy inputs give numpy results with the
reference's dtypes (trim_dataset's assembled arrays are float64, as its
np.zeros-based assembly makes them); tensor inputs stay on the device.  h5 loading and
checkpointing (prep_data.py:11-93, :265-316) are not part of this module.
"""
import numpy as np
import torch

from . import _lib
from ._lib import check, ptr, stream


def _dev_f(x):
    """Device tensor of x, float32 unless x is float64."""
    t = x if isinstance(x, torch.Tensor) else torch.from_numpy(np.ascontiguousarray(x))
    if t.dtype not in (torch.float32, torch.float64):
        t = t.float()
    return t.to(_lib.device()).contiguous()


def _ret(t, like, dtype=None):
    if isinstance(like, torch.Tensor):
        return t
    a = t.cpu().numpy()
    return a.astype(dtype, copy=False) if dtype is not None else a


def bin_data(arr, bin_factor, det_shape=None):
    """Bin detector data by averaging bin_factor x bin_factor blocks (prep_data.py:197-226).

    arr: (n_images, n_panels, panel_x, panel_y), or (n_images, 1, n_pixels) with
    det_shape = (n_panels, panel_x, panel_y); returns the binned array in the same
    layout.  float32, in numpy's order of operations (bit-identical means)."""
    a = _dev_f(arr)
    n = a.shape[0]
    if det_shape is not None:
        a = a.reshape((n,) + tuple(det_shape))
    assert a.shape[2] % bin_factor == 0
    assert a.shape[3] % bin_factor == 0
    if a.dtype != torch.float32:
        raise TypeError("bin_data on the GPU takes float32 data (load_h5 stores float32, prep_data.py:54-55)")
    npan, nx, ny = a.shape[1], a.shape[2], a.shape[3]
    b = int(bin_factor)
    out = torch.empty((n, npan, nx // b, ny // b), dtype=a.dtype, device=a.device)
    check(_lib.lib().cmtip_bin_images(ptr(a), n * npan, nx, ny, b, ptr(out), stream()), "cmtip_bin_images")
    if det_shape is not None:
        out = out.reshape(n, 1, -1)
    return _ret(out, arr)


def bin_pixel_index_map(arr, bin_factor):
    """Binned pixel index map (prep_data.py:229-242): (n_panels, x / b, y / b, 2)."""
    a = (arr if isinstance(arr, torch.Tensor) else torch.from_numpy(np.ascontiguousarray(arr))).to(
        _lib.device(), torch.int64).contiguous()
    npan, px, py = a.shape[0], a.shape[1], a.shape[2]
    b = int(bin_factor)
    out = torch.empty((npan, px // b, py // b, 2), dtype=torch.int64, device=a.device)
    check(_lib.lib().cmtip_bin_index_map(ptr(a), npan, px, py, b, ptr(out), stream()), "cmtip_bin_index_map")
    return out if isinstance(arr, torch.Tensor) else out.cpu().numpy().astype(np.asarray(arr).dtype, copy=False)


def _select(pixel_position_reciprocal, res_limit):
    """Device indices of the pixels of panel 0 whose resolution exceeds res_limit
    (prep_data.py:166-167)."""
    s = _dev_f(pixel_position_reciprocal).float()
    npan, P = s.shape[1], s.shape[2]
    flags = torch.empty(P, dtype=torch.uint8, device=s.device)
    idx = torch.empty(P, dtype=torch.int64, device=s.device)
    cnt = torch.zeros(1, dtype=torch.int64, device=s.device)
    check(_lib.lib().cmtip_resolution_select(ptr(s), npan * P, P, float(res_limit), ptr(flags), ptr(idx), ptr(cnt),
                                             stream()), "cmtip_resolution_select")
    return idx[:int(cnt.item())]


def _gather(a, idx):
    """a[..., idx] along the last axis (a: device tensor, rows contiguous)."""
    lead = tuple(a.shape[:-1])
    rows = int(np.prod(lead)) if lead else 1
    out = torch.empty(lead + (idx.numel(),), dtype=a.dtype, device=a.device)
    check(_lib.lib().cmtip_gather_columns(ptr(a), a.element_size(), rows, a.shape[-1], ptr(idx), idx.numel(),
                                          ptr(out), stream()), "cmtip_gather_columns")
    return out


def clip_data(arr, pixel_position_reciprocal, res_limit):
    """Keep the pixels whose resolution (from panel 0's positions) exceeds res_limit
    (prep_data.py:155-169): arr[:, :, indices]."""
    idx = _select(pixel_position_reciprocal, res_limit)
    return _ret(_gather(_dev_f(arr), idx), arr)


def clip_dataset(pixel_position_reciprocal, intensities, res_limit):
    """clip_data on the intensities and on the positions (prep_data.py:172-194)."""
    if res_limit == 0:
        return pixel_position_reciprocal, intensities
    idx = _select(pixel_position_reciprocal, res_limit)
    return (_ret(_gather(_dev_f(pixel_position_reciprocal), idx), pixel_position_reciprocal),
            _ret(_gather(_dev_f(intensities), idx), intensities))


def _assembled_index(index_map):
    """(flat source pixel of every assembled detector position or -1, assembled shape)
    from the index map (n_panels, x, y, 2) (assemble_image_stack_batch, prep_data.py:245-262)."""
    im = np.asarray(index_map.cpu() if isinstance(index_map, torch.Tensor) else index_map)
    X, Y = int(im[:, :, :, 0].max()) + 1, int(im[:, :, :, 1].max()) + 1
    d = torch.from_numpy(np.ascontiguousarray(im, dtype=np.int64)).to(_lib.device())
    src = torch.empty(X * Y, dtype=torch.int64, device=d.device)
    check(_lib.lib().cmtip_assemble_index(ptr(d), im.shape[0] * im.shape[1] * im.shape[2], X, Y, ptr(src), stream()),
          "cmtip_assemble_index")
    return src, X, Y


def assemble_image_stack_batch(image_stack, index_map):
    """Assemble panels into 2-D images through the index map (prep_data.py:245-262):
    (n, n_panels, x, y) -> (n, X, Y), zero where no pixel lands (float64 for numpy input,
    as the reference's np.zeros makes it)."""
    src, X, Y = _assembled_index(index_map)
    a = _dev_f(image_stack)
    n = a.shape[0]
    out = _gather(a.reshape(n, -1), src).reshape(n, X, Y)
    return _ret(out, image_stack, np.float64)


def trim_dataset(pixel_index_map, pixel_position_reciprocal, intensities, det_shape, res_limit):
    """Trim the assembled detector to the square whose corner pixel has the requested
    resolution, keeping only positions a pixel lands on (prep_data.py:96-152).  Returns
    (pixel_position_reciprocal (3, 1, K), intensities (n, 1, K))."""
    if res_limit == 0:
        return pixel_position_reciprocal, intensities
    ppr_h = np.asarray(pixel_position_reciprocal.cpu() if isinstance(pixel_position_reciprocal, torch.Tensor)
                       else pixel_position_reciprocal)
    data_max_res = 1e10 / np.linalg.norm(ppr_h, axis=0).max()
    if res_limit < data_max_res:
        print("Warning: requested resolution exceeds data resolution.")
        return pixel_position_reciprocal, intensities
    src, X, Y = _assembled_index(pixel_index_map)
    # assembled positions (small: 3 x X x Y) decide the trimming width on the host
    s = _dev_f(ppr_h).reshape(3, -1)
    pos = _gather(s, src).double().reshape(3, X, Y).cpu().numpy()
    center = np.array(0.5 * np.array([X, Y])).astype(int)
    min_pix = 20
    img_width = range(min_pix, np.min(center))
    norms = np.linalg.norm(pos, axis=0)
    res = np.array([1e10 / norms[center[0] - i, center[1] - i] for i in img_width])
    nkeep = int(np.argmin(np.abs(res - res_limit))) + min_pix
    win = src.reshape(X, Y)[center[0] - nkeep:center[0] + nkeep, center[1] - nkeep:center[1] + nkeep].reshape(-1)
    keep = win[win >= 0].contiguous()
    n = intensities.shape[0]
    it = _gather(_dev_f(intensities).reshape(n, -1), keep).reshape(n, 1, -1)
    pp = _gather(s, keep).reshape(3, 1, -1)
    return _ret(pp, pixel_position_reciprocal, np.float64), _ret(it, intensities, np.float64)
