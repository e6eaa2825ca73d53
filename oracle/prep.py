"""The data layer restated with numpy (TEST INFRASTRUCTURE ONLY).

Follows cmtip/prep_data.py:96-262: trim_dataset (:96-152), clip_data (:155-169),
clip_dataset (:172-194), bin_data (:197-226), bin_pixel_index_map (:229-242) and
assemble_image_stack_batch (:245-262).  The h5 loading and checkpointing of that file
are outside the hot path and not restated.
"""
import numpy as np


def assemble_image_stack_batch(image_stack, index_map):
    """prep_data.py:245-262."""
    index_max_x = np.max(index_map[:, :, :, 0]) + 1
    index_max_y = np.max(index_map[:, :, :, 1]) + 1
    stack_num = image_stack.shape[0]
    panel_num = image_stack.shape[1]
    image = np.zeros((stack_num, index_max_x, index_max_y))
    for l in range(panel_num):
        image[:, index_map[l, :, :, 0], index_map[l, :, :, 1]] = image_stack[:, l, :, :]
    return image


def trim_dataset(pixel_index_map, pixel_position_reciprocal, intensities, det_shape, res_limit):
    """prep_data.py:96-152 (the warning print of :117 omitted)."""
    if res_limit == 0:
        return pixel_position_reciprocal, intensities
    data_max_res = 1e10 / np.linalg.norm(pixel_position_reciprocal, axis=0).max()
    if res_limit < data_max_res:
        return pixel_position_reciprocal, intensities
    intensities = intensities.reshape((intensities.shape[0],) + det_shape)
    mask = np.expand_dims(np.ones_like(intensities[0]), axis=0)
    mask = assemble_image_stack_batch(mask, pixel_index_map)
    intensities = assemble_image_stack_batch(intensities, pixel_index_map)
    pixel_position_reciprocal = pixel_position_reciprocal.reshape((3,) + det_shape)
    pixel_position_reciprocal = assemble_image_stack_batch(pixel_position_reciprocal, pixel_index_map)
    assembled_shape = np.array([pixel_index_map[:, :, :, 0].max() + 1, pixel_index_map[:, :, :, 1].max() + 1])
    center = np.array(0.5 * assembled_shape).astype(int)
    min_pix = 20
    img_width = range(min_pix, np.min(center))
    norms = np.linalg.norm(pixel_position_reciprocal, axis=0)
    res = np.array([1e10 / norms[center[0] - i, center[1] - i] for i in img_width])
    nkeep = np.argmin(np.abs(res - res_limit)) + min_pix
    sx = slice(center[0] - nkeep, center[0] + nkeep)
    sy = slice(center[1] - nkeep, center[1] + nkeep)
    intensities = intensities[:, sx, sy]
    mask = mask[:, sx, sy]
    pixel_position_reciprocal = pixel_position_reciprocal[:, sx, sy]
    intensities = intensities[:, np.where(mask == 1)[1], np.where(mask == 1)[2]]
    pixel_position_reciprocal = pixel_position_reciprocal[:, np.where(mask == 1)[1], np.where(mask == 1)[2]]
    return np.expand_dims(pixel_position_reciprocal, axis=1), np.expand_dims(intensities, axis=1)


def clip_data(arr, pixel_position_reciprocal, res_limit):
    """prep_data.py:155-169."""
    resolution = 1.0 / np.linalg.norm(pixel_position_reciprocal, axis=0) * 1e10
    indices = np.where(resolution[0] > res_limit)[0]
    return arr[:, :, indices]


def clip_dataset(pixel_position_reciprocal, intensities, res_limit):
    """prep_data.py:172-194."""
    if res_limit == 0:
        return pixel_position_reciprocal, intensities
    intensities = clip_data(intensities, pixel_position_reciprocal, res_limit)
    pixel_position_reciprocal = clip_data(pixel_position_reciprocal, pixel_position_reciprocal, res_limit)
    return pixel_position_reciprocal, intensities


def bin_data(arr, bin_factor, det_shape=None):
    """prep_data.py:197-226."""
    if det_shape is not None:
        arr = np.array([arr[i].reshape(det_shape) for i in range(arr.shape[0])])
    assert arr.shape[2] % bin_factor == 0
    assert arr.shape[3] % bin_factor == 0
    binned_arr = arr.reshape(arr.shape[0], arr.shape[1], int(arr.shape[2] / bin_factor), bin_factor,
                             int(arr.shape[3] / bin_factor), bin_factor).mean(-1).mean(3)
    if det_shape is not None:
        binned_arr = binned_arr.reshape((binned_arr.shape[0], 1) + (np.prod(np.array(binned_arr.shape[1:])),))
    return binned_arr


def bin_pixel_index_map(arr, bin_factor):
    """prep_data.py:229-242."""
    arr = np.moveaxis(arr, -1, 0)
    arr = np.minimum(arr[..., ::bin_factor, :], arr[..., 1::bin_factor, :])
    arr = np.minimum(arr[..., ::bin_factor], arr[..., 1::bin_factor])
    arr = arr // bin_factor
    return np.moveaxis(arr, 0, -1)
