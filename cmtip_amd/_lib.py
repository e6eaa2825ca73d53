"""ctypes binding of ``libcmtip_amd.so`` (the C ABI declared in include/cmtip_amd.h).

torch is imported before the library is loaded: both link the HIP runtime under the
soname ``libamdhip64.so.7``, so the library binds to the runtime torch already
loaded and device pointers and streams are shared between them.

There is no CPU path: every entry point runs HIP kernels, and loading fails loudly
when the library has not been built or no GPU is visible.
"""
import ctypes
import os

import torch

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libcmtip_amd.so")

CMTIP_GRID_REAL = 1
CMTIP_PTS_REAL = 2
CMTIP_PTS_ONES = 4
CMTIP_PTS_IMAG = 8
CMTIP_EXEC_ADJOINT = 16
CMTIP_PTS_BOUND = 32

_vp = ctypes.c_void_p
_i = ctypes.c_int
_i64 = ctypes.c_int64
_d = ctypes.c_double
_f = ctypes.c_float
_u64 = ctypes.c_ulonglong
_pi = ctypes.POINTER(ctypes.c_int)
_pd = ctypes.POINTER(ctypes.c_double)

# name -> argument types; every function returns int status unless listed in _RESTYPE
SIGNATURES = {
    "cmtip_last_error": [],
    "cmtip_version": [],
    "cmtip_nufft_makeplan": [_i, _i, _i, _d, ctypes.POINTER(_vp)],
    "cmtip_nufft_info": [_vp, _pi, _pi, _pd],
    "cmtip_nufft_setpts": [_vp, _i64, _vp, _vp, _vp, _vp],
    "cmtip_nufft_setpts_strengths": [_vp, _i64, _vp, _vp, _vp, _vp, _vp],
    "cmtip_nufft_setpts_rotated": [_vp, _vp, _i64, _vp, _i64, _f, _vp, _vp],
    "cmtip_nufft_points_max_abs": [_vp, ctypes.POINTER(ctypes.c_float), _vp],
    "cmtip_nufft_execute": [_vp, _vp, _vp, _i, _vp, _f, _vp],
    "cmtip_nufft_destroy": [_vp],
    "cmtip_nufft_set_bins": [_vp, _i, _i, _i],
    "cmtip_nufft_get_bins": [_vp, _pi, _pi, _pi],
    "cmtip_nufft_autotune": [_vp, _vp, _vp, _i, _vp, _f, _vp, _i, _i, _f, _vp, _pi, _vp],
    "cmtip_slice_set_route": [_vp, _i],
    "cmtip_slice_get_route": [_vp, _pi, _pi],
    "cmtip_slice_timer": [_vp, _i, ctypes.POINTER(ctypes.c_float), _pi],
    "cmtip_rotate_coords": [_vp, _i64, _vp, _i64, _f, _vp, _vp, _vp, _vp],
    "cmtip_slice_prepare": [_vp, _vp, _vp, _vp],
    "cmtip_slice_real": [_vp, _vp, _i64, _vp, _vp, _i64, _f, _vp, _vp, _vp, _vp],
    "cmtip_weighted_norms": [_vp, _i64, _i64, _vp, _vp, _vp],
    "cmtip_match_prepare_images": [_vp, _i64, _i64, _vp, _vp, _vp, _vp],
    "cmtip_image_stats": [_vp, _i64, _i64, _vp, _vp, _vp, _vp],
    "cmtip_image_stats_noalloc": [_vp, _i64, _i64, _vp, _vp, _vp],
    "cmtip_phaser_create": [_i, ctypes.POINTER(_vp)],
    "cmtip_phaser_destroy": [_vp],
    "cmtip_phase_setup": [_vp, _vp, _vp, _vp, _i, _vp],
    "cmtip_phase_get_amplitudes": [_vp, _vp, _vp],
    "cmtip_phase_set_amplitudes": [_vp, _vp, _vp],
    "cmtip_phase_iterate": [_vp, _vp, _vp, _i, _i, _d, _d, _vp],
    "cmtip_phase_shrink_wrap": [_vp, _vp, _d, _d, _vp, _vp],
    "cmtip_phase_recenter": [_vp, _vp, _vp, _pi, _vp],
    "cmtip_phase_ac": [_vp, _vp, _vp, _vp],
    "cmtip_bin_images": [_vp, _i64, _i, _i, _i, _vp, _vp],
    "cmtip_resolution_select": [_vp, _i64, _i64, _d, _vp, _vp, _vp, _vp],
    "cmtip_gather_columns": [_vp, _i, _i64, _i64, _vp, _i64, _vp, _vp],
    "cmtip_assemble_index": [_vp, _i64, _i64, _i64, _vp, _vp],
    "cmtip_bin_index_map": [_vp, _i64, _i, _i, _i, _vp, _vp],
    "cmtip_match_gamma": [_vp, ctypes.c_double, _vp, ctypes.c_double, _i64, _vp, _vp],
    "cmtip_center_rows": [_vp, _i64, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp],
    "cmtip_match_prepare_images_centered": [_vp, _i64, _i64, _vp, _vp, _vp, _vp, _vp],
    "cmtip_match_prepare_images_norms": [_vp, _i64, _i64, _vp, _vp, _vp, _vp, _vp, _vp],
    "cmtip_fill_u64": [_vp, _i64, _u64, _vp],
    "cmtip_match_argmin_f32": [_vp, _i64, _i64, _vp, _i64, _i64, _vp, _vp, _vp, _vp, _vp],
    "cmtip_eudist_f32": [_vp, _i64, _vp, _i64, _i64, _vp, _vp, _vp, _vp],
    "cmtip_match_topk_bf16": [_vp, _i64, _i64, _vp, _i64, _i64, _vp, _vp, _vp, _i, _vp, _vp],
    "cmtip_match_cross_bf16": [_vp, _i64, _vp, _i64, _i64, _vp, _vp],
    "cmtip_match_topk_cross": [_vp, _i64, _i64, _i64, _vp, _vp, _vp, _i, _vp, _vp],
    "cmtip_match_rescore": [_vp, _i64, _vp, _i64, _i64, _vp, _vp, _i, _vp, _vp, _vp],
    "cmtip_image_stats_masked": [_vp, _vp, _i64, _i64, _vp, _vp, _vp, _vp],
    "cmtip_match_prepare_masked": [_vp, _vp, _i64, _i64, _vp, _vp, _vp, _vp, _vp, _vp],
    "cmtip_center_rows_sq": [_vp, _i64, _i64, _vp, _vp, _vp, _vp, _vp],
    "cmtip_match_rescore_masked": [_vp, _i64, _vp, _vp, _i64, _i64, _vp, _vp, _i, _vp, _vp, _vp],
    "cmtip_lds_read_probe": [_i, _i, _vp, _vp],
    "cmtip_lds_atomic_probe": [_i, _i, _vp, _vp],
    "cmtip_acsolver_create": [_i, _i, ctypes.POINTER(_vp)],
    "cmtip_acsolver_destroy": [_vp],
    "cmtip_ac_support": [_vp, _i, _vp, _vp],
    "cmtip_ac_support_ex": [_vp, _i, _i, _vp, _vp],
    "cmtip_acsolver_set_support": [_vp, _vp],
    "cmtip_acsolver_set_kernel": [_vp, _vp, _vp],
    "cmtip_acsolver_set_spectrum": [_vp, _vp, _vp],
    "cmtip_gen_F_antisupport": [_i, _vp, _vp],
    "cmtip_acsolver_set_antisupport": [_vp, _vp, _vp],
    "cmtip_core_problem_convolution": [_vp, _vp, _vp, _vp],
    "cmtip_fourier_reg": [_vp, _vp, _vp, _vp],
    "cmtip_acsolver_matvec": [_vp, _vp, _vp, _d, _d, _d, _vp],
    "cmtip_acsolver_cg": [_vp, _vp, _vp, _d, _d, _d, _i, _d, _pi, _pi, _vp],
}
_RESTYPE = {"cmtip_last_error": ctypes.c_char_p}

_LIB = None


class CmtipError(RuntimeError):
    """A C-ABI call returned a non-zero status."""


def load(path=LIB_PATH):
    """Load the library and declare every entry point (no GPU needed)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(path):
            raise ImportError(
                f"{path} is missing: build the HIP library first (python -m cmtip_amd.build)")
        L = ctypes.CDLL(path)
        for name, args in SIGNATURES.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = _RESTYPE.get(name, ctypes.c_int)
        _LIB = L
    return _LIB


def lib():
    """The loaded library, after checking that a GPU is visible."""
    if not torch.cuda.is_available():
        raise RuntimeError("cmtip_amd runs on a ROCm GPU only: no GPU is visible to torch")
    return load()


def check(status, what=""):
    if status != 0:
        msg = load().cmtip_last_error()
        msg = msg.decode() if msg else ""
        raise CmtipError(f"{what} failed with status {status}: {msg}")


def stream():
    """The current torch stream as a hipStream_t."""
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def ptr(t):
    """Device pointer of a tensor (or NULL for None)."""
    if t is None:
        return None
    return ctypes.c_void_p(t.data_ptr())


def device():
    return torch.device("cuda", torch.cuda.current_device())
