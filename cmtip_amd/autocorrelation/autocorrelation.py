"""Autocorrelation solve on the GPU (cmtip/autocorrelation/autocorrelation.py).

Signatures and defaults follow the reference; extra knobs are keyword-only.  The
two type-1 spreads run once per solve; each CG iteration applies the Toeplitz normal
operator (five FFT passes on the L^3 embedding) and the Fourier-corner regulariser
(five passes on M^3), with FP64 CG scalars, all on the device.

Vectors are real float32 (the reference's CG vectors are complex with zero
imaginary part, asserted at :45-46, :102-103, :241-242); use_recip_sym=False is not
supported by the solver.
"""
import ctypes

import numpy as np
import torch

from .. import _arrays, _lib, synth
from .._lib import check, ptr, stream
from ..nufft.adjoint import _adjoint, assert_points_in_range
from ..nufft.plan import DEFAULT_EPS, get_plan

_SOLVERS = {}
_ANTISUPPORT = {}


class ACSolver:
    """Device state of one W x = d system: FFT plans, kernel spectrum, antisupport,
    support and CG work vectors (setup_linops, autocorrelation.py:113-173)."""

    def __init__(self, M, L=None):
        self.M = int(M)
        self.L = int(L) if L is not None else 2 * self.M
        h = ctypes.c_void_p()
        check(_lib.lib().cmtip_acsolver_create(self.M, self.L, ctypes.byref(h)), "cmtip_acsolver_create")
        self._h = h
        self.support = None

    def set_support(self, support):
        self.support = _arrays.to_dev(support, torch.float32).contiguous()
        check(_lib.lib().cmtip_acsolver_set_support(self._h, ptr(self.support)), "cmtip_acsolver_set_support")

    def set_kernel(self, ugrid_conv):
        k = _arrays.to_dev(ugrid_conv, torch.float32)
        check(_lib.lib().cmtip_acsolver_set_kernel(self._h, ptr(k), stream()), "cmtip_acsolver_set_kernel")

    def set_spectrum(self, F_ugrid_conv_):
        F = F_ugrid_conv_.real if isinstance(F_ugrid_conv_, torch.Tensor) else np.real(F_ugrid_conv_)
        F = _arrays.to_dev(F, torch.float32)
        check(_lib.lib().cmtip_acsolver_set_spectrum(self._h, ptr(F), stream()), "cmtip_acsolver_set_spectrum")

    def set_antisupport(self, F_antisupport):
        m = _arrays.to_dev(F_antisupport, torch.float32)
        check(_lib.lib().cmtip_acsolver_set_antisupport(self._h, ptr(m), stream()), "cmtip_acsolver_set_antisupport")

    def conv(self, u):
        out = torch.empty(self.M ** 3, dtype=torch.float32, device=u.device)
        check(_lib.lib().cmtip_core_problem_convolution(self._h, ptr(u), ptr(out), stream()),
              "cmtip_core_problem_convolution")
        return out

    def reg(self, u):
        out = torch.empty(self.M ** 3, dtype=torch.float32, device=u.device)
        check(_lib.lib().cmtip_fourier_reg(self._h, ptr(u), ptr(out), stream()), "cmtip_fourier_reg")
        return out

    def matvec(self, u, alambda, rlambda, flambda):
        out = torch.empty(self.M ** 3, dtype=torch.float32, device=u.device)
        check(_lib.lib().cmtip_acsolver_matvec(self._h, ptr(u), ptr(out), float(alambda), float(rlambda),
                                               float(flambda), stream()), "cmtip_acsolver_matvec")
        return out

    def cg(self, d, x, alambda, rlambda, flambda, maxiter, rtol=1e-5):
        """scipy.sparse.linalg.cg on W from x (overwritten); returns (iterations, info)."""
        it, info = ctypes.c_int(), ctypes.c_int()
        check(_lib.lib().cmtip_acsolver_cg(self._h, ptr(d), ptr(x), float(alambda), float(rlambda), float(flambda),
                                           int(maxiter), float(rtol), ctypes.byref(it), ctypes.byref(info),
                                           stream()), "cmtip_acsolver_cg")
        return it.value, info.value

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value and _lib is not None and _lib._LIB is not None:
            _lib._LIB.cmtip_acsolver_destroy(h)
            self._h = None


def get_solver(M, L=None):
    key = (int(M), int(L) if L is not None else 2 * int(M), torch.cuda.current_device())
    s = _SOLVERS.get(key)
    if s is None:
        s = ACSolver(M, L)
        _SOLVERS[key] = s
    return s


def gen_nonuniform_positions(orientations, pixel_position_reciprocal, *, recip_extent=None):
    """Rotated reciprocal-space positions H, K, L (autocorrelation.py:12-29).

    Returns three float32 arrays of shape (n_images,) + pixel shape[1:], computed on the
    device in FP64 and rounded once (the reference's einsum + astype(float32) at
    :27 and :210).  With ``recip_extent`` they come back already scaled to
    x = H / recip_extent * pi (:133-135), which is what the solver uses.
    """
    as_torch = _arrays.is_torch(orientations, pixel_position_reciprocal)
    q = _arrays.to_dev(orientations, torch.float32).reshape(-1, 4)
    if q.shape[0] == 0:
        q = torch.tensor([[1.0, 0.0, 0.0, 0.0]], dtype=torch.float32, device=q.device)  # identity (:26)
    shp = tuple(pixel_position_reciprocal.shape[1:])
    s = _arrays.to_dev(pixel_position_reciprocal, torch.float64).reshape(3, -1).contiguous()
    P = s.shape[1]
    n = q.shape[0]
    H = torch.empty(n * P, dtype=torch.float32, device=q.device)
    K = torch.empty_like(H)
    L = torch.empty_like(H)
    ext = 0.0 if recip_extent is None else float(np.float32(
        recip_extent.item() if isinstance(recip_extent, torch.Tensor) else recip_extent))
    check(_lib.lib().cmtip_rotate_coords(ptr(q), n, ptr(s), P, ext, ptr(H), ptr(K), ptr(L), stream()),
          "cmtip_rotate_coords")
    H, K, L = (t.reshape((n,) + shp) for t in (H, K, L))
    return tuple(_arrays.out(t, as_torch) for t in (H, K, L))


def _antisupport_dev(M):
    """gen_F_antisupport(M) as a float32 (0 / 1) device tensor, built once per (M, device)."""
    key = (int(M), torch.cuda.current_device())
    out = _ANTISUPPORT.get(key)
    if out is None:
        out = torch.empty((M, M, M), dtype=torch.float32, device=_lib.device())
        check(_lib.lib().cmtip_gen_F_antisupport(int(M), ptr(out), stream()), "cmtip_gen_F_antisupport")
        _ANTISUPPORT[key] = out
    return out


def gen_F_antisupport(M):
    """Boolean (M, M, M) mask, True outside the sphere of radius pi (autocorrelation.py:66-85)."""
    return (_antisupport_dev(M) > 0.5).cpu().numpy()


def _support_of(ac_estimate, M):
    """(ac_smoothed > 1e-12) with ac_smoothed = gaussian_filter(ac_estimate, 0.5) (:218-219)."""
    is_f32 = (ac_estimate.dtype == torch.float32) if isinstance(ac_estimate, torch.Tensor) else (
        np.asarray(ac_estimate).dtype == np.float32)
    a = _arrays.to_dev(ac_estimate, torch.float64)
    sup = torch.empty((M, M, M), dtype=torch.float32, device=a.device)
    check(_lib.lib().cmtip_ac_support_ex(ptr(a), int(M), 1 if is_f32 else 0, ptr(sup), stream()),
          "cmtip_ac_support_ex")
    return sup


def core_problem_convolution(uvect, M, F_ugrid_conv_, M_ups, ac_support, use_recip_sym=True):
    """Toeplitz normal operator (autocorrelation.py:32-63) with the spectrum F_ugrid_conv_
    of an M_ups^3 embedding; flattened float32 result."""
    if use_recip_sym:
        assert _arrays.all_real(uvect)
    as_torch = _arrays.is_torch(uvect, F_ugrid_conv_)
    s = get_solver(M, M_ups)
    s.set_spectrum(F_ugrid_conv_)
    s.set_support(ac_support if ac_support is not None else np.ones((M,) * 3, np.float32))
    u = _arrays.to_dev(uvect.real if np.iscomplexobj(uvect) and not isinstance(uvect, torch.Tensor) else uvect,
                       torch.float32).reshape(-1)
    return _arrays.out(s.conv(u), as_torch)


def fourier_reg(uvect, support, F_antisupport, M, use_recip_sym):
    """Projection onto the Fourier corners (autocorrelation.py:88-110); flattened float32."""
    as_torch = _arrays.is_torch(uvect, support)
    s = get_solver(M)
    s.set_antisupport(F_antisupport)
    s.set_support(support)
    u = _arrays.to_dev(uvect.real if np.iscomplexobj(uvect) and not isinstance(uvect, torch.Tensor) else uvect,
                       torch.float32).reshape(-1)
    out = s.reg(u)
    s.set_antisupport(torch.as_tensor(gen_F_antisupport(M), dtype=torch.float32))
    return _arrays.out(out, as_torch)


class WOperator:
    """W of setup_linops (autocorrelation.py:150-160) as a device operator."""

    def __init__(self, solver, alambda, rlambda, flambda):
        self.solver = solver
        self.alambda, self.rlambda, self.flambda = alambda, rlambda, flambda
        self.shape = (solver.M ** 3, solver.M ** 3)
        self.dtype = np.float32

    def matvec(self, uvect):
        as_torch = _arrays.is_torch(uvect)
        u = _arrays.to_dev(uvect, torch.float32).reshape(-1)
        return _arrays.out(self.solver.matvec(u, self.alambda, self.rlambda, self.flambda), as_torch)


def setup_linops(H, K, L, data, ac_support, weights, x0, M, Mtot, N, reciprocal_extent, alambda, rlambda, flambda,
                 use_gpu=False, use_recip_sym=True, *, L_embed=None, eps=DEFAULT_EPS, scaled=False, autotune=False,
                 rot=None):
    """(W, d) of W x = d (autocorrelation.py:113-173); d is a float32 device vector.

    With ``scaled`` the coordinates are already x = H / reciprocal_extent * pi.  With
    ``autotune`` the two spreads pick their bin and launch settings with the NUFFT
    autotuner (first call per shape; later calls reuse the choice).  ``weights`` None
    means unit weights.  ``rot`` (quats, s, reciprocal_extent) as device tensors: the
    points are the rotated pixels gen_nonuniform_positions would return, binned straight
    from the rotations (H, K, L are then not used and may be None; no autotune)."""
    if not use_recip_sym:
        raise NotImplementedError("the GPU solver keeps real vectors (use_recip_sym=True)")
    if rot is not None:
        H_ = K_ = L_ = None
    elif scaled:
        H_, K_, L_ = (_arrays.to_dev(t, torch.float32).reshape(-1) for t in (H, K, L))
    else:
        ext = np.float32(reciprocal_extent)
        H_, K_, L_ = (_arrays.to_dev(t, torch.float32).reshape(-1) / float(ext) * float(np.float32(np.pi))
                      for t in (H, K, L))
    solver = get_solver(M, L_embed)
    Lm = solver.L
    sup = _arrays.to_dev(ac_support, torch.float32)
    solver.set_support(sup)
    solver.set_antisupport(_antisupport_dev(M))  # the mask stays on the device
    # kernel: real type-1 of ones at M_ups modes over (2M)^3 (:141-148; 1/(2M)^3 also when L > 2M)
    ugrid_conv = _adjoint(None, H_, K_, L_, Lm, True, None, eps, scale=1.0 / float(2 * M) ** 3, ones=True,
                          autotune=autotune, rot=rot, check_range=False)
    solver.set_kernel(ugrid_conv)
    # rhs: d = alambda * A^H (data * weights) + rlambda * x0 (:162-171)
    nuvect_Db = _arrays.to_dev(data, torch.float32).reshape(-1)
    if weights is not None:  # None: unit weights, the product is the data itself
        nuvect_Db = nuvect_Db * _arrays.to_dev(weights, torch.float32).reshape(-1)
    uvect_ADb = _adjoint(nuvect_Db, H_, K_, L_, M, True, sup, eps, autotune=autotune, rot=rot,
                         check_range=False).reshape(-1)
    has_nan = torch.isnan(uvect_ADb).any()
    # the range checks of both spreads (adjoint.py:29,66) after both are queued: one
    # synchronisation for the two checks and the NaN test
    assert_points_in_range(get_plan(1, Lm, +1, eps), get_plan(1, M, +1, eps))
    if bool(has_nan):
        print("Warning: nans in the adjoint calculation; intensities may be too large", flush=True)
    x0d = _arrays.to_dev(x0, torch.float32).reshape(-1)
    d = alambda * uvect_ADb + rlambda * x0d
    return WOperator(solver, alambda, rlambda, flambda), d


def solve_ac(generation, pixel_position_reciprocal, reciprocal_extent, slices_, M, orientations=None,
             ac_estimate=None, use_gpu=False, use_recip_sym=True, *, maxiter=100, rlambda=None, flambda=1e3,
             alambda=1, L=None, eps=DEFAULT_EPS, rtol=1e-5, seed=None, return_info=False, autotune=False):
    """Autocorrelation (M, M, M) from images and orientations (autocorrelation.py:176-248).

    Defaults follow the reference (alpha 1, rho = M^3/N/1000, phi 1e3, maxiter 100,
    rtol 1e-5).  ``L`` picks the Toeplitz embedding (even, >= 2M; default 2M);
    ``autotune`` tunes the spreads' bin settings (see ``setup_linops``).  The
    result dtype follows the reference's CG dtype: float64 when the initial guess is
    float64 (including the zero guess of generation 0), else float32.
    """
    if not use_recip_sym:
        raise NotImplementedError("the GPU solver keeps real vectors (use_recip_sym=True)")
    as_torch = _arrays.is_torch(slices_, ac_estimate)
    Mtot = M ** 3
    N_images = slices_.shape[0]
    N = int(np.prod(slices_.shape))
    if orientations is None:
        # the reference draws sk.get_random_quat(N_images) (:208); skopi is not available
        orientations = synth.random_quaternions(N_images, seed if seed is not None else np.random.randint(2 ** 31))
    ext = np.float32(reciprocal_extent.item() if isinstance(reciprocal_extent, torch.Tensor) else reciprocal_extent)
    dev = _lib.device()
    q_d = _arrays.to_dev(orientations, torch.float32).reshape(-1, 4)
    s_d = _arrays.to_dev(pixel_position_reciprocal, torch.float64).reshape(3, -1).contiguous()
    rot = None
    if autotune:
        # the autotuner bins the points again per setting: coordinate arrays (12 B per point)
        H_, K_, L_ = gen_nonuniform_positions(q_d, s_d, recip_extent=ext)
        H_, K_, L_ = (t.reshape(-1) for t in (H_, K_, L_))
    else:
        # the spreads bin the rotated pixels straight from (orientations, pixel positions):
        # the coordinates gen_nonuniform_positions would write, never stored
        H_ = K_ = L_ = None
        rot = (q_d, s_d, float(ext))
    data = _arrays.to_dev(slices_, torch.float32).reshape(-1)
    if ac_estimate is None:
        ac_support = torch.ones((M,) * 3, dtype=torch.float32, device=dev)
        x0 = torch.zeros(Mtot, dtype=torch.float32, device=dev)
        out_dtype = np.float64
    else:
        ac_support = _support_of(ac_estimate, M)
        # the reference masks the caller's estimate in place (:220)
        if isinstance(ac_estimate, torch.Tensor):
            ac_estimate *= ac_support.to(ac_estimate.device, ac_estimate.dtype)
        else:
            ac_estimate *= ac_support.cpu().numpy().astype(ac_estimate.dtype)
        x0 = _arrays.to_dev(ac_estimate, torch.float32).reshape(-1).clone()
        est_dtype = ac_estimate.dtype
        out_dtype = np.float32 if est_dtype in (np.float32, torch.float32) else np.float64
    weights = None  # the reference's np.ones(N) (:221): multiplying by it is the identity
    if rlambda is None:
        rlambda = Mtot / N / 1000
    W, d = setup_linops(H_, K_, L_, data, ac_support, weights, x0, M, Mtot, N, ext, alambda, rlambda, flambda,
                        use_gpu=True, use_recip_sym=True, L_embed=L, eps=eps, scaled=True, autotune=autotune,
                        rot=rot)
    x = x0.clone()
    iters, info = W.solver.cg(d, x, alambda, rlambda, flambda, maxiter, rtol)
    ac = x.reshape((M,) * 3)
    if as_torch:
        res = ac.to(torch.float64) if out_dtype == np.float64 else ac
    else:
        res = ac.cpu().numpy().astype(out_dtype)
    if return_info:
        return res, {"iterations": iters, "info": info}
    return res


def solve_ac_mult(generation, pixel_position_reciprocal, reciprocal_extent, slices_, M, n_conformations,
                  orientations=None, conformations=None, ac_estimate=None, use_gpu=False, use_recip_sym=True, **kw):
    """One autocorrelation per conformation (n_conformations, M, M, M), float64.

    The single-GPU share of solve_ac_mpi_mult (autocorrelation_mpi.py:193-243): images
    with no conformation assigned are given a random one (np.random.randint, :218-219),
    then conformation nc is solved from the images and orientations assigned to it
    (``slices_[conformations == nc]``, :234-243) with ``ac_estimate[nc]`` as its
    estimate.  Keyword arguments are passed on to solve_ac (maxiter, rlambda, ...).
    Orientations must be given when skopi's random draw (:222) is wanted otherwise;
    without them solve_ac draws seeded uniform ones per conformation."""
    n_img = slices_.shape[0]
    if conformations is None:
        conformations = np.random.randint(low=0, high=n_conformations, size=n_img)
    conf = np.asarray(conformations.cpu() if isinstance(conformations, torch.Tensor) else conformations)
    if ac_estimate is None:
        ac_estimate = n_conformations * [None]
    as_torch = _arrays.is_torch(slices_)
    out = []
    for nc in range(n_conformations):
        sel = np.flatnonzero(conf == nc)
        idx = torch.as_tensor(sel, device=slices_.device) if isinstance(slices_, torch.Tensor) else sel
        ori = None
        if orientations is not None:
            ori = orientations[torch.as_tensor(sel, device=orientations.device)] if isinstance(
                orientations, torch.Tensor) else np.asarray(orientations)[sel]
        est = ac_estimate[nc]
        res = solve_ac(generation, pixel_position_reciprocal, reciprocal_extent, slices_[idx], M, orientations=ori,
                       ac_estimate=est, use_gpu=use_gpu, use_recip_sym=use_recip_sym, **kw)
        out.append(torch.as_tensor(res, dtype=torch.float64) if as_torch else np.asarray(res, dtype=np.float64))
    return torch.stack(out) if as_torch else np.stack(out)
