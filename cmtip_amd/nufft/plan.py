"""NUFFT plan objects over the C ABI (the cufinufft v1 plan the reference builds per
call, cmtip/nufft/forward.py:77-81 and cmtip/nufft/adjoint.py:73-76).

Plans own their fine grid and FFT scratch on the device.  ``get_plan`` keeps one plan
per (type, modes, sign, eps, device) so repeated calls reuse the allocations; plans
are not thread-safe.
"""
import ctypes

import torch

from .. import _lib
from .._lib import check, ptr, stream

DEFAULT_EPS = 1e-6


class NufftPlan:
    def __init__(self, nufft_type, n_modes, iflag, eps=DEFAULT_EPS):
        L = _lib.lib()
        h = ctypes.c_void_p()
        check(L.cmtip_nufft_makeplan(int(nufft_type), int(n_modes), int(iflag), float(eps), ctypes.byref(h)),
              "cmtip_nufft_makeplan")
        self._h = h
        self.type = int(nufft_type)
        self.n_modes = int(n_modes)
        self.iflag = int(iflag)
        self.eps = float(eps)
        self._pts = None
        self._pts_key = None
        self.slice_owner = None
        w, n, beta = ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
        check(L.cmtip_nufft_info(h, ctypes.byref(w), ctypes.byref(n), ctypes.byref(beta)), "cmtip_nufft_info")
        self.w, self.n_fine, self.beta = w.value, n.value, beta.value

    def setpts(self, x, y, z, force=False):
        """x, y, z: float32 device tensors of equal length (kept alive by the plan).

        Registering the points this plan already holds is skipped unless force: same
        storage (the plan keeps the previous tensors alive, so an equal address is the
        same memory), same length and the same version counter (no torch in-place write
        since).  Writes that bypass torch's version counter (a kernel writing through a
        raw pointer) are not seen: register such points with force=True."""
        n = x.numel()
        assert y.numel() == n and z.numel() == n
        for t in (x, y, z):
            assert t.dtype == torch.float32 and t.is_cuda and t.is_contiguous()
        try:
            key = tuple((t.data_ptr(), t.numel(), t._version) for t in (x, y, z))
        except RuntimeError:  # inference tensors carry no version counter
            key = None
        if not force and key is not None and key == self._pts_key:
            return
        self._pts = (x, y, z)
        self._pts_key = None
        check(_lib.lib().cmtip_nufft_setpts(self._h, n, ptr(x), ptr(y), ptr(z), stream()), "cmtip_nufft_setpts")
        self._pts_key = key

    def setpts_strengths(self, x, y, z, c):
        """Type-1 plans: register the points with their real strengths c (float32, one per
        point), carried in the bin-sorted records; execute with CMTIP_PTS_BOUND (c None)."""
        n = x.numel()
        assert y.numel() == n and z.numel() == n and c.numel() == n
        for t in (x, y, z, c):
            assert t.dtype == torch.float32 and t.is_cuda and t.is_contiguous()
        self._pts = (x, y, z, c)
        self._pts_key = None  # the records hold strengths: the next setpts bins again
        check(_lib.lib().cmtip_nufft_setpts_strengths(self._h, n, ptr(x), ptr(y), ptr(z), ptr(c), stream()),
              "cmtip_nufft_setpts_strengths")

    def setpts_rotated(self, quats, s, recip_extent, c=None, force=False):
        """Register the rotated pixels of ``gen_nonuniform_positions(quats, s,
        recip_extent=...)`` (quats: (n, 4) float32, s: (3, P) float64, device tensors;
        point o * P + p) without materialising their coordinate arrays; c (type 1,
        optional): real strengths bound as by ``setpts_strengths``.  Without strengths,
        registering the rotations this plan already holds is skipped unless force (same
        rule as ``setpts``)."""
        n, P = quats.shape[0], s.shape[1]
        assert quats.dtype == torch.float32 and quats.is_cuda and quats.is_contiguous() and quats.shape[1] == 4
        assert s.dtype == torch.float64 and s.is_cuda and s.is_contiguous() and s.shape[0] == 3
        key = None
        if c is not None:
            assert c.dtype == torch.float32 and c.is_cuda and c.is_contiguous() and c.numel() == n * P
        else:
            try:
                key = ("rotated", quats.data_ptr(), quats._version, s.data_ptr(), s._version, n, P,
                       float(recip_extent))
            except RuntimeError:  # inference tensors carry no version counter
                key = None
        if not force and key is not None and key == self._pts_key:
            return
        self._pts = (quats, s, c)
        self._pts_key = None
        check(_lib.lib().cmtip_nufft_setpts_rotated(self._h, ptr(quats), n, ptr(s), P, float(recip_extent), ptr(c),
                                                    stream()), "cmtip_nufft_setpts_rotated")
        self._pts_key = key

    def points_max_abs(self):
        """Largest |coordinate| of the registered points (NaN as inf), recorded by setpts."""
        v = ctypes.c_float()
        check(_lib.lib().cmtip_nufft_points_max_abs(self._h, ctypes.byref(v), stream()), "cmtip_nufft_points_max_abs")
        return v.value

    def execute(self, c, f, flags, support=None, scale=1.0):
        check(_lib.lib().cmtip_nufft_execute(self._h, ptr(c), ptr(f), int(flags), ptr(support), float(scale),
                                             stream()), "cmtip_nufft_execute")

    def set_bins(self, bin_size=0, chunk=0, wg_per_cu=0):
        """Bin edge, points per chunk and workgroups per CU (0 = default), used from the
        next setpts on."""
        self._pts_key = None  # the next setpts bins again with these settings
        check(_lib.lib().cmtip_nufft_set_bins(self._h, int(bin_size), int(chunk), int(wg_per_cu)),
              "cmtip_nufft_set_bins")

    def get_bins(self):
        b, c, w = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        check(_lib.lib().cmtip_nufft_get_bins(self._h, ctypes.byref(b), ctypes.byref(c), ctypes.byref(w)),
              "cmtip_nufft_get_bins")
        return b.value, c.value, w.value

    def autotune(self, c, f, flags, support=None, scale=1.0, candidates=None, reps=3, setpts_weight=0.0):
        """Time execute(c, f, flags, support, scale) over candidate (bin_size, chunk,
        wg_per_cu) settings with the registered points and keep the fastest (cost =
        setpts_weight * setpts + execute).  Returns (best, table) with table rows
        (bin_size, chunk, wg_per_cu, setpts_ms, execute_ms); -1 ms marks candidates
        that do not fit.  The outputs of execute are overwritten."""
        assert self._pts is not None, "autotune needs setpts first"
        if candidates is None:
            # staged search: bin edge at the default chunk and width, then chunk and width
            bs, t1 = self.autotune(c, f, flags, support, scale, spread_candidates(self.w, self.type), reps,
                                   setpts_weight)
            best, t2 = self.autotune(c, f, flags, support, scale,
                                     [(bs[0], ch, wpc) for ch in (1024, 2048, 4096) for wpc in (0, 1, 2, 4)],
                                     reps, setpts_weight)
            return best, t1 + t2
        cands = list(candidates)
        n = len(cands)
        arr = (ctypes.c_int * (3 * n))(*[int(v) for cand in cands for v in cand])
        times = (ctypes.c_float * (2 * n))()
        best = ctypes.c_int(-1)
        check(_lib.lib().cmtip_nufft_autotune(self._h, ptr(c), ptr(f), int(flags), ptr(support), float(scale), arr,
                                              n, int(reps), float(setpts_weight), times, ctypes.byref(best),
                                              stream()), "cmtip_nufft_autotune")
        table = [tuple(cands[i]) + (times[2 * i], times[2 * i + 1]) for i in range(n)]
        return tuple(cands[best.value]), table

    def slice_prepare(self, ac, support=None, owner=None):
        """Loads ac into the plan's slicing grid; ``owner`` (an integer token) is recorded as
        the holder of that grid, so users sharing the plan can tell whose volume it is."""
        self.slice_owner = None
        check(_lib.lib().cmtip_slice_prepare(self._h, ptr(ac), ptr(support), stream()), "cmtip_slice_prepare")
        self.slice_owner = owner

    def set_slice_route(self, route):
        """0 automatic, 1 fused tiles, 2 binned points, 3 binned tiles."""
        check(_lib.lib().cmtip_slice_set_route(self._h, int(route)), "cmtip_slice_set_route")

    def slice_route(self):
        """(requested route, route used by the last slice_real)."""
        r, u = ctypes.c_int(), ctypes.c_int()
        check(_lib.lib().cmtip_slice_get_route(self._h, ctypes.byref(r), ctypes.byref(u)), "cmtip_slice_get_route")
        return r.value, u.value

    def slice_real(self, quats, s, P, recip_extent, out, out_bf16=None, moments=None, order=None):
        self._pts_key = None  # the slicing routes reuse the plan's bin buffers
        n_orient = quats.shape[0]
        check(_lib.lib().cmtip_slice_real(self._h, ptr(quats), n_orient, ptr(s), ptr(order), int(P),
                                          float(recip_extent), ptr(out), ptr(out_bf16), ptr(moments), stream()),
              "cmtip_slice_real")

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value and _lib is not None and _lib._LIB is not None:
            _lib._LIB.cmtip_nufft_destroy(h)
            self._h = None


def spread_candidates(w, nufft_type):
    """Candidate (bin_size, 0, 0) settings: subgrid edges 10..24 cells (bin edge =
    subgrid edge - w + 1) whose LDS subgrid fits a workgroup in every execute form of
    the plan type."""
    out = []
    per_cell = 16 if nufft_type == 1 else 4  # 64-bit pairs (complex spread) / float (interpolation)
    reserve = 16384 if nufft_type == 1 else 256  # the kernels' static LDS
    for S in range(10, 25, 2):
        bs = S - w + 1
        if bs >= 2 and S * S * (S + 1) * per_cell + reserve <= 160 * 1024:
            out.append((bs, 0, 0))
    return out


_CACHE = {}


def get_plan(nufft_type, n_modes, iflag, eps=DEFAULT_EPS, tag=None):
    """Cached plan per (type, modes, sign, eps, device, tag); users that keep state in
    a plan between calls (a prepared slicing grid) take their own tag."""
    key = (int(nufft_type), int(n_modes), int(iflag), float(eps), torch.cuda.current_device(), tag)
    p = _CACHE.get(key)
    if p is None:
        p = NufftPlan(nufft_type, n_modes, iflag, eps)
        _CACHE[key] = p
    return p


def clear_plans():
    _CACHE.clear()
    _TUNED.clear()


# tuned (bin_size, chunk, wg_per_cu) per (plan type, modes, eps, execute flags, points)
_TUNED = {}


def tuned_setpts(plan, x, y, z, c, f, flags, support=None, scale=1.0):
    """setpts with autotuned bin settings: the first call for a (plan, flags, number of
    points) runs the autotuner (overwriting the execute outputs), later calls reuse its
    choice.  The plan keeps the tuned settings."""
    key = (plan.type, plan.n_modes, plan.eps, int(flags), x.numel(), torch.cuda.current_device())
    best = _TUNED.get(key)
    if best is not None:
        plan.set_bins(*best)
        plan.setpts(x, y, z)
        return best
    plan.setpts(x, y, z)
    best, _ = plan.autotune(c, f, flags, support, scale)
    _TUNED[key] = best
    return best
