"""Orientation matching on the GPU (cmtip/alignment/align.py).

Signatures and defaults follow the reference; extra knobs are keyword-only.  The
slices of every reference orientation are computed by the fused slicing kernel
(coordinates generated in-kernel from the quaternions), the weighted distances and
their argmin by the MFMA matching kernels; nothing is copied to the host until the
matched quaternions are returned.

Reference orientations: the reference draws them with skopi's get_uniform_quat
(align.py:108), which is not available; pass ``ref_orientations`` to use a given
set, otherwise a seeded uniform set (cmtip_amd.synth.random_quaternions) is drawn.
"""
import collections
import ctypes
import hashlib
import itertools

import numpy as np
import torch

from .. import _arrays, _lib, synth
from .._lib import check, ptr, stream
from ..nufft.plan import DEFAULT_EPS, get_plan

EMPTY_KEY = -1  # all ones bits: the largest unsigned 64-bit key


def generate_weights(pixel_position_reciprocal, order=0):
    """w = (1e-10 |s|)^-order / sum (align.py:11-24); shape of |s| over axis 0."""
    as_torch = _arrays.is_torch(pixel_position_reciprocal)
    s = _arrays.to_dev(pixel_position_reciprocal, torch.float32)
    s_magnitudes = torch.linalg.vector_norm(s, dim=0) * 1e-10
    weights = 1.0 / (s_magnitudes ** order)
    weights = weights / torch.sum(weights)
    return _arrays.out(weights, as_torch)


def _norms(X, w2):
    out = torch.empty(X.shape[0], dtype=torch.float32, device=X.device)
    check(_lib.lib().cmtip_weighted_norms(ptr(X), X.shape[0], X.shape[1], ptr(w2), ptr(out), stream()),
          "cmtip_weighted_norms")
    return out


def calc_eudist(model_slices, slices):
    """Euclidean distance matrix (n_model, n_slices), float32 (align.py:27-36; sklearn
    euclidean_distances), with the cross term on MFMA."""
    as_torch = _arrays.is_torch(model_slices, slices)
    A = _arrays.to_dev(model_slices, torch.float32)
    B = _arrays.to_dev(slices, torch.float32)
    A = A.reshape(A.shape[0], -1)
    B = B.reshape(B.shape[0], -1)
    assert A.shape[1] == B.shape[1]
    ones = torch.ones(A.shape[1], dtype=torch.float32, device=A.device)
    an, bn = _norms(A, ones), _norms(B, ones)
    D = torch.empty((A.shape[0], B.shape[0]), dtype=torch.float32, device=A.device)
    check(_lib.lib().cmtip_eudist_f32(ptr(A), A.shape[0], ptr(B), B.shape[0], A.shape[1], ptr(an), ptr(bn), ptr(D),
                                      stream()), "cmtip_eudist_f32")
    return _arrays.out(D, as_torch)


def _s64(pixel_position_reciprocal):
    s = _arrays.to_dev(pixel_position_reciprocal, torch.float64)
    return s.reshape(3, -1).contiguous()


CROSS_BLOCK_ELEMS = 1 << 28  # FP32 cross-term block of the library GEMM path (1 GiB)

_GEOMETRY = collections.OrderedDict()


def _geometry(pixel_position_reciprocal, tile_pixels):
    """Device copy of the pixel vectors (3, P) float64 and their pixel order.  Host
    arrays are cached by content (a few detectors), so repeated alignment calls reuse
    the same device tensors and the library's route probe for them."""
    key = None
    if not isinstance(pixel_position_reciprocal, torch.Tensor):
        a = np.ascontiguousarray(pixel_position_reciprocal)
        key = (a.shape, a.dtype.str, hashlib.blake2b(a.tobytes(), digest_size=16).digest(), bool(tile_pixels),
               torch.cuda.current_device())
        hit = _GEOMETRY.get(key)
        if hit is not None:
            _GEOMETRY.move_to_end(key)
            return hit
    s = _s64(pixel_position_reciprocal)
    geo = (s, pixel_order(s) if tile_pixels else None)
    if key is not None:
        _GEOMETRY[key] = geo
        while len(_GEOMETRY) > 4:
            _GEOMETRY.popitem(last=False)
    return geo


_WEIGHTS2 = collections.OrderedDict()


def _weights2(pixel_position_reciprocal, order):
    """Squared pixel weights w^2 (P,) float32 on the device (generate_weights squared).
    Host arrays are cached by content and order like _geometry, so repeated calls make
    no host-to-device copy and no reduction for them."""
    key = None
    if not isinstance(pixel_position_reciprocal, torch.Tensor):
        a = np.ascontiguousarray(pixel_position_reciprocal)
        key = (a.shape, a.dtype.str, hashlib.blake2b(a.tobytes(), digest_size=16).digest(), float(order),
               torch.cuda.current_device())
        hit = _WEIGHTS2.get(key)
        if hit is not None:
            _WEIGHTS2.move_to_end(key)
            return hit
    w = generate_weights(_arrays.to_dev(pixel_position_reciprocal, torch.float32), order=order).reshape(-1)
    w2 = (w * w).contiguous()
    if key is not None:
        _WEIGHTS2[key] = w2
        while len(_WEIGHTS2) > 4:
            _WEIGHTS2.popitem(last=False)
    return w2


def _ext32(reciprocal_extent):
    if isinstance(reciprocal_extent, torch.Tensor):
        reciprocal_extent = reciprocal_extent.item()
    return float(np.float32(reciprocal_extent))


def _spread_bits(v):
    v = (v | (v << 8)) & 0x00FF00FF
    v = (v | (v << 4)) & 0x0F0F0F0F
    v = (v | (v << 2)) & 0x33333333
    return (v | (v << 1)) & 0x55555555


def pixel_order(s):
    """Pixels in Morton order of their panel coordinates: runs of 64 consecutive
    entries are compact 8 x 8 blocks of a square panel, so the slicing kernel can stage
    their fine-grid footprint in LDS.  s: (3, P) device tensor of reciprocal vectors.

    Pixels on the Ewald sphere |s + z/lambda| = 1/lambda have 1/lambda =
    -|s|^2 / (2 s_z), and (s_x, s_y) / (s_z + 1/lambda) is the gnomonic (flat-panel)
    position, linear in the pixel index of a planar detector; other geometries fall
    back to (s_x, s_y)."""
    P = s.shape[1]
    side = max(1, int(np.ceil(np.sqrt(P))))
    sd = s.double()
    s2 = (sd * sd).sum(0)
    sel = sd[2].abs() > 1e-12 * s2.max().sqrt().clamp(min=1e-300)
    xy = sd[:2]
    if bool(sel.any()):
        k = float((-s2[sel] / (2 * sd[2][sel])).median())
        den = sd[2] + k
        if k > 0 and bool((den > 0).all()):
            xy = sd[:2] / den
    lo = xy.min(dim=1, keepdim=True).values
    span = (xy.max(dim=1, keepdim=True).values - lo).clamp(min=1e-300)
    q = torch.floor((xy - lo) / span * (side - 1) + 0.5).long().clamp(0, 0xFFFF)
    code = _spread_bits(q[0]) | (_spread_bits(q[1]) << 1)
    return torch.argsort(code, stable=True).to(torch.int32).contiguous()


_SLICER_IDS = itertools.count(1)


class Slicer:
    """Slices of one autocorrelation at many orientations (compute_slices, align.py:39-65).

    The fine grid is prepared once per autocorrelation; ``slice`` then writes
    Re(T2(ac)) at the rotated pixels of a batch of quaternions.  Pixels are processed
    in ``pixel_order`` (the output layout does not change).  The library picks the
    fused-tile or the binned route per detector geometry (``plan.set_slice_route``
    fixes one); both give the same values.  Slicers of the same M share one plan: a
    Slicer whose volume is no longer the one in the plan's grid loads it again before
    slicing.
    """

    def __init__(self, ac, pixel_position_reciprocal, reciprocal_extent, eps=DEFAULT_EPS, tile_pixels=True):
        self.ac = _arrays.to_dev(ac, torch.float32)
        M = self.ac.shape[0]
        self.plan = get_plan(2, M, -1, eps, tag="slice")
        self.s, self.order = _geometry(pixel_position_reciprocal, tile_pixels)
        self.P = self.s.shape[1]
        self.ext = _ext32(reciprocal_extent)
        self._token = next(_SLICER_IDS)
        self.plan.slice_prepare(self.ac, owner=self._token)

    def slice(self, quats, out=None, out_bf16=None, moments=None):
        if self.plan.slice_owner != self._token:
            self.plan.slice_prepare(self.ac, owner=self._token)
        q = _arrays.to_dev(quats, torch.float32).reshape(-1, 4)
        if out is None:
            out = torch.empty((q.shape[0], self.P), dtype=torch.float32, device=q.device)
        assert out.is_contiguous() and out.numel() == q.shape[0] * self.P
        self.plan.slice_real(q, self.s, self.P, self.ext, out, out_bf16, moments, order=self.order)
        return out


def compute_slices(orientations, pixel_position_reciprocal, reciprocal_extent, ac, use_gpu=False, *,
                   eps=DEFAULT_EPS):
    """Flattened model slices (n_orient * P,) float32 (align.py:39-65)."""
    as_torch = _arrays.is_torch(orientations, ac)
    sl = Slicer(ac, pixel_position_reciprocal, reciprocal_extent, eps)
    out = sl.slice(orientations)
    return _arrays.out(out.reshape(-1), as_torch)


def reference_orientations(n_ref_orientations, seed=0):
    return synth.random_quaternions(n_ref_orientations, seed)


class _Matcher:
    """State shared by one alignment call: images, weights and per-image best keys.

    Model slices and images are centred on the mean image c before the cross term
    (d2 is unchanged; the FP32 / BF16 expanded form then does not cancel).  With
    per-image masks (1 = valid pixel) the cross term runs over 2P columns,
    [y^2, y] . [-m w^2 / 2, m w^2 (b - c)] with y = gamma a - c, or 3P in BF16 with y^2
    split into two bfloat16 parts."""

    def __init__(self, pixel_position_reciprocal, slices_, order, precision, topk, masks=None, cross="lib"):
        assert cross in ("lib", "fused")
        self.cross = cross
        self.P = pixel_position_reciprocal.shape[-1]
        B = _arrays.to_dev(slices_, torch.float32).reshape(-1, self.P)
        self.B = B
        self.n_img = B.shape[0]
        L = _lib.lib()
        self.w2 = _weights2(pixel_position_reciprocal, order)
        self.precision = precision
        self.k = int(topk)
        dev = B.device
        self.c = torch.empty(self.P, dtype=torch.float32, device=dev)
        self.bn = torch.empty(self.n_img, dtype=torch.float32, device=dev)
        self.mask = None
        if masks is not None:
            self.mask = _arrays.to_dev(masks, torch.uint8).reshape(self.n_img, self.P).contiguous()
            colsum = torch.empty(3 * self.P, dtype=torch.float64, device=dev)
            mom = torch.empty(3, dtype=torch.float64, device=dev)
            check(L.cmtip_image_stats_masked(ptr(B), ptr(self.mask), self.n_img, self.P, ptr(colsum), ptr(self.c),
                                             ptr(mom), stream()), "cmtip_image_stats_masked")
            self.img_mom, self.img_count = mom, 0.0  # the valid-pixel count is mom[2]
            # float32 cross term over [y^2, y]; bfloat16 over [hi(y^2), lo(y^2), y]
            self.K = (2 if precision == "fp32" else 3) * self.P
            self.Bw = torch.empty((self.n_img, self.K), dtype=torch.float32, device=dev) \
                if precision == "fp32" else None
            self.Bw16 = torch.empty((self.n_img, self.K), dtype=torch.int16, device=dev) \
                if precision == "bf16" else None
            check(L.cmtip_match_prepare_masked(ptr(B), ptr(self.mask), self.n_img, self.P, ptr(self.w2), ptr(self.c),
                                               ptr(self.Bw), ptr(self.Bw16), ptr(self.bn), stream()),
                  "cmtip_match_prepare_masked")
            return
        colsum = torch.empty(2 * self.P, dtype=torch.float64, device=dev)
        mom = torch.empty(2, dtype=torch.float64, device=dev)
        check(L.cmtip_image_stats(ptr(B), self.n_img, self.P, ptr(colsum), ptr(self.c), ptr(mom), stream()),
              "cmtip_image_stats")
        self.K = self.P
        self.Bw = torch.empty_like(B) if precision == "fp32" else None
        self.Bw16 = torch.empty(B.shape, dtype=torch.int16, device=dev) if precision == "bf16" else None
        check(L.cmtip_match_prepare_images_norms(ptr(B), self.n_img, self.P, ptr(self.w2), ptr(self.c),
                                                 ptr(self.Bw), ptr(self.Bw16), ptr(self.bn), stream()),
              "cmtip_match_prepare_images_norms")
        self.img_mom, self.img_count = mom, float(B.numel())

    def run(self, slicer, ref, nbatches, verbose=False):
        """Per-image best key over all references (int64 holding the unsigned key)."""
        n_ref = ref.shape[0]
        quot = n_ref // nbatches
        dev = self.B.device
        L = _lib.lib()
        masked = self.mask is not None
        best = torch.full((self.n_img,), EMPTY_KEY, dtype=torch.int64, device=dev)
        an = torch.zeros(quot, dtype=torch.float32, device=dev)
        A_all = torch.empty((n_ref if self.precision == "bf16" else quot, self.P), dtype=torch.float32, device=dev)
        if self.precision == "bf16":
            cand = torch.full((self.n_img, self.k), EMPTY_KEY, dtype=torch.int64, device=dev)
            gamma_ref = torch.empty(n_ref, dtype=torch.float64, device=dev)
            A16 = torch.empty((quot, self.K), dtype=torch.int16, device=dev)
            blk = max(1, min(quot, CROSS_BLOCK_ELEMS // max(self.n_img, 1)))
            Dt = None
        elif masked:
            A2 = torch.empty((quot, self.K), dtype=torch.float32, device=dev)
        gamma_one = torch.empty(1, dtype=torch.float64, device=dev)
        for nb in range(nbatches):
            start, end = nb * quot, (nb + 1) * quot
            mom = torch.zeros(2, dtype=torch.float64, device=dev)
            Ab = A_all[start:end] if self.precision == "bf16" else A_all
            slicer.slice(ref[start:end], out=Ab, moments=mom)
            # data/model std ratio (align.py:122-124), rounded to float32 as numpy keeps it
            gamma = gamma_ref[start:end] if self.precision == "bf16" else gamma_one
            check(L.cmtip_match_gamma(ptr(self.img_mom), self.img_count, ptr(mom), float(Ab.numel()),
                                      gamma.numel(), ptr(gamma), stream()), "cmtip_match_gamma")
            if verbose:
                print(f"Data/Model std ratio: {float(gamma[0]):.7g}.", flush=True)
            if self.precision == "bf16":
                if masked:
                    check(L.cmtip_center_rows_sq(ptr(Ab), quot, self.P, ptr(gamma), ptr(self.c), None, ptr(A16),
                                                 stream()), "cmtip_center_rows_sq")
                else:
                    check(L.cmtip_center_rows(ptr(Ab), quot, self.P, ptr(gamma), ptr(self.c), ptr(self.w2), None,
                                              ptr(A16), ptr(an), stream()), "cmtip_center_rows")
                if self.cross == "fused":
                    check(L.cmtip_match_topk_bf16(ptr(A16), quot, start, ptr(self.Bw16), self.n_img, self.K,
                                                  ptr(an), ptr(self.bn), None, self.k, ptr(cand), stream()),
                          "cmtip_match_topk_bf16")
                else:
                    # library GEMM into an FP32 block of at most CROSS_BLOCK_ELEMS, then top-k
                    if Dt is None:
                        Dt = torch.empty(blk * self.n_img, dtype=torch.float32, device=dev)
                    for r0 in range(0, quot, blk):
                        r1 = min(quot, r0 + blk)
                        check(L.cmtip_match_cross_bf16(ptr(A16[r0:r1]), r1 - r0, ptr(self.Bw16), self.n_img,
                                                       self.K, ptr(Dt), stream()), "cmtip_match_cross_bf16")
                        check(L.cmtip_match_topk_cross(ptr(Dt), r1 - r0, start + r0, self.n_img, ptr(an[r0:r1]),
                                                       ptr(self.bn), None, self.k, ptr(cand), stream()),
                              "cmtip_match_topk_cross")
            else:
                if masked:
                    check(L.cmtip_center_rows_sq(ptr(Ab), quot, self.P, ptr(gamma), ptr(self.c), ptr(A2), None,
                                                 stream()), "cmtip_center_rows_sq")
                    Aop = A2
                else:
                    check(L.cmtip_center_rows(ptr(Ab), quot, self.P, ptr(gamma), ptr(self.c), ptr(self.w2), ptr(Ab),
                                              None, ptr(an), stream()), "cmtip_center_rows")
                    Aop = Ab
                check(L.cmtip_match_argmin_f32(ptr(Aop), quot, start, ptr(self.Bw), self.n_img, self.K, ptr(an),
                                               ptr(self.bn), None, ptr(best), stream()),
                      "cmtip_match_argmin_f32")
        if self.precision == "bf16":
            check(L.cmtip_match_rescore_masked(ptr(A_all), n_ref, ptr(self.B), ptr(self.mask), self.n_img, self.P,
                                               ptr(self.w2), ptr(gamma_ref), self.k, ptr(cand), ptr(best), stream()),
                  "cmtip_match_rescore_masked")
        return best


def _round_up_refs(n_ref_orientations, nbatches):
    quot, remainder = divmod(n_ref_orientations, nbatches)
    if remainder != 0:
        n_ref_orientations = (quot + 1) * nbatches
    return n_ref_orientations


def _ref_set(n_ref_orientations, ref_orientations, true_orientations, seed):
    if ref_orientations is None:
        ref = reference_orientations(n_ref_orientations, seed)
    else:
        ref = np.asarray(ref_orientations.cpu() if isinstance(ref_orientations, torch.Tensor) else ref_orientations,
                         dtype=np.float32)
        assert ref.shape == (n_ref_orientations, 4), "ref_orientations must hold n_ref_orientations rows"
    if true_orientations is not None:
        t = np.asarray(true_orientations.cpu() if isinstance(true_orientations, torch.Tensor)
                       else true_orientations, dtype=np.float32)
        # align.py:109-111: the true quaternions replace the tail, then a global shuffle
        ref = np.vstack((ref[:-len(t)], t))
        np.random.shuffle(ref)
    return np.ascontiguousarray(ref, dtype=np.float32)


def key_index(best):
    """Reference index of a packed best key (low 32 bits)."""
    return best & 0xFFFFFFFF


def key_d2(best):
    """Squared distance of a packed best key (high 32 bits, float32 bits)."""
    return (best >> 32).to(torch.int32).view(torch.float32)


def match_orientations(generation, pixel_position_reciprocal, reciprocal_extent, slices_, ac, n_ref_orientations,
                       nbatches=1, order=0, use_gpu=False, true_orientations=None, *, ref_orientations=None,
                       precision="fp32", topk=None, eps=DEFAULT_EPS, seed=0, verbose=False, return_index=False,
                       masks=None, cross="lib"):
    """Quaternions (n_images, 4) float32 of the best-matching reference slices (align.py:68-132).

    precision "fp32": cross term in FP32 on MFMA, argmin fused.  "bf16": cross term in
    BF16 on MFMA with FP32 accumulation, the ``topk`` best per image rescored in FP32;
    ``cross`` "lib" takes the cross term from the hipBLASLt GEMM then reduces it to the
    top-k, "fused" runs this library's MFMA kernel with the top-k in its epilogue (both
    keep up to 16 candidates; the library GEMM is the faster of the two at these shapes).
    masks (n_images, P), 1 = valid pixel: distances over each image's valid pixels only
    (the data std is then taken over valid pixels); not in the reference.  topk defaults
    to 8, and to 16 for masked images in BF16: masking the bright
    centre shrinks the gaps between the best references about 100 x, so more of them
    fall inside the BF16 error (at 10000 x 10000 masked images the 8 best missed the
    true best for 1 image in 10^4, the 16 best for none).
    """
    assert precision in ("fp32", "bf16")
    if topk is None:
        topk = 16 if (masks is not None and precision == "bf16") else 8
    n_ref_orientations = _round_up_refs(n_ref_orientations, nbatches)
    ref = _ref_set(n_ref_orientations, ref_orientations, true_orientations, seed)
    as_torch = _arrays.is_torch(slices_, ac)
    slicer = Slicer(ac, pixel_position_reciprocal, reciprocal_extent, eps)
    m = _Matcher(pixel_position_reciprocal, slices_, order, precision, topk, masks, cross)
    ref_dev = torch.from_numpy(ref).to(m.B.device)
    best = m.run(slicer, ref_dev, nbatches, verbose)
    index = key_index(best)
    q = ref_dev[index]
    if return_index:
        return _arrays.out(q, as_torch), _arrays.out(index, as_torch)
    return _arrays.out(q, as_torch)


def match_orientations_mult(generation, pixel_position_reciprocal, reciprocal_extent, slices_, ac_mult,
                            n_ref_orientations, nbatches=1, order=0, use_gpu=False, true_orientations=None, *,
                            ref_orientations=None, precision="fp32", topk=8, eps=DEFAULT_EPS, seed=0,
                            verbose=False):
    """(quaternions, conformations) over several autocorrelation volumes (align.py:135-213)."""
    assert precision in ("fp32", "bf16")
    n_ref_orientations = _round_up_refs(n_ref_orientations, nbatches)
    ref = _ref_set(n_ref_orientations, ref_orientations, true_orientations, seed)
    as_torch = _arrays.is_torch(slices_, ac_mult)
    m = _Matcher(pixel_position_reciprocal, slices_, order, precision, topk)
    ref_dev = torch.from_numpy(ref).to(m.B.device)
    bests = []
    for ac in ac_mult:
        slicer = Slicer(ac, pixel_position_reciprocal, reciprocal_extent, eps)
        bests.append(m.run(slicer, ref_dev, nbatches, verbose))
    bests = torch.stack(bests)  # (n_conf, n_img)
    d2 = key_d2(bests)
    # argmin over conformations of the per-volume minimum distance (first on ties, :206)
    conformations = torch.argmin(d2, dim=0)
    idx = key_index(bests.gather(0, conformations[None, :])[0])
    q = ref_dev[idx]
    return _arrays.out(q, as_torch), _arrays.out(conformations, as_torch)


def match_orientations_batch(generation, pixel_position_reciprocal, reciprocal_extent, slices_, ac,
                             n_ref_orientations, nbatches=1, batch_size=400, order=0, true_orientations=None, *,
                             ref_orientations=None, precision="fp32", topk=8, eps=DEFAULT_EPS, seed=0):
    """Aligns images in batches of batch_size (align.py:216-266); returns (n_images, 4) float64
    as the reference's np.zeros accumulator does (:243)."""
    n = slices_.shape[0]
    out = np.zeros((n, 4))
    quot, remainder = divmod(n, batch_size)
    if quot == 0:
        quot = 1
    for i in range(quot):
        start, end = i * batch_size, (i + 1) * batch_size
        if i == quot - 1:
            end = n
        q = match_orientations(generation, pixel_position_reciprocal, reciprocal_extent, slices_[start:end], ac,
                               n_ref_orientations, nbatches=nbatches, order=order,
                               true_orientations=true_orientations, ref_orientations=ref_orientations,
                               precision=precision, topk=topk, eps=eps, seed=seed)
        out[start:end] = q.cpu().numpy() if isinstance(q, torch.Tensor) else q
    return out
