"""Writes the golden fixtures of this directory from the FP64 oracle.

The reference holds no golden vectors (its tests simulate their inputs on the fly)
and cannot run here, so these vectors are the oracle's outputs on small seeded
inputs: exact sums where the size allows, the FP64 ES route at eps 1e-12 elsewhere.
The oracle itself is pinned against closed forms and brute force in
tests/test_oracle_*.py; the fixtures freeze its outputs so that a change in the
oracle (tests/test_golden.py) or in the HIP path (tests/test_gpu_golden.py) shows up
against fixed numbers.

    python tests/golden/make_fixtures.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from cmtip_amd import synth  # noqa: E402
from oracle import ac as OAC  # noqa: E402
from oracle import align as OA  # noqa: E402
from oracle import geom  # noqa: E402
from oracle import nufft as O  # noqa: E402


def nufft_m21():
    """Exact type 2 of a real grid and exact type 1 of complex strengths, M = 21,
    3000 points given in float32 (the precision the GPU path reads)."""
    rng = np.random.default_rng(101)
    M, N = 21, 3000
    x, y, z = (rng.uniform(-np.pi, np.pi, N).astype(np.float32) for _ in range(3))
    f = rng.standard_normal((M,) * 3).astype(np.float32)
    c = (rng.standard_normal(N) + 1j * rng.standard_normal(N)).astype(np.complex64)
    X, Y, Z = (v.astype(np.float64) for v in (x, y, z))
    t2 = O.nudft2(f.astype(np.complex128), X, Y, Z)
    t1 = O.nudft1(c.astype(np.complex128), X, Y, Z, M)
    return dict(M=M, x=x, y=y, z=z, f=f, c=c, t2=t2, t1=t1)


def slices_and_match_m21():
    """Exact slices of an M = 21 autocorrelation at 16 orientations on a 24 x 24 panel,
    and the FP64 distance table of 48 references against 24 noisy images."""
    M = 21
    s, ext = synth.detector(24)
    ac = synth.gaussian_ac(M, 40, seed=0)
    q_sl = synth.random_quaternions(16, seed=3)
    slices = OA.compute_slices(q_sl, s, ext, ac, exact=True).reshape(16, -1)
    q_ref = synth.random_quaternions(48, seed=4)
    q_img = np.vstack([q_ref[:12], synth.random_quaternions(12, seed=5)]).astype(np.float32)
    clean = OA.compute_slices(q_img, s, ext, ac, exact=True).reshape(24, -1)
    rng = np.random.default_rng(6)
    imgs = (np.maximum(clean, 0) * 50 + rng.standard_normal(clean.shape) * 0.05).astype(np.float32)
    D, ratios = OA.distances(q_ref, s, ext, imgs, ac)
    return dict(M=M, s=s, ext=np.float32(ext), ac=ac, q_sl=q_sl, slices=slices, q_ref=q_ref, imgs=imgs,
                D=D, ratio=np.array(ratios), argmin=np.argmin(D, axis=0), ok=OA.tie_band_ok(D))


def ac_m9():
    """Toeplitz normal operator, corner projection and a 5-iteration CG solve at M = 9."""
    M = 9
    s, ext = synth.detector(12)
    q = synth.random_quaternions(40, seed=7)
    acm = synth.gaussian_ac(M, 12, seed=8)
    x, y, z = geom.rotated_coords(q, s, ext)
    F = OAC.kernel_spectrum(x, y, z, M, exact=True)
    rng = np.random.default_rng(9)
    S = (rng.random((M,) * 3) > 0.2).astype(np.float32)
    u = rng.standard_normal(M ** 3).astype(np.float32)
    conv = OAC.core_problem_convolution(u.astype(np.float64), M, F, 2 * M, S)
    anti = OAC.gen_F_antisupport(M)
    reg = OAC.fourier_reg(u.astype(np.float64), S, anti, M)
    data = OA.compute_slices(q, s, ext, acm, exact=True).astype(np.float32).reshape(40, 1, -1)
    ac5, iters, _, _, _ = OAC.solve(q, s, ext, data, M, maxiter=5, rlambda=1e-4, exact=True, rtol=0.0)
    return dict(M=M, s=s, ext=np.float32(ext), q=q, F=F, S=S, u=u, conv=conv, reg=reg, data=data, ac5=ac5,
                iters=np.int64(iters))


def geom_fixture():
    """Rotation matrices of 8 quaternions and the alignment errors between two sets."""
    q1 = synth.random_quaternions(8, seed=10)
    q2 = synth.random_quaternions(8, seed=11)
    R = np.stack([geom.quaternion2rot3d(q) for q in q1])
    err = geom.alignment_error(q1, q2)
    return dict(q1=q1, q2=q2, R=R, err=np.asarray(err))


FIXTURES = {"nufft_m21": nufft_m21, "slices_match_m21": slices_and_match_m21, "ac_m9": ac_m9,
            "geom": geom_fixture}

if __name__ == "__main__":
    for name, fn in FIXTURES.items():
        path = os.path.join(HERE, name + ".npz")
        np.savez_compressed(path, **fn())
        print(path, os.path.getsize(path), "bytes")
