"""Autocorrelation solve (autocorrelation.py:12-248) against the FP64 oracle."""
import numpy as np
import pytest
import torch
from scipy.ndimage import gaussian_filter

from cmtip_amd import autocorrelation as A
from cmtip_amd import synth
from oracle import ac as OAC
from oracle import geom
from oracle import nufft as O
from tests.gpu_util import dev, rel_l2, say

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("M", [5, 41, 81])
def test_antisupport_exact(M):
    say(f"gen_F_antisupport M={M}")
    assert np.array_equal(A.gen_F_antisupport(M), OAC.gen_F_antisupport(M))


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_support_exact(dtype):
    M = 21
    ac = synth.gaussian_ac(M, 20, seed=2).astype(np.float64)
    # hard zeros outside a ball so the filtered volume crosses the 1e-12 threshold
    c = np.arange(M) - M // 2
    r = np.sqrt(c[:, None, None] ** 2 + c[None, :, None] ** 2 + c[None, None, :] ** 2)
    ac[r > 6] = 0.0
    ac = ac.astype(dtype)
    say(f"support dtype={dtype.__name__}")
    got = A.autocorrelation._support_of(ac, M).cpu().numpy()
    want = (gaussian_filter(ac, 0.5) > 1e-12).astype(np.float32)
    assert np.array_equal(got, want)


def _geometry(M=11, n_side=16, n_img=60, seed=0):
    s, ext = synth.detector(n_side)
    q = synth.random_quaternions(n_img, seed=seed)
    acm = synth.gaussian_ac(M, 20, seed=seed + 1)
    x, y, z = geom.rotated_coords(q, s, ext)
    return s, ext, q, acm, x, y, z


@pytest.mark.parametrize("M", [7, 11])
def test_core_problem_convolution(M):
    s, ext, q, acm, x, y, z = _geometry(M)
    F = OAC.kernel_spectrum(x, y, z, M, eps=1e-12)
    rng = np.random.default_rng(1)
    S = (rng.random((M,) * 3) > 0.2).astype(np.float32)
    u = rng.standard_normal(M ** 3).astype(np.float32)
    say(f"core_problem_convolution M={M}")
    got = A.core_problem_convolution(u, M, F, 2 * M, S)
    want = OAC.core_problem_convolution(u.astype(np.float64), M, F, 2 * M, S)
    assert rel_l2(got, want) < 1e-5


@pytest.mark.parametrize("M", [7, 8, 11])
def test_fourier_reg(M):
    rng = np.random.default_rng(M)
    S = (rng.random((M,) * 3) > 0.2).astype(np.float32)
    u = rng.standard_normal(M ** 3).astype(np.float32)
    anti = OAC.gen_F_antisupport(M)
    say(f"fourier_reg M={M}")
    got = A.fourier_reg(u, S, anti, M, True)
    want = OAC.fourier_reg(u.astype(np.float64), S, anti, M)
    assert rel_l2(got, want) < 1e-5


def test_setup_linops_vs_oracle():
    M = 11
    s, ext, q, acm, x, y, z = _geometry(M)
    data = O.nufft2(acm.astype(np.float64), x, y, z, eps=1e-12).real
    N = data.size
    sup = np.ones((M,) * 3, np.float32)
    x0 = np.zeros(M ** 3)
    rl = M ** 3 / N / 1000
    say("setup_linops M=11")
    W, d = A.setup_linops(torch.as_tensor(x), torch.as_tensor(y), torch.as_tensor(z), data.astype(np.float32), sup,
                          np.ones(N, np.float32), x0, M, M ** 3, N, ext, 1, rl, 1e3, scaled=True)
    d_want = OAC.rhs(x, y, z, data, M, sup, x0, 1, rl)
    assert rel_l2(d, d_want) < 1e-5
    F = OAC.kernel_spectrum(x, y, z, M)
    anti = OAC.gen_F_antisupport(M)
    u = np.random.default_rng(0).standard_normal(M ** 3)
    want = OAC.core_problem_convolution(u, M, F, 2 * M, sup) + rl * u + 1e3 * OAC.fourier_reg(u, sup, anti, M)
    assert rel_l2(W.matvec(u.astype(np.float32)), want) < 1e-5


def test_setup_linops_range_check():
    # the range assertion of adjoint.py:29,66, read once after both spreads are queued,
    # still refuses points outside [-3 pi, 3 pi) (one coordinate moved to 3.5 pi)
    M = 11
    s, ext, q, acm, x, y, z = _geometry(M)
    N = x.size
    x = x.copy()
    x[N // 2] = 3.5 * np.pi
    say("setup_linops range check M=11")
    with pytest.raises(AssertionError):
        A.setup_linops(torch.as_tensor(x), torch.as_tensor(y), torch.as_tensor(z), np.ones(N, np.float32),
                       np.ones((M,) * 3, np.float32), None, np.zeros(M ** 3), M, M ** 3, N, ext, 1, 1e-3, 1e3,
                       scaled=True)


@pytest.mark.parametrize("with_estimate", [False, True])
def test_solve_ac_vs_oracle(with_estimate):
    M = 11
    s, ext, q, acm, x, y, z = _geometry(M, n_img=80)
    data = O.nufft2(acm.astype(np.float64), x, y, z, eps=1e-12).real.reshape(len(q), 1, -1).astype(np.float32)
    est = acm.astype(np.float64) if with_estimate else None
    say(f"solve_ac M=11 estimate={with_estimate}")
    got, info = A.solve_ac(0, s, ext, data, M, orientations=q,
                           ac_estimate=None if est is None else est.copy(), return_info=True)
    want, its, sup, d, F = OAC.solve(q, s, ext, data, M, ac_estimate=est)
    assert got.dtype == np.float64 and got.shape == (M, M, M)
    # the GPU solution solves the oracle's system to the requested tolerance (FP32 slack)
    anti = OAC.gen_F_antisupport(M)
    rl = M ** 3 / data.size / 1000

    def W(u):
        return OAC.core_problem_convolution(u, M, F, 2 * M, sup) + rl * u + 1e3 * OAC.fourier_reg(u, sup, anti, M)
    res = np.linalg.norm(W(got.ravel()) - d) / np.linalg.norm(d)
    assert res < 5e-5
    assert abs(info["iterations"] - its) <= max(2, its // 5)
    assert rel_l2(got, want) < 1e-3


def test_cg_iterates_track_oracle():
    M = 11
    s, ext, q, acm, x, y, z = _geometry(M, n_img=80)
    data = O.nufft2(acm.astype(np.float64), x, y, z, eps=1e-12).real.reshape(len(q), 1, -1).astype(np.float32)
    for k in (1, 3, 6):
        say(f"solve_ac maxiter={k}")
        got, info = A.solve_ac(0, s, ext, data, M, orientations=q, maxiter=k, return_info=True)
        want, its, _, _, _ = OAC.solve(q, s, ext, data, M, maxiter=k)
        assert info["iterations"] == its == k
        assert rel_l2(got, want) < 1e-4


def _solver_system(M=11):
    s, ext, q, acm, x, y, z = _geometry(M, n_img=80)
    data = O.nufft2(acm.astype(np.float64), x, y, z, eps=1e-12).real
    sup = np.ones((M,) * 3, np.float32)
    rl = M ** 3 / data.size / 1000
    W, d = A.setup_linops(torch.as_tensor(x), torch.as_tensor(y), torch.as_tensor(z), data.astype(np.float32), sup,
                          np.ones(data.size, np.float32), np.zeros(M ** 3), M, M ** 3, data.size, ext, 1, rl, 1e3,
                          scaled=True)
    return W, d, rl


def _random_w(M, L, seed, positive=False):
    """A solver with a random real spectrum (not symmetric), a random antisupport mask
    and a random support, and the oracle's W = alpha conv + rho I + phi fourier_reg
    (autocorrelation.py:150-155) for the same terms."""
    rng = np.random.default_rng(seed)
    F = rng.random((L,) * 3) if positive else rng.standard_normal((L,) * 3)
    anti = rng.random((M,) * 3) > 0.5
    S = (rng.random((M,) * 3) > 0.3).astype(np.float32)
    solver = A.ACSolver(M, L)
    solver.set_support(torch.as_tensor(S))
    solver.set_spectrum(F.astype(np.float32))
    solver.set_antisupport(anti.astype(np.float32))
    F32 = F.astype(np.float32).astype(np.float64)

    def W(u, alpha, rho, phi):
        u = np.asarray(u, dtype=np.float64)
        return (alpha * OAC.core_problem_convolution(u, M, F32, L, S) + rho * u
                + phi * OAC.fourier_reg(u, S, anti, M))
    return solver, W


@pytest.mark.parametrize("M,L", [(11, 22), (9, 18), (10, 30), (11, 24), (12, 26)])
def test_w_operator_half_spectrum_vs_oracle(M, L):
    # the CG operator runs on half the spectrum with a symmetrised multiplier; L a
    # multiple of M (fold factor 2 and 3) folds fourier_reg into it, other L (24 and 26
    # here) run fourier_reg on the M grid after it
    solver, W = _random_w(M, L, seed=M * 100 + L)
    u = np.random.default_rng(L).standard_normal(M ** 3).astype(np.float32)
    say(f"W matvec M={M} L={L}")
    for alpha, rho, phi in ((0.7, 1e-2, 3.0), (1.0, 0.0, 0.0), (0.0, 0.5, 1e3)):
        got = solver.matvec(torch.as_tensor(u, device=dev()), alpha, rho, phi).cpu().numpy()
        assert rel_l2(got, W(u, alpha, rho, phi)) < 1e-5


@pytest.mark.parametrize("M,L", [(10, 20), (10, 30), (10, 24)])
def test_cg_solves_w_system_for_each_embedding(M, L):
    # the device CG (fused direction update, p.q partials in the last pass, folded or
    # separate fourier_reg) converges to the solution of the oracle's system
    solver, W = _random_w(M, L, seed=7 * L, positive=True)
    d = np.random.default_rng(5).standard_normal(M ** 3).astype(np.float32)
    x = torch.zeros(M ** 3, device=dev())
    say(f"cg M={M} L={L}")
    alpha, rho, phi = 1.0, 0.05, 2.0
    it, info = solver.cg(torch.as_tensor(d, device=dev()), x, alpha, rho, phi, 300, 1e-6)
    assert info == 0 and 0 < it < 300
    res = np.linalg.norm(W(x.cpu().numpy(), alpha, rho, phi) - d) / np.linalg.norm(d)
    assert res < 1e-5


def test_cg_graph_replay_matches_direct_launches(monkeypatch):
    W, d, rl = _solver_system()
    out = {}
    for mode in ("1", "0"):
        monkeypatch.setenv("CMTIP_CG_GRAPH", mode)
        x = torch.zeros_like(d)
        say(f"cg graph={mode}")
        out[mode] = (W.solver.cg(d, x, 1.0, rl, 1e3, 17, 1e-5), x.clone())
    assert out["1"][0] == out["0"][0]
    assert torch.equal(out["1"][1], out["0"][1])


def test_cg_edge_cases():
    W, d, rl = _solver_system()
    say("cg b = 0 and maxiter = 0")
    x = torch.full_like(d, 3.0)
    it, info = W.solver.cg(torch.zeros_like(d), x, 1.0, rl, 1e3, 10, 1e-5)
    assert (it, info) == (0, 0) and bool((x == 0).all())  # scipy returns b itself
    x = torch.full_like(d, 0.5)
    it, info = W.solver.cg(d, x, 1.0, rl, 1e3, 0, 1e-5)
    assert (it, info) == (0, 0)  # maxiter 0: no iteration, info = maxiter = 0
    assert bool((x == 0.5).all())
    x = torch.zeros_like(d)
    it, info = W.solver.cg(d, x, 1.0, rl, 1e3, 3, 1e-12)
    assert (it, info) == (3, 3)  # loop exhausted: info = maxiter


def test_solve_ac_mult_per_conformation():
    # solve_ac_mpi_mult (autocorrelation_mpi.py:193-243) on one GPU: conformation nc is
    # solved from slices_[conformations == nc] with its own orientations and estimate
    M, n_side = 11, 16
    s, ext = synth.detector(n_side)
    acs = [synth.gaussian_ac(M, 20, seed=41), synth.gaussian_ac(M, 20, seed=42)]
    q = synth.random_quaternions(90, seed=43)
    conf = np.random.default_rng(44).integers(0, 2, 90)
    x, y, z = geom.rotated_coords(q, s, ext)
    P = n_side * n_side
    data = np.stack([O.nufft2(acs[c].astype(np.float64), x[i * P:(i + 1) * P], y[i * P:(i + 1) * P],
                              z[i * P:(i + 1) * P], eps=1e-12).real for i, c in enumerate(conf)])
    data = data.reshape(90, 1, P).astype(np.float32)
    est = [acs[0].astype(np.float64), None]
    say("solve_ac_mult M=11, two conformations")
    got = A.solve_ac_mult(0, s, ext, data, M, 2, orientations=q, conformations=conf,
                          ac_estimate=[None if e is None else e.copy() for e in est])
    assert got.shape == (2, M, M, M) and got.dtype == np.float64
    for nc in range(2):
        sel = conf == nc
        alone = A.solve_ac(0, s, ext, data[sel], M, orientations=q[sel],
                           ac_estimate=None if est[nc] is None else est[nc].copy())
        # the same solve as solve_ac on the subset; the spreads' float-atomic flushes
        # make repeated solves differ in the last bits, which W's conditioning amplifies
        assert rel_l2(got[nc], alone) < 1e-3
        want, its, sup, d, F = OAC.solve(q[sel], s, ext, data[sel], M, ac_estimate=est[nc])
        anti = OAC.gen_F_antisupport(M)
        rl = M ** 3 / data[sel].size / 1000

        def W(u):
            return (OAC.core_problem_convolution(u, M, F, 2 * M, sup) + rl * u
                    + 1e3 * OAC.fourier_reg(u, sup, anti, M))
        assert np.linalg.norm(W(got[nc].ravel()) - d) / np.linalg.norm(d) < 5e-5
        assert rel_l2(got[nc], want) < 1e-3


def test_cg_graph_kept_across_calls(monkeypatch):
    # the captured block of CG iterations is kept across calls with the same rho and phi
    # and captured again when they change; every call gives the iterate and iteration
    # count of direct launches bit for bit (fixed-order reductions), including maxiter not
    # a multiple of the block, a converging rtol and repeated calls
    M = 11
    s, ext, q, acm, x, y, z = _geometry(M)
    data = O.nufft2(acm.astype(np.float64), x, y, z, eps=1e-12).real
    N = data.size
    rl = M ** 3 / N / 1000
    W, dvec = A.setup_linops(torch.as_tensor(x), torch.as_tensor(y), torch.as_tensor(z), data.astype(np.float32),
                             np.ones((M,) * 3, np.float32), None, np.zeros(M ** 3), M, M ** 3, N, ext, 1, rl, 1e3,
                             scaled=True)
    solver = W.solver
    cases = [(rl, 12, 0.0), (rl, 12, 0.0), (10 * rl, 7, 0.0), (rl, 40, 1e-5), (rl, 12, 0.0)]
    say("CG graph reuse")
    got = {}
    for graphs in ("1", "0"):
        monkeypatch.setenv("CMTIP_CG_GRAPH", graphs)
        for i, (r, maxiter, rtol) in enumerate(cases):
            xx = torch.zeros(M ** 3, device=dvec.device)
            it, info = solver.cg(dvec, xx, 1.0, r, 1e3, maxiter, rtol)
            got.setdefault(i, []).append((it, info, xx.cpu()))
    for i, ((it1, info1, x1), (it0, info0, x0)) in got.items():
        assert (it1, info1) == (it0, info0), cases[i]
        assert torch.equal(x1, x0), cases[i]
    assert torch.equal(got[0][0][2], got[1][0][2]) and torch.equal(got[0][0][2], got[4][0][2])
