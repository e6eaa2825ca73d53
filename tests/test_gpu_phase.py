"""Phasing on the GPU (cmtip/phasing/phase.py:11-219) against the FP64 oracle
(oracle/phase.py): the amplitude set-up, single ER / HIO steps and short loops, the
projection, shrink-wrap, recentring, the phased autocorrelation, and whole phase()
runs from the same random start."""
import numpy as np
import pytest
import torch

from cmtip_amd import phasing, synth

from oracle import phase as OP
from tests.gpu_util import dev, rel_l2, say

pytestmark = pytest.mark.gpu


def _ac(M, seed=0):
    return synth.gaussian_ac(M, 60, seed=seed, radius_frac=0.15)


def _state(M, seed=0):
    ac = _ac(M, seed)
    ac_filt_, amp, mask = OP.amplitudes_of(ac.astype(np.float64))
    sup = OP.create_support_(ac_filt_, M, int((M - 1) / 4))
    rho = sup * np.random.default_rng(seed + 10).random(sup.shape)
    return ac, ac_filt_, amp, mask, sup, rho


@pytest.mark.parametrize("M", [41, 81, 30])
def test_setup_amplitudes_and_support(M):
    ac, ac_filt_, amp, mask, sup, _ = _state(M)
    say(f"phase setup M={M}")
    ph = phasing.Phaser(M)
    filt, s = ph.setup(ac, int((M - 1) / 4))
    a = ph.amplitudes().cpu().numpy()
    # the FP64 filter is scipy's arithmetic in scipy's order, rounded once to float32
    assert np.array_equal(filt.cpu().numpy(), ac_filt_.astype(np.float32))
    assert np.array_equal(s.cpu().numpy() > 0, sup)
    assert a.ravel()[0] < 0  # the DC entry is masked (kept by the projection)
    # intensities |F| = amplitudes^2 to FP32 FFT accuracy; the square root magnifies
    # the absolute FP32 error of the smallest intensities (measured ~6e-4 rel L2 on the
    # amplitudes themselves)
    assert rel_l2(a.ravel()[1:] ** 2, amp.ravel()[1:] ** 2) < 1e-6
    assert rel_l2(a.ravel()[1:], amp.ravel()[1:]) < 2e-3


@pytest.mark.parametrize("M", [41, 30])
def test_single_steps_and_projection(M):
    ac, ac_filt_, amp, mask, sup, rho = _state(M, seed=1)
    d = dev()
    say(f"ER / HIO steps M={M}")
    mod_w, star_w = OP.step_phase(rho, amp, mask, sup)
    mod_g, star_g = phasing.step_phase(torch.from_numpy(rho).to(d), torch.from_numpy(amp).to(d),
                                       torch.from_numpy(mask).to(d), torch.from_numpy(sup).to(d))
    assert rel_l2(mod_g, mod_w) < 1e-5
    # support_star agrees wherever rho_mod is not within FP32 noise of zero
    clear = np.abs(mod_w) > 1e-5 * np.abs(mod_w).max()
    assert np.array_equal(star_g.cpu().numpy()[clear], star_w[clear])
    for name in ("ER", "HIO"):
        want = rho.copy()
        got = rho.copy()
        if name == "ER":
            OP.ER(want, amp, mask, sup, np.inf)
            phasing.ER(got, amp, mask, sup, np.inf)
        else:
            OP.HIO(0.3, want, amp, mask, sup, np.inf)
            phasing.HIO(0.3, got, amp, mask, sup, np.inf)
        assert got.dtype == np.float64
        assert rel_l2(got, want) < 1e-5, name
    # short loops track the FP64 iterates, the differences growing with the steps as the
    # iteration amplifies them (HIO more than ER)
    want, got = rho.copy(), rho.copy()
    for _ in range(5):
        OP.ER(want, amp, mask, sup, np.inf)
    phasing.ER_loop(5, got, amp, mask, sup, np.inf)
    e_er = rel_l2(got, want)
    for _ in range(5):
        OP.HIO(0.3, want, amp, mask, sup, np.inf)
    phasing.HIO_loop(5, 0.3, got, amp, mask, sup, np.inf)
    e_hio = rel_l2(got, want)
    say(f"  5 ER rel L2 {e_er:.2e}; then 5 HIO {e_hio:.2e}")
    assert e_er < 1e-3
    assert e_hio < 2e-2


def test_rho_max_thresholds():
    M = 30
    ac, ac_filt_, amp, mask, sup, rho = _state(M, seed=2)
    say("ER / HIO with a finite rho_max")
    mod, _ = OP.step_phase(rho, amp, mask, sup)
    rmax = float(np.quantile(mod, 0.9))
    for name in ("ER", "HIO"):
        want, got = rho.copy(), rho.copy()
        if name == "ER":
            OP.ER(want, amp, mask, sup, rmax)
            phasing.ER(got, amp, mask, sup, rmax)
        else:
            OP.HIO(0.3, want, amp, mask, sup, rmax)
            phasing.HIO(0.3, got, amp, mask, sup, rmax)
        clear = np.abs(mod - rmax) > 1e-5 * np.abs(mod).max()
        assert rel_l2(got[clear], want[clear]) < 1e-5, name


@pytest.mark.parametrize("M", [41, 30])
def test_shrink_wrap(M):
    ac, ac_filt_, amp, mask, sup, rho = _state(M, seed=3)
    for _ in range(10):
        OP.ER(rho, amp, mask, sup, np.inf)
    rho32 = rho.astype(np.float32).astype(np.float64)  # both sides from the same values
    say(f"shrink_wrap M={M}")
    want = np.zeros_like(sup)
    g = OP.shrink_wrap(5e-2, 1, rho32.copy(), want)
    got = np.zeros_like(sup)
    phasing.shrink_wrap(5e-2, 1, rho32.copy(), got)
    thr = np.abs(rho32).max() * 5e-2
    clear = np.abs(g - thr) > 1e-12 * thr
    assert clear.mean() > 0.99
    assert np.array_equal(got[clear], want[clear])


def test_recenter_and_phased_ac():
    M = 41
    ac, ac_filt_, amp, mask, sup, rho = _state(M, seed=4)
    rho = np.roll(rho, (3, -5, 7), axis=(0, 1, 2))
    sup = np.roll(sup, (3, -5, 7), axis=(0, 1, 2))
    rho32 = rho.astype(np.float32).astype(np.float64)
    say("recenter + phased autocorrelation")
    want_r, want_s = rho32.copy(), sup.copy()
    vect = OP.recenter(want_r, want_s, M)
    ph = phasing.Phaser(M)
    r = torch.from_numpy(rho32).float().to(dev())
    s = torch.from_numpy(sup).float().to(dev())
    got_v = ph.recenter(r, s)
    np.testing.assert_array_equal(got_v, vect)
    np.testing.assert_array_equal(r.cpu().numpy().astype(np.float64), want_r)
    np.testing.assert_array_equal(s.cpu().numpy() > 0, want_s)
    acp = ph.ac(r).cpu().numpy()
    assert rel_l2(acp, OP.phased_ac(want_r)) < 1e-5


@pytest.mark.parametrize("M,sched", [(41, (1, 3, 2)), (41, (3, 20, 10))])
def test_phase_from_the_same_start(M, sched):
    # whole phase() runs from the same random start: the short schedule tracks the FP64
    # iterates; the longer one (chaotic HIO) is compared through its result, the
    # autocorrelation of the phased density, which does not depend on the density's
    # translation or handedness
    n_it, nER, nHIO = sched
    ac, ac_filt_, amp, mask, sup, rho = _state(M, seed=5)
    say(f"phase M={M} n_iterations={n_it} nER={nER} nHIO={nHIO}")
    acp_w, sup_w, rho_w = OP.phase(0, ac.astype(np.float64), rho_=rho.copy(), n_iterations=n_it, nER=nER, nHIO=nHIO)
    acp_g, sup_g, rho_g = phasing.phase(0, ac, rho_=rho.copy(), n_iterations=n_it, nER=nER, nHIO=nHIO)
    assert acp_g.dtype == np.float64 and sup_g.dtype == np.bool_ and rho_g.shape == (M, M, M)
    cc = np.corrcoef(acp_g.ravel(), acp_w.ravel())[0, 1]
    target = np.fft.fftshift(ac_filt_)
    cc_t = np.corrcoef(acp_g.ravel(), target.ravel())[0, 1]
    cc_w = np.corrcoef(acp_w.ravel(), target.ravel())[0, 1]
    say(f"  cc(gpu, oracle) {cc:.6f}; vs the filtered ac: gpu {cc_t:.6f}, oracle {cc_w:.6f}; "
        f"rho rel L2 {rel_l2(rho_g, rho_w):.2e}")
    if n_it == 1:
        assert rel_l2(rho_g, rho_w) < 1e-3
        assert rel_l2(acp_g, acp_w) < 1e-3
    else:
        assert cc > 0.9995
        assert cc_t > cc_w - 1e-3
