"""Benchmark of the cmtip hot path on one MI355X.

Workload (default, "c2"): the full-iteration configuration at M = 81 — a Gaussian-atom
autocorrelation (1000 atoms) on an 81^3 grid, 5000 images of 128 x 128 pixels
(16384 q-vectors, one flat panel) and 10000 reference orientations.  A step is one
orientation-matching pass: prepare the fine grid from the autocorrelation, slice all
10000 references (type-2 NUFFT at 163.84 M rotated pixels), per-batch std scaling,
the weighted distance cross term on MFMA and the per-image argmin, in FP32 (the
workload's stated precision).  The headline value is images aligned per second
(5000 images per step) over --steps timed steps.

Also measured in the same run (printed in the JSON line): the BF16 form of the step
(BF16 cross term, FP32 top-8 rescoring) over the same number of steps and its ratio to
the FP32 form, one full reconstruction generation (matching + autocorrelation solve
with 20 CG iterations), type-2 and type-1 NUFFT throughput (Mpts/s), the CG iteration
time and the FFT passes' bandwidth, roofline entries for the dominant kernel and the
slicing kernel, and a CPU baseline (the oracle on a bounded sample, on this host's
cores).
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

CONFIGS = {
    # name: (M, atoms, detector side, n_img, n_ref)
    "c2": (81, 1000, 128, 5000, 10000),
    "c1": (41, 200, 64, 256, 1000),
}
HBM_PEAK_GBS = 8000.0
BF16_PEAK_TFS = 2500.0
FP32_MFMA_PEAK_TFS = 157.3


def recorded_traffic(kernel):
    """HBM bytes per launch from the rocprofv3 FETCH_SIZE / WRITE_SIZE runs recorded in
    profiles/traffic.json (each counter scaled by its copy-kernel calibration)."""
    try:
        rec = json.load(open(os.path.join(ROOT, "profiles", "traffic.json")))[kernel]
    except (OSError, KeyError, ValueError):
        return None
    return 1024.0 * (rec["fetch_size_kib"] * rec["fetch_scale"] + rec["write_size_kib"] * rec["write_scale"])


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=5)
    p.add_argument("--warmup", type=int, default=2)
    p.add_argument("--config", default="c2", choices=sorted(CONFIGS))
    p.add_argument("--precision", default="fp32", choices=["fp32", "bf16"],
                   help="precision of the headline step (the workload's own: FP32 distances and argmin); the "
                        "other form is timed beside it")
    p.add_argument("--nbatches", type=int, default=1)
    p.add_argument("--cross", default="lib", choices=["lib", "fused"],
                   help="BF16 cross term of --part match: the library GEMM + top-k pass, or the fused kernel")
    p.add_argument("--no-extras", action="store_true", help="time the step only")
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--part", default="all",
                   choices=["all", "slice", "match", "nufft", "traffic", "solve", "autograd", "c4", "c5"],
                   help="slice: one slicing pass of the workload (for counter runs); match: --steps matching "
                        "steps of the workload at --precision after one untimed step; nufft: setpts / execute "
                        "timing of both transforms; traffic: a 1 GiB copy (counter calibration) then one slicing "
                        "pass; autograd: forward + backward of the differentiable slice (2048 orientations); c4 / "
                        "c5: matching throughput at the larger workloads (see large_main)")
    return p.parse_args()


class Timer:
    """HIP-event timing of one kernel family on torch's current stream."""

    def __init__(self):
        self.pairs = []

    def wrap(self, fn):
        import torch

        def w(*a, **k):
            s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            s.record()
            r = fn(*a, **k)
            e.record()
            self.pairs.append((s, e))
            return r
        return w

    def reset(self):
        self.pairs = []

    def mean_ms(self):
        if not self.pairs:
            return float("nan")
        return sum(s.elapsed_time(e) for s, e in self.pairs) / len(self.pairs)


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus above 1 is not supported"}))
        sys.exit(2)
    import torch

    from cmtip_amd import _lib, alignment, nufft, synth
    from cmtip_amd._lib import check
    from cmtip_amd.alignment import align as AL
    from cmtip_amd.autocorrelation import autocorrelation as AC

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    torch.cuda.set_device(0)
    dev = torch.device("cuda", 0)
    M, atoms, side, n_img, n_ref = CONFIGS[args.config]
    if args.part in ("c4", "c5"):
        return large_main(args, dev)
    if args.part != "all":
        return part_main(args, M, atoms, side, n_img, n_ref, dev)
    P = side * side
    log(f"[bench] config {args.config}: M={M} atoms={atoms} P={P} n_img={n_img} n_ref={n_ref}")
    t0 = time.time()
    ac = synth.gaussian_ac(M, atoms, seed=0)
    s, ext = synth.detector(side)
    ref = synth.random_quaternions(n_ref, seed=0)
    q_img = synth.random_quaternions(n_img, seed=1)
    ac_d = torch.from_numpy(ac).to(dev)
    ref_d = torch.from_numpy(ref).to(dev)
    # synthetic images: slices at the image orientations, Poisson at 10 photons/pixel
    sl = alignment.Slicer(ac_d, s, ext)
    clean = sl.slice(q_img)
    torch.manual_seed(2)
    scale = 10.0 / float(clean.clamp(min=0).mean())
    imgs = torch.poisson(clean.clamp(min=0) * scale).float()
    del clean
    torch.cuda.synchronize()
    log(f"[bench] inputs ready in {time.time() - t0:.1f}s")

    slice_t, gemm_t = Timer(), Timer()
    plan_slice = sl.plan.slice_real
    sl.plan.slice_real = slice_t.wrap(plan_slice)
    L = _lib.lib()
    orig_topk, orig_f32 = L.cmtip_match_topk_bf16, L.cmtip_match_argmin_f32

    def step(precision):
        slicer = AL.Slicer(ac_d, s, ext)
        slicer.plan.slice_real = slice_t.wrap(plan_slice)
        m = AL._Matcher(s, imgs, 0, precision, 8)
        best = m.run(slicer, ref_d, args.nbatches)
        return ref_d[AL.key_index(best)]

    # wrap the GEMM entry points with event timing
    class _Wrapped:
        def __init__(self, fn):
            self.fn = fn

        def __call__(self, *a):
            return gemm_t.wrap(self.fn)(*a)

    L.cmtip_match_topk_bf16 = _Wrapped(orig_topk)
    L.cmtip_match_argmin_f32 = _Wrapped(orig_f32)
    # the library-GEMM form: cross term (timed with the GEMMs) and its top-k pass
    topk_t = Timer()
    L.cmtip_match_cross_bf16 = _Wrapped(L.cmtip_match_cross_bf16)
    L.cmtip_match_topk_cross = topk_t.wrap(L.cmtip_match_topk_cross)

    kernel_ms = {}

    def timed(precision, steps, warmup):
        for _ in range(warmup):
            step(precision)
        torch.cuda.synchronize()
        slice_t.reset()
        gemm_t.reset()
        topk_t.reset()
        # HIP events around the main slicing kernel, recorded by the library on its stream
        check(L.cmtip_slice_timer(sl.plan._h, 1, None, None), "cmtip_slice_timer")
        t = time.perf_counter()
        for i in range(steps):
            q = step(precision)
        torch.cuda.synchronize()
        el = time.perf_counter() - t
        tot, cnt = ctypes.c_float(), ctypes.c_int()
        check(L.cmtip_slice_timer(sl.plan._h, 0, ctypes.byref(tot), ctypes.byref(cnt)), "cmtip_slice_timer")
        kernel_ms["slice_kernel"] = tot.value / max(cnt.value, 1)
        kernel_ms["route"] = sl.plan.slice_route()[1]
        return el, q

    # the configuration's own precision (FP32 distances and argmin) is the headline; the
    # BF16 form runs the same step over the same number of steps beside it
    other = "bf16" if args.precision == "fp32" else "fp32"
    forms = {}
    for prec in (args.precision, other):
        log(f"[bench] {prec}: warmup {args.warmup}, timing {args.steps} steps")
        el, q = timed(prec, args.steps, args.warmup)
        forms[prec] = {"el": el, "q": q, "ms_per_step": el / args.steps * 1e3, "images_per_s": n_img * args.steps / el,
                       "slice_ms": slice_t.mean_ms(), "cross_ms": gemm_t.mean_ms(),
                       "topk_ms": topk_t.mean_ms() if topk_t.pairs else 0.0,
                       "slice_kernel_ms": kernel_ms["slice_kernel"], "route": kernel_ms["route"]}
        f = forms[prec]
        log(f"[bench] {prec}: {f['images_per_s']:.1f} images/s, {f['ms_per_step']:.2f} ms/step (slice "
            f"{f['slice_ms']:.2f} ms, cross term {f['cross_ms']:.2f} ms, top-k {f['topk_ms']:.2f} ms)")
        if args.no_extras:
            break
    head = forms[args.precision]
    value, ms_step = head["images_per_s"], head["ms_per_step"]

    N = n_ref * P
    n_fine = sl.plan.n_fine

    def slice_roofline(f):
        route, k_ms = f["route"], f["slice_kernel_ms"]
        kname = {1: "slice_real_kernel", 2: "binned points route", 3: "slice_tiles_kernel"}[route]
        n_tiles = n_ref * ((P + 63) // 64)
        bytes_launch = 4 * N + 24 * P + 16 * n_ref + 4 * n_fine ** 3 + (4 * n_tiles if route == 3 else 0)
        achieved = bytes_launch / (k_ms * 1e-3) / 1e9
        roof = {"kernel": kname, "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": recorded_traffic(kname)
                if args.config == "c2" and args.nbatches == 1 else None,
                "per_unit": "4 B per rotated pixel + 24 B per pixel + 16 B per orientation + fine grid"
                            + (" + 4 B per tile" if route == 3 else ""),
                "bytes_per_launch": bytes_launch, "avg_ms": k_ms, "slicing_call_ms": f["slice_ms"]}
        if route == 3 and sl.plan.w == 7:
            # the kernel's real limit: LDS reads, 49 rows of 7 taps per pixel read as
            # b128 + b64 + b32 (28 B), against this GPU's conflict-free b128 read rate
            lds_bytes = 49 * 28 * N
            lds_peak = lds_read_ceiling(dev)
            roof["lds"] = {"achieved": lds_bytes / (k_ms * 1e-3) / 1e12, "peak": lds_peak, "unit": "TB/s",
                           "frac": lds_bytes / (k_ms * 1e-3) / 1e12 / lds_peak, "bytes_per_launch": lds_bytes,
                           "per_unit": "49 x 28 B of LDS reads per rotated pixel",
                           "peak_source": "cmtip_lds_read_probe: conflict-free ds_read_b128 on this GPU"}
        return roof

    def cross_roofline(prec, f):
        flops = 2.0 * n_img * (n_ref // args.nbatches) * P
        peak = BF16_PEAK_TFS if prec == "bf16" else FP32_MFMA_PEAK_TFS
        achieved = flops / (f["cross_ms"] * 1e-3) / 1e12
        kname = "match_argmin_f32_kernel" if prec == "fp32" else "hipBLASLt bf16 GEMM (cmtip_match_cross_bf16)"
        return {"kernel": kname, "bound": "mfma", "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
                "frac": achieved / peak, "traffic": recorded_traffic(kname)
                if args.config == "c2" and args.nbatches == 1 else None,
                "per_unit": "2 flop per (image, reference, pixel)", "flop_per_launch": flops,
                "avg_ms": f["cross_ms"]}

    # roofline of the step's dominant kernel (the longer of the cross term and the slicing)
    roof = (cross_roofline(args.precision, head) if head["cross_ms"] >= head["slice_kernel_ms"]
            else slice_roofline(head))

    result = {
        "metric": "images aligned/s + type-1/2 NUFFT Mpts/s at M=81; BF16/autograd vs plain FP32",
        "value": value, "unit": "images/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms_step, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32" if args.precision == "fp32" else "bf16 cross term (fp32 accumulation) + fp32 top-8 rescoring",
        "data": "synthetic (seeded Gaussian-atom autocorrelation, flat-panel detector, Poisson images)",
        "config": {"workload": f"orientation matching at M={M}: {n_img} images x {n_ref} references x {P} pixels, "
                               f"{args.precision} distances + argmin",
                   "M": M, "n_images": n_img, "n_references": n_ref, "pixels": P, "nbatches": args.nbatches,
                   "precision": args.precision, "nufft_eps": nufft.DEFAULT_EPS},
        "breakdown_ms": {"slice": head["slice_ms"], "cross_term": head["cross_ms"], "topk": head["topk_ms"]},
        "roofline": roof,
        "roofline_slice": slice_roofline(head),
    }

    if not args.no_extras:
        f = forms[other]
        result[other] = {"images_per_s": f["images_per_s"], "ms_per_step": f["ms_per_step"], "steps": args.steps,
                         "warmup": args.warmup,
                         "breakdown_ms": {"slice": f["slice_ms"], "cross_term": f["cross_ms"], "topk": f["topk_ms"]},
                         "roofline_cross_term": cross_roofline(other, f)}
        if other == "fp32":
            result["roofline_cross_term_fp32"] = result[other]["roofline_cross_term"]
        bf, fp = forms["bf16"], forms["fp32"]
        # the BF16 form against the plain FP32 form of the same step (same slicing), and
        # of the cross term alone; the dense MFMA peaks bound the latter at 2500 / 157.3
        result["bf16_over_fp32_form"] = bf["images_per_s"] / fp["images_per_s"]
        result["bf16_over_fp32_cross_term"] = fp["cross_ms"] / max(bf["cross_ms"] + bf["topk_ms"], 1e-9)
        result["bf16_over_fp32_peak_bound"] = BF16_PEAK_TFS / FP32_MFMA_PEAK_TFS
        result["assignment_agreement_fp32_vs_bf16"] = float((bf["q"] == fp["q"]).all(dim=1).float().mean())
        sl.plan.slice_real = plan_slice
        # one generation of the reconstruction loop (reconstruct.py:72-87): FP32 matching of
        # the images against the references, then the autocorrelation solve from the assigned
        # orientations (kernel spread at 2M modes, rhs spread at M, support from the current
        # estimate, 20 CG iterations, Tikhonov weight 1e-4)
        log("[bench] full iteration: match + solve_ac")

        def generation():
            qa = step(args.precision)
            acs = AC.solve_ac(1, s_d, ext, imgs.reshape(n_img, 1, P), M, orientations=qa, ac_estimate=ac_d.clone(),
                              maxiter=20, rtol=0.0, rlambda=1e-4)
            return acs
        s_d = torch.from_numpy(s).to(dev)
        generation()
        torch.cuda.synchronize()
        reps = 3
        t = time.perf_counter()
        for _ in range(reps):
            generation()
        torch.cuda.synchronize()
        t_gen = (time.perf_counter() - t) / reps
        result["full_iteration"] = {"ms": t_gen * 1e3, "match_ms": ms_step, "solve_ms": t_gen * 1e3 - ms_step,
                                    "images_per_s": n_img / t_gen, "reps": reps, "cg_iterations": 20,
                                    "note": f"{args.precision} matching step + solve_ac (both type-1 spreads over "
                                            f"{n_img * P} points, support, 20 CG iterations); warm plans"}
        log(f"[bench] full iteration {result['full_iteration']}")
        # type-2 with coordinates in HBM (the image q-vectors, 81.92 M points) and type-1
        log("[bench] NUFFT throughput")
        x, y, z = AC.gen_nonuniform_positions(torch.from_numpy(q_img).to(dev), torch.from_numpy(s).to(dev),
                                              recip_extent=ext)
        x, y, z = (t.reshape(-1) for t in (x, y, z))
        Np = x.numel()

        def tm(fn, reps=3):
            fn()
            torch.cuda.synchronize()
            t = time.perf_counter()
            for _ in range(reps):
                fn()
            torch.cuda.synchronize()
            return (time.perf_counter() - t) / reps

        data = imgs.reshape(-1)
        # plan-level timing: setpts (binning) once per point set, execute per transform
        from cmtip_amd._lib import CMTIP_GRID_REAL, CMTIP_PTS_ONES, CMTIP_PTS_REAL
        res_nufft = {"points": Np, "note": "Mpts_s over execute; with_setpts adds one binning of the points"}
        for typ, modes, flags, label in ((2, M, CMTIP_GRID_REAL | CMTIP_PTS_REAL, "type2_real"),
                                         (2, M, CMTIP_GRID_REAL, "type2_complex"),
                                         (1, M, CMTIP_GRID_REAL | CMTIP_PTS_REAL, "type1_real_M81"),
                                         (1, 2 * M, CMTIP_GRID_REAL | CMTIP_PTS_ONES, "type1_ones_M162")):
            p = nufft.get_plan(typ, modes, -1 if typ == 2 else 1)
            ts = tm(lambda: p.setpts(x, y, z, force=True))
            if typ == 2:
                o = torch.empty(Np, device=dev, dtype=torch.float32 if flags & CMTIP_PTS_REAL else torch.complex64)
                te = tm(lambda: p.execute(o, ac_d, flags))
            else:
                o = torch.empty((modes,) * 3, device=dev)
                te = tm(lambda: p.execute(None if flags & CMTIP_PTS_ONES else data, o, flags, None,
                                          1.0 / modes ** 3))
            res_nufft[label] = {"setpts_ms": ts * 1e3, "execute_ms": te * 1e3, "Mpts_s": Np / te / 1e6,
                                "Mpts_s_with_setpts": Np / (te + ts) / 1e6}
        # the solve's rhs form: real strengths bound to the point records at setpts
        from cmtip_amd._lib import CMTIP_PTS_BOUND
        p = nufft.get_plan(1, M, 1)
        ts = tm(lambda: p.setpts_strengths(x, y, z, data))
        o = torch.empty((M,) * 3, device=dev)
        te = tm(lambda: p.execute(None, o, CMTIP_GRID_REAL | CMTIP_PTS_REAL | CMTIP_PTS_BOUND, None, 1.0 / M ** 3))
        res_nufft["type1_real_M81_bound"] = {"setpts_ms": ts * 1e3, "execute_ms": te * 1e3, "Mpts_s": Np / te / 1e6,
                                             "Mpts_s_with_setpts": Np / (te + ts) / 1e6}
        t2 = res_nufft["type2_real"]["execute_ms"] * 1e-3
        t1 = res_nufft["type1_real_M81"]["execute_ms"] * 1e-3
        # 20 B per point + 4 M^3 per transform (real form), execute only
        res_nufft["type2_plus_type1_real_hbm_frac"] = ((20 * Np + 4 * M ** 3) * 2 / (t2 + t1) / 1e9) / HBM_PEAK_GBS
        # the limits these kernels meet: LDS reads (interpolation: 49 rows x 28 B per point
        # at w = 7) and 64-bit LDS atomics (spreading: w^3 per point)
        w = nufft.get_plan(2, M, -1).w
        if w == 7:
            read_peak, atomic_peak = lds_read_ceiling(dev), lds_atomic_ceiling(dev)
            rd = 49 * 28 * Np / t2 / 1e12
            at = w ** 3 * Np / t1 / 1e9
            res_nufft["lds"] = {"type2_real_read_TBps": rd, "read_ceiling_TBps": read_peak,
                                "type2_read_frac": rd / read_peak, "type1_real_atomic_Gops": at,
                                "atomic_ceiling_Gops": atomic_peak, "type1_atomic_frac": at / atomic_peak,
                                "note": "ceilings from cmtip_lds_read_probe / cmtip_lds_atomic_probe on this GPU"}
        result["nufft"] = res_nufft
        log(f"[bench] nufft {result['nufft']}")
        # CG iteration time of the autocorrelation solve (20 iterations)
        solver = AC.get_solver(M)
        solver.set_support(torch.ones((M,) * 3, device=dev))
        solver.set_kernel(nufft.adjoint_ones(x, y, z, 2 * M, scale=1.0 / (2 * M) ** 3))
        d = nufft.adjoint_gpu(data, x, y, z, M).reshape(-1)
        rl = M ** 3 / data.numel() / 1000

        def cg20():
            xx = torch.zeros(M ** 3, device=dev)
            solver.cg(d, xx, 1.0, rl, 1e3, 20, 0.0)
        tcg = tm(cg20, reps=2)
        result["solve"] = {"cg_iterations": 20, "cg_ms_per_iteration": tcg / 20 * 1e3}
        # the FFT axis passes: one application of the CG operator W (five axis_pass_kernel
        # launches on half the spectrum, fourier_reg folded in), and of the full-spectrum
        # Toeplitz operator (core_problem_convolution), timed with HIP events
        u = torch.rand(M ** 3, device=dev)

        def ev_time(fn, reps=20):
            fn()
            ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
            ev[0].record()
            for _ in range(reps):
                fn()
            ev[1].record()
            torch.cuda.synchronize()
            return ev[0].elapsed_time(ev[1]) / reps * 1e-3
        t_w = ev_time(lambda: solver.matvec(u, 1.0, rl, 1e3))
        t_conv = ev_time(lambda: solver.conv(u))
        wb = w_pass_bytes(M, solver.L)
        cb = conv_pass_bytes(M, solver.L)
        result["fft"] = {"kernel": "axis_pass_kernel", "operator": "W (CG operator, half spectrum)", "M": M,
                         "L": solver.L, "passes": 5, "ms": t_w * 1e3, "bytes_per_application": wb,
                         "achieved_GBps": wb / t_w / 1e9, "frac_hbm": wb / t_w / 1e9 / HBM_PEAK_GBS,
                         "per_unit": "each pass reads and writes its lines once (8 B per complex element, 4 B per "
                                     "real one) over L/2+1 positions of the last axis, plus the multiplier "
                                     "(4 L^2 (L/2+1)), the support (4 M^3, twice) and u for the rho term",
                         "conv_full_spectrum": {"ms": t_conv * 1e3, "bytes_per_application": cb,
                                                "achieved_GBps": cb / t_conv / 1e9}}
        log(f"[bench] fft {result['fft']}")
        # differentiable slicing: forward + backward for 2048 orientations x P pixels
        from cmtip_amd import autograd as AG
        nb = min(2048, n_ref)
        qb = ref_d[:nb].clone().requires_grad_(True)
        acg = ac_d.clone().requires_grad_(True)
        target = imgs[:nb].reshape(nb, -1) if nb <= n_img else imgs[:1].expand(nb, -1)

        def fwdbwd(kind):
            qb.grad = None
            acg.grad = None
            model = AG.slice_quaternions(acg, qb, s, ext)
            loss = AG.l2_loss(model, target) if kind == "l2" else AG.poisson_nll(model, target)
            loss.backward()
        fwdbwd("l2")  # plans and the library GEMM choice of the first calls stay out of the timing
        t_l2 = tm(lambda: fwdbwd("l2"), reps=5)
        t_pn = tm(lambda: fwdbwd("poisson"), reps=5)
        result["autograd"] = {"orientations": nb, "points": nb * P, "fwd_bwd_ms_l2": t_l2 * 1e3,
                              "fwd_bwd_ms_poisson": t_pn * 1e3, "Mpts_s_l2": nb * P / t_l2 / 1e6}
        log(f"[bench] autograd {result['autograd']}")
        result["phasing"] = bench_phasing(ac, args.no_cpu_baseline)
        log(f"[bench] phasing {result['phasing']}")
        result["data_layer"] = bench_data_layer(dev, args.no_cpu_baseline)
        log(f"[bench] data layer {result['data_layer']}")

    if not args.no_cpu_baseline:
        result["cpu_baseline"] = cpu_baseline(ac, s, ext, ref, imgs.cpu().numpy(), n_img, n_ref, P)
    print(json.dumps(result), flush=True)


def part_main(args, M, atoms, side, n_img, n_ref, dev):
    """Single-purpose runs used under rocprofv3 (see --part)."""
    import torch

    from cmtip_amd import alignment, nufft, synth
    from cmtip_amd._lib import CMTIP_GRID_REAL, CMTIP_PTS_ONES, CMTIP_PTS_REAL
    from cmtip_amd.autocorrelation import autocorrelation as AC

    if args.part == "traffic":
        a = torch.ones(256 * 1024 * 1024, device=dev)
        b = torch.empty_like(a)
        b.copy_(a)  # 1 GiB read + 1 GiB written: the FETCH_SIZE / WRITE_SIZE scale
        torch.cuda.synchronize()
        del a, b
    ac = torch.from_numpy(synth.gaussian_ac(M, atoms, seed=0)).to(dev)
    s, ext = synth.detector(side)
    if args.part == "autograd":
        # forward (type 2) + backward (type 1 spread and three ik-weighted type 2 passes)
        # of the L2 loss over 2048 orientations x P pixels, one warm-up and args.steps timed
        from cmtip_amd import autograd as AG
        nb = 2048
        q = torch.from_numpy(synth.random_quaternions(nb, seed=0)).to(dev).requires_grad_(True)
        acg = ac.clone().requires_grad_(True)
        target = alignment.Slicer(ac, s, ext).slice(torch.from_numpy(synth.random_quaternions(nb, seed=1)).to(dev))

        def fwdbwd():
            q.grad = None
            acg.grad = None
            AG.l2_loss(AG.slice_quaternions(acg, q, s, ext), target).backward()
        fwdbwd()
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(args.steps):
            fwdbwd()
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t) / args.steps * 1e3
        res = {"part": "autograd", "M": M, "orientations": nb, "points": nb * side * side, "fwd_bwd_ms": ms}
        log(f"[bench] autograd {res}")
        print(json.dumps(res), flush=True)
        return
    if args.part == "match":
        from cmtip_amd.alignment import align as AL
        P = side * side
        s_d = s
        ref = torch.from_numpy(synth.random_quaternions(n_ref, seed=0)).to(dev)
        q_img = synth.random_quaternions(n_img, seed=1)
        clean = alignment.Slicer(ac, s, ext).slice(q_img).clamp_(min=0)
        torch.manual_seed(2)
        imgs = torch.poisson(clean * (10.0 / float(clean.mean())))
        del clean

        def one():
            m = AL._Matcher(s_d, imgs, 0, args.precision, 8, cross=args.cross)
            return m.run(AL.Slicer(ac, s_d, ext), ref, args.nbatches)
        one()
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(args.steps):
            one()
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t) / args.steps * 1e3
        res = {"part": "match", "precision": args.precision, "M": M, "images": n_img, "references": n_ref,
               "pixels": P, "ms_per_step": ms, "images_per_s": n_img / ms * 1e3}
        log(f"[bench] match {res}")
        print(json.dumps(res), flush=True)
        return
    if args.part in ("slice", "traffic"):
        ref = synth.random_quaternions(n_ref, seed=0)
        sl = alignment.Slicer(ac, s, ext)
        out = sl.slice(ref)
        for _ in range(max(0, args.steps - 1)):
            sl.slice(ref, out=out)
        torch.cuda.synchronize()
        log(f"[bench] slicing pass done ({n_ref} references)")
        return
    # nufft / solve: the image q-vectors
    q = synth.random_quaternions(n_img, seed=1)
    x, y, z = AC.gen_nonuniform_positions(torch.from_numpy(q).to(dev), torch.from_numpy(s).to(dev), recip_extent=ext)
    x, y, z = (t.reshape(-1).contiguous() for t in (x, y, z))
    N = x.numel()
    data = torch.rand(N, device=dev)

    def tm(fn, reps=5):
        fn()
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(reps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t) / reps

    if args.part == "solve":
        # the autocorrelation solve of the workload: kernel and rhs spreads, then CG
        imgs = alignment.Slicer(ac, s, ext).slice(q)
        t_calls = []
        for _ in range(2):  # the first call also creates the plans and the solver (cached afterwards)
            t = time.perf_counter()
            acs, info = AC.solve_ac(0, s, ext, imgs.reshape(n_img, 1, -1), M, orientations=q, maxiter=20,
                                    rtol=0.0, rlambda=1e-4, return_info=True)
            torch.cuda.synchronize()
            t_calls.append(time.perf_counter() - t)
        t_total = t_calls[-1]
        solver = AC.get_solver(M)
        d = torch.rand(M ** 3, device=dev)
        xx = torch.zeros(M ** 3, device=dev)
        t_cg = tm(lambda: solver.cg(d, xx.zero_(), 1.0, 1e-4, 1e3, 20, 0.0), reps=3)
        res = {"part": "solve", "M": M, "points": N, "solve_ms": t_total * 1e3, "first_call_ms": t_calls[0] * 1e3,
               "iterations": info["iterations"], "cg_ms_per_iteration": t_cg / 20 * 1e3}
        log(f"[bench] solve {res}")
        print(json.dumps(res), flush=True)
        return
    res = {"points": N}
    for typ, modes, flags, label in ((2, M, CMTIP_GRID_REAL | CMTIP_PTS_REAL, "type2_real"),
                                     (2, M, CMTIP_GRID_REAL, "type2_complex"),
                                     (1, M, CMTIP_GRID_REAL | CMTIP_PTS_REAL, "type1_real"),
                                     (1, 2 * M, CMTIP_GRID_REAL | CMTIP_PTS_ONES, "type1_ones_2M")):
        p = nufft.get_plan(typ, modes, -1 if typ == 2 else 1)
        ts = tm(lambda: p.setpts(x, y, z, force=True), 3)
        if typ == 2:
            out = torch.empty(N, device=dev, dtype=torch.float32 if flags & CMTIP_PTS_REAL else torch.complex64)
            te = tm(lambda: p.execute(out, ac, flags))
        else:
            out = torch.empty((modes,) * 3, device=dev)
            te = tm(lambda: p.execute(data, out, flags, None, 1.0 / modes ** 3))
        res[label] = {"setpts_ms": ts * 1e3, "execute_ms": te * 1e3, "Mpts_s": N / te / 1e6}
        log(f"[bench] {label}: setpts {ts * 1e3:.2f} ms, execute {te * 1e3:.2f} ms, {N / te / 1e6:.0f} Mpts/s")
    # real strengths bound to the point records at setpts (the solve's rhs form)
    from cmtip_amd._lib import CMTIP_PTS_BOUND
    p = nufft.get_plan(1, M, 1)
    ts = tm(lambda: p.setpts_strengths(x, y, z, data), 3)
    out = torch.empty((M,) * 3, device=dev)
    te = tm(lambda: p.execute(None, out, CMTIP_GRID_REAL | CMTIP_PTS_REAL | CMTIP_PTS_BOUND, None, 1.0 / M ** 3))
    res["type1_real_bound"] = {"setpts_ms": ts * 1e3, "execute_ms": te * 1e3, "Mpts_s": N / te / 1e6}
    log(f"[bench] type1_real_bound: setpts {ts * 1e3:.2f} ms, execute {te * 1e3:.2f} ms, {N / te / 1e6:.0f} Mpts/s")
    print(json.dumps({"part": "nufft", "M": M, **res}), flush=True)


def w_pass_bytes(M, L):
    """Bytes the five axis passes of the CG operator W move at sizes (M, L) on half the
    spectrum (H = L/2 + 1 positions of the last axis): axis 2 takes u and the support
    to (M, M, H) complex, axis 1 to (M, L, H), axis 0 forward x multiplier x inverse
    keeps M, axis 1 back to (M, M, H), axis 2 to the real M^3 result with the support
    and the rho u term."""
    H = L // 2 + 1
    m3, mmh, mlh, llh = M ** 3, M * M * H, M * L * H, L * L * H
    p1 = 4 * m3 + 4 * m3 + 8 * mmh
    p2 = 8 * mmh + 8 * mlh
    p3 = 8 * mlh + 4 * llh + 8 * mlh
    p4 = 8 * mlh + 8 * mmh
    p5 = 8 * mmh + 4 * m3 + 4 * m3 + 4 * m3
    return p1 + p2 + p3 + p4 + p5


def conv_pass_bytes(M, L):
    """Bytes the five axis passes of the Toeplitz operator move at sizes (M, L): axis 2
    pads M -> L (real u and the support in, complex out), axis 1 likewise, axis 0
    forward x spectrum x inverse keeps M, then the inverse passes crop back to M^3."""
    m3, m2l, ml2, l3 = M ** 3, M * M * L, M * L * L, L ** 3
    p1 = 4 * m3 + 4 * m3 + 8 * m2l
    p2 = 8 * m2l + 8 * ml2
    p3 = 8 * ml2 + 4 * l3 + 8 * ml2
    p4 = 8 * ml2 + 8 * m2l
    p5 = 8 * m2l + 4 * m3 + 4 * m3
    return p1 + p2 + p3 + p4 + p5


def _probe_best_ms(fn, dev, wg, iters):
    import torch

    from cmtip_amd._lib import check, ptr, stream

    out = torch.empty(wg * 512, device=dev)
    check(fn(wg, iters, ptr(out), stream()), "LDS probe")
    best = None
    for _ in range(3):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        check(fn(wg, iters, ptr(out), stream()), "LDS probe")
        b.record()
        torch.cuda.synchronize()
        ms = a.elapsed_time(b)
        best = ms if best is None else min(best, ms)
    return best


def lds_read_ceiling(dev):
    """TB/s of conflict-free ds_read_b128 over the whole GPU (cmtip_amd's probe kernel:
    2 workgroups of 512 threads per CU, 64 KB of LDS each), best of 3 timed launches."""
    import torch

    from cmtip_amd import _lib

    wg = 2 * torch.cuda.get_device_properties(dev).multi_processor_count
    iters = 20000
    ms = _probe_best_ms(_lib.lib().cmtip_lds_read_probe, dev, wg, iters)
    return wg * 512 * 16.0 * iters / (ms * 1e-3) / 1e12


def lds_atomic_ceiling(dev):
    """G ops/s of 64-bit LDS atomic adds over the whole GPU (the same launch shape, each
    lane on its own counter), best of 3 timed launches."""
    import torch

    from cmtip_amd import _lib

    wg = 2 * torch.cuda.get_device_properties(dev).multi_processor_count
    iters = 20000
    ms = _probe_best_ms(_lib.lib().cmtip_lds_atomic_probe, dev, wg, iters)
    return wg * 512.0 * iters / (ms * 1e-3) / 1e9


def large_main(args, dev):
    """Whole match_orientations calls at the larger workloads, BF16 cross term with top-8
    FP32 rescoring; one warm-up call, then timed calls; then the plain FP32 form and the
    BF16 form on the same (at most 20000) images, for their ratio.  c4: M = 81 (1000 atoms), 20000
    images x 50000 references x 128 x 128 pixels, Poisson mean 5, log1p-scaled.  c5:
    M = 151 (5000 atoms), 100000 images x 100000 references with a random circular
    beamstop mask per image, at 128 x 128 and at 96 x 96 pixels, Poisson mean 10."""
    import torch

    from cmtip_amd import alignment, synth

    if args.part == "c4":
        cases, steps = [(81, 1000, 128, 20000, 50000, False)], max(1, args.steps)
    else:
        cases, steps = [(151, 5000, 128, 100000, 100000, True), (151, 5000, 96, 100000, 100000, True)], 1
    out = {"part": args.part, "precision": "bf16 (fp32 beside it)", "steps": steps, "cases": []}
    for M, atoms, side, n_img, n_ref, masked in cases:
        P = side * side
        log(f"[bench] {args.part}: M={M} side={side} images={n_img} references={n_ref} masked={masked}")
        s, ext = synth.detector(side)
        ac = torch.from_numpy(synth.gaussian_ac(M, atoms, seed=0)).to(dev)
        ref = synth.random_quaternions(n_ref, seed=0)
        q_img = torch.from_numpy(synth.random_quaternions(n_img, seed=1)).to(dev)
        clean = alignment.Slicer(ac, s, ext).slice(q_img).clamp_(min=0)
        torch.manual_seed(2)
        masks = None
        if masked:
            imgs = torch.poisson(clean * (10.0 / float(clean.mean())))
            g = torch.Generator(device=dev).manual_seed(side)
            cy = torch.randint(side // 3, 2 * side // 3, (n_img, 1), device=dev, generator=g)
            cx = torch.randint(side // 3, 2 * side // 3, (n_img, 1), device=dev, generator=g)
            r = torch.rand((n_img, 1), device=dev, generator=g) * (side / 5 - 2) + 2
            pix = torch.arange(P, device=dev)
            masks = (((pix // side)[None, :] - cy) ** 2 + ((pix % side)[None, :] - cx) ** 2 > r * r).to(torch.uint8)
            del cy, cx, r, pix
        else:
            imgs = torch.log1p(torch.poisson(clean * (5.0 / float(clean.mean()))))
        del clean
        torch.cuda.synchronize()

        def call(precision, n):
            return alignment.match_orientations(0, s, ext, imgs[:n], ac, n_ref, ref_orientations=ref,
                                                precision=precision, masks=None if masks is None else masks[:n],
                                                return_index=True)[1]

        def timed_call(precision, n, reps):
            call(precision, n)
            torch.cuda.synchronize()
            log(f"[bench] {args.part}: {precision} warm-up call done ({n} images)")
            t = time.perf_counter()
            for _ in range(reps):
                idx = call(precision, n)
            torch.cuda.synchronize()
            return (time.perf_counter() - t) / reps, idx

        el, _ = timed_call("bf16", n_img, steps)
        K = 3 * P if masked else P
        case = {"M": M, "pixels": P, "images": n_img, "references": n_ref, "masked": masked,
                "ms_per_call": el * 1e3, "images_per_s": n_img / el,
                "cross_term_tflop": 2.0 * n_img * n_ref * K / 1e12,
                "whole_call_tflop_per_s": 2.0 * n_img * n_ref * K / el / 1e12}
        # the plain FP32 form (FP32 MFMA cross term over P columns, 2P with masks) on the
        # first n_cmp images, and the BF16 form on the same images: the ratio where the
        # cross term is nearly all of the work
        n_cmp = min(n_img, 20000)
        el32, idx32 = timed_call("fp32", n_cmp, 1)
        el16, idx16 = timed_call("bf16", n_cmp, 1) if n_cmp != n_img else (el, call("bf16", n_cmp))
        K32 = 2 * P if masked else P
        case["fp32"] = {"images": n_cmp, "ms_per_call": el32 * 1e3, "images_per_s": n_cmp / el32,
                        "cross_term_tflop": 2.0 * n_cmp * n_ref * K32 / 1e12,
                        "whole_call_tflop_per_s": 2.0 * n_cmp * n_ref * K32 / el32 / 1e12,
                        "bf16_ms_per_call_same_images": el16 * 1e3,
                        "bf16_over_fp32": el32 / el16,
                        "assignment_agreement": float((torch.as_tensor(idx32) == torch.as_tensor(idx16)).float()
                                                      .mean())}
        log(f"[bench] {args.part}: {case}")
        out["cases"].append(case)
        del imgs, masks
        torch.cuda.empty_cache()
    print(json.dumps(out), flush=True)


def cpu_baseline(ac, s, ext, ref, imgs, n_img, n_ref, P):
    """The oracle (FP64 ES NUFFT at eps 1e-6 + sklearn distances + argmin) on a bounded
    sample, extrapolated linearly to the full workload."""
    from sklearn.metrics.pairwise import euclidean_distances

    from oracle import geom
    from oracle import nufft as O

    cores = len(os.sched_getaffinity(0))
    n_s = 256
    x, y, z = geom.rotated_coords(ref[:n_s], s, ext)
    t = time.perf_counter()
    model = O.nufft2(ac.astype(np.float64), x, y, z, eps=1e-6).real.reshape(n_s, P).astype(np.float32)
    t_slice = time.perf_counter() - t
    m_s = min(2000, len(imgs))
    t = time.perf_counter()
    D = euclidean_distances(model, imgs[:m_s])
    np.argmin(D, axis=0)
    t_match = time.perf_counter() - t
    per_slice = t_slice / n_s
    per_pair = t_match / (n_s * m_s)
    total = per_slice * n_ref + per_pair * n_ref * n_img
    log(f"[bench] cpu: slice {per_slice * 1e3:.1f} ms/ref, match {per_pair * 1e9:.1f} ns/pair")
    parts = cpu_solve_parts(ac.shape[0], s, ext, n_s)
    return {"value": n_img / total, "unit": "images/s", "cores": cores, "kind": "port",
            "sample": f"oracle FP64 ES type-2 (eps 1e-6, C/OpenMP gridding + numpy FFT) for {n_s} references "
                      f"and sklearn euclidean_distances + argmin of those {n_s} x {m_s} images, extrapolated "
                      f"linearly to {n_ref} references x {n_img} images",
            "threads_openmp": O.num_threads(), "solve_parts": parts}


def bench_data_layer(dev, no_cpu):
    """The data layer on the workload's image stack: bin_data of 5000 images of 256 x 256
    pixels by 2 (prep_data.py:197-226) and clip_data of the binned stack at 12 A
    (:155-169), HIP events over 5 calls each; CPU: the numpy oracle on 200 images."""
    import torch

    from cmtip_amd import prep_data as PD
    from cmtip_amd import synth

    n, side = 5000, 256
    g = torch.Generator(device=dev).manual_seed(4)
    raw = torch.poisson(torch.rand((n, 1, side * side), device=dev, generator=g) * 10, generator=g)
    s, ext = synth.detector(side // 2)
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
    binned = PD.bin_data(raw, 2, det_shape=(1, side, side))
    clipped = PD.clip_data(binned, s, 12.0)
    torch.cuda.synchronize()
    reps = 5
    ev[0].record()
    for _ in range(reps):
        PD.bin_data(raw, 2, det_shape=(1, side, side))
    ev[1].record()
    for _ in range(reps):
        PD.clip_data(binned, s, 12.0)
    ev[2].record()
    torch.cuda.synchronize()
    t_bin = ev[0].elapsed_time(ev[1]) / reps * 1e-3
    t_clip = ev[1].elapsed_time(ev[2]) / reps * 1e-3
    b_bin = 4 * raw.numel() + 4 * binned.numel()
    b_clip = 4 * clipped.numel() * 2 + 3 * 4 * s.shape[-1]
    out = {"bin": {"images": n, "in_pixels": side * side, "factor": 2, "ms": t_bin * 1e3, "bytes": b_bin,
                   "GBps": b_bin / t_bin / 1e9, "frac_hbm": b_bin / t_bin / 1e9 / HBM_PEAK_GBS},
           "clip": {"images": n, "kept_pixels": clipped.shape[-1], "of": binned.shape[-1], "ms": t_clip * 1e3,
                    "bytes": b_clip, "GBps": b_clip / t_clip / 1e9, "frac_hbm": b_clip / t_clip / 1e9 / HBM_PEAK_GBS,
                    "note": "time includes the pixel selection (one synchronisation for the kept count)"}}
    if not no_cpu:
        from oracle import prep as OPR
        sub = raw[:200].cpu().numpy()
        t = time.perf_counter()
        bs = OPR.bin_data(sub, 2, det_shape=(1, side, side))
        tb = time.perf_counter() - t
        t = time.perf_counter()
        OPR.clip_data(bs, s, 12.0)
        tc = time.perf_counter() - t
        out["cpu_baseline"] = {"bin_ms_extrapolated": tb * n / 200 * 1e3, "clip_ms_extrapolated": tc * n / 200 * 1e3,
                               "cores": 1, "kind": "port", "sample": "numpy oracle on 200 of the 5000 images"}
    return out


def phase_step_bytes(M):
    """HBM bytes of one ER / HIO step's five axis passes on an M^3 volume: real density
    in (4 B) and complex out (8 B); two complex read + write passes; the amplitude pass
    (complex in and out plus 4 B of amplitudes); the last pass reads the complex volume,
    the density and the support and writes the density."""
    m3 = M ** 3
    return m3 * ((4 + 8) + 16 + (8 + 4 + 8) + 16 + (8 + 4 + 4 + 4))


def bench_phasing(ac, no_cpu):
    """phase() (phasing/phase.py:165-219) with the reference's schedule: 10 x (50 ER +
    25 HIO + 50 ER + shrink-wrap) + 50 ER, then recentring and the phased
    autocorrelation, on the benchmark's autocorrelation; one untimed call, then one
    timed.  CPU: the oracle (numpy complex128 FFTs) over 25 steps, extrapolated."""
    import torch

    from cmtip_amd import phasing

    M = ac.shape[0]
    np.random.seed(0)
    phasing.phase(0, ac, n_iterations=1, nER=2, nHIO=2)
    torch.cuda.synchronize()
    np.random.seed(0)
    t = time.perf_counter()
    acp, sup, rho = phasing.phase(0, ac)
    el = time.perf_counter() - t
    steps = 10 * (50 + 25 + 50) + 50
    # the ER / HIO steps alone, timed with HIP events
    ph = phasing.get_phaser(M)
    r = torch.from_numpy(rho).float().cuda()
    sp = torch.from_numpy(sup.astype(np.float32)).cuda()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    ev[0].record()
    ph.iterate(r, sp, 1, 100, 0.0, np.inf)
    ev[1].record()
    torch.cuda.synchronize()
    step_ms = ev[0].elapsed_time(ev[1]) / 100
    b = phase_step_bytes(M)
    out = {"M": M, "schedule": "10 x (50 ER + 25 HIO + 50 ER + shrink-wrap) + 50 ER", "steps": steps,
           "phase_ms": el * 1e3, "step_ms": step_ms, "step_bytes": b, "step_GBps": b / (step_ms * 1e-3) / 1e9,
           "step_frac_hbm": b / (step_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
           "ac_phased_cc_vs_input": float(np.corrcoef(acp.ravel(), np.asarray(ac, np.float64).ravel())[0, 1])}
    if not no_cpu:
        from oracle import phase as OP
        ac_filt_, amp, mask = OP.amplitudes_of(np.asarray(ac, np.float64))
        sup0 = OP.create_support_(ac_filt_, M, int((M - 1) / 4))
        rho0 = sup0 * np.random.default_rng(0).random(sup0.shape)
        n = 25
        t = time.perf_counter()
        for _ in range(n):
            OP.ER(rho0, amp, mask, sup0, np.inf)
        per = (time.perf_counter() - t) / n
        out["cpu_baseline"] = {"step_ms": per * 1e3, "phase_ms_extrapolated": per * steps * 1e3, "cores": 1,
                               "kind": "port", "sample": f"oracle ER steps (numpy complex128 FFTs) x {n}, "
                                                         f"extrapolated to {steps} steps"}
    return out


def cpu_solve_parts(M, s, ext, n_o):
    """The oracle's type-1 (real strengths, M modes, FP64 ES at eps 1e-6) over the
    q-vectors of n_o image orientations, and one application of the normal operator of
    the CG solve (Toeplitz embedding at 2M + the corner projection, numpy complex128
    FFTs), beside the GPU's type1_real_M81 and cg_ms_per_iteration."""
    from cmtip_amd import synth
    from oracle import ac as OAC
    from oracle import geom
    from oracle import nufft as O

    q = synth.random_quaternions(n_o, seed=1)
    x, y, z = geom.rotated_coords(q, s, ext)
    c = np.random.default_rng(0).random(len(x))
    t = time.perf_counter()
    O.nufft1(c, x, y, z, M, eps=1e-6)
    t1 = time.perf_counter() - t
    F = OAC.kernel_spectrum(x[:65536], y[:65536], z[:65536], M, eps=1e-6)
    anti = OAC.gen_F_antisupport(M)
    sup = np.ones((M,) * 3)
    u = np.random.default_rng(1).random(M ** 3)
    t = time.perf_counter()
    for _ in range(2):
        OAC.core_problem_convolution(u, M, F, 2 * M, sup) + 1e-4 * u + 1e3 * OAC.fourier_reg(u, sup, anti, M)
    t_mv = (time.perf_counter() - t) / 2
    log(f"[bench] cpu: type-1 {len(x) / t1 / 1e6:.2f} Mpts/s, normal operator {t_mv * 1e3:.0f} ms")
    return {"type1_real_Mpts_s": len(x) / t1 / 1e6, "type1_points": len(x),
            "normal_operator_ms": t_mv * 1e3,
            "note": "type-1 over the q-vectors of the first image orientations; normal operator = the CG "
                    "iteration's matrix-vector product (its vector updates excluded)"}


if __name__ == "__main__":
    main()
