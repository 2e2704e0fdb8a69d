"""Benchmark of the chi2 + gradient evaluation on one MI355X.

One step = one chi2 + gradient evaluation of the whole workload with every
input already resident in device memory (patch pixels and metadata, the
proposals); outputs stay on the device.  Prints one JSON line.

The headline (value, ms_per_step, roofline) is the batched workload: 512
uneven patches (1-20 sources, 1-8 bands x 1-4 exposures, square stamps of
side 48-256 px, 3-6 PSF Gaussians, 9 radii) evaluated in one persistent
launch, whose "vs plain" figure is the same 512 patches evaluated by one
plain launch each.  The line also carries:
  plain     one patch of 8 bands x 2 exposures x 128x128 px, 15 sources, 9
            radii x 4 PSF Gaussians (1.416e8 pixel-Gaussian pairs), one
            launch per evaluation
  single_plain   1 band x 1 exposure x 64x64 px, 1 source, 9 x 3 Gaussians
  split_tiles    patches of fewer tiles than workgroup slots (16-144 tiles of
            540-1080 Gaussians), whole tiles against the split tiles that
            HIPProposer.balance plans (pieces of a tile share its Gaussians)
  parity    the multi-band patch's gradients and chi2 against the reference
            CPU kernel and the float64 oracle (computed once, untimed)
  autograd, ensemble, valu_probe, cpu_baseline(_parallel)

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""

import argparse
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

METRIC = "pixel·Gaussian evals/s (chi2+grad), batched vs plain, % of FP32 VALU peak"
FLOP_PER_EVAL = 117          # FP32 flop per pair inside the cutoff, both passes (FMA = 2)
FLOP_OUTSIDE = 22            # flop per pair beyond it: dx, dy, F dx, dx^T F dx in each pass
PEAK_FP32 = 157.3e12         # MI355X FP32 vector peak, FLOP/s
TRAFFIC_FILE = os.path.join(REPO, "profiles", "chi2_grad_kernel_multi_band_traffic.txt")
BATCH_TRAFFIC_FILE = os.path.join(REPO, "profiles", "chi2_grad_kernel_batched_uneven_traffic.txt")
VALU_FILE = os.path.join(REPO, "profiles", "chi2_grad_kernel_multi_band_valu.txt")
BATCH_VALU_FILE = os.path.join(REPO, "profiles", "chi2_grad_kernel_batched_uneven_valu.txt")


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20,
                    help="timed batched launches (the headline)")
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--plain-steps", type=int, default=200,
                    help="timed evaluations of the plain multi-band patch")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--config", type=int, default=None, help="kernel config (default: tuned)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--copy-mb", type=int, default=0,
                    help="also run one device copy of this many MiB (counter calibration)")
    ap.add_argument("--batch-patches", type=int, default=512,
                    help="uneven patches of the batched (headline) workload")
    ap.add_argument("--no-parity", action="store_true", help="skip the parity object")
    ap.add_argument("--no-probe", action="store_true", help="skip the VALU rate probe")
    ap.add_argument("--no-single", action="store_true",
                    help="skip the single-patch plain measurement")
    ap.add_argument("--no-split", action="store_true",
                    help="skip the split-tile measurement on few-tile patches")
    ap.add_argument("--no-autograd", action="store_true",
                    help="skip the autograd + leapfrog measurement")
    ap.add_argument("--ensemble-patches", type=int, default=64,
                    help="patches of the HMC ensemble measurement (0 skips it)")
    ap.add_argument("--ensemble-chains", type=int, default=64, help="chains per patch")
    ap.add_argument("--leapfrog-steps", type=int, default=32)
    return ap.parse_args()


def build_workload(name, seed):
    from forcepho_amd import scenes
    from forcepho_amd.patches import SimplePatch
    spec = scenes.spec_multi(seed) if name == "multi" else scenes.spec_single(seed)
    patch, scene, theta = scenes.build_patch(spec, SimplePatch, scenes.render_on_device)
    desc = ("multi-band plain: 8 bands x 2 exposures x 128x128 px, 15 sources, "
            "9 radii x 4 PSF Gaussians" if name == "multi" else
            "single plain: 1 band x 1 exposure x 64x64 px, 1 source, 9 radii x 3 PSF Gaussians")
    return patch, scene, desc


def valu_probe():
    """Measured FP32 FMA and v_exp_f32 rates of the device (fpx_probe_valu):
    8 dependent chains per lane (each wavefront waits on its results: a
    latency figure) and 32 FMA / 16 exp chains at 8 wavefronts per SIMD (the
    issue rate the VALU reaches), as operations/s over the launch, wavefront
    instructions per CU per cycle at MI355X's 2.4 GHz peak clock, and the
    shader clock the wavefronts saw (s_memtime against the 100 MHz
    s_memrealtime)."""
    import torch
    from forcepho_amd._lib import lib, check
    import ctypes
    n_cu = ctypes.c_int32()
    check(lib().fpx_device_info(ctypes.byref(n_cu), None, None), "fpx_device_info")
    res = {}
    for kind, name, iters, blocks_per_cu in ((0, "fma_chains8", 8192, 8), (1, "exp_chains8", 2048, 8),
                                             (2, "fma_chains32", 2000, 32),
                                             (3, "exp_chains16", 400, 32)):
        n_threads = 256 * blocks_per_cu * n_cu.value
        out = torch.empty(n_threads, dtype=torch.float32, device="cuda")
        ops = ctypes.c_double()
        s = torch.cuda.current_stream()
        check(lib().fpx_probe_valu(kind, iters, out.data_ptr(), n_threads, ctypes.byref(ops),
                                   s.cuda_stream), "probe")
        torch.cuda.synchronize()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(s)
        check(lib().fpx_probe_valu(kind, iters, out.data_ptr(), n_threads, ctypes.byref(ops),
                                   s.cuda_stream), "probe")
        b.record(s)
        torch.cuda.synchronize()
        v = out.view(-1, 64).cpu().numpy()
        cyc, ref = np.median(v[:, 0]), np.median(v[:, 1])
        rate = ops.value / (a.elapsed_time(b) * 1e-3)
        res[name + "_per_s"] = rate
        res[name + "_wave_inst_per_cu_cycle"] = rate / 64 / n_cu.value / 2.4e9
        res[name + "_clock_ghz"] = float(cyc / ref * 0.1)
    res["fma_tflops"] = 2 * res["fma_chains32_per_s"] / 1e12
    res["exp_per_s"] = res["exp_chains16_per_s"]
    return res


def roofline(n_pairs, n_inside, seconds):
    """FP32 VALU roofline of one evaluation: the reference algorithm's flops
    (117 per pair inside the cutoff, 22 per pair beyond it, where the
    per-pixel loops stop after the exponent) over the kernel time."""
    flops = n_inside * FLOP_PER_EVAL + max(n_pairs - n_inside, 0) * FLOP_OUTSIDE
    achieved = flops / seconds
    return {"bound": "valu", "achieved": achieved / 1e12, "peak": PEAK_FP32 / 1e12,
            "unit": "TFLOP/s", "frac": achieved / PEAK_FP32,
            "kernel_ms": seconds * 1e3, "pairs": int(n_pairs), "pairs_inside_cutoff": int(n_inside),
            "flop_per_pair_inside": FLOP_PER_EVAL, "flop_per_pair_outside": FLOP_OUTSIDE,
            "frac_if_117_per_pair": n_pairs * FLOP_PER_EVAL / seconds / PEAK_FP32}


def roofline_line(base, traffic_bytes, seconds, n_inside, probe):
    """The roofline object plus the exponential throughput (two v_exp_f32 per
    pair inside the cutoff) against the probe's measured rate, and the HBM
    bandwidth the committed traffic figure implies."""
    line = dict(base, traffic=traffic_bytes)
    line["exp_per_s"] = 2.0 * n_inside / seconds
    if probe is not None:
        line["exp_frac_of_measured_peak"] = line["exp_per_s"] / probe["exp_per_s"]
    if traffic_bytes is not None:
        line["hbm_gb_per_s"] = traffic_bytes / seconds / 1e9
        line["hbm_frac"] = line["hbm_gb_per_s"] / 8000.0
    return line


def profile_record(cfg, path):
    """The JSON record of kernel configuration cfg in a committed rocprof
    summary (None if the file or that configuration's record is absent)."""
    import ctypes
    from forcepho_amd._lib import lib, check
    w, p = ctypes.c_int32(), ctypes.c_int32()
    check(lib().fpx_config_info(cfg, ctypes.byref(w), ctypes.byref(p), None), "fpx_config_info")
    key = f"{w.value}x{p.value}"
    try:
        with open(path) as f:
            rows = [json.loads(l) for l in f if l.startswith("{")]
    except OSError:
        return None
    for r in rows:
        if r.get("kernel") == key:
            return r
    return None


def traffic(cfg, path=TRAFFIC_FILE):
    """HBM bytes per launch of a workload for kernel configuration cfg, from
    its committed FETCH_SIZE / WRITE_SIZE summary (the multi-band workload's
    by default; None if that configuration was not profiled)."""
    r = profile_record(cfg, path)
    return None if r is None else r["hbm_bytes_per_launch"]


def executed(cfg, seconds, path=VALU_FILE):
    """FP32 flop and exponentials the kernel issues per launch (rocprof's
    SQ_INSTS_VALU_{FMA,MUL,ADD,TRANS}_F32 x 64 lanes, from the committed
    summary of the workload) over this run's kernel time: the work the
    moment-form kernel actually does, next to the reference algorithm's
    flops that the roofline credits.  None if not profiled."""
    r = profile_record(cfg, path)
    if r is None:
        return None
    tflops = r["executed_flop_per_launch"] / seconds / 1e12
    return {"flop_per_launch": r["executed_flop_per_launch"], "tflops": tflops,
            "frac": tflops * 1e12 / PEAK_FP32, "exp_per_launch": r["exp_per_launch"],
            "exp_per_s": r["exp_per_launch"] / seconds,
            "valu_wave_insts_per_launch": r["valu"]}


def build_uneven(n_patches, seed, with_scenes=False):
    """n_patches independent uneven patches: 1-20 sources, 1-8 bands, 1-4
    exposures per band, square stamps of side 48-256 px, 3-6 PSF Gaussians,
    9 radii (scenes.spec_uneven); data = model + noise."""
    from forcepho_amd import scenes
    from forcepho_amd.patches import SimplePatch
    rng = np.random.default_rng(seed)
    patches, proposals, scns = [], [], []
    for _ in range(n_patches):
        spec = scenes.spec_uneven(rng)
        patch, scene, _ = scenes.build_patch(spec, SimplePatch, scenes.render_on_device)
        patch.return_residual = False
        patch.send_to_device()
        prop = scene.get_proposal().copy()
        prop["q"] *= np.float32(0.97)
        patches.append(patch)
        proposals.append(prop)
        scns.append(scene)
    if with_scenes:
        return patches, proposals, scns
    return patches, proposals


def single_plain_measurement(args):
    """The single-patch plain launch path: one 64x64 exposure, one source,
    9 radii x 3 PSF Gaussians, evaluated back-to-back with device-resident
    inputs (launch-bound), and through evaluate_proposal (host proposal in,
    results out)."""
    import torch
    from forcepho_amd.proposal import HIPProposer
    patch, scene, desc = build_workload("single", args.seed)
    patch.return_residual = False
    patch.send_to_device()
    proposal = scene.get_proposal()
    raw = np.ascontiguousarray(proposal).view(np.uint8).reshape(-1)
    pdev = torch.from_numpy(raw.copy()).cuda()
    prop = HIPProposer(patch=patch)
    prop.autotune(pdev, patch, repeats=20)
    n = patch.n_evals()
    launch_s = time_region(lambda: prop.launch(pdev, patch), 200)
    prop.evaluate_proposal(proposal, patch=patch)
    t0 = time.perf_counter()
    for _ in range(200):
        prop.evaluate_proposal(proposal, patch=patch)
    api_s = (time.perf_counter() - t0) / 200
    return {"workload": desc, "evals_per_step": n,
            "launch_evals_per_s": n / launch_s, "launch_us": launch_s * 1e6,
            "evaluate_proposal_evals_per_s": n / api_s, "evaluate_proposal_us": api_s * 1e6}


def autograd_measurement(args, patch, scene):
    """The multi-band patch behind torch.autograd (PatchModel): chi2 value +
    gradient calls per second (the fused chi2, one chi2-mode launch, and the
    chi2 built by torch from the differentiable residual image: a model-only
    launch forward, the gradient pass backward), and a leapfrog trajectory
    of --leapfrog-steps steps driven by each (hmc.leapfrog, the inner
    integrator of forcepho's HMC), in unit-scaled parameters."""
    import torch
    from forcepho_amd.autograd import PatchModel
    from forcepho_amd.hmc import leapfrog
    model = PatchModel(patch, scene)
    theta = scene.get_all_source_params().copy()
    nb = int(patch.n_bands)
    pix_deg = 0.03 / 3600.0
    units = np.tile(np.concatenate([np.abs(theta.reshape(-1, nb + 6)[:, :nb]).mean(0).clip(1e-3)
                                    * 0.01, [0.1 * pix_deg, 0.1 * pix_deg, 0.01, 0.01, 0.01,
                                             0.001]]), model.n_sources)
    q0 = theta / units
    q0.reshape(-1, nb + 6)[:, nb + 2] *= 0.99

    def U_and_grad(q, chi2=model.chi2):
        t = torch.tensor(q * units, dtype=torch.float64, device=model.proposer.device,
                         requires_grad=True)
        u = 0.5 * chi2(t)
        u.backward()
        return u.item(), t.grad.cpu().numpy() * units

    def per_call(chi2):
        for _ in range(3):
            U_and_grad(q0, chi2)
        n = 20
        t0 = time.perf_counter()
        for _ in range(n):
            U_and_grad(q0, chi2)
        return (time.perf_counter() - t0) / n

    call_s = per_call(model.chi2)
    resid_s = per_call(model.chi2_from_residual)
    p0 = np.random.default_rng(args.seed).normal(size=len(q0))

    def trajectory(chi2):
        f = lambda q: U_and_grad(q, chi2)
        leapfrog(q0, p0, f, eps=0.01, n_steps=2)
        t0 = time.perf_counter()
        _, _, _, calls = leapfrog(q0, p0, f, eps=0.01, n_steps=args.leapfrog_steps)
        return time.perf_counter() - t0, calls

    traj_s, calls = trajectory(model.chi2)
    traj_resid_s, _ = trajectory(model.chi2_from_residual)
    post_dev_s, post_host_s = posterior_measurement(patch, scene, theta, nb)
    n_evals = patch.n_evals()
    return {"workload": "torch.autograd chi2 of the multi-band patch (fused: one chi2-mode "
                        "launch gives value and gradient)",
            "value_and_grad_per_s": 1.0 / call_s, "ms_per_value_and_grad": call_s * 1e3,
            "residual_path_ms_per_value_and_grad": resid_s * 1e3,
            "residual_path_trajectory_s": traj_resid_s,
            "fastposterior_ms_per_call": post_dev_s * 1e3,
            "fastposterior_host_records_ms_per_call": post_host_s * 1e3,
            "evals_per_s": n_evals / call_s,
            "leapfrog_steps": args.leapfrog_steps, "gradient_calls": calls,
            "trajectory_s": traj_s, "trajectories_per_s": 1.0 / traj_s}


def ensemble_bounds(theta, nb):
    """Prior box of one patch's parameters around theta (fluxes x0.5-1.5,
    +-3 pixels, q in [0.2, 1], pa in [-2, 2], sersic in [0.8, 5.5], rh in
    [0.03, 0.4] arcsec)."""
    t = theta.reshape(-1, nb + 6)
    lo, hi = t.copy(), t.copy()
    lo[:, :nb], hi[:, :nb] = t[:, :nb] * 0.5, t[:, :nb] * 1.5 + 1e-3
    pix = 0.03 / 3600.0
    lo[:, nb:nb + 2] -= 3 * pix
    hi[:, nb:nb + 2] += 3 * pix
    lo[:, nb + 2], hi[:, nb + 2] = 0.2, 1.0
    lo[:, nb + 3], hi[:, nb + 3] = -2.0, 2.0
    lo[:, nb + 4], hi[:, nb + 4] = 0.8, 5.5
    lo[:, nb + 5], hi[:, nb + 5] = 0.03, 0.4
    return lo.reshape(-1), hi.reshape(-1)


def ensemble_measurement(args):
    """--ensemble-chains chains on each of --ensemble-patches uneven patches,
    all positions on the device, one --leapfrog-steps leapfrog trajectory of
    every chain recorded once into a graph (batched persistent chi2 +
    gradient launches, bounded transform, mixture amplitudes, momentum and
    position updates) and replayed."""
    import torch
    from forcepho_amd.ensemble import HMCEnsemble
    t0 = time.perf_counter()
    patches, _, scns = build_uneven(args.ensemble_patches, args.seed + 2, with_scenes=True)
    bounds, start = [], []
    thetas = [sc.get_all_source_params().copy() for sc in scns]
    for patch, th in zip(patches, thetas):
        bounds.append(ensemble_bounds(th, int(patch.n_bands)))
    ens = HMCEnsemble(patches, scns, args.ensemble_chains, bounds)
    for pi, c in ens.batch.slot_map:
        t = thetas[pi].copy()
        nb = int(patches[pi].n_bands)
        t.reshape(-1, nb + 6)[:, nb + 2] *= 0.98 - 0.001 * c
        start.append(t)
    ens.set_theta(start)
    ens.potential()                    # source records of the start positions
    build_s = time.perf_counter() - t0
    ens.batch.autotune(repeats=1, configs=[2, 6, 7], persistent=(True,))
    gen = torch.Generator(device=ens.device)
    gen.manual_seed(args.seed)
    # per-chain step sizes (untimed): the starting heuristic, two 8-step
    # rounds, then two rounds of the recorded trajectory itself
    t1 = time.perf_counter()
    ens.eps.fill_(0.005)
    ens.find_step_size(generator=gen)
    ens.adapt_step_size(2, 8, generator=gen)
    t2 = time.perf_counter()
    ens.capture(args.leapfrog_steps)
    capture_s = time.perf_counter() - t2
    ens.adapt_step_size(2, args.leapfrog_steps, generator=gen, use_graph=True)
    torch.cuda.synchronize()
    adapt_s = time.perf_counter() - t1 - capture_s
    p0 = torch.randn(ens.p.shape, generator=gen, dtype=torch.float64,
                     device=ens.device) * ens.fmask
    ens.p.copy_(p0)
    traj_s = time_region(ens.graph.replay, 1)
    dH = ((ens.U + 0.5 * (ens.p * ens.p).sum(dim=(1, 2)))
          - (ens.U_start + 0.5 * (p0 * p0).sum(dim=(1, 2)))).cpu().numpy()
    accept = np.where(np.isfinite(dH), np.minimum(1.0, np.exp(-np.clip(dH, -50, None))), 0.0)
    n_evals = ens.batch.n_evals() * (args.leapfrog_steps + 1)
    small = small_ensemble_measurement(args)
    return {"workload": f"{ens.n_slots} chains ({args.ensemble_chains} per patch) on "
                        f"{len(patches)} uneven patches, graph-captured {args.leapfrog_steps}-step "
                        "leapfrog, state on the device",
            "chains": ens.n_slots, "trajectory_s": traj_s,
            "trajectories_per_s": ens.n_slots / traj_s,
            "evals_per_s": n_evals / traj_s,
            "kernel_config": "{}{}".format(ens.batch.config, "p" if ens.batch.persistent else ""),
            "mean_accept_prob": float(accept.mean()),
            "finite_energy_fraction": float(np.isfinite(dH).mean()),
            "median_abs_energy_error": float(np.median(np.abs(dH))),
            "host_build_s": build_s, "step_size_adaptation_s": adapt_s,
            "warmup_and_capture_s": capture_s, "small_regions": small}


def small_ensemble_measurement(args, n_patches=16, chains=4, side=64, n_bands=3, n_sources=4):
    """A dispatcher-sized ensemble (16 small regions of 64x64 px in 3 bands
    with 4 sources, 4 chains each): time per leapfrog step of the recorded
    trajectory against the batched chi2 launch alone, i.e. what the record
    build, transform, potential and updates add per step; and a whole HMC
    transition (momenta, trajectory, Metropolis test) against the trajectory
    alone."""
    import torch
    from forcepho_amd import scenes
    from forcepho_amd.patches import SimplePatch
    from forcepho_amd.ensemble import HMCEnsemble
    rng = np.random.default_rng(args.seed + 5)
    patches, scns, thetas = [], [], []
    for _ in range(n_patches):
        spec = scenes.draw_patch_spec(rng, n_bands, [1] * n_bands, [side] * n_bands, n_sources,
                                      [4] * n_bands)
        patch, scene, theta = scenes.build_patch(spec, SimplePatch, scenes.render_on_device)
        patch.return_residual = False
        patch.send_to_device()
        patches.append(patch)
        scns.append(scene)
        thetas.append(theta)
    ens = HMCEnsemble(patches, scns, chains, [ensemble_bounds(t, n_bands) for t in thetas])
    start = []
    for pi, c in ens.batch.slot_map:
        t = thetas[pi].copy()
        t.reshape(-1, n_bands + 6)[:, n_bands + 2] *= 0.98 - 0.001 * c
        start.append(t)
    ens.set_theta(start)
    ens.potential()
    ens.batch.autotune(repeats=3)
    gen = torch.Generator(device=ens.device)
    gen.manual_seed(args.seed)
    ens.eps.fill_(0.005)
    ens.find_step_size(generator=gen)
    ens.adapt_step_size(4, 8, generator=gen)
    ens.p.copy_(torch.randn(ens.p.shape, generator=gen, dtype=torch.float64,
                            device=ens.device) * ens.fmask)
    ens.capture(args.leapfrog_steps)
    traj_s = time_region(ens.graph.replay, 10)
    launch_s = time_region(ens.batch.launch, 50)
    ens.hmc_step(args.leapfrog_steps, generator=gen)
    hmc_s = time_region(lambda: ens.hmc_step(args.leapfrog_steps, generator=gen), 10)
    steps = args.leapfrog_steps + 1
    return {"workload": f"{n_patches} regions of {side}x{side} px, {n_bands} bands, "
                        f"{n_sources} sources, {chains} chains each, graph-captured "
                        f"{args.leapfrog_steps}-step leapfrog",
            "chains": ens.n_slots, "trajectory_s": traj_s,
            "trajectories_per_s": ens.n_slots / traj_s,
            "us_per_step": traj_s / steps * 1e6, "batch_launch_us": launch_s * 1e6,
            "hmc_transition_ms": hmc_s * 1e3, "trajectory_ms": traj_s * 1e3}


def posterior_measurement(patch, scene, theta, nb, n=50):
    """forcepho's own API on the multi-band patch: seconds per
    FastPosterior.lnprob_and_grad call (bounded transform around theta),
    with the records built on the device and with Scene.get_proposal on the
    host."""
    from forcepho_amd.model import FastPosterior, BoundedTransform
    from forcepho_amd.proposal import HIPProposer
    tr = BoundedTransform(*ensemble_bounds(theta, nb))
    z0 = tr.inverse_transform(theta.reshape(-1))
    out = []
    saved, patch.return_residual = patch.return_residual, False   # chi2 and gradient only
    for device_records in (None, False):
        post = FastPosterior(HIPProposer(patch=patch), patch, scene=scene, transform=tr,
                             device_records=device_records)
        for i in range(3):
            post.lnprob_and_grad(z0 + 1e-9 * (i + 1))
        t0 = time.perf_counter()
        for i in range(n):
            post.lnprob_and_grad(z0 + 1e-9 * (i + 10))     # a new position every call
        out.append((time.perf_counter() - t0) / n)
    patch.return_residual = saved
    return out[0], out[1]


def time_region(fn, steps):
    """Seconds per call of ``fn`` over ``steps`` back-to-back calls,
    bracketed by device synchronisations (HIP events on the current stream)."""
    import torch
    s = torch.cuda.current_stream()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    a.record(s)
    for _ in range(steps):
        fn()
    b.record(s)
    torch.cuda.synchronize()
    return a.elapsed_time(b) * 1e-3 / steps


def timed_region(fn, steps, warmup):
    """The timed region of the contract: warmup untimed calls, then exactly
    ``steps`` back-to-back calls between two HIP events on the launch
    stream, synchronised on both sides (no per-step markers: they would add
    their own time to every step).  Returns (seconds, wall seconds)."""
    import torch
    for _ in range(warmup):
        fn()
    s = torch.cuda.current_stream()
    start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    start.record(s)
    for _ in range(steps):
        fn()
    end.record(s)
    torch.cuda.synchronize()
    return start.elapsed_time(end) * 1e-3, time.perf_counter() - t0


def batched_measurement(args, plain_prop):
    """The headline: --batch-patches uneven patches evaluated in one
    persistent launch (autotuned tile shape), timed over --steps launches;
    then the tile kernel alone (its average launch time feeds the roofline),
    and the same patches evaluated by one plain launch each (device-resident
    inputs, the plain proposer's configuration) and through evaluate_proposal
    per patch (host proposal in, results out)."""
    import torch
    from forcepho_amd.batch import BatchEvaluator
    t0 = time.perf_counter()
    patches, proposals = build_uneven(args.batch_patches, args.seed + 1)
    build_s = time.perf_counter() - t0
    ev = BatchEvaluator(patches, chains=1)
    ev.set_proposals(proposals)
    tuned = ev.autotune(repeats=2)
    n_evals = ev.n_evals()
    n_inside = int(ev.count_pairs().sum())
    region, wall = timed_region(ev.launch, args.steps, args.warmup)
    batch_s = region / args.steps
    for (chi2, d) in ev.results():
        assert np.all(np.isfinite(chi2)) and np.all(np.isfinite(d))
    # the dominant kernel alone: back-to-back tile launches, then one output pass
    kernel_s = time_region(lambda: ev.launch(outputs=False), max(2, args.steps // 2))
    ev.finalize()
    torch.cuda.synchronize()

    # one plain launch per patch, the plain proposer's kernel configuration
    pdevs = [torch.from_numpy(np.ascontiguousarray(p).view(np.uint8).reshape(-1).copy()).cuda()
             for p in proposals]

    def sequential():
        for patch, pd in zip(patches, pdevs):
            plain_prop.launch(pd, patch)
    sequential()
    seq_s = time_region(sequential, 2)

    def sequential_api():
        for patch, pr in zip(patches, proposals):
            plain_prop.evaluate_proposal(pr, patch=patch)
    sequential_api()
    t1 = time.perf_counter()
    sequential_api()
    api_s = time.perf_counter() - t1
    sides = [int(np.sqrt(p.original_sizes[0])) for p in patches]
    desc = (f"{len(patches)} uneven patches (1-20 sources, 1-8 bands x 1-4 exposures, "
            f"48-256 px stamps, 3-6 PSF Gaussians, 9 radii) in one persistent launch")
    # the committed counter summaries describe the profiled workload (512 patches, seed 0) only
    profiled = args.batch_patches == 512 and args.seed == 0
    line = roofline(n_evals, n_inside, kernel_s)
    line["traffic"] = traffic(ev.config, BATCH_TRAFFIC_FILE) if profiled else None
    if line["traffic"] is not None:
        line["hbm_gb_per_s"] = line["traffic"] / kernel_s / 1e9
        line["hbm_frac"] = line["hbm_gb_per_s"] / 8000.0
    line["executed"] = executed(ev.config, kernel_s, BATCH_VALU_FILE) if profiled else None
    return {
        "value": n_evals / batch_s,
        "region_s": region,
        "wall_s": wall,
        "ms_per_step": batch_s * 1e3,
        "config": {"workload": desc, "n_patches": len(patches), "evals_per_step": n_evals,
                   "pixels": int(sum(int(p.npix) for p in patches)),
                   "stamp_side_range": [min(sides), max(sides)], "path": "batched persistent",
                   "kernel_config": "{}{}".format(ev.config, "p" if ev.persistent else ""),
                   "config_ms": {f"{c}{'p' if p else ''}": t * 1e3 for (c, p), t in tuned.items()},
                   "host_build_s": build_s},
        "roofline": line,
        "vs_plain": {
            "workload": "the same patches, one plain launch each (device-resident inputs)",
            "sequential_plain_launches": {"evals_per_s": n_evals / seq_s, "ms": seq_s * 1e3},
            "sequential_evaluate_proposal": {"evals_per_s": n_evals / api_s, "ms": api_s * 1e3},
            "speedup_vs_sequential_plain": seq_s / batch_s,
            "speedup_vs_sequential_evaluate_proposal": api_s / batch_s},
    }


def plain_measurement(args):
    """The multi-band patch (8 bands x 2 exposures x 128x128 px, 15
    sources) evaluated one plain launch at a time: autotuned tile shape, then
    --plain-steps back-to-back evaluations, then the tile kernel alone."""
    import torch
    from forcepho_amd.proposal import HIPProposer
    patch, scene, desc = build_workload("multi", args.seed)
    patch.return_residual = False
    patch.send_to_device()
    proposal = scene.get_proposal()
    proposal["q"] *= np.float32(0.97)   # off the truth: gradients are generic
    n_evals = patch.n_evals()
    raw = np.ascontiguousarray(proposal).view(np.uint8).reshape(-1)
    pdev = torch.from_numpy(raw.copy()).cuda()
    prop = HIPProposer(patch=patch)
    configs = [args.config] if args.config is not None else None
    tuned = prop.autotune(pdev, patch, repeats=10, configs=configs)
    region, wall = timed_region(lambda: prop.launch(pdev, patch), args.plain_steps, 10)
    kernel_s = time_region(lambda: prop.launch(pdev, patch, outputs=False), args.plain_steps)
    prop.finalize()
    chi2, derivs = prop.evaluate_proposal(proposal, patch=patch)
    assert np.all(np.isfinite(chi2)) and np.all(np.isfinite(derivs))
    n_inside = prop.count_pairs(pdev, patch)
    # the committed counter summaries describe the multi-band workload at seed 0 only
    profiled = args.seed == 0
    out = {"workload": desc, "evals_per_step": n_evals, "path": "plain",
           "kernel_config": f"{prop.config}{'p' if prop.persistent else ''}",
           "config_ms": {f"{c}{'p' if p else ''}": t * 1e3 for (c, p), t in tuned.items()},
           "evals_per_s": n_evals * args.plain_steps / region,
           "ms_per_step": region / args.plain_steps * 1e3, "steps": args.plain_steps,
           "roofline": dict(roofline_line(roofline(n_evals, n_inside, kernel_s),
                                          traffic(prop.config) if profiled else None,
                                          kernel_s, n_inside, None),
                            executed=executed(prop.config, kernel_s) if profiled else None)}
    return out, prop, patch, scene, proposal


# few-tile patches of the split measurement: (description, spec_multi keywords)
FEW_TILE_PATCHES = [
    ("2 bands x 1 exposure x 64x64 px, 15 sources, 9 radii x 4 PSF Gaussians",
     dict(side=64, n_bands=2, n_exp=1, n_sources=15, n_psf_comp=4)),
    ("1 band x 1 exposure x 64x64 px, 10 sources, 9 radii x 6 PSF Gaussians",
     dict(side=64, n_bands=1, n_exp=1, n_sources=10, n_psf_comp=6)),
    ("4 bands x 1 exposure x 96x96 px, 20 sources, 9 radii x 6 PSF Gaussians",
     dict(side=96, n_bands=4, n_exp=1, n_sources=20, n_psf_comp=6)),
]


def split_measurement(args, steps=200):
    """Patches with fewer tiles than the persistent grid has workgroup slots:
    one evaluation per launch (8 x 4 tiles, persistent), timed with whole
    tiles and with the split table HIPProposer.balance plans (the heavy
    tiles shared by several workgroups, each over its share of the
    Gaussians); the results of the two compared once, untimed."""
    import torch
    from forcepho_amd import scenes
    from forcepho_amd.patches import SimplePatch
    from forcepho_amd.proposal import HIPProposer
    out = []
    for desc, kw in FEW_TILE_PATCHES:
        spec = scenes.spec_multi(args.seed + 1, **kw)
        patch, scene, _ = scenes.build_patch(spec, SimplePatch, scenes.render_on_device)
        patch.return_residual = False
        patch.send_to_device()
        proposal = scene.get_proposal()
        proposal["q"] *= np.float32(0.97)
        raw = np.ascontiguousarray(proposal).view(np.uint8).reshape(-1)
        pdev = torch.from_numpy(raw.copy()).cuda()
        prop = HIPProposer(patch=patch, config=5, persistent=True)
        row = {"workload": desc, "evals_per_step": patch.n_evals()}
        results = {}
        for label, split in (("whole", False), ("split", True)):
            prop.balance(proposal, patch, split=split)
            table = patch.split_tiles(prop.pix_per_tile) if split else None
            n_items = table[1] if table is not None else patch.device_tiles(prop.pix_per_tile)[1]
            results[label] = prop.evaluate_proposal(pdev, patch=patch)
            for _ in range(10):
                prop.launch(pdev, patch)
            t = time_region(lambda: prop.launch(pdev, patch), steps)
            # forcepho's call: numpy proposal in, numpy results out
            prop.evaluate_proposal(proposal, patch=patch)
            t0 = time.perf_counter()
            for _ in range(steps):
                prop.evaluate_proposal(proposal, patch=patch)
            api = (time.perf_counter() - t0) / steps
            row[label] = {"items": int(n_items), "us": t * 1e6,
                          "evals_per_s": row["evals_per_step"] / t,
                          "evaluate_proposal_us": api * 1e6}
        row["speedup"] = row["whole"]["us"] / row["split"]["us"]
        row["evaluate_proposal_speedup"] = (row["whole"]["evaluate_proposal_us"]
                                            / row["split"]["evaluate_proposal_us"])
        (c0, g0), (c1, g1) = results["whole"], results["split"]
        scale = np.abs(g0).max(axis=(1, 2), keepdims=True) + 1e-30
        row["split_vs_whole"] = {"chi2_rel_max": float(np.max(np.abs(c1 - c0) / np.abs(c0))),
                                 "grad_rel_of_band_max": float(np.max(np.abs(g1 - g0) / scale))}
        out.append(row)
    return out


def parity_measurement():
    """Parity of the multi-band workload against its bar (chi2 1e-6 relative,
    gradients 1e-4 of each band's largest, against the reference CPU kernel), once and
    untimed: the patch of tests/golden/make_fullsize.py (the benchmark's
    scene with data rendered on the host, every q scaled by 0.97) evaluated
    by the HIP kernel, by the reference's CPU kernel (oracle/_ref, the
    checker) and from the committed float64 oracle values of the same patch.
    Relative chi2 errors, and gradient errors per entry and relative to the
    largest |gradient| of the band, as p50 / p99 / max, for GPU vs the
    reference kernel, GPU vs float64 and the reference kernel vs float64."""
    import importlib.util
    import oracle
    from forcepho_amd.proposal import HIPProposer
    path = os.path.join(REPO, "tests", "golden", "make_fullsize.py")
    spec = importlib.util.spec_from_file_location("make_fullsize", path)
    F = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(F)
    patch, proposal = F.build()
    fp64 = np.load(F.OUT)
    if not np.allclose(F.data_digest(patch), fp64["digest"], rtol=1e-12):
        return {"error": "patch differs from the committed float64 fixture"}
    patch.return_residual = False
    patch.send_to_device()
    prop = HIPProposer(patch=patch)
    c_gpu, d_gpu = prop.evaluate_proposal(proposal, patch=patch)
    mod = oracle.reference_module()
    if mod is None:
        return {"error": "the reference CPU kernel is not built (oracle/_ref)"}
    c_ref, d_ref, _ = oracle.evaluate_reference(patch, proposal, module=mod)
    d_ref = d_ref.astype(np.float64)
    c64, g64 = fp64["chi2"], fp64["dchi2"]

    def q3(x):
        x = np.asarray(x, dtype=np.float64).reshape(-1)
        return {"p50": float(np.percentile(x, 50)), "p99": float(np.percentile(x, 99)),
                "max": float(x.max())}

    def entry(g, h):
        return np.abs(g - h) / np.where(h != 0, np.abs(h), np.inf)

    def band(g, h):
        m = np.abs(h).max(axis=(1, 2), keepdims=True)
        return np.abs(g - h) / np.where(m > 0, m, np.inf)

    pairs = {"gpu_vs_reference": (d_gpu, d_ref, c_gpu, c_ref),
             "gpu_vs_fp64": (d_gpu, g64, c_gpu, c64),
             "reference_vs_fp64": (d_ref, g64, c_ref, c64)}
    return {"workload": "multi-band patch (8 bands x 2 exposures x 128x128 px, 15 sources), "
                        "host-rendered data, proposal q x 0.97",
            "bar": "chi2 <= 1e-6 relative, gradients <= 1e-4 of the band's largest |dchi2/dp|, "
                   "against the reference CPU kernel",
            "chi2_rel_max": {k: float(np.max(np.abs(c - h) / np.abs(h)))
                             for k, (_, _, c, h) in pairs.items()},
            "grad_rel_per_entry": {k: q3(entry(g, h)) for k, (g, h, _, _) in pairs.items()},
            "grad_rel_of_band_max": {k: q3(band(g, h)) for k, (g, h, _, _) in pairs.items()}}


def cpu_model():
    """The host CPU's model name (/proc/cpuinfo)."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(patch, proposal, n_evals, budget_s=12.0):
    """The reference CPU kernel (compiled from its sources into oracle/_ref),
    or the C restatement if that is absent, single-threaded, on whole
    evaluations of the same workload (one call per band, as forcepho's
    CPUProposer does) until about budget_s of CPU time."""
    import oracle
    mod = oracle.reference_module()
    kind = "reference" if mod is not None else "port"
    fn = ((lambda: oracle.evaluate_reference(patch, proposal, module=mod)) if mod is not None
          else (lambda: oracle.evaluate_cpu(patch, proposal)))
    reps, t0 = 0, time.perf_counter()
    while True:
        fn()
        reps += 1
        el = time.perf_counter() - t0
        if el >= budget_s or el / reps * (reps + 1) > 2.5 * budget_s:
            break
    return {"value": n_evals * reps / el, "unit": "evals/s", "cores": 1, "kind": kind,
            "sample": f"{reps} full evaluation(s) of the benchmark workload, "
                      f"{el:.1f} s on one host core"}


def cpu_baseline_parallel(patch, proposal, n_evals, budget_s=6.0):
    """The C restatement of the reference CPU kernel (bit-identical to it on
    the test fixtures) on every host thread of this process's CPU share at
    once (OMP_NUM_THREADS, else the CPU count), each thread evaluating the
    whole workload repeatedly; ctypes releases the GIL during the kernel
    calls.  Aggregate evals/s over the wall time until every thread ends."""
    import threading
    import oracle
    threads = max(1, int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count() or 1))
    counts = [0] * threads
    oracle.evaluate_cpu(patch, proposal)        # loads the library once
    t0 = time.perf_counter()
    stop = t0 + budget_s

    def work(i):
        while time.perf_counter() < stop:
            oracle.evaluate_cpu(patch, proposal)
            counts[i] += 1

    pool = [threading.Thread(target=work, args=(i,)) for i in range(threads)]
    for t in pool:
        t.start()
    for t in pool:
        t.join()
    el = time.perf_counter() - t0
    return {"value": n_evals * sum(counts) / el, "unit": "evals/s", "cores": threads,
            "kind": "port", "sample": f"{sum(counts)} full evaluation(s) of the benchmark "
                                      f"workload on {threads} threads in {el:.1f} s"}


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on exactly one GPU (asked for {args.gpus}); the path shards "
              "by patch and chain with no exchange, and multi-GPU runs are not supported",
              file=sys.stderr)
        sys.exit(2)
    import torch
    torch.cuda.set_device(0)
    if args.copy_mb:
        # a plain copy kernel of known size: its FETCH_SIZE / WRITE_SIZE
        # counters calibrate the chi2 kernel's
        a = torch.ones(args.copy_mb * 2 ** 18, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        del a, b

    plain, plain_prop, patch, scene, proposal = plain_measurement(args)
    print(f"plain multi-band: {plain['ms_per_step']:.4f} ms per evaluation", file=sys.stderr,
          flush=True)
    if args.batch_patches > 0:
        head = batched_measurement(args, plain_prop)
        print(f"batched: {head['ms_per_step']:.3f} ms per launch", file=sys.stderr, flush=True)
    else:
        # a run of the plain workload alone (counter collection): the plain
        # measurement stands in for the headline
        head = {"value": plain["evals_per_s"], "ms_per_step": plain["ms_per_step"],
                "config": {"workload": plain["workload"],
                           "evals_per_step": plain["evals_per_step"]},
                "wall_s": None, "roofline": plain["roofline"], "vs_plain": {}}
    probe = None if args.no_probe else valu_probe()
    if probe is not None:
        for r in (head["roofline"], plain["roofline"]):
            n_in = r["pairs_inside_cutoff"]
            r["exp_per_s"] = 2.0 * n_in / (r["kernel_ms"] * 1e-3)
            r["exp_frac_of_measured_peak"] = r["exp_per_s"] / probe["exp_per_s"]
    single = None if args.no_single else single_plain_measurement(args)
    if single is not None:
        head["vs_plain"]["speedup_vs_single_patch_plain_launches"] = (
            head["value"] / single["launch_evals_per_s"])
    parity = None if args.no_parity else parity_measurement()
    split = None if args.no_split else split_measurement(args)
    ensemble = ensemble_measurement(args) if args.ensemble_patches > 0 else None
    autograd = None if args.no_autograd else autograd_measurement(args, patch, scene)
    line = {
        "metric": METRIC,
        "value": head["value"],
        "unit": "evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (seeded scenes, model + Gaussian noise at S/N~50)",
        "config": dict(head["config"], wall_s=head["wall_s"]),
        "roofline": head["roofline"],
        "vs_plain": head["vs_plain"],
        "plain": plain,
        "single_plain": single,
        "split_tiles": split,
        "parity": parity,
        "autograd": autograd,
        "ensemble": ensemble,
        "valu_probe": probe,
    }
    if not args.no_cpu:
        model = cpu_model()
        n_evals = patch.n_evals()
        line["cpu_baseline"] = dict(cpu_baseline(patch, proposal, n_evals), cpu_model=model)
        line["cpu_baseline_parallel"] = dict(cpu_baseline_parallel(patch, proposal, n_evals),
                                             cpu_model=model)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
