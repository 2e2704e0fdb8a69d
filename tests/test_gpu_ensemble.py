"""Device-resident HMC ensemble: potential and gradient against the host
Posterior path, graph-captured trajectories against eager ones, and the
gradient error distribution against the float64 oracle."""

import numpy as np
import pytest

import fpx_testutil as U
import oracle
from forcepho_amd import scenes

pytestmark = pytest.mark.gpu


def _bounds(theta, nb):
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


@pytest.fixture(scope="module")
def ensemble():
    from forcepho_amd.patches import SimplePatch
    from forcepho_amd.ensemble import HMCEnsemble
    rng = np.random.default_rng(77)
    pats, scs, bnds, thetas = [], [], [], []
    for k in range(3):
        nb = [1, 3, 2][k]
        spec = scenes.draw_patch_spec(rng, nb, [1] * nb, [40 + 8 * k] * nb, [1, 4, 2][k],
                                      [3] * nb)
        patch, scene, theta = scenes.build_patch(spec, SimplePatch, oracle.render_model)
        patch.send_to_device()
        pats.append(patch)
        scs.append(scene)
        bnds.append(_bounds(theta, nb))
        thetas.append(theta)
    chains = [2, 1, 2]
    ens = HMCEnsemble(pats, scs, chains, bnds)
    start = []
    for pi, c in ens.batch.slot_map:
        t = thetas[pi].copy()
        nb = int(pats[pi].n_bands)
        t.reshape(-1, nb + 6)[:, nb + 2] *= 0.97 - 0.02 * c
        start.append(t)
    ens.set_theta(start)
    return ens, pats, scs, bnds, start


def test_potential_matches_host_posterior(ensemble):
    from forcepho_amd.model import FastPosterior
    from forcepho_amd.proposal import HIPProposer
    ens, pats, scs, bnds, start = ensemble
    Ud, gd = ens.potential()
    Ud, gd = Ud.cpu().numpy(), gd.cpu().numpy()
    for s, (pi, _) in enumerate(ens.batch.slot_map):
        patch = pats[pi]
        nb = int(patch.n_bands)
        post = FastPosterior(HIPProposer(patch=patch), patch, scene=scs[pi],
                             lower=bnds[pi][0], upper=bnds[pi][1], device_records=False)
        z = post.transform.inverse_transform(start[s])
        lnp, grad = post.lnprob_and_grad(z)
        assert Ud[s] == pytest.approx(-lnp, rel=1e-6)
        g = np.concatenate([gd[s, :, :nb], gd[s, :, ens.NB:]], axis=1)[:patch.n_sources].reshape(-1)
        np.testing.assert_allclose(g, -grad, rtol=1e-4, atol=1e-5 * np.abs(grad).max())


def test_graph_replay_matches_eager(ensemble):
    import torch
    ens = ensemble[0]
    z0 = ens.z.clone()
    gen = torch.Generator(device=ens.device)
    gen.manual_seed(5)
    p0 = torch.randn(ens.p.shape, generator=gen, dtype=torch.float64, device=ens.device) * ens.fmask
    ens.eps.fill_(0.01)
    ens.p.copy_(p0)
    ens.trajectory(8, use_graph=False)
    z_eager, p_eager = ens.z.clone(), ens.p.clone()
    for rep in range(3):
        ens.z.copy_(z0)
        ens.p.copy_(p0)
        ens.trajectory(8, use_graph=True)
        torch.cuda.synchronize()
        np.testing.assert_allclose(ens.z.cpu().numpy(), z_eager.cpu().numpy(), rtol=1e-9, atol=1e-12)
        np.testing.assert_allclose(ens.p.cpu().numpy(), p_eager.cpu().numpy(), rtol=1e-9, atol=1e-12)
        ens.potential()           # eager work between replays
    ens.z.copy_(z0)


def test_trajectory_energy_and_gradient_errors(ensemble):
    import torch
    ens, pats, scs, bnds, start = ensemble
    gen = torch.Generator(device=ens.device)
    gen.manual_seed(9)
    ens.eps.fill_(0.005)
    acc = [ens.hmc_step(32, generator=gen) for _ in range(3)]
    assert sum(int(a.sum()) for a in acc) >= 1
    # gradient errors of the chi2 kernel at the chains' final positions
    theta = ens.theta()
    errs = []
    for s, (pi, _) in enumerate(ens.batch.slot_map):
        scs[pi].set_all_source_params(theta[s])
        prop = scs[pi].get_proposal()
        from forcepho_amd.proposal import HIPProposer
        _, d = HIPProposer(patch=pats[pi]).evaluate_proposal(prop, patch=pats[pi])
        _, g64, mass, smass = U.fp64_reference(pats[pi], prop)
        errs.append(U.grad_errors(d, g64, mass, smass).reshape(-1))
    errs = np.concatenate(errs)
    print("gradient error / tolerance quantiles (50, 90, 99, 100%):",
          np.percentile(errs, [50, 90, 99, 100]))
    assert errs.max() <= 1.0
    assert np.median(errs) <= 0.1


def test_uneven_ensemble_gradient_error_distribution():
    """The benchmark's HMC ensemble workload in small: its uneven patch mix
    (the five patches with the fewest pairs among the 64 that the benchmark
    draws), two chains each, one graph-captured 32-step trajectory of every
    chain; then the batched kernel's gradient at every chain's position
    against the float64 oracle evaluated on the same float32 records.  Prints
    the distribution of |error| / tolerance and of the error relative to the
    largest gradient entry of the band, next to the same for the reference's
    CPU kernel on the same records where it is built; every entry within
    tolerance."""
    import torch
    from forcepho_amd.patches import SimplePatch
    from forcepho_amd.ensemble import HMCEnsemble
    from forcepho_amd.kernel_limits import MAXSOURCES
    from forcepho_amd.sources import source_struct_dtype
    rng = np.random.default_rng(2)          # the benchmark's ensemble draws (seed 0 + 2)
    specs = [scenes.spec_uneven(rng) for _ in range(64)]
    sizes = []
    for k, spec in enumerate(specs):
        p = SimplePatch()
        p.build_patch(spec.stamps, spec.scene, sky_zero=spec.sky_zero)
        sizes.append((p.n_evals(), k))
    pick = [k for _, k in sorted(sizes)[:5]]
    pats, scs, bnds, thetas = [], [], [], []
    for k in pick:
        patch, scene, theta = scenes.build_patch(specs[k], SimplePatch, oracle.render_model)
        patch.send_to_device()
        pats.append(patch)
        scs.append(scene)
        bnds.append(_bounds(theta, int(patch.n_bands)))
        thetas.append(theta)
    ens = HMCEnsemble(pats, scs, 2, bnds)
    start = []
    for pi, c in ens.batch.slot_map:
        t = thetas[pi].copy()
        nb = int(pats[pi].n_bands)
        t.reshape(-1, nb + 6)[:, nb + 2] *= 0.98 - 0.001 * c
        start.append(t)
    ens.set_theta(start)
    gen = torch.Generator(device=ens.device)
    gen.manual_seed(4)
    ens.eps.fill_(0.005)
    print(f"patches {pick}: pairs {[int(s) for s, _ in sorted(sizes)[:5]]}", flush=True)
    ens.hmc_step(32, generator=gen)
    ens.potential()                          # the batched launch at the final positions
    torch.cuda.synchronize()
    recs = ens.batch.proposals.cpu().numpy().view(source_struct_dtype).reshape(-1, MAXSOURCES)
    resp = ens.resp.cpu().numpy()
    ref_mod = oracle.reference_module()
    errs, rel, ref_errs, ref_rel = [], [], [], []
    for s, (pi, _) in enumerate(ens.batch.slot_map):
        ns, nb = int(pats[pi].n_sources), int(pats[pi].n_bands)
        d = resp[s, :nb, :ns].astype(np.float64)
        prop = recs[s, :ns].copy()
        _, g64, mass, smass = U.fp64_reference(pats[pi], prop)
        scale = np.abs(g64).max(axis=(1, 2), keepdims=True)
        errs.append(U.grad_errors(d, g64, mass, smass).reshape(-1))
        rel.append((np.abs(d - g64) / scale).reshape(-1))
        if ref_mod is not None:
            dr = oracle.evaluate_reference(pats[pi], prop, module=ref_mod)[1].astype(np.float64)
            ref_errs.append(U.grad_errors(dr, g64, mass, smass).reshape(-1))
            ref_rel.append((np.abs(dr - g64) / scale).reshape(-1))
    errs, rel = np.concatenate(errs), np.concatenate(rel)
    q = [50, 90, 99, 100]
    print("gradient |error| / tolerance quantiles", q, np.percentile(errs, q), flush=True)
    print("gradient |error| / max|g| of the band quantiles", q, np.percentile(rel, q), flush=True)
    if ref_mod is not None:
        ref_errs, ref_rel = np.concatenate(ref_errs), np.concatenate(ref_rel)
        print("reference CPU kernel: |error| / tolerance", np.percentile(ref_errs, q),
              "|error| / max|g|", np.percentile(ref_rel, q), flush=True)
        assert ref_errs.max() <= 1.0
    assert errs.max() <= 1.0
    assert np.median(errs) <= 0.1


def _torch_trajectory(ens, n_steps):
    """The leapfrog of HMCEnsemble written out in torch operations (theta,
    Jacobian, potential, gradient and updates as tensor expressions; records
    from theta, the batched launch and the stacking as library calls)."""
    import torch
    from forcepho_amd.device_scene import stack_grad_device

    def potential():
        sz = torch.sigmoid(ens.z)
        rng = ens.hi - ens.lo
        jac = rng * sz * (1.0 - sz)
        ens.records.write((ens.lo + sz * rng).contiguous())
        ens.batch.launch()
        g_theta, chi2 = stack_grad_device(ens.resp, ens.S, ens.NB, chi2=ens.chi2)
        U = 0.5 * chi2 - (torch.log(jac) * ens.fmask).sum(dim=(1, 2))
        return U, (0.5 * g_theta * jac - (1.0 - 2.0 * sz)) * ens.fmask

    U, g = potential()
    ens.p.sub_(0.5 * ens.eps * g)
    for i in range(n_steps):
        ens.z.add_(ens.eps * ens.p * ens.fmask)
        U, g = potential()
        if i < n_steps - 1:
            ens.p.sub_(ens.eps * g)
    ens.p.sub_(0.5 * ens.eps * g)
    return U, g


def test_fused_leapfrog_matches_torch_formulation(ensemble):
    """fpx_build_sources from z with bounds and fpx_leapfrog_update give the
    potential, gradient and 8-step trajectory of the same maps written as
    torch operations."""
    import torch
    ens = ensemble[0]
    gen = torch.Generator(device=ens.device)
    gen.manual_seed(21)
    z0 = ens.z.clone()
    p0 = torch.randn(ens.p.shape, generator=gen, dtype=torch.float64, device=ens.device) * ens.fmask
    ens.eps.fill_(0.004)
    U_f, g_f = ens.potential()
    ens.p.copy_(p0)
    ens.trajectory(8, use_graph=False)
    z_f, p_f = ens.z.clone(), ens.p.clone()
    ens.z.copy_(z0)
    ens.p.copy_(p0)
    U_t, g_t = _torch_trajectory(ens, 0)[0], None
    ens.z.copy_(z0)
    ens.p.copy_(p0)
    _torch_trajectory(ens, 8)
    z_t, p_t = ens.z.clone(), ens.p.clone()
    ens.z.copy_(z0)
    ens.p.copy_(p0)
    U0 = ens.potential()[0]
    np.testing.assert_allclose(U_f.cpu().numpy(), U0.cpu().numpy(), rtol=1e-12)   # atomic order
    np.testing.assert_allclose(U_f.cpu().numpy(), U_t.cpu().numpy(), rtol=1e-9)
    for a, b in ((z_f, z_t), (p_f, p_t)):
        a, b = a.cpu().numpy(), b.cpu().numpy()
        np.testing.assert_allclose(a, b, rtol=1e-6, atol=1e-6 * np.abs(b).max())
    assert np.abs(z_f.cpu().numpy() - z0.cpu().numpy()).max() > 0


def test_trajectory_end_point_potentials(ensemble):
    """The Metropolis energies come from the trajectory's own first and last
    evaluations: U_start is U at the start point and U at the end point,
    eager and replayed from the recorded graph alike."""
    import torch
    ens = ensemble[0]
    gen = torch.Generator(device=ens.device)
    gen.manual_seed(33)
    z0 = ens.z.clone()
    p0 = torch.randn(ens.p.shape, generator=gen, dtype=torch.float64, device=ens.device) * ens.fmask
    ens.eps.fill_(0.004)
    U0 = ens.potential()[0].clone()
    ends = []
    for use_graph in (False, True):
        ens.z.copy_(z0)
        ens.p.copy_(p0)
        ens.trajectory(6, use_graph=use_graph)
        Us, Ue = ens.U_start.clone(), ens.U.clone()
        U1 = ens.potential()[0].clone()
        np.testing.assert_allclose(Us.cpu().numpy(), U0.cpu().numpy(), rtol=1e-12)
        np.testing.assert_allclose(Ue.cpu().numpy(), U1.cpu().numpy(), rtol=1e-12)
        ends.append(ens.z.clone())
    np.testing.assert_allclose(ends[0].cpu().numpy(), ends[1].cpu().numpy(), rtol=1e-12, atol=1e-14)
    ens.z.copy_(z0)
    with pytest.raises(ValueError):
        ens.trajectory(0, use_graph=False)


def test_metropolis_kernel_matches_torch():
    """fpx_hmc_metropolis against the test written in torch: accept where
    ln u < H0 - H1 with H = U + |p|^2 / 2, rejected chains back to z0, a NaN
    energy rejects."""
    import torch
    from forcepho_amd._lib import lib, check
    rng = np.random.default_rng(17)
    n, m = 37, 53
    dev = dict(dtype=torch.float64, device="cuda")
    U0 = torch.tensor(rng.normal(0, 2, n), **dev)
    U1 = U0 + torch.tensor(rng.normal(0, 1, n), **dev)
    U1[3] = float("nan")
    p0 = torch.tensor(rng.normal(size=(n, m)), **dev)
    p1 = p0 + torch.tensor(rng.normal(0, 0.1, (n, m)), **dev)
    u = torch.tensor(rng.uniform(size=n), **dev)
    z0 = torch.tensor(rng.normal(size=(n, m)), **dev)
    z = torch.tensor(rng.normal(size=(n, m)), **dev)
    z_in = z.clone()
    acc = torch.empty(n, dtype=torch.int32, device="cuda")
    check(lib().fpx_hmc_metropolis(U0.data_ptr(), U1.data_ptr(), p0.data_ptr(), p1.data_ptr(),
                                   u.data_ptr(), z0.data_ptr(), n, m, z.data_ptr(),
                                   acc.data_ptr(), torch.cuda.current_stream().cuda_stream),
          "fpx_hmc_metropolis")
    H0 = U0 + 0.5 * (p0 * p0).sum(dim=1)
    H1 = U1 + 0.5 * (p1 * p1).sum(dim=1)
    want = (torch.log(u) < (H0 - H1)).cpu().numpy()
    got = acc.cpu().numpy().astype(bool)
    assert 0 < want.sum() < n and not want[3]
    np.testing.assert_array_equal(got, want)
    expect_z = torch.where(torch.tensor(want, device="cuda")[:, None], z_in, z0)
    np.testing.assert_array_equal(z.cpu().numpy(), expect_z.cpu().numpy())


def test_full_size_ensemble_every_slot_and_energy():
    """The HMC ensemble workload at its stated size: the benchmark's 64
    uneven patches (seed 0 + 2; 1-20 sources, 1-8 bands x 1-4 exposures,
    48-256 px stamps) x 64 chains = 4096 slots, per-chain step sizes found
    and adapted as the benchmark does, then one graph-captured 32-step
    leapfrog trajectory of every chain replayed once (littlemcmc's per-chain
    loop over FastPosterior.evaluate, forcepho/fitting.py:254-341 and
    forcepho/model.py:302-387, as one device-resident ensemble).

    Checks: (1) the trajectory's recorded end-point potential equals a fresh
    evaluation at the end positions, and the energy errors H1 - H0 are finite
    with a bounded distribution; (2) at the end positions every one of the
    4096 slots' batched chi2 and gradient equals that slot's records
    evaluated alone with the same tile shape (chi2 to 1e-7 relative,
    gradients to 1e-3 of the slot's largest entry: the bounds of the
    512-patch batch test); (3) on the five patches with the fewest pairs,
    the first and last chain against the C restatement (chi2) and the
    float64 oracle (gradients)."""
    import time
    import torch
    from forcepho_amd.ensemble import HMCEnsemble
    from forcepho_amd.kernel_limits import MAXSOURCES
    from forcepho_amd.patches import SimplePatch
    from forcepho_amd.proposal import HIPProposer
    from forcepho_amd.sources import source_struct_dtype
    t0 = time.time()
    rng = np.random.default_rng(2)
    pats, scs, bnds, thetas = [], [], [], []
    for _ in range(64):
        patch, scene, theta = scenes.build_patch(scenes.spec_uneven(rng), SimplePatch,
                                                 scenes.render_on_device)
        patch.return_residual = False
        patch.send_to_device()
        pats.append(patch)
        scs.append(scene)
        thetas.append(scene.get_all_source_params().copy())
        bnds.append(_bounds(thetas[-1], int(patch.n_bands)))
    ens = HMCEnsemble(pats, scs, 64, bnds, config=6, persistent=True)
    assert ens.n_slots == 4096
    start = []
    for pi, c in ens.batch.slot_map:
        t = thetas[pi].copy()
        nb = int(pats[pi].n_bands)
        t.reshape(-1, nb + 6)[:, nb + 2] *= 0.999 - 0.0001 * c    # near the truth
        start.append(t)
    ens.set_theta(start)
    gen = torch.Generator(device=ens.device)
    gen.manual_seed(0)
    print(f"[gpu] ensemble of {ens.n_slots} chains built ({time.time() - t0:.1f} s)", flush=True)
    # per-chain step sizes: the starting heuristic, two 8-step rounds, then
    # (with the recorded graph) two rounds of the measured 32 steps
    ens.eps.fill_(0.005)
    eps = ens.find_step_size(generator=gen).cpu().numpy()
    ens.adapt_step_size(2, 8, generator=gen)
    ens.capture(32)
    print(f"[gpu] trajectory recorded ({time.time() - t0:.1f} s)", flush=True)
    ens.adapt_step_size(2, 32, generator=gen, use_graph=True)
    print("[gpu] step sizes: heuristic quantiles (0, 50, 100%)",
          np.percentile(eps, [0, 50, 100]), "adapted",
          np.percentile(ens.eps.view(-1).cpu().numpy(), [0, 50, 100]),
          f"({time.time() - t0:.1f} s)", flush=True)
    ens.p.copy_(torch.randn(ens.p.shape, generator=gen, dtype=torch.float64,
                            device=ens.device) * ens.fmask)
    p0 = ens.p.clone()
    U0 = ens.potential()[0].clone()
    ens.graph.replay()
    torch.cuda.synchronize()
    print(f"[gpu] trajectory replayed ({time.time() - t0:.1f} s)", flush=True)
    U_start, U_end, p_end = ens.U_start.clone(), ens.U.clone(), ens.p.clone()
    np.testing.assert_allclose(U_start.cpu().numpy(), U0.cpu().numpy(), rtol=1e-10)
    U_fresh = ens.potential()[0]                       # batched launch at the end positions
    torch.cuda.synchronize()
    np.testing.assert_allclose(U_end.cpu().numpy(), U_fresh.cpu().numpy(), rtol=1e-10)
    H0 = (U0 + 0.5 * (p0 * p0).sum(dim=(1, 2))).cpu().numpy()
    H1 = (U_end + 0.5 * (p_end * p_end).sum(dim=(1, 2))).cpu().numpy()
    dH = H1 - H0
    q = [50, 90, 99, 100]
    print("[gpu] |H1 - H0| quantiles", q, np.percentile(np.abs(dH), q),
          "mean acceptance", float(np.mean(np.minimum(1.0, np.exp(-dH)))), flush=True)

    # every slot against its own launch with the same tile shape
    recs = ens.batch.proposals.cpu().numpy().view(source_struct_dtype).reshape(-1, MAXSOURCES)
    batched = ens.batch.results()
    plain = HIPProposer(config=ens.batch.config, persistent=ens.batch.persistent)
    chi2_dev, grad_dev = np.zeros(ens.n_slots), np.zeros(ens.n_slots)
    for s, (pi, _) in enumerate(ens.batch.slot_map):
        ns = int(pats[pi].n_sources)
        c, d = plain.evaluate_proposal(recs[s, :ns].copy(), patch=pats[pi])
        cb, db = batched[s]
        chi2_dev[s] = np.max(np.abs(cb - c) / np.abs(c))
        grad_dev[s] = np.max(np.abs(db - d)) / np.abs(d).max()
    print(f"[gpu] 4096 slots vs alone: chi2 {chi2_dev.max():.2e}, gradients {grad_dev.max():.2e} "
          f"({time.time() - t0:.1f} s)", flush=True)
    assert chi2_dev.max() <= 1e-7 and grad_dev.max() <= 1e-3

    # the smallest patches against the checkers
    sizes = np.array([p.n_evals() for p in pats])
    for pi in np.argsort(sizes)[:5]:
        for c in (0, 63):
            s = ens.batch.slot_map.index((int(pi), c))
            ns = int(pats[pi].n_sources)
            prop = recs[s, :ns].copy()
            c32, _, _ = oracle.evaluate_cpu(pats[pi], prop)
            assert np.max(np.abs(batched[s][0] - c32) / np.abs(c32)) <= U.CHI2_RTOL, (pi, c)
            _, g64, mass, smass = U.fp64_reference(pats[pi], prop)
            assert U.grad_errors(batched[s][1], g64, mass, smass).max() <= 1.0, (pi, c)
    print(f"[gpu] smallest patches vs checkers done ({time.time() - t0:.1f} s)", flush=True)
    # the energy error distribution of the trajectory near the truth with
    # adapted step sizes: almost every chain finite, most errors small
    finite = np.isfinite(dH)
    assert finite.mean() >= 0.99
    assert np.median(np.abs(dH[finite])) <= 1.0
    assert np.mean(np.where(finite, np.minimum(1.0, np.exp(-np.clip(dH, -50, None))), 0)) >= 0.6
