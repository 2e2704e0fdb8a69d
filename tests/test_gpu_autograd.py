"""torch.autograd over the patch model: the residual image's backward (the
kernel's gradient pass with dL/dresidual as pixel weights) and the fused chi2
(value and gradient of one chi2-mode launch) against each other, the float64
oracle, the reference's own CPU kernel and central differences; a 32-step
leapfrog driven by them."""

import numpy as np
import pytest

import fpx_testutil as U
import oracle
from forcepho_amd import scenes
from forcepho_amd.model import stack_grad

pytestmark = pytest.mark.gpu


def _model(side, seed=0, n_bands=8, n_sources=15):
    from forcepho_amd.patches import SimplePatch
    from forcepho_amd.autograd import PatchModel
    spec = scenes.spec_multi(seed, side=side, n_bands=n_bands, n_sources=n_sources)
    render = oracle.render_model if side <= 48 else scenes.render_on_device
    patch, scene, theta = scenes.build_patch(spec, SimplePatch, render, return_residual=True)
    patch.send_to_device()
    return PatchModel(patch, scene), theta


def _offset(theta, n_bands):
    t = theta.copy().reshape(-1, n_bands + 6)
    t[:, n_bands + 2] *= 0.95      # q
    t[:, n_bands + 3] += 0.05      # pa
    t[:, :n_bands] *= 1.03         # fluxes
    return t.reshape(-1)


def test_autograd_matches_kernel_gradient_and_fp64():
    import torch
    model, theta = _model(side=32)
    th = _offset(theta, 8)
    t = torch.tensor(th, dtype=torch.float64, requires_grad=True)
    chi2 = model.chi2_from_residual(t)
    chi2.backward()
    g_auto = t.grad.numpy()
    t2 = torch.tensor(th, dtype=torch.float64, requires_grad=True)
    fused = model.chi2(t2)
    (3.0 * fused).backward()
    g_fused = t2.grad.numpy() / 3.0
    # the chi2-mode kernel gradient of the same proposal
    _, c_band, d = model.forward_residual(model.proposal(th))
    g_kernel = stack_grad(d, model.n_sources, model.n_bands)
    np.testing.assert_allclose(chi2.item(), c_band.sum(), rtol=1e-6)
    np.testing.assert_allclose(fused.item(), c_band.sum(), rtol=1e-12)
    np.testing.assert_allclose(g_fused, g_kernel, rtol=1e-12, atol=0)
    # float64 reference of both
    c64, g64, mass, smass = U.fp64_reference(model.patch, model.proposal(th))
    g64s = stack_grad(g64, model.n_sources, model.n_bands)
    ms = stack_grad(mass, model.n_sources, model.n_bands)
    ss = stack_grad(smass, model.n_sources, model.n_bands)
    tol = U.GRAD_RTOL * np.abs(g64s) + U.GRAD_FLOOR * ms + U.MODEL_EPS * ss
    assert np.all(np.abs(g_auto - g64s) <= tol)
    assert np.all(np.abs(g_kernel - g64s) <= tol)
    assert np.all(np.abs(g_fused - g64s) <= tol)
    # the reference's own CPU kernel on the same proposal (where it is built)
    mod = oracle.reference_module()
    if mod is not None:
        c_ref, d_ref, _ = oracle.evaluate_reference(model.patch, model.proposal(th), module=mod)
        g_ref = stack_grad(d_ref.astype(np.float64), model.n_sources, model.n_bands)
        np.testing.assert_allclose(fused.item(), c_ref.sum(), rtol=1e-6)
        assert np.all(np.abs(g_ref - g64s) <= tol)
        # directly against the reference: every entry within its float64
        # tolerance, and within 1e-4 of the largest |gradient| (one scale for
        # the stacked vector, whose band fluxes and shared shape parameters
        # are all in it)
        for g in (g_auto, g_fused):
            assert np.all(np.abs(g - g_ref) <= tol)
            assert np.abs(g - g_ref).max() <= U.GRAD_BANDMAX_RTOL * np.abs(g_ref).max()


def test_autograd_central_differences_full_size():
    """The 8-band x 2-exposure x 128x128, 15-source patch: autograd gradient
    against central differences of the kernel's own chi2 (step 1e-4 in
    unit-scaled parameters).  The differences carry the float32 rounding of
    the residual image: each pixel's chi2 term (r^2 - d^2) ierr^2 moves by
    about 2 |r ierr| |m ierr| 2^-24, so the chi2 difference has a rounding
    floor of a few times sqrt(sum of squares) of that, which enters the
    tolerance next to a 2% relative term."""
    import torch
    model, theta = _model(side=128)
    nb = 8
    th = _offset(theta, nb)
    t = torch.tensor(th, dtype=torch.float64, requires_grad=True)
    r = model.residual(t)
    chi2 = ((r.double() * model.ierr) ** 2 - (model.data * model.ierr) ** 2).sum()
    chi2.backward()
    g = t.grad.numpy()
    t2 = torch.tensor(th, dtype=torch.float64, requires_grad=True)
    model.chi2(t2).backward()
    g_fused = t2.grad.numpy()
    rr = r.detach().double() * model.ierr
    mm = (model.data - r.detach().double()) * model.ierr
    noise = 4.0 * 2.0 ** -24 * float(torch.sqrt(((2 * rr.abs() * mm.abs()) ** 2).sum()))
    pix_deg = 0.03 / 3600.0
    units = np.tile(np.concatenate([np.abs(th[:nb]).clip(1e-3), [pix_deg, pix_deg, 1, 1, 1, 0.1]]),
                    model.n_sources)
    h = 1e-4
    rng = np.random.default_rng(0)
    idx = rng.choice(len(th), size=24, replace=False)
    for i in idx:
        tp, tm = th.copy(), th.copy()
        tp[i] += h * units[i]
        tm[i] -= h * units[i]
        with torch.no_grad():
            fd = (model.chi2(torch.tensor(tp)).item() - model.chi2(torch.tensor(tm)).item()) / (2 * h * units[i])
        tol = 2e-2 * abs(g[i]) + 2 * noise / (2 * h * units[i])
        assert abs(fd - g[i]) <= tol, (i, fd, g[i], tol)
        assert abs(fd - g_fused[i]) <= tol, (i, fd, g_fused[i], tol)


def test_leapfrog_trajectory_conserves_energy():
    """32 leapfrog steps of HMC on -lnL = chi2/2 in unit-scaled parameters,
    gradients from autograd: the Hamiltonian stays put and the trajectory
    reverses."""
    import torch
    from forcepho_amd.hmc import leapfrog, hamiltonian
    model, theta = _model(side=48, n_bands=4, n_sources=6)
    nb = 4
    pix_deg = 0.03 / 3600.0
    units = np.tile(np.concatenate([np.abs(theta[:nb]).clip(1e-3) * 0.01,
                                    [0.1 * pix_deg, 0.1 * pix_deg, 0.01, 0.01, 0.01, 0.001]]),
                    model.n_sources)
    th0 = _offset(theta, nb)

    def U_and_grad(q):
        t = torch.tensor(q * units, dtype=torch.float64, requires_grad=True)
        u = 0.5 * model.chi2(t)
        u.backward()
        return u.item(), t.grad.numpy() * units

    rng = np.random.default_rng(2)
    q0 = th0 / units
    p0 = rng.normal(size=len(q0))
    U0, _ = U_and_grad(q0)
    q1, p1, U1, calls = leapfrog(q0, p0, U_and_grad, eps=0.02, n_steps=32)
    assert calls == 33
    H0, H1 = hamiltonian(U0, p0), hamiltonian(U1, p1)
    assert abs(H1 - H0) < 0.05 * max(1.0, abs(np.sum(p0 ** 2)))
    q2, p2, U2, _ = leapfrog(q1, -p1, U_and_grad, eps=0.02, n_steps=32)
    np.testing.assert_allclose(q2, q0, atol=1e-3)


def test_device_records_match_host_proposal():
    """The device-built Source records (DeviceRecords) equal Galaxy.proposal's
    float32 fields, and autograd through them matches the host-record path."""
    import torch
    from forcepho_amd.sources import source_struct_dtype
    model, theta = _model(side=32, n_bands=3, n_sources=4)
    assert model.records is not None
    th = _offset(theta, 3)
    host = model.proposal(th)
    dev = model.device_proposal(torch.tensor(th, dtype=torch.float64))
    rec = dev.cpu().numpy().view(source_struct_dtype)[:model.n_sources]
    for f in ("ra", "dec", "q", "pa", "nsersic", "rh"):
        np.testing.assert_array_equal(rec[f], host[f])
    np.testing.assert_array_equal(rec["fluxes"], host["fluxes"])
    for f in ("mixture_amplitudes", "damplitude_drh", "damplitude_dnsersic"):
        np.testing.assert_allclose(rec[f], host[f], rtol=2e-6, atol=1e-7 * np.abs(host[f]).max())
    grads = []
    for device_path in (True, False):
        saved = model.records
        if not device_path:
            model.records = None
        t = torch.tensor(th, dtype=torch.float64, device="cuda", requires_grad=True)
        c = model.chi2(t)
        c.backward()
        grads.append((c.item(), t.grad.cpu().numpy()))
        model.records = saved
    np.testing.assert_allclose(grads[0][0], grads[1][0], rtol=1e-9)
    g = grads[1][1]
    np.testing.assert_allclose(grads[0][1], g, rtol=1e-4, atol=1e-5 * np.abs(g).max())


def test_graph_replays_match_eager():
    """The recorded graphs (record build, launch, gradient stacking) of the
    chi2, the residual image and the backward pass, replayed for several
    parameter vectors with eager launches of the same proposer between the
    replays, give the eager results of each; a change of kernel
    configuration records the graphs again."""
    import torch
    model, theta = _model(side=48, n_bands=4, n_sources=6)
    nb = 4
    rng = np.random.default_rng(5)
    thetas = []
    for k in range(4):
        t = _offset(theta, nb).reshape(-1, nb + 6)
        t[:, :nb] *= 1 + 0.05 * rng.normal(size=(len(t), nb))
        t[:, nb + 2] = np.clip(t[:, nb + 2] * (1 + 0.05 * rng.normal(size=len(t))), 0.2, 0.99)
        thetas.append(t.reshape(-1))
    weight = torch.tensor(rng.normal(size=model.patch.npix), dtype=torch.float32, device="cuda")

    def run(th):
        t = torch.tensor(th, dtype=torch.float64, device="cuda")
        v, g = model.chi2_and_grad(t)
        r = model.residual_device(t)
        w = model.vjp_theta(t, weight * (1.0 + float(th[0])))
        return v.item(), g.cpu().numpy(), r.cpu().numpy(), w.cpu().numpy()

    def same(a, b):
        np.testing.assert_allclose(a[0], b[0], rtol=1e-12)
        for x, y in ((a[1], b[1]), (a[3], b[3])):
            np.testing.assert_allclose(x, y, rtol=1e-6, atol=1e-6 * np.abs(y).max())
        np.testing.assert_array_equal(a[2], b[2])

    model.use_graph = False
    eager = [run(th) for th in thetas]
    model.use_graph = True
    for _ in range(2):
        for th, ref in zip(thetas, eager):
            got = run(th)
            model.forward_residual(model.proposal(thetas[0]))     # eager work between replays
            same(got, ref)
    assert set(model._graphs) == {"_chi2_body", "_residual_body", "_vjp_body"}
    graph = model._graphs["_chi2_body"]["graph"]
    cfg = model.proposer.config
    model.proposer.set_config(2 if cfg != 2 else 1)
    got = run(thetas[1])
    assert model._graphs["_chi2_body"]["graph"] is not graph
    model.use_graph = False
    same(got, run(thetas[1]))
    model.proposer.set_config(cfg)


def test_model_only_launch():
    """A model-only launch (FPX_SLOT_MODEL_ONLY) writes the residual image
    of the full chi2 launch bit for bit and its chi2, with a zero gradient."""
    model, theta = _model(side=48, n_bands=4, n_sources=6)
    p, patch = model.proposer, model.patch
    prop = model.proposal(_offset(theta, 4))
    resid_full, chi2_full, resp_full = model.forward_residual(prop)
    assert np.abs(resp_full).max() > 0
    resid_model = model.forward_residual(prop, outputs=False)
    np.testing.assert_array_equal(resid_model.cpu().numpy(), resid_full.cpu().numpy())
    np.testing.assert_allclose(p._chi2[:4].cpu().numpy(), chi2_full, rtol=1e-12)
    assert not p._resp[:4 * 30 * 7].cpu().numpy().any()
    with pytest.raises(ValueError):
        p.launch(prop, patch, pixel_weight=resid_model, model_only=True)


def test_graphs_follow_swap_on_device():
    """After swap_on_device (the device data becomes the last residual) the
    recorded residual graph is recorded again: its output equals the eager
    launch bit for bit and is the residual of the swapped data, 2 r - d."""
    import torch
    model, theta = _model(side=32, n_bands=2, n_sources=3)
    t = torch.tensor(_offset(theta, 2), dtype=torch.float64, device="cuda")
    r0 = model.residual_device(t).cpu().numpy().astype(np.float64)
    model.patch.swap_on_device()
    r1 = model.residual_device(t).cpu().numpy()
    model.use_graph = False
    r1e = model.residual_device(t).cpu().numpy()
    np.testing.assert_array_equal(r1, r1e)
    d = model.patch.data.astype(np.float64)
    valid = model.patch.ierr > 0
    scale = np.abs(d).max()
    assert np.abs(r1[valid] - (2 * r0[valid] - d[valid])).max() <= 1e-5 * scale
