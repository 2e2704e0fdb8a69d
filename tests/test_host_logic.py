"""Host logic around the kernel (CPU only): parameter vectors, transforms,
amplitude derivatives, stack_grad and the Posterior plumbing (driven by a
CPU stand-in proposer built on the checker, test-only)."""


import numpy as np
import pytest

import oracle
from forcepho_amd import scenes
from forcepho_amd.model import (BoundedTransform, FastPosterior, Transform, stack_grad,
                                sigmoid, logit)
from forcepho_amd.patches import StampPatch
from forcepho_amd.sources import LogNormalMixtureAmplitudes, SplineMixtureAmplitudes


def test_stack_grad():
    rng = np.random.default_rng(0)
    d = rng.normal(size=(3, 2, 7))
    g = stack_grad(d, 2, 3).reshape(2, 9)
    for s in range(2):
        np.testing.assert_allclose(g[s, :3], d[:, s, 0])
        np.testing.assert_allclose(g[s, 3:], d[:, s, 1:].sum(axis=0))


def test_bounded_transform():
    lo, hi = np.array([0.0, -2.0, 1.0]), np.array([1.0, 2.0, 5.0])
    t = BoundedTransform(lo, hi)
    z = np.array([0.3, -1.2, 2.0])
    q = t.transform(z)
    assert np.all((q > lo) & (q < hi))
    np.testing.assert_allclose(t.inverse_transform(q), z, rtol=1e-12)
    h = 1e-6
    fd = (t.transform(z + h) - t.transform(z - h)) / (2 * h)
    np.testing.assert_allclose(t.jacobian(z), fd, rtol=1e-7)
    fdl = np.array([(t.lndetjac(z + h * e) - t.lndetjac(z - h * e)) / (2 * h) for e in np.eye(3)])
    np.testing.assert_allclose(t.lndetjac_grad(z), fdl, rtol=1e-6)
    assert sigmoid(logit(0.25)) == pytest.approx(0.25)


def test_amplitude_model_derivatives():
    m = LogNormalMixtureAmplitudes()
    for n, rh in [(1.2, 0.06), (3.3, 0.15), (4.9, 0.29)]:
        A = m.amplitudes(n, rh)
        assert A.sum() == pytest.approx(1.0) and np.all(A >= 0)
        dn, dr = m.derivatives(n, rh)
        h = 1e-6
        np.testing.assert_allclose(dn, (m.amplitudes(n + h, rh) - m.amplitudes(n - h, rh)) / (2 * h),
                                   rtol=1e-5, atol=1e-9)
        np.testing.assert_allclose(dr, (m.amplitudes(n, rh + h) - m.amplitudes(n, rh - h)) / (2 * h),
                                   rtol=1e-5, atol=1e-7)


def test_spline_amplitudes_reproduce_table():
    """On grid nodes a cubic RectBivariateSpline with s=0 interpolates the
    table; its derivative matches central differences of itself."""
    m = LogNormalMixtureAmplitudes()
    ng, rg = np.linspace(1, 5, 9), np.linspace(0.04, 0.32, 15)
    table = np.array([[m.amplitudes(n, r) for r in rg] for n in ng])
    sp = SplineMixtureAmplitudes(ng, rg, table, m.radii, smoothing=0)
    np.testing.assert_allclose(sp.amplitudes(ng[3], rg[5]), table[3, 5], atol=1e-12)
    dn, dr = sp.derivatives(2.2, 0.1)
    h = 1e-5
    np.testing.assert_allclose(dr, (sp.amplitudes(2.2, 0.1 + h) - sp.amplitudes(2.2, 0.1 - h)) / (2 * h),
                               rtol=1e-4, atol=1e-8)


def test_scene_parameter_round_trip():
    spec = scenes.spec_multi(4, side=32, n_bands=3, n_sources=4)
    sc = spec.scene
    theta = sc.get_all_source_params()
    assert len(theta) == 4 * (3 + 6)
    theta2 = theta + 0.01
    sc.set_all_source_params(theta2)
    np.testing.assert_allclose(sc.get_all_source_params(), theta2)
    prop = sc.get_proposal()
    assert prop.dtype.itemsize == 264 and len(prop) == 4
    np.testing.assert_allclose(prop["fluxes"][:, :3], theta2.reshape(4, 9)[:, :3], rtol=1e-6)
    assert np.all(prop["fluxes"][:, 3:] == 0)


class CheckerProposer:
    """Test-only stand-in for the device proposer: the C restatement."""

    def evaluate_proposal(self, proposal, patch=None, **kw):
        chi2, d, _ = oracle.evaluate_cpu(patch, proposal)
        return chi2, d.astype(np.float64)


def test_posterior_plumbing_matches_finite_differences():
    rng = np.random.default_rng(12)
    spec = scenes.draw_patch_spec(rng, 2, [1, 1], [32, 32], 2, [2, 2])
    patch, scene, theta = scenes.build_patch(spec, StampPatch, oracle.render_model)
    lo = theta - np.abs(theta) * 0.5 - 1e-7
    hi = theta + np.abs(theta) * 0.5 + 1e-7
    model = FastPosterior(CheckerProposer(), patch, scene=scene, lower=lo, upper=hi)
    z = model.transform.inverse_transform(theta * 1.02 + 1e-8)
    lnp, grad = model.lnprob_and_grad(z)
    assert model.ncall == 1
    lnp2, grad2 = model.lnprob_and_grad(z)       # cached
    assert model.ncall == 1 and lnp2 == lnp
    dlnp, dnum = model.check_grad(z, delta=1e-3)
    np.testing.assert_allclose(dlnp, dnum, rtol=3e-2, atol=1e-3 * np.abs(dlnp).max())
    nll, nllg = model.nll(z)
    assert nll == -lnp and np.array_equal(nllg, -grad)


def test_posterior_needs_transform_or_bounds():
    with pytest.raises(ValueError):
        FastPosterior(CheckerProposer(), patch=None, scene=object())
    FastPosterior(CheckerProposer(), patch=None, scene=object(), transform=Transform(3))


def test_optimize_one_band_matches_lstsq():
    """The weighted normal equations give the least-squares solution of
    (X w)^T f = y w, and recover exact fluxes from noise-free data."""
    from forcepho_amd.fitting import optimize_one_band
    rng = np.random.default_rng(5)
    X = rng.random((3, 200))
    w = 0.5 + rng.random(200)
    f_true = np.array([2.0, -1.0, 0.5])
    y = f_true @ X + 0.01 * rng.normal(size=200) / w
    flux, prec = optimize_one_band(X, w, y)
    ref = np.linalg.lstsq((X * w).T, y * w, rcond=None)[0]
    np.testing.assert_allclose(flux[:, 0], ref, rtol=1e-10)
    np.testing.assert_allclose(prec, (X * w) @ (X * w).T, rtol=1e-12)
    flux0, _ = optimize_one_band(X, w, f_true @ X)
    np.testing.assert_allclose(flux0[:, 0], f_true, rtol=1e-10)
    # a fixed-source model is subtracted first
    fixed = rng.random(200)
    flux1, _ = optimize_one_band(X, w, f_true @ X + fixed, fixedX=fixed)
    np.testing.assert_allclose(flux1[:, 0], f_true, rtol=1e-10)


def test_evaluator_release_waits_for_capture_to_end(monkeypatch):
    """A bound evaluator released while a graph is being captured (a patch
    collected mid-capture) is kept and released with the next release outside
    the capture, since fpx_release synchronises the device."""
    from forcepho_amd.patches import device_patch as dp
    released = []

    class FakeLib:
        def fpx_release(self, ev):
            released.append(ev)

    monkeypatch.setattr(dp, "lib", lambda: FakeLib())
    monkeypatch.setattr(dp, "_DEFERRED", [])
    monkeypatch.setattr(dp, "_capturing", lambda: True)
    dp._release_handles([1, 2])
    assert released == [] and dp._DEFERRED == [1, 2]
    monkeypatch.setattr(dp, "_capturing", lambda: False)
    dp._release_handles([3])
    assert sorted(released) == [1, 2, 3] and dp._DEFERRED == []
