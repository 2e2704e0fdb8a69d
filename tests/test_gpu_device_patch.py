"""HIPPatchMixin semantics: data/residual swap (tests/cswap_test.py:79-90 of
the reference), fixed-source subtraction, design matrix and residual
unpacking, all through the HIP path."""

import copy

import numpy as np
import pytest

import oracle
from forcepho_amd import scenes

pytestmark = pytest.mark.gpu


def _patch(seed=31, return_residual=True):
    from forcepho_amd.patches import SimplePatch
    rng = np.random.default_rng(seed)
    spec = scenes.draw_patch_spec(rng, 2, [1, 2], [40, 36], 3, [3, 2])
    patch, scene, theta = scenes.build_patch(spec, SimplePatch, oracle.render_model,
                                             return_residual=return_residual)
    return patch, scene


def test_swap_on_device():
    patch, scene = _patch()
    patch.send_to_device()
    prop = patch.get_proposer()
    proposal = scene.get_proposal()
    chi, dchi, residual = prop.evaluate_proposal(proposal, patch=patch, unpack=False)
    data_original = patch.retrieve_array("data")
    patch.pack_meta(scene)
    patch.swap_on_device()
    data_post = patch.retrieve_array("data")
    resid_post = patch.retrieve_array("residual")
    np.testing.assert_array_equal(data_post, residual)
    np.testing.assert_array_equal(resid_post, data_original)
    chi, dchi, residual_new = prop.evaluate_proposal(proposal, patch=patch, unpack=False)
    model = data_original - residual
    scale = np.abs(data_original).max()
    assert np.abs(residual_new - (data_original - 2 * model)).max() <= 1e-5 * scale


def test_residual_unpacking_and_images():
    patch, scene = _patch(seed=32)
    patch.send_to_device()
    prop = patch.get_proposer()
    chi, dchi, resids = prop.evaluate_proposal(scene.get_proposal(), patch=patch, unpack=True)
    assert len(resids) == patch.n_exp
    assert [len(r) for r in resids] == list(patch.exposure_N)
    _, _, r_ref = oracle.evaluate_cpu(patch, scene.get_proposal(), return_residual=True)
    flat = np.concatenate(resids)
    assert np.abs(flat - r_ref).max() <= 2e-6 * np.abs(patch.data).max()
    imgs = prop.unpack_residuals(flat, patch, reshape=True)
    assert imgs[0].shape == tuple(patch.original_shapes[0])


def test_subtract_fixed_and_design_matrix():
    """Fixed sources are subtracted from the device data; the design matrix
    columns are unit-flux models of each active source."""
    patch, scene = _patch(seed=33)
    fixed = scene.subscene([0])
    active = scene.subscene([1, 2])
    data0 = patch.data.copy()
    proposer, sc = patch.subtract_fixed(fixed, active=active)
    assert sc is active
    # device data now = data - model(fixed); compare with the checker's model
    chk = copy.copy(patch)
    chk.pack_meta(fixed)
    chk.data = data0
    model_fixed = oracle.render_model(chk, fixed.get_proposal())
    got = patch.retrieve_array("data")
    assert np.abs(got - (data0 - model_fixed)).max() <= 2e-6 * np.abs(data0).max()
    patch.pack_meta(active)
    Xes, ys = patch.design_matrix(active, proposer=proposer)
    assert len(Xes) == patch.n_bands
    for b in range(patch.n_bands):
        assert Xes[b].shape == (2, patch.band_N_pix[b])
    # X . flux reproduces the active model
    chk2 = copy.copy(patch)
    chk2.pack_meta(active)
    m_active = oracle.render_model(chk2, active.get_proposal())
    split = np.split(m_active, np.cumsum(patch.band_N_pix[:-1]))
    for b in range(patch.n_bands):
        flux = np.array([s.flux[b] for s in active.sources])
        np.testing.assert_allclose(flux @ Xes[b], split[b], rtol=0,
                                   atol=2e-6 * np.abs(m_active).max())


def test_optimize_fluxes():
    """forcepho/fitting.py:589-620: after subtracting a fixed source the ML
    fluxes of the active sources solve the weighted normal equations of the
    kernel-rendered design matrix, and lie within 5 sigma of the truth."""
    from forcepho_amd.fitting import optimize_fluxes
    patch, scene = _patch(seed=34)
    fixed = scene.subscene([0])
    active = scene.subscene([1, 2])
    patch.subtract_fixed(fixed, active=active)
    fluxes, precisions, Xes, ys, ws = optimize_fluxes(patch, active, return_all=True)
    assert len(fluxes) == patch.n_bands
    for b in range(patch.n_bands):
        w = ws[b].astype(np.float64)
        ref = np.linalg.lstsq((Xes[b] * w).T, ys[b].astype(np.float64) * w, rcond=None)[0]
        np.testing.assert_allclose(fluxes[b], ref, rtol=1e-8)
        truth = np.array([s.flux[b] for s in active.sources])
        sigma = np.sqrt(np.diag(np.linalg.inv(precisions[b])))
        assert (np.abs(fluxes[b] - truth) < 5 * sigma).all(), (fluxes[b], truth, sigma)
