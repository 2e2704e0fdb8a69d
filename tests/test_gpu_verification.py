"""The reference's end-to-end verification and swap tests on the HIP kernel.

* tests/verification/verify_reference.py: the FP64 slow-model image of one
  Sersic galaxy, packed by FITSPatch, through ``prepare_model`` and
  ``FastPosterior``: residual within 5e-7 on every real pixel (:107),
  repeatable lnp and gradient, sum (r ierr)^2 within 2e-7 of zero and the
  chi2 identity (:131-143), finite-difference gradients at an offset point
  (:156-164).  The gradient at that point is also held to the FP64 oracle at
  the usual parity tolerance.
* tests/cswap_test.py:40-90: the reference's noisy test image with its HDF5
  PSF store: after ``swap_on_device`` the device data is the first residual
  and the device residual the original data; evaluating again gives
  data - 2 model.
"""

import numpy as np
import pytest

from fpx_testutil import (verification_patch, verification_amplitudes, fp64_reference,
                          TEST_IMAGE, TEST_IMAGE_PSF, GRAD_RTOL, GRAD_FLOOR, MODEL_EPS)

pytestmark = pytest.mark.gpu


def _model():
    from forcepho_amd.patches import FITSDevicePatch
    patch, catalog = verification_patch(FITSDevicePatch, return_residual=True)
    model, q = patch.prepare_model(active=catalog)
    return patch, model, q


def test_verification_residual():
    patch, model, q = _model()
    print("verification residual", flush=True)
    residuals = model.residuals(q, unpack=False)
    valid = patch.xpix >= 0
    assert np.abs(residuals[valid]).max() < 5e-7


def test_verification_chi2():
    patch, model, q = _model()
    valid = patch.xpix >= 0
    z = model.transform.inverse_transform(q)
    lnp, lnp_grad = model.lnprob_and_grad(z)
    model.evaluate(z)
    assert np.allclose(lnp, model._lnp)
    assert np.allclose(lnp_grad, model._lnp_grad)
    residual = model.residuals(q, unpack=False).astype(np.float64)
    ie = patch.ierr.astype(np.float64)
    chi2 = (residual * ie) ** 2
    assert np.allclose(0, chi2[valid].sum(), atol=2e-7)
    chi2 -= (patch.data * ie) ** 2
    assert np.allclose(chi2[valid].sum(), -2 * lnp, rtol=1e-7, atol=1e-7)


def test_verification_gradients():
    patch, model, q = _model()
    qoff = q.copy()
    qoff[3:5] = 0.8, 1.0      # q and pa away from the maximum
    zoff = model.transform.inverse_transform(qoff)
    delta = [1e-3, 1e-6, 1e-6, 1e-3, 1e-1, 1e-3, 1e-3]
    dlnp, dlnp_num = model.check_grad(zoff, delta=delta)
    assert np.allclose(dlnp, dlnp_num, rtol=1e-1)
    # the same point against the FP64 oracle at the parity tolerance
    model.scene.set_all_source_params(qoff)
    prop = model.scene.get_proposal()
    chi2, derivs = model.proposer.evaluate_proposal(prop, patch=patch)[:2]
    c64, g64, mass, smass = fp64_reference(patch, prop)
    assert np.allclose(chi2, c64, rtol=1e-6)
    bound = GRAD_RTOL * np.abs(g64) + GRAD_FLOOR * mass + MODEL_EPS * smass
    assert (np.abs(derivs - g64) <= bound).all(), np.abs(derivs - g64) / bound


def test_cswap_on_reference_test_image():
    """tests/cswap_test.py:40-90 with the reference's test image, its PSF
    store (HDF5) and the whole image as the region."""
    from forcepho_amd.formats import fits
    from forcepho_amd.patches import FITSDevicePatch
    cat = fits.getdata(TEST_IMAGE, -1)
    bands = fits.getheader(TEST_IMAGE, -1)["FILTERS"].split(",")
    patcher = FITSDevicePatch(fitsfiles=[TEST_IMAGE], psfstore=TEST_IMAGE_PSF,
                              splinedata=verification_amplitudes("rect"),
                              return_residual=True)
    patcher.build_patch(None, None, allbands=bands)
    scenes = [patcher.set_scene(cat) for _ in range(2)]
    proposer = patcher.get_proposer()
    patcher.return_residual = True
    patcher.pack_meta(scenes[0])
    patcher.send_to_device()
    print("cswap first evaluation", flush=True)
    chi, dchi_dp, residual = proposer.evaluate_proposal(scenes[0].get_proposal(),
                                                        patch=patcher, unpack=False)
    residual_original = residual
    data_original = patcher.data.copy()
    patcher.pack_meta(scenes[1])
    patcher.swap_on_device()
    data_postswap = patcher.retrieve_array("data")
    residual_postswap = patcher.retrieve_array("residual")
    assert data_postswap is not residual_original
    assert residual_postswap is not data_original
    assert np.allclose(residual_original, data_postswap)
    assert np.allclose(data_original, residual_postswap)
    chi, dchi_dp, residual_new = proposer.evaluate_proposal(scenes[1].get_proposal(),
                                                            patch=patcher, unpack=False)
    model = data_original - residual_original
    assert np.allclose(residual_new, data_original - 2 * model)
    # the catalog is the truth of the noisy image (noise from extension 1):
    # reduced chi2 of the first residual is 1 within 5 sigma (sqrt(2/N))
    valid = patcher.xpix >= 0
    n = valid.sum()
    red = np.sum((residual_original.astype(np.float64) * patcher.ierr)[valid] ** 2) / n
    assert abs(red - 1.0) < 5 * np.sqrt(2.0 / n)
    assert model[valid].sum() == pytest.approx(float(cat["CLEAR"][0]), abs=0.01)


def _bounds(theta, nb, q_hi=1.2):
    t = theta.reshape(-1, nb + 6)
    lo, hi = t.copy(), t.copy()
    lo[:, :nb], hi[:, :nb] = t[:, :nb] * 0.5, t[:, :nb] * 1.5 + 1e-3
    pix = 0.03 / 3600.0
    lo[:, nb:nb + 2] -= 3 * pix
    hi[:, nb:nb + 2] += 3 * pix
    lo[:, nb + 2], hi[:, nb + 2] = 0.3, q_hi
    lo[:, nb + 3], hi[:, nb + 3] = -2.0, 2.0
    lo[:, nb + 4], hi[:, nb + 4] = 0.8, 5.5
    lo[:, nb + 5], hi[:, nb + 5] = 0.02, 0.35
    return lo.reshape(-1), hi.reshape(-1)


def test_ensemble_on_real_data_with_spline_amplitudes():
    """HMCEnsemble builds Source records on the device from the reference's
    spline table (DeviceSplineAmplitudes) for two FITS patches; its potential
    and gradient equal the host FastPosterior, whose records come from scipy
    (Galaxy.proposal)."""
    from forcepho_amd.formats import fits
    from forcepho_amd.patches import FITSDevicePatch
    from forcepho_amd.ensemble import HMCEnsemble
    from forcepho_amd.model import FastPosterior
    from forcepho_amd.proposal import HIPProposer
    amps = verification_amplitudes("smooth")
    p1, cat1 = verification_patch(FITSDevicePatch, amplitudes=amps)
    p2 = FITSDevicePatch(fitsfiles=[TEST_IMAGE], psfstore=TEST_IMAGE_PSF, splinedata=amps)
    p2.build_patch(None, None, allbands=["CLEAR"])
    pats, scs, bnds, start = [], [], [], []
    for patch, cat in ((p1, cat1), (p2, fits.getdata(TEST_IMAGE, -1))):
        scene = patch.set_scene(cat)
        patch.pack_meta(scene)
        patch.send_to_device()
        theta = scene.get_all_source_params().copy()
        pats.append(patch)
        scs.append(scene)
        bnds.append(_bounds(theta, 1))
        start.append(theta)
    ens = HMCEnsemble(pats, scs, [2, 2], bnds)
    assert ens.records.spline is not None
    thetas = []
    for pi, c in ens.batch.slot_map:
        t = start[pi].copy()
        t[2] *= 0.95 - 0.05 * c          # q
        t[4] *= 1.0 + 0.1 * c            # sersic
        t[5] *= 1.1 - 0.2 * c            # rhalf
        thetas.append(t)
    ens.set_theta(thetas)
    print("ensemble potential", flush=True)
    Ud, gd = (x.cpu().numpy() for x in ens.potential())
    for s, (pi, _) in enumerate(ens.batch.slot_map):
        patch = pats[pi]
        post = FastPosterior(HIPProposer(patch=patch), patch, scene=scs[pi],
                             lower=bnds[pi][0], upper=bnds[pi][1], device_records=False)
        lnp, grad = post.lnprob_and_grad(post.transform.inverse_transform(thetas[s]))
        assert Ud[s] == pytest.approx(-lnp, rel=1e-6)
        g = np.concatenate([gd[s, :, :1], gd[s, :, ens.NB:]], axis=1)[:patch.n_sources].reshape(-1)
        np.testing.assert_allclose(g, -grad, rtol=1e-4, atol=1e-5 * np.abs(grad).max())


def test_autograd_records_from_spline_table():
    """PatchModel keeps real-data scenes on the device path: the records
    come from the device spline evaluation, and chi2 and its autograd
    gradient equal the host-record path."""
    import torch
    from forcepho_amd.patches import FITSDevicePatch
    from forcepho_amd.autograd import PatchModel
    patch, cat = verification_patch(FITSDevicePatch)
    scene = patch.set_scene(cat)
    patch.pack_meta(scene)
    model = PatchModel(patch, scene)
    assert model.records is not None and model.records.spline is not None
    theta = scene.get_all_source_params().copy()
    theta[2:6] = 0.85, 0.4, 2.6, 0.12
    out = []
    for device_path in (True, False):
        saved = model.records
        if not device_path:
            model.records = None
        t = torch.tensor(theta, dtype=torch.float64, device="cuda", requires_grad=True)
        c = model.chi2(t)
        c.backward()
        out.append((c.item(), t.grad.cpu().numpy()))
        model.records = saved
    np.testing.assert_allclose(out[0][0], out[1][0], rtol=1e-9)
    g = out[1][1]
    np.testing.assert_allclose(out[0][1], g, rtol=1e-4, atol=1e-5 * np.abs(g).max())
