"""forcepho's Posterior API (FastPosterior.lnprob_and_grad, model.py:302-340)
with the Source records built on the device (fpx_build_sources) against the
same posterior with the records built on the host by Scene.get_proposal:
the log-normal amplitudes of the synthetic multi-band patch and the
reference's spline table on its verification image."""

import numpy as np
import pytest

import fpx_testutil as U
from forcepho_amd import scenes

pytestmark = pytest.mark.gpu


def _bounds(theta, nb):
    t = theta.reshape(-1, nb + 6)
    lo, hi = t.copy(), t.copy()
    lo[:, :nb], hi[:, :nb] = t[:, :nb] * 0.5, t[:, :nb] * 1.5 + 1e-3
    pix = 0.03 / 3600.0
    lo[:, nb:nb + 2] -= 3 * pix
    hi[:, nb:nb + 2] += 3 * pix
    lo[:, nb + 2], hi[:, nb + 2] = 0.2, 1.05
    lo[:, nb + 3], hi[:, nb + 3] = -2.0, 2.0
    lo[:, nb + 4], hi[:, nb + 4] = 0.8, 5.5
    lo[:, nb + 5], hi[:, nb + 5] = 0.03, 0.4
    return lo.reshape(-1), hi.reshape(-1)


def _compare(patch, scene, theta, nb, seed):
    from forcepho_amd.model import FastPosterior, BoundedTransform
    from forcepho_amd.proposal import HIPProposer
    lo, hi = _bounds(theta, nb)
    tr = BoundedTransform(lo, hi)
    dev = FastPosterior(HIPProposer(patch=patch), patch, scene=scene, transform=tr)
    host = FastPosterior(HIPProposer(patch=patch), patch, scene=scene, transform=tr,
                         device_records=False)
    rng = np.random.default_rng(seed)
    z0 = tr.inverse_transform(theta)
    for k in range(3):
        z = z0 + 0.05 * rng.normal(size=z0.shape)
        lp_d, g_d = dev.lnprob_and_grad(z)
        lp_h, g_h = host.lnprob_and_grad(z)
        assert dev._dev is not None and host._dev is None
        np.testing.assert_allclose(lp_d, lp_h, rtol=1e-6)
        np.testing.assert_allclose(g_d, g_h, rtol=1e-4, atol=1e-5 * np.abs(g_h).max())


def test_device_records_posterior_lognormal():
    from forcepho_amd.patches import SimplePatch
    spec = scenes.spec_multi(3, side=48, n_bands=4, n_sources=6)
    patch, scene, theta = scenes.build_patch(spec, SimplePatch, scenes.render_on_device)
    patch.send_to_device()
    _compare(patch, scene, theta, 4, 1)


def test_device_records_posterior_spline_table():
    from forcepho_amd.patches import FITSDevicePatch
    patch, cat = U.verification_patch(FITSDevicePatch, spline_type="smooth")
    scene = patch.set_scene(cat)
    patch.pack_meta(scene)
    patch.send_to_device()
    theta = scene.get_all_source_params().copy()
    theta[1 + 2] = 0.9          # q inside the box
    _compare(patch, scene, theta, 1, 2)
