"""The reference's end-to-end verification case on the CPU checkers.

forcepho's tests/verification/verify_reference.py packs the FP64 image of one
Sersic galaxy (rendered by the reference's slow model,
make_reference.py:34-47) with FITSPatch and requires the fast model's
residual to stay within 5e-7 on every real pixel (:107), sum (r ierr)^2 to
stay within 2e-7 of zero and the kernel's chi2 to equal
sum((r ierr)^2 - (d ierr)^2) (:137-143).  Here the patch is packed by this
package's FITSPatch from the committed fixtures and evaluated by the FP64
oracle, the C restatement and (where built) the reference's own CPU kernel.
These pin the FITS/WCS/PSF/spline ingestion and the packing end to end
against the reference's independent FP64 renderer; the GPU kernel is held to
the same bars in test_gpu_verification.py.

The image was rendered with SmoothBivariateSpline amplitudes
(forcepho/sources.py:640-642): they reproduce it to 4.0e-7 while the
RectBivariateSpline form of the same table (:633-638) misses by 2.1e-4.
"""

import numpy as np
import pytest

import oracle
import oracle.fp64 as fp64
from fpx_testutil import verification_amplitudes, verification_patch, VERIF_DIR
from forcepho_amd.patches import FITSPatch

RESID_TOL = 5e-7      # verify_reference.py:107
CHI2_ZERO_TOL = 2e-7  # verify_reference.py:138


def _packed(spline_type="smooth", pixel_order="superpixel"):
    patch, catalog = verification_patch(FITSPatch, spline_type=spline_type,
                                        pixel_order=pixel_order)
    scene = patch.set_scene(catalog)
    patch.pack_meta(scene)
    return patch, scene


def test_fits_patch_packing():
    patch, scene = _packed()
    assert patch.n_bands == 1 and patch.n_exp == 1 and patch.bandlist == ["F200W"]
    assert patch.npix == 2048 and patch.exposure_N[0] == 2048
    assert int(patch.original_sizes[0]) == 63 * 32
    valid = patch.xpix >= 0
    assert valid.sum() == 2016 and (patch.ierr[~valid] == 0).all() and (patch.ypix[~valid] == -1).all()
    # every pixel of the 63 x 32 image once
    key = patch.xpix[valid].astype(int) * 32 + patch.ypix[valid].astype(int)
    assert np.array_equal(np.sort(key), np.arange(2016))
    # first wavefront = one 8x8 square
    assert patch.xpix[:64].max() - patch.xpix[:64].min() == 7
    # ierr = 1 / (dFLUX/dF200W image) = 1 / image (flux 1)
    np.testing.assert_allclose(patch.ierr[valid] * patch.data[valid], 1.0, rtol=1e-6)
    # source at the image centre, 0.03"/px, 36 PSF Gaussians = 9 radii x 4
    np.testing.assert_allclose(patch.crpix[0, 0], [31.0, 15.5], atol=1e-4)
    np.testing.assert_allclose(patch.D[0, 0], np.eye(2) / 0.03, rtol=1e-6)
    assert patch.n_psf_per_source[0] == 36 and patch.n_radii == 9
    # entries grouped by centre: each PSF component's 9 radius bins together
    assert (patch.psfgauss["sersic_bin"] == np.tile(np.arange(9), 4)).all()
    assert (np.diff(patch.psfgauss["xcen"].reshape(4, 9), axis=1) == 0).all()
    assert scene.sources[0].sersic == 2.2 and scene.sources[0].rhalf == 0.1


def test_rect_spline_interpolates_table_nodes():
    """An interpolating (s=0) RectBivariateSpline returns the table at its
    nodes: pins the sersic-major reshape and the per-radius indexing."""
    t = np.load(VERIF_DIR + "/sersic_mog_smooth0.0150.npz")
    amps = verification_amplitudes("rect")
    for k in (0, 7, 23, 49):
        np.testing.assert_allclose(amps.amplitudes(t["nsersic"][k], t["rh"][k]),
                                   t["amplitudes"][k], atol=1e-12)


@pytest.mark.parametrize("order", ["superpixel", "flat"])
def test_fp64_oracle_reproduces_reference_image(order):
    patch, scene = _packed(pixel_order=order)
    valid = patch.xpix >= 0
    chi2, _, resid = fp64.evaluate(patch, scene.get_proposal())
    assert np.abs(resid[valid]).max() < RESID_TOL
    ie = patch.ierr.astype(np.float64)
    assert np.sum((resid * ie)[valid] ** 2) < CHI2_ZERO_TOL
    np.testing.assert_allclose(chi2[0], np.sum(((resid * ie) ** 2 - (patch.data * ie) ** 2)[valid]),
                               rtol=1e-12)


def test_rect_spline_does_not_reproduce_the_image():
    """Which spline form rendered the image: the grid form is off by 2e-4."""
    patch, scene = _packed(spline_type="rect")
    valid = patch.xpix >= 0
    _, _, resid = fp64.evaluate(patch, scene.get_proposal())
    assert np.abs(resid[valid]).max() > 1e-4


def _check_bars(patch, chi2, resid):
    valid = patch.xpix >= 0
    r = resid.astype(np.float64)
    assert np.abs(r[valid]).max() < RESID_TOL
    ie = patch.ierr.astype(np.float64)
    assert np.sum((r * ie)[valid] ** 2) < CHI2_ZERO_TOL
    host = np.sum(((r * ie) ** 2 - (patch.data * ie) ** 2)[valid])
    assert np.isclose(chi2[0], host, rtol=1e-7, atol=1e-7)


def test_c_restatement_passes_verification():
    patch, scene = _packed()
    chi2, _, resid = oracle.evaluate_cpu(patch, scene.get_proposal(), return_residual=True)
    _check_bars(patch, chi2, resid)


def test_reference_cpu_kernel_passes_verification():
    if oracle.reference_module() is None:
        pytest.skip("reference CPU kernel not built here")
    patch, scene = _packed()
    prop = scene.get_proposal()
    chi2, derivs, resid = oracle.evaluate_reference(patch, prop, return_residual=True)
    _check_bars(patch, chi2, resid)
    c2, d2, r2 = oracle.evaluate_cpu(patch, prop, return_residual=True)
    assert np.array_equal(chi2, c2) and np.array_equal(derivs, d2) and np.array_equal(resid, r2)
