"""Shared helpers of the test suite: golden fixtures and tolerances."""

import glob
import os

import numpy as np

from forcepho_amd.patches import PatchBase, HIPPatchMixin, PSF_DTYPE
from forcepho_amd.proposal import source_struct_dtype

GOLDEN_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

# tolerance of the parity bar
CHI2_RTOL = 1e-6     # chi2 relative to the reference CPU kernel
N_CONFIGS = 9        # kernel configurations (tile shapes) the library defines
GRAD_RTOL = 1e-4     # gradient relative error ...
# ... above a floor of this fraction of the summed magnitude sum|w dC/dtheta|.
# Near the MAX_EXP_ARG cutoff the float32 exponent e ~ 36 is the result of
# about three roundings (~6e-6 absolute), so exp(-e/2) carries ~3e-6
# relative error in any float32 evaluation, the reference's included; the
# float32 scale constant -log2(e)/2 adds ~7e-7.  Gradients that cancel (the
# sersic and rh terms sum mixture-amplitude derivatives, which add to zero
# over radii) sit on that floor.
GRAD_FLOOR = 4e-6
# ... and above the propagation of the float32 model's own rounding: the
# model (up to ~100 sigma per pixel) enters the residual (~1 sigma), so a
# relative model error eps moves the gradient by up to
# eps * sum 2 ierr^2 |m| |dC/dtheta|.  Both float32 implementations, the
# reference's CPU kernel and this kernel, show the same error on the entries
# this term governs.
MODEL_EPS = 1e-6


# The parity bar against the reference CPU kernel itself: |g - g_ref| at most
# this fraction of the largest |g_ref| of the same band (per-entry relative
# errors are unbounded where a gradient crosses zero; the reference's own
# float32 gradients sit within 1.5e-5 of each band's largest float64 entry on
# every fixture and 4.3e-5 on the full-size multi-band patch).
GRAD_BANDMAX_RTOL = 1e-4


def bandmax_rel(g, g_ref):
    """|g - g_ref| over the largest |g_ref| of the same band (axis 0)."""
    g = np.asarray(g, dtype=np.float64)
    g_ref = np.asarray(g_ref, dtype=np.float64)
    axes = tuple(range(1, g_ref.ndim))
    m = np.abs(g_ref).max(axis=axes, keepdims=True) if axes else np.abs(g_ref)
    return np.abs(g - g_ref) / np.where(m > 0, m, 1e-300)


def fp64_reference(patch, proposal):
    """float64 chi2, gradient, the noise-weighted magnitude sum|w dC/dtheta|
    and the signal-weighted magnitude sum 2 ierr^2 |d| |dC/dtheta|."""
    import oracle.fp64 as fp64
    c64, g64, _, mass = fp64.evaluate(patch, proposal, mass=True)
    sw = 2.0 * patch.ierr.astype(np.float64) ** 2 * np.abs(patch.data.astype(np.float64))
    smass = fp64.evaluate(patch, proposal, pixel_weight=sw, mass=True)[3]
    return c64, g64, mass, smass


class GoldenPatch(PatchBase, HIPPatchMixin):
    """A packed patch loaded from a fixture."""

    def __init__(self, npz, return_residual=False):
        super().__init__(return_residual=return_residual)
        for name in ["xpix", "ypix", "data", "ierr", "exposure_start", "exposure_N",
                     "band_start", "band_N", "rad2", "D", "crpix", "crval", "CW", "G",
                     "n_psf_per_source", "psfgauss_start", "original_sizes"]:
            setattr(self, name, np.array(npz[name]))
        self.psfgauss = np.array(npz["psfgauss"]).view(PSF_DTYPE).copy()
        self.n_sources = int(npz["n_sources"])
        self.n_radii = int(npz["n_radii"])
        self.n_bands = int(npz["n_bands"])
        self.n_exp = len(self.exposure_N)


def golden_names():
    return sorted(os.path.basename(p)[:-4] for p in glob.glob(os.path.join(GOLDEN_DIR, "*.npz")))


def load_golden(name, return_residual=False):
    """(patch, proposal, expected dict) of fixture ``name``."""
    npz = np.load(os.path.join(GOLDEN_DIR, name + ".npz"))
    patch = GoldenPatch(npz, return_residual=return_residual)
    proposal = np.array(npz["proposal"]).view(source_struct_dtype).copy()
    expected = {k: np.array(npz[k]) for k in ("chi2", "dchi2", "residual")}
    return patch, proposal, expected


def grad_errors(g, g64, mass, smass=None):
    """Per-entry error of a gradient against the float64 value, in units of
    its tolerance GRAD_RTOL |g64| + GRAD_FLOOR mass + MODEL_EPS smass
    (<= 1 passes)."""
    g = np.asarray(g, dtype=np.float64)
    tol = GRAD_RTOL * np.abs(g64) + GRAD_FLOOR * mass
    if smass is not None:
        tol = tol + MODEL_EPS * smass
    tol = np.where(tol > 0, tol, 1e-300)
    return np.abs(g - g64) / tol


# ---------------------------------------------------------------------------
# The reference's end-to-end verification case
# (tests/verification/verify_reference.py), fixtures written by
# tests/golden/make_verification.py.
VERIF_DIR = os.path.join(GOLDEN_DIR, "verification")
VERIF_IMAGE = os.path.join(VERIF_DIR, "reference_f200w_sersic2.2_rhalf0.10.fits")
TEST_IMAGE = os.path.join(VERIF_DIR, "test_image_data.fits")
TEST_IMAGE_PSF = os.path.join(VERIF_DIR, "test_image_psf.h5")


def verification_amplitudes(spline_type="smooth"):
    """Mixture amplitude model of the verification image: splines of the
    sersic_mog_model.smooth=0.0150 table.  The image was rendered with the
    SmoothBivariateSpline form (forcepho/sources.py:640-642)."""
    from forcepho_amd.sources import SplineMixtureAmplitudes
    t = np.load(os.path.join(VERIF_DIR, "sersic_mog_smooth0.0150.npz"))
    return SplineMixtureAmplitudes.from_table(t["nsersic"], t["rh"], t["amplitudes"],
                                              t["radii"], spline_type=spline_type)


def verification_psfstore(n_radii=9):
    """The F200W 4-Gaussian PSF the image header names, for every radius bin."""
    from forcepho_amd.patches import PSFStore
    t = np.load(os.path.join(VERIF_DIR, "psf_f200w_ng4m0.npz"))
    mix = np.zeros(len(t["amp"]), dtype=[(k, np.float64) for k in
                                         ("amp", "xcr", "ycr", "Cxx", "Cyy", "Cxy")])
    for k in mix.dtype.names:
        mix[k] = t[k]
    return PSFStore.from_mixtures({"F200W": mix}, n_radii=n_radii)


def verification_patch(cls, spline_type="smooth", pixel_order="superpixel", amplitudes=None,
                       **kwargs):
    """(patch, catalog) of the verification image packed by ``cls`` (a
    FITSPatch subclass): whole image, pixel inverse errors 1/(dFLUX/dF200W)
    from extension 1, as verify_reference.py:73-79 configures it."""
    from forcepho_amd.formats import fits
    amps = amplitudes if amplitudes is not None else verification_amplitudes(spline_type)
    patch = cls(fitsfiles=[VERIF_IMAGE], psfstore=verification_psfstore(len(amps.radii)),
                splinedata=amps, pixel_order=pixel_order, **kwargs)
    catalog = fits.getdata(VERIF_IMAGE, -1)
    bands = [b.upper() for b in fits.getheader(VERIF_IMAGE, -1)["FILTERS"].split(",")]
    patch.build_patch(region=None, allbands=bands)
    return patch, catalog
