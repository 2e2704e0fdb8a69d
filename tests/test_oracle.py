"""Pins of the checkers themselves (CPU only).

* the C restatement reproduces every golden fixture (written by the
  reference's compiled CPU kernel) bit for bit, and fresh scenes too where
  the reference build is present;
* the float64 oracle agrees with the fixtures to float32 rounding;
* closed forms and finite differences fix what the fixtures leave open:
  a lone circular Gaussian evaluated analytically, flux normalisation, the
  chi2 convention, zero gradient at the noise-free truth, central
  differences of chi2 for every parameter, rotation invariance of a round
  source and source-permutation equivariance.
"""

import copy

import numpy as np
import pytest

import fpx_testutil as U
import oracle
import oracle.fp64 as fp64
from forcepho_amd import scenes
from forcepho_amd.patches import StampPatch, Stamp, LinearWCS, PSF_DTYPE
from forcepho_amd.sources import Galaxy, Scene, LogNormalMixtureAmplitudes

NAMES = U.golden_names()


def test_fixtures_present():
    assert len(NAMES) >= 10


@pytest.mark.parametrize("name", NAMES)
def test_c_restatement_bit_exact_on_fixture(name):
    patch, proposal, exp = U.load_golden(name)
    chi2, dchi2, resid = oracle.evaluate_cpu(patch, proposal, return_residual=True)
    assert np.array_equal(chi2, exp["chi2"])
    assert np.array_equal(dchi2, exp["dchi2"])
    assert np.array_equal(resid, exp["residual"])


@pytest.mark.parametrize("seed", [21, 22])
def test_c_restatement_bit_exact_vs_reference_build(seed):
    mod = oracle.reference_module()
    if mod is None:
        pytest.skip("reference CPU kernel not built here (oracle/_ref)")
    rng = np.random.default_rng(seed)
    spec = scenes.draw_patch_spec(rng, 2, [2, 1], [40, 36], 5, [3, 4])
    patch, scene, _ = scenes.build_patch(spec, StampPatch, oracle.render_model)
    prop = scene.get_proposal()
    prop["pa"] += np.float32(0.3)
    a = oracle.evaluate_cpu(patch, prop, return_residual=True)
    b = oracle.evaluate_reference(patch, prop, return_residual=True, module=mod)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


@pytest.mark.parametrize("name", NAMES)
def test_fp64_oracle_vs_fixture(name):
    patch, proposal, exp = U.load_golden(name)
    c64, g64, mass, smass = U.fp64_reference(patch, proposal)
    r64 = fp64.evaluate(patch, proposal)[2]
    rel = np.abs(c64 - exp["chi2"]) / np.abs(exp["chi2"]).clip(1e-300)
    assert rel.max() < 1e-7
    # the reference's float32 gradients lie inside the float32 tolerance band
    assert U.grad_errors(exp["dchi2"], g64, mass, smass).max() <= 1.0
    dscale = np.abs(patch.data).max()
    assert np.abs(r64 - exp["residual"]).max() < 2e-6 * dscale


# ---------------------------------------------------------------------------
# closed forms
# ---------------------------------------------------------------------------
def _lone_gaussian_patch(n=48, radius=0.06, sig_psf=1.3, flux=7.0, G=1.7, q=1.0, pa=0.0,
                         pix_scale=0.03, offset=(0.3, -0.2)):
    """One source with a single mixture radius and one circular PSF Gaussian."""
    ra0, dec0 = scenes.SKY_CENTER
    s_deg = pix_scale / 3600.0
    wcs = LinearWCS((ra0, dec0), ((n - 1) / 2.0, (n - 1) / 2.0), np.eye(2) * s_deg)
    psf = np.zeros(1, dtype=PSF_DTYPE)
    psf["amp"], psf["Cxx"], psf["Cyy"], psf["sersic_bin"] = 1.0, sig_psf ** 2, sig_psf ** 2, 0
    st = Stamp(0, np.zeros((n, n)), np.ones((n, n)), wcs, psf, photocounts=G)
    amp_model = LogNormalMixtureAmplitudes(np.array([radius]))
    g = Galaxy(filters=["b0"], amplitude_model=amp_model)
    g.flux = np.array([flux])
    g.q, g.pa, g.sersic, g.rhalf = q, pa, 2.0, 0.1
    cosd = np.cos(np.deg2rad(dec0))
    g.ra = ra0 + offset[0] * s_deg / cosd
    g.dec = dec0 + offset[1] * s_deg
    scene = Scene([g])
    patch = StampPatch(pixel_order="flat")
    patch.build_patch([st], scene, sky_zero=np.array([ra0, dec0]))
    return patch, scene


def test_lone_gaussian_closed_form():
    n, radius, sig_psf, flux, G = 48, 0.06, 1.3, 7.0, 1.7
    patch, scene = _lone_gaussian_patch(n, radius, sig_psf, flux, G)
    prop = scene.get_proposal()
    assert prop["mixture_amplitudes"][0, 0] == pytest.approx(1.0)
    # closed form from the packed (float32) inputs promoted to float64
    f64 = np.float64
    D = patch.D[0, 0].astype(f64)
    assert D[0, 1] == 0 and D[1, 0] == 0 and D[0, 0] == D[1, 1]
    s2 = f64(patch.rad2[0]) * D[0, 0] ** 2 + f64(patch.psfgauss["Cxx"][0])
    sky = np.array([prop["ra"][0], prop["dec"][0]], dtype=f64) - patch.crval[0, 0].astype(f64)
    x0, y0 = patch.CW[0, 0].astype(f64) @ sky + patch.crpix[0, 0].astype(f64)
    assert x0 == pytest.approx((n - 1) / 2.0 + 0.3, abs=1e-4)
    assert s2 == pytest.approx((radius / 0.03) ** 2 + sig_psf ** 2, rel=1e-5)
    flux, G = f64(prop["fluxes"][0, 0]), f64(patch.G[0])
    xx, yy = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    r2 = (xx - x0) ** 2 + (yy - y0) ** 2
    amp = G * flux / (2 * np.pi * s2)
    ea = r2 / s2
    H = 1.0 + ((r2 / s2 ** 2) - 2.0 / s2) / 24.0     # (vx^2 + vy^2 - fxx - fyy)/24
    expect = np.where(ea > 36.0, 0.0, amp * np.exp(-0.5 * ea) * H).reshape(-1)
    model64 = -fp64.evaluate(patch, prop)[2][:n * n]
    np.testing.assert_allclose(model64, expect, rtol=1e-9, atol=1e-14 * amp)
    model32 = oracle.render_model(patch, prop)[:n * n]
    np.testing.assert_allclose(model32, expect, rtol=3e-6, atol=1e-7 * amp)


def test_flux_normalisation():
    """A well-sampled Gaussian on a large stamp sums to flux * G."""
    flux, G = 5.0, 2.0
    patch, scene = _lone_gaussian_patch(n=96, radius=0.15, sig_psf=2.0, flux=flux, G=G,
                                        q=0.6, pa=0.4)
    model = -fp64.evaluate(patch, scene.get_proposal())[2]
    assert model.sum() == pytest.approx(flux * G, rel=2e-4)


def test_chi2_convention_and_zero_gradient_at_truth():
    """Data equal to the model: chi2 = -sum (d ierr)^2 and the gradient
    vanishes (verify_reference.py:136-143 conventions)."""
    rng = np.random.default_rng(5)
    spec = scenes.draw_patch_spec(rng, 2, [1, 1], [40, 40], 3, [2, 3])
    for nz in spec.noise:
        nz[...] = 0.0
    patch, scene, _ = scenes.build_patch(spec, StampPatch, oracle.render_model)
    prop = scene.get_proposal()
    c64, g64, r64, mass = fp64.evaluate(patch, prop, mass=True)
    d = patch.data.astype(np.float64) * patch.ierr
    per_band = [x.sum() for x in np.split(d * d, np.cumsum(patch.band_N_pix[:-1]))]
    np.testing.assert_allclose(c64, -np.array(per_band), rtol=1e-6)
    # the data are the float32-rounded model, so the gradient is rounding
    # noise: bound it by the gradient scale of a relative model change
    scale_w = 2.0 * patch.ierr.astype(np.float64) ** 2 * np.abs(patch.data)
    scale = fp64.evaluate(patch, prop, pixel_weight=scale_w, mass=True)[3]
    assert np.all(np.abs(g64) <= 1e-5 * scale + 1e-12)
    c32, g32, _ = oracle.evaluate_cpu(patch, prop)
    np.testing.assert_allclose(c32, c64, rtol=1e-6)


_SRC64 = np.dtype([(n, np.float64, s) if s else (n, np.float64)
                   for n, s in [("ra", 0), ("dec", 0), ("q", 0), ("pa", 0), ("nsersic", 0),
                                ("rh", 0), ("fluxes", 30), ("mixture_amplitudes", 10),
                                ("damplitude_drh", 10), ("damplitude_dnsersic", 10)]])


def _proposal64(scene):
    """The scene's proposal in float64 (no float32 quantisation of the
    parameters, so finite differences can use small steps)."""
    out = np.zeros(len(scene.sources), dtype=_SRC64)
    for i, g in enumerate(scene.sources):
        dn, dr = g.amplitude_model.derivatives(g.sersic, g.rhalf)
        out[i]["ra"], out[i]["dec"], out[i]["q"], out[i]["pa"] = g.ra, g.dec, g.q, g.pa
        out[i]["nsersic"], out[i]["rh"] = g.sersic, g.rhalf
        out[i]["fluxes"][:g.n_bands] = g.flux
        out[i]["mixture_amplitudes"][:g.n_gauss] = g.amplitudes
        out[i]["damplitude_drh"][:g.n_gauss] = dr
        out[i]["damplitude_dnsersic"][:g.n_gauss] = dn
    return out


def _chi2_at(patch, scene, theta):
    sc = copy.deepcopy(scene)
    sc.set_all_source_params(theta)
    return fp64.evaluate(patch, _proposal64(sc))[0].sum()


def test_central_differences_every_parameter():
    """Analytic float64 gradient (through the Jacobian and the amplitude
    derivatives) against central differences of chi2."""
    rng = np.random.default_rng(9)
    spec = scenes.draw_patch_spec(rng, 2, [1, 1], [36, 36], 2, [2, 2])
    patch, scene, theta = scenes.build_patch(spec, StampPatch, oracle.render_model)
    theta = theta.copy()
    nb = 2
    theta[nb + 2] *= 0.9        # q of source 0 off the truth
    theta[nb + 3] += 0.2        # pa
    theta[nb + 5] *= 1.1        # rhalf
    scene.set_all_source_params(theta)
    _, g64, _ = fp64.evaluate(patch, _proposal64(scene))
    from forcepho_amd.model import stack_grad
    grad = stack_grad(g64, 2, nb)
    # steps: flux, flux, ra [deg], dec [deg], q, pa, sersic, rhalf
    steps = [1e-4, 1e-4, 1e-9, 1e-9, 1e-5, 1e-5, 1e-5, 1e-6]
    for i in range(len(theta)):
        h = steps[i % (nb + 6)]
        tp, tm = theta.copy(), theta.copy()
        tp[i] += h
        tm[i] -= h
        fd = (_chi2_at(patch, scene, tp) - _chi2_at(patch, scene, tm)) / (2 * h)
        assert fd == pytest.approx(grad[i], rel=2e-4, abs=1e-6 * np.abs(grad).max()), i


def test_round_source_has_no_pa_gradient():
    patch, scene = _lone_gaussian_patch(q=1.0, pa=0.3)
    rng = np.random.default_rng(0)
    patch.data = rng.normal(size=patch.npix).astype(np.float32)
    patch.data[patch.ierr == 0] = 0.0
    _, g64, _, mass = fp64.evaluate(patch, scene.get_proposal(), mass=True)
    assert abs(g64[0, 0, 4]) <= 1e-12 * mass[0, 0, 4] + 1e-15


def test_source_permutation_equivariance():
    patch, proposal, exp = U.load_golden("uneven_s6")
    perm = np.array([2, 0, 3, 1])
    c0, g0, _ = fp64.evaluate(patch, proposal)
    pp = copy.deepcopy(patch)
    pp.D, pp.CW = patch.D[:, perm], patch.CW[:, perm]
    pp.crpix, pp.crval = patch.crpix[:, perm], patch.crval[:, perm]
    c1, g1, _ = fp64.evaluate(pp, proposal[perm])
    np.testing.assert_allclose(c1, c0, rtol=1e-12)
    np.testing.assert_allclose(g1, g0[:, perm], rtol=1e-9, atol=1e-12 * np.abs(g0).max())


def test_reference_cpu_kernel_inside_band_at_full_size():
    """At the full multi-band benchmark size the C restatement of the
    reference CPU kernel (bit-identical to it) stays inside the tolerance
    band around the committed float64 values, as on the small fixtures."""
    import sys
    sys.path.insert(0, U.GOLDEN_DIR)
    import make_fullsize as F
    ref = np.load(F.OUT)
    patch, proposal = F.build()
    np.testing.assert_allclose(F.data_digest(patch), ref["digest"], rtol=1e-12)
    c32, d32, _ = oracle.evaluate_cpu(patch, proposal)
    assert np.max(np.abs(c32 - ref["chi2"]) / np.abs(ref["chi2"])) <= 1e-6
    assert U.grad_errors(d32, ref["dchi2"], ref["mass"], ref["smass"]).max() <= 1.0


@pytest.mark.parametrize("field", ["q", "pa", "ra", "dec"])
def test_nan_geometry_gives_nan_chi2(field):
    """A NaN in one source's position, axis ratio or angle makes the
    reference CPU kernel's chi2 NaN in every band: its cutoff test
    `exparg > MAX_EXP_ARG` (common_kernel.cc:73,108) is false for NaN, so the
    NaN reaches every pixel's model.  (rh and sersic enter only through the
    host's mixture amplitudes.)"""
    patch, proposal, _ = U.load_golden("multi_small_s2", return_residual=False)
    p = proposal.copy()
    p[field][3] = np.nan
    runs = [oracle.evaluate_cpu(patch, p)]
    mod = oracle.reference_module()
    if mod is not None:
        runs.append(oracle.evaluate_reference(patch, p, module=mod))
    for chi2, dchi2, _ in runs:
        assert np.all(np.isnan(chi2))
        assert not np.all(np.isfinite(dchi2))
