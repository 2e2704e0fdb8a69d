"""The one-GPU dispatcher on a synthetic field: a FITS image of four groups
of galaxies rendered by the checker from a truth catalog plus noise; the
SuperScene starts from a perturbed catalog.  All regions are sampled in
batched launches and checked in; afterwards nothing is held, every source
has its iterations, and the catalog fits the image better than the start
(chi2 of the whole image, evaluated by the kernel)."""

import numpy as np
import pytest

import oracle
from fpx_testutil import verification_amplitudes, verification_psfstore

pytestmark = pytest.mark.gpu

BAND = "F200W"
SCALE = 0.03 / 3600.0
RA0, DEC0 = 53.1, -27.8


def _header(nx, ny):
    return {"CTYPE1": "RA---TAN", "CTYPE2": "DEC--TAN", "CRPIX1": nx / 2 + 0.5,
            "CRPIX2": ny / 2 + 0.5, "CRVAL1": RA0, "CRVAL2": DEC0, "PC1_1": -SCALE,
            "PC2_2": SCALE, "FILTER": BAND}


def _truth_catalog():
    from forcepho_amd.superscene import sourcecat_dtype
    rng = np.random.default_rng(11)
    centers = [(-1.4, -1.4), (1.4, -1.3), (-1.3, 1.4), (1.3, 1.3)]   # arcsec
    rows = []
    for cx, cy in centers:
        for k in range(2):
            dx, dy = cx + rng.uniform(-0.25, 0.25), cy + rng.uniform(-0.25, 0.25)
            rows.append((dx, dy))
    cat = np.zeros(len(rows), dtype=sourcecat_dtype(bands=[BAND]))
    cat["ra"] = RA0 + np.array([r[0] for r in rows]) / 3600 / np.cos(np.deg2rad(DEC0))
    cat["dec"] = DEC0 + np.array([r[1] for r in rows]) / 3600
    cat["q"] = rng.uniform(0.6, 0.9, len(rows))
    cat["pa"] = rng.uniform(-1.0, 1.0, len(rows))
    cat["sersic"] = rng.uniform(1.5, 3.5, len(rows))
    cat["rhalf"] = rng.uniform(0.06, 0.15, len(rows))
    cat[BAND] = rng.uniform(3.0, 6.0, len(rows))
    cat["roi"] = 3 * cat["rhalf"]
    return cat


def _render_field(tmp_path, cat, nx=128, ny=128, sigma=0.002):
    """Truth image through the C checker of a whole-image FITSPatch."""
    from forcepho_amd.formats import fits
    from forcepho_amd.patches import FITSPatch
    hdr = _header(nx, ny)
    blank = str(tmp_path / "blank.fits")
    fits.write_images(blank, [np.zeros((ny, nx), np.float32),
                              np.full((ny, nx), sigma, np.float32)], [hdr, None])
    amps = verification_amplitudes("rect")
    p = FITSPatch(fitsfiles=[blank], psfstore=verification_psfstore(), splinedata=amps)
    p.build_patch(None, None, allbands=[BAND])
    scene = p.set_scene(cat)
    p.pack_meta(scene)
    model = oracle.render_model(p, scene.get_proposal())
    valid = p.xpix >= 0
    img = np.zeros((nx, ny))
    img[p.xpix[valid].astype(int), p.ypix[valid].astype(int)] = model[valid]
    rng = np.random.default_rng(12)
    img = img + rng.normal(0, sigma, img.shape)
    path = str(tmp_path / "field.fits")
    fits.write_images(path, [img.T.astype(np.float32), np.full((ny, nx), sigma, np.float32)],
                      [hdr, None])
    return path, amps


def _chi2(path, amps, cat):
    from forcepho_amd.patches import FITSDevicePatch
    p = FITSDevicePatch(fitsfiles=[path], psfstore=verification_psfstore(), splinedata=amps)
    p.build_patch(None, None, allbands=[BAND])
    scene = p.set_scene(cat)
    p.pack_meta(scene)
    p.send_to_device()
    chi2, _ = p.get_proposer().evaluate_proposal(scene.get_proposal(), patch=p)
    ie = p.ierr.astype(np.float64)
    return float(chi2.sum() + np.sum((p.data * ie) ** 2))


def test_dispatcher_fits_all_regions(tmp_path):
    from forcepho_amd.superscene import LinkedSuperScene
    from forcepho_amd.dispatcher import PatchDispatcher
    from forcepho_amd.patches import FITSDevicePatch
    truth = _truth_catalog()
    path, amps = _render_field(tmp_path, truth)
    start = truth.copy()
    start[BAND] *= 1.25
    start["rhalf"] *= 1.15
    start["q"] = np.clip(start["q"] * 1.1, 0.3, 0.95)
    sc = LinkedSuperScene(start.copy(), [BAND], roi=start["roi"], target_niter=10,
                          maxactive_per_patch=4, boundary_radius=1.0, maxradius=1.2,
                          minradius=0.5, buffer=0.3, rng=np.random.default_rng(1),
                          bounds_kwargs=dict(n_sig_flux=20.0, sqrtq_range=(0.3, 1.0),
                                             rhalf_range=(0.03, 0.3)))
    store = verification_psfstore()

    def make_patch(region):
        p = FITSDevicePatch(fitsfiles=[path], psfstore=store, splinedata=amps,
                            return_residual=True)
        p.build_patch(region, None, allbands=[BAND])
        return p

    disp = PatchDispatcher(sc, make_patch, max_regions=8, chains=2, n_leapfrog=12,
                           n_warmup=25, n_samples=10, step_size=0.05, seed=3)
    print("dispatcher run", flush=True)
    rounds = disp.run(max_rounds=6)
    assert rounds >= 1 and not sc.undone
    assert sc.n_active == 0 and sc.n_fixed == 0
    assert (sc.sourcecat["is_valid"] == 1).all() and (sc.sourcecat["is_active"] == 0).all()
    assert (sc.sourcecat["n_patch"] >= 1).all() and (sc.sourcecat["n_iter"] >= 10).all()
    # several regions sampled together in the first round
    assert disp.history[0]["n_regions"] >= 2
    final = sc.sourcecat
    for c in ("ra", "dec", "q", "pa", "sersic", "rhalf", BAND):
        assert np.all(np.isfinite(final[c]))
    c_start, c_final, c_truth = (_chi2(path, amps, c) for c in (start, final, truth))
    print("chi2 start %.1f final %.1f truth %.1f" % (c_start, c_final, c_truth), flush=True)
    assert c_final < c_start
    assert c_final - c_truth < 0.25 * (c_start - c_truth)
    err0 = np.abs(start[BAND] / truth[BAND] - 1).mean()
    err1 = np.abs(final[BAND] / truth[BAND] - 1).mean()
    assert err1 < 0.5 * err0
