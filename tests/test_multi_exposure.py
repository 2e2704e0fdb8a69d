"""FITSPatch over several FITS files, bands and WCSes (forcepho's real-data
case, pixel_patch.py:387-484, patch.py:504-560): two F200W exposures (one
rotated by 20 degrees and shifted) and one F444W exposure at twice the pixel
scale.  Each image is rendered by the C restatement through a patch of that
file alone; the tests then pack all three together and check

* the packing: band order, exposures per band, the pixel-scale determinant
  and the rotation of each exposure's D matrix;
* the astrometry against the images themselves: the flux-weighted centroid
  of each source's rendered light sits where the image's own TAN WCS puts the
  source (plus the PSF's mean offset);
* the model: the combined patch reproduces every rendered pixel at the
  catalog parameters (zero residual), so each exposure keeps its own WCS, PSF
  and pixels through the packing."""

import numpy as np
import pytest

import fpx_testutil as U
import oracle
from fpx_testutil import verification_amplitudes
from forcepho_amd.formats import fits
from forcepho_amd.formats.wcs import TanWCS

RA0, DEC0 = 53.16, -27.78
ARCSEC = 1.0 / 3600.0
IMAGES = [  # (band, pixel scale arcsec, rotation deg, shift px, nx, ny)
    ("F200W", 0.03, 0.0, (0.0, 0.0), 96, 80),
    ("F200W", 0.03, 20.0, (3.3, -2.7), 88, 96),
    ("F444W", 0.06, 0.0, (0.4, 0.2), 48, 48),
]
SOURCES = [  # (dx arcsec, dy arcsec, q, pa, sersic, rhalf)
    (-0.55, -0.35, 0.7, 0.4, 2.0, 0.10),
    (0.60, 0.45, 0.85, -0.6, 3.0, 0.12),
]


def _header(band, scale, rot, shift, nx, ny):
    s, t = scale * ARCSEC, np.deg2rad(rot)
    return {"CTYPE1": "RA---TAN", "CTYPE2": "DEC--TAN",
            "CRPIX1": nx / 2 + 0.5 + shift[0], "CRPIX2": ny / 2 + 0.5 + shift[1],
            "CRVAL1": RA0, "CRVAL2": DEC0,
            "CD1_1": -s * np.cos(t), "CD1_2": s * np.sin(t),
            "CD2_1": s * np.sin(t), "CD2_2": s * np.cos(t), "FILTER": band}


def _catalog():
    from forcepho_amd.superscene import sourcecat_dtype
    cat = np.zeros(len(SOURCES), dtype=sourcecat_dtype(bands=["F200W", "F444W"]))
    for i, (dx, dy, q, pa, n, rh) in enumerate(SOURCES):
        cat[i]["ra"] = RA0 + dx * ARCSEC / np.cos(np.deg2rad(DEC0))
        cat[i]["dec"] = DEC0 + dy * ARCSEC
        cat[i]["q"], cat[i]["pa"], cat[i]["sersic"], cat[i]["rhalf"] = q, pa, n, rh
        cat[i]["F200W"], cat[i]["F444W"] = 4.0 + i, 6.0 - i
        cat[i]["roi"] = 3 * rh
    return cat


def _psfstore():
    from forcepho_amd.patches import PSFStore
    base = np.zeros(2, dtype=[(k, np.float64) for k in ("amp", "xcr", "ycr", "Cxx", "Cyy", "Cxy")])
    base["amp"] = [0.7, 0.3]
    base["xcr"], base["ycr"] = [0.05, -0.1], [0.0, 0.08]
    base["Cxx"], base["Cyy"], base["Cxy"] = [1.2, 4.0], [1.0, 3.5], [0.1, 0.0]
    wide = base.copy()
    wide["Cxx"] *= 0.6
    wide["Cyy"] *= 0.6
    return PSFStore.from_mixtures({"F200W": base, "F444W": wide}, n_radii=9), base, wide


def _patch(paths, bands, store, amps):
    from forcepho_amd.patches import FITSPatch
    p = FITSPatch(fitsfiles=paths, psfstore=store, splinedata=amps)
    p.build_patch(None, None, allbands=bands)
    return p


@pytest.fixture(scope="module")
def field(tmp_path_factory):
    """Paths of the three rendered images (zero noise), the catalog, the
    PSF store and its mixtures, the amplitude model."""
    tmp = tmp_path_factory.mktemp("multi")
    cat = _catalog()
    store, base, wide = _psfstore()
    amps = verification_amplitudes("rect")
    paths, images = [], []
    for k, (band, scale, rot, shift, nx, ny) in enumerate(IMAGES):
        hdr = _header(band, scale, rot, shift, nx, ny)
        blank = str(tmp / f"blank{k}.fits")
        fits.write_images(blank, [np.zeros((ny, nx), np.float32), np.full((ny, nx), 0.01, np.float32)],
                          [hdr, None])
        p = _patch([blank], [band], store, amps)
        scene = p.set_scene(cat)
        p.pack_meta(scene)
        model = oracle.render_model(p, scene.get_proposal())
        valid = p.xpix >= 0
        img = np.zeros((nx, ny))
        img[p.xpix[valid].astype(int), p.ypix[valid].astype(int)] = model[valid]
        path = str(tmp / f"image{k}.fits")
        fits.write_images(path, [img.T.astype(np.float32), np.full((ny, nx), 0.01, np.float32)],
                          [hdr, None])
        paths.append(path)
        images.append(img)
    return paths, images, cat, store, (base, wide), amps


def test_multi_file_packing(field):
    paths, _, cat, store, _, amps = field
    p = _patch(paths, ["F200W", "F444W"], store, amps)
    scene = p.set_scene(cat)
    p.pack_meta(scene)
    assert p.bandlist == ["F200W", "F444W"] and p.n_bands == 2 and p.n_exp == 3
    assert list(p.band_N) == [2, 1] and list(p.band_start) == [0, 2]
    assert list(p.exposure_start) == [0, int(p.exposure_N[0]),
                                      int(p.exposure_N[0] + p.exposure_N[1])]
    for e, (band, scale, rot, _, nx, ny) in enumerate(IMAGES):
        assert int(p.original_sizes[e]) == nx * ny
        for s in range(len(SOURCES)):
            D = p.D[e, s].astype(np.float64)
            # pixels per arcsec: |det D| = 1 / scale^2 (local TAN scale)
            assert abs(np.linalg.det(D)) == pytest.approx(1 / scale ** 2, rel=2e-4)
    # the rotated exposure's D (sky -> pixels, the inverse of the header's CD)
    # is the first one's turned by -20 degrees
    R = p.D[1, 0].astype(np.float64) @ np.linalg.inv(p.D[0, 0].astype(np.float64))
    assert np.degrees(np.arctan2(R[1, 0], R[0, 0])) == pytest.approx(-20.0, abs=0.01)
    assert abs(np.linalg.det(R)) == pytest.approx(1.0, rel=1e-4)
    assert p.n_psf_per_source[0] == p.n_psf_per_source[1] == 2 * 9


def test_rendered_centroids_follow_each_wcs(field, tmp_path):
    """Centroid of each source's light, rendered one source at a time into a
    600x600 image with the WCS of each of the three images (reference pixel
    moved by 250 px so that every mixture component lies inside), against
    that WCS's pixel position of the source plus the flux-weighted PSF
    offset."""
    _, _, cat, store, mixes, amps = field
    pad, big = 250, 600
    for k, (band, scale, rot, shift, nx, ny) in enumerate(IMAGES):
        mix = mixes[0] if band == "F200W" else mixes[1]
        off = np.array([np.sum(mix["amp"] * mix["xcr"]), np.sum(mix["amp"] * mix["ycr"])])
        hdr = _header(band, scale, rot, shift, nx, ny)
        hdr["CRPIX1"] += pad
        hdr["CRPIX2"] += pad
        path = str(tmp_path / f"big{k}.fits")
        fits.write_images(path, [np.zeros((big, big), np.float32),
                                 np.full((big, big), 0.01, np.float32)], [hdr, None])
        wcs = TanWCS(fits.getheader(path))
        for s in range(len(SOURCES)):
            p = _patch([path], [band], store, amps)
            scene = p.set_scene(cat[s:s + 1])
            p.pack_meta(scene)
            model = oracle.render_model(p, scene.get_proposal())
            valid = p.xpix >= 0
            w = model[valid]
            cx = np.sum(w * p.xpix[valid]) / np.sum(w)
            cy = np.sum(w * p.ypix[valid]) / np.sum(w)
            x, y = wcs.all_world2pix(cat[s]["ra"], cat[s]["dec"], 0)
            assert pad < float(x) < big - pad and pad < float(y) < big - pad
            # measured: ~2e-5 px
            assert np.hypot(cx - (float(x) + off[0]), cy - (float(y) + off[1])) < 1e-3, (k, s)


def test_multi_file_model_reproduces_every_image(field):
    paths, images, cat, store, _, amps = field
    p = _patch(paths, ["F200W", "F444W"], store, amps)
    scene = p.set_scene(cat)
    p.pack_meta(scene)
    _, _, resid = oracle.evaluate_cpu(p, scene.get_proposal(), return_residual=True)
    valid = p.xpix >= 0
    scale = max(np.abs(im).max() for im in images)
    assert np.abs(resid[valid]).max() <= 2e-6 * scale
    # and every exposure's pixels are its own image's
    for e in range(3):
        sl = slice(int(p.exposure_start[e]), int(p.exposure_start[e] + p.exposure_N[e]))
        v = p.xpix[sl] >= 0
        got = p.data[sl][v]
        want = images[e][p.xpix[sl][v].astype(int), p.ypix[sl][v].astype(int)]
        np.testing.assert_allclose(got, want.astype(np.float32), rtol=0, atol=0)


@pytest.mark.gpu
def test_gpu_multi_file_patch_parity(field):
    """The HIP kernel on the three-file patch, off the truth: chi2 against the
    C restatement, gradients against the float64 oracle within the float32
    tolerance, residual image against the C restatement."""
    from forcepho_amd.patches import FITSDevicePatch
    paths, _, cat, store, _, amps = field
    p = FITSDevicePatch(fitsfiles=paths, psfstore=store, splinedata=amps, return_residual=True)
    p.build_patch(None, None, allbands=["F200W", "F444W"])
    scene = p.set_scene(cat)
    p.pack_meta(scene)
    prop = scene.get_proposal()
    prop["q"] *= np.float32(0.93)
    prop["fluxes"] *= np.float32(1.05)
    p.send_to_device()
    chi2, derivs, resid = p.get_proposer().evaluate_proposal(prop, patch=p, unpack=False)
    c32, _, r32 = oracle.evaluate_cpu(p, prop, return_residual=True)
    assert np.max(np.abs(chi2 - c32) / np.abs(c32)) <= U.CHI2_RTOL
    _, g64, mass, smass = U.fp64_reference(p, prop)
    assert U.grad_errors(derivs, g64, mass, smass).max() <= 1.0
    valid = p.xpix >= 0
    assert np.max(np.abs(resid - r32)[valid]) <= 2e-6 * np.abs(p.data).max()
