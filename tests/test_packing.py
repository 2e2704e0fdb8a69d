"""Packing invariants of StampPatch / Patch.pack_meta (CPU only), following
forcepho/patches/static_patch.py:238-307, pixel_patch.py:473-484 and
patch.py:244-398."""

import numpy as np
import pytest

from forcepho_amd import scenes
from forcepho_amd.patches import StampPatch, superpixel_order


def _zero_render(patch, proposal):
    return np.zeros(patch.npix)


@pytest.fixture(scope="module")
def packed():
    rng = np.random.default_rng(3)
    spec = scenes.draw_patch_spec(rng, 3, [2, 1, 3], [37, 64, 50], 4, [3, 2, 5])
    patch, scene, theta = scenes.build_patch(spec, StampPatch, _zero_render)
    return patch, scene, spec


def test_exposure_index_arrays(packed):
    patch, _, spec = packed
    assert patch.n_exp == 6 and patch.n_bands == 3
    assert np.all(patch.exposure_N % 64 == 0)
    np.testing.assert_array_equal(patch.exposure_start[1:],
                                  patch.exposure_start[:-1] + patch.exposure_N[:-1])
    assert patch.exposure_start[-1] + patch.exposure_N[-1] == patch.npix
    np.testing.assert_array_equal(patch.band_N, [2, 1, 3])
    np.testing.assert_array_equal(patch.band_start, [0, 2, 3])
    assert patch.band_start.dtype == np.int16 and patch.exposure_N.dtype == np.int32


def test_padding_pixels(packed):
    patch, _, _ = packed
    for e in range(patch.n_exp):
        s, n = patch.exposure_start[e], patch.original_sizes[e]
        pad = slice(s + n, s + patch.exposure_N[e])
        assert np.all(patch.ierr[pad] == 0) and np.all(patch.data[pad] == 0)
        assert np.all(patch.xpix[pad] == -1) and np.all(patch.ypix[pad] == -1)
        assert np.all(patch.ierr[s:s + n] > 0)


def test_superpixel_order_is_permutation_and_unpacks():
    for nx, ny in [(8, 8), (37, 50), (64, 16)]:
        o = superpixel_order(nx, ny)
        assert np.array_equal(np.sort(o), np.arange(nx * ny))
    rng = np.random.default_rng(0)
    spec = scenes.draw_patch_spec(rng, 1, [1], [37], 1, [2])
    img = rng.normal(size=(37, 37)).astype(np.float32)
    spec.stamps[0].pixel_values = img
    patch = StampPatch()
    patch.build_patch(spec.stamps, spec.scene, sky_zero=spec.sky_zero)
    np.testing.assert_array_equal(patch.unpack_image(patch.data, 0), img)
    x = patch.unpack_image(patch.xpix, 0)
    np.testing.assert_array_equal(x, np.repeat(np.arange(37)[:, None], 37, axis=1))


def test_psf_blocks(packed):
    patch, _, spec = packed
    ns = patch.n_sources
    for e in range(patch.n_exp):
        n = patch.n_psf_per_source[patch.bands[e]]
        blk = patch.psfgauss[patch.psfgauss_start[e]:patch.psfgauss_start[e] + ns * n]
        for s in range(ns):
            np.testing.assert_array_equal(blk[s * n:(s + 1) * n], spec.stamps[e].psf)
    np.testing.assert_array_equal(patch.n_psf_per_source, [9 * 3, 9 * 2, 9 * 5])
    assert len(patch.psfgauss) == sum(patch.n_psf_per_source[patch.bands]) * ns


def test_astrometry(packed):
    patch, scene, spec = packed
    for e in range(patch.n_exp):
        for i, src in enumerate(scene.sources):
            # crval is the source's own position (offset from the reference)
            np.testing.assert_allclose(patch.crval[e, i], [src.ra, src.dec], rtol=1e-6,
                                       atol=1e-12)
            D = patch.D[e, i]
            np.testing.assert_allclose(D, np.eye(2) / 0.03, rtol=1e-6)
            cosd = np.cos(np.deg2rad(scenes.SKY_CENTER[1] + src.dec))
            np.testing.assert_allclose(patch.CW[e, i], np.diag([3600 / 0.03 * cosd, 3600 / 0.03]),
                                       rtol=1e-5)


def test_pack_meta_twice_does_not_shift_twice(packed):
    patch, scene, _ = packed
    before = scene.get_all_source_params().copy()
    crval = patch.crval.copy()
    patch.pack_meta(scene)
    np.testing.assert_array_equal(scene.get_all_source_params(), before)
    np.testing.assert_array_equal(patch.crval, crval)


def test_n_evals_counts_real_pixels(packed):
    patch, _, _ = packed
    n = 0
    for e in range(patch.n_exp):
        n += patch.original_sizes[e] * patch.n_sources * patch.n_psf_per_source[patch.bands[e]]
    assert patch.n_evals() == n


def test_group_by_centre_is_a_stable_grouping():
    """PSF entries with one centre become adjacent, groups in order of first
    appearance, entries within a group in their original order; the block is
    a permutation of the input."""
    from forcepho_amd.patches.patch import group_by_centre, PSF_DTYPE
    rng = np.random.default_rng(4)
    n_comp, n_bins = 5, 7
    b = np.zeros(n_comp * n_bins, dtype=PSF_DTYPE)
    cx, cy = rng.normal(size=n_comp), rng.normal(size=n_comp)
    comp = np.tile(np.arange(n_comp), n_bins)            # bin-major, as a store lists them
    b["xcen"], b["ycen"] = cx[comp], cy[comp]
    b["sersic_bin"] = np.repeat(np.arange(n_bins), n_comp)
    b["amp"] = np.arange(len(b))
    g = group_by_centre(b)
    assert sorted(g["amp"].tolist()) == list(range(len(b)))
    assert (g["xcen"].reshape(n_comp, n_bins) == cx[:, None].astype(np.float32)).all()
    assert (g["sersic_bin"].reshape(n_comp, n_bins) == np.arange(n_bins)).all()
    # already grouped input is unchanged
    assert np.array_equal(group_by_centre(g), g)
