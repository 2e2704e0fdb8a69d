"""SuperScene checkout / check-in invariants (forcepho/superscene.py:229-349,
495-748), the scene-coordinate frame, friends-of-friends grouping and the
bounds helpers (:805-965).

Invariants: a checkout marks its active sources active and invalid and its
fixed sources invalid and updates the counts; no source is active in two
regions at once; a checkout near an active source is refused; check-in
writes the new parameters back, counts the iterations and patches, and
releases every source; running regions until ``undone`` is False gives every
source the target iterations and leaves nothing checked out.
"""

import json

import numpy as np
import pytest

from forcepho_amd.superscene import (SuperScene, LinkedSuperScene, sourcecat_dtype,
                                     make_bounds, bounds_vectors, check_bounds,
                                     flux_bounds, sky_offsets)
from forcepho_amd.sources import Galaxy


def _catalog(n=150, seed=0, field=40.0, bands=("F1", "F2")):
    rng = np.random.default_rng(seed)
    cat = np.zeros(n, dtype=sourcecat_dtype(bands=list(bands)))
    ra0, dec0 = 53.1, -27.8
    cat["ra"] = ra0 + rng.uniform(-field, field, n) / 3600 / np.cos(np.deg2rad(dec0))
    cat["dec"] = dec0 + rng.uniform(-field, field, n) / 3600
    cat["rhalf"] = rng.uniform(0.05, 0.3, n)
    cat["q"] = rng.uniform(0.5, 0.95, n)
    cat["pa"] = rng.uniform(-1, 1, n)
    cat["sersic"] = rng.uniform(1.5, 4, n)
    for b in bands:
        cat[b] = rng.lognormal(1, 0.5, n)
    cat["roi"] = 5 * cat["rhalf"]
    cat["id"] = np.arange(n) + 1000
    return cat, list(bands)


def test_sky_offsets_frame():
    ra0, dec0 = 53.1, -27.8
    lon, lat = sky_offsets(ra0, dec0, ra0, dec0)
    assert abs(lon) < 1e-9 and abs(lat) < 1e-9
    # along the meridian the latitude offset is the declination offset
    lon, lat = sky_offsets(ra0, dec0 + 0.01, ra0, dec0)
    assert abs(lon) < 1e-9 and lat == pytest.approx(36.0, rel=1e-12)
    # a small step in ra is cos(dec) shorter on the sky
    lon, lat = sky_offsets(ra0 + 1e-4, dec0, ra0, dec0)
    assert lon == pytest.approx(0.36 * np.cos(np.deg2rad(dec0)), rel=1e-6)
    assert abs(lat) < 1e-6
    # the frame is a rotation: angular separations are preserved
    rng = np.random.default_rng(1)
    ra = ra0 + rng.uniform(-0.05, 0.05, 20)
    dec = dec0 + rng.uniform(-0.05, 0.05, 20)
    lon, lat = sky_offsets(ra, dec, ra0, dec0)
    lo, la = np.deg2rad(lon / 3600), np.deg2rad(lat / 3600)
    r, d, r0, d0 = map(np.deg2rad, (ra, dec, ra0, dec0))
    def hav(dlat, dlon, c1, c2):       # haversine separation
        return 2 * np.arcsin(np.sqrt(np.sin(dlat / 2) ** 2 + c1 * c2 * np.sin(dlon / 2) ** 2))
    sep_sky = hav(d - d0, r - r0, np.cos(d), np.cos(d0))
    sep_frame = hav(la, lo, np.cos(la), 1.0)
    np.testing.assert_allclose(sep_frame, sep_sky, rtol=1e-9)


def _assert_state(sc, out):
    """Flags and counts agree with the regions currently checked out: a
    source is active iff one region holds it active, and invalid iff any
    region holds it (active or fixed)."""
    active = [int(i) for _, a, _ in out for i in a["source_index"]]
    fixed = [int(i) for _, _, f in out if f is not None for i in f["source_index"]]
    assert len(active) == len(set(active)), "a source is active in two regions"
    assert sc.n_active == len(active) and sc.n_fixed == len(fixed)
    is_active = np.zeros(sc.n_sources, bool)
    is_active[active] = True
    assert np.array_equal(sc.sourcecat["is_active"] != 0, is_active)
    invalid = is_active.copy()
    invalid[fixed] = True
    assert np.array_equal(sc.sourcecat["is_valid"] == 0, invalid)


@pytest.mark.parametrize("linked", [False, True])
def test_checkout_checkin_invariants(linked):
    cat, bands = _catalog()
    kw = dict(target_niter=3, maxactive_per_patch=6, boundary_radius=4.0, maxradius=3.0,
              rng=np.random.default_rng(2))
    sc = (LinkedSuperScene(cat.copy(), bands, roi=cat["roi"], **kw) if linked
          else SuperScene(cat.copy(), bands, **kw))
    out = []
    for seed in range(40):
        before = sc.sourcecat.copy()
        region, active, fixed = sc.checkout_region(seed_index=seed)
        if active is None:
            # refused because a source near the seed is active: no side effects
            assert sc.n_active > 0
            assert np.array_equal(sc.sourcecat, before)
            _assert_state(sc, out)
            continue
        assert 1 <= len(active) <= sc.maxactive
        # the region encloses its active sources
        c = np.array(sc.sky_to_scene(region.ra, region.dec))
        d = np.hypot(*(sc.scene_coordinates[active["source_index"]] - c).T)
        assert np.all(d <= region.radius * 3600 + 1e-9)
        out.append((region, active, fixed))
        _assert_state(sc, out)
        # a second checkout seeded on an active source is refused
        assert sc.checkout_region(seed_index=int(active["source_index"][0]))[1] is None
    assert len(out) >= 3
    # check in half of them with new parameters
    for region, active, fixed in out[::2]:
        new = active.copy()
        new["F1"] *= 2.0
        new["rhalf"] = 0.123
        sc.checkin_region(new, fixed, niter=2, taskID="t%d" % len(sc.patchlog))
        idx = active["source_index"]
        np.testing.assert_array_equal(sc.sourcecat["F1"][idx], 2.0 * active["F1"])
        np.testing.assert_array_equal(sc.sourcecat["rhalf"][idx], 0.123)
        assert (sc.sourcecat["n_iter"][idx] == 2).all() and (sc.sourcecat["n_patch"][idx] == 1).all()
    out = out[1::2]
    _assert_state(sc, out)
    for region, active, fixed in out:
        sc.checkin_region(active, fixed, niter=0)
    _assert_state(sc, [])
    assert sc.n_active == 0 and sc.n_fixed == 0
    with pytest.raises(ValueError):
        sc.checkin_region(out[0][1], out[0][2])      # not checked out any more
    assert len(sc.patchlog) == len(set(sc.patchlog)) > 0


@pytest.mark.parametrize("linked", [False, True])
def test_run_until_done(linked):
    """Random checkouts with several regions out at a time until every source
    has target_niter iterations."""
    cat, bands = _catalog(n=80, seed=3, field=25.0)
    kw = dict(target_niter=2, maxactive_per_patch=5, boundary_radius=3.0, maxradius=2.5,
              rng=np.random.default_rng(4))
    sc = (LinkedSuperScene(cat.copy(), bands, roi=cat["roi"], **kw) if linked
          else SuperScene(cat.copy(), bands, **kw))
    out, n_checkins = [], 0
    for it in range(5000):
        if not sc.undone:
            break
        if len(out) < 4:
            region, active, fixed = sc.checkout_region()
            if active is not None:
                out.append((region, active, fixed))
                _assert_state(sc, out)
                continue
        if out:
            region, active, fixed = out.pop(0)
            sc.checkin_region(active, fixed, niter=1)
            n_checkins += 1
    for region, active, fixed in out:
        sc.checkin_region(active, fixed, niter=1)
    assert not sc.undone
    assert sc.n_active == 0 and sc.n_fixed == 0
    assert (sc.sourcecat["is_valid"] == 1).all() and (sc.sourcecat["is_active"] == 0).all()
    assert (sc.sourcecat["n_iter"] >= 2).all()


def test_friends_of_friends_groups():
    cat, bands = _catalog(n=120, seed=5, field=20.0)
    sc = LinkedSuperScene(cat.copy(), bands, roi=cat["roi"], maxactive_per_patch=8)
    gid = sc.make_group_catalog()
    assert (gid >= 0).all()
    xy = sc.scene_coordinates
    linked = np.hypot(*(xy[:, None, :] - xy[None, :, :]).transpose(2, 0, 1)) < (
        sc.roi[:, None] + sc.roi[None, :])
    # groups are the connected components of the overlap graph
    for g in np.unique(gid):
        members = np.where(gid == g)[0]
        grown = sc.grow_source(int(members[0]))
        assert sorted(grown) == sorted(members.tolist()) and len(set(grown)) == len(grown)
        outside = np.where(gid != g)[0]
        assert not linked[np.ix_(members, outside)].any()
    # a symmetric relation: j in group(i) <=> i in group(j)
    for i in range(0, 120, 17):
        for j in sc.grow_source(i):
            assert i in sc.grow_source(j)


def test_linked_checkout_keeps_groups_whole():
    cat, bands = _catalog(n=100, seed=6, field=15.0)
    sc = LinkedSuperScene(cat.copy(), bands, roi=cat["roi"], maxactive_per_patch=30,
                          strict=True, boundary_radius=3.0)
    gid = sc.make_group_catalog()
    region, active, fixed = sc.checkout_region(seed_index=0)
    assert active is not None
    groups = set(gid[active["source_index"]])
    for g in groups:
        if (gid == g).sum() <= sc.maxactive:
            assert set(np.where(gid == g)[0]) <= set(active["source_index"].tolist())


def test_bounds_helpers():
    cat, bands = _catalog(n=5, seed=7)
    b = make_bounds(cat, bands, n_pix=2, pixscale=0.03)
    assert (b["source_index"] == cat["source_index"]).all()
    np.testing.assert_allclose(b["dec"][:, 1] - b["dec"][:, 0], 2 * 0.06 / 3600)
    np.testing.assert_allclose((b["ra"][:, 1] - b["ra"][:, 0]) * np.cos(np.deg2rad(cat["dec"])),
                               2 * 0.06 / 3600)
    check_bounds(cat, b)
    ref = np.array([53.0, -27.0])
    lo, hi = bounds_vectors(b, bands, reference_coordinates=ref)
    assert lo.shape == (5 * 8,)
    cols = bands + list(Galaxy.SHAPE_COLS)
    k = cols.index("ra")
    np.testing.assert_allclose(lo.reshape(5, 8)[:, k], b["ra"][:, 0] - 53.0)
    np.testing.assert_allclose(hi.reshape(5, 8)[:, cols.index("sersic")], 5.0)
    bad = cat.copy()
    bad["q"][2] = 1.5
    with pytest.raises(AssertionError):
        check_bounds(bad, b)
    # heuristic flux bounds contain the flux and widen with unc1
    f = np.array([0.001, 1.0, 100.0])
    lo1, hi1 = flux_bounds(f, 1.0)
    lo2, hi2 = flux_bounds(f, 2.0)
    assert (lo1 < f).all() and (hi1 > f).all() and (lo2 <= lo1).all() and (hi2 >= hi1).all()
    # from a precision matrix: unc1 sigma either side, at least flux / snr_max
    prec = np.diag([1e4, 1.0])
    lo, hi = flux_bounds(np.array([10.0, 10.0]), 3.0, snr_max=10, precisions=prec)
    np.testing.assert_allclose(lo, [10 - 3 * 1.0, 10 - 3 * 1.0])
    np.testing.assert_allclose(hi, [10 + 3 * 1.0, 10 + 3 * 1.0])


def test_writeout_round_trip(tmp_path):
    from forcepho_amd.formats import fits
    cat, bands = _catalog(n=20, seed=8)
    sc = SuperScene(cat.copy(), bands, statefile=str(tmp_path / "scene.fits"))
    region, active, fixed = sc.checkout_region(seed_index=3)
    sc.checkin_region(active, fixed, niter=5, taskID=7)
    sc.writeout()
    back = fits.getdata(str(tmp_path / "scene.fits"), 1)
    assert fits.getheader(str(tmp_path / "scene.fits"), 1)["FILTERS"] == "F1,F2"
    for name in cat.dtype.names:
        np.testing.assert_array_equal(back[name], sc.sourcecat[name])
    logs = json.load(open(str(tmp_path / "scene_log.json")))
    assert logs["patchlog"] == ["7"]
    assert sorted(int(k) for k in logs["sourcelog"]) == sorted(active["source_index"].tolist())
