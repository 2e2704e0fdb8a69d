"""The catalog of all sources and the regions checked out of it (mirrors
forcepho/superscene.py).

A ``SuperScene`` holds every source of a field in one structured catalog
(``sourcecat_dtype``) and hands out *regions*: a circle on the sky, the
active sources to fit in it and the fixed sources around it.  Checked-out
sources cannot be checked out again until their region is checked in with
new parameters (superscene.py:229-349); many disjoint regions can therefore
be fitted at the same time -- on one MI355X, as one batched launch over all
of their patches (``dispatcher.PatchDispatcher``).

``LinkedSuperScene`` groups sources whose radii of influence overlap
(friends of friends) and never splits a group between active and fixed when
it can avoid it (superscene.py:495-748).  ``make_bounds``/``bounds_vectors``
give the prior box of each source (:805-899), ``flux_bounds`` its flux
range (:922-965).

Scene coordinates are offsets in arcsec along longitude and latitude of a
frame whose origin is the catalog's median position (an offset frame with no
rotation, as astropy's ``skyoffset_frame`` builds, superscene.py:163-227),
computed here with the equivalent 3-d rotation R_y(-dec0) R_z(ra0) of unit
vectors.

Deliberate differences, all about several regions being out at once:
``is_valid`` is cleared while *any* checked-out region holds a source (a
count per source), where the reference re-validates a source when the first
region holding it checks in, even if another region still holds it, active
or fixed (superscene.py:316-320); seed weights (``exp_weight``,
``sigmoid_weight``) give zero weight to sources that are held, the documented
intent of superscene.py:470-492 (the reference applies a bitwise not to the
integer flag there); and ``grow_scene`` refuses a scene whose
friends-of-friends groups reach a source that another region holds active,
so concurrently checked-out regions never share an active source.
"""

import json

import numpy as np

from .region import CircularRegion
from .sources import Galaxy

__all__ = ["REQUIRED_COLUMNS", "SuperScene", "LinkedSuperScene", "sourcecat_dtype",
           "make_bounds", "bounds_vectors", "check_bounds", "flux_bounds", "sky_offsets"]

REQUIRED_COLUMNS = ("ra", "dec", "rhalf", "source_index", "is_active", "is_valid",
                    "n_iter", "n_patch")


def sourcecat_dtype(source_type=np.float64, bands=()):
    """dtype of a source catalog row (forcepho/superscene.py:751-764)."""
    dt = [(t, np.int32) for t in ("id", "source_index", "is_active", "is_valid",
                                  "n_iter", "n_patch")]
    dt += [("roi", np.float32)]
    dt += [(c, source_type) for c in Galaxy.SHAPE_COLS]
    dt += [(c, source_type) for c in bands]
    return np.dtype(dt)


def sky_offsets(ra, dec, ra0, dec0):
    """(lon, lat) in arcsec of (ra, dec) in the offset frame centred on
    (ra0, dec0) with no rotation; all angles in degrees."""
    ra, dec = np.deg2rad(np.asarray(ra, float)), np.deg2rad(np.asarray(dec, float))
    a0, d0 = np.deg2rad(ra0), np.deg2rad(dec0)
    x = np.cos(dec) * np.cos(ra)
    y = np.cos(dec) * np.sin(ra)
    z = np.sin(dec)
    # rotate the centre to lon = 0 (about z), then to lat = 0 (about y)
    x1 = x * np.cos(a0) + y * np.sin(a0)
    y1 = -x * np.sin(a0) + y * np.cos(a0)
    x2 = x1 * np.cos(d0) + z * np.sin(d0)
    z2 = -x1 * np.sin(d0) + z * np.cos(d0)
    lon = np.arctan2(y1, x2)
    lat = np.arcsin(np.clip(z2, -1.0, 1.0))
    return np.rad2deg(lon) * 3600.0, np.rad2deg(lat) * 3600.0


class SuperScene:
    """All sources of a field, and the bookkeeping of checked-out regions.

    Parameters
    ----------
    sourcecat : structured array with REQUIRED_COLUMNS, the shape columns and
        one flux column per band
    bands : list of band (flux column) names
    target_niter : iterations every source should receive (``undone``)
    maxactive_fraction : checkouts stop being ``sparse`` above this fraction
        of active sources
    maxactive_per_patch, nscale, boundary_radius, maxradius, minradius :
        region geometry (arcsec), as superscene.py:58-82
    """

    sourcecat = None
    n_sources = 0

    def __init__(self, sourcecat=None, bands=None, statefile="superscene.fits",
                 target_niter=200, maxactive_fraction=0.1, maxactive_per_patch=20, nscale=3,
                 boundary_radius=8.0, maxradius=6.0, minradius=1, bounds_kwargs={},
                 rng=None):
        self.statefilename = statefile
        self.bands = list(bands)
        self.shape_cols = list(Galaxy.SHAPE_COLS)
        self.rng = rng if rng is not None else np.random.default_rng()
        self.n_active = 0
        self.n_fixed = 0
        self.maxactive_fraction = maxactive_fraction
        self.target_niter = target_niter
        self.maxradius = maxradius
        self.minradius = minradius
        self.maxactive = maxactive_per_patch
        self.boundary_radius = boundary_radius
        self.nscale = nscale
        if sourcecat is not None:
            self.bounds_kwargs = bounds_kwargs
            self.ingest(sourcecat, **bounds_kwargs)

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.writeout()

    def writeout(self, filename=None):
        """Write the catalog (FITS table) and the check-in logs (JSON)."""
        from .formats.fits import write_table
        filename = filename or self.statefilename
        write_table(filename, self.sourcecat, {"FILTERS": ",".join(self.bands)})
        with open(filename.replace(".fits", "_log.json"), "w") as fobj:
            json.dump(dict(sourcelog=self.sourcelog, patchlog=self.patchlog), fobj)

    @property
    def sparse(self):
        return self.n_active / self.n_sources < self.maxactive_fraction

    @property
    def undone(self):
        return bool(np.any(np.abs(self.sourcecat["n_iter"]) < self.target_niter))

    @property
    def parameter_columns(self):
        return self.bands + self.shape_cols

    @property
    def cat_dtype(self):
        return self.sourcecat.dtype

    def ingest(self, sourcecat, **bounds_kwargs):
        """Set the catalog, build the KD-tree of scene coordinates, the bounds
        catalog and identity covariance matrices (superscene.py:115-134)."""
        from scipy.spatial import cKDTree
        self.set_catalog(sourcecat)
        self.kdt = cKDTree(self.scene_coordinates)
        self.bounds_catalog = make_bounds(self.sourcecat, self.bands,
                                          shapenames=self.shape_cols, **bounds_kwargs)
        n_param = len(self.parameter_columns)
        self.covariance_matrices = np.tile(np.eye(n_param), (self.n_sources, 1, 1))
        self.sourcelog = {}
        self.patchlog = []

    def check_bounds(self):
        check_bounds(self.sourcecat, self.bounds_catalog)

    def set_catalog(self, sourcecat):
        for c in REQUIRED_COLUMNS + tuple(self.bands):
            assert c in sourcecat.dtype.names, f"required column {c} is not present."
        self.sourcecat = sourcecat
        self.n_sources = len(sourcecat)
        self.sourcecat["source_index"][:] = np.arange(self.n_sources)
        self.sourcecat["is_valid"][:] = True
        self.sourcecat["is_active"][:] = False
        self._held = np.zeros(self.n_sources, dtype=np.int64)   # regions holding each source
        self.original = sourcecat.copy()

    def sky_to_scene(self, ra, dec):
        """Scene coordinates (lon, lat) in arcsec of sky positions (degrees)."""
        if not hasattr(self, "_scene_center"):
            self._scene_center = (float(np.median(self.sourcecat["ra"])),
                                  float(np.median(self.sourcecat["dec"])))
        return sky_offsets(ra, dec, *self._scene_center)

    @property
    def scene_coordinates(self):
        """(n_sources, 2) scene coordinates of the original positions."""
        if not hasattr(self, "_scene_coordinates"):
            x, y = self.sky_to_scene(self.sourcecat["ra"], self.sourcecat["dec"])
            self._scene_coordinates = np.array([x, y]).T
        return self._scene_coordinates

    def _mark_out(self, active_inds, fixed_inds):
        self.sourcecat["is_active"][active_inds] = True
        np.add.at(self._held, np.asarray(active_inds, dtype=int), 1)
        np.add.at(self._held, np.asarray(fixed_inds, dtype=int), 1)
        self.sourcecat["is_valid"][:] = self._held == 0
        self.n_active += len(active_inds)
        self.n_fixed += len(fixed_inds)

    def checkout_region(self, seed_index=-1):
        """(region, active rows, fixed rows) of a circular scene around a
        seed source, or (center, None, None) if an active source is within
        the boundary radius (superscene.py:229-267)."""
        cra, cdec = self.draw_center(seed_index=seed_index)
        center = self.sky_to_scene(cra, cdec)
        radius, active_inds, fixed_inds = self.get_circular_scene(center)
        if radius is None:
            return center, None, None
        region = CircularRegion(cra, cdec, radius / 3600.0)
        self._mark_out(active_inds, fixed_inds)
        return region, self.sourcecat[active_inds], self.sourcecat[fixed_inds]

    def checkin_region(self, active, fixed=None, niter=1, block_covs=None, new_bounds=None,
                       taskID=None, flush=False):
        """Write back the active sources' parameters, release them and the
        fixed sources, count iterations and patches (superscene.py:269-349)."""
        active_inds = np.asarray(active["source_index"]).astype(int)
        fixed_inds = (np.asarray(fixed["source_index"]).astype(int) if fixed is not None
                      else np.array([], dtype=int))
        release = np.bincount(np.concatenate([active_inds, fixed_inds]),
                              minlength=self.n_sources)
        if np.any(self.sourcecat["is_active"][active_inds] == 0) or np.any(release > self._held):
            raise ValueError("checked in a source that was not checked out")
        for f in self.parameter_columns:
            self.sourcecat[f][active_inds] = active[f]
        self.sourcecat["n_iter"][active_inds] += niter
        self.sourcecat["is_active"][active_inds] = False
        self.sourcecat["n_patch"][active_inds] += int(niter > 0)
        self._held -= release
        self.sourcecat["is_valid"][:] = self._held == 0
        self.n_active -= len(active_inds)
        self.n_fixed -= len(fixed_inds)
        if new_bounds is not None:
            self.bounds_catalog[active_inds] = new_bounds[:]
        if block_covs is not None:
            self.covariance_matrices[active_inds] = block_covs
        if taskID is not None:
            pid = str(taskID)
            for k in active_inds:
                self.sourcelog.setdefault(int(k), []).append(pid)
            self.patchlog.append(pid)
        if flush:
            self.writeout()

    def bounds_and_covs(self, sourceIDs):
        """Bounds rows and the block-diagonal covariance of some sources."""
        from scipy.linalg import block_diag
        bounds = self.bounds_catalog[sourceIDs]
        cov = block_diag(*self.covariance_matrices[sourceIDs])
        return bounds, cov

    def reset(self):
        self._held[:] = 0
        self.sourcecat["is_valid"][:] = True
        self.sourcecat["is_active"][:] = False
        self.sourcecat["n_iter"][:] = 0
        self.sourcecat["n_patch"][:] = 0
        self.n_active = 0
        self.n_fixed = 0

    def get_circular_scene(self, center):
        """(radius arcsec, active indices, fixed indices) of a circular scene
        about ``center`` (scene coordinates, arcsec) (superscene.py:372-440).

        A source reaches ``nscale`` half-light radii beyond its distance from
        the centre.  The sources that reach least far, as many as reach less
        than ``maxradius`` and at most ``maxactive``, are active; the radius
        is the furthest of their reaches, at least ``minradius``.  Of the
        others, those whose near edge (distance less ``nscale`` half-light
        radii) lies inside the radius are fixed, at most ``maxactive`` and at
        least the nearest-reaching one.  All three are None when a source
        within ``boundary_radius`` is active elsewhere or none fits."""
        near = np.asarray(self.kdt.query_ball_point(center, self.boundary_radius), dtype=int)
        if near.size == 0 or self.sourcecat["is_active"][near].any():
            return None, None, None
        sep = np.hypot(*(self.scene_coordinates[near] - center).T)
        halo = self.nscale * self.sourcecat["rhalf"][near]
        rank = np.argsort(sep + halo)
        reach, edge = (sep + halo)[rank], (sep - halo)[rank]
        k = min(self.maxactive, int(np.count_nonzero(reach < self.maxradius)))
        if k == 0:
            return None, None, None
        radius = max(self.minradius, reach[k - 1])
        rest = np.arange(k, len(rank))
        fixed = rest[edge[k:] < radius][:self.maxactive]
        if fixed.size == 0:
            fixed = rest[:1]
        return radius, near[rank[:k]], near[rank[fixed]]

    def draw_center(self, seed_index=-1):
        """(ra, dec) of a seed source: the given one, or a random draw with
        ``seed_weight`` (superscene.py:442-465)."""
        if seed_index is None or seed_index < 0:
            k = self.rng.choice(self.n_sources, p=self.seed_weight())
        else:
            k = seed_index
        seed = self.sourcecat[k]
        return seed["ra"], seed["dec"]

    def seed_weight(self):
        return self.exp_weight()

    def _available(self):
        return (self.sourcecat["is_valid"] != 0).astype(float)

    def exp_weight(self):
        """Available sources, favouring those with fewer iterations."""
        w = self._available()
        n = np.abs(self.sourcecat["n_iter"])
        w = w * np.exp((n.mean() - n) / 20.0)
        if w.sum() == 0:
            return np.full(self.n_sources, 1.0 / self.n_sources)
        return w / w.sum()

    def sigmoid_weight(self):
        """Available sources, weight falling past target_niter."""
        w = self._available()
        x = 20.0 * (1 - self.sourcecat["n_iter"] / self.target_niter) - 1.0
        w = w / (1 + np.exp(-x))
        if w.sum() == 0:
            return np.full(self.n_sources, 1.0 / self.n_sources)
        return w / w.sum()


class LinkedSuperScene(SuperScene):
    """SuperScene whose regions keep friends-of-friends groups together.

    Extra parameters
    ----------------
    buffer : arcsec added around the circle enclosing the active sources
    roi : per-source radius of influence in arcsec (default rhalf)
    strict : leave slots empty rather than split a group (except the first)
    """

    def __init__(self, *args, buffer=0.3, roi=None, strict=False, **kwargs):
        super().__init__(*args, **kwargs)
        self.buffer = buffer
        self.strict = strict
        self.roi = np.asarray(self.sourcecat["rhalf"] if roi is None else roi, dtype=float)

    def find_overlaps(self, center, radius, seed_index=-1, sort=False):
        """Sources other than seed_index with |x_j - center| < radius + roi_j
        (superscene.py:677-711), optionally sorted by that ratio."""
        kinds = np.array(self.kdt.query_ball_point(center, radius + self.boundary_radius),
                         dtype=int)
        if len(kinds) == 0:
            return kinds
        d = self.scene_coordinates[kinds] - center
        metric = np.hypot(*d.T) / (radius + self.roi[kinds])
        keep = (metric < 1) & (metric >= 0) & (kinds != seed_index)
        out, m = kinds[keep], metric[keep]
        return out[np.argsort(m, kind="stable")] if sort else out

    def grow_source(self, index):
        """Every source linked to ``index`` through overlapping radii of
        influence, ``index`` first (superscene.py:529-565)."""
        members = [int(index)]
        seen = {int(index)}
        frontier = [int(index)]
        while frontier:
            nxt = []
            for s in frontier:
                for b in self.find_overlaps(self.scene_coordinates[s], self.roi[s], seed_index=s):
                    b = int(b)
                    if b not in seen:
                        seen.add(b)
                        nxt.append(b)
            members += nxt
            frontier = nxt
        return members

    def grow_scene(self, seed_index, max_fixed=None):
        """(region, active indices, fixed indices) of a scene grown from a
        seed (superscene.py:567-626); (seed position, None, None) when a
        source near the seed, or one the growth would take, is active
        elsewhere.

        Whole friends-of-friends groups join in order of their nearest
        member's distance from the seed until one does not fit into
        ``maxactive``: that group then fills the remaining room if the scene
        is still empty or the scene is not ``strict``, and the growth stops.
        The region is centred on the mean position of the active sources and
        reaches ``buffer`` beyond the furthest radius of influence (at least
        ``minradius``); fixed sources are the nearest others within 1.2
        times the region radius, at most ``max_fixed`` (default
        ``maxactive``)."""
        seed = self.original[seed_index]
        busy = ((seed["ra"], seed["dec"]), None, None)
        near = self.find_overlaps(self.scene_coordinates[seed_index], self.boundary_radius,
                                  sort=True)
        if self.sourcecat["is_active"][near].any():
            return busy
        chosen, taken = [], set()
        for s in near:
            if s in taken:
                continue
            group = [int(m) for m in self.grow_source(s) if m not in taken]
            room = self.maxactive - len(chosen)
            if len(group) > room:
                if not (self.strict and chosen):
                    chosen += group[:room]
                break
            chosen += group
            taken.update(group)
        if self.sourcecat["is_active"][chosen].any():
            return busy
        rows = self.sourcecat[chosen]
        ra0, dec0 = float(np.mean(rows["ra"])), float(np.mean(rows["dec"]))
        middle = np.array(self.sky_to_scene(ra0, dec0))
        spread = np.hypot(*(self.scene_coordinates[chosen] - middle).T)
        radius = max(self.minradius, float(np.max(spread + self.roi[chosen] + self.buffer)))
        limit = self.maxactive if max_fixed is None else max_fixed
        others = [int(i) for i in self.find_overlaps(middle, 1.2 * radius, sort=True)
                  if i not in chosen]
        return CircularRegion(ra0, dec0, radius / 3600.0), chosen, others[:limit]

    def checkout_region(self, seed_index=-1, max_fixed=None):
        """(region, active rows, fixed rows or None); ((seed position,
        seed index), None, None) when the seed's neighbourhood is busy
        (superscene.py:628-675)."""
        if seed_index is None or seed_index < 0:
            seed_index = int(self.rng.choice(self.n_sources, p=self.seed_weight()))
        region, active_inds, fixed_inds = self.grow_scene(seed_index, max_fixed)
        if active_inds is None:
            return (region, seed_index), None, None
        fixed = self.sourcecat[fixed_inds] if len(fixed_inds) else None
        self._mark_out(np.asarray(active_inds, dtype=int), np.asarray(fixed_inds, dtype=int))
        return region, self.sourcecat[active_inds], fixed

    def overlap_circle(self, ra, dec, radius):
        """Catalog rows whose radius of influence meets a circle (radius in
        arcsec)."""
        return self.sourcecat[self.find_overlaps(np.array(self.sky_to_scene(ra, dec)), radius)]

    def make_group_catalog(self):
        """Friends-of-friends group index of every source."""
        gid = np.full(self.n_sources, -1)
        g = 0
        for i in range(self.n_sources):
            if gid[i] >= 0:
                continue
            gid[self.grow_source(i)] = g
            g += 1
        return gid


def make_bounds(active, filternames, shapenames=Galaxy.SHAPE_COLS, unccat=None,
                n_sig_flux=5.0, dpos=None, n_pix=2, pixscale=0.03,
                sqrtq_range=(0.4, 1.0), pa_range=(-0.6 * np.pi, 0.6 * np.pi),
                rhalf_range=(0.03, 0.3), sersic_range=(1.0, 5.0)):
    """Parameter bounds of each source, a structured array with a
    ``source_index`` field and a (lower, upper) pair per parameter
    (superscene.py:805-881).

    Positions: the catalog position plus or minus ``dpos`` arcsec (default
    ``n_pix`` pixels of ``pixscale``; the ra half-width divided by cos dec).
    Shapes: the same ranges for every source.  Fluxes: plus or minus
    ``n_sig_flux`` times the uncertainty catalog's value (no less than 5% of
    the flux) when ``unccat`` is given, else ``flux_bounds``."""
    n = len(active)
    half = np.full(2, n_pix * pixscale) if dpos is None else dpos
    names = list(filternames) + list(shapenames)
    out = np.zeros(n, dtype=[("source_index", np.int32)] + [(c, np.float64, (2,)) for c in names])
    out["source_index"] = active["source_index"]
    ranges = {"q": sqrtq_range, "pa": pa_range, "sersic": sersic_range, "rhalf": rhalf_range}
    for name, r in ranges.items():
        out[name] = r
    widths = {"ra": half[0] / 3600.0 / np.cos(np.deg2rad(active["dec"])),
              "dec": np.full(n, half[1] / 3600.0)}
    for name, w in widths.items():
        out[name] = np.stack([active[name] - w, active[name] + w], axis=-1)
    for b in filternames:
        f = np.asarray(active[b], dtype=np.float64)
        if unccat is None:
            lo, hi = flux_bounds(f, n_sig_flux)
        else:
            w = n_sig_flux * np.maximum(unccat[b], np.abs(0.05 * f))
            lo, hi = f - w, f + w
        out[b] = np.stack([lo, hi], axis=-1)
    return out


def bounds_vectors(bounds_cat, filternames, shapenames=Galaxy.SHAPE_COLS,
                   reference_coordinates=(0.0, 0.0)):
    """Flat lower and upper bound vectors in the Scene parameter order
    (fluxes, ra, dec, q, pa, sersic, rhalf per source), positions relative to
    the patch reference coordinates (superscene.py:884-899)."""
    shift = {"ra": reference_coordinates[0], "dec": reference_coordinates[1]}
    names = list(filternames) + list(shapenames)
    table = np.stack([bounds_cat[c] - shift.get(c, 0.0) for c in names], axis=1)
    return table[..., 0].reshape(-1), table[..., 1].reshape(-1)


def check_bounds(sourcecat, boundscat):
    """Raise AssertionError unless every parameter with a (lower, upper)
    bound lies strictly between them, row by row (superscene.py:902-919)."""
    for name in boundscat.dtype.names:
        pair = np.asarray(boundscat[name])
        if pair.ndim != 2 or pair.shape[1] != 2:
            continue
        value = np.asarray(sourcecat[name])
        outside = np.flatnonzero(~((value > pair[:, 0]) & (value < pair[:, 1])))
        if outside.size:
            raise AssertionError(f"{name}: rows {outside.tolist()} are not strictly inside "
                                 "their bounds")


def flux_bounds(flux, unc1, snr_max=10, precisions=None):
    """(lower, upper) flux bounds of each source (superscene.py:922-965).

    With a precision matrix: ``unc1`` standard deviations either side, the
    deviation no smaller than flux / ``snr_max``, the upper bound at least
    ``unc1`` deviations.  Without: a half-width of ``unc1`` times the square
    root of |flux| (as if the noise at flux 1 were ``unc1``); fluxes at or
    below f0 = (unc1 / 2)^2 get a lower bound of at most -f0, every upper
    bound is at least f0, and a zero flux (no width) gets lower bound 0 and
    the largest upper bound of the band."""
    flux = np.asarray(flux, dtype=float)
    if precisions is not None:
        sd = np.maximum(np.sqrt(np.diag(np.linalg.pinv(precisions))), flux / snr_max)
        return flux - unc1 * sd, np.maximum(flux + unc1 * sd, unc1 * sd)
    f0 = (unc1 / 2.0) ** 2
    with np.errstate(divide="ignore", invalid="ignore"):
        width = np.abs(flux) / (np.sqrt(np.abs(flux)) / unc1)
    lo = np.where(flux <= f0, np.minimum(flux - width, -f0), flux - width)
    hi = np.maximum(flux + width, f0)
    lo = np.where(np.isfinite(lo), lo, 0.0)
    hi = np.where(np.isfinite(hi), hi, np.nanmax(hi))
    return lo, hi
