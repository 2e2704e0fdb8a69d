"""PSF Gaussian mixtures per band (mirrors forcepho/patches/storage.py:436-489).

A PSF store holds, for each band, an array ``parameters`` of shape
(n_locations, n_psf_per_source) of PSF Gaussians (amp, xcen, ycen, Cxx, Cyy,
Cxy, sersic_bin), optionally ``detector_locations`` (2, n_locations), and the
band attribute ``n_psf_per_source``.  Each row holds, for every Sersic radius
bin, the PSF Gaussians convolved with that bin's mixture component.  The
reference reads the store with h5py; here an HDF5 store is read with
``forcepho_amd.formats.hdf5``, or a store is built in memory from one mixture
per band (``PSFStore.from_mixtures``: the same Gaussians for every radius bin,
which is what a PSF fitted without radius dependence means).
"""

import numpy as np

from .patch import PSF_COLS, PSF_DTYPE

__all__ = ["PSFStore", "psf_records"]


def psf_records(mixture, sersic_bin=0):
    """PSF_DTYPE records of a Gaussian mixture table with columns amp,
    xcen|xcr, ycen|ycr, Cxx, Cyy, Cxy (a structured array)."""
    mixture = np.asarray(mixture)
    names = mixture.dtype.names
    out = np.zeros(len(mixture), dtype=PSF_DTYPE)
    for c in PSF_COLS:
        src = c
        if c not in names:
            src = {"xcen": "xcr", "ycen": "ycr"}.get(c, c)
        out[c] = mixture[src]
    out["sersic_bin"] = sersic_bin
    return out


class PSFStore:
    """PSF mixtures by band.

    Parameters
    ----------
    h5file : str or None
        HDF5 file with a group per band holding ``parameters`` (and optionally
        ``detector_locations``) and the attribute ``n_psf_per_source``.
    """

    def __init__(self, h5file=None):
        self.h5file = h5file
        self._bands = {}

    @classmethod
    def from_mixtures(cls, mixtures, n_radii):
        """A store with one location per band: ``mixtures`` maps band name ->
        table of PSF Gaussians (see ``psf_records``), replicated for each of
        the n_radii Sersic radius bins (sersic_bin = 0 .. n_radii - 1)."""
        store = cls(None)
        for band, mix in mixtures.items():
            base = psf_records(mix)
            rows = np.concatenate([psf_records(base, sersic_bin=r) for r in range(n_radii)])
            store._bands[band] = {"parameters": rows[None, :], "detector_locations": None,
                                  "n_psf_per_source": len(rows)}
        return store

    def _band(self, band):
        if band not in self._bands:
            if not self.h5file:
                raise KeyError("PSF store has no band %r" % band)
            from ..formats.hdf5 import File
            with File(self.h5file) as f:
                g = f[band]
                pars = np.asarray(g["parameters"][:])
                locs = np.asarray(g["detector_locations"][:]) if "detector_locations" in g else None
                n = int(g.attrs.get("n_psf_per_source", pars.shape[-1]))
            if pars.ndim == 1:
                pars = pars[None, :]
            self._bands[band] = {"parameters": pars, "detector_locations": locs,
                                 "n_psf_per_source": n}
        return self._bands[band]

    def n_psf_per_source(self, band):
        return int(self._band(band)["n_psf_per_source"])

    def lookup(self, band, xy=None):
        """PSF Gaussians (n_psf_per_source,) of the location nearest to xy
        (detector pixels), or of the first location when xy is None or the
        store has no locations (forcepho/patches/storage.py:455-468)."""
        b = self._band(band)
        choose = 0
        locs = b["detector_locations"]
        if xy is not None and locs is not None:
            x, y = xy
            xp, yp = locs
            choose = int(np.argmin(np.hypot(x - xp, y - yp)))
        rows = b["parameters"][choose]
        out = np.zeros(len(rows), dtype=PSF_DTYPE)
        for c in PSF_DTYPE.names:
            out[c] = rows[c]
        return out

    def get_local_psf(self, band="F090W", source=None, wcs=None):
        """The PSF Gaussians for a source in a band; position dependent only
        when a wcs is given (forcepho/patches/storage.py:470-485)."""
        xy = None
        if wcs is not None and source is not None:
            xy = wcs.all_world2pix(source.ra, source.dec, 0)
        return self.lookup(band, xy=xy)
