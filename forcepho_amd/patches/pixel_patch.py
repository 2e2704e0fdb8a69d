"""Patches packed from exposure images on disk (mirrors
forcepho/patches/pixel_patch.py).

``PixelPatch.build_patch(region, sourcecat, allbands)`` finds the exposures
that overlap a sky region, packs their pixels band by band and, given a
source catalog, the scene metadata (pixel_patch.py:52-122).  Packing follows
pixel_patch.py:124-202 and the padding rule of :473-484 -- every exposure
padded to a multiple of 64 pixels (one wavefront) with flux = 0, ierr = 0,
x = y = -1; non-finite or negative inverse errors are masked to 0
(:449-451).  One difference in the packed order: within an exposure the
selected pixels are sorted into 16x16 squares of 8x8 squares
(``pixel_order="superpixel"``, the default) so that a wavefront's 64 pixels
are close together; ``pixel_order="flat"`` keeps the reference's x-major
order.  Results do not depend on the order beyond float32 summation order;
xpix/ypix of the packed arrays describe it.

``FITSPatch`` reads science images (extension ``sci_ext``, band from the
FILTER card), inverse errors from the ``unc_ext`` uncertainty image (or a
constant ``unc``, or ``snr / flux``), and the TAN WCS of the science header
(pixel_patch.py:349-484).
"""

import numpy as np

from .patch import Patch
from .storage import PSFStore

__all__ = ["PixelPatch", "FITSPatch", "JWST_BANDS", "superpixel_sort"]

JWST_BANDS = ["F090W", "F115W", "F150W", "F200W",
              "F277W", "F335M", "F356W", "F410M", "F444W"]


def superpixel_sort(x, y, block=16, sub=8):
    """Permutation putting pixels (x, y) into block x block squares (x-major),
    each as sub x sub squares (x-major), each x-major."""
    x = np.asarray(x, dtype=np.int64)
    y = np.asarray(y, dtype=np.int64)
    return np.lexsort((y % sub, x % sub, (y % block) // sub, (x % block) // sub,
                       y // block, x // block))


class PixelPatch(Patch):
    """Packs pixel data found by ``find_exposures``/``find_pixels``
    (implemented by subclasses) and scene metadata from a source catalog.

    Parameters
    ----------
    psfstore : str (HDF5 PSF store) or PSFStore
    splinedata : str, HDF5 table of Sersic mixture amplitudes
        (sources.SplineMixtureAmplitudes.from_file), or an amplitude model
    spline_type : "rect" (the reference's default) or "smooth"
    """

    def __init__(self, psfstore="", splinedata="", spline_smoothing=None,
                 return_residual=False, meta_dtype=np.float32, pix_dtype=np.float32,
                 debug=0, pixel_order="superpixel", spline_type="rect"):
        super().__init__(return_residual=return_residual, meta_dtype=meta_dtype,
                         pix_dtype=pix_dtype)
        self.psfstore = psfstore if isinstance(psfstore, PSFStore) else PSFStore(psfstore)
        self.splinedata = splinedata
        self.spline_smoothing = spline_smoothing
        self.spline_type = spline_type
        self.debug = debug
        assert pixel_order in ("superpixel", "flat")
        self.pixel_order = pixel_order
        self.background_offsets = None
        self.max_snr = None

    def build_patch(self, region=None, sourcecat=None, allbands=JWST_BANDS,
                    tweak_background=False):
        """Find and pack the pixels of every exposure overlapping ``region``
        (None: whole images) in the bands ``allbands``, and the metadata of
        ``sourcecat`` if given (forcepho/patches/pixel_patch.py:52-122)."""
        self.hdrs, self.wcses, self.epaths, self.band_names = self.find_exposures(region, allbands)
        if len(self.epaths) == 0:
            raise ValueError("No exposures in the specified bands overlap the region")
        if tweak_background:
            self.background_offsets = [hdr.get(tweak_background, 0.0) for hdr in self.hdrs]
        else:
            self.background_offsets = None
        allbands = list(allbands)
        band_ids = [allbands.index(b) for b in self.band_names]
        assert (np.diff(band_ids) >= 0).all(), "Exposures must be sorted by band"
        u, n = np.unique(band_ids, return_counts=True)
        self.uniq_bands, self.n_exp_per_band = u, n
        self.bandlist = [allbands[i] for i in u]
        self.n_bands = len(u)
        self.n_exp = len(self.hdrs)
        self.bands = np.searchsorted(u, band_ids).astype(np.int32)
        self.pack_pix(region=region)
        if sourcecat is not None:
            scene = self.set_scene(sourcecat)
            self.pack_meta(scene)
        self._dirty_data = False

    def pack_pix(self, region, dtype=None):
        """Concatenate the exposures' pixels (forcepho/patches/pixel_patch.py:124-202)."""
        dtype = dtype or self.pix_dtype
        w = self.warp_size
        self.band_start = np.empty(self.n_bands, dtype=np.int16)
        self.band_N = np.zeros(self.n_bands, dtype=np.int16)
        self.exposure_start = np.empty(self.n_exp, dtype=np.int32)
        self.exposure_N = np.empty(self.n_exp, dtype=np.int32)
        self.original_sizes = np.zeros(self.n_exp, dtype=np.int64)
        data, ierr, xpix, ypix = [], [], [], []
        i = 0
        for e, wcs in enumerate(self.wcses):
            flux, ie, xp, yp = self.find_pixels(self.epaths[e], wcs, region)
            n_real = flux.size
            assert n_real > 0, "There were no valid pixels in exposure {}".format(self.epaths[e])
            if self.pixel_order == "superpixel":
                order = superpixel_sort(xp, yp)
                flux, ie, xp, yp = flux[order], ie[order], xp[order], yp[order]
            if self.background_offsets is not None:
                flux = flux - self.background_offsets[e]
            pad = (-n_real) % w
            flux = np.append(flux, np.zeros(pad))
            ie = np.append(ie, np.zeros(pad))
            xp = np.append(xp, np.full(pad, -1))
            yp = np.append(yp, np.full(pad, -1))
            self.band_N[self.bands[e]] += 1
            self.exposure_start[e] = i
            self.exposure_N[e] = flux.size
            self.original_sizes[e] = n_real
            data.append(flux)
            ierr.append(ie)
            xpix.append(xp)
            ypix.append(yp)
            i += flux.size
        self.data = np.concatenate(data).astype(dtype)
        self.ierr = np.concatenate(ierr).astype(dtype)
        self.xpix = np.concatenate(xpix).astype(dtype)
        self.ypix = np.concatenate(ypix).astype(dtype)
        assert self.data.shape[0] == i
        self.band_start[0] = 0
        self.band_start[1:] = np.cumsum(self.band_N)[:-1]
        if self.max_snr:
            to_cap = (self.ierr > 0) & (self.data * self.ierr > self.max_snr)
            self.ierr[to_cap] = self.max_snr / self.data[to_cap]

    def set_scene(self, sourcecat, splinedata=None):
        """A Scene of Galaxies from a catalog with SHAPE_COLS and one flux
        column per band of ``bandlist`` (forcepho/patches/patch.py:400-425)."""
        from ..sources import Scene
        return Scene.from_catalog(sourcecat, filternames=self.bandlist,
                                  splinedata=splinedata if splinedata is not None
                                  else self.splinedata,
                                  spline_smoothing=self.spline_smoothing,
                                  spline_type=self.spline_type)

    def pack_meta(self, scene, sky_zero=None):
        """Scene metadata, with each exposure's PSF Gaussians from the store
        (forcepho/patches/patch.py:244-398)."""
        self.psfs = [self.psfstore.get_local_psf(band=self.band_names[e])
                     for e in range(self.n_exp)]
        self.photocounts = np.ones(self.n_exp, dtype=self.meta_dtype)
        super().pack_meta(scene, sky_zero=sky_zero)

    def find_exposures(self, region, bandlist):
        raise NotImplementedError

    def find_pixels(self, epath, wcs, region):
        raise NotImplementedError


class FITSPatch(PixelPatch):
    """Pixels and WCS from FITS images.

    Parameters
    ----------
    fitsfiles : list of str
    sci_ext, unc_ext : extensions of the science and uncertainty images
    """

    def __init__(self, fitsfiles=(), psfstore="", splinedata="", spline_smoothing=None,
                 return_residual=False, sci_ext=0, unc_ext=1, meta_dtype=np.float32,
                 pix_dtype=np.float32, debug=0, pixel_order="superpixel", spline_type="rect"):
        super().__init__(psfstore=psfstore, splinedata=splinedata,
                         spline_smoothing=spline_smoothing, return_residual=return_residual,
                         meta_dtype=meta_dtype, pix_dtype=pix_dtype, debug=debug,
                         pixel_order=pixel_order, spline_type=spline_type)
        self.fitsfiles = list(fitsfiles)
        self.sci_ext = sci_ext
        self.unc_ext = unc_ext
        self.snr = None
        self.unc = None

    def find_exposures(self, region, bandlist, do_check=True):
        """Headers, WCSes, paths and bands of the images in ``bandlist`` that
        overlap the region, in band order (pixel_patch.py:387-436)."""
        from ..formats import fits
        from ..formats.wcs import TanWCS
        hdrs_all = [fits.getheader(f, self.sci_ext) for f in self.fitsfiles]
        imbands = [str(h["FILTER"]).strip() for h in hdrs_all]
        epaths, bands, hdrs, wcses = [], [], [], []
        for band in bandlist:
            for path, hdr, b in zip(self.fitsfiles, hdrs_all, imbands):
                if b != band:
                    continue
                wcs = TanWCS(hdr)
                if region is not None and do_check:
                    bra, bdec = region.bounding_box
                    nx, ny = hdr["NAXIS1"], hdr["NAXIS2"]
                    bx, by = wcs.all_world2pix(bra, bdec, self.wcs_origin)
                    inim = np.any((bx > 0) & (bx < nx) & (by > 0) & (by < ny))
                    around = (nx < bx.max()) & (bx.min() < 0) & (ny < by.max()) & (by.min() < 0)
                    if not (inim or around):
                        continue
                epaths.append(path)
                wcses.append(wcs)
                bands.append(b)
                hdrs.append(hdr)
        return hdrs, wcses, epaths, bands

    def find_pixels(self, epath, wcs, region):
        """flux, inverse error, x, y of the image pixels in the region
        (pixel_patch.py:438-471; padding is added by pack_pix)."""
        from ..formats import fits
        flux = np.array(fits.getdata(epath, self.sci_ext), dtype=np.float64).T
        with np.errstate(divide="ignore", invalid="ignore"):
            if self.snr is not None:
                ie = self.snr / flux
            elif self.unc is not None:
                ie = np.zeros_like(flux) + 1.0 / self.unc
            else:
                ie = 1.0 / np.array(fits.getdata(epath, self.unc_ext), dtype=np.float64).T
        bad = ~np.isfinite(flux) | ~np.isfinite(ie) | (ie < 0)
        ie[bad] = 0
        flux[bad] = 0
        nx, ny = flux.shape
        yp, xp = np.meshgrid(np.arange(ny), np.arange(nx))
        if region is None:
            sel = (slice(None), slice(None))
        else:
            offsets = np.array([(0, 0), (1, 0), (1, 1), (0, 1)])
            lower_left = np.array([xp, yp])
            corners = offsets[:, :, None, None] + lower_left[None, :, :, :]
            corners = corners.transpose(2, 3, 0, 1)
            sel = region.contains(corners[..., 0], corners[..., 1], wcs, origin=self.wcs_origin)
        return (flux[sel].reshape(-1), ie[sel].reshape(-1),
                xp[sel].reshape(-1), yp[sel].reshape(-1))
