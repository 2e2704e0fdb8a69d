"""Host-side data model of a patch: pixel and metadata arrays packed in the
byte layout the kernel reads (include/forcepho_amd.h FpxPatch).

Mirrors forcepho/patches/patch.py: ``PatchBase.pdt`` (:44-73) is the same
struct dtype field for field, ``build_patch`` (:122-198) allocates the same
arrays, and ``Patch.pack_meta`` (:244-272) fills source-dependent metadata in
the same order (astrometry per exposure per source, flux calibration, PSF
Gaussians ordered [exposure][source][n_psf_per_source]).
"""

import numpy as np

__all__ = ["PatchBase", "Patch", "PSF_COLS", "PSF_DTYPE", "LinearWCS", "scale_at_sky",
           "group_by_centre"]

# PSF Gaussian record, forcepho/patches/storage.py:27-28
PSF_COLS = ["amp", "xcen", "ycen", "Cxx", "Cyy", "Cxy"]
PSF_DTYPE = np.dtype([(c, np.float32) for c in PSF_COLS] + [("sersic_bin", np.int32)])


class PatchBase:
    """Packs pixel data and scene metadata into flat arrays for the device.

    Parameters
    ----------
    return_residual : bool
        Whether the device writes a residual image on each evaluation.
    pix_dtype, meta_dtype : numpy dtype
        float32 (the kernel reads float32).
    """

    ptr_dtype = np.uintp
    pdt = np.dtype([("data", ptr_dtype),
                    ("ierr", ptr_dtype),
                    ("xpix", ptr_dtype),
                    ("ypix", ptr_dtype),
                    ("residual", ptr_dtype),
                    ("exposure_start", ptr_dtype),
                    ("exposure_N", ptr_dtype),
                    ("band_start", ptr_dtype),
                    ("band_N", ptr_dtype),
                    ("n_sources", np.int32),
                    ("n_radii", np.int32),
                    ("rad2", ptr_dtype),
                    ("D", ptr_dtype),
                    ("crpix", ptr_dtype),
                    ("crval", ptr_dtype),
                    ("CW", ptr_dtype),
                    ("G", ptr_dtype),
                    ("n_psf_per_source", ptr_dtype),
                    ("psfgauss", ptr_dtype),
                    ("psfgauss_start", ptr_dtype)],
                   align=True)
    patch_struct_dtype = pdt
    meta_names = ["n_sources", "n_radii", "rad2",
                  "D", "crpix", "crval", "CW", "G",
                  "n_psf_per_source", "psfgauss", "psfgauss_start"]

    # exposures are padded to a multiple of this many pixels (one wavefront)
    warp_size = 64

    def __init__(self, return_residual=False, pix_dtype=np.float32, meta_dtype=np.float32):
        self.pix_dtype = pix_dtype
        self.meta_dtype = meta_dtype
        self.return_residual = return_residual
        self._dirty_data = True

    @property
    def npix(self):
        try:
            return self.xpix.shape[0]
        except AttributeError:
            return 0

    @property
    def size(self):
        s = 0
        for name in self.patch_struct_dtype.names:
            try:
                s += getattr(self, name).nbytes
            except AttributeError:
                pass
        return s

    def clear(self):
        for name in self.patch_struct_dtype.names:
            try:
                delattr(self, name)
            except AttributeError:
                pass

    def free(self):
        pass

    def build_patch(self, npix=1, n_sources=1, n_radii=1, n_bands=1, n_exp=1):
        """Allocate every packed array with its final shape and dtype
        (forcepho/patches/patch.py:122-198)."""
        self.n_exp = n_exp
        self.n_bands = n_bands
        self.xpix = np.zeros(npix, dtype=self.pix_dtype)
        self.ypix = np.zeros(npix, dtype=self.pix_dtype)
        self.data = np.zeros(npix, dtype=self.pix_dtype)
        self.ierr = np.zeros(npix, dtype=self.pix_dtype)
        self.band_start = np.empty(n_bands, dtype=np.int16)
        self.band_N = np.zeros(n_bands, dtype=np.int16)
        self.exposure_start = np.empty(n_exp, dtype=np.int32)
        self.exposure_N = np.empty(n_exp, dtype=np.int32)
        self.n_sources = n_sources
        self.n_radii = n_radii
        self.rad2 = np.empty(n_radii, dtype=self.meta_dtype)
        self.D = np.empty((n_exp, n_sources, 2, 2), dtype=self.meta_dtype)
        self.CW = np.empty((n_exp, n_sources, 2, 2), dtype=self.meta_dtype)
        self.crpix = np.empty((n_exp, n_sources, 2), dtype=self.meta_dtype)
        self.crval = np.empty((n_exp, n_sources, 2), dtype=self.meta_dtype)
        self.G = np.empty(n_exp, dtype=self.meta_dtype)
        self.n_psf_per_source = np.zeros(n_bands, dtype=np.int32)
        self.psfgauss = np.empty(0, dtype=PSF_DTYPE)
        self.psfgauss_start = np.zeros(n_exp, dtype=np.int32)

    # ------------------------------------------------------------------
    # helpers shared by subclasses
    # ------------------------------------------------------------------
    def split_pix(self, attr):
        """Split a pixel array into per-exposure arrays (padding included)."""
        arr = getattr(self, attr)
        return np.split(arr, np.cumsum(self.exposure_N[:-1]))

    def split_band(self, attr):
        """Split a pixel array into per-band arrays."""
        arr = getattr(self, attr)
        return np.split(arr, np.cumsum(self.band_N_pix[:-1]))

    @property
    def band_N_pix(self):
        out = np.zeros(len(self.band_N), dtype=np.int64)
        for b, (s, n) in enumerate(zip(self.band_start, self.band_N)):
            out[b] = self.exposure_N[s:s + n].sum()
        return out

    def n_evals(self):
        """Pixel x ImageGaussian pairs one evaluation processes, counting real
        (unpadded) pixels only."""
        real = getattr(self, "original_sizes", self.exposure_N)
        n = 0
        for b in range(len(self.band_N)):
            s, k = int(self.band_start[b]), int(self.band_N[b])
            n += int(np.sum(real[s:s + k])) * int(self.n_sources) * int(self.n_psf_per_source[b])
        return n


def group_by_centre(block):
    """PSF entries reordered so that entries with the same centre (xcen,
    ycen) are adjacent: groups in order of their first entry, entries within
    a group in their original order (a stable permutation).  A store lists
    entries radius bin by radius bin; every bin of a PSF component shares
    the component's centre."""
    block = np.asarray(block)
    if len(block) < 2:
        return block.copy()
    key = np.stack([block["xcen"].astype(np.float32), block["ycen"].astype(np.float32)], axis=1)
    _, first, inverse = np.unique(key, axis=0, return_index=True, return_inverse=True)
    lead = first[np.asarray(inverse).reshape(-1)]
    return block[np.argsort(lead, kind="stable")]


class LinearWCS:
    """A local linear world<->pixel map of one exposure: pixel = crpix +
    CW (sky - crval_sky), CW = 3600 D diag(cos dec0, 1), D = (3600 CD)^-1.

    This is the linearisation the kernel itself uses (Patch.CW / crpix /
    crval), so scenes rendered through it are exact under the kernel's model.

    Parameters
    ----------
    crval : (ra0, dec0) degrees
    crpix : (x0, y0) pixels (0-based)
    cd : 2x2 degrees per pixel
    """

    has_distortion = False

    def __init__(self, crval, crpix, cd):
        self.crval = np.asarray(crval, dtype=np.float64)
        self.crpix = np.asarray(crpix, dtype=np.float64)
        self.pixel_scale_matrix = np.asarray(cd, dtype=np.float64)

    def all_world2pix(self, ra, dec, origin=0):
        _, CW = self._mats()
        d = np.array([ra - self.crval[0], dec - self.crval[1]])
        return self.crpix + CW @ d + origin

    def _mats(self):
        D = np.linalg.inv(self.pixel_scale_matrix * 3600.0)
        W = np.eye(2)
        W[0, 0] = np.cos(np.deg2rad(self.crval[1]))
        return D, (D * 3600.0) @ W


def scale_at_sky(sky, wcs, dpix=1.0, origin=1, make_approx=False):
    """CW (pix/deg, with the cos(dec) factor) and D (pix/arcsec) at ``sky``
    for a WCS without distortion (forcepho/patches/patch.py:554-558)."""
    ra, dec = sky
    W = np.eye(2)
    W[0, 0] = np.cos(np.deg2rad(dec))
    D_mat = np.linalg.inv(wcs.pixel_scale_matrix * 3600.0)
    CW_mat = np.matmul(D_mat * 3600.0, W)
    return CW_mat, D_mat


class Patch(PatchBase):
    """PatchBase plus metadata packing from a Scene, per-exposure WCSes and
    per-exposure PSF mixtures (forcepho/patches/patch.py:201-398).

    Required attributes (set by the pixel packer): ``bandlist``, ``n_bands``,
    ``n_exp``, ``n_exp_per_band``, ``wcses`` [n_exp], ``bands`` [n_exp] (band
    index of each exposure), ``psfs`` [n_exp] (PSF_DTYPE arrays of
    n_psf_per_source(band) Gaussians each), ``photocounts`` [n_exp] (G).
    """

    def __init__(self, return_residual=False, meta_dtype=np.float32, pix_dtype=np.float32):
        super().__init__(return_residual=return_residual, meta_dtype=meta_dtype,
                         pix_dtype=pix_dtype)
        self.patch_reference_coordinates = np.zeros(2)
        self.wcs_origin = 0

    def pack_meta(self, scene, sky_zero=None):
        """Pack all source-dependent metadata for ``scene``
        (forcepho/patches/patch.py:244-272)."""
        self.patch_reference_coordinates = self.zerocoords(scene, sky_zero=sky_zero)
        self.n_sources = len(scene.sources)
        self._pack_source_metadata(scene)
        self._pack_astrometry(self.wcses, scene)
        self._pack_fluxcal()
        self._pack_psf(scene)
        self.scene = scene

    def _pack_source_metadata(self, scene):
        self.n_sources = len(scene.sources)
        self.n_radii = len(scene.sources[0].radii)
        self.rad2 = np.asarray(scene.sources[0].radii, dtype=np.float64) ** 2
        self.rad2 = self.rad2.astype(self.meta_dtype)

    def _pack_astrometry(self, wcses, scene):
        dt = self.meta_dtype
        self.D = np.empty((self.n_exp, self.n_sources, 2, 2), dtype=dt)
        self.CW = np.empty((self.n_exp, self.n_sources, 2, 2), dtype=dt)
        self.crpix = np.empty((self.n_exp, self.n_sources, 2), dtype=dt)
        self.crval = np.empty((self.n_exp, self.n_sources, 2), dtype=dt)
        ra0, dec0 = self.patch_reference_coordinates
        for j, wcs in enumerate(wcses):
            for i, s in enumerate(scene.sources):
                ssky = np.array([s.ra + ra0, s.dec + dec0])
                CW_mat, D_mat = scale_at_sky(ssky, wcs)
                self.D[j, i] = D_mat
                self.CW[j, i] = CW_mat
                self.crval[j, i] = ssky - self.patch_reference_coordinates
                self.crpix[j, i] = wcs.all_world2pix(ssky[0], ssky[1], self.wcs_origin)

    def _pack_fluxcal(self):
        self.G = np.asarray(getattr(self, "photocounts", np.ones(self.n_exp)),
                            dtype=self.meta_dtype).copy()

    def _pack_psf(self, scene):
        """psfgauss ordered [exposure][source][n_psf_per_source(band)]
        (forcepho/patches/patch.py:359-398).  Every source of an exposure gets
        that exposure's PSF block, its entries grouped by centre
        (``group_by_centre``): the kernel's sum over entries does not depend on
        their order, and it computes pixel offsets once per run of entries
        with one centre."""
        n_psf = np.array([len(self.psfs[self.band_start[b]]) for b in range(self.n_bands)],
                         dtype=np.int32)
        self.n_psf_per_source = n_psf
        per_exp = n_psf[self.bands]
        self.psfgauss_start = np.zeros(self.n_exp, dtype=np.int32)
        total = int((per_exp * self.n_sources).sum())
        self.psfgauss = np.empty(total, dtype=PSF_DTYPE)
        s = 0
        for e in range(self.n_exp):
            self.psfgauss_start[e] = s
            block = group_by_centre(np.asarray(self.psfs[e], dtype=PSF_DTYPE))
            assert len(block) == per_exp[e], "PSF size must be constant within a band"
            for _ in range(self.n_sources):
                self.psfgauss[s:s + len(block)] = block
                s += len(block)
        assert s == total

    def zerocoords(self, scene, sky_zero=None):
        """Shift source coordinates to offsets from a reference point (median
        of the scene unless given) to keep float32 coordinates precise
        (forcepho/patches/patch.py:427-459)."""
        prior = getattr(scene, "sky_zero", None)
        if prior is not None:
            # already shifted (packing the same scene twice must not shift twice)
            self.patch_reference_coordinates = prior
            return prior
        if sky_zero is None:
            sky_zero = np.array([np.median([s.ra for s in scene.sources]),
                                 np.median([s.dec for s in scene.sources])])
        zero = np.array(sky_zero, dtype=np.float64)
        self.patch_reference_coordinates = zero
        for source in scene.sources:
            source.ra -= zero[0]
            source.dec -= zero[1]
        scene.sky_zero = zero
        return zero

    def unzerocoords(self, scene):
        zero = self.patch_reference_coordinates
        for source in scene.sources:
            source.ra += zero[0]
            source.dec += zero[1]
        scene.sky_zero = None
        return scene
