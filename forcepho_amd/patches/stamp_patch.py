"""Pack a list of image stamps (one per exposure) into a Patch.

Mirrors the pixel packing of forcepho/patches/static_patch.py:238-307 and the
padding rule of forcepho/patches/pixel_patch.py:473-484:
  * exposures are concatenated in band order (band ids must be sorted);
  * each exposure is padded to a multiple of ``warp_size`` (64, one CDNA
    wavefront) with flux = 0, ierr = 0 and x = y = -1;
  * band_start = cumsum(band_N) shifted, exposure_start = running offset.
Within an exposure pixels are emitted in superpixel order: 16x16 squares
made of four 8x8 squares, so that one wavefront of 64 lanes covers an 8x8
footprint and a 256-pixel tile a 16x16 one; ``pixel_order="flat"`` keeps
plain row-major order.
"""

import numpy as np

from .patch import Patch, PSF_DTYPE

__all__ = ["Stamp", "StampPatch", "superpixel_order"]


class Stamp:
    """One exposure of pixel data.

    Attributes
    ----------
    band : int          band id (exposures are grouped by it)
    pixel_values : (nx, ny) float
    ierr : (nx, ny) float   inverse uncertainty, 0 masks a pixel
    wcs : LinearWCS
    psf : PSF_DTYPE array   the exposure's PSF mixture, tagged with radius bins
    photocounts : float     flux calibration G of the exposure
    """

    def __init__(self, band, pixel_values, ierr, wcs, psf, photocounts=1.0):
        self.band = int(band)
        self.pixel_values = np.asarray(pixel_values, dtype=np.float32)
        self.ierr = np.asarray(ierr, dtype=np.float32)
        assert self.pixel_values.shape == self.ierr.shape
        self.wcs = wcs
        self.psf = np.asarray(psf, dtype=PSF_DTYPE)
        self.photocounts = float(photocounts)

    @property
    def shape(self):
        return self.pixel_values.shape

    @property
    def npix(self):
        return self.pixel_values.size


def superpixel_order(nx, ny, block=16, sub=8):
    """Flat (row-major) indices of an (nx, ny) image visited in block x block
    squares (row-major), each visited as sub x sub squares (row-major), each
    of those row-major: a run of 64 packed pixels is one 8x8 square and a run
    of 256 one 16x16 square."""
    idx = np.arange(nx * ny).reshape(nx, ny)
    out = []
    for bx in range(0, nx, block):
        for by in range(0, ny, block):
            blk = idx[bx:bx + block, by:by + block]
            for sx in range(0, blk.shape[0], sub):
                for sy in range(0, blk.shape[1], sub):
                    out.append(blk[sx:sx + sub, sy:sy + sub].reshape(-1))
    return np.concatenate(out)


class StampPatch(Patch):
    """Patch built from Stamps; ``build_patch(stamps, scene)`` packs pixels and
    (if a scene is given) metadata."""

    def __init__(self, return_residual=False, pixel_order="superpixel", **kwargs):
        super().__init__(return_residual=return_residual, **kwargs)
        assert pixel_order in ("superpixel", "flat")
        self.pixel_order = pixel_order

    def build_patch(self, stamps, scene=None, sky_zero=None):
        band_ids = [st.band for st in stamps]
        assert (np.diff(band_ids) >= 0).all(), "Stamps must be sorted by band"
        uniq, counts = np.unique(band_ids, return_counts=True)
        self.uniq_bands = uniq
        self.n_exp_per_band = counts
        self.n_bands = len(uniq)
        self.n_exp = len(stamps)
        self.bandlist = [int(b) for b in uniq]
        self.bands = np.searchsorted(uniq, band_ids).astype(np.int32)
        self.wcses = [st.wcs for st in stamps]
        self.psfs = [st.psf for st in stamps]
        self.photocounts = np.array([st.photocounts for st in stamps], dtype=np.float32)
        self.pack_pix(stamps)
        if scene is not None:
            self.pack_meta(scene, sky_zero=sky_zero)
        self._dirty_data = False

    def pack_pix(self, stamps):
        w = self.warp_size
        sizes = np.array([st.npix for st in stamps], dtype=np.int64)
        padded = (sizes + w - 1) // w * w
        total = int(padded.sum())
        dt = self.pix_dtype
        self.original_shapes = np.array([st.shape for st in stamps], dtype=np.int64)
        self.original_sizes = sizes
        self.xpix = np.full(total, -1, dtype=dt)
        self.ypix = np.full(total, -1, dtype=dt)
        self.data = np.zeros(total, dtype=dt)
        self.ierr = np.zeros(total, dtype=dt)
        self.band_start = np.empty(self.n_bands, dtype=np.int16)
        self.band_N = np.zeros(self.n_bands, dtype=np.int16)
        self.exposure_start = np.empty(self.n_exp, dtype=np.int32)
        self.exposure_N = np.empty(self.n_exp, dtype=np.int32)
        self.pixel_index = []
        i = 0
        for e, st in enumerate(stamps):
            self.band_N[self.bands[e]] += 1
            nx, ny = st.shape
            order = (superpixel_order(nx, ny) if self.pixel_order == "superpixel"
                     else np.arange(nx * ny))
            self.pixel_index.append(order)
            xx, yy = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
            n = nx * ny
            self.exposure_start[e] = i
            self.exposure_N[e] = padded[e]
            self.xpix[i:i + n] = xx.reshape(-1)[order]
            self.ypix[i:i + n] = yy.reshape(-1)[order]
            self.data[i:i + n] = st.pixel_values.reshape(-1)[order]
            self.ierr[i:i + n] = st.ierr.reshape(-1)[order]
            i += int(padded[e])
        assert i == total
        self.band_start[0] = 0
        self.band_start[1:] = np.cumsum(self.band_N)[:-1]

    def unpack_image(self, flat, e):
        """Exposure e of a flat pixel array as an (nx, ny) image."""
        nx, ny = self.original_shapes[e]
        start = self.exposure_start[e]
        vals = np.asarray(flat)[start:start + nx * ny]
        img = np.zeros(nx * ny, dtype=vals.dtype)
        img[self.pixel_index[e]] = vals
        return img.reshape(nx, ny)
