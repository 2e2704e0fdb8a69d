"""Gnomonic (TAN) world coordinate systems read from FITS headers.

forcepho takes each exposure's astrometry from an ``astropy.wcs.WCS`` of the
image header and uses three of its methods: ``all_world2pix`` (source pixel
positions, ``patches/patch.py:335``), ``all_pix2world`` and
``pixel_scale_matrix`` (the local linearisation ``scale_at_sky``,
``patches/patch.py:504-560``), plus ``has_distortion``.  ``TanWCS``
provides the same methods for the ``RA---TAN``/``DEC--TAN`` headers of
undistorted images, from the gnomonic projection of Calabretta & Greisen
(2002, A&A 395, 1077), section 5.1.3, with the native pole at LONPOLE = 180
degrees (the default for zenithal projections):

    (xi, eta) = CD (p - CRPIX)                          [degrees]
    ra  = ra0 + atan2(xi, cos dec0 - eta sin dec0)
    dec = atan2(sin dec0 + eta cos dec0, hypot(xi, cos dec0 - eta sin dec0))

with CD = diag(CDELT) PC (or the CDi_j cards), and its inverse

    cos c = sin dec sin dec0 + cos dec cos dec0 cos(ra - ra0)
    xi  = cos dec sin(ra - ra0) / cos c
    eta = (sin dec cos dec0 - cos dec sin dec0 cos(ra - ra0)) / cos c.

Pixel coordinates follow astropy's ``origin`` argument: 0 for 0-based, 1 for
the 1-based FITS convention.  Headers with distortion terms (SIP, TPV, ...)
are refused rather than approximated.
"""

import numpy as np

__all__ = ["TanWCS", "WCSError"]


class WCSError(ValueError):
    """The header does not describe an undistorted TAN projection."""


class TanWCS:
    """World <-> pixel map of an undistorted gnomonic projection."""

    has_distortion = False

    def __init__(self, header):
        ctype = (str(header.get("CTYPE1", "")).strip(), str(header.get("CTYPE2", "")).strip())
        if ctype != ("RA---TAN", "DEC--TAN"):
            raise WCSError("only RA---TAN / DEC--TAN headers are supported, got %s" % (ctype,))
        for key in ("A_ORDER", "B_ORDER", "PV1_1", "PV2_1", "PV2_2"):
            if key in header:
                raise WCSError("distortion keyword %s is not supported" % key)
        lonpole = float(header.get("LONPOLE", 180.0))
        if lonpole != 180.0:
            raise WCSError("LONPOLE %g is not supported" % lonpole)
        self.crpix = np.array([float(header.get("CRPIX1", 0.0)), float(header.get("CRPIX2", 0.0))])
        self.crval = np.array([float(header["CRVAL1"]), float(header["CRVAL2"])])
        if "CD1_1" in header or "CD2_2" in header:
            cd = np.array([[header.get("CD1_1", 0.0), header.get("CD1_2", 0.0)],
                           [header.get("CD2_1", 0.0), header.get("CD2_2", 0.0)]], dtype=np.float64)
        else:
            pc = np.array([[header.get("PC1_1", 1.0), header.get("PC1_2", 0.0)],
                           [header.get("PC2_1", 0.0), header.get("PC2_2", 1.0)]], dtype=np.float64)
            cdelt = np.array([header.get("CDELT1", 1.0), header.get("CDELT2", 1.0)],
                             dtype=np.float64)
            cd = cdelt[:, None] * pc
        self.cd = cd
        self._cd_inv = np.linalg.inv(cd)

    @property
    def pixel_scale_matrix(self):
        """Degrees per pixel, CD = diag(CDELT) PC."""
        return self.cd.copy()

    def all_pix2world(self, x, y, origin):
        """Pixel coordinates -> (ra, dec) in degrees."""
        x = np.asarray(x, dtype=np.float64)
        y = np.asarray(y, dtype=np.float64)
        dx = x + (1 - origin) - self.crpix[0]
        dy = y + (1 - origin) - self.crpix[1]
        xi = np.deg2rad(self.cd[0, 0] * dx + self.cd[0, 1] * dy)
        eta = np.deg2rad(self.cd[1, 0] * dx + self.cd[1, 1] * dy)
        ra0, dec0 = np.deg2rad(self.crval)
        den = np.cos(dec0) - eta * np.sin(dec0)
        ra = ra0 + np.arctan2(xi, den)
        dec = np.arctan2(np.sin(dec0) + eta * np.cos(dec0), np.hypot(xi, den))
        return np.rad2deg(ra) % 360.0, np.rad2deg(dec)

    def all_world2pix(self, ra, dec, origin):
        """(ra, dec) in degrees -> pixel coordinates."""
        ra = np.deg2rad(np.asarray(ra, dtype=np.float64))
        dec = np.deg2rad(np.asarray(dec, dtype=np.float64))
        ra0, dec0 = np.deg2rad(self.crval)
        dra = ra - ra0
        cosc = np.sin(dec) * np.sin(dec0) + np.cos(dec) * np.cos(dec0) * np.cos(dra)
        xi = np.rad2deg(np.cos(dec) * np.sin(dra) / cosc)
        eta = np.rad2deg((np.sin(dec) * np.cos(dec0) - np.cos(dec) * np.sin(dec0) * np.cos(dra))
                         / cosc)
        ci = self._cd_inv
        px = ci[0, 0] * xi + ci[0, 1] * eta + self.crpix[0] - (1 - origin)
        py = ci[1, 0] * xi + ci[1, 1] * eta + self.crpix[1] - (1 - origin)
        return px, py
