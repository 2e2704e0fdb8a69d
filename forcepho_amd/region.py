"""Sky regions that select the pixels of a patch (mirrors forcepho/region.py).

``contains(xcorners, ycorners, wcs, origin)`` takes the pixel coordinates of
the four corners of each (super-)pixel, shape (..., 4), and returns the
``np.where`` indices of those with at least one corner inside the region
(forcepho/region.py:91-118, 168-192).  ``bounding_box`` gives the ra, dec of
four corners of a box around the region, shape (2, 4).
"""

import numpy as np

__all__ = ["Region", "CircularRegion", "RectangularRegion"]


class Region:
    """Base class: contains nothing."""

    def contains(self, xcorners, ycorners, wcs, origin=0):
        return np.where(np.zeros(np.shape(xcorners)[:-1], dtype=bool))

    @property
    def bounding_box(self):
        return None


class CircularRegion(Region):
    """A circle of ``radius`` degrees about (ra, dec) degrees.  The radius is
    converted to pixels along the declination axis of the exposure's WCS
    (square pixels assumed, as forcepho/region.py:110-114 does)."""

    def __init__(self, ra, dec, radius):
        self.ra, self.dec, self.radius = ra, dec, radius
        self.shape = "CircularRegion"

    def contains(self, xcorners, ycorners, wcs, origin=0):
        xc, yc = wcs.all_world2pix(self.ra, self.dec, origin)
        xr, yr = wcs.all_world2pix(self.ra, self.dec + self.radius, origin)
        r2 = (xc - xr) ** 2 + (yc - yr) ** 2
        d2 = (xc - np.asarray(xcorners)) ** 2 + (yc - np.asarray(ycorners)) ** 2
        return np.where(np.any(d2 < r2, axis=-1))

    @property
    def bounding_box(self):
        dra = self.radius / np.cos(np.deg2rad(self.dec))
        ddec = self.radius
        return np.array([(self.ra - dra, self.dec - ddec), (self.ra + dra, self.dec - ddec),
                         (self.ra + dra, self.dec + ddec), (self.ra - dra, self.dec + ddec)]).T


class RectangularRegion(Region):
    """ra_min < ra < ra_max and dec_min < dec < dec_max (degrees)."""

    def __init__(self, ra_min, ra_max, dec_min, dec_max):
        self.ra_min, self.ra_max = ra_min, ra_max
        self.dec_min, self.dec_max = dec_min, dec_max
        self.shape = "RectangularRegion"

    def contains(self, xcorners, ycorners, wcs, origin=0):
        ra, dec = wcs.all_pix2world(np.asarray(xcorners), np.asarray(ycorners), origin)
        valid = ((ra > self.ra_min) & (ra < self.ra_max) &
                 (dec > self.dec_min) & (dec < self.dec_max))
        return np.where(np.any(valid, axis=-1))

    @property
    def bounding_box(self):
        return np.array([(self.ra_min, self.dec_min), (self.ra_max, self.dec_min),
                         (self.ra_max, self.dec_max), (self.ra_min, self.dec_max)]).T
