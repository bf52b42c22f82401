"""Sky regions that select exposures and superpixels for a patch
(forcepho/region.py).

A region answers two questions for the pixel packers: a celestial bounding
box (to find exposures that may overlap it) and, given an exposure's WCS
and the pixel corners of its superpixels, which superpixels it contains (a
superpixel is in when any of its corners is).
"""

import numpy as np

from .patches.wcs import WCS

__all__ = ["Region", "CircularRegion", "RectangularRegion",
           "polygons_overlap", "polygon_contains"]


class Region:
    """Base class (forcepho/region.py:16-62)."""

    def contains(self, x, y, wcs):
        return np.zeros(len(x), dtype=bool)

    @property
    def bounding_box(self):
        return None

    def overlaps(self, hdr):
        """Whether this region's bounding box meets the celestial bounding
        box of an image with WCS and size from ``hdr``."""
        wcs = WCS(hdr)
        nx, ny = hdr["NAXIS1"], hdr["NAXIS2"]
        px = np.array([0, nx, nx, 0], dtype=np.float64)
        py = np.array([0, 0, ny, ny], dtype=np.float64)
        ira, idec = wcs.pixel_to_world_values(px, py)
        im_bb = RectangularRegion(ira.min(), ira.max(), idec.min(), idec.max())
        return polygons_overlap(self, im_bb)


class CircularRegion(Region):
    """A circle on the sky (forcepho/region.py:65-136): centre (ra, dec)
    and radius, all in degrees."""

    def __init__(self, ra, dec, radius):
        self.ra = ra
        self.dec = dec
        self.radius = radius
        self.shape = "CircularRegion"

    def contains(self, xcorners, ycorners, wcs):
        """Indices (2-tuple from ``np.where``) of the superpixels with a
        corner inside the circle; ``xcorners``/``ycorners`` are
        (nsuper_x, nsuper_y, 4) pixel coordinates.  The radius is measured
        in pixels along the exposure's Dec direction at the centre."""
        xc, yc = wcs.world_to_pixel_values(self.ra, self.dec)
        xr, yr = wcs.world_to_pixel_values(self.ra, self.dec + self.radius)
        r2 = (xc - xr) ** 2 + (yc - yr) ** 2
        d2 = (xc - xcorners) ** 2 + (yc - ycorners) ** 2
        return np.where(np.any(d2 < r2, axis=-1))

    @property
    def bounding_box(self):
        """(2, 4) ra, dec of the corners (SW, SE, NE, NW) of the RA/Dec
        aligned box circumscribing the circle."""
        dra = self.radius / np.cos(np.deg2rad(self.dec))
        ddec = self.radius
        return np.array([(self.ra - dra, self.dec - ddec), (self.ra + dra, self.dec - ddec),
                         (self.ra + dra, self.dec + ddec), (self.ra - dra, self.dec + ddec)]).T


class RectangularRegion(Region):
    """An RA/Dec aligned box (forcepho/region.py:139-211), degrees."""

    def __init__(self, ra_min, ra_max, dec_min, dec_max):
        self.ra_min = ra_min
        self.ra_max = ra_max
        self.dec_min = dec_min
        self.dec_max = dec_max
        self.shape = "RectangularRegion"

    def contains(self, xcorners, ycorners, wcs):
        """Indices of the superpixels with a corner strictly inside the box."""
        ra, dec = wcs.pixel_to_world_values(xcorners, ycorners)
        valid = ((ra > self.ra_min) & (ra < self.ra_max) &
                 (dec > self.dec_min) & (dec < self.dec_max))
        return np.where(np.any(valid, axis=-1))

    @property
    def bounding_box(self):
        return np.array([(self.ra_min, self.dec_min), (self.ra_max, self.dec_min),
                         (self.ra_max, self.dec_max), (self.ra_min, self.dec_max)]).T


def polygons_overlap(a, b):
    """Whether the (convex, consistently ordered) bounding polygons of
    regions ``a`` and ``b`` intersect, by the separating-axis test over the
    edge normals of both (forcepho/region.py:214-262)."""
    ax, ay = a.bounding_box
    bx, by = b.bounding_box
    for X, Y in (a.bounding_box, b.bounding_box):
        n = len(X)
        nxt = (np.arange(n) + 1) % n
        # edge normals
        ux, uy = Y[nxt] - Y, X - X[nxt]
        pa = ux[:, None] * ax[None, :] + uy[:, None] * ay[None, :]
        pb = ux[:, None] * bx[None, :] + uy[:, None] * by[None, :]
        if np.any((pa.max(axis=1) < pb.min(axis=1)) | (pb.max(axis=1) < pa.min(axis=1))):
            return False
    return True


def polygon_contains(poly, points):
    """Whether each point lies on the same side of every edge of ``poly``
    ((nside, 2) vertices, either orientation); ``points`` is (npoint, 2)
    (forcepho/region.py:265-302)."""
    poly = np.asarray(poly, dtype=np.float64)
    points = np.atleast_2d(np.asarray(points, dtype=np.float64))
    edges = np.roll(poly, -1, axis=0) - poly                 # (nside, 2)
    rel = points[:, None, :] - poly[None, :, :]              # (npoint, nside, 2)
    cross = edges[None, :, 0] * rel[..., 1] - edges[None, :, 1] * rel[..., 0]
    s = np.sign(cross)
    return np.all(s == s[:, :1], axis=1)
