"""Sky regions that choose the exposures and superpixels of a patch
(forcepho/region.py).

The pixel packers ask a region two things:

* its celestial bounding polygon (``bounding_box``, corners SW, SE, NE, NW
  as a (2, 4) array of ra, dec), to keep only exposures whose footprint
  meets it (``overlaps``);
* which superpixels of an exposure it takes (``contains``): given the pixel
  coordinates of every superpixel's corners, (nsuper_x, nsuper_y, 4) arrays,
  the ``np.where`` indices of the superpixels with at least one corner
  inside.

Each region here states a single per-corner predicate, ``corner_mask``; the
base class reduces it over the corners.  Both shipped regions have RA/Dec
aligned extents, from which their bounding polygon follows.
"""

import numpy as np

from .patches.wcs import WCS

__all__ = ["Region", "CircularRegion", "RectangularRegion",
           "polygons_overlap", "polygon_contains"]


def _aligned_box(ra_lo, ra_hi, dec_lo, dec_hi):
    """(2, 4) ra, dec of an RA/Dec aligned box, corners SW, SE, NE, NW."""
    return np.array([[ra_lo, ra_hi, ra_hi, ra_lo],
                     [dec_lo, dec_lo, dec_hi, dec_hi]], dtype=np.float64)


class Region:
    """A patch footprint on the sky (forcepho/region.py:16-62)."""

    def __init__(self, *args, **kwargs):
        pass

    def corner_mask(self, xcorners, ycorners, wcs):
        """Boolean array (the shape of the corner arrays): corner inside."""
        return np.zeros(np.shape(xcorners), dtype=bool)

    def contains(self, xcorners, ycorners, wcs):
        """Indices of the superpixels with at least one corner inside."""
        mask = self.corner_mask(np.asarray(xcorners), np.asarray(ycorners), wcs)
        if mask.ndim == 1:
            return mask
        return np.nonzero(mask.any(axis=-1))

    @property
    def extent(self):
        """(ra_lo, ra_hi, dec_lo, dec_hi) of the aligned bounding box."""
        return None

    @property
    def bounding_box(self):
        ext = self.extent
        return None if ext is None else _aligned_box(*ext)

    def overlaps(self, hdr):
        """Does the bounding polygon meet the RA/Dec aligned box around an
        image whose WCS and NAXIS1 x NAXIS2 size come from ``hdr``?"""
        wcs = WCS(hdr)
        w, h = float(hdr["NAXIS1"]), float(hdr["NAXIS2"])
        ra, dec = wcs.pixel_to_world_values(np.array([0.0, w, w, 0.0]),
                                            np.array([0.0, 0.0, h, h]))
        image = RectangularRegion(np.min(ra), np.max(ra), np.min(dec), np.max(dec))
        return polygons_overlap(self, image)


class CircularRegion(Region):
    """A circle of ``radius`` degrees of arc about (ra, dec) in degrees
    (forcepho/region.py:65-136).

    In an exposure the circle becomes the pixel circle about the centre's
    pixel position whose radius is the pixel distance from the centre to the
    point ``radius`` north of it; a corner is inside when strictly closer."""

    def __init__(self, ra, dec, radius):
        self.ra = ra
        self.dec = dec
        self.radius = radius
        self.shape = "CircularRegion"

    def corner_mask(self, xcorners, ycorners, wcs):
        cx, cy = wcs.world_to_pixel_values(self.ra, self.dec)
        nx, ny = wcs.world_to_pixel_values(self.ra, self.dec + self.radius)
        r_pix2 = np.square(nx - cx) + np.square(ny - cy)
        return np.square(xcorners - cx) + np.square(ycorners - cy) < r_pix2

    @property
    def extent(self):
        half_ra = self.radius / np.cos(np.deg2rad(self.dec))
        return (self.ra - half_ra, self.ra + half_ra,
                self.dec - self.radius, self.dec + self.radius)


class RectangularRegion(Region):
    """The RA/Dec aligned box ra_min < ra < ra_max, dec_min < dec < dec_max
    in degrees (forcepho/region.py:139-211); a corner is inside when its sky
    position is strictly inside."""

    def __init__(self, ra_min, ra_max, dec_min, dec_max):
        self.ra_min = ra_min
        self.ra_max = ra_max
        self.dec_min = dec_min
        self.dec_max = dec_max
        self.shape = "RectangularRegion"

    def corner_mask(self, xcorners, ycorners, wcs):
        ra, dec = wcs.pixel_to_world_values(xcorners, ycorners)
        in_ra = (self.ra_min < ra) & (ra < self.ra_max)
        in_dec = (self.dec_min < dec) & (dec < self.dec_max)
        return in_ra & in_dec

    @property
    def extent(self):
        return (self.ra_min, self.ra_max, self.dec_min, self.dec_max)


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
