"""World coordinate systems for the packers.

The packers need, per exposure, the sky-to-pixel map at each source
(``world_to_pixel_values``) and its local linear approximation
(:func:`scale_at_sky`, forcepho/slow/stamp.py:200-256).  The reference
takes these from astropy's WCS; this package does not depend on astropy, so this
module implements the gnomonic (TAN) projection of the FITS WCS standard
(Calabretta & Greisen 2002, A&A 395, 1077, eq. 54 with the rotation of
eqs. 2 and 5) with a CD (or PC x CDELT) matrix, in the equivalent
tangent-plane form (standard coordinates xi, eta, rotated by LONPOLE - 180
degrees), which stays exact near the reference point; and
:class:`AffineWCS`, the linear map the synthetic workloads use.

Pixel coordinates of the ``*_values`` methods are 0-based, as astropy's
(the FITS CRPIX is 1-based).  SIP or other distortion terms are not
implemented: a header that carries them is refused rather than read
without them.
"""

import numpy as np

__all__ = ["AffineWCS", "TanWCS", "WCS", "scale_at_sky", "local_affine"]

_R2D = 180.0 / np.pi


class AffineWCS:
    """A linear world coordinate system without distortion.

    pixel = crpix + CD^-1 [(ra - ra0) cos(dec0), dec - dec0]  (0-based pixels)
    where ``cd`` (degrees per pixel) plays the role of astropy's
    ``pixel_scale_matrix``.
    """

    has_distortion = False

    def __init__(self, crpix, crval, cd):
        self.crpix = np.asarray(crpix, dtype=np.float64)
        self.crval = np.asarray(crval, dtype=np.float64)
        self.cd = np.asarray(cd, dtype=np.float64)
        self._cdinv = np.linalg.inv(self.cd)

    @property
    def pixel_scale_matrix(self):
        return self.cd

    def world_to_pixel_values(self, ra, dec):
        dra = (np.asarray(ra) - self.crval[0]) * np.cos(np.deg2rad(self.crval[1]))
        ddec = np.asarray(dec) - self.crval[1]
        ci = self._cdinv
        x = ci[0, 0] * dra + ci[0, 1] * ddec + self.crpix[0]
        y = ci[1, 0] * dra + ci[1, 1] * ddec + self.crpix[1]
        return x, y

    def pixel_to_world_values(self, x, y):
        dx = np.asarray(x, dtype=np.float64) - self.crpix[0]
        dy = np.asarray(y, dtype=np.float64) - self.crpix[1]
        dra = self.cd[0, 0] * dx + self.cd[0, 1] * dy
        ddec = self.cd[1, 0] * dx + self.cd[1, 1] * dy
        return self.crval[0] + dra / np.cos(np.deg2rad(self.crval[1])), self.crval[1] + ddec

    def all_world2pix(self, sky, origin):
        sky = np.atleast_2d(sky)
        x, y = self.world_to_pixel_values(sky[:, 0], sky[:, 1])
        return np.stack([x, y], axis=-1) + origin

    def all_pix2world(self, pix, origin):
        pix = np.atleast_2d(pix) - origin
        ra, dec = self.pixel_to_world_values(pix[:, 0], pix[:, 1])
        return np.stack([ra, dec], axis=-1)


class TanWCS:
    """Gnomonic projection, RA---TAN / DEC--TAN.

    Parameters
    ----------
    crpix : (2,) reference pixel, FITS 1-based
    crval : (2,) (ra, dec) of the reference point, degrees
    cd : (2, 2) degrees per pixel (CD_ij = PC_ij CDELT_i)
    lonpole : native longitude of the celestial pole, degrees (180 unless
        the reference point is the pole)
    naxis : optional (NAXIS1, NAXIS2) image size
    """

    has_distortion = False

    def __init__(self, crpix, crval, cd, lonpole=None, naxis=None):
        self.crpix = np.asarray(crpix, dtype=np.float64)
        self.crval = np.asarray(crval, dtype=np.float64)
        self.cd = np.asarray(cd, dtype=np.float64).reshape(2, 2)
        self._cdinv = np.linalg.inv(self.cd)
        if lonpole is None:
            lonpole = 0.0 if self.crval[1] >= 90.0 else 180.0
        self.lonpole = float(lonpole)
        self.naxis = None if naxis is None else tuple(int(n) for n in naxis)

    @classmethod
    def from_header(cls, hdr):
        """From FITS WCS keywords (axes 1 and 2; further axes ignored)."""
        c1 = str(hdr.get("CTYPE1", "")).strip().upper()
        c2 = str(hdr.get("CTYPE2", "")).strip().upper()
        if not (c1.startswith("RA") and c2.startswith("DEC")):
            raise ValueError(f"celestial axes RA/DEC expected, got {c1!r}, {c2!r}")
        if c1.endswith("-SIP") or c2.endswith("-SIP") or "A_ORDER" in hdr:
            raise NotImplementedError("SIP distortion terms are not implemented")
        if not (c1.endswith("-TAN") and c2.endswith("-TAN")):
            raise NotImplementedError(f"projection {c1[5:]!r} is not implemented (TAN only)")
        crpix = [float(hdr["CRPIX1"]), float(hdr["CRPIX2"])]
        crval = [float(hdr["CRVAL1"]), float(hdr["CRVAL2"])]
        if any(k in hdr for k in ("CD1_1", "CD1_2", "CD2_1", "CD2_2")):
            cd = [[float(hdr.get("CD1_1", 0.0)), float(hdr.get("CD1_2", 0.0))],
                  [float(hdr.get("CD2_1", 0.0)), float(hdr.get("CD2_2", 0.0))]]
        else:
            pc = np.array([[float(hdr.get("PC1_1", 1.0)), float(hdr.get("PC1_2", 0.0))],
                           [float(hdr.get("PC2_1", 0.0)), float(hdr.get("PC2_2", 1.0))]])
            cdelt = np.array([float(hdr.get("CDELT1", 1.0)), float(hdr.get("CDELT2", 1.0))])
            cd = pc * cdelt[:, None]
        naxis = None
        if "NAXIS1" in hdr and "NAXIS2" in hdr:
            naxis = (hdr["NAXIS1"], hdr["NAXIS2"])
        return cls(crpix, crval, cd, lonpole=hdr.get("LONPOLE", None), naxis=naxis)

    def to_header(self):
        """The WCS keywords of this projection."""
        h = {"CTYPE1": "RA---TAN", "CTYPE2": "DEC--TAN",
             "CRPIX1": float(self.crpix[0]), "CRPIX2": float(self.crpix[1]),
             "CRVAL1": float(self.crval[0]), "CRVAL2": float(self.crval[1]),
             "CD1_1": float(self.cd[0, 0]), "CD1_2": float(self.cd[0, 1]),
             "CD2_1": float(self.cd[1, 0]), "CD2_2": float(self.cd[1, 1]),
             "LONPOLE": self.lonpole}
        if self.naxis is not None:
            h["NAXIS1"], h["NAXIS2"] = self.naxis
        return h

    @property
    def pixel_scale_matrix(self):
        return self.cd

    def _rotate(self, xi, eta, sign):
        """Rotate tangent-plane coordinates by sign x (LONPOLE - 180 deg)."""
        t = np.deg2rad(self.lonpole - 180.0) * sign
        if t == 0.0:
            return xi, eta
        c, s = np.cos(t), np.sin(t)
        return xi * c - eta * s, eta * c + xi * s

    def pixel_to_world_values(self, x, y):
        """(ra, dec) in degrees of 0-based pixel coordinates."""
        dx = np.asarray(x, dtype=np.float64) + 1.0 - self.crpix[0]
        dy = np.asarray(y, dtype=np.float64) + 1.0 - self.crpix[1]
        xi = self.cd[0, 0] * dx + self.cd[0, 1] * dy     # intermediate world, deg
        eta = self.cd[1, 0] * dx + self.cd[1, 1] * dy
        xi, eta = self._rotate(np.deg2rad(xi), np.deg2rad(eta), -1.0)
        # inverse gnomonic projection in tangent-plane form (no arcsin, so
        # exact near the reference point)
        a0, d0 = np.deg2rad(self.crval)
        den = np.cos(d0) - eta * np.sin(d0)
        ra = a0 + np.arctan2(xi, den)
        dec = np.arctan2(np.sin(d0) + eta * np.cos(d0), np.hypot(xi, den))
        return np.mod(np.rad2deg(ra), 360.0), np.rad2deg(dec)

    def world_to_pixel_values(self, ra, dec):
        """0-based pixel coordinates of (ra, dec) in degrees."""
        a = np.deg2rad(np.asarray(ra, dtype=np.float64))
        d = np.deg2rad(np.asarray(dec, dtype=np.float64))
        a0, d0 = np.deg2rad(self.crval)
        da = a - a0
        cosc = np.sin(d0) * np.sin(d) + np.cos(d0) * np.cos(d) * np.cos(da)
        xi = np.cos(d) * np.sin(da) / cosc
        eta = (np.cos(d0) * np.sin(d) - np.sin(d0) * np.cos(d) * np.cos(da)) / cosc
        xi, eta = self._rotate(xi, eta, 1.0)
        xi, eta = np.rad2deg(xi), np.rad2deg(eta)
        ci = self._cdinv
        x = ci[0, 0] * xi + ci[0, 1] * eta + self.crpix[0] - 1.0
        y = ci[1, 0] * xi + ci[1, 1] * eta + self.crpix[1] - 1.0
        return x, y

    def all_world2pix(self, sky, origin):
        sky = np.atleast_2d(sky)
        x, y = self.world_to_pixel_values(sky[:, 0], sky[:, 1])
        return np.stack([x, y], axis=-1) + origin

    def all_pix2world(self, pix, origin):
        pix = np.atleast_2d(np.asarray(pix, dtype=np.float64)) - origin
        ra, dec = self.pixel_to_world_values(pix[:, 0], pix[:, 1])
        return np.stack([ra, dec], axis=-1)


def WCS(hdr):
    """A WCS object from a FITS header (TAN projection)."""
    return TanWCS.from_header(hdr)


def scale_at_sky(sky, wcs, dpix=1.0, origin=1, make_approx=False):
    """Local linear approximation of ``wcs`` at ``sky`` = (ra, dec) degrees
    (forcepho/slow/stamp.py:200-256).

    Returns (CW, D): CW (pixels per degree) maps (dra, ddec) in degrees to
    (dx, dy) pixels; D is pixels per arcsec of (RA offset on the sky, Dec).
    Without distortion (and ``make_approx`` false) D = (3600 CD)^-1 and
    CW = 3600 D diag(cos dec, 1); otherwise both come from finite
    differences of ``dpix`` pixels at the source.
    """
    ra, dec = sky
    if getattr(wcs, "has_distortion", False) or make_approx:
        pos0_sky = np.array([ra, dec], dtype=np.float64)
        pos0_pix = wcs.all_world2pix([pos0_sky], origin)[0]
        pos1_sky = wcs.all_pix2world([pos0_pix + np.array([dpix, 0.0])], origin)[0]
        pos2_sky = wcs.all_pix2world([pos0_pix + np.array([0.0, dpix])], origin)[0]
        St = np.array([pos1_sky - pos0_sky, pos2_sky - pos0_sky])
        CW_mat = np.linalg.solve(St, np.eye(2) * dpix).T
        Winv = np.eye(2)
        Winv[0, 0] = 1.0 / np.cos(np.deg2rad(pos0_sky[-1]))
        D_mat = CW_mat @ Winv / 3600.0
    else:
        W = np.eye(2)
        W[0, 0] = np.cos(np.deg2rad(dec))
        D_mat = np.linalg.inv(wcs.pixel_scale_matrix * 3600.0)
        CW_mat = (D_mat * 3600.0) @ W
    return CW_mat, D_mat


def local_affine(wcs, ra, dec):
    """:func:`scale_at_sky` at many sky positions at once: (CW, D), each
    [n, 2, 2], for ``ra``, ``dec`` arrays of length n (degrees).  Without
    distortion D is the same at every position and CW differs only by
    cos(dec); with distortion each position takes the finite-difference
    branch."""
    ra = np.atleast_1d(np.asarray(ra, dtype=np.float64))
    dec = np.atleast_1d(np.asarray(dec, dtype=np.float64))
    n = len(ra)
    if getattr(wcs, "has_distortion", False):
        pairs = [scale_at_sky((ra[i], dec[i]), wcs) for i in range(n)]
        return (np.array([p[0] for p in pairs]).reshape(n, 2, 2),
                np.array([p[1] for p in pairs]).reshape(n, 2, 2))
    D = np.linalg.inv(np.asarray(wcs.pixel_scale_matrix, dtype=np.float64) * 3600.0)
    CW = np.empty((n, 2, 2))
    CW[:, :, 0] = 3600.0 * D[None, :, 0] * np.cos(np.deg2rad(dec))[:, None]
    CW[:, :, 1] = 3600.0 * D[None, :, 1]
    return CW, np.broadcast_to(D, (n, 2, 2)).copy()
