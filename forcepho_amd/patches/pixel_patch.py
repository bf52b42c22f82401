"""Patches packed from survey exposures for a sky region
(forcepho/patches/pixel_patch.py).

A :class:`PixelPatch` is built in two steps.  ``find_exposures(region,
bandlist)`` lists the exposures that overlap the region (their headers,
WCSs, paths and bands, in band order); ``find_pixels(epath, wcs, region)``
gives each exposure's selected pixels as flat (flux, ierr, x, y) arrays.
The pixel lists are then laid out by :func:`.patch.layout_exposures`, the
same routine that packs synthetic images, and the scene metadata comes from
the exposures' WCSs, a PSF store and optional header flux calibrations.

Where the pixels come from:

* :class:`StorePatch` -- a :class:`.storage.PixelStore` of superpixels and
  a :class:`.storage.MetaStore` of headers; a superpixel is taken when one
  of its corners is in the region.  The selection made while looking for
  exposures is kept and reused for the pixels.
* :class:`FITSPatch` -- FITS files with science and uncertainty extensions;
  single pixels, padded to the warp size with ierr = 0 and x = y = -1.
* :class:`DirectPatch` -- pixel arrays and headers handed over in
  dictionaries.
"""

import functools

import numpy as np

from . import fits
from .patch import Patch, layout_exposures
from .storage import MetaStore, PixelStore, PSFStore
from .wcs import WCS

__all__ = ["PixelPatch", "StorePatch", "FITSPatch", "DirectPatch", "JWST_BANDS"]

JWST_BANDS = ["F090W", "F115W", "F150W", "F200W", "F277W", "F335M", "F356W",
              "F410M", "F444W"]

# unit-square corners of a pixel, counter-clockwise from its origin
_PIXEL_CORNERS = np.array([(0.0, 0.0), (1.0, 0.0), (1.0, 1.0), (0.0, 1.0)])


def _pixels_of_box(bx, by, nx, ny):
    """Is any of the sky box's corners, projected to pixels (bx, by),
    strictly inside an nx x ny image?"""
    bx, by = np.asarray(bx), np.asarray(by)
    return bool(np.any((0 < bx) & (bx < nx) & (0 < by) & (by < ny)))


def _box_surrounds(bx, by, nx, ny):
    """Do the projected sky box corners enclose the whole image?"""
    bx, by = np.asarray(bx), np.asarray(by)
    return bool((bx.min() < 0) & (nx < bx.max()) & (by.min() < 0) & (ny < by.max()))


@functools.lru_cache(maxsize=16)
def _pixel_grid(nx, ny):
    """x, y index grids of an [nx, ny] image (x the first index) and the
    pixel-corner coordinates, (nx, ny, 4) each."""
    x, y = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
    xc = x[..., None] + _PIXEL_CORNERS[:, 0]
    yc = y[..., None] + _PIXEL_CORNERS[:, 1]
    return x, y, xc, yc


class PixelPatch(Patch):
    """Region-driven packing (pixel_patch.py:23-213).  Subclasses provide
    :meth:`find_exposures` and :meth:`find_pixels`."""

    def __init__(self, psfstore="", splinedata=None, return_residual=False,
                 meta_dtype=np.float32, pix_dtype=np.float32, debug=0, spline_smoothing=None):
        super().__init__(splinedata=splinedata or None, return_residual=return_residual,
                         meta_dtype=meta_dtype, pix_dtype=pix_dtype,
                         spline_smoothing=spline_smoothing)
        if isinstance(psfstore, str):
            psfstore = PSFStore(psfstore) if psfstore else None
        self.psfstore = psfstore
        self.debug = debug
        self.background_offsets = None
        self.max_snr = None
        self.tweakphot = None

    # -- building -----------------------------------------------------------
    def build_patch(self, region=None, sourcecat=None, allbands=JWST_BANDS,
                    tweak_background=False, tweakphot=None):
        """Pack the pixels of every exposure in ``allbands`` that overlaps
        ``region`` and, given a source catalog, the scene metadata
        (pixel_patch.py:52-121).  ``tweak_background`` names a header
        keyword whose value is subtracted from that exposure's pixels;
        ``tweakphot`` names the header keyword of a flux calibration G."""
        found = self.find_exposures(region, allbands)
        self.hdrs, self.wcses, self.epaths, self.bands = (list(f) for f in found)
        if not self.epaths:
            raise ValueError("No exposures in the specified bands overlap the region")
        rank = [list(allbands).index(b) for b in self.bands]
        if np.any(np.diff(rank) < 0):
            raise ValueError("Exposures must be sorted by band")
        self.uniq_bands, self.n_exp_per_band = np.unique(rank, return_counts=True)
        self.bandlist = [allbands[i] for i in self.uniq_bands]
        self.n_bands = len(self.bandlist)
        self.n_exp = len(self.epaths)
        self.background_offsets = ([h.get(tweak_background, 0.0) for h in self.hdrs]
                                   if tweak_background else None)
        self.tweakphot = tweakphot
        self.pack_pix(region=region)
        if sourcecat is not None:
            self.pack_meta(self.set_scene(sourcecat))
        self._dirty_data = False

    def _exposure_pixels(self, e, region):
        """(flux, ierr, x, y) of exposure e, background offset applied."""
        flux, ierr, x, y = (np.asarray(a).reshape(-1)
                            for a in self.find_pixels(self.epaths[e], self.wcses[e], region))
        if flux.size == 0:
            raise ValueError(f"There were no valid pixels in exposure {self.epaths[e]}")
        if self.debug > 0 and not np.isfinite(flux * ierr).all():
            raise ValueError(f"There were non-finite pixels in exposure {self.epaths[e]}")
        if self.background_offsets is not None:
            flux = flux - self.background_offsets[e]
        return flux, ierr, x, y

    def pack_pix(self, region, dtype=None):
        """Lay out every exposure's pixels (pixel_patch.py:123-199)."""
        band_index = [self.bandlist.index(b) for b in self.bands]
        chunks = [self._exposure_pixels(e, region) for e in range(len(self.epaths))]
        self.__dict__.update(layout_exposures(chunks, band_index,
                                              dtype=dtype or self.pix_dtype))
        if self.max_snr:
            self._cap_snr(self.max_snr)
        self._dirty_data = True

    def _cap_snr(self, max_snr=None):
        """Lower ierr so that no pixel's flux * ierr exceeds ``max_snr``."""
        if not max_snr:
            return
        snr = self.data * self.ierr
        over = (self.ierr > 0) & (snr > max_snr)
        self.ierr = np.where(over, max_snr / np.where(over, self.data, 1), self.ierr) \
            .astype(self.ierr.dtype)

    def pack_meta(self, scene):
        """PSFs from the PSF store (its first detector location, as
        forcepho/patches/patch.py:357-396 asks the store without a WCS), the
        flux calibration G of each exposure from header key ``tweakphot``
        (1 without one, patch.py:340-355), then the scene metadata."""
        if self.psfstore is None:
            raise ValueError("a PSF store is needed to pack metadata")
        self.psfs = [self.psfstore.get_local_psf(band=b) for b in self.bands]
        self.Gs = [float(h[self.tweakphot]) for h in self.hdrs] if self.tweakphot else None
        super().pack_meta(scene)

    # -- sources of exposures and pixels -------------------------------------
    def find_exposures(self, region, bandlist):
        """(hdrs, wcses, epaths, bands) of the overlapping exposures."""
        raise NotImplementedError

    def find_pixels(self, epath, wcs, region):
        """(flux, ierr, xpix, ypix) of the exposure's pixels in the region."""
        raise NotImplementedError


class StorePatch(PixelPatch):
    """Pixels from a :class:`PixelStore`, headers and WCSs from a
    :class:`MetaStore` (pixel_patch.py:216-313)."""

    def __init__(self, pixelstore="", metastore="", psfstore="", splinedata=None,
                 return_residual=False, meta_dtype=np.float32, pix_dtype=np.float32,
                 debug=0, spline_smoothing=None):
        super().__init__(psfstore=psfstore, splinedata=splinedata,
                         return_residual=return_residual, meta_dtype=meta_dtype,
                         pix_dtype=pix_dtype, debug=debug,
                         spline_smoothing=spline_smoothing)
        self.metastore = metastore if isinstance(metastore, MetaStore) else MetaStore(metastore)
        self.pixelstore = pixelstore if isinstance(pixelstore, PixelStore) \
            else PixelStore(pixelstore)
        self._selection = {}

    def _superpixels_in(self, wcs, region):
        """(ix, iy) of the superpixels with a corner in the region."""
        corners = self.pixelstore.superpixel_corners()
        return region.contains(corners[..., 0], corners[..., 1], wcs)

    def find_exposures(self, region, bandlist):
        """Exposures onto which a corner of the region's bounding box
        projects and which have a superpixel in the region; the superpixel
        selection is kept for :meth:`find_pixels`."""
        nx, ny = np.zeros(2) + self.pixelstore.nside_full
        box_ra, box_dec = region.bounding_box
        self._selection = {}
        found = []
        for band in bandlist:
            for expID, hdr in self.metastore.headers.get(band, {}).items():
                wcs = self.metastore.get_wcs(band, expID)
                if not _pixels_of_box(*wcs.world_to_pixel_values(box_ra, box_dec), nx, ny):
                    continue
                sel = self._superpixels_in(wcs, region)
                if len(sel[0]) == 0:
                    continue
                epath = f"{band}/{expID}"
                self._selection[epath] = sel
                found.append((hdr, wcs, epath, band))
        return tuple(list(col) for col in zip(*found)) if found else ([], [], [], [])

    def find_pixels(self, epath, wcs, region):
        """Flux, ierr and pixel coordinates of the selected superpixels,
        each [n_superpixels, s^2]."""
        sel = self._selection.get(epath)
        if sel is None:
            sel = self._superpixels_in(wcs, region)
        s2 = self.pixelstore.super_pixel_size ** 2
        block = self.pixelstore.data[epath + "/data"][sel[0], sel[1], :]
        return (block[:, :s2], block[:, s2:], self.pixelstore.xpix[sel],
                self.pixelstore.ypix[sel])


class FITSPatch(PixelPatch):
    """Pixels and headers from FITS files with science and uncertainty image
    extensions (pixel_patch.py:316-483).  ``snr`` or ``unc``, when set,
    replace the uncertainty extension by flux/snr or a constant."""

    def __init__(self, fitsfiles=(), psfstore="", splinedata=None, return_residual=False,
                 sci_ext="SCI", unc_ext="ERR", meta_dtype=np.float32, pix_dtype=np.float32,
                 debug=0, spline_smoothing=None):
        super().__init__(psfstore=psfstore, splinedata=splinedata,
                         return_residual=return_residual, meta_dtype=meta_dtype,
                         pix_dtype=pix_dtype, debug=debug,
                         spline_smoothing=spline_smoothing)
        self.fitsfiles = list(fitsfiles)
        self.sci_ext = sci_ext
        self.unc_ext = unc_ext
        self.snr = None
        self.unc = None
        self.warp_size = 64

    def _overlaps(self, hdr, wcs, region):
        """A corner of the region's bounding box inside the image, or the
        box around the whole image."""
        box = wcs.world_to_pixel_values(*region.bounding_box)
        nx, ny = hdr["NAXIS1"], hdr["NAXIS2"]
        return _pixels_of_box(*box, nx, ny) or _box_surrounds(*box, nx, ny)

    def find_exposures(self, region, bandlist, do_check=True):
        headers = [(str(f), fits.getheader(f, self.sci_ext)) for f in self.fitsfiles]
        found = []
        for band in bandlist:
            for path, hdr in headers:
                if hdr["FILTER"] != band:
                    continue
                wcs = WCS(hdr)
                if region is not None and do_check and not self._overlaps(hdr, wcs, region):
                    continue
                found.append((hdr, wcs, path, str(hdr["FILTER"])))
        return tuple(list(col) for col in zip(*found)) if found else ([], [], [], [])

    def _inverse_errors(self, epath, flux):
        if self.snr is not None:
            return self.snr / flux
        if self.unc is not None:
            return np.full_like(flux, 1.0 / self.unc)
        unc = np.array(fits.getdata(epath, self.unc_ext), dtype=np.float64).T
        with np.errstate(divide="ignore"):
            return 1.0 / unc

    def find_pixels(self, epath, wcs, region):
        """Pixels with a corner in the region, bad ones (non-finite, or
        ierr < 0) with flux = ierr = 0, padded to the warp size.  FITS arrays
        are [NAXIS2, NAXIS1] = [y, x]; they are transposed so that the first
        index is x."""
        flux = np.array(fits.getdata(epath, self.sci_ext), dtype=np.float64).T
        with np.errstate(divide="ignore", invalid="ignore"):
            ierr = self._inverse_errors(epath, flux)
            good = np.isfinite(flux) & np.isfinite(ierr) & (ierr >= 0)
        flux, ierr = np.where(good, flux, 0.0), np.where(good, ierr, 0.0)
        x, y, xc, yc = _pixel_grid(*flux.shape)
        sel = (Ellipsis,) if region is None else region.contains(xc, yc, wcs)
        take = [a[sel].reshape(-1) for a in (flux, ierr, x, y)]
        pad = (-take[0].size) % self.warp_size
        fill = (0.0, 0.0, -1, -1)
        return tuple(np.concatenate([a, np.full(pad, f, dtype=a.dtype)])
                     for a, f in zip(take, fill))


class DirectPatch(PixelPatch):
    """Pixels and headers supplied as dictionaries keyed by exposure name
    (pixel_patch.py:486-550): ``pixel_dict[epath]`` has ``flux, ierr, xpix,
    ypix`` arrays, ``header_dict[epath]`` a header with FILTER and a WCS."""

    def __init__(self, pixel_dict=None, header_dict=None, psfstore="", splinedata=None,
                 return_residual=False, meta_dtype=np.float32, pix_dtype=np.float32,
                 debug=0, spline_smoothing=None):
        super().__init__(psfstore=psfstore, splinedata=splinedata,
                         return_residual=return_residual, meta_dtype=meta_dtype,
                         pix_dtype=pix_dtype, debug=debug,
                         spline_smoothing=spline_smoothing)
        self.pixel_dict = pixel_dict or {}
        self.header_dict = header_dict or {}

    def find_exposures(self, region, bandlist, do_check=True):
        """Every supplied exposure in ``bandlist``, in band order."""
        found = [(hdr, WCS(hdr), epath, band) for band in bandlist
                 for epath, hdr in self.header_dict.items() if hdr["FILTER"] == band]
        return tuple(list(col) for col in zip(*found)) if found else ([], [], [], [])

    def find_pixels(self, epath, wcs, region):
        p = self.pixel_dict[epath]
        return p["flux"], p["ierr"], p["xpix"], p["ypix"]
