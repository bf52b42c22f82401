"""Pixel packers that find exposures and pixels for a sky region
(forcepho/patches/pixel_patch.py).

:class:`PixelPatch` packs, for a region, every overlapping exposure's
contained pixels into the device layout of :class:`.patch.Patch`, and the
scene metadata from the exposures' WCSs, headers and a PSF store.  The
subclasses say where exposures and pixels come from:

* :class:`StorePatch` -- a :class:`.storage.PixelStore` and
  :class:`.storage.MetaStore` (superpixels whose corners fall in the
  region; 64-pixel superpixels need no padding);
* :class:`FITSPatch` -- FITS files with SCI and ERR image extensions
  (single pixels, padded to the warp size with ierr = 0, x = y = -1);
* :class:`DirectPatch` -- pixel arrays and headers given in dictionaries.
"""

import numpy as np

from . import fits
from .patch import Patch
from .storage import MetaStore, PixelStore, PSFStore
from .wcs import WCS

__all__ = ["PixelPatch", "StorePatch", "FITSPatch", "DirectPatch", "JWST_BANDS"]

JWST_BANDS = ["F090W", "F115W", "F150W", "F200W", "F277W", "F335M", "F356W",
              "F410M", "F444W"]


class PixelPatch(Patch):
    """Region-driven packing (pixel_patch.py:23-213).  Subclasses implement
    :meth:`find_exposures` and :meth:`find_pixels`."""

    def __init__(self, psfstore="", splinedata=None, return_residual=False,
                 meta_dtype=np.float32, pix_dtype=np.float32, debug=0):
        super().__init__(splinedata=splinedata or None, return_residual=return_residual,
                         meta_dtype=meta_dtype, pix_dtype=pix_dtype)
        self.psfstore = PSFStore(psfstore) if isinstance(psfstore, str) and psfstore \
            else (psfstore or None)
        self.debug = debug
        self.background_offsets = None
        self.max_snr = None
        self.tweakphot = None

    def build_patch(self, region=None, sourcecat=None, allbands=JWST_BANDS,
                    tweak_background=False, tweakphot=None):
        """Find the exposures overlapping ``region`` (sorted by band in the
        order of ``allbands``), pack their pixels, and, with a source
        catalog, the scene metadata (pixel_patch.py:52-121)."""
        self.hdrs, self.wcses, self.epaths, self.bands = self.find_exposures(region, allbands)
        if len(self.epaths) == 0:
            raise ValueError("No exposures in the specified bands overlap the region")
        if tweak_background:
            self.background_offsets = [h.get(tweak_background, 0.0) for h in self.hdrs]
        else:
            self.background_offsets = None
        self.tweakphot = tweakphot
        allbands = list(allbands)
        band_ids = [allbands.index(b) for b in self.bands]
        assert (np.diff(band_ids) >= 0).all(), "Exposures must be sorted by band"
        u, n = np.unique(band_ids, return_counts=True)
        self.uniq_bands, self.n_exp_per_band = u, n
        self.bandlist = [allbands[i] for i in self.uniq_bands]
        self.n_bands = len(self.uniq_bands)
        self.n_exp = len(self.hdrs)
        self.pack_pix(region=region)
        if sourcecat is not None:
            scene = self.set_scene(sourcecat)
            self.pack_meta(scene)
        self._dirty_data = False

    def pack_pix(self, region, dtype=None):
        """Concatenate every exposure's pixels (pixel_patch.py:123-199)."""
        dtype = dtype or self.pix_dtype
        self.band_start = np.empty(self.n_bands, dtype=np.int16)
        self.band_N = np.zeros(self.n_bands, dtype=np.int16)
        self.exposure_start = np.empty(self.n_exp, dtype=np.int32)
        self.exposure_N = np.empty(self.n_exp, dtype=np.int32)
        self.band_N_pix = np.empty(self.n_bands, dtype=np.int32)
        data, ierr, xpix, ypix = [], [], [], []
        b, i = 0, 0
        for e, wcs in enumerate(self.wcses):
            flux, ie, xp, yp = self.find_pixels(self.epaths[e], wcs, region)
            n_pix = np.size(flux)
            if n_pix == 0:
                raise ValueError(f"There were no valid pixels in exposure {self.epaths[e]}")
            if self.debug > 0 and not np.all(np.isfinite(np.asarray(flux) * np.asarray(ie))):
                raise ValueError(f"There were non-finite pixels in exposure {self.epaths[e]}")
            if self.background_offsets is not None:
                flux = np.asarray(flux) - self.background_offsets[e]
            if e > 0 and self.bands[e] != self.bands[e - 1]:
                b += 1
            self.band_N[b] += 1
            self.exposure_start[e] = i
            self.exposure_N[e] = n_pix
            data.append(np.asarray(flux).reshape(-1))
            ierr.append(np.asarray(ie).reshape(-1))
            xpix.append(np.asarray(xp).reshape(-1))
            ypix.append(np.asarray(yp).reshape(-1))
            i += n_pix
        self.data = np.concatenate(data).astype(dtype)
        self.ierr = np.concatenate(ierr).astype(dtype)
        self.xpix = np.concatenate(xpix).astype(dtype)
        self.ypix = np.concatenate(ypix).astype(dtype)
        self.band_start[0] = 0
        self.band_start[1:] = np.cumsum(self.band_N)[:-1]
        starts = self.exposure_start[self.band_start]
        self.band_N_pix[:-1] = np.diff(starts)
        self.band_N_pix[-1] = self.npix - self.band_N_pix[:-1].sum()
        if self.max_snr:
            self._cap_snr(self.max_snr)
        self._dirty_data = True

    def _cap_snr(self, max_snr=None):
        if max_snr:
            cap = (self.ierr > 0) & (self.data * self.ierr > max_snr)
            self.ierr[cap] = max_snr / self.data[cap]

    def pack_meta(self, scene):
        """Per-exposure PSFs from the PSF store (first detector location, as
        forcepho/patches/patch.py:357-396 asks the store without a WCS) and
        flux calibrations from the header key ``tweakphot`` (default 1,
        patch.py:340-355), then the scene metadata."""
        if self.psfstore is None:
            raise ValueError("a PSF store is needed to pack metadata")
        self.psfs = [self.psfstore.get_local_psf(band=b) for b in self.bands]
        if self.tweakphot:
            self.Gs = [float(h[self.tweakphot]) for h in self.hdrs]
        else:
            self.Gs = None
        super().pack_meta(scene)

    def find_exposures(self, region, bandlist):
        raise NotImplementedError

    def find_pixels(self, epath, wcs, region):
        raise NotImplementedError


class StorePatch(PixelPatch):
    """Pixels from a :class:`PixelStore`, headers and WCSs from a
    :class:`MetaStore` (pixel_patch.py:216-313)."""

    def __init__(self, pixelstore="", metastore="", psfstore="", splinedata=None,
                 return_residual=False, meta_dtype=np.float32, pix_dtype=np.float32,
                 debug=0):
        super().__init__(psfstore=psfstore, splinedata=splinedata,
                         return_residual=return_residual, meta_dtype=meta_dtype,
                         pix_dtype=pix_dtype, debug=debug)
        self.metastore = metastore if isinstance(metastore, MetaStore) else MetaStore(metastore)
        self.pixelstore = pixelstore if isinstance(pixelstore, PixelStore) \
            else PixelStore(pixelstore)

    def find_exposures(self, region, bandlist):
        """Exposures with a corner of the region's bounding box inside the
        image and at least one superpixel in the region.  WCSs are built from
        the stored headers on first use."""
        imsize = np.zeros(2) + self.pixelstore.nside_full
        corners = self.pixelstore.superpixel_corners()
        bra, bdec = region.bounding_box
        epaths, bands, hdrs, wcses = [], [], [], []
        for band in bandlist:
            for expID in self.metastore.headers.get(band, {}):
                wcs = self.metastore.get_wcs(band, expID)
                bx, by = wcs.world_to_pixel_values(bra, bdec)
                inim = np.any((bx > 0) & (bx < imsize[0]) & (by > 0) & (by < imsize[1]))
                if not inim:
                    continue
                sx, sy = region.contains(corners[..., 0], corners[..., 1], wcs)
                if len(sx) == 0:
                    continue
                wcses.append(wcs)
                epaths.append(f"{band}/{expID}")
                bands.append(band)
                hdrs.append(self.metastore.headers[band][expID])
        return hdrs, wcses, epaths, bands

    def find_pixels(self, epath, wcs, region):
        """(flux, ierr, xpix, ypix), each [n_superpixels, s^2], of the
        superpixels with a corner in the region."""
        s2 = self.pixelstore.super_pixel_size ** 2
        corners = self.pixelstore.superpixel_corners()
        sx, sy = region.contains(corners[..., 0], corners[..., 1], wcs)
        data = self.pixelstore.data[epath + "/data"]
        sel = data[sx, sy, :]
        return (sel[:, :s2], sel[:, s2:], self.pixelstore.xpix[sx, sy, :],
                self.pixelstore.ypix[sx, sy, :])


class FITSPatch(PixelPatch):
    """Pixels and headers from FITS files with science and uncertainty
    image extensions (pixel_patch.py:316-483)."""

    def __init__(self, fitsfiles=(), psfstore="", splinedata=None, return_residual=False,
                 sci_ext="SCI", unc_ext="ERR", meta_dtype=np.float32, pix_dtype=np.float32,
                 debug=0):
        super().__init__(psfstore=psfstore, splinedata=splinedata,
                         return_residual=return_residual, meta_dtype=meta_dtype,
                         pix_dtype=pix_dtype, debug=debug)
        self.fitsfiles = list(fitsfiles)
        self.sci_ext = sci_ext
        self.unc_ext = unc_ext
        self.snr = None
        self.unc = None
        self.warp_size = 64

    def find_exposures(self, region, bandlist, do_check=True):
        files = np.array(self.fitsfiles)
        hdr_all = [fits.getheader(f, self.sci_ext) for f in files]
        imbands = np.array([h["FILTER"] for h in hdr_all])
        epaths, bands, hdrs, wcses = [], [], [], []
        for band in bandlist:
            for i in np.where(imbands == band)[0]:
                hdr = hdr_all[i]
                wcs = WCS(hdr)
                if region is not None and do_check:
                    bra, bdec = region.bounding_box
                    nx, ny = hdr["NAXIS1"], hdr["NAXIS2"]
                    bx, by = wcs.world_to_pixel_values(bra, bdec)
                    inim = np.any((bx > 0) & (bx < nx) & (by > 0) & (by < ny))
                    around = ((nx < bx.max()) & (bx.min() < 0) &
                              (ny < by.max()) & (by.min() < 0))
                    if not (inim or around):
                        continue
                epaths.append(str(files[i]))
                wcses.append(wcs)
                bands.append(str(imbands[i]))
                hdrs.append(hdr)
        return hdrs, wcses, epaths, bands

    def find_pixels(self, epath, wcs, region):
        """Pixels with a corner in the region, padded to the warp size.  The
        FITS array is [NAXIS2, NAXIS1] = [y, x]; it is transposed so the
        first index is x."""
        flux = np.array(fits.getdata(epath, self.sci_ext), dtype=np.float64).T
        if self.snr is not None:
            ie = self.snr / flux
        elif self.unc is not None:
            ie = np.zeros_like(flux) + 1.0 / self.unc
        else:
            with np.errstate(divide="ignore"):
                ie = 1.0 / np.array(fits.getdata(epath, self.unc_ext), dtype=np.float64).T
        bad = ~np.isfinite(flux) | ~np.isfinite(ie) | (ie < 0)
        ie[bad] = 0
        flux[bad] = 0
        nx, ny = flux.shape
        xp, yp = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
        if region is None:
            sel = (slice(None), slice(None))
        else:
            off = np.array([(0, 0), (1, 0), (1, 1), (0, 1)], dtype=np.float64)
            xc = xp[..., None] + off[:, 0]
            yc = yp[..., None] + off[:, 1]
            sel = region.contains(xc, yc, wcs)
        flux, ie = flux[sel].reshape(-1), ie[sel].reshape(-1)
        xp, yp = xp[sel].reshape(-1), yp[sel].reshape(-1)
        pad = (-len(flux)) % self.warp_size
        if pad:
            flux = np.append(flux, np.zeros(pad))
            ie = np.append(ie, np.zeros(pad))
            xp = np.append(xp, np.full(pad, -1))
            yp = np.append(yp, np.full(pad, -1))
        return flux, ie, xp, yp


class DirectPatch(PixelPatch):
    """Pixels and headers supplied as dictionaries keyed by exposure name
    (pixel_patch.py:486-550): ``pixel_dict[epath]`` has ``flux, ierr, xpix,
    ypix`` arrays, ``header_dict[epath]`` a header with FILTER and a WCS."""

    def __init__(self, pixel_dict=None, header_dict=None, psfstore="", splinedata=None,
                 return_residual=False, meta_dtype=np.float32, pix_dtype=np.float32,
                 debug=0):
        super().__init__(psfstore=psfstore, splinedata=splinedata,
                         return_residual=return_residual, meta_dtype=meta_dtype,
                         pix_dtype=pix_dtype, debug=debug)
        self.pixel_dict = pixel_dict or {}
        self.header_dict = header_dict or {}

    def find_exposures(self, region, bandlist, do_check=True):
        """Every supplied exposure in ``bandlist``, sorted by band."""
        hdrs, wcses, epaths, bands = [], [], [], []
        for band in bandlist:
            for epath, hdr in self.header_dict.items():
                if hdr["FILTER"] != band:
                    continue
                epaths.append(epath)
                hdrs.append(hdr)
                wcses.append(WCS(hdr))
                bands.append(band)
        return hdrs, wcses, epaths, bands

    def find_pixels(self, epath, wcs, region):
        p = self.pixel_dict[epath]
        return p["flux"], p["ierr"], p["xpix"], p["ypix"]
