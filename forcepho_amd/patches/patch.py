"""Patch packers: pixel data in superpixel order and scene metadata.

Mirrors forcepho/patches/patch.py (``PatchBase``, ``Patch.pack_meta`` and its
helpers, lines 20-499) and the pixel layout of
forcepho/patches/static_patch.py:238-307 / pixel_patch.py:124-202:

* all exposures concatenated into 1-D float32 ``data, ierr, xpix, ypix``,
  band-sorted; every exposure padded to a multiple of ``warp_size`` with
  ierr = 0, data = 0 and x = y = -1 (pixel_patch.py:472-483);
* ``band_start/band_N`` int16 [n_bands], ``exposure_start/exposure_N`` int32
  [n_exp];
* per (exposure, source): D (pixels per arcsec), CW (pixels per degree),
  crpix, crval, from :func:`scale_at_sky` (forcepho/slow/stamp.py:200-256);
* ``psfgauss`` laid out [exposure][source][n_psf_per_source] with
  ``psfgauss_start[exposure]`` (patch.py:357-396).  The kernels read the
  block at ``psfgauss_start[e]`` for every source, as the reference kernels
  do (compute_gaussians_kernel.cu:77).
"""

import numpy as np

from ..limits import PSF_COLS, psf_struct_dtype
from ..sources import Galaxy, Scene
from .wcs import AffineWCS, scale_at_sky

__all__ = ["PatchBase", "Patch", "AffineWCS", "scale_at_sky",
           "superpixel_order", "pack_exposures"]


def superpixel_order(nx, ny, sx, sy):
    """Flat indices (into an [nx, ny] image) in superpixel order: sx-by-sy
    blocks taken row-major over the block grid, each block flattened
    row-major (forcepho/patches/storage.py:238-270).  Edge blocks are
    truncated when sx or sy does not divide the image."""
    NX = -(-nx // sx) * sx
    NY = -(-ny // sy) * sy
    ii, jj = np.meshgrid(np.arange(NX), np.arange(NY), indexing="ij")
    ii = ii.reshape(NX // sx, sx, NY // sy, sy).transpose(0, 2, 1, 3).reshape(-1)
    jj = jj.reshape(NX // sx, sx, NY // sy, sy).transpose(0, 2, 1, 3).reshape(-1)
    keep = (ii < nx) & (jj < ny)
    return ii[keep] * ny + jj[keep]


def pack_exposures(images, ierrs, bands, superpixel=(8, 8), warp_size=64,
                   dtype=np.float32):
    """Concatenate exposures (each [nx, ny], pixel (i, j) at x=i, y=j) in
    superpixel order with padding.  ``bands`` gives each exposure's band
    index (non-decreasing).  Returns a dict of the PatchBase pixel arrays."""
    bands = np.asarray(bands)
    assert np.all(np.diff(bands) >= 0), "Exposures must be sorted by band"
    n_exp = len(images)
    n_bands = int(bands.max()) + 1
    sizes = [im.size for im in images]
    padded = [(n + warp_size - 1) // warp_size * warp_size for n in sizes]
    total = int(np.sum(padded))
    out = {k: np.zeros(total, dtype=dtype) for k in ("data", "ierr", "xpix", "ypix")}
    out["exposure_start"] = np.zeros(n_exp, dtype=np.int32)
    out["exposure_N"] = np.zeros(n_exp, dtype=np.int32)
    out["band_N"] = np.zeros(n_bands, dtype=np.int16)
    i = 0
    for e, (im, ie) in enumerate(zip(images, ierrs)):
        nx, ny = im.shape
        order = superpixel_order(nx, ny, superpixel[0], superpixel[1])
        xp, yp = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
        n = im.size
        out["exposure_start"][e] = i
        out["exposure_N"][e] = padded[e]
        out["data"][i:i + n] = im.reshape(-1)[order]
        out["ierr"][i:i + n] = ie.reshape(-1)[order]
        out["xpix"][i:i + n] = xp.reshape(-1)[order]
        out["ypix"][i:i + n] = yp.reshape(-1)[order]
        out["xpix"][i + n:i + padded[e]] = -1
        out["ypix"][i + n:i + padded[e]] = -1
        out["band_N"][bands[e]] += 1
        i += padded[e]
    out["band_start"] = np.zeros(n_bands, dtype=np.int16)
    out["band_start"][1:] = np.cumsum(out["band_N"])[:-1]
    out["band_N_pix"] = np.array([np.sum(np.array(padded)[bands == b])
                                  for b in range(n_bands)], dtype=np.int32)
    out["original_shapes"] = np.array([im.shape for im in images])
    out["original_sizes"] = np.array(sizes)
    return out


class PatchBase:
    """Pixel and metadata arrays of one patch in the layout the device
    expects (forcepho/patches/patch.py:20-197)."""

    ptr_dtype = np.uintp
    # field list of the reference struct (forcepho/src/patch.cc:23-100)
    pdt = np.dtype([('data', ptr_dtype), ('ierr', ptr_dtype), ('xpix', ptr_dtype),
                    ('ypix', ptr_dtype), ('residual', ptr_dtype),
                    ('exposure_start', ptr_dtype), ('exposure_N', ptr_dtype),
                    ('band_start', ptr_dtype), ('band_N', ptr_dtype),
                    ('n_sources', np.int32), ('n_radii', np.int32),
                    ('rad2', ptr_dtype), ('D', ptr_dtype), ('crpix', ptr_dtype),
                    ('crval', ptr_dtype), ('CW', ptr_dtype), ('G', ptr_dtype),
                    ('n_psf_per_source', ptr_dtype), ('psfgauss', ptr_dtype),
                    ('psfgauss_start', ptr_dtype)], align=True)
    patch_struct_dtype = pdt
    meta_names = ["n_sources", "n_radii", "rad2", "D", "crpix", "crval", "CW", "G",
                  "n_psf_per_source", "psfgauss", "psfgauss_start"]

    def __init__(self, return_residual=False, pix_dtype=np.float32,
                 meta_dtype=np.float32):
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
            arr = getattr(self, name, None)
            if isinstance(arr, np.ndarray):
                s += arr.nbytes
        return s

    def clear(self):
        for name in self.patch_struct_dtype.names:
            if hasattr(self, name):
                delattr(self, name)

    def free(self):
        pass

    def build_patch(self, npix=1, n_sources=1, n_radii=1, n_bands=1, n_exp=1):
        """Allocate empty arrays with the shapes and dtypes of the device
        layout (forcepho/patches/patch.py:121-197)."""
        self.n_exp = n_exp
        self.n_bands = n_bands
        self.xpix = np.zeros(npix, dtype=self.pix_dtype)
        self.ypix = np.zeros(npix, dtype=self.pix_dtype)
        self.data = np.zeros(npix, dtype=self.pix_dtype)
        self.ierr = np.zeros(npix, dtype=self.pix_dtype)
        self.band_start = np.zeros(n_bands, dtype=np.int16)
        self.band_N = np.zeros(n_bands, dtype=np.int16)
        self.exposure_start = np.zeros(n_exp, dtype=np.int32)
        self.exposure_N = np.zeros(n_exp, dtype=np.int32)
        self.n_sources = n_sources
        self.n_radii = n_radii
        self.rad2 = np.zeros(n_radii, dtype=self.meta_dtype)
        self.D = np.zeros((n_exp, n_sources, 2, 2), dtype=self.meta_dtype)
        self.CW = np.zeros((n_exp, n_sources, 2, 2), dtype=self.meta_dtype)
        self.crpix = np.zeros((n_exp, n_sources, 2), dtype=self.meta_dtype)
        self.crval = np.zeros((n_exp, n_sources, 2), dtype=self.meta_dtype)
        self.G = np.ones(n_exp, dtype=self.meta_dtype)
        self.n_psf_per_source = np.zeros(n_bands, dtype=np.int32)
        self.psfgauss = np.zeros(0, dtype=psf_struct_dtype)
        self.psfgauss_start = np.zeros(n_exp, dtype=np.int32)

    def split_pix(self, attr):
        return np.split(getattr(self, attr), np.cumsum(self.exposure_N[:-1]))

    def split_band(self, attr):
        return np.split(getattr(self, attr), np.cumsum(self.band_N_pix[:-1]))


class Patch(PatchBase):
    """PatchBase plus metadata packing from a Scene, per-exposure WCSs and
    PSF mixtures (forcepho/patches/patch.py:200-499).

    Attributes that must be set before :meth:`pack_meta`: ``bandlist``
    [n_bands], ``bands`` [n_exp] (band name of each exposure), ``wcses``
    [n_exp] (objects with ``pixel_scale_matrix`` and
    ``world_to_pixel_values``), ``psfs`` [n_exp] (structured arrays of
    ``psf_struct_dtype``, n_psf_per_source rows each, constant per band) and
    optionally ``Gs`` [n_exp] flux calibrations (default 1).
    """

    def __init__(self, splinedata=None, return_residual=False,
                 meta_dtype=np.float32, pix_dtype=np.float32):
        super().__init__(return_residual=return_residual, meta_dtype=meta_dtype,
                         pix_dtype=pix_dtype)
        self.splinedata = splinedata
        self.patch_reference_coordinates = np.zeros(2)

    # -- pixels
    def pack_pix(self, images, ierrs, bands, superpixel=(8, 8), warp_size=64):
        """Pack exposure images (each [nx, ny]) in superpixel order."""
        out = pack_exposures(images, ierrs, bands, superpixel=superpixel,
                             warp_size=warp_size, dtype=self.pix_dtype)
        for k, v in out.items():
            setattr(self, k, v)
        self.n_exp = len(images)
        self.n_bands = len(self.band_N)
        self._dirty_data = True

    # -- metadata
    def pack_meta(self, scene):
        self.patch_reference_coordinates = self.zerocoords(scene)
        self.n_sources = len(scene.sources)
        self._pack_source_metadata(scene)
        self._pack_astrometry(self.wcses, scene)
        self._pack_fluxcal()
        self._pack_psf(scene)
        self.scene = scene

    def _pack_source_metadata(self, scene):
        self.n_sources = scene.n_active
        self.n_radii = len(scene.sources[0].radii)
        self.rad2 = np.empty(self.n_radii, dtype=self.meta_dtype)
        self.rad2[:] = scene.sources[0].radii ** 2

    def _pack_astrometry(self, wcses, scene):
        dt = self.meta_dtype
        ns = self.n_sources
        self.D = np.empty((self.n_exp, ns, 2, 2), dtype=dt)
        self.CW = np.empty((self.n_exp, ns, 2, 2), dtype=dt)
        self.crpix = np.empty((self.n_exp, ns, 2), dtype=dt)
        self.crval = np.empty((self.n_exp, ns, 2), dtype=dt)
        ra0, dec0 = self.patch_reference_coordinates
        for j, wcs in enumerate(wcses):
            for i, s in enumerate(scene.sources):
                ssky = np.array([s.ra + ra0, s.dec + dec0])
                CW_mat, D_mat = scale_at_sky(ssky, wcs)
                self.D[j, i] = D_mat
                self.CW[j, i] = CW_mat
                self.crval[j, i] = ssky - self.patch_reference_coordinates
                self.crpix[j, i] = wcs.world_to_pixel_values(ssky[0], ssky[1])

    def _pack_fluxcal(self):
        G = getattr(self, "Gs", None)
        self.G = np.ones(self.n_exp, dtype=self.meta_dtype)
        if G is not None:
            self.G[:] = G

    def _pack_psf(self, scene):
        """[exposure][source][n_psf_per_source] (patch.py:357-396)."""
        band_index = [self.bandlist.index(b) for b in self.bands]
        self.n_psf_per_source = np.zeros(self.n_bands, dtype=np.int32)
        for e, psf in enumerate(self.psfs):
            b = band_index[e]
            if self.n_psf_per_source[b] == 0:
                self.n_psf_per_source[b] = len(psf)
            assert self.n_psf_per_source[b] == len(psf), \
                "n_psf_per_source must be constant within a band"
        self.psfgauss_start = np.zeros(self.n_exp, dtype=np.int32)
        n_psfgauss = int(sum(len(p) for p in self.psfs) * self.n_sources)
        self.psfgauss = np.empty(n_psfgauss, dtype=psf_struct_dtype)
        s = 0
        for e, psf in enumerate(self.psfs):
            self.psfgauss_start[e] = s
            for i, source in enumerate(scene.sources):
                N = len(psf)
                self.psfgauss[s:s + N] = psf
                s += N
        assert s == n_psfgauss

    def set_scene(self, sourcecat, splinedata=None):
        return Scene(catalog=sourcecat, filternames=self.bandlist,
                     splinedata=splinedata or self.splinedata)

    def zerocoords(self, scene, sky_zero=None):
        """Subtract the median source position (or ``sky_zero``) from the
        source coordinates (forcepho/patches/patch.py:425-457)."""
        if sky_zero is None:
            sky_zero = np.array([np.median([s.ra for s in scene.sources]),
                                 np.median([s.dec for s in scene.sources])])
        zero = np.array(sky_zero, dtype=np.float64)
        self.patch_reference_coordinates = zero
        for source in scene.sources:
            source.ra -= zero[0]
            source.dec -= zero[1]
        return zero

    def unzerocoords(self, scene):
        zero = self.patch_reference_coordinates
        for source in scene.sources:
            source.ra += zero[0]
            source.dec += zero[1]
        return scene
