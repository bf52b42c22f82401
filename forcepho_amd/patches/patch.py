"""Patch packing: the host arrays a device patch is built from.

What the device needs is fixed by the reference's patch struct
(forcepho/src/patch.cc:23-100) and its packers (forcepho/patches/patch.py:
121-396, static_patch.py:238-307, pixel_patch.py:124-202):

* pixels -- all exposures concatenated, grouped by band, each exposure in
  superpixel order and padded to a multiple of the warp size with
  ierr = data = 0 and x = y = -1: ``data, ierr, xpix, ypix`` (float32),
  ``exposure_start/exposure_N`` (int32 [n_exp]), ``band_start/band_N``
  (int16 [n_bands]);
* scene metadata -- per (exposure, source) the local affine map of the
  exposure's WCS at the source: D (pixels per arcsec), CW (pixels per
  degree), crpix, crval; per exposure the flux calibration G; the squared
  Sersic radii; the PSF mixture of every (exposure, source) pair, laid out
  [exposure][source][n_psf_per_source] with ``psfgauss_start[exposure]``
  (the kernels read the block at ``psfgauss_start[e]`` for every source,
  as compute_gaussians_kernel.cu:77 does).

Here the pixel layout of an image shape is computed once and cached
(synthetic and survey patches repeat a few shapes), the astrometry of an
exposure is evaluated for all sources at once, and the array set is
described by one field table that allocation and the device descriptor
share.
"""

import functools

import numpy as np

from ..limits import psf_struct_dtype
from ..sources import Scene
from .wcs import AffineWCS, local_affine, scale_at_sky

__all__ = ["PatchBase", "Patch", "AffineWCS", "scale_at_sky",
           "superpixel_order", "pack_exposures", "layout_exposures"]


# --------------------------------------------------------------- pixels

@functools.lru_cache(maxsize=256)
def _layout(nx, ny, sx, sy):
    """Read-only (order, x, y) of an [nx, ny] image: ``order`` lists flat
    pixel indices (x * ny + y) superpixel by superpixel -- sx-by-sy blocks
    visited row-major over the block grid, each block row-major -- with
    edge blocks cut at the image border; x, y are float32 coordinates of
    those pixels (x the first image index)."""
    bx, by = -(-nx // sx), -(-ny // sy)
    # block (p, q), in-block offset (u, v) -> pixel (p sx + u, q sy + v)
    p = np.arange(bx)[:, None, None, None]
    q = np.arange(by)[None, :, None, None]
    u = np.arange(sx)[None, None, :, None]
    v = np.arange(sy)[None, None, None, :]
    x = np.broadcast_to(p * sx + u, (bx, by, sx, sy)).reshape(-1)
    y = np.broadcast_to(q * sy + v, (bx, by, sx, sy)).reshape(-1)
    inside = (x < nx) & (y < ny)
    x, y = x[inside], y[inside]
    order = x * ny + y
    xf, yf = x.astype(np.float32), y.astype(np.float32)
    for a in (order, xf, yf):
        a.setflags(write=False)
    return order, xf, yf


def superpixel_order(nx, ny, sx, sy):
    """Flat indices into an [nx, ny] image in superpixel order (the order of
    forcepho/patches/storage.py:238-270, edge blocks truncated)."""
    return _layout(int(nx), int(ny), int(sx), int(sy))[0].copy()


def _index_arrays(sizes, bands):
    """exposure_start/exposure_N (int32), band_start/band_N (int16) and
    band_N_pix (int32) for exposures of ``sizes`` pixels in ``bands``."""
    sizes = np.asarray(sizes, dtype=np.int64)
    bands = np.asarray(bands, dtype=np.int64)
    if np.any(np.diff(bands) < 0):
        raise ValueError("Exposures must be sorted by band")
    n_bands = int(bands.max()) + 1 if len(bands) else 0
    band_N = np.bincount(bands, minlength=n_bands)
    return {"exposure_start": np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int32),
            "exposure_N": sizes.astype(np.int32),
            "band_N": band_N.astype(np.int16),
            "band_start": np.concatenate([[0], np.cumsum(band_N)[:-1]]).astype(np.int16),
            "band_N_pix": np.bincount(bands, weights=sizes,
                                      minlength=n_bands).astype(np.int32)}


def layout_exposures(chunks, bands, dtype=np.float32):
    """Concatenate per-exposure pixel lists into the device arrays.

    ``chunks`` holds one (flux, ierr, x, y) tuple of 1-D arrays per
    exposure, already in the order and padding wanted; ``bands`` gives each
    exposure's band index, non-decreasing.  Returns ``data, ierr, xpix,
    ypix``, ``exposure_start/exposure_N`` (int32), ``band_start/band_N``
    (int16, exposures per band) and ``band_N_pix`` (int32, pixels per
    band)."""
    sizes = [np.size(c[0]) for c in chunks]
    out = _index_arrays(sizes, bands)
    total = int(np.sum(sizes))
    for k, name in enumerate(("data", "ierr", "xpix", "ypix")):
        a = np.empty(total, dtype)
        for c, s0, n in zip(chunks, out["exposure_start"], sizes):
            a[s0:s0 + n] = np.asarray(c[k]).reshape(-1)
        out[name] = a
    return out


def pack_exposures(images, ierrs, bands, superpixel=(8, 8), warp_size=64,
                   dtype=np.float32, shapes=None):
    """Pack exposure images (each [nx, ny], pixel (i, j) at x = i, y = j)
    into the device pixel arrays: each exposure in superpixel order, padded
    to a multiple of ``warp_size`` with flux = ierr = 0 and x = y = -1.
    ``bands`` gives each exposure's band index and must be non-decreasing.
    An image or ierr may be a scalar, a uniform image whose shape is then
    ``shapes[e]``.  Returns the arrays of :func:`layout_exposures` plus
    ``original_shapes`` and ``original_sizes``.  Every output element is
    written once: the superpixel gather goes straight into its slice."""
    def shape_of(e):
        for a in (images[e], ierrs[e]):
            if np.ndim(a) == 2:
                return np.shape(a)
        return tuple(shapes[e])
    shapes = np.array([shape_of(e) for e in range(len(images))],
                      dtype=np.int64).reshape(-1, 2)
    sizes = shapes.prod(axis=1)
    padded = -(-sizes // warp_size) * warp_size
    out = _index_arrays(padded, bands)
    total = int(padded.sum())
    data, ierr = np.empty(total, dtype), np.empty(total, dtype)
    xpix, ypix = np.empty(total, dtype), np.empty(total, dtype)
    for e, (s0, n, n_pad) in enumerate(zip(out["exposure_start"], sizes, padded)):
        order, x, y = _layout(int(shapes[e, 0]), int(shapes[e, 1]), *superpixel)
        sl, pad = slice(int(s0), int(s0 + n)), slice(int(s0 + n), int(s0 + n_pad))
        for dst, src in ((data, images[e]), (ierr, ierrs[e])):
            if np.ndim(src) == 0:
                dst[sl] = src
            else:
                src = np.asarray(src).reshape(-1)
                if src.dtype == dtype:
                    np.take(src, order, out=dst[sl])
                else:
                    dst[sl] = src[order]
            dst[pad] = 0
        xpix[sl], ypix[sl] = x, y
        xpix[pad] = ypix[pad] = -1
    out.update(data=data, ierr=ierr, xpix=xpix, ypix=ypix, original_shapes=shapes,
               original_sizes=sizes)
    return out


# ------------------------------------------------------------ field table

# name -> (dtype key, shape as a function of the counts).  "pix" and "meta"
# resolve to the patch's pixel and metadata dtypes.
_FIELDS = {
    "xpix": ("pix", lambda c: (c["npix"],)),
    "ypix": ("pix", lambda c: (c["npix"],)),
    "data": ("pix", lambda c: (c["npix"],)),
    "ierr": ("pix", lambda c: (c["npix"],)),
    "band_start": (np.int16, lambda c: (c["n_bands"],)),
    "band_N": (np.int16, lambda c: (c["n_bands"],)),
    "exposure_start": (np.int32, lambda c: (c["n_exp"],)),
    "exposure_N": (np.int32, lambda c: (c["n_exp"],)),
    "rad2": ("meta", lambda c: (c["n_radii"],)),
    "D": ("meta", lambda c: (c["n_exp"], c["n_sources"], 2, 2)),
    "CW": ("meta", lambda c: (c["n_exp"], c["n_sources"], 2, 2)),
    "crpix": ("meta", lambda c: (c["n_exp"], c["n_sources"], 2)),
    "crval": ("meta", lambda c: (c["n_exp"], c["n_sources"], 2)),
    "G": ("meta", lambda c: (c["n_exp"],)),
    "n_psf_per_source": (np.int32, lambda c: (c["n_bands"],)),
    "psfgauss": (psf_struct_dtype, lambda c: (0,)),
    "psfgauss_start": (np.int32, lambda c: (c["n_exp"],)),
}


def _struct_dtype(ptr):
    """The reference's patch struct of device pointers (patch.cc:23-100)."""
    head = ("data", "ierr", "xpix", "ypix", "residual", "exposure_start", "exposure_N",
            "band_start", "band_N")
    tail = ("rad2", "D", "crpix", "crval", "CW", "G", "n_psf_per_source", "psfgauss",
            "psfgauss_start")
    return np.dtype([(n, ptr) for n in head] + [("n_sources", np.int32),
                                                 ("n_radii", np.int32)] +
                    [(n, ptr) for n in tail], align=True)


class PatchBase:
    """The array set of one patch in the device layout
    (forcepho/patches/patch.py:20-197)."""

    ptr_dtype = np.uintp
    # kept as a description of the field set; the HIP library builds its own
    pdt = _struct_dtype(np.uintp)
    patch_struct_dtype = pdt
    meta_names = ["n_sources", "n_radii", "rad2", "D", "crpix", "crval", "CW", "G",
                  "n_psf_per_source", "psfgauss", "psfgauss_start"]

    def __init__(self, return_residual=False, pix_dtype=np.float32, meta_dtype=np.float32):
        self.pix_dtype = pix_dtype
        self.meta_dtype = meta_dtype
        self.return_residual = return_residual
        self._dirty_data = True

    def _dtype(self, key):
        return {"pix": self.pix_dtype, "meta": self.meta_dtype}.get(key, key) \
            if isinstance(key, str) else key

    @property
    def npix(self):
        """Pixels of the patch (padding included): the length of the host
        pixel arrays, or, once they were dropped after an upload, of the
        device copy."""
        x = getattr(self, "xpix", None)
        if x is not None:
            return int(np.shape(x)[0])
        return int(getattr(self, "_device_npix", 0))

    @property
    def size(self):
        """Bytes of the host arrays of the field set."""
        return sum(getattr(self, n).nbytes for n in _FIELDS
                   if isinstance(getattr(self, n, None), np.ndarray))

    def clear(self):
        for n in list(_FIELDS) + ["n_sources", "n_radii"]:
            self.__dict__.pop(n, None)

    def free(self):
        pass

    def build_patch(self, npix=1, n_sources=1, n_radii=1, n_bands=1, n_exp=1):
        """Allocate zeroed arrays of every field at these counts (G = 1)."""
        counts = dict(npix=npix, n_sources=n_sources, n_radii=n_radii, n_bands=n_bands,
                      n_exp=n_exp)
        for name, (dt, shape) in _FIELDS.items():
            setattr(self, name, np.zeros(shape(counts), dtype=self._dtype(dt)))
        self.G[:] = 1
        self.n_exp, self.n_bands = n_exp, n_bands
        self.n_sources, self.n_radii = n_sources, n_radii

    def split_pix(self, attr):
        """One array per exposure (padding included)."""
        return np.split(getattr(self, attr), np.cumsum(self.exposure_N)[:-1])

    def split_band(self, attr):
        """One array per band."""
        return np.split(getattr(self, attr), np.cumsum(self.band_N_pix)[:-1])


# ------------------------------------------------------------- metadata

def _psf_block(psfs, n_sources, n_bands, band_of_exposure):
    """psfgauss [exposure][source][n_psf] and psfgauss_start from one PSF
    mixture per exposure; n_psf_per_source per band (constant in a band)."""
    n_psf = np.zeros(n_bands, dtype=np.int32)
    for e, psf in enumerate(psfs):
        b = band_of_exposure[e]
        if n_psf[b] and n_psf[b] != len(psf):
            raise ValueError("n_psf_per_source must be constant within a band")
        n_psf[b] = len(psf)
    rows = [np.tile(np.asarray(psf).astype(psf_struct_dtype), n_sources) for psf in psfs]
    lengths = np.array([len(r) for r in rows], dtype=np.int64)
    start = np.concatenate([[0], np.cumsum(lengths)[:-1]]).astype(np.int32)
    return np.concatenate(rows) if rows else np.zeros(0, psf_struct_dtype), start, n_psf


class Patch(PatchBase):
    """PatchBase plus the scene metadata (forcepho/patches/patch.py:
    200-499).

    Set before :meth:`pack_meta`: ``bandlist`` [n_bands], ``bands`` [n_exp]
    (band name of every exposure), ``wcses`` [n_exp] (objects with
    ``world_to_pixel_values``, and ``pixel_scale_matrix`` or the
    ``all_world2pix``/``all_pix2world`` pair), ``psfs`` [n_exp] (arrays of
    ``psf_struct_dtype``, the same length within a band) and optionally
    ``Gs`` [n_exp] flux calibrations (default 1).
    """

    def __init__(self, splinedata=None, return_residual=False,
                 meta_dtype=np.float32, pix_dtype=np.float32, spline_smoothing=None):
        super().__init__(return_residual=return_residual, meta_dtype=meta_dtype,
                         pix_dtype=pix_dtype)
        self.splinedata = splinedata
        self.spline_smoothing = spline_smoothing
        self.patch_reference_coordinates = np.zeros(2)

    # -- pixels
    def pack_pix(self, images, ierrs, bands, superpixel=(8, 8), warp_size=64, shapes=None):
        """Pack exposure images (each [nx, ny]) in superpixel order;
        ``bands`` is each exposure's band index (scalar images and ierrs
        are uniform, of ``shapes[e]``)."""
        self.__dict__.update(pack_exposures(images, ierrs, bands, superpixel=superpixel,
                                            warp_size=warp_size, dtype=self.pix_dtype,
                                            shapes=shapes))
        self.n_exp = len(images)
        self.n_bands = len(self.band_N)
        self._dirty_data = True

    # -- metadata
    def pack_meta(self, scene):
        """Metadata of ``scene`` against the exposures: the scene's source
        positions become offsets from their median (``zerocoords``), then
        every per-source and per-exposure array is rebuilt."""
        zero = self.zerocoords(scene)
        dt = self.meta_dtype
        srcs = scene.sources
        self.n_sources = scene.n_active
        self.n_radii = len(srcs[0].radii)
        self.rad2 = (np.asarray(srcs[0].radii, dtype=np.float64) ** 2).astype(dt)
        ra = np.array([s.ra for s in srcs], dtype=np.float64) + zero[0]
        dec = np.array([s.dec for s in srcs], dtype=np.float64) + zero[1]
        n_exp = len(self.wcses)
        self.D = np.empty((n_exp, len(srcs), 2, 2), dtype=dt)
        self.CW = np.empty((n_exp, len(srcs), 2, 2), dtype=dt)
        self.crpix = np.empty((n_exp, len(srcs), 2), dtype=dt)
        self.crval = np.empty((n_exp, len(srcs), 2), dtype=dt)
        offsets = np.stack([ra - zero[0], dec - zero[1]], axis=-1)
        for e, wcs in enumerate(self.wcses):
            CW, D = local_affine(wcs, ra, dec)
            self.CW[e], self.D[e] = CW, D
            self.crval[e] = offsets
            self.crpix[e] = np.stack(wcs.world_to_pixel_values(ra, dec), axis=-1)
        G = getattr(self, "Gs", None)
        self.G = np.ones(n_exp, dtype=dt) if G is None else np.asarray(G, dtype=dt).copy()
        band_of = [self.bandlist.index(b) for b in self.bands]
        self.psfgauss, self.psfgauss_start, self.n_psf_per_source = _psf_block(
            self.psfs, len(srcs), self.n_bands, band_of)
        self.scene = scene

    def set_scene(self, sourcecat, splinedata=None):
        """A Scene of Galaxy sources from a catalog with this patch's bands
        (amplitude splines of ``splinedata`` with this patch's
        ``spline_smoothing``, forcepho/patches/patch.py:392-422)."""
        return Scene(catalog=sourcecat, filternames=self.bandlist,
                     splinedata=splinedata or self.splinedata,
                     spline_smoothing=self.spline_smoothing)

    def zerocoords(self, scene, sky_zero=None):
        """Make the scene's source coordinates offsets from ``sky_zero``
        (default: the median source position), which becomes the patch
        reference, so that FP32 offsets keep their precision
        (forcepho/patches/patch.py:425-457).  Returns the reference."""
        if sky_zero is None:
            pos = np.array([(s.ra, s.dec) for s in scene.sources], dtype=np.float64)
            sky_zero = np.median(pos, axis=0)
        zero = np.array(sky_zero, dtype=np.float64)
        for s in scene.sources:
            s.ra, s.dec = s.ra - zero[0], s.dec - zero[1]
        self.patch_reference_coordinates = zero
        return zero

    def unzerocoords(self, scene):
        """Undo :meth:`zerocoords` on the scene."""
        ra0, dec0 = self.patch_reference_coordinates
        for s in scene.sources:
            s.ra, s.dec = s.ra + ra0, s.dec + dec0
        return scene
