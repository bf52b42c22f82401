"""Exposure data on disk: pixel, metadata and PSF stores
(forcepho/patches/storage.py).

The reference keeps pixels and PSFs in HDF5 files and headers in a JSON
file.  This package does not depend on h5py, so the two array stores here are
directory trees with the same group layout, one ``.npy`` file per dataset
(read memory-mapped, never unpickled) and a JSON file per attribute set:

* :class:`PixelStore` -- ``<root>/<band>/<expID>/data.npy`` of shape
  (nsuper_x, nsuper_y, 2 s^2): per superpixel the s x s fluxes, then the
  s x s inverse errors, each block flattened row-major (storage.py:60-299);
  store attributes ``nside_full`` and ``super_pixel_size``;
* :class:`MetaStore` -- a JSON file {band: {expID: header string}}, the
  reference's own format (storage.py:302-522);
* :class:`PSFStore` -- ``<root>/<band>/parameters.npy`` [n_locations,
  n_psf_per_source] of :data:`PSF_DTYPE`, ``detector_locations.npy``
  (2, n_locations) and the band attribute ``n_psf_per_source``
  (storage.py:525-578).

Paths and attribute names follow the HDF5 layout, so a store converted
from the reference's files with any HDF5 reader loads unchanged.
"""

import json
import os
from collections import namedtuple

import numpy as np

from ..region import polygon_contains
from .fits import Header
from .wcs import WCS

__all__ = ["PixelStore", "MetaStore", "PSFStore", "ImageNameSet", "ImageSet",
           "header_to_id", "PSF_COLS", "PSF_DTYPE", "DirStore"]

ImageNameSet = namedtuple("ImageNames", ["im", "err", "mask", "bkg"])
ImageSet = namedtuple("Images", ["im", "ierr", "mask", "bkg", "hdr", "band", "expID", "names"])

EXP_FMT = "{}/{}"
PSF_COLS = ["amp", "xcen", "ycen", "Cxx", "Cyy", "Cxy"]
PSF_DTYPE = np.dtype([(c, np.float32) for c in PSF_COLS] + [("sersic_bin", np.int32)])


def header_to_id(hdr, name, framedir=None, split_by_detector=()):
    """(band, expID) of an exposure from its header FILTER (plus the JWST
    module letter for bands in ``split_by_detector``) and file name
    (storage.py:37-57)."""
    filt = hdr["FILTER"]
    if filt in split_by_detector:
        band = f"{filt.upper()}{hdr['DETECTOR'][3].upper()}"
    else:
        band = filt.upper()
    expID = os.path.basename(name).replace(".fits", "")
    if framedir:
        assert framedir in os.path.dirname(name)
        rpath = os.path.dirname(name).replace(framedir, "")
        if rpath.startswith("/"):
            rpath = rpath[1:]
        expID = os.path.join(rpath, expID)
    return band, expID


# ----------------------------------------------------------- directory store
class _Attrs(dict):
    """Attributes of a group or dataset, saved to JSON on every change."""

    def __init__(self, path):
        self._path = path
        data = {}
        if os.path.exists(path):
            with open(path) as f:
                data = json.load(f)
        super().__init__(data)

    def __setitem__(self, key, value):
        if isinstance(value, np.ndarray):
            value = value.tolist()
        elif isinstance(value, np.generic):
            value = value.item()
        super().__setitem__(key, value)
        os.makedirs(os.path.dirname(self._path) or ".", exist_ok=True)
        with open(self._path, "w") as f:
            json.dump(dict(self), f, indent=1, sort_keys=True)


class DirStore:
    """A group of an HDF5-like tree kept in a directory: subgroups are
    subdirectories, a dataset ``name`` is ``name.npy`` (attributes in
    ``name.attrs.json``), group attributes in ``attrs.json``."""

    def __init__(self, root):
        self.root = root

    @property
    def attrs(self):
        return _Attrs(os.path.join(self.root, "attrs.json"))

    def _path(self, key):
        return os.path.join(self.root, *[k for k in key.split("/") if k])

    def __contains__(self, key):
        p = self._path(key)
        return os.path.isdir(p) or os.path.exists(p + ".npy")

    def __getitem__(self, key):
        p = self._path(key)
        if os.path.exists(p + ".npy"):
            return _Dataset(p)
        if os.path.isdir(p):
            return DirStore(p)
        raise KeyError(key)

    def keys(self):
        if not os.path.isdir(self.root):
            return []
        out = []
        for n in sorted(os.listdir(self.root)):
            if n.endswith(".npy"):
                out.append(n[:-4])
            elif os.path.isdir(os.path.join(self.root, n)):
                out.append(n)
        return out

    def create_group(self, key):
        p = self._path(key)
        os.makedirs(p, exist_ok=True)
        return DirStore(p)

    def create_dataset(self, key, data):
        p = self._path(key)
        os.makedirs(os.path.dirname(p), exist_ok=True)
        np.save(p + ".npy", np.asarray(data), allow_pickle=False)
        if os.path.exists(p + ".attrs.json"):
            os.remove(p + ".attrs.json")
        return _Dataset(p)

    def delete(self, key):
        import shutil
        p = self._path(key)
        if os.path.isdir(p):
            shutil.rmtree(p)
        for ext in (".npy", ".attrs.json"):
            if os.path.exists(p + ext):
                os.remove(p + ext)


class _Dataset:
    def __init__(self, path):
        self._path = path
        self._arr = None

    @property
    def array(self):
        if self._arr is None:
            self._arr = np.load(self._path + ".npy", mmap_mode="r", allow_pickle=False)
        return self._arr

    @property
    def attrs(self):
        return _Attrs(self._path + ".attrs.json")

    @property
    def shape(self):
        return self.array.shape

    @property
    def dtype(self):
        return self.array.dtype

    def __getitem__(self, idx):
        return np.array(self.array[idx])

    def __len__(self):
        return len(self.array)


# ----------------------------------------------------------------- pixels
def clean_exposure(im, ierr, bkg=None, mask=None, bitmask=None):
    """(flux, ierr) in FP64 of one exposure ready for the store.

    flux = im - bkg.  A pixel is bad where flux or ierr is not finite, where
    ierr < 0, or where ``mask`` flags it (with ``bitmask``: where any of
    those bits is set in the integer mask, otherwise where the mask is
    true).  Bad pixels get ierr = 0 and, as flux, the background that was
    subtracted there (0 where it is not finite), so that they still sample
    the background (storage.py:170-189)."""
    im = np.asarray(im, dtype=np.float64)
    ierr = np.asarray(ierr, dtype=np.float64)
    sky = np.zeros_like(im) if bkg is None else np.asarray(bkg, dtype=np.float64)
    flux = im - sky
    with np.errstate(invalid="ignore"):
        bad = ~(np.isfinite(flux) & np.isfinite(ierr) & (ierr >= 0))
    if mask is not None:
        flagged = np.asarray(mask)
        flagged = (np.bitwise_and(flagged, bitmask) != 0) if bitmask else flagged.astype(bool)
        bad |= flagged
    fill = np.where(np.isfinite(sky), sky, 0.0)
    return np.where(bad, fill, flux), np.where(bad, 0.0, ierr)


def _exposure_attrs(fluxconv, units, bitmask, names):
    """Attributes of a stored exposure: calibration, the bitmask applied and
    the file names (a string, or the string fields of a named tuple)."""
    attrs = {"counts_to_flux": fluxconv, "flux_units": units}
    if bitmask:
        attrs["bitmask_applied"] = int(bitmask)
    if isinstance(names, str):
        attrs["image_name"] = names
    elif names is not None and hasattr(names, "_asdict"):
        attrs.update({k: v for k, v in names._asdict().items() if isinstance(v, str)})
    return attrs


class PixelStore:
    """Exposure pixels in superpixel order (storage.py:60-299).

    Parameters
    ----------
    root : directory of the store (created if absent)
    nside_full : int or (nx, ny), image size in pixels, a multiple of
        ``super_pixel_size``; read from the store when it exists
    super_pixel_size : pixels along a superpixel side
    """

    def __init__(self, root, nside_full=2048, super_pixel_size=8, pix_dtype=np.float32):
        self.root = root
        self._tree = DirStore(root)
        attrs = self._tree.attrs
        if "nside_full" in attrs and "super_pixel_size" in attrs:
            self.nside_full = np.array(attrs["nside_full"])
            self.super_pixel_size = int(attrs["super_pixel_size"])
        else:
            os.makedirs(root, exist_ok=True)
            self.nside_full = np.array(nside_full)
            self.super_pixel_size = int(super_pixel_size)
            attrs["nside_full"] = self.nside_full
            attrs["super_pixel_size"] = self.super_pixel_size
        self.pix_dtype = pix_dtype
        self.nside_super = self.nside_full / self.super_pixel_size
        self.xpix, self.ypix = self.pixel_coordinates()

    def _imsize(self, imsize=None):
        if imsize is None:
            imsize = np.zeros(2, dtype=int) + self.nside_full
        return np.asarray(imsize, dtype=int).reshape(2)

    def superpixelize(self, im, ierr, pix_dtype=None):
        """(nsuper_x, nsuper_y, 2 s^2): each s x s block of ``im`` then of
        ``ierr``, flattened row-major."""
        s = self.super_pixel_size
        nx, ny = im.shape
        assert nx % s == 0 and ny % s == 0, "image size must be a multiple of the superpixel"
        dt = pix_dtype or self.pix_dtype

        def blocks(a):
            return np.asarray(a).reshape(nx // s, s, ny // s, s).transpose(0, 2, 1, 3) \
                .reshape(nx // s, ny // s, s * s)
        return np.concatenate([blocks(im), blocks(ierr)], axis=-1).astype(dt)

    def pixel_coordinates(self, imsize=None):
        """(xpix, ypix), each (nsuper_x, nsuper_y, s^2) int16: the pixel
        coordinates of every pixel in superpixel order (x = first image
        index)."""
        nx, ny = self._imsize(imsize)
        xx, yy = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
        packed = self.superpixelize(xx, yy, pix_dtype=np.int32)
        s2 = self.super_pixel_size ** 2
        return packed[..., :s2].astype(np.int16), packed[..., s2:].astype(np.int16)

    def superpixel_corners(self, imsize=None):
        """(nsuper_x, nsuper_y, 4, 2): pixel (x, y) of the corners
        (0,0), (1,0), (1,1), (0,1) x (s-1) from each superpixel's first
        pixel."""
        if imsize is None:
            xpix, ypix = self.xpix, self.ypix
        else:
            xpix, ypix = self.pixel_coordinates(imsize=imsize)
        ll = np.stack([xpix[:, :, 0], ypix[:, :, 0]], axis=-1).astype(np.float64)
        off = np.array([(0, 0), (1, 0), (1, 1), (0, 1)], dtype=np.float64) * \
            (self.super_pixel_size - 1)
        return ll[:, :, None, :] + off[None, None, :, :]

    def flux_calibration(self, hdr):
        """(counts-to-nJy factor, units) from the header's ABMAG zero point
        (storage.py:272-282)."""
        if "ABMAG" in hdr:
            zp = hdr["ABMAG"]
            return 1e9 * 10 ** (0.4 * (8.9 - zp)), "nJy"
        return 1.0, "counts"

    def add_exposure(self, imset, bitmask=None, do_fluxcal=False):
        """Store one exposure as ``<band>/<expID>/data`` (storage.py:148-232):
        background-subtracted, bad pixels given ierr = 0 (see
        :func:`clean_exposure`), optionally flux-calibrated, in superpixel
        order, with the calibration and the image names as attributes.  An
        existing entry of the same exposure is replaced."""
        flux, ierr = clean_exposure(imset.im, imset.ierr, bkg=imset.bkg, mask=imset.mask,
                                    bitmask=bitmask)
        fluxconv, units = self.flux_calibration(imset.hdr) if do_fluxcal else (1.0, "image")
        if do_fluxcal:
            flux *= fluxconv
            ierr /= fluxconv
        shape = np.array(flux.shape)
        if np.any(shape % self.super_pixel_size):
            raise ValueError("image size must be a multiple of the superpixel size")
        if np.any(shape != self._imsize()):
            raise ValueError("Image is not the expected size")
        packed = self.superpixelize(flux, ierr)
        if not np.isfinite(packed).all():
            raise ValueError(f"There were non-finite pixels in exposure {imset.expID}")
        group = EXP_FMT.format(imset.band, imset.expID)
        self._tree.delete(group)
        attrs = self._tree.create_dataset(group + "/data", packed).attrs
        for key, value in _exposure_attrs(fluxconv, units, bitmask, imset.names).items():
            attrs[key] = value

    @property
    def data(self):
        """The store tree: ``data["<band>/<expID>/data"][...]``."""
        return self._tree

    def close(self):
        pass


# --------------------------------------------------------------- metadata
def _nested_map(tree, fn):
    """{band: {expID: fn(value)}} of a two-level dict."""
    return {band: {exp: fn(v) for exp, v in exps.items()} for band, exps in tree.items()}


def _image_corners(hdr):
    """Pixel (x, y) of the four corners of an NAXIS1 x NAXIS2 image, in
    order round the image."""
    nx, ny = float(hdr["NAXIS1"]), float(hdr["NAXIS2"])
    return np.array([0.0, 0.0, nx, nx]), np.array([0.0, ny, ny, 0.0])


def _s_region_polygon(hdr):
    """(4, 2) ra, dec of the header's S_REGION POLYGON, or None."""
    if "S_REGION" not in hdr:
        return None
    values = [float(v) for v in str(hdr["S_REGION"]).split()[2:]]
    return np.array(values).reshape(4, 2)


class MetaStore:
    """FITS headers of the exposures by band and expID, with their WCSs
    built on first use (storage.py:302-522).  On disk it is the reference's
    JSON file of header strings, {band: {expID: header}}."""

    def __init__(self, metastorefile=None, load_all=False):
        self.headers = {} if not metastorefile else self.read_from_file(metastorefile)
        self.wcs = {}
        if load_all:
            self.populate_wcs()

    def populate_wcs(self):
        self.wcs = _nested_map(self.headers, WCS)

    def get_wcs(self, band, expID):
        """The WCS of one exposure, built from its header once."""
        cache = self.wcs.setdefault(band, {})
        if expID not in cache:
            cache[expID] = WCS(self.headers[band][expID])
        return cache[expID]

    def add_exposure(self, imset):
        hdr = imset.hdr if isinstance(imset.hdr, Header) else Header(imset.hdr)
        self.headers.setdefault(imset.band, {})[imset.expID] = hdr

    def write_to_file(self, filename):
        assert ".json" in filename
        with open(filename, "w") as f:
            json.dump(_nested_map(self.headers, lambda h: h.tostring()), f)

    def read_from_file(self, filename):
        with open(filename) as f:
            return _nested_map(json.load(f), Header.fromstring)

    def covers(self, band, expID, sky):
        """Does the exposure's footprint contain sky = (ra, dec)?  The
        S_REGION polygon, when the header has one, rules exposures out
        cheaply; then the sky polygon of the image corners decides (it
        avoids WCS errors far from the image)."""
        hdr = self.headers[band][expID]
        point = np.atleast_2d(sky)
        rough = _s_region_polygon(hdr)
        if rough is not None and not polygon_contains(rough, point)[0]:
            return False
        ra, dec = self.get_wcs(band, expID).pixel_to_world_values(*_image_corners(hdr))
        return bool(polygon_contains(np.stack([ra, dec], axis=-1), point)[0])

    def find_exposures(self, sky, bandlist):
        """(expIDs, bands) of the exposures in ``bandlist`` that cover sky
        (storage.py:415-476)."""
        hits = [(expID, band) for band in bandlist for expID in self.headers.get(band, {})
                if self.covers(band, expID, sky)]
        return [h[0] for h in hits], [h[1] for h in hits]

    @property
    def nexp(self):
        return sum(len(v) for v in self.headers.values())


# -------------------------------------------------------------------- PSFs
class PSFStore:
    """PSF Gaussian mixtures per band and detector location
    (storage.py:525-578)."""

    def __init__(self, root):
        self.root = root
        self._tree = DirStore(root)

    @property
    def data(self):
        return self._tree

    def add_band(self, band, parameters, detector_locations=None, n_psf_per_source=None):
        """Write a band: ``parameters`` [n_locations, n_psf_per_source] of
        PSF_DTYPE (a 1-D array is one location), ``detector_locations``
        (2, n_locations) pixel (x, y)."""
        p = np.asarray(parameters)
        if p.ndim == 1:
            p = p[None, :]
        p = p.astype(PSF_DTYPE)
        self._tree.create_dataset(f"{band}/parameters", p)
        if detector_locations is None:
            detector_locations = np.zeros((2, p.shape[0]))
        self._tree.create_dataset(f"{band}/detector_locations",
                                  np.asarray(detector_locations, dtype=np.float64))
        self._tree[band].attrs["n_psf_per_source"] = int(n_psf_per_source or p.shape[1])

    def n_psf_per_source(self, band):
        return int(self._tree[band].attrs["n_psf_per_source"])

    def lookup(self, band, xy=None):
        """Parameters (n_psf_per_source rows) at the detector location
        nearest to pixel ``xy``; the first location without one."""
        choose = 0
        if xy is not None:
            xp, yp = self._tree[band]["detector_locations"][:]
            choose = int(np.argmin(np.hypot(xy[0] - xp, xy[1] - yp)))
        return self._tree[band]["parameters"][choose]

    def get_local_psf(self, band="F090W", source=None, wcs=None):
        xy = None
        if wcs is not None and source is not None:
            xy = wcs.world_to_pixel_values(source.ra, source.dec)
        return self.lookup(band, xy=xy)
