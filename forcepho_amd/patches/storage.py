"""Exposure data on disk: pixel, metadata and PSF stores
(forcepho/patches/storage.py).

The reference keeps pixels and PSFs in HDF5 files and headers in a JSON
file.  h5py is not part of this build, so the two array stores here are
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
        """Store one exposure (storage.py:148-232): subtract the background,
        mask non-finite or negative-ierr pixels (and, with ``bitmask``, mask
        pixels with any of those bits set) to ierr = 0 with the subtracted
        background as their flux, optionally flux-calibrate, superpixelize
        and write ``<band>/<expID>/data``."""
        hdr = imset.hdr
        band, expID = imset.band, imset.expID
        im = np.array(imset.im, dtype=np.float64)
        ierr = np.array(imset.ierr, dtype=np.float64)
        if imset.bkg is not None:
            bkg = np.asarray(imset.bkg, dtype=np.float64)
            im = im - bkg
        else:
            bkg = np.zeros_like(im)
        mask = ~(np.isfinite(ierr) & np.isfinite(im) & (ierr >= 0))
        if imset.mask is not None:
            pmask = np.asarray(imset.mask)
            if bitmask:
                pmask = np.bitwise_and(pmask, bitmask) != 0
            mask = mask | pmask.astype(bool)
        ierr[mask] = 0
        im[mask] = 0
        gb = np.isfinite(bkg)
        im[mask & gb] = bkg[mask & gb]
        if do_fluxcal:
            fluxconv, unitname = self.flux_calibration(hdr)
            im *= fluxconv
            ierr *= 1.0 / fluxconv
        else:
            fluxconv, unitname = 1.0, "image"
        imsize = np.array(im.shape)
        if np.any(imsize % self.super_pixel_size):
            raise ValueError("image size must be a multiple of the superpixel size")
        if np.any(imsize != self._imsize()):
            raise ValueError("Image is not the expected size")
        sp = self.superpixelize(im, ierr)
        if not np.all(np.isfinite(sp)):
            raise ValueError(f"There were non-finite pixels in exposure {expID}")
        path = EXP_FMT.format(band, expID)
        self._tree.delete(path)
        dset = self._tree.create_dataset(path + "/data", sp)
        a = dset.attrs
        a["counts_to_flux"] = fluxconv
        a["flux_units"] = unitname
        if bitmask:
            a["bitmask_applied"] = int(bitmask)
        names = imset.names
        if isinstance(names, str):
            a["image_name"] = names
        elif names is not None and hasattr(names, "_fields"):
            for f, v in zip(names._fields, names):
                if isinstance(v, str):
                    a[f] = v

    @property
    def data(self):
        """The store tree: ``data["<band>/<expID>/data"][...]``."""
        return self._tree

    def close(self):
        pass


# --------------------------------------------------------------- metadata
class MetaStore:
    """FITS headers of the exposures, by band and expID, and their WCSs
    (storage.py:302-522)."""

    def __init__(self, metastorefile=None, load_all=False):
        self.headers = {}
        self.wcs = {}
        if metastorefile:
            self.headers = self.read_from_file(metastorefile)
            if load_all:
                self.populate_wcs()

    def populate_wcs(self):
        self.wcs = {}
        for band, exps in self.headers.items():
            self.wcs[band] = {e: WCS(h) for e, h in exps.items()}

    def get_wcs(self, band, expID):
        """The WCS of one exposure, built on first use and cached."""
        w = self.wcs.setdefault(band, {})
        if expID not in w:
            w[expID] = WCS(self.headers[band][expID])
        return w[expID]

    def add_exposure(self, imset):
        band, expID = imset.band, imset.expID
        hdr = imset.hdr if isinstance(imset.hdr, Header) else Header(imset.hdr)
        self.headers.setdefault(band, {})[expID] = hdr

    def write_to_file(self, filename):
        """{band: {expID: header string}} as JSON (the reference format)."""
        assert ".json" in filename
        out = {b: {e: h.tostring() for e, h in exps.items()} for b, exps in self.headers.items()}
        with open(filename, "w") as f:
            json.dump(out, f)

    def read_from_file(self, filename):
        with open(filename) as f:
            s = json.load(f)
        return {b: {e: Header.fromstring(h) for e, h in exps.items()} for b, exps in s.items()}

    def find_exposures(self, sky, bandlist):
        """(expIDs, bands) of the exposures in ``bandlist`` whose image
        footprint contains ``sky`` = (ra, dec) (storage.py:415-476): a rough
        S_REGION polygon test when the header has one, then the polygon of
        the image corners through the WCS."""
        epaths, bands = [], []
        for band in bandlist:
            for expID, hdr in self.headers.get(band, {}).items():
                if "S_REGION" in hdr:
                    poly = np.array([float(p) for p in str(hdr["S_REGION"]).split()[2:]])
                    if not polygon_contains(poly.reshape(4, 2), np.atleast_2d(sky))[0]:
                        continue
                wcs = self.get_wcs(band, expID)
                nx, ny = hdr["NAXIS1"], hdr["NAXIS2"]
                px = np.array([0, 0, nx, nx], dtype=np.float64)
                py = np.array([0, ny, ny, 0], dtype=np.float64)
                bra, bdec = wcs.pixel_to_world_values(px, py)
                if polygon_contains(np.stack([bra, bdec], axis=-1), np.atleast_2d(sky))[0]:
                    epaths.append(expID)
                    bands.append(band)
        return epaths, bands

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
