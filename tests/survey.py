"""A small synthetic survey on disk for the store-backed packers (test code
only): pixel, metadata and PSF stores and FITS files of the same
exposures, rendered from a known scene with the FP64 oracle."""
import os

import numpy as np

from forcepho_amd import synth
from forcepho_amd.patches import Patch, fits
from forcepho_amd.patches.storage import ImageSet, MetaStore, PixelStore, PSFStore
from forcepho_amd.patches.wcs import TanWCS
from forcepho_amd.sources import Scene

BANDS = ["F090W", "F200W"]
ALLBANDS = ["F070W", "F090W", "F150W", "F200W"]
NSIDE = 64
SUPER = 8
CENTER = (53.16, -27.78)
SCALE = 0.06 / 3600          # degrees per pixel
SIGMA = (0.05, 0.08)         # noise per band


def _catalog(rng):
    ns = 3
    cat = np.zeros(ns, dtype=[("ra", "f8"), ("dec", "f8"), ("q", "f8"), ("pa", "f8"),
                              ("sersic", "f8"), ("rhalf", "f8")] + [(b, "f8") for b in BANDS])
    cosd = np.cos(np.deg2rad(CENTER[1]))
    off = rng.uniform(-0.4, 0.4, size=(ns, 2))           # arcsec
    cat["ra"] = CENTER[0] + off[:, 0] / 3600 / cosd
    cat["dec"] = CENTER[1] + off[:, 1] / 3600
    cat["q"] = rng.uniform(0.4, 0.9, ns)
    cat["pa"] = rng.uniform(-1.2, 1.2, ns)
    cat["sersic"] = rng.uniform(1.2, 4.0, ns)
    cat["rhalf"] = rng.uniform(0.08, 0.25, ns)
    for b in BANDS:
        cat[b] = rng.uniform(20.0, 60.0, ns)
    return cat


def make_survey(root, seed=7):
    """Write the stores and FITS files under ``root``; returns a dict with
    the catalog, per-exposure (band, expID, wcs, image, ierr) and paths."""
    from helpers import oracle_render
    rng = np.random.default_rng(seed)
    cat = _catalog(rng)
    exposures = []
    psfs = {}
    for b, band in enumerate(BANDS):
        psfs[band] = synth._psf_rows(rng, 3 + b, (1.2, 2.2), True)
        for e in range(2):
            t = np.deg2rad(rng.uniform(-8, 8))
            R = np.array([[np.cos(t), -np.sin(t)], [np.sin(t), np.cos(t)]])
            cd = R @ np.diag([-SCALE, SCALE])
            crpix = np.array([NSIDE / 2 + 0.5, NSIDE / 2 + 0.5]) + rng.uniform(-3, 3, 2)
            wcs = TanWCS(crpix, CENTER, cd, naxis=(NSIDE, NSIDE))
            exposures.append(dict(band=band, expID=f"exp{b}{e}", wcs=wcs))

    # render the scene into full images through a Patch of whole exposures
    full = Patch()
    full.bandlist = BANDS
    full.bands = [x["band"] for x in exposures]
    full.wcses = [x["wcs"] for x in exposures]
    full.psfs = [psfs[x["band"]] for x in exposures]
    zeros = [np.zeros((NSIDE, NSIDE), np.float32) for _ in exposures]
    ierrs = [np.full((NSIDE, NSIDE), 1.0 / SIGMA[BANDS.index(x["band"])], np.float32)
             for x in exposures]
    full.pack_pix(zeros, ierrs, [BANDS.index(x["band"]) for x in exposures],
                  superpixel=(SUPER, SUPER))
    scene = Scene(catalog=cat, filternames=BANDS)
    full.pack_meta(scene)
    model = oracle_render(full, scene.get_proposal())
    for e, x in enumerate(exposures):
        s0, n = full.exposure_start[e], full.exposure_N[e]
        im = np.zeros((NSIDE, NSIDE))
        xp = full.xpix[s0:s0 + n].astype(int)
        yp = full.ypix[s0:s0 + n].astype(int)
        ok = xp >= 0
        im[xp[ok], yp[ok]] = model[s0:s0 + n][ok]
        sig = SIGMA[BANDS.index(x["band"])]
        x["image"] = (im + rng.normal(0, sig, im.shape)).astype(np.float32)
        x["ierr"] = np.full(im.shape, 1.0 / sig, np.float32)

    paths = dict(pixel=os.path.join(root, "pixels"), meta=os.path.join(root, "meta.json"),
                 psf=os.path.join(root, "psf"), fits=[])
    pstore = PixelStore(paths["pixel"], nside_full=NSIDE, super_pixel_size=SUPER)
    mstore = MetaStore()
    for x in exposures:
        hdr = fits.Header(dict(x["wcs"].to_header(), FILTER=x["band"], ABMAG=27.0))
        iset = ImageSet(im=x["image"], ierr=x["ierr"], mask=None, bkg=None, hdr=hdr,
                        band=x["band"], expID=x["expID"], names=None)
        pstore.add_exposure(iset)
        mstore.add_exposure(iset)
        # FITS: arrays are [NAXIS2, NAXIS1] = [y, x]
        fpath = os.path.join(root, f"{x['expID']}.fits")
        sci_hdr = dict(hdr.items(), EXTNAME="SCI")
        fits.writeto(fpath, [(None, None), (sci_hdr, x["image"].T),
                             ({"EXTNAME": "ERR"}, (1.0 / x["ierr"]).T)])
        paths["fits"].append(fpath)
    mstore.write_to_file(paths["meta"])
    psfstore = PSFStore(paths["psf"])
    for band in BANDS:
        # two detector locations; the packer asks for the first
        rows = psfs[band]
        other = rows.copy()
        other["amp"] *= 0.5
        psfstore.add_band(band, np.stack([rows, other]),
                          detector_locations=[[10.0, 50.0], [10.0, 50.0]])
    return dict(catalog=cat, exposures=exposures, paths=paths, psfs=psfs)
