"""The reference's shipped test image (test code only):
``tests/golden/sersic1.0_rhalf0.150_snr050_noise1_data.fits`` is
forcepho's ``tests/data/test_image`` fixture -- a GalSim rendering of one
n = 1, r_h = 0.15 arcsec Sersic source of flux 1 on a 100 x 100 TAN image,
HDU 0 = model + noise, HDU 1 = the uncertainty, HDU 2 = the noise, last HDU
= the one-row catalog (FILTERS in its header).  Its PSF file
(``sersic1.0_rhalf0.150_snr050_noise1_psf.h5``, group CLEAR, dataset
``parameters`` [1, 9]) holds one circular Gaussian (amp 1, Cxx = Cyy = 6.25,
Cxy = 0, no offset) for each of the 9 Sersic radius bins, as h5dump shows;
:func:`psf_store` writes those rows."""
import os

import numpy as np

from forcepho_amd.patches import fits
from forcepho_amd.patches.storage import PSF_DTYPE, PSFStore
from forcepho_amd.region import CircularRegion

PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                    "sersic1.0_rhalf0.150_snr050_noise1_data.fits")


def catalog():
    cat = fits.getdata(PATH, -1)
    bands = fits.getheader(PATH, -1)["FILTERS"].split(",")
    return cat, bands


def psf_store(root):
    rows = np.zeros(9, dtype=PSF_DTYPE)
    rows["amp"] = 1.0
    rows["Cxx"] = rows["Cyy"] = 2.5 ** 2
    rows["sersic_bin"] = np.arange(9)
    PSFStore(root).add_band("CLEAR", rows)
    return root


def region(cat, n_rhalf=5.0):
    """The reference test's region: 5 half-light radii around the source."""
    return CircularRegion(cat["ra"][0], cat["dec"][0], n_rhalf * cat["rhalf"][0] / 3600)


def galsim_model_at(xpix, ypix):
    """The fixture's noise-free model (HDU 0 - HDU 2) at packed pixels."""
    m = (fits.getdata(PATH, 0).astype(np.float64) - fits.getdata(PATH, 2)).T   # [x, y]
    return m[np.asarray(xpix, dtype=int), np.asarray(ypix, dtype=int)]
