"""forcepho's verification image (test code only).

``tests/golden/reference_2021Nov30_f200w_sersic2.2_rhalf0.10.fits`` is
forcepho's ``tests/verification/data/reference-2021Nov30_f200w_sersic=2.2_
rhalf=0.10.fits``, written by its make_reference.py with the pure-python
renderer (``forcepho.slow``: PostageStamp.render of one Galaxy through
gaussmodel._compute_gaussian with the second-order pixel term and no exp
cutoff).  HDU 0 is the model of one n = 2.2, r_h = 0.10 arcsec Sersic source
of flux 1 (q = 1, pa = 0) centred on pixel (31, 15.5) of a 63 x 32, 0.03
arcsec TAN image; HDU 1-7 are dFLUX/d(F200W, ra, dec, q, pa, sersic, rhalf);
the last HDU is the one-row catalog.  ``verification_f200w_inputs.npz`` holds
the 4-Gaussian F200W PSF table the image header names and the Sersic-mixture
table make_reference.py interpolates (both written by
``tests/golden/make_verification_golden.py``).

The renderer's astrometry is its PostageStamp's, not the image's TAN
projection: a source lands at dpix_dsky (sky - crval) + crpix with crpix = 0,
crval = (53, -27) and dpix_dsky = the local linear map of the TAN WCS at
1-based pixel (0, 0) (slow/stamp.py:137-161, 200-256, gaussmodel.py:95-110),
and the scale matrix is exactly I / 0.03 pixels per arcsec
(make_reference.py:91-98).  :func:`renderer_geometry` gives those, so that a
patch can be put in the renderer's own coordinates.
"""
import os

import numpy as np

from forcepho_amd.patches import FITSPatch, fits
from forcepho_amd.sources import Scene
from forcepho_amd.patches.storage import PSF_DTYPE, PSFStore
from forcepho_amd.patches.wcs import WCS, scale_at_sky
from forcepho_amd.region import CircularRegion

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
PATH = os.path.join(GOLDEN, "reference_2021Nov30_f200w_sersic2.2_rhalf0.10.fits")
INPUTS = os.path.join(GOLDEN, "verification_f200w_inputs.npz")
PARAMS = ["F200W", "ra", "dec", "q", "pa", "sersic", "rhalf"]


def catalog():
    cat = fits.getdata(PATH, -1)
    bands = [b.strip().upper() for b in fits.getheader(PATH, -1)["FILTERS"].split(",")]
    return cat, bands


def images():
    """The rendered model and its 7 derivative images, indexed [x, y]."""
    model = np.asarray(fits.getdata(PATH, 0), dtype=np.float64).T
    grads = np.stack([np.asarray(fits.getdata(PATH, k), dtype=np.float64).T
                      for k in range(1, 8)])
    units = [fits.getheader(PATH, k)["BUNIT"].strip() for k in range(1, 8)]
    return model, grads, units


def psf_store(root):
    """The F200W table for every one of the 9 Sersic radius bins: the
    renderer convolves each mixture component with the same PSF."""
    f = np.load(INPUTS)
    n_radii = len(f["radii"])
    rows = np.zeros(4 * n_radii, dtype=PSF_DTYPE)
    for j, (src, dst) in enumerate([("amp", "amp"), ("xcr", "xcen"), ("ycr", "ycen"),
                                    ("Cxx", "Cxx"), ("Cyy", "Cyy"), ("Cxy", "Cxy")]):
        rows[dst] = np.tile(f["psf_" + src], n_radii)
    rows["sersic_bin"] = np.repeat(np.arange(n_radii), 4)
    PSFStore(root).add_band("F200W", rows)
    return root


def renderer_geometry():
    """(CW, D, pixel position of the catalog source) of the renderer's
    stamp."""
    cat, _ = catalog()
    wcs = WCS(fits.getheader(PATH, 0))
    sky0 = wcs.all_pix2world([[0.0, 0.0]], 1)[0]
    CW, _ = scale_at_sky(sky0, wcs)
    D = np.eye(2) / 0.03
    crval = np.array([53.0, -27.0])
    src = np.array([cat["ra"][0], cat["dec"][0]], dtype=np.float64)
    return CW, D, CW @ (src - crval)


def scene(spline_type="smooth"):
    """The catalog source.  The image's mixture amplitudes are those of
    SmoothBivariateSpline(nsersic, rh, A_m, s=None) on the table (the
    non-"rect" branch of forcepho/sources.py:634-642): a least-squares fit to
    the image over the 9 component amplitudes gives them to 7e-8, while
    the interpolating "rect" spline is off by up to 9e-3."""
    cat, bands = catalog()
    return Scene(catalog=cat, filternames=bands, splinedata=INPUTS, spline_type=spline_type)


def build(root, patch_class=FITSPatch, geometry="renderer", unc=1.0, spline_type="smooth"):
    """The image packed as a one-exposure patch (ierr = 1/unc) with the
    catalog scene; ``geometry="renderer"`` puts the source where the
    renderer put it (its CW, D and position), ``"tan"`` keeps the packer's
    own TAN astrometry."""
    cat, bands = catalog()
    p = patch_class(fitsfiles=[PATH], psfstore=psf_store(os.path.join(root, "psf")),
                    splinedata=INPUTS, sci_ext=0)
    p.unc = unc
    region = CircularRegion(cat["ra"][0], cat["dec"][0], 3.0 / 3600)
    p.build_patch(region, None, allbands=bands)
    p.pack_meta(scene(spline_type))
    if geometry == "renderer":
        CW, D, pos = renderer_geometry()
        p.CW[:, 0] = CW
        p.D[:, 0] = D
        # the scene's position relative to the patch reference is (0, 0), so
        # the pixel position is crpix itself
        assert np.all(np.asarray(p.crval[:, 0]) == 0.0)
        p.crpix[:, 0] = pos
    return p
