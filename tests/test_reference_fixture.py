"""forcepho's shipped test image read through the FITS packer: the oracle's
model of its catalog source matches the image's own (GalSim) model."""
import numpy as np

import oracle
from fixture_image import PATH, catalog, galsim_model_at, psf_store, region
from forcepho_amd.patches import FITSPatch, fits


def test_catalog_and_wcs_of_the_fixture():
    cat, bands = catalog()
    assert bands == ["CLEAR"] and len(cat) == 1
    assert cat["ra"][0] == 53.0 and cat["dec"][0] == -27.0
    assert cat["sersic"][0] == 1.0 and cat["rhalf"][0] == 0.15 and cat["CLEAR"][0] == 1.0
    from forcepho_amd.patches.wcs import WCS
    w = WCS(fits.getheader(PATH, 0))
    x, y = w.world_to_pixel_values(53.0, -27.0)
    assert abs(x - 49.5) < 1e-9 and abs(y - 49.5) < 1e-9


def test_oracle_model_matches_the_fixture_model(tmp_path):
    cat, bands = catalog()
    p = FITSPatch(fitsfiles=[PATH], psfstore=psf_store(str(tmp_path / "psf")),
                  sci_ext=0, unc_ext=1)
    p.build_patch(region(cat), cat, allbands=bands)
    chi2, _, model = oracle.evaluate(p, p.scene.get_proposal(), model=True)
    real = p.xpix >= 0
    gs = galsim_model_at(p.xpix[real], p.ypix[real])
    peak = gs.max()
    # the mixture-of-Gaussians Sersic, PSF convolution and TAN astrometry
    # reproduce an independent renderer's image to 1% of its peak
    assert np.abs(model[real] - gs).max() < 0.01 * peak
    assert abs(model[real].sum() / gs.sum() - 1.0) < 0.01
    # chi2 at the truth is the chi2 of the noise
    n = int(real.sum())
    full = chi2[0] + np.sum((p.data.astype(np.float64) * p.ierr) ** 2)
    noise = fits.getdata(PATH, 2).T[p.xpix[real].astype(int), p.ypix[real].astype(int)]
    expect = np.sum((noise * p.ierr[real]) ** 2)
    assert abs(full - expect) < 0.02 * expect
