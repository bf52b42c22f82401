"""forcepho's shipped test image on the HIP path: lnlike and gradient
against the FP64 oracle, the rendered model against the image's own
(GalSim) model, and the data/residual swap of forcepho's tests/cswap_test.py
(lines 52-90): after the swap the device data is the first residual and a
second evaluation leaves data - 2 model."""
import numpy as np
import pytest

import oracle
from fixture_image import PATH, catalog, galsim_model_at, psf_store, region
from forcepho_amd.patches import SimplePatch
from helpers import CHI2_RTOL, chi2_relerr, grad_mismatch

pytestmark = pytest.mark.gpu


def test_fixture_lnlike_grad_and_swap(hip, tmp_path):
    cat, bands = catalog()
    p = SimplePatch(fitsfiles=[PATH], psfstore=psf_store(str(tmp_path / "psf")),
                    sci_ext=0, unc_ext=1)
    p.return_residual = True
    p.build_patch(region(cat), cat, allbands=bands)
    scenes = [p.set_scene(cat), p.set_scene(cat)]
    proposer = p.get_proposer()
    p.pack_meta(scenes[0])
    p.send_to_device()
    try:
        prop = scenes[0].get_proposal()
        chi2, dchi2, res0 = proposer.evaluate_proposal(prop, patch=p, unpack=False)
        c_ref, d_ref = oracle.evaluate(p, prop)
        assert chi2_relerr(chi2, c_ref).max() <= CHI2_RTOL
        assert not grad_mismatch(dchi2, d_ref).any()
        data0 = p.data.astype(np.float64).copy()
        model = data0 - res0
        real = p.xpix >= 0
        gs = galsim_model_at(p.xpix[real], p.ypix[real])
        assert np.abs(model[real] - gs).max() < 0.01 * gs.max()
        # swap: the residual becomes the data, the data the residual
        p.pack_meta(scenes[1])
        p.swap_on_device()
        d_dev = p.retrieve_array("data")
        r_dev = p.retrieve_array("residual")
        assert np.allclose(d_dev, res0) and np.allclose(r_dev, data0)
        prop = scenes[1].get_proposal()
        _, _, res1 = proposer.evaluate_proposal(prop, patch=p, unpack=False)
        assert np.allclose(res1, data0 - 2 * model, atol=1e-6 * np.abs(data0).max())
    finally:
        p.free()
