"""The store-backed and FITS packers feed the HIP path: lnlike and gradient
of JadesPatch / SimplePatch patches (TAN WCS, superpixels or padded single
pixels) against the FP64 oracle."""
import numpy as np
import pytest

import oracle
from forcepho_amd import synth
from forcepho_amd.patches import JadesPatch, SimplePatch
from forcepho_amd.region import CircularRegion
from helpers import CHI2_RTOL, chi2_relerr, grad_mismatch, proposal_for
from survey import ALLBANDS, CENTER, make_survey

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def survey(tmp_path_factory):
    return make_survey(str(tmp_path_factory.mktemp("survey_gpu")))


def _check(patch, label):
    scene = patch.scene
    theta = scene.get_all_source_params().copy()
    props = [theta] + list(synth.make_proposals(theta, patch.n_bands, 3, config=2, patch_id=9))
    patch.send_to_device()
    try:
        for k, t in enumerate(props):
            prop = proposal_for(scene, t)
            chi2, dchi2 = patch.get_proposer().evaluate_proposal(prop, patch=patch)[:2]
            c_ref, d_ref = oracle.evaluate(patch, prop)
            rel = chi2_relerr(chi2, c_ref)
            bad = grad_mismatch(dchi2, d_ref)
            print(f"{label} proposal {k}: chi2 rel err {rel.max():.2e}, "
                  f"grad mismatches {bad.sum()}/{bad.size}", flush=True)
            assert rel.max() <= CHI2_RTOL, (label, k, rel)
            assert not bad.any(), (label, k, np.argwhere(bad)[:5])
    finally:
        patch.free()


def test_jadespatch_lnlike_and_grad(hip, survey):
    p = JadesPatch(pixelstore=survey["paths"]["pixel"], metastore=survey["paths"]["meta"],
                   psfstore=survey["paths"]["psf"])
    p.build_patch(CircularRegion(CENTER[0], CENTER[1], 1.0 / 3600), survey["catalog"],
                  allbands=ALLBANDS)
    _check(p, "JadesPatch")


def test_simplepatch_lnlike_and_grad(hip, survey):
    p = SimplePatch(fitsfiles=survey["paths"]["fits"], psfstore=survey["paths"]["psf"])
    p.build_patch(CircularRegion(CENTER[0], CENTER[1], 0.8 / 3600), survey["catalog"],
                  allbands=ALLBANDS)
    _check(p, "SimplePatch")
