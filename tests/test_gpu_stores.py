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


def test_prepare_model_and_fit(hip, survey):
    """The user workflow of forcepho: pack a region from FITS files, wrap it
    in a posterior with ``prepare_model`` (no fixed sources, no bounds), and
    its lnp and gradient at the catalog match the FP64 oracle; BFGS from the
    catalog does not lose posterior probability."""
    from forcepho_amd.fitting import run_opt
    p = SimplePatch(fitsfiles=survey["paths"]["fits"], psfstore=survey["paths"]["psf"])
    p.build_patch(CircularRegion(CENTER[0], CENTER[1], 0.8 / 3600), survey["catalog"],
                  allbands=ALLBANDS)
    try:
        model, q = p.prepare_model(active=survey["catalog"])
        lnp, grad = model.lnprob_and_grad(q)
        prop = proposal_for(model.scene, q)
        c_ref, d_ref = oracle.evaluate(p, prop)
        ll_ref = -0.5 * c_ref.sum()
        g_ref = -0.5 * oracle.stack_grad(d_ref)
        print(f"lnp {lnp:.6f} oracle {ll_ref:.6f}", flush=True)
        assert abs(lnp - ll_ref) <= CHI2_RTOL * abs(ll_ref)
        err = np.abs(grad - g_ref)
        assert np.all((err <= 1e-4 * np.abs(g_ref)) | (err <= 1e-6 * np.abs(g_ref).max()))
        res, sci = run_opt(model, q, maxiter=20)
        lnp_opt = model.lnprob(res.chain[0])
        print(f"BFGS: lnp {lnp:.3f} -> {lnp_opt:.3f} in {sci.nit} iterations", flush=True)
        assert lnp_opt >= lnp - 1e-6 * abs(lnp)
    finally:
        p.free()
