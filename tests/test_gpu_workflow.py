"""Device-side residual workflow: render, fixed-source subtraction and the
flux design matrix (forcepho/patches/device_patch.py:113-254) against the
FP64 oracle."""
import numpy as np
import pytest

import oracle
from forcepho_amd import synth
from forcepho_amd.synth import PatchSpec
from helpers import oracle_render

pytestmark = pytest.mark.gpu

SPEC = PatchSpec(n_sources=6, n_bands=2, n_exp_per_band=[2, 1], side=48, n_psf=[3, 4])


def _model(patch, catalog):
    """FP64 oracle model image of ``catalog`` on ``patch`` (flat)."""
    scene = patch.set_scene(catalog)
    patch.pack_meta(scene)
    return oracle.evaluate(patch, scene.get_proposal(), model=True)[2]


def _close(a, b, scale):
    return np.max(np.abs(np.asarray(a, dtype=np.float64) - b)) <= 2e-6 * scale


def test_render_matches_oracle_model(hip):
    patch, scene, theta = synth.make_patch(SPEC, config=43, patch_id=0)   # data = 0
    cat = patch.truth_catalog
    patch.return_residual = True
    patch.send_to_device()
    images = patch.render(cat)
    ref = _model(patch, cat)
    ref_split = np.split(ref, np.cumsum(patch.exposure_N)[:-1])
    scale = np.abs(ref).max()
    assert len(images) == len(patch.exposure_N)
    for im, r in zip(images, ref_split):
        assert _close(im, r, scale)
    patch.free()


def test_subtract_fixed_leaves_residual_of_fixed_sources(hip):
    patch, scene, theta = synth.make_patch(SPEC, config=43, patch_id=1, render=oracle_render)
    cat = patch.truth_catalog
    data = patch.data.astype(np.float64)
    fixed, active = cat[:4], cat[4:]
    proposer, ascene = patch.subtract_fixed(fixed, active=active, maxactive=3)
    # device data = data - model(fixed), two blocks of fixed sources
    d_dev = patch.retrieve_array("data")
    m_fixed = _model(patch, fixed)
    scale = np.abs(data).max()
    assert _close(d_dev, data - m_fixed, scale)
    # the device metadata now describe the active sources
    patch.pack_meta(ascene)
    chi2, dchi2, res = proposer.evaluate_proposal(ascene.get_proposal(), patch=patch,
                                                  unpack=False)
    m_all = _model(patch, cat)
    assert _close(res, data - m_all, scale)
    patch.free()


def test_design_matrix_reproduces_active_models(hip):
    patch, scene, theta = synth.make_patch(SPEC, config=43, patch_id=2, render=oracle_render)
    cat = patch.truth_catalog
    data = patch.data.astype(np.float64)
    fixed, active = cat[:2], cat[2:]
    patch.subtract_fixed(fixed, active=active)
    Xes, ys = patch.design_matrix(active=active)
    m_fixed = _model(patch, fixed)
    ys_ref = np.split(data - m_fixed, np.cumsum(patch.band_N_pix[:-1]))
    scale = np.abs(data).max()
    for j, b in enumerate(patch.bandlist):
        assert Xes[j].shape == (len(active), patch.band_N_pix[j])
        assert _close(ys[j], ys_ref[j], scale)
        # fluxes . X = the active sources' model in that band
        m_act = np.split(_model(patch, active), np.cumsum(patch.band_N_pix[:-1]))[j]
        fl = np.array([a[b] for a in active])
        assert _close(fl @ Xes[j], m_act, scale)
        # each row is one source's unit-flux model
        for i in range(len(active)):
            mi = np.split(_model(patch, active[i:i + 1]), np.cumsum(patch.band_N_pix[:-1]))[j]
            assert _close(Xes[j][i] * fl[i], mi, scale)
    patch.free()


def test_optimize_fluxes_recovers_true_fluxes(hip):
    """Noiseless data: the conditional ML fluxes are the true fluxes and the
    precision matrix is the weighted Gram matrix of the design matrix."""
    from forcepho_amd.fitting import optimize_fluxes

    def noiseless(patch, proposal):
        return oracle.evaluate(patch, proposal, model=True)[2]

    patch, scene, theta = synth.make_patch(SPEC, config=43, patch_id=3)
    cat = patch.truth_catalog
    patch.data = noiseless(patch, scene.get_proposal()).astype(np.float32)
    patch.subtract_fixed(None, active=cat)
    fluxes, precisions, Xes, ys, ws = optimize_fluxes(patch, cat, return_all=True)
    assert len(fluxes) == len(patch.bandlist)
    for j, b in enumerate(patch.bandlist):
        true = np.array([c[b] for c in cat])
        assert np.allclose(fluxes[j], true, rtol=2e-4, atol=0), (b, fluxes[j], true)
        Xp = Xes[j] * ws[j]
        assert np.allclose(precisions[j], Xp @ Xp.T, rtol=1e-10, atol=0)
    patch.free()
