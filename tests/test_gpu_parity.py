"""Parity of the HIP path against the FP64 oracle (and the compiled
reference CPU kernel where it is built) on configurations 1 and 2."""
import os

import numpy as np
import pytest

import oracle
from forcepho_amd import synth
from forcepho_amd.synth import PatchSpec
from helpers import (CHI2_RTOL, Fixture, chi2_relerr, grad_mismatch, load_fixture,
                     make_config_patch, oracle_render, proposal_for)

pytestmark = pytest.mark.gpu

FIXTURE_C2 = os.path.join(os.path.dirname(__file__), "golden", "config2_proposals.npz")


@pytest.fixture(scope="module")
def c1(hip):
    patch, scene, theta = make_config_patch(1, return_residual=True)
    patch.send_to_device()
    yield patch, scene, theta
    patch.free()


@pytest.fixture(scope="module")
def c2(hip):
    patch, scene, theta = make_config_patch(2)
    patch.send_to_device()
    yield patch, scene, theta
    patch.free()


def _check(patch, proposal, label):
    chi2, dchi2 = patch.get_proposer().evaluate_proposal(proposal, patch=patch)[:2]
    c_ref, d_ref = oracle.evaluate(patch, proposal)
    rel = chi2_relerr(chi2, c_ref)
    bad = grad_mismatch(dchi2, d_ref)
    print(f"{label}: chi2 rel err max {rel.max():.3e}; grad mismatches {bad.sum()}/{bad.size}",
          flush=True)
    assert rel.max() <= CHI2_RTOL, (label, rel)
    assert not bad.any(), (label, np.argwhere(bad)[:5])
    return chi2, dchi2


def test_config1_truth_and_proposals(c1):
    patch, scene, theta = c1
    props = synth.make_proposals(theta, 1, 5, config=1)
    _check(patch, proposal_for(scene, theta), "c1 truth")
    for k, t in enumerate(props):
        _check(patch, proposal_for(scene, t), f"c1 proposal {k}")


def test_config1_residual(c1):
    patch, scene, theta = c1
    prop = proposal_for(scene, theta)
    chi2, dchi2, res = patch.get_proposer().evaluate_proposal(prop, patch=patch, unpack=False)
    _, _, r_ref = oracle.evaluate(patch, prop, residual=True)
    scale = np.abs(patch.data).max()
    assert np.abs(res - r_ref).max() <= 2e-6 * scale


def test_config1_against_reference_kernel(c1):
    if oracle.reference_module() is None:
        pytest.skip("compiled reference kernel not present")
    patch, scene, theta = c1
    prop = proposal_for(scene, theta)
    chi2, dchi2 = patch.get_proposer().evaluate_proposal(prop, patch=patch)[:2]
    c_r, d_r = oracle.ReferenceKernel(patch).evaluate(prop)
    assert chi2_relerr(chi2, c_r).max() <= CHI2_RTOL
    assert not grad_mismatch(dchi2, d_r).any()


def test_config2_truth(c2):
    patch, scene, theta = c2
    _check(patch, proposal_for(scene, theta), "c2 truth")


@pytest.mark.parametrize("mode", ["mixed", "mixed_centre"])
def test_config2_100_proposals_against_oracle_fixtures(c2, mode):
    """The 100 seeded proposals of configuration 2 against FP64 oracle chi2
    and gradients committed in tests/golden/config2_proposals.npz (written
    by tests/golden/make_scale_fixtures.py on the same seeded patch), in the
    default mode and in the exact-centre mode."""
    patch, scene, theta = c2
    fx = load_fixture(FIXTURE_C2)
    assert Fixture.sha256(patch.data) == fx["data_sha256"], \
        "configuration 2's data differ from the fixture's; regenerate the fixtures"
    props = [proposal_for(scene, t) for t in fx["theta"]]
    assert Fixture.sha256(np.concatenate(props)) == fx["proposals_sha256"]
    proposer = patch.get_proposer()
    worst_c, worst_g = 0.0, 0.0
    patch.set_precision(mode)
    try:
        for k, prop in enumerate(props):
            chi2, dchi2 = proposer.evaluate_proposal(prop, patch=patch)[:2]
            c_ref, d_ref = fx["chi2"][k], fx["dchi2"][k]
            rel = chi2_relerr(chi2, c_ref)
            bad = grad_mismatch(dchi2, d_ref)
            scale = np.abs(d_ref).reshape(len(d_ref), -1).max(axis=1)[:, None, None]
            worst_c = max(worst_c, rel.max())
            worst_g = max(worst_g, float(np.max(np.abs(dchi2 - d_ref) / scale)))
            assert rel.max() <= CHI2_RTOL, (k, rel)
            assert not bad.any(), (k, np.argwhere(bad)[:5])
    finally:
        patch.set_precision("mixed")
    print(f"c2 ({mode}), 100 proposals: chi2 rel err <= {worst_c:.2e}; gradient err <= "
          f"{worst_g:.2e} of the band max; all 112000 components within tolerance",
          flush=True)


def test_determinism(c2):
    patch, scene, theta = c2
    prop = proposal_for(scene, theta)
    a = patch.get_proposer().evaluate_proposal(prop, patch=patch)
    b = patch.get_proposer().evaluate_proposal(prop, patch=patch)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


def test_config2_culling_bitwise(c2):
    patch, scene, theta = c2
    prop = proposal_for(scene, synth.make_proposals(theta, 8, 1, config=2, patch_id=5)[0])
    patch.set_launch(cull=1)
    a = patch.get_proposer().evaluate_proposal(prop, patch=patch)
    patch.set_launch(cull=0)
    b = patch.get_proposer().evaluate_proposal(prop, patch=patch)
    patch.set_launch(cull=1)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


def test_gaussians_beyond_lds_mask_capacity(hip):
    """128 sources x 72 image Gaussians = 9216 Gaussians in one exposure, more
    than the 8192 whose per-slot masks fit in LDS: the Gaussians past that
    capacity are evaluated on every slot their box allows in pass 2."""
    spec = PatchSpec(n_sources=128, n_bands=1, n_exp_per_band=1, side=48, n_psf=8)
    patch, scene, theta = synth.make_patch(spec, config=41, patch_id=0, render=oracle_render)
    assert patch.n_sources * int(patch.n_psf_per_source[0]) > 8192
    patch.send_to_device()
    try:
        _check(patch, proposal_for(scene, theta), "9216 Gaussians truth")
        t = synth.make_proposals(theta, 1, 1, config=41)[0]
        _check(patch, proposal_for(scene, t), "9216 Gaussians proposal")
    finally:
        patch.free()


def test_large_pixel_coordinates(hip):
    """Configuration 2's geometry with every pixel coordinate (and the WCS
    reference pixel) moved by +4000, as in a patch cut from a large detector
    frame: the parity must not depend on where the patch sits."""
    base = synth.config_spec(2)
    spec = PatchSpec(n_sources=base.n_sources, n_bands=base.n_bands,
                     n_exp_per_band=base.n_exp_per_band, side=base.side, n_psf=base.n_psf,
                     superpixel=base.superpixel, warp_size=base.warp_size,
                     mask_frac=base.mask_frac,
                     psf_per_exposure=base.psf_per_exposure,
                     psf_elliptical=base.psf_elliptical, psf_sigma=base.psf_sigma,
                     dither=base.dither, pix_offset=4000.0)
    patch, scene, theta = synth.make_patch(spec, config=44, patch_id=0, render=oracle_render)
    assert patch.xpix.max() > 4000
    patch.send_to_device()
    try:
        _check(patch, proposal_for(scene, theta), "offset 4000 truth")
        t = synth.make_proposals(theta, base.n_bands, 1, config=44)[0]
        _check(patch, proposal_for(scene, t), "offset 4000 proposal")
    finally:
        patch.free()


def test_offset_psf_components(hip):
    """PSF mixture components centred off the source (as real PSF fits are),
    per-exposure flux calibrations G != 1, several bands, per-exposure
    elliptical PSFs and dithers: the drop-in and
    the batched theta path against the FP64 oracle."""
    from forcepho_amd.batch import DeviceSplines, PatchBatch
    import torch
    spec = PatchSpec(n_sources=5, n_bands=3, n_exp_per_band=[2, 1, 2], side=64,
                     n_psf=[3, 2, 4], psf_per_exposure=True, psf_elliptical=True,
                     psf_offset=1.5, dither=0.5, mask_frac=0.02)
    patch, scene, theta = synth.make_patch(spec, config=45, patch_id=0, render=oracle_render)
    assert abs(patch.psfgauss["xcen"]).max() > 0.5
    patch.G = np.array([0.9, 1.1, 1.0, 1.2, 0.95], dtype=np.float32)   # calibrations
    patch.send_to_device()
    batch = PatchBatch([patch])
    splines = DeviceSplines()
    try:
        for k, t in enumerate(np.vstack([theta[None],
                                         synth.make_proposals(theta, 3, 3, config=45)])):
            _check(patch, proposal_for(scene, t), f"offset PSF proposal {k}")
            tt = torch.as_tensor(t, dtype=torch.float64, device="cuda")
            ll, g = batch.lnlike_and_grad(tt, splines)
            ll_ref, g_ref = oracle.lnlike_theta(patch, t)
            g = g.cpu().numpy()
            err = np.abs(g - g_ref)
            assert abs(ll.item() - ll_ref) <= CHI2_RTOL * abs(ll_ref)
            assert np.all((err <= 1e-4 * np.abs(g_ref)) | (err <= 1e-6 * np.abs(g_ref).max()))
    finally:
        splines.free()
        batch.free()
        patch.free()
