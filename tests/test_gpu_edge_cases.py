"""Edge cases on the HIP path against the FP64 oracle: a source far outside
the patch (every Gaussian culled: exact zeros), a zero-flux source (only its
flux derivatives survive), a fully masked exposure (contributes nothing),
and the amplitude branches for nsersic below the table and rh < 0
(forcepho/sources.py:656-694: one-hot amplitudes, zero derivatives) on the
device parameter path."""
import numpy as np
import pytest
import torch

import oracle
from forcepho_amd import synth
from forcepho_amd.batch import DeviceSplines, PatchBatch
from forcepho_amd.limits import source_struct_dtype
from forcepho_amd.synth import PatchSpec
from helpers import CHI2_RTOL, chi2_relerr, grad_mismatch, oracle_render, proposal_for

pytestmark = pytest.mark.gpu


def _patch(seed, **kw):
    spec = PatchSpec(n_sources=4, n_bands=2, n_exp_per_band=[2, 1], side=64, n_psf=3,
                     mask_frac=0.02, **kw)
    return synth.make_patch(spec, config=seed, patch_id=0, render=oracle_render)


def _compare(patch, prop, label):
    chi2, dchi2 = patch.get_proposer().evaluate_proposal(prop, patch=patch)[:2]
    c_ref, d_ref = oracle.evaluate(patch, prop)
    print(f"{label}: chi2 rel err {chi2_relerr(chi2, c_ref).max():.2e}", flush=True)
    assert chi2_relerr(chi2, c_ref).max() <= CHI2_RTOL, label
    assert not grad_mismatch(dchi2, d_ref).any(), label
    return np.asarray(dchi2), d_ref, np.asarray(chi2)


def test_outside_and_zero_flux_sources(hip):
    patch, scene, theta = _patch(61)
    nb = len(patch.band_N)
    t = theta.copy().reshape(4, nb + 6)
    t[1, nb] += 0.02            # ra: ~2400 pixels away, beyond every cutoff
    t[2, :nb] = 0.0             # zero flux in every band
    patch.send_to_device()
    try:
        dchi2, d_ref, _ = _compare(patch, proposal_for(scene, t.reshape(-1)), "outside/zero")
        # the far source: every Gaussian culled on the device, every pair
        # beyond the cutoff in the oracle -- both exactly zero
        assert np.all(dchi2[:, 1, :] == 0.0) and np.all(d_ref[:, 1, :] == 0.0)
        # the zero-flux source: shape derivatives vanish, flux ones do not
        assert np.all(dchi2[:, 2, 1:] == 0.0)
        assert np.all(np.abs(dchi2[:, 2, 0]) > 0.0)
    finally:
        patch.free()


def test_fully_masked_exposure(hip):
    """An exposure with ierr = 0 everywhere: the evaluation matches the
    oracle, and changing that exposure's data changes no output bit."""
    patch, scene, theta = _patch(62)
    s0, n0 = int(patch.exposure_start[1]), int(patch.exposure_N[1])
    patch.ierr = patch.ierr.copy()
    patch.ierr[s0:s0 + n0] = 0.0          # exposure 1 (band 0) carries no weight
    prop = proposal_for(scene, theta)
    out = []
    for shift in (0.0, 1.0e3):
        patch.data = patch.data.copy()
        patch.data[s0:s0 + n0] += np.float32(shift)
        patch.send_to_device()
        try:
            out.append(_compare(patch, prop, f"masked exposure, data + {shift}"))
        finally:
            patch.free()
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][2], out[1][2])


def test_amplitude_edge_branches_on_device(hip):
    """The data are re-rendered from the edge-branch parameters (the truth's
    noise kept), so the gradient is that of a model fitting its data.  Far
    from its data, a point-like model's position derivative is a sum that
    nearly cancels over the pixels, and the bright-term FP64 centre against
    the gradient pass's FP32 centre (half an ulp apart) moves it by up to
    3e-4 relative -- a limit of the mixed mode on that ill-conditioned
    component, measured separately."""
    patch, scene, theta = _patch(63)
    nb = len(patch.band_N)
    from forcepho_amd.spline import get_splines
    nmin = get_splines().sersic_range[0]
    t = theta.copy().reshape(4, nb + 6)
    t[0, nb + 4] = nmin - 0.1              # below the table: one-hot at the nearest radius
    t[1, nb + 5] = -0.01                   # rh < 0: one-hot at radius 0
    t = t.reshape(-1)
    real = patch.xpix >= 0
    shift = (oracle.render(patch, proposal_for(scene, t)) -
             oracle.render(patch, proposal_for(scene, theta)))
    patch.data = np.where(real, patch.data + shift, 0.0).astype(np.float32)
    patch.send_to_device()
    batch = PatchBatch([patch])
    splines = DeviceSplines()
    try:
        tt = torch.as_tensor(t, dtype=torch.float64, device="cuda")
        batch.theta_to_proposals(tt, splines)
        torch.cuda.synchronize()
        dev = batch.proposals.cpu().numpy().view(source_struct_dtype)
        host = proposal_for(scene, t)
        for k in (0, 1):
            a = dev["mixture_amplitudes"][k]
            assert np.array_equal(a, host["mixture_amplitudes"][k])
            assert np.count_nonzero(a) == 1 and a.max() == 1.0
            assert not dev["damplitude_drh"][k].any() and not dev["damplitude_dnsersic"][k].any()
        ll, g = batch.lnlike_and_grad(tt, splines)
        ll_ref, g_ref = oracle.lnlike_theta(patch, t)
        g = g.cpu().numpy()
        err = np.abs(g - g_ref)
        ok = (err <= 1e-4 * np.abs(g_ref)) | (err <= 1e-6 * np.abs(g_ref).max())
        for i in np.nonzero(~ok)[0]:
            print(f"component {i}: device {g[i]:.9e} oracle {g_ref[i]:.9e}", flush=True)
        assert abs(ll.item() - ll_ref) <= CHI2_RTOL * abs(ll_ref)
        assert ok.all()
        # nsersic and rh derivatives of the two edge sources are zero
        g = g.reshape(4, nb + 6)
        assert g[0, nb + 4] == 0.0 and g[1, nb + 4] == 0.0 and g[1, nb + 5] == 0.0
    finally:
        splines.free()
        batch.free()
        patch.free()


def _edge_theta(theta, nb):
    from forcepho_amd.spline import get_splines
    t = theta.copy().reshape(4, nb + 6)
    t[0, nb + 4] = get_splines().sersic_range[0] - 0.1   # below the table
    t[1, nb + 5] = -0.01                                   # rh < 0
    return t.reshape(-1)


def test_exact_centre_mode_on_a_near_cancelling_position_derivative(hip):
    """The amplitude edge branches against data rendered from the truth: the
    point-like source 1 (rh < 0) sits away from its data, and its position
    derivative is a sum that nearly cancels over the pixels.  In the mixed
    mode the FP64 bright terms and the FP32 terms see centres up to half an
    FP32 ulp apart, which moves such a component by up to ~3e-4 relative;
    the mixed_centre mode takes every pixel offset from the FP64 centre, and
    every component meets the stated tolerance."""
    patch, scene, theta = _patch(63)
    nb = len(patch.band_N)
    t = _edge_theta(theta, nb)
    patch.send_to_device()
    splines = DeviceSplines()
    ll_ref, g_ref = oracle.lnlike_theta(patch, t)
    tol = lambda g: (np.abs(g - g_ref) <= 1e-4 * np.abs(g_ref)) | \
        (np.abs(g - g_ref) <= 1e-6 * np.abs(g_ref).max())
    try:
        tt = torch.as_tensor(t, dtype=torch.float64, device="cuda")
        out = {}
        for mode in ("mixed", "mixed_centre"):
            batch = PatchBatch([patch], precision=mode)
            try:
                ll, g = batch.lnlike_and_grad(tt, splines)
                out[mode] = (ll.item(), g.cpu().numpy())
            finally:
                batch.free()
        for mode, (ll, g) in out.items():
            nz = g_ref != 0
            rel = np.abs(g - g_ref)[nz] / np.abs(g_ref)[nz]
            print(f"{mode}: lnlike rel err {abs(ll - ll_ref) / abs(ll_ref):.2e}, "
                  f"{int((~tol(g)).sum())} components outside tolerance, "
                  f"worst rel err of the nonzero ones {rel.max():.2e}", flush=True)
        ll, g = out["mixed_centre"]
        assert abs(ll - ll_ref) <= CHI2_RTOL * abs(ll_ref)
        assert tol(g).all()
    finally:
        splines.free()
        patch.free()
