"""forcepho's verification image on the HIP path: the drop-in residual
against the reference renderer's image (the reference's own validation asks
5e-7, verify_reference.py:90-114), lnlike and gradient against the FP64
oracle and against the image's derivative images, and the parameter-vector
path with the smooth-spline amplitudes (shared knots, evaluated on the
device) against the oracle fed scipy's SmoothBivariateSpline amplitudes."""
import warnings

import numpy as np
import pytest
import torch

import oracle
import verification_image as V
from forcepho_amd.autograd import LnLike
from forcepho_amd.batch import DeviceSplines, PatchBatch
from forcepho_amd.patches import SimplePatch
from forcepho_amd.spline import get_splines
from helpers import CHI2_RTOL, chi2_relerr, grad_mismatch

pytestmark = [pytest.mark.gpu, pytest.mark.filterwarnings("ignore::UserWarning")]


def _noisy(p, model, sigma, seed):
    real = p.xpix >= 0
    xi, yi = p.xpix[real].astype(int), p.ypix[real].astype(int)
    rng = np.random.default_rng(seed)
    p.data = np.zeros_like(p.data)
    p.data[real] = model[xi, yi] + rng.normal(0.0, sigma, real.sum())
    p.ierr = np.where(real, 1.0 / sigma, 0.0).astype(np.float32)
    return real, xi, yi


def test_residual_of_the_reference_image(hip, tmp_path):
    model, _, _ = V.images()
    p = V.build(str(tmp_path), patch_class=SimplePatch)
    p.return_residual = True
    real = p.xpix >= 0
    xi, yi = p.xpix[real].astype(int), p.ypix[real].astype(int)
    p.send_to_device()
    try:
        prop = p.scene.get_proposal()
        chi2, dchi2, res = p.get_proposer().evaluate_proposal(prop, patch=p, unpack=False)
        # the image minus the HIP model: 2e-8 from the FP32 position
        # metadata, the rest FP32 rounding of the data and the terms
        assert np.abs(res[real]).max() < 5e-8
        assert np.array_equal(p.data[real], model[xi, yi].astype(np.float32))
        c_ref, d_ref = oracle.evaluate(p, prop)
        assert chi2_relerr(chi2, c_ref).max() <= CHI2_RTOL
    finally:
        p.free()


def test_gradient_against_oracle_and_derivative_images(hip, tmp_path):
    model, grads, _ = V.images()
    p = V.build(str(tmp_path), patch_class=SimplePatch)
    p.return_residual = True
    sigma = 1e-3
    real, xi, yi = _noisy(p, model, sigma, seed=7)
    p.send_to_device()
    try:
        prop = p.scene.get_proposal()
        chi2, dchi2, res = p.get_proposer().evaluate_proposal(prop, patch=p, unpack=False)
        c_ref, d_ref = oracle.evaluate(p, prop)
        assert chi2_relerr(chi2, c_ref).max() <= CHI2_RTOL
        assert not grad_mismatch(dchi2, d_ref).any()
        g = np.asarray(dchi2, dtype=np.float64)[0, 0]
        expect = np.array([-2.0 * np.sum(np.asarray(res, np.float64)[real] * grads[j][xi, yi])
                           / sigma ** 2 for j in range(7)])
        keep = [0, 1, 2, 3, 5, 6]     # pa: zero for a q = 1 source
        assert np.all(np.abs(g[keep] - expect[keep]) <= 1e-4 * np.abs(expect[keep])), \
            (g, expect)
        assert abs(g[4]) <= 1e-6 * np.abs(expect).max()
    finally:
        p.free()


def _smooth_sources(theta):
    """FP64 source records of theta ([flux, ra, dec, q, pa, n, rh]) with
    amplitudes from scipy's SmoothBivariateSpline on the table."""
    from scipy.interpolate import SmoothBivariateSpline
    f = np.load(V.INPUTS)
    n, rh = theta[5], theta[6]
    fits = [SmoothBivariateSpline(f["nsersic"], f["rh"], f["amplitudes"][:, m])
            for m in range(f["amplitudes"].shape[1])]
    return [{"fluxes": theta[:1], "ra": theta[1], "dec": theta[2], "q": theta[3],
             "pa": theta[4],
             "mixture_amplitudes": np.array([s(n, rh)[0, 0] for s in fits]),
             "damplitude_dnsersic": np.array([s(n, rh, dx=1)[0, 0] for s in fits]),
             "damplitude_drh": np.array([s(n, rh, dy=1)[0, 0] for s in fits])}]


def test_theta_path_with_smooth_splines(hip, tmp_path):
    model, _, _ = V.images()
    p = V.build(str(tmp_path), patch_class=SimplePatch)
    _noisy(p, model, 1e-3, seed=8)
    p.send_to_device()
    batch = PatchBatch([p])
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        splines = DeviceSplines(splines=get_splines(V.INPUTS, spline_type="smooth"))
    try:
        truth = np.asarray(p.scene.get_all_source_params(), dtype=np.float64)
        rng = np.random.default_rng(9)
        for k in range(4):
            t = truth.copy()
            if k:
                t += rng.normal(0.0, [0.02, 2e-6, 2e-6, 0.02, 0.05, 0.05, 0.005])
                t[3] = min(t[3], 0.99)
            tt = torch.as_tensor(t, dtype=torch.float64, device="cuda").requires_grad_(True)
            ll = LnLike.apply(tt, batch, splines)
            ll.sum().backward()
            chi2, dchi2 = oracle.evaluate64(p, _smooth_sources(t))
            ll_ref, g_ref = -0.5 * chi2.sum(), -0.5 * oracle.stack_grad(dchi2)
            assert abs(ll.item() - ll_ref) <= CHI2_RTOL * abs(ll_ref), (k, ll.item(), ll_ref)
            g = tt.grad.cpu().numpy()
            err = np.abs(g - g_ref)
            assert np.all((err <= 1e-4 * np.abs(g_ref)) | (err <= 1e-6 * np.abs(g_ref).max())), \
                (k, g, g_ref)
    finally:
        splines.free()
        batch.free()
        p.free()
