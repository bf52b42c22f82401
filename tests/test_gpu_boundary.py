"""Boundary behaviour of the C ABI on the device: the compat Response
output (forcepho/src/proposal.cu:33-39) equals the native gradients; the
device status word records invalid proposal values that only the device
sees; stream-ordered calls refuse to rebuild a stale plan while a graph is
being captured and refuse a batch whose member changed its source count;
residual writes can be switched off and on."""
import ctypes

import numpy as np
import pytest
import torch

import oracle
from forcepho_amd import _lib, synth
from forcepho_amd.batch import DeviceSplines, PatchBatch
from forcepho_amd.limits import MAXSOURCES, NPARAMS, source_struct_dtype
from forcepho_amd.sources import Scene
from helpers import (CHI2_RTOL, chi2_relerr, grad_mismatch, make_small_patches,
                     oracle_render, proposal_for)

pytestmark = pytest.mark.gpu

STATE = 4   # FPHO_E_STATE
LIMIT = 2   # FPHO_E_LIMIT


@pytest.fixture(scope="module")
def patches(hip):
    ps = make_small_patches(config=41)
    for p, _, _ in ps:
        p.return_residual = True
        p.send_to_device()
    yield ps
    for p, _, _ in ps:
        p.free()


def _props(ps):
    return [proposal_for(sc, th) for _, sc, th in ps]


def test_response_layout_equals_native_gradients(patches):
    p, sc, th = patches[2]
    prop = proposal_for(sc, th)
    pr = p.get_proposer()
    p.set_residual_writes(False)
    chi2, d = pr.evaluate_proposal(prop, patch=p)
    c2, resp = pr.evaluate_response(prop, patch=p)
    ns = p.n_sources
    r = resp.view(np.float32).reshape(len(p.band_N), MAXSOURCES, NPARAMS)
    assert np.array_equal(chi2, c2)
    assert np.array_equal(r[:, :ns], d)
    assert not r[:, ns:].any()
    # stream ordered, device pointers: the same bytes
    dev = torch.device("cuda", 0)
    prop_dev = torch.from_numpy(np.ascontiguousarray(prop, dtype=source_struct_dtype)
                                .view(np.uint8).copy()).to(dev)
    chi2_dev = torch.zeros(len(p.band_N), dtype=torch.float64, device=dev)
    resp_dev = torch.full((len(p.band_N) * resp.dtype.itemsize,), 7, dtype=torch.uint8,
                          device=dev)
    _lib.check(_lib.lib().fpho_evaluate_response_async(
        p.device_patch, _lib.dptr(prop_dev), _lib.dptr(chi2_dev), _lib.dptr(resp_dev),
        ctypes.c_void_p(int(torch.cuda.current_stream().cuda_stream))))
    torch.cuda.synchronize()
    assert np.array_equal(chi2_dev.cpu().numpy(), chi2)
    assert np.array_equal(resp_dev.cpu().numpy(), resp.view(np.uint8).reshape(-1))
    p.set_residual_writes(True)


def test_response_layout_rejects_more_than_maxsources(hip):
    spec = synth.PatchSpec(n_sources=MAXSOURCES + 1, n_bands=1, n_exp_per_band=1, side=32,
                           n_psf=3)
    p, sc, th = synth.make_patch(spec, config=42, patch_id=0)
    p.send_to_device()
    try:
        with pytest.raises(_lib.FphoError) as e:
            p.get_proposer().evaluate_response(sc.get_proposal(), patch=p)
        assert e.value.code == LIMIT
        # the native layout has no such limit
        c, d = p.get_proposer().evaluate_proposal(sc.get_proposal(), patch=p)
        assert d.shape == (1, MAXSOURCES + 1, NPARAMS) and np.isfinite(c).all()
    finally:
        p.free()


def test_status_word_records_invalid_device_proposals(patches):
    ps = patches[:2]
    batch = PatchBatch([p for p, _, _ in ps])
    props = _props(ps)
    batch.set_proposals(props)
    batch.evaluate()
    assert batch.status() == 0
    bad = [x.copy() for x in props]
    bad[1]["q"][0] = -0.5
    batch.set_proposals(bad)
    batch.evaluate()
    w = batch.status(clear=True)
    print("status after q < 0:", hex(w), _lib.describe_status(w), flush=True)
    assert w & 0x2
    assert batch.status() == 0          # cleared by the read before
    # a NaN parameter reaches the device through theta
    splines = DeviceSplines()
    theta = torch.as_tensor(np.concatenate([th for _, _, th in ps]), dtype=torch.float64,
                            device="cuda")
    batch.lnlike_and_grad(theta, splines)
    assert batch.status() == 0
    theta[3] = float("nan")
    batch.lnlike_and_grad(theta, splines)
    w = batch.status()
    print("status after NaN theta:", hex(w), _lib.describe_status(w), flush=True)
    assert w & 0x10 and w & 0x1
    splines.free()
    batch.free()


def test_stale_plan_is_not_rebuilt_inside_a_capture(patches):
    ps = patches[:2]
    batch = PatchBatch([p for p, _, _ in ps])
    batch.set_proposals(_props(ps))
    batch.evaluate()
    torch.cuda.synchronize()
    ps[0][0].replace_device_meta_ptrs()       # same scene, new generation: plan stale
    g = torch.cuda.CUDAGraph()
    s = torch.cuda.Stream()
    with pytest.raises(_lib.FphoError) as e:
        with torch.cuda.graph(g, stream=s):
            batch.evaluate()
    assert e.value.code == STATE
    # outside a capture the plan is rebuilt and the results are those of the
    # plain drop-in
    chi2, dchi2 = batch.evaluate()
    torch.cuda.synchronize()
    p, sc, th = ps[0]
    p.set_residual_writes(False)
    cp, dp = p.get_proposer().evaluate_proposal(proposal_for(sc, th), patch=p)
    p.set_residual_writes(True)
    assert np.array_equal(batch.chi2_host(0), cp)
    batch.free()


def test_batch_refuses_a_member_whose_source_count_changed(hip):
    ps = make_small_patches(config=43)[:3]
    for p, _, _ in ps:
        p.send_to_device()
    try:
        batch = PatchBatch([p for p, _, _ in ps])
        p, sc, th = ps[0]
        assert p.n_sources > 1
        fewer = Scene(catalog=p.truth_catalog[:1], filternames=p.bandlist)
        p.pack_meta(fewer)
        p.replace_device_meta_ptrs()
        with pytest.raises(_lib.FphoError) as e:
            batch.evaluate()
        assert e.value.code == STATE and b"new batch" in _lib.lib().fpho_last_error()
        batch.free()
    finally:
        for p, _, _ in ps:
            p.free()


def test_residual_writes_switch(patches):
    p, sc, th = patches[0]
    pr = p.get_proposer()
    prop = proposal_for(sc, th)
    c, d, res = pr.evaluate_proposal(prop, patch=p, unpack=False)
    p.data = p.retrieve_array("data")
    _, _, r_ref = oracle.evaluate(p, prop, residual=True)
    scale = np.abs(p.data).max()
    assert np.max(np.abs(res - r_ref)) <= 1e-5 * scale
    p.set_residual_writes(False)
    moved = prop.copy()
    moved["fluxes"] *= 1.5
    out = pr.evaluate_proposal(moved, patch=p)
    assert len(out) == 2                       # no residual returned
    assert np.array_equal(p.retrieve_array("residual"), res)   # nothing written
    p.set_residual_writes(True)
    _, _, res2 = pr.evaluate_proposal(moved, patch=p, unpack=False)
    _, _, r_ref2 = oracle.evaluate(p, moved, residual=True)
    assert np.max(np.abs(res2 - r_ref2)) <= 1e-5 * scale
    assert not np.array_equal(res2, res)
