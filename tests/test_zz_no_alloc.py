"""The zero-allocation entry point fpho_evaluate_async_noalloc, through the C
ABI: bit for bit the results of fpho_evaluate_async on two calls in a row
(the second on a smaller patch, so the shared scratch array holds the first
call's contents), the stale-descriptor refusal and the size limit."""
import ctypes

import numpy as np
import pytest
import torch

from forcepho_amd import _lib, synth
from forcepho_amd.limits import source_struct_dtype
from forcepho_amd.synth import PatchSpec
from helpers import make_config_patch, proposal_for

pytestmark = pytest.mark.gpu

SCRATCH_BYTES = 64 << 20


def _dev(a):
    return torch.as_tensor(np.ascontiguousarray(a)).cuda()


def _outs(patch):
    nb, ns = len(patch.band_N), int(patch.n_sources)
    return (torch.zeros(nb, dtype=torch.float64, device="cuda"),
            torch.zeros(nb * ns * 7, dtype=torch.float64, device="cuda"))


def _call(fn, patch, prop_dev, chi2, dchi2):
    stream = ctypes.c_void_p(int(torch.cuda.current_stream().cuda_stream))
    return fn(patch.device_patch, _lib.dptr(prop_dev), _lib.dptr(chi2), _lib.dptr(dchi2),
              stream)


def _both(patch, prop):
    """(async chi2, dchi2), (noalloc chi2, dchi2) for one proposal."""
    lib = _lib.lib()
    prop_dev = _dev(np.asarray(prop, dtype=source_struct_dtype).view(np.uint8))
    a = _outs(patch)
    _lib.check(_call(lib.fpho_evaluate_async, patch, prop_dev, *a))
    b = _outs(patch)
    _lib.check(_call(lib.fpho_evaluate_async_noalloc, patch, prop_dev, *b))
    torch.cuda.synchronize()
    return [x.cpu().numpy() for x in a], [x.cpu().numpy() for x in b]


def test_noalloc_equals_async_on_two_calls(hip):
    big, sc_big, th_big = make_config_patch(2)
    small, sc_small, th_small = make_config_patch(1)
    big.send_to_device()
    small.send_to_device()
    try:
        pb = proposal_for(sc_big, synth.make_proposals(th_big, 8, 1, config=2)[0])
        ps = proposal_for(sc_small, synth.make_proposals(th_small, 1, 1, config=1)[0])
        for label, patch, prop in (("config 2", big, pb), ("config 1 (smaller)", small, ps)):
            (c_a, d_a), (c_n, d_n) = _both(patch, prop)
            print(f"{label}: chi2 {c_a}, noalloc equal: {np.array_equal(c_a, c_n)}",
                  flush=True)
            assert np.array_equal(c_a, c_n) and np.array_equal(d_a, d_n), label
            assert np.all(np.isfinite(c_n)) and np.any(d_n != 0)
    finally:
        big.free()
        small.free()


def test_noalloc_refuses_stale_descriptors(hip):
    patch, scene, theta = make_config_patch(1)
    patch.send_to_device()
    try:
        prop_dev = _dev(np.asarray(proposal_for(scene, theta), dtype=source_struct_dtype)
                        .view(np.uint8))
        chi2, dchi2 = _outs(patch)
        lib = _lib.lib()
        # never evaluated: no descriptors yet
        assert _call(lib.fpho_evaluate_async_noalloc, patch, prop_dev, chi2, dchi2) == 4
        _lib.check(_call(lib.fpho_evaluate_async, patch, prop_dev, chi2, dchi2))
        _lib.check(_call(lib.fpho_evaluate_async_noalloc, patch, prop_dev, chi2, dchi2))
        patch.replace_device_meta_ptrs()      # new metadata: descriptors stale
        assert _call(lib.fpho_evaluate_async_noalloc, patch, prop_dev, chi2, dchi2) == 4
        torch.cuda.synchronize()
    finally:
        patch.free()


def test_noalloc_rejects_work_buffers_beyond_the_scratch(hip):
    """100 sources, 24 exposures of 256^2 pixels, one wave tile per 64 pixels:
    image Gaussians and tile partials need more than the 64 MiB scratch."""
    spec = PatchSpec(n_sources=100, n_bands=1, n_exp_per_band=24, side=256, n_psf=6)
    patch, scene, theta = synth.make_patch(spec, config=47, patch_id=0)
    patch.send_to_device()
    try:
        patch.set_launch(pix_per_lane=1, waves_per_group=1, persistent=0)
        n_gauss = 24 * 100 * 54
        n_tiles = 24 * 256 * 256 // 64
        need = n_gauss * (128 + 64 + 16) + n_tiles * 100 * 7 * 4 + n_tiles * 8
        assert need > SCRATCH_BYTES
        prop_dev = _dev(np.asarray(proposal_for(scene, theta), dtype=source_struct_dtype)
                        .view(np.uint8))
        chi2, dchi2 = _outs(patch)
        lib = _lib.lib()
        _lib.check(_call(lib.fpho_evaluate_async, patch, prop_dev, chi2, dchi2))
        rc = _call(lib.fpho_evaluate_async_noalloc, patch, prop_dev, chi2, dchi2)
        msg = lib.fpho_last_error().decode()
        print("limit:", rc, msg, flush=True)
        assert rc == 2 and "scratch" in msg
        torch.cuda.synchronize()
    finally:
        patch.free()
