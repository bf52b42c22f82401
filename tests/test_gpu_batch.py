"""Batched persistent evaluation: per patch identical to the plain drop-in,
independent of the launch shape within tolerance, and within the stated
tolerance of the FP64 oracle."""
import numpy as np
import pytest
import torch

import oracle
from forcepho_amd.batch import PatchBatch
from helpers import (CHI2_RTOL, chi2_relerr, grad_mismatch, make_small_patches,
                     proposal_for)

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def patches(hip):
    ps = make_small_patches()
    for p, _, _ in ps:
        p.send_to_device()
    yield ps
    for p, _, _ in ps:
        p.free()


def _batch_eval(batch, props):
    batch.set_proposals(props)
    batch.evaluate()
    torch.cuda.synchronize()
    return batch.chi2_host(), batch.dchi2_host()


def test_batch_equals_plain_drop_in(patches):
    props = [proposal_for(sc, th) for _, sc, th in patches]
    batch = PatchBatch([p for p, _, _ in patches])
    chi2, dchi2 = _batch_eval(batch, props)
    for (p, sc, th), prop, c, d in zip(patches, props, chi2, dchi2):
        cp, dp = p.get_proposer().evaluate_proposal(prop, patch=p)
        assert np.array_equal(c, cp)
        assert np.array_equal(d.astype(np.float32), dp)
    batch.free()


def test_persistent_and_grid_launch_identical(patches):
    props = [proposal_for(sc, th) for _, sc, th in patches]
    batch = PatchBatch([p for p, _, _ in patches])
    batch.set_launch(persistent=1)
    c1, d1 = _batch_eval(batch, props)
    batch.set_launch(persistent=0)
    c2, d2 = _batch_eval(batch, props)
    for a, b in zip(c1 + d1, c2 + d2):
        assert np.array_equal(a, b)
    batch.free()


@pytest.mark.parametrize("K,W", [(1, 1), (2, 2), (4, 4), (8, 2)])
def test_launch_shapes_within_tolerance(patches, K, W):
    props = [proposal_for(sc, th) for _, sc, th in patches]
    batch = PatchBatch([p for p, _, _ in patches])
    batch.set_launch(pix_per_lane=K, waves_per_group=W, persistent=1)
    chi2, dchi2 = _batch_eval(batch, props)
    for (p, sc, th), prop, c, d in zip(patches, props, chi2, dchi2):
        c_ref, d_ref = oracle.evaluate(p, prop)
        assert chi2_relerr(c, c_ref).max() <= CHI2_RTOL
        assert not grad_mismatch(d, d_ref).any()
    batch.free()


@pytest.mark.parametrize("mode", ["fp32", "fp64_reduce", "mixed", "mixed_centre"])
def test_precision_modes(patches, mode):
    props = [proposal_for(sc, th) for _, sc, th in patches]
    batch = PatchBatch([p for p, _, _ in patches], precision=mode)
    chi2, dchi2 = _batch_eval(batch, props)
    worst = 0.0
    for (p, sc, th), prop, c, d in zip(patches, props, chi2, dchi2):
        c_ref, d_ref = oracle.evaluate(p, prop)
        worst = max(worst, chi2_relerr(c, c_ref).max())
        if mode in ("mixed", "mixed_centre"):
            assert not grad_mismatch(d, d_ref).any()
        else:
            scale = np.abs(d_ref).max()
            assert np.abs(d - d_ref).max() <= 1e-3 * scale
    print(f"precision {mode}: worst chi2 rel err {worst:.2e}", flush=True)
    assert worst <= (CHI2_RTOL if mode != "fp32" else 1e-5)
    batch.free()


def test_box_culling_changes_no_bit(patches):
    """Skipping (tile, Gaussian) pairs whose cutoff box misses the tile gives
    results identical to evaluating every pair."""
    props = [proposal_for(sc, th) for _, sc, th in patches]
    batch = PatchBatch([p for p, _, _ in patches])
    for mode in ("mixed", "fp64_reduce", "mixed_centre"):
        batch.set_precision(mode)
        batch.set_launch(cull=1)
        c1, d1 = _batch_eval(batch, props)
        batch.set_launch(cull=0)
        c0, d0 = _batch_eval(batch, props)
        for a, b in zip(c1 + d1, c0 + d0):
            assert np.array_equal(a, b), mode
    batch.free()
