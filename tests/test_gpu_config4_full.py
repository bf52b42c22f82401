"""Configuration 4 at its full size (512 uneven patches in one batched
launch, the autotuned launch shape): two sweeps agree bit for bit; sampled
patches evaluated alone through the plain drop-in, at the batch's tile size,
agree with their batch results bit for bit; and the FP64 oracle, through the
committed fixtures of tests/golden/config4_patches.npz, checks the three
patches with the most pairs (up to 3.3e9 pairs, 5.6 M pixels), the three
around the median pair count and the patch with the most sources, with the
three smallest patches checked against the oracle live."""
import os

import numpy as np
import pytest
import torch

import bench
import oracle
from forcepho_amd import synth
from forcepho_amd.autotune import load_or_default
from forcepho_amd.batch import PatchBatch
from forcepho_amd.limits import source_struct_dtype
from helpers import (CHI2_RTOL, chi2_relerr, fixture_patches, grad_mismatch, load_fixture,
                     proposal_for)

pytestmark = pytest.mark.gpu

FIXTURE = os.path.join(os.path.dirname(__file__), "golden", "config4_patches.npz")


@pytest.fixture(scope="module")
def full(hip):
    specs = synth.config4_specs(512)
    fx = load_fixture(FIXTURE)
    print("rendering the fixture patches with the oracle", flush=True)
    prebuilt = fixture_patches(specs, fx, config=4)
    print("building 512 patches", flush=True)
    patches = bench.build_patches(specs, log_every=10 ** 6, prebuilt=prebuilt)
    props = [proposal_for(sc, synth.make_proposals(th, len(p.band_N), 1, config=4,
                                                   patch_id=i)[0])
             for i, (p, sc, th) in enumerate(patches)]
    batch = PatchBatch([p for p, _, _ in patches])
    launch = load_or_default("batch")
    batch.set_launch(**launch)
    yield specs, patches, props, batch, launch, fx
    batch.free()
    for p, _, _ in patches:
        p.free()


def _sweep(batch, props):
    batch.set_proposals(props)
    batch.evaluate()
    torch.cuda.synchronize()
    return batch.chi2_host(), batch.dchi2_host()


def test_full_sweep_is_deterministic(full):
    specs, patches, props, batch, launch, fx = full
    c1, d1 = _sweep(batch, props)
    c2, d2 = _sweep(batch, props)
    assert all(np.array_equal(a, b) for a, b in zip(c1, c2))
    assert all(np.array_equal(a, b) for a, b in zip(d1, d2))


def test_sampled_patches_equal_the_plain_drop_in(full):
    specs, patches, props, batch, launch, fx = full
    chi2, dchi2 = _sweep(batch, props)
    rng = np.random.default_rng(11)
    for i in rng.choice(len(patches), 4, replace=False):
        p = patches[i][0]
        p.set_launch(pix_per_lane=launch["pix_per_lane"],
                     waves_per_group=launch["waves_per_group"], persistent=0,
                     groups_per_cu=launch["groups_per_cu"])
        p.return_residual = False   # the residual stays on the device
        cp, dp = p.get_proposer().evaluate_proposal(props[i], patch=p)
        print(f"patch {i}: plain vs batch", flush=True)
        assert np.array_equal(chi2[i], cp)
        assert np.array_equal(dchi2[i].astype(np.float32), dp)


def test_largest_median_and_crowded_patches_against_oracle_fixtures(full):
    """FP64 oracle chi2 and gradients of the 7 fixture patches (the regime
    where FP32 sums over millions of pixels lose the most) against the
    batched sweep."""
    specs, patches, props, batch, launch, fx = full
    chi2, dchi2 = _sweep(batch, props)
    for k, i in enumerate(fx["patch"]):
        i = int(i)
        assert fx["proposal_sha256"][k] == fx.sha256(props[i]), \
            f"patch {i}: the proposal differs from the fixture's; regenerate the fixtures"
        c_ref, d_ref = fx[f"chi2_{i}"], fx[f"dchi2_{i}"]
        rel = chi2_relerr(chi2[i], c_ref)
        bad = grad_mismatch(dchi2[i], d_ref)
        print(f"patch {i}: {specs[i].n_sources} sources, {specs[i].n_pairs():.3e} pairs, "
              f"chi2 rel err {rel.max():.2e}, grad mismatches {bad.sum()}/{bad.size}",
              flush=True)
        assert rel.max() <= CHI2_RTOL
        assert not bad.any(), np.argwhere(bad)[:5]


def test_smallest_patches_against_the_oracle(full):
    specs, patches, props, batch, launch, fx = full
    chi2, dchi2 = _sweep(batch, props)
    for i in np.argsort([s.n_pairs() for s in specs])[:3]:
        p = patches[i][0]
        p.data = p.retrieve_array("data")
        c_ref, d_ref = oracle.evaluate(p, np.asarray(props[i], dtype=source_struct_dtype))
        p.data = None
        rel = chi2_relerr(chi2[i], c_ref)
        bad = grad_mismatch(dchi2[i], d_ref)
        print(f"patch {i}: chi2 rel err {rel.max():.2e}, grad mismatches {bad.sum()}/{bad.size}",
              flush=True)
        assert rel.max() <= CHI2_RTOL
        assert not bad.any()
