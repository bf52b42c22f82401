"""Configuration 4 at its full size (512 uneven patches in one batched
launch, the autotuned launch shape), through properties that do not depend
on size: two sweeps agree bit for bit; sampled patches evaluated alone
through the plain drop-in, at the batch's tile size, agree with their batch
results bit for bit; the three smallest patches agree with the FP64 oracle
within the stated tolerance."""
import numpy as np
import pytest
import torch

import bench
import oracle
from forcepho_amd import synth
from forcepho_amd.autotune import load_or_default
from forcepho_amd.batch import PatchBatch
from forcepho_amd.limits import source_struct_dtype
from helpers import CHI2_RTOL, chi2_relerr, grad_mismatch, proposal_for

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def full(hip):
    specs = synth.config4_specs(512)
    print("building 512 patches", flush=True)
    patches = bench.build_patches(specs, log_every=10 ** 6)
    props = [proposal_for(sc, synth.make_proposals(th, len(p.band_N), 1, config=4,
                                                   patch_id=i)[0])
             for i, (p, sc, th) in enumerate(patches)]
    batch = PatchBatch([p for p, _, _ in patches])
    launch = load_or_default("batch")
    batch.set_launch(**launch)
    yield specs, patches, props, batch, launch
    batch.free()
    for p, _, _ in patches:
        p.free()


def _sweep(batch, props):
    batch.set_proposals(props)
    batch.evaluate()
    torch.cuda.synchronize()
    return batch.chi2_host(), batch.dchi2_host()


def test_full_sweep_is_deterministic(full):
    specs, patches, props, batch, launch = full
    c1, d1 = _sweep(batch, props)
    c2, d2 = _sweep(batch, props)
    assert all(np.array_equal(a, b) for a, b in zip(c1, c2))
    assert all(np.array_equal(a, b) for a, b in zip(d1, d2))


def test_sampled_patches_equal_the_plain_drop_in(full):
    specs, patches, props, batch, launch = full
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


def test_smallest_patches_against_the_oracle(full):
    specs, patches, props, batch, launch = full
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
