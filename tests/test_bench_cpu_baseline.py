"""bench.py's CPU baselines (the compiled reference kernel, one process and
several processes) on small uneven patches: both run, count the sample's
pairs, and give the same evaluations as a direct call (CPU only; the device
read-back is replaced by the host data)."""
import numpy as np
import pytest

import bench
import oracle
from forcepho_amd import synth
from helpers import small_uneven_specs


@pytest.fixture(scope="module")
def patches():
    if oracle.reference_module() is None:
        pytest.skip("compiled reference kernel not built (make -C oracle ref)")
    specs = small_uneven_specs()
    out = []
    for i, sp in enumerate(specs):
        p, sc, th = synth.make_patch(sp, config=4, patch_id=i)
        p.data = np.random.default_rng(i).normal(size=p.xpix.shape).astype(np.float32)
        host = p.data
        p.retrieve_array = lambda name, host=host: host
        out.append((p, sc, th))
    return specs, out


def test_one_process(patches):
    specs, ps = patches
    c = bench.cpu_baseline(ps, specs, [th for _, _, th in ps], 0.01)
    assert c["patches"] >= 1 and c["pairs_per_s"] > 0
    assert c["sample_pairs"] == sum(specs[i].n_pairs() for i in
                                    np.random.default_rng(1234).permutation(len(ps))[:c["patches"]])


def test_all_cores(patches):
    specs, ps = patches
    c = bench.cpu_baseline_all_cores(ps, specs, [th for _, _, th in ps], 0.01, 2)
    assert c["workers"] == 2 and c["pairs_per_s"] > 0 and c["seconds"] > 0
    assert 1 <= c["patches"] <= len(ps) and c["sample_pairs"] > 0


def test_worker_matches_direct_call(patches):
    """A worker's unit (patch fields and proposal) evaluates as the patch
    does, and a patch without its host pixels is refused, not read."""
    specs, ps = patches
    p, sc, th = ps[0]
    p.data = p.retrieve_array("data")   # the baselines drop the host copy after use
    fields = {name: getattr(p, name) for name in bench._REF_FIELDS}
    prop = sc.get_proposal().copy()
    assert bench._ref_worker([(fields, prop)]) > 0
    c1, d1 = oracle.ReferenceKernel(p).evaluate(prop)
    import types
    c2, d2 = oracle.ReferenceKernel(types.SimpleNamespace(**fields)).evaluate(prop)
    assert np.array_equal(c1, c2) and np.array_equal(d1, d2)
    with pytest.raises(ValueError):
        oracle.ReferenceKernel(types.SimpleNamespace(**dict(fields, data=None)))
