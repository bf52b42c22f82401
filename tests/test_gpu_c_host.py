"""libfpho_hip.so driven by a plain C program, with no Python or torch in its
process (tests/c_host/evaluate_patch.c: fpho_init, fpho_patch_create,
fpho_evaluate and the compat fpho_evaluate_response per proposal,
fpho_patch_free -- the calls a C binding of forcepho's
GPUProposer.evaluate_proposal makes, forcepho/proposal.py:96-152).
Its results must equal the Python drop-in's bit for bit (one library, a
fixed tile partition and reduction order) and match the FP64 oracle.  The
program is built by __graft_entry__.build() with the library."""
import os
import subprocess

import numpy as np
import pytest

import oracle
from forcepho_amd import synth
from forcepho_amd.limits import MAXSOURCES, psf_struct_dtype, source_struct_dtype
from helpers import CHI2_RTOL, chi2_relerr, grad_mismatch, make_config_patch, proposal_for

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST = os.path.join(ROOT, "tests", "c_host", "evaluate_patch")


def _dump(patch, props, d):
    """Write the packed patch and the proposals as the raw files the C host
    reads: the arrays the Python drop-in hands to fpho_patch_create, after
    its casts to the fpho_patch_desc dtypes."""
    patch._desc()
    keep, patch._desc_keep = patch._desc_keep, None
    suffix = {np.dtype(np.float32): ".f32", np.dtype(np.int16): ".i16",
              np.dtype(np.int32): ".i32", psf_struct_dtype: ".bin"}
    for name, arr in keep.items():
        arr.tofile(os.path.join(d, name + suffix[arr.dtype]))
    stacked = np.stack([np.asarray(p).view(source_struct_dtype).reshape(-1) for p in props])
    np.ascontiguousarray(stacked).tofile(os.path.join(d, "proposals.bin"))
    with open(os.path.join(d, "meta.txt"), "w") as f:
        f.write(f"{int(patch.npix)} {len(patch.band_N)} {len(patch.exposure_N)} "
                f"{int(patch.n_sources)} {int(patch.n_radii)} {len(keep['psfgauss'])} "
                f"{len(props)}\n")


@pytest.mark.parametrize("config", [1, 2])
def test_c_host_equals_python_dropin(hip, tmp_path, config):
    assert os.access(HOST, os.X_OK), f"{HOST} is not built: run __graft_entry__.build()"
    patch, scene, theta = make_config_patch(config)
    nb, ns = len(patch.band_N), int(patch.n_sources)
    thetas = [theta] + list(synth.make_proposals(theta, nb, 2, config=config, patch_id=7))
    props = [proposal_for(scene, t) for t in thetas]
    _dump(patch, props, str(tmp_path))
    print(f"config {config}: C host on {len(props)} proposals, {ns} sources, {nb} bands",
          flush=True)
    run = subprocess.run([HOST, str(tmp_path)], capture_output=True, text=True, timeout=120)
    print(run.stdout, run.stderr, flush=True)
    assert run.returncode == 0, run.stderr
    chi2_c = np.fromfile(tmp_path / "chi2.f64", dtype=np.float64).reshape(len(props), nb)
    dchi2_c = np.fromfile(tmp_path / "dchi2.f32",
                          dtype=np.float32).reshape(len(props), nb, ns, 7)

    patch.send_to_device()
    try:
        proposer = patch.get_proposer()
        for k, prop in enumerate(props):
            chi2, dchi2 = proposer.evaluate_proposal(prop, patch=patch)[:2]
            assert np.array_equal(chi2_c[k], chi2), k
            assert np.array_equal(dchi2_c[k], dchi2), k
        # the compat Response output: rows [0, n_sources) of each band equal
        # the compact gradients bit for bit, the rows past them are zero
        chi2_r = np.fromfile(tmp_path / "chi2_response.f64",
                             dtype=np.float64).reshape(len(props), nb)
        resp = np.fromfile(tmp_path / "response.bin", dtype=np.float32).reshape(
            len(props), nb, MAXSOURCES, 7)
        assert np.array_equal(chi2_r, chi2_c)
        assert np.array_equal(resp[:, :, :ns], dchi2_c)
        assert not resp[:, :, ns:].any()
        c_ref, d_ref = oracle.evaluate(patch, props[0])
        assert chi2_relerr(chi2_c[0], c_ref).max() <= CHI2_RTOL
        assert not grad_mismatch(dchi2_c[0], d_ref).any()
    finally:
        patch.free()
