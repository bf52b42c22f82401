"""Shared builders and tolerance checks for the tests (test code only)."""
import numpy as np

import oracle
from forcepho_amd import synth

# Stated tolerances: chi2 relative error <= 1e-6 against the FP64 oracle;
# each gradient component relative error <= 1e-4 or absolute error
# <= 1e-6 * max |gradient| (of that band).
CHI2_RTOL = 1e-6
GRAD_RTOL = 1e-4
GRAD_ATOL_FRAC = 1e-6


def oracle_render(patch, proposal):
    return oracle.render(patch, proposal)


def make_config_patch(config, patch_id=0, return_residual=False):
    spec = synth.config_spec(config)
    return synth.make_patch(spec, config=config, patch_id=patch_id,
                            return_residual=return_residual, render=oracle_render)


def proposal_for(scene, theta):
    scene.set_all_source_params(theta)
    return scene.get_proposal().copy()


def grad_mismatch(d, d_ref):
    """Boolean mask of gradient components outside the stated tolerance."""
    d = np.asarray(d, dtype=np.float64)
    d_ref = np.asarray(d_ref, dtype=np.float64)
    scale = np.abs(d_ref).reshape(d_ref.shape[0], -1).max(axis=1)
    atol = GRAD_ATOL_FRAC * scale[:, None, None]
    err = np.abs(d - d_ref)
    return (err > GRAD_RTOL * np.abs(d_ref)) & (err > atol)


def chi2_relerr(c, c_ref):
    return np.abs(np.asarray(c) - c_ref) / np.abs(c_ref)


def small_uneven_specs():
    """Four small uneven patches (configuration-4 shape, sizes the oracle
    evaluates in well under a second)."""
    from forcepho_amd.synth import PatchSpec
    return [
        PatchSpec(n_sources=3, n_bands=2, n_exp_per_band=[2, 1], side=48, n_psf=[3, 4]),
        PatchSpec(n_sources=1, n_bands=1, n_exp_per_band=1, side=32, n_psf=3),
        PatchSpec(n_sources=6, n_bands=3, n_exp_per_band=[1, 3, 2], side=64, n_psf=[5, 3, 6],
                  psf_per_exposure=True, psf_elliptical=True, dither=1.5),
        PatchSpec(n_sources=2, n_bands=1, n_exp_per_band=4, side=40, n_psf=4, mask_frac=0.05),
    ]


def make_small_patches(config=40):
    from forcepho_amd import synth
    out = []
    for i, spec in enumerate(small_uneven_specs()):
        out.append(synth.make_patch(spec, config=config, patch_id=i, render=oracle_render))
    return out


# ----------------------------------------------------------- oracle fixtures
class Fixture(dict):
    """Arrays of a committed oracle fixture (tests/golden/*.npz, written by
    tests/golden/make_scale_fixtures.py) and the digest it stores inputs by."""

    @staticmethod
    def sha256(a):
        import hashlib
        return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def load_fixture(path):
    with np.load(path, allow_pickle=False) as f:
        return Fixture({k: f[k] for k in f.files})


def fixture_patches(specs, fx, config):
    """{index: (patch, scene, theta)} of the fixture's patches, rebuilt from
    their seeds with data = FP64 oracle model of the truth + seeded noise;
    each patch's float32 data must hash to the fixture's digest."""
    out = {}
    for k, i in enumerate(fx["patch"]):
        i = int(i)
        patch, scene, theta = synth.make_patch(specs[i], config=config, patch_id=i,
                                               render=oracle_render)
        assert Fixture.sha256(patch.data) == fx["data_sha256"][k], \
            f"patch {i}: rebuilt data differ from the fixture's; regenerate the fixtures"
        out[i] = (patch, scene, theta)
    return out
