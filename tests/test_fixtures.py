"""The committed FP64 oracle fixtures (tests/golden/*_patches.npz,
config2_proposals.npz, config3_theta.npz) that the GPU tests compare
against: their inputs rebuild from the seeds, sampled entries equal the
oracle recomputed here, the finite differences agree with the analytic
gradient, and the leapfrog trajectory is a leapfrog trajectory."""
import os
import sys

import numpy as np
import pytest

import oracle
from forcepho_amd import synth
from helpers import Fixture, fixture_patches, load_fixture, make_config_patch, proposal_for

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
sys.path.insert(0, GOLDEN)
import make_scale_fixtures as msf  # noqa: E402


def test_config4_selection_and_smallest_fixture_patch():
    specs = synth.config4_specs(512)
    fx = load_fixture(os.path.join(GOLDEN, "config4_patches.npz"))
    assert [int(i) for i in fx["patch"]] == msf.config4_selection(specs)
    pairs = np.array([s.n_pairs() for s in specs])
    assert set(np.argsort(pairs)[-3:]) <= set(int(i) for i in fx["patch"])
    assert max(s.n_sources for s in specs) in [specs[int(i)].n_sources for i in fx["patch"]]
    # recompute the cheapest fixture patch
    i = int(fx["patch"][np.argmin(fx["n_pairs"])])
    k = list(fx["patch"]).index(i)
    patch, scene, theta = fixture_patches(specs, {"patch": np.array([i]),
                                                  "data_sha256": fx["data_sha256"][k:k + 1]},
                                          config=4)[i]
    t = synth.make_proposals(theta, len(patch.band_N), 1, config=4, patch_id=i)[0]
    prop = proposal_for(scene, t)
    assert Fixture.sha256(prop) == fx["proposal_sha256"][k]
    c, d = oracle.evaluate(patch, prop)
    np.testing.assert_allclose(c, fx[f"chi2_{i}"], rtol=1e-12)
    np.testing.assert_allclose(d, fx[f"dchi2_{i}"], rtol=1e-10, atol=1e-12 * np.abs(d).max())


def test_config5_selection_has_a_crowded_patch():
    specs = synth.config5_specs(4096)
    fx = load_fixture(os.path.join(GOLDEN, "config5_patches.npz"))
    chosen = [int(i) for i in fx["patch"]]
    assert chosen == msf.config5_selection(specs)
    assert any(specs[i].crowded and specs[i].n_sources == 60 for i in chosen)
    assert fx[f"dchi2_{chosen[3]}"].shape == (12, 60, 7)


def test_config2_fixture_entries_equal_the_oracle():
    fx = load_fixture(os.path.join(GOLDEN, "config2_proposals.npz"))
    patch, scene, theta = make_config_patch(2)
    assert Fixture.sha256(patch.data) == fx["data_sha256"]
    assert np.array_equal(fx["theta"], synth.make_proposals(theta, 8, 100, config=2))
    for k in (0, 63):
        c, d = oracle.evaluate(patch, proposal_for(scene, fx["theta"][k]))
        np.testing.assert_allclose(c, fx["chi2"][k], rtol=1e-12)
        np.testing.assert_allclose(d, fx["dchi2"][k], rtol=1e-10,
                                   atol=1e-12 * np.abs(d).max())


def test_config3_finite_differences_agree_with_the_analytic_gradient():
    fx = load_fixture(os.path.join(GOLDEN, "config3_theta.npz"))
    for k in range(len(fx["fd"])):
        g = fx["grad"][k][fx["fd_components"][k]]
        scale = np.abs(fx["grad"][k]).max()
        err = np.abs(fx["fd"][k] - g)
        assert np.all((err <= 2e-5 * np.abs(g)) | (err <= 1e-7 * scale)), k
    # every parameter kind (8 fluxes + 6 shape) has two components
    kinds = fx["fd_components"] % 14
    assert all(np.array_equal(np.sort(r), np.repeat(np.arange(14), 2)) for r in kinds)


def test_config3_leapfrog_fixture_is_a_leapfrog_trajectory():
    """Reversibility: from the end point with the momentum negated, the
    same steps (through the FP64 oracle) return to the start."""
    fx = load_fixture(os.path.join(GOLDEN, "config3_theta.npz"))
    patch, scene, theta = make_config_patch(2)
    lo, hi = synth.theta_bounds(theta, 8)
    post = msf.sampling_posterior(patch, lo, hi)
    eps, n = float(fx["lf_eps"]), int(fx["lf_steps"])
    z, p = fx["lf_z"].copy(), -fx["lf_p"]
    lnp, g = post(z)
    assert abs(lnp - fx["lf_lnp"]) <= 1e-12 * abs(lnp)
    p = p + 0.5 * eps * g
    for i in range(n):
        z = z + eps * p
        lnp, g = post(z)
        p = p + (eps if i < n - 1 else 0.5 * eps) * g
    scale = np.abs(fx["lf_z"] - fx["lf_z0"]).max()
    assert np.abs(z - fx["lf_z0"]).max() <= 1e-8 * scale
    assert np.abs(-p - fx["lf_p0"]).max() <= 1e-8 * np.abs(fx["lf_p0"]).max()
