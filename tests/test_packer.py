"""Host-side packers, layouts and transforms (CPU only)."""
import numpy as np
import pytest

from forcepho_amd import limits, synth
from forcepho_amd.model import BoundedTransform, Transform, stack_grad
from forcepho_amd.patches import AffineWCS, pack_exposures, scale_at_sky, superpixel_order


def test_struct_sizes():
    assert limits.source_struct_dtype.itemsize == 264
    assert limits.response_struct_dtype.itemsize == 840
    assert limits.psf_struct_dtype.itemsize == 28
    # field offsets of class Source (forcepho/src/proposal.cu:13-27)
    off = {k: v[1] for k, v in limits.source_struct_dtype.fields.items()}
    assert off["fluxes"] == 24 and off["mixture_amplitudes"] == 144
    assert off["damplitude_drh"] == 184 and off["damplitude_dnsersic"] == 224


@pytest.mark.parametrize("nx,ny,sx,sy", [(16, 16, 8, 8), (12, 20, 4, 8), (13, 7, 4, 4)])
def test_superpixel_order_is_permutation(nx, ny, sx, sy):
    o = superpixel_order(nx, ny, sx, sy)
    assert np.array_equal(np.sort(o), np.arange(nx * ny))
    # the first block is the top-left sx-by-sy corner, row-major
    first = o[:min(sx, nx) * min(sy, ny)]
    ii, jj = np.divmod(first, ny)
    assert ii.max() < sx and jj.max() < sy
    assert np.all(np.diff(first[:min(sy, ny)]) == 1)


def test_pack_exposures_padding():
    rng = np.random.default_rng(0)
    ims = [rng.normal(size=(10, 10)).astype(np.float32),
           rng.normal(size=(8, 8)).astype(np.float32)]
    ies = [np.ones_like(i) for i in ims]
    out = pack_exposures(ims, ies, [0, 1], superpixel=(4, 4), warp_size=32)
    assert list(out["exposure_N"]) == [128, 64]
    assert list(out["exposure_start"]) == [0, 128]
    assert list(out["band_start"]) == [0, 1] and list(out["band_N"]) == [1, 1]
    assert np.all(out["ierr"][100:128] == 0) and np.all(out["xpix"][100:128] == -1)
    # pixel values land where their coordinates say
    x, y = out["xpix"][:100].astype(int), out["ypix"][:100].astype(int)
    assert np.array_equal(out["data"][:100], ims[0][x, y])


def test_scale_at_sky_matches_affine_wcs():
    th = np.deg2rad(3.0)
    R = np.array([[np.cos(th), -np.sin(th)], [np.sin(th), np.cos(th)]])
    D = R / 0.03
    wcs = AffineWCS([31.5, 31.5], (53.0, -27.0), np.linalg.inv(D) / 3600.0)
    CW, Dm = scale_at_sky((53.0, -27.0), wcs)
    assert np.allclose(Dm, D)
    assert np.allclose(CW, 3600 * D @ np.diag([np.cos(np.deg2rad(-27.0)), 1.0]))
    # CW is the derivative of world_to_pixel_values
    h = 1e-7
    x0 = np.array(wcs.world_to_pixel_values(53.0, -27.0))
    x1 = np.array(wcs.world_to_pixel_values(53.0 + h, -27.0))
    assert np.allclose((x1 - x0) / h, CW[:, 0], rtol=1e-6)


def test_stack_grad_order():
    nb, ns = 3, 2
    d = np.arange(nb * ns * 7, dtype=float).reshape(nb, ns, 7)
    g = stack_grad(d, ns, nb).reshape(ns, nb + 6)
    for s in range(ns):
        assert np.array_equal(g[s, :nb], d[:, s, 0])
        assert np.array_equal(g[s, nb:], d[:, s, 1:].sum(axis=0))


def test_bounded_transform():
    lo, hi = np.array([0.0, -1.0, 2.0]), np.array([1.0, 1.0, 5.0])
    t = BoundedTransform(lo, hi)
    z = np.array([-0.7, 0.1, 2.3])
    q = t.transform(z)
    assert np.all((q > lo) & (q < hi))
    assert np.allclose(t.inverse_transform(q), z)
    h = 1e-6
    fd = (t.transform(z + h) - t.transform(z - h)) / (2 * h)
    assert np.allclose(t.jacobian(z), fd, rtol=1e-7)
    # d ln|J| / dz
    fdj = (t.lndetjac(z + h * np.eye(3)[1]) - t.lndetjac(z - h * np.eye(3)[1])) / (2 * h)
    assert np.isclose(t.lndetjac_grad(z)[1], fdj, rtol=1e-6)
    assert np.array_equal(Transform(3).transform(z), z)


def test_config_specs_sizes():
    s2 = synth.config_spec(2)
    assert s2.n_exp == 32 and s2.npix_unpadded == 524288
    assert s2.n_pairs() == 524288 * 20 * 36
    specs = synth.config4_specs(16)
    assert all(1 <= s.n_sources <= 30 and 1 <= s.n_bands <= 10 for s in specs)
    assert all(32 <= s.side <= 256 and np.all(s.n_exp_per_band <= 16) for s in specs)


def test_config5_specs_crowded_tail():
    specs = synth.config5_specs(4096)
    crowded = [s for s in specs if s.crowded]
    assert len(specs) == 4096 and len(crowded) == 205
    for s in crowded:
        assert s.n_sources == 60 and s.n_bands == 12 and s.n_exp == 24 and s.side == 256
    # the rest follow the configuration-4 distribution
    rest = [s for s in specs if not s.crowded]
    assert max(s.n_sources for s in rest) <= 30 and max(s.n_bands for s in rest) <= 10
    npix = sum(s.npix_unpadded for s in specs)
    assert 4.5e9 < npix < 5.5e9   # about 39 GB of data + ierr (f32)
    # seeded: the same call gives the same patches
    again = synth.config5_specs(4096)
    assert [s.n_pairs() for s in again] == [s.n_pairs() for s in specs]
