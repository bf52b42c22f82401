"""Sersic amplitude splines: this package's own bicubic interpolant against
scipy's RectBivariateSpline (the reference's third-party algorithm,
forcepho/sources.py:623-694) -- golden vectors committed under tests/golden
plus a live comparison -- and the edge branches."""
import os

import numpy as np
import pytest

from forcepho_amd.spline import SersicSplines, notaknot_knots

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "spline_golden.npz")


@pytest.fixture(scope="module")
def sp():
    return SersicSplines()


def test_golden_vectors(sp):
    g = np.load(GOLDEN)
    for (n, r), v, dn, dr in zip(g["points"], g["value"], g["d_nsersic"], g["d_rh"]):
        a, an, ar = sp.evaluate(n, r)
        assert np.allclose(a, v, rtol=0, atol=1e-12)
        assert np.allclose(an, dn, rtol=0, atol=1e-11)
        assert np.allclose(ar, dr, rtol=0, atol=1e-10)


def test_live_scipy(sp):
    from scipy.interpolate import RectBivariateSpline
    s0 = RectBivariateSpline(sp.ngrid, sp.rgrid, sp.table[:, :, 3], s=None)
    tx, ty = s0.get_knots()
    assert np.allclose(tx, sp.tx) and np.allclose(ty, sp.ty)
    assert np.allclose(s0.get_coeffs().reshape(sp.coeffs[3].shape), sp.coeffs[3],
                       rtol=0, atol=1e-12)


def test_interpolates_table(sp):
    for i, j in [(0, 0), (7, 33), (54, 100)]:
        a = sp.amplitudes(sp.ngrid[i], sp.rgrid[j])
        assert np.allclose(a, sp.table[i, j], atol=1e-12)


def test_edge_branches(sp):
    a, dn, dr = sp.evaluate(0.5, 0.1)          # nsersic below the table
    assert a.sum() == 1.0 and a[np.argmin(np.abs(0.1 - sp.radii))] == 1.0
    assert not dn.any() and not dr.any()
    a, dn, dr = sp.evaluate(2.0, -0.01)        # negative rh
    assert a[0] == 1.0 and a.sum() == 1.0 and not dn.any() and not dr.any()


def test_knots_notaknot():
    x = np.arange(10.0)
    t = notaknot_knots(x)
    assert len(t) == len(x) + 4
    assert np.all(t[:4] == 0) and np.all(t[-4:] == 9) and np.all(t[4:-4] == x[2:-2])
