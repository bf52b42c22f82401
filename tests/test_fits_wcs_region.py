"""FITS headers and files, the TAN world coordinate system, scale_at_sky and
sky regions (host code of the store-backed packers)."""
import numpy as np
import pytest

from forcepho_amd.patches import fits
from forcepho_amd.patches.wcs import AffineWCS, TanWCS, WCS, scale_at_sky
from forcepho_amd.region import (CircularRegion, RectangularRegion, polygon_contains,
                                 polygons_overlap)


# ------------------------------------------------------------------ FITS
def test_header_cards_round_trip():
    h = fits.Header()
    h["SIMPLE"] = True
    h["BITPIX"] = -32
    h["FILTER"] = ("F200W", "the band")
    h["OBJECT"] = "O'Brien field"
    h["CRVAL1"] = 53.1625
    h["CD1_1"] = -8.5e-6
    h["EXPTIME"] = 1374
    h["FLAG"] = False
    h.add_comment("a comment card")
    h.add_history("made by a test")
    s = h.tostring()
    assert len(s) % 2880 == 0
    # fixed-format columns of the FITS standard
    assert s[:80] == "SIMPLE  =                    T" + " " * 50
    assert s[80:110] == "BITPIX  =                  -32"
    g = fits.Header.fromstring(s)
    for k in ("SIMPLE", "BITPIX", "FILTER", "OBJECT", "CRVAL1", "CD1_1", "EXPTIME", "FLAG"):
        assert g[k] == h[k], k
        assert type(g[k]) is type(h[k]), k
    assert g.cards[2][2] == "the band"
    assert ("COMMENT", "a comment card", "") in g.cards
    assert ("HISTORY", "made by a test", "") in g.cards
    with pytest.raises(KeyError):
        g["NOPE"]


def test_header_parses_d_exponent_and_blank_values():
    card = lambda s: f"{s:<80}"
    s = card("GAIN    =             1.5D+00 / e/ADU") + card("BLANKV  =") + card("END")
    g = fits.Header.fromstring(s)
    assert g["GAIN"] == 1.5 and g["BLANKV"] is None


def test_fits_file_round_trip(tmp_path):
    rng = np.random.default_rng(1)
    sci = rng.normal(size=(5, 7)).astype(np.float32)      # [NAXIS2, NAXIS1]
    err = rng.uniform(1, 2, size=(5, 7))
    cnt = rng.integers(-100, 100, size=(3, 4)).astype(np.int16)
    path = str(tmp_path / "x.fits")
    fits.writeto(path, [({"ORIGIN": "test"}, None),
                        ({"EXTNAME": "SCI", "FILTER": "F200W"}, sci),
                        ({"EXTNAME": "ERR"}, err),
                        ({"EXTNAME": "CNT", "BZERO": 10.0, "BSCALE": 2.0}, cnt)])
    raw = open(path, "rb").read()
    assert len(raw) % 2880 == 0
    h = fits.getheader(path, "SCI")
    assert h["NAXIS1"] == 7 and h["NAXIS2"] == 5 and h["FILTER"] == "F200W"
    assert h["XTENSION"] == "IMAGE"
    np.testing.assert_array_equal(fits.getdata(path, "SCI"), sci)
    np.testing.assert_array_equal(fits.getdata(path, 2), err)
    np.testing.assert_array_equal(fits.getdata(path, "CNT"), cnt * 2.0 + 10.0)
    assert fits.getdata(path, 0) is None
    assert fits.getheader(path)["ORIGIN"] == "test"
    with pytest.raises(KeyError):
        fits.getdata(path, "NOPE")


# ------------------------------------------------------------------- WCS
def _tan(rot_deg=20.0, scale=0.03 / 3600, crval=(53.16, -27.78), crpix=(1024.5, 980.0)):
    t = np.deg2rad(rot_deg)
    R = np.array([[np.cos(t), -np.sin(t)], [np.sin(t), np.cos(t)]])
    cd = R @ np.diag([-scale, scale])
    return TanWCS(crpix, crval, cd)


def spherical_rotation_pixel(w, ra, dec):
    """RA---TAN by the native-spherical route of the FITS WCS paper
    (Calabretta & Greisen 2002, eqs. 5 and 54): an independent statement of
    the projection (the module uses tangent-plane coordinates)."""
    a, d = np.deg2rad(ra), np.deg2rad(dec)
    a0, d0 = np.deg2rad(w.crval)
    phi_p = np.deg2rad(w.lonpole)
    theta = np.arcsin(np.sin(d) * np.sin(d0) + np.cos(d) * np.cos(d0) * np.cos(a - a0))
    phi = phi_p + np.arctan2(-np.cos(d) * np.sin(a - a0),
                             np.sin(d) * np.cos(d0) - np.cos(d) * np.sin(d0) * np.cos(a - a0))
    r = np.rad2deg(1.0 / np.tan(theta))
    p = np.linalg.solve(w.cd, np.stack([r * np.sin(phi), -r * np.cos(phi)]))
    return p[0] + w.crpix[0] - 1, p[1] + w.crpix[1] - 1


def test_tan_reference_point_and_closed_forms():
    w = _tan()
    x, y = w.world_to_pixel_values(*w.crval)
    assert abs(x - (w.crpix[0] - 1)) < 1e-9 and abs(y - (w.crpix[1] - 1)) < 1e-9
    # due north by t: eta = tan(t) exactly, xi = 0
    u = TanWCS((1.0, 1.0), (10.0, 40.0), np.diag([-1e-4, 1e-4]))
    for t in (0.01, 0.3, 2.0):
        x, y = u.world_to_pixel_values(10.0, 40.0 + t)
        assert abs(x) < 1e-9
        assert abs(y - np.rad2deg(np.tan(np.deg2rad(t))) / 1e-4) < 1e-6 * abs(y)


def test_tan_matches_spherical_rotation_and_round_trips():
    w = _tan()
    rng = np.random.default_rng(2)
    ra = w.crval[0] + rng.uniform(-0.5, 0.5, 200) / np.cos(np.deg2rad(w.crval[1]))
    dec = w.crval[1] + rng.uniform(-0.5, 0.5, 200)
    x, y = w.world_to_pixel_values(ra, dec)
    gx, gy = spherical_rotation_pixel(w, ra, dec)
    assert np.max(np.abs(x - gx)) < 1e-5 and np.max(np.abs(y - gy)) < 1e-5
    # a LONPOLE other than 180 rotates the tangent plane
    v = TanWCS(w.crpix, w.crval, w.cd, lonpole=150.0)
    vx, vy = v.world_to_pixel_values(ra, dec)
    gx, gy = spherical_rotation_pixel(v, ra, dec)
    assert np.max(np.abs(vx - gx)) < 1e-5 and np.max(np.abs(vy - gy)) < 1e-5
    vx3, vy3 = v.world_to_pixel_values(*v.pixel_to_world_values(vx, vy))
    assert np.max(np.abs(vx3 - vx)) < 1e-7 and np.max(np.abs(vy3 - vy)) < 1e-7
    r2, d2 = w.pixel_to_world_values(x, y)
    assert np.max(np.abs(r2 - ra)) * 3600 < 1e-7 and np.max(np.abs(d2 - dec)) * 3600 < 1e-7
    px = rng.uniform(0, 2048, 100)
    py = rng.uniform(0, 2048, 100)
    x3, y3 = w.world_to_pixel_values(*w.pixel_to_world_values(px, py))
    assert np.max(np.abs(x3 - px)) < 1e-7 and np.max(np.abs(y3 - py)) < 1e-7
    # origin convention of the all_* methods
    p1 = w.all_world2pix([[ra[0], dec[0]]], 1)[0]
    assert np.allclose(p1, [x[0] + 1, y[0] + 1])
    assert np.allclose(w.all_pix2world([p1], 1)[0], [ra[0], dec[0]], atol=1e-10)


def test_tan_header_forms_and_refusals():
    w = _tan()
    h = fits.Header(w.to_header())
    w2 = WCS(h)
    assert np.allclose(w2.cd, w.cd) and np.allclose(w2.crpix, w.crpix)
    # PC x CDELT form gives the same CD matrix
    h3 = fits.Header({"CTYPE1": "RA---TAN", "CTYPE2": "DEC--TAN", "CRPIX1": 10.0, "CRPIX2": 20.0,
                      "CRVAL1": 150.0, "CRVAL2": 2.0, "CDELT1": -2e-5, "CDELT2": 2e-5,
                      "PC1_1": 0.8, "PC1_2": -0.6, "PC2_1": 0.6, "PC2_2": 0.8})
    w3 = WCS(h3)
    assert np.allclose(w3.cd, [[0.8 * -2e-5, -0.6 * -2e-5], [0.6 * 2e-5, 0.8 * 2e-5]])
    sip = fits.Header(dict(w.to_header(), A_ORDER=3))
    with pytest.raises(NotImplementedError):
        WCS(sip)
    sin = fits.Header(dict(w.to_header(), CTYPE1="RA---SIN", CTYPE2="DEC--SIN"))
    with pytest.raises(NotImplementedError):
        WCS(sin)


def test_scale_at_sky_branches():
    w = _tan(rot_deg=-35.0)
    sky = w.crval
    CW, D = scale_at_sky(sky, w)
    CWn, Dn = scale_at_sky(sky, w, make_approx=True, dpix=0.5)
    assert np.allclose(CW, CWn, rtol=1e-6)
    assert np.allclose(D, Dn, rtol=1e-6)
    # D is pixels per arcsec of on-sky offset: CW = 3600 D diag(cos dec, 1)
    W = np.diag([np.cos(np.deg2rad(sky[1])), 1.0])
    assert np.allclose(CW, 3600 * D @ W)
    # CW maps a small (dra, ddec) to the pixel offset
    dra, ddec = 2e-5, -1e-5
    x0, y0 = w.world_to_pixel_values(*sky)
    x1, y1 = w.world_to_pixel_values(sky[0] + dra, sky[1] + ddec)
    assert np.allclose(CW @ [dra, ddec], [x1 - x0, y1 - y0], rtol=1e-5)
    # the affine WCS agrees with TAN to first order at the reference point
    a = AffineWCS(w.crpix - 1, w.crval, w.cd)
    assert np.allclose(scale_at_sky(sky, a)[0], CW)


# ---------------------------------------------------------------- regions
def test_polygon_vectors_from_region_docs():
    # the examples of forcepho/region.py:232-247 and 278-291
    a = RectangularRegion(58, 59, 26, 27)
    assert polygons_overlap(a, RectangularRegion(57, 60, 25, 28))
    assert not polygons_overlap(a, RectangularRegion(60, 62, 25, 28))
    assert polygons_overlap(a, RectangularRegion(58.5, 62, 25, 28))
    points = np.atleast_2d([[0.5, 0.5], [0.5, 1.1], [-0.1, 0.5]])
    poly = np.array([[0, 0], [1., 0], [1, 1], [0, 1]])
    assert list(polygon_contains(poly, points)) == [True, False, False]
    poly = np.array([[0, 0], [0., 1], [1, 1], [1, 0]])
    assert list(polygon_contains(poly, points)) == [True, False, False]
    poly = np.array([[0, 0], [0, 3.0], [1, 1], [1, 0]])
    assert list(polygon_contains(poly, points)) == [True, True, False]


def test_circular_region_selects_superpixels_with_a_corner_inside():
    w = _tan(rot_deg=10.0, crpix=(32.5, 32.5), scale=0.1 / 3600)
    s, n = 8, 8
    ll = np.stack(np.meshgrid(np.arange(n) * s, np.arange(n) * s, indexing="ij"), -1)
    off = np.array([(0, 0), (1, 0), (1, 1), (0, 1)]) * (s - 1)
    corners = ll[:, :, None, :] + off
    ra, dec = w.pixel_to_world_values(20.3, 40.6)
    reg = CircularRegion(ra, dec, 1.2 / 3600)        # 12 pixels
    sx, sy = reg.contains(corners[..., 0], corners[..., 1], w)
    d = np.hypot(corners[..., 0] - 20.3, corners[..., 1] - 40.6)
    want = np.where(np.any(d < 12.0, axis=-1))
    # the radius is measured along Dec in pixels: 12 up to TAN distortion
    assert set(zip(sx, sy)) == set(zip(*want))
    bb = reg.bounding_box
    assert bb.shape == (2, 4)
    assert np.allclose(bb[1], dec + np.array([-1, -1, 1, 1]) * 1.2 / 3600)
    rect = RectangularRegion(bb[0].min(), bb[0].max(), bb[1].min(), bb[1].max())
    assert polygons_overlap(reg, rect)
