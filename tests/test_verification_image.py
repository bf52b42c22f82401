"""forcepho's verification image (tests/verification/data of the reference,
made by its pure-python renderer) reproduced by the oracle: the model to
2e-8 where the reference's own validation asks 5e-7
(verify_reference.py:90-114), the gradient against the image's seven
derivative images, the chi2 convention of verify_reference.py:116-143, and
the spline branch the image was made with (SmoothBivariateSpline)."""
import warnings

import numpy as np
import pytest

import oracle
import verification_image as V
from forcepho_amd.spline import SersicSplines, common_knots, insert_knot, refine

pytestmark = pytest.mark.filterwarnings("ignore::UserWarning")


@pytest.fixture(scope="module")
def ref():
    return V.images()


def _pixels(p):
    real = p.xpix >= 0
    return real, p.xpix[real].astype(int), p.ypix[real].astype(int)


def test_fixture_contents(ref):
    model, grads, units = ref
    cat, bands = V.catalog()
    assert bands == ["F200W"] and model.shape == (63, 32) and grads.shape == (7, 63, 32)
    assert units == [f"dFLUX/d{p}_0" for p in V.PARAMS]
    assert cat["sersic"][0] == 2.2 and cat["rhalf"][0] == 0.1 and cat["F200W"][0] == 1.0
    assert cat["q"][0] == 1.0 and cat["pa"][0] == 0.0
    # the renderer's stamp puts the source 3.8e-5 pixel off the TAN position
    # of the catalog, (31, 15.5)
    _, _, pos = V.renderer_geometry()
    assert np.allclose(pos, [31.0, 15.5], atol=5e-5) and abs(pos[0] - 31.0) > 3e-5


def test_smooth_splines_equal_fitpack():
    """The package's smooth-spline branch (FITPACK construction, own
    evaluation on shared knots) against scipy's SmoothBivariateSpline, and
    the least-squares amplitudes of the image against it."""
    from scipy.interpolate import SmoothBivariateSpline
    f = np.load(V.INPUTS)
    sp = SersicSplines(V.INPUTS, spline_type="smooth")
    rng = np.random.default_rng(3)
    for x, y in np.c_[rng.uniform(1, 5, 25), rng.uniform(0.03, 0.3, 25)]:
        v, dn, dr = sp.evaluate(x, y)
        for m in range(sp.n_amp):
            s = SmoothBivariateSpline(f["nsersic"], f["rh"], f["amplitudes"][:, m])
            assert abs(v[m] - s(x, y)[0, 0]) <= 1e-13
            assert abs(dn[m] - s(x, y, dx=1)[0, 0]) <= 1e-12
            assert abs(dr[m] - s(x, y, dy=1)[0, 0]) <= 1e-11


def test_image_was_made_with_the_smooth_branch(tmp_path, ref):
    """Fit the 9 mixture amplitudes to the image by least squares (one
    oracle render per component): they are the smooth spline's to 1e-6 and
    far from the interpolating spline's."""
    model, _, _ = ref
    p = V.build(str(tmp_path))
    real, xi, yi = _pixels(p)
    prop = p.scene.get_proposal()
    cols = []
    for m in range(9):
        q = prop.copy()
        q["mixture_amplitudes"][0][:] = 0.0
        q["mixture_amplitudes"][0][m] = 1.0
        cols.append(oracle.render(p, q, cutoff=False)[real])
    fit = np.linalg.lstsq(np.array(cols).T, model[xi, yi], rcond=None)[0]
    smooth = SersicSplines(V.INPUTS, spline_type="smooth").amplitudes(2.2, 0.1)
    rect = SersicSplines(V.INPUTS).amplitudes(2.2, 0.1)
    assert np.abs(fit - smooth).max() < 1e-6
    assert np.abs(fit - rect).max() > 1e-3


@pytest.mark.parametrize("cutoff", [False, True])
def test_oracle_reproduces_the_reference_image(tmp_path, ref, cutoff):
    model, _, _ = ref
    # in the renderer's own coordinates: the remaining 2e-8 is the FP32
    # pixel position of the metadata (half an ulp of 31 is 1e-6 pixel)
    p = V.build(str(tmp_path / "r"))
    real, xi, yi = _pixels(p)
    _, _, m = oracle.evaluate(p, p.scene.get_proposal(), cutoff=cutoff, model=True)
    assert np.abs(m[real] - model[xi, yi]).max() < 5e-8
    # with the packer's exact TAN position the renderer's 3.8e-5 pixel
    # linearisation offset shows, still inside the reference's 5e-7
    p = V.build(str(tmp_path / "t"), geometry="tan")
    _, _, m = oracle.evaluate(p, p.scene.get_proposal(), cutoff=cutoff, model=True)
    assert np.abs(m[real] - model[xi, yi]).max() < 5e-7


def test_oracle_gradient_against_the_derivative_images(tmp_path, ref):
    """With data = image + noise, dchi2/dtheta_j = -2 sum r ierr^2 dC/dtheta_j:
    the oracle's gradient against the same sums over the image's derivative
    images (with the oracle's residual)."""
    model, grads, _ = ref
    p = V.build(str(tmp_path))
    real, xi, yi = _pixels(p)
    sigma = 1e-3
    rng = np.random.default_rng(7)
    p.data = np.zeros_like(p.data)
    p.data[real] = model[xi, yi] + rng.normal(0.0, sigma, real.sum())
    p.ierr = np.where(real, 1.0 / sigma, 0.0).astype(np.float32)
    for cutoff in (False, True):
        _, dchi2, res = oracle.evaluate(p, p.scene.get_proposal(), cutoff=cutoff,
                                        residual=True)
        g = dchi2[0, 0]
        expect = np.array([-2.0 * np.sum(res[real] * grads[j][xi, yi]) / sigma ** 2
                           for j in range(7)])
        scale = np.abs(expect).max()
        # pa: a q = 1 source does not change under rotation, both are zero
        assert abs(g[4]) <= 1e-9 * scale and abs(expect[4]) <= 1e-9 * scale
        keep = [0, 1, 2, 3, 5, 6]
        assert np.all(np.abs(g[keep] - expect[keep]) <= 1e-4 * np.abs(expect[keep])), \
            (g, expect)


def test_chi2_convention_at_the_truth(tmp_path, ref):
    """verify_reference.py:116-143: at the truth the residual chi2 is ~0 and
    the kernel's chi2 is sum (r^2 - d^2) ierr^2."""
    p = V.build(str(tmp_path))
    real, _, _ = _pixels(p)
    chi2, _, res = oracle.evaluate(p, p.scene.get_proposal(), residual=True)
    assert np.sum((res[real] * p.ierr[real]) ** 2) < 2e-7
    d = p.data[real].astype(np.float64) * p.ierr[real]
    full = np.sum((res[real] * p.ierr[real]) ** 2) - np.sum(d ** 2)
    assert np.isclose(chi2[0], full, rtol=1e-10, atol=1e-12)


def test_knot_insertion_keeps_the_spline():
    """Boehm insertion (the merge of the smooth branch's per-component knots)
    leaves a random bicubic spline unchanged, also at repeated knots."""
    from forcepho_amd.spline import basis, knot_span
    rng = np.random.default_rng(11)
    tx = np.r_[[0.0] * 4, 0.3, 0.7, [1.0] * 4]
    ty = np.r_[[0.0] * 4, 0.5, [2.0] * 4]
    c = rng.normal(size=(len(tx) - 4, len(ty) - 4))

    def ev(tx, ty, c, x, y):
        lx, ly = knot_span(tx, x), knot_span(ty, y)
        Nx, _ = basis(tx, lx, x)
        Ny, _ = basis(ty, ly, y)
        return Nx @ c[lx - 3:lx + 1, ly - 3:ly + 1] @ Ny

    target_x = common_knots([tx, np.r_[[0.0] * 4, 0.3, 0.3, 0.5, [1.0] * 4]])
    assert list(target_x) == [0.0] * 4 + [0.3, 0.3, 0.5, 0.7] + [1.0] * 4
    tx2, c2 = refine(tx, c, target_x, 0)
    ty2, c2 = insert_knot(ty, c2, 1.25, 1)
    assert np.array_equal(tx2, target_x) and c2.shape == (len(tx2) - 4, len(ty2) - 4)
    for x, y in rng.uniform([0, 0], [1, 2], size=(50, 2)):
        assert abs(ev(tx, ty, c, x, y) - ev(tx2, ty2, c2, x, y)) <= 1e-13


def test_smoothing_rect_branch_merges_knots():
    """A smoothing RectBivariateSpline gives each component its own knots;
    merged onto shared knots the package's splines still equal scipy's."""
    from scipy.interpolate import RectBivariateSpline
    f = np.load(V.INPUTS)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        sp = SersicSplines(V.INPUTS, spline_smoothing=1e-4)
        ng, rg = np.unique(f["nsersic"]), np.unique(f["rh"])
        ref = [RectBivariateSpline(ng, rg, f["amplitudes"][:, m].reshape(5, 10), s=1e-4)
               for m in range(sp.n_amp)]
    assert len({len(s.get_knots()[1]) for s in ref}) > 1
    for x, y in [(1.3, 0.05), (2.2, 0.1), (4.9, 0.29), (3.0, 0.2)]:
        v, dn, dr = sp.evaluate(x, y)
        assert np.allclose(v, [s(x, y)[0, 0] for s in ref], rtol=0, atol=1e-13)
        assert np.allclose(dr, [s(x, y, dy=1)[0, 0] for s in ref], rtol=0, atol=1e-11)
