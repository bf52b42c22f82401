"""Sersic-mixture amplitudes a_m(nsersic, rh) and their derivatives.

The reference builds one bicubic spline per mixture component
(forcepho/sources.py:623-645) and evaluates value, d/dnsersic and d/drh
(forcepho/sources.py:656-694).  By default (``spline_type="rect"``, no
smoothing) that is ``scipy.interpolate.RectBivariateSpline(ngrid, rgrid, A_m,
s=None)``: the unique bicubic interpolating spline with not-a-knot end
conditions, knots ``[x0]*4 + x[2:-2] + [x_last]*4`` along each axis (FITPACK
regrid).  This module computes that spline itself -- the B-spline
coefficients solve the collocation system along each axis, C = Bn^-1 A Br^-T
-- and evaluates it in FP64 with arguments clamped to the knot range, as
FITPACK's bispev does.

The reference's other choices, a smoothing ``RectBivariateSpline`` (a
``spline_smoothing`` s > 0) and ``SmoothBivariateSpline(n, r, A_m, s)`` (any
other ``spline_type``), place knots by FITPACK's regrid / surfit smoothing
search.  Those are built by the same scipy (FITPACK) calls the reference
makes; each component's knots and coefficients are then re-expressed on the
union of all components' knots by exact knot insertion (Boehm's algorithm),
so that every component shares one (tx, ty) pair.  The shared knots and the
coefficients are what the device-side evaluation uses (fpho_spline_create).
"""

import os

import numpy as np

DEFAULT_SPLINEDATA = os.path.join(os.path.dirname(os.path.abspath(__file__)),
                                  "data", "sersic_splinedata_large.npz")


def notaknot_knots(x):
    """Cubic not-a-knot knot vector for interpolation at sites ``x``."""
    x = np.asarray(x, dtype=np.float64)
    return np.concatenate([[x[0]] * 4, x[2:-2], [x[-1]] * 4])


def knot_span(t, x):
    """Largest l in [3, len(t)-5] with t[l] <= x (x inside [t[3], t[-4]])."""
    n_t = len(t)
    l = int(np.searchsorted(t, x, side="right")) - 1
    return min(max(l, 3), n_t - 5)


def basis(t, l, x):
    """Cubic B-spline values N_{l-3..l}(x) and derivatives (Cox-de Boor)."""
    N = np.zeros(4)
    left = np.zeros(4)
    right = np.zeros(4)
    N[0] = 1.0
    N2 = None
    for j in range(1, 4):
        left[j] = x - t[l + 1 - j]
        right[j] = t[l + j] - x
        saved = 0.0
        for r in range(j):
            tmp = N[r] / (right[r + 1] + left[j - r])
            N[r] = saved + right[r + 1] * tmp
            saved = left[j - r] * tmp
        N[j] = saved
        if j == 2:
            N2 = N[:3].copy()
    dN = np.zeros(4)
    for r in range(4):
        i = l - 3 + r
        a = N2[r - 1] / (t[i + 3] - t[i]) if r >= 1 else 0.0
        b = N2[r] / (t[i + 4] - t[i + 1]) if r <= 2 else 0.0
        dN[r] = 3.0 * (a - b)
    return N, dN


def collocation_matrix(x, t):
    """B[i, j] = N_j(x_i) for the cubic spline space with knots t."""
    m = len(x)
    B = np.zeros((m, m))
    for i, xi in enumerate(x):
        l = knot_span(t, xi)
        N, _ = basis(t, l, xi)
        B[i, l - 3:l + 1] = N
    return B


def insert_knot(t, c, u, axis=0):
    """Boehm's insertion of knot ``u`` (t[3] <= u < t[-4]) into the cubic
    spline with knots ``t`` and coefficients ``c`` along ``axis``: the same
    function on knots t + {u}.  Returns (new knots, new coefficients)."""
    t = np.asarray(t, dtype=np.float64)
    c = np.moveaxis(np.asarray(c, dtype=np.float64), axis, 0)
    k = int(np.searchsorted(t, u, side="right")) - 1      # t[k] <= u < t[k+1]
    q = np.empty((c.shape[0] + 1,) + c.shape[1:])
    q[:k - 2] = c[:k - 2]
    for i in range(k - 2, k + 1):
        a = (u - t[i]) / (t[i + 3] - t[i])
        q[i] = a * c[i] + (1.0 - a) * c[i - 1]
    q[k + 1:] = c[k:]
    return np.insert(t, k + 1, u), np.moveaxis(q, 0, axis)


def common_knots(knots):
    """The smallest knot vector holding every vector of ``knots`` (all with
    the same end knots), each value with its largest multiplicity."""
    vals = np.unique(np.concatenate(knots))
    mult = [max(int(np.sum(t == v)) for t in knots) for v in vals]
    return np.repeat(vals, mult)


def refine(t, c, target, axis):
    """Coefficients of the spline (t, c) on the finer knot vector ``target``
    along ``axis`` (each knot of ``target`` missing from ``t`` inserted)."""
    for v in np.unique(target):
        for _ in range(int(np.sum(target == v)) - int(np.sum(t == v))):
            t, c = insert_knot(t, c, v, axis)
    return t, c


class SersicSplines:
    """Bicubic splines of the Sersic-mixture amplitude table.

    Parameters
    ----------
    splinedata : str
        ``.npz`` with ``nsersic`` (float32, as stored in the reference's HDF5
        table), ``rh``, ``amplitudes`` [n_n * n_r, n_amp] (nsersic-major) and
        ``radii`` [n_amp] (arcsec).
    spline_smoothing : float or None
        FITPACK's smoothing factor s (None: the reference's default).
    spline_type : str
        ``"rect"`` for RectBivariateSpline on the (nsersic, rh) grid,
        anything else for SmoothBivariateSpline on the scattered table, as
        forcepho/sources.py:634-642 chooses.
    """

    def __init__(self, splinedata=DEFAULT_SPLINEDATA, spline_smoothing=None,
                 spline_type="rect"):
        with np.load(splinedata) as f:
            n = f["nsersic"]
            r = f["rh"]
            A = f["amplitudes"]
            self.radii = np.asarray(f["radii"], dtype=np.float64)
        # np.unique of the float32 column, promoted to float64 the way
        # RectBivariateSpline promotes its inputs
        self.ngrid = np.unique(n).astype(np.float64)
        self.rgrid = np.unique(r).astype(np.float64)
        self.n_amp = A.shape[1]
        self.spline_type, self.spline_smoothing = spline_type, spline_smoothing
        if spline_type == "rect" and not spline_smoothing:
            self._interpolate(A)
        else:
            self._fitpack(n, r, A, spline_type, spline_smoothing)
        self.sersic_range = (float(np.min(n)), float(np.max(n)))
        self.rh_range = (float(np.min(r)), float(np.max(r)))

    def _interpolate(self, A):
        nn, nr = len(self.ngrid), len(self.rgrid)
        self.table = np.asarray(A, dtype=np.float64).reshape(nn, nr, self.n_amp)
        self.tx = notaknot_knots(self.ngrid)
        self.ty = notaknot_knots(self.rgrid)
        Bn = collocation_matrix(self.ngrid, self.tx)
        Br = collocation_matrix(self.rgrid, self.ty)
        coeffs = np.empty((self.n_amp, nn, nr))
        for m in range(self.n_amp):
            tmp = np.linalg.solve(Bn, self.table[:, :, m])
            coeffs[m] = np.linalg.solve(Br, tmp.T).T
        self.coeffs = np.ascontiguousarray(coeffs)

    def _fitpack(self, n, r, A, spline_type, s):
        from scipy.interpolate import RectBivariateSpline, SmoothBivariateSpline
        nn, nr = len(self.ngrid), len(self.rgrid)
        fits = []
        for m in range(self.n_amp):
            if spline_type == "rect":
                f = RectBivariateSpline(self.ngrid, self.rgrid,
                                        np.asarray(A[:, m], dtype=np.float64).reshape(nn, nr),
                                        s=s)
            else:
                f = SmoothBivariateSpline(np.asarray(n, dtype=np.float64),
                                          np.asarray(r, dtype=np.float64),
                                          np.asarray(A[:, m], dtype=np.float64), s=s)
            tx, ty = f.get_knots()
            fits.append((tx, ty, f.get_coeffs().reshape(len(tx) - 4, len(ty) - 4)))
        self.tx = common_knots([f[0] for f in fits])
        self.ty = common_knots([f[1] for f in fits])
        coeffs = np.empty((self.n_amp, len(self.tx) - 4, len(self.ty) - 4))
        for m, (tx, ty, c) in enumerate(fits):
            _, c = refine(tx, c, self.tx, 0)
            _, coeffs[m] = refine(ty, c, self.ty, 1)
        self.coeffs = np.ascontiguousarray(coeffs)

    def _eval(self, nsersic, rh):
        x = min(max(float(nsersic), self.tx[3]), self.tx[-4])
        y = min(max(float(rh), self.ty[3]), self.ty[-4])
        lx, ly = knot_span(self.tx, x), knot_span(self.ty, y)
        Nx, dNx = basis(self.tx, lx, x)
        Ny, dNy = basis(self.ty, ly, y)
        c = self.coeffs[:, lx - 3:lx + 1, ly - 3:ly + 1]
        val = np.einsum("mab,a,b->m", c, Nx, Ny)
        dn = np.einsum("mab,a,b->m", c, dNx, Ny)
        dr = np.einsum("mab,a,b->m", c, Nx, dNy)
        return val, dn, dr

    def evaluate(self, nsersic, rh):
        """(amplitudes, d/dnsersic, d/drh), each [n_amp], with the edge
        branches of forcepho/sources.py:656-694."""
        zeros = np.zeros(self.n_amp)
        if nsersic < self.sersic_range[0]:
            amp = zeros.copy()
            amp[np.argmin(np.abs(rh - self.radii))] = 1.0
            return amp, zeros.copy(), zeros.copy()
        if rh < 0:
            amp = zeros.copy()
            amp[0] = 1.0
            return amp, zeros.copy(), zeros.copy()
        return self._eval(nsersic, rh)

    def amplitudes(self, nsersic, rh):
        return self.evaluate(nsersic, rh)[0]

    def damplitude_dsersic(self, nsersic, rh):
        return self.evaluate(nsersic, rh)[1]

    def damplitude_drh(self, nsersic, rh):
        return self.evaluate(nsersic, rh)[2]


_CACHE = {}


def get_splines(splinedata=DEFAULT_SPLINEDATA, spline_smoothing=None, spline_type="rect"):
    """Shared SersicSplines instance per (table file, smoothing, type)."""
    key = (os.path.abspath(splinedata), spline_smoothing, spline_type)
    if key not in _CACHE:
        _CACHE[key] = SersicSplines(splinedata, spline_smoothing=spline_smoothing,
                                    spline_type=spline_type)
    return _CACHE[key]
