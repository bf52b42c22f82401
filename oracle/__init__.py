"""Oracle for forcepho_amd -- TEST INFRASTRUCTURE ONLY.

Only tests/, ``__graft_entry__.smoke()`` and the cpu_baseline leg of bench.py
use this package, and only as the checker; the product never imports it.

Contents
--------
* ``evaluate`` -- FP64 C restatement (fpho_oracle.c) of the reference
  likelihood and gradient: chi2 = sum (r^2 - d^2) ierr^2 per band and
  dchi2/d(flux, ra, dec, q, pa, nsersic, rh) per band and source.
* ``render`` -- the same restatement's FP64 model only (builds test data).
* ``evaluate_np`` -- the same quantities in vectorised numpy FP64, written
  from forcepho/slow/gaussmodel.py:63-147,273-476 and
  forcepho/slow/likelihood.py:219-252 (used for small cases and to check the
  C restatement).
* ``reference_evaluate`` -- the reference's own CPU kernel
  (forcepho/src/compute_gaussians_kernel.cc) compiled from its sources by
  ``oracle/Makefile`` into ``oracle/_ref`` (FP32 arithmetic, FP64 chi2,
  sequential FP32 gradient sums), driven with the reference's struct
  layouts (forcepho/patches/patch.py:44-72, forcepho/proposal.py:229-244).

Pinning (tests/test_oracle.py): closed forms (isotropic Gaussians, flux
linearity, model normalisation), central finite differences in FP64 of every
gradient component, numpy vs C restatements, and agreement with the
compiled reference kernel at FP32 tolerance.  tests/test_verification_image.py
pins it against the reference's verification image (its pure-python
renderer's model and seven derivative images): the model to 2e-8, the
gradient to 2.4e-5 relative.
"""

import ctypes
import glob
import importlib.util
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REF_DIR = os.path.join(HERE, "_ref")

SOURCE_DTYPE = np.dtype([('ra', 'f4'), ('dec', 'f4'), ('q', 'f4'), ('pa', 'f4'),
                         ('nsersic', 'f4'), ('rh', 'f4'), ('fluxes', 'f4', 30),
                         ('mixture_amplitudes', 'f4', 10), ('damplitude_drh', 'f4', 10),
                         ('damplitude_dnsersic', 'f4', 10)], align=True)
PSF_DTYPE = np.dtype([("amp", "f4"), ("xcen", "f4"), ("ycen", "f4"), ("Cxx", "f4"),
                      ("Cyy", "f4"), ("Cxy", "f4"), ("sersic_bin", "i4")])

_fp = ctypes.POINTER(ctypes.c_float)
_dp = ctypes.POINTER(ctypes.c_double)
_i16 = ctypes.POINTER(ctypes.c_int16)
_i32 = ctypes.POINTER(ctypes.c_int32)


class _OraclePatch(ctypes.Structure):
    _fields_ = [("npix", ctypes.c_int64), ("data", _fp), ("ierr", _fp), ("xpix", _fp),
                ("ypix", _fp), ("n_bands", ctypes.c_int32), ("n_exp", ctypes.c_int32),
                ("band_start", _i16), ("band_N", _i16), ("exposure_start", _i32),
                ("exposure_N", _i32), ("n_sources", ctypes.c_int32),
                ("n_radii", ctypes.c_int32), ("rad2", _fp), ("D", _fp), ("crpix", _fp),
                ("crval", _fp), ("CW", _fp), ("G", _fp), ("n_psf_per_source", _i32),
                ("psfgauss", ctypes.c_void_p), ("n_psfgauss", ctypes.c_int64),
                ("psfgauss_start", _i32)]


_LIBH = None


def load():
    global _LIBH
    if _LIBH is None:
        if not os.path.exists(LIB):
            raise ImportError(f"{LIB} missing: run `make -C oracle`")
        h = ctypes.CDLL(LIB)
        h.oracle_evaluate.argtypes = [ctypes.POINTER(_OraclePatch), _fp, ctypes.c_int, _dp,
                                      _dp, _dp, _dp]
        h.oracle_evaluate.restype = ctypes.c_int
        h.oracle_evaluate64.argtypes = [ctypes.POINTER(_OraclePatch), _dp, ctypes.c_int, _dp,
                                        _dp, _dp, _dp]
        h.oracle_evaluate64.restype = ctypes.c_int
        h.oracle_image_gaussians.argtypes = [ctypes.POINTER(_OraclePatch), _fp, ctypes.c_int,
                                             ctypes.c_int, _dp]
        h.oracle_image_gaussians.restype = ctypes.c_int
        _LIBH = h
    return _LIBH


def _arrays(patch):
    """Contiguous copies of the patch arrays with the device dtypes."""
    a = {}
    for name in ("data", "ierr", "xpix", "ypix", "psfgauss"):
        if getattr(patch, name, None) is None:
            raise ValueError(f"patch.{name} is not on the host")
    for name, dt in [("data", np.float32), ("ierr", np.float32), ("xpix", np.float32),
                     ("ypix", np.float32), ("band_start", np.int16), ("band_N", np.int16),
                     ("exposure_start", np.int32), ("exposure_N", np.int32),
                     ("rad2", np.float32), ("D", np.float32), ("crpix", np.float32),
                     ("crval", np.float32), ("CW", np.float32), ("G", np.float32),
                     ("n_psf_per_source", np.int32), ("psfgauss_start", np.int32)]:
        a[name] = np.ascontiguousarray(getattr(patch, name), dtype=dt).reshape(-1)
    psf = np.ascontiguousarray(patch.psfgauss)
    a["psfgauss"] = psf.view(PSF_DTYPE) if psf.dtype.itemsize == 28 else psf.astype(PSF_DTYPE)
    return a


def _desc(patch, a):
    d = _OraclePatch()
    d.npix = len(a["data"])
    for k in ("data", "ierr", "xpix", "ypix", "rad2", "D", "crpix", "crval", "CW", "G"):
        setattr(d, k, a[k].ctypes.data_as(_fp))
    for k in ("band_start", "band_N"):
        setattr(d, k, a[k].ctypes.data_as(_i16))
    for k in ("exposure_start", "exposure_N", "n_psf_per_source", "psfgauss_start"):
        setattr(d, k, a[k].ctypes.data_as(_i32))
    d.n_bands = len(a["band_N"])
    d.n_exp = len(a["exposure_N"])
    d.n_sources = int(patch.n_sources)
    d.n_radii = int(patch.n_radii)
    d.psfgauss = ctypes.c_void_p(a["psfgauss"].ctypes.data)
    d.n_psfgauss = len(a["psfgauss"])
    return d


def _proposal(proposal):
    p = np.ascontiguousarray(proposal)
    assert p.dtype.itemsize == 264
    return p


def evaluate(patch, proposal, cutoff=True, residual=False, model=False):
    """FP64 chi2 [n_bands], dchi2 [n_bands, n_sources, 7] (and residual /
    model [npix] when asked) from the C restatement."""
    a = _arrays(patch)
    d = _desc(patch, a)
    p = _proposal(proposal)
    nb, ns = len(a["band_N"]), int(patch.n_sources)
    chi2 = np.zeros(nb)
    dchi2 = np.zeros((nb, ns, 7))
    res = np.zeros(len(a["data"])) if residual else None
    mod = np.zeros(len(a["data"])) if model else None
    rc = load().oracle_evaluate(ctypes.byref(d), p.ctypes.data_as(_fp), int(bool(cutoff)),
                                chi2.ctypes.data_as(_dp), dchi2.ctypes.data_as(_dp),
                                res.ctypes.data_as(_dp) if res is not None else None,
                                mod.ctypes.data_as(_dp) if mod is not None else None)
    if rc != 0:
        raise RuntimeError("oracle_evaluate failed")
    out = [chi2, dchi2]
    if residual:
        out.append(res)
    if model:
        out.append(mod)
    return tuple(out)


def render(patch, proposal, cutoff=True):
    """FP64 model [npix] of the proposal over the packed pixels (no
    gradient; the data array may hold anything)."""
    a = _arrays(patch)
    d = _desc(patch, a)
    p = _proposal(proposal)
    chi2 = np.zeros(len(a["band_N"]))
    mod = np.zeros(len(a["data"]))
    rc = load().oracle_evaluate(ctypes.byref(d), p.ctypes.data_as(_fp), int(bool(cutoff)),
                                chi2.ctypes.data_as(_dp), None, None, mod.ctypes.data_as(_dp))
    if rc != 0:
        raise RuntimeError("oracle_evaluate failed")
    return mod


def image_gaussians(patch, proposal, exposure):
    """[n_sources*n_psf, 48]: amp, x, y, fxx, fyy, fxy, Jacobian 7x6 (FP64)."""
    a = _arrays(patch)
    d = _desc(patch, a)
    b = int(np.searchsorted(np.cumsum(a["band_N"]), exposure, side="right"))
    ng = int(patch.n_sources) * int(a["n_psf_per_source"][b])
    out = np.zeros((ng, 48))
    load().oracle_image_gaussians(ctypes.byref(d), _proposal(proposal).ctypes.data_as(_fp),
                                  int(exposure), b, out.ctypes.data_as(_dp))
    return out


# ----------------------------------------------------------- numpy restatement

def _rot(t):
    c, s = np.cos(t), np.sin(t)
    return np.array([[c, -s], [s, c]])


def gaussians_np(patch, proposal, exposure):
    """Image Gaussians of one exposure in FP64, following
    forcepho/slow/gaussmodel.py:convert_to_gaussians/_get_gaussian_gradients.
    Returns (params [ng, 6] as amp, x, y, fxx, fyy, fxy; jac [ng, 7, 6])."""
    a = _arrays(patch)
    ns = int(patch.n_sources)
    b = int(np.searchsorted(np.cumsum(a["band_N"]), exposure, side="right"))
    npsf = int(a["n_psf_per_source"][b])
    psf = a["psfgauss"][a["psfgauss_start"][exposure]:a["psfgauss_start"][exposure] + npsf]
    D = a["D"].reshape(-1, ns, 2, 2)[exposure].astype(np.float64)
    CW = a["CW"].reshape(-1, ns, 2, 2)[exposure].astype(np.float64)
    crpix = a["crpix"].reshape(-1, ns, 2)[exposure].astype(np.float64)
    crval = a["crval"].reshape(-1, ns, 2)[exposure].astype(np.float64)
    G = float(a["G"][exposure])
    params, jacs = [], []
    for s in range(ns):
        src = proposal[s]
        q, pa = float(src["q"]), float(src["pa"])
        R = _rot(pa)
        dR = np.array([[-np.sin(pa), -np.cos(pa)], [np.cos(pa), -np.sin(pa)]])
        S = np.diag([1.0 / q, q])
        dS = np.diag([-1.0 / q ** 2, 1.0])
        T = D[s] @ R @ S
        dTq = D[s] @ R @ dS
        dTp = D[s] @ dR @ S
        sky = np.array([src["ra"], src["dec"]], dtype=np.float64)
        smean = CW[s] @ (sky - crval[s]) + crpix[s]
        flux = float(src["fluxes"][b])
        for g in psf:
            m = int(g["sersic_bin"])
            scov = float(a["rad2"][m]) * np.eye(2)
            pcov = np.array([[g["Cxx"], g["Cxy"]], [g["Cxy"], g["Cyy"]]], dtype=np.float64)
            Sigma = T @ scov @ T.T + pcov
            F = np.linalg.inv(Sigma)
            detF = np.linalg.det(F)
            am, al = float(src["mixture_amplitudes"][m]), float(g["amp"])
            norm = G * al * np.sqrt(detF) / (2 * np.pi)
            K = flux * am * norm
            dSq = T @ scov @ dTq.T + dTq @ scov @ T.T
            dSp = T @ scov @ dTp.T + dTp @ scov @ T.T
            dFq = -F @ dSq @ F
            dFp = -F @ dSp @ F
            dA_dq = K / 2 * np.trace(Sigma @ dFq)
            dA_dpa = K / 2 * np.trace(Sigma @ dFp)
            jac = np.zeros((7, 6))
            jac[0, 0] = am * norm
            jac[1, 1:3] = CW[s][0, 0], CW[s][1, 0]
            jac[2, 1:3] = CW[s][0, 1], CW[s][1, 1]
            jac[3] = [dA_dq, 0, 0, dFq[0, 0], dFq[1, 1], dFq[1, 0]]
            jac[4] = [dA_dpa, 0, 0, dFp[0, 0], dFp[1, 1], dFp[1, 0]]
            jac[5, 0] = flux * norm * float(src["damplitude_dnsersic"][m])
            jac[6, 0] = flux * norm * float(src["damplitude_drh"][m])
            params.append([K, smean[0] + g["xcen"], smean[1] + g["ycen"],
                           F[0, 0], F[1, 1], F[1, 0]])
            jacs.append(jac)
    return np.array(params, dtype=np.float64), np.array(jacs)


def pixel_values_np(par, xp, yp, cutoff=True):
    """Counts and d(counts)/d(A, x, y, Fxx, Fyy, Fxy) of one Gaussian over
    pixels (forcepho/slow/gaussmodel.py:351-476, second order, no det)."""
    amp, xc, yc, fxx, fyy, fxy = par
    dx, dy = xp - xc, yp - yc
    vx = fxx * dx + fxy * dy
    vy = fyy * dy + fxy * dx
    ea = dx * vx + dy * vy
    Gp = np.exp(-0.5 * ea)
    H = 1 + (vx * vx + vy * vy - fxx - fyy) / 24.
    C = amp * Gp * H
    ch = amp * Gp
    grad = np.array([Gp * H,
                     C * vx - ch * (fxx * vx + fxy * vy) / 12.,
                     C * vy - ch * (fyy * vy + fxy * vx) / 12.,
                     -0.5 * C * dx * dx - ch * (1. - 2. * dx * vx) / 24.,
                     -0.5 * C * dy * dy - ch * (1. - 2. * dy * vy) / 24.,
                     -1.0 * C * dx * dy + ch * (dy * vx + dx * vy) / 12.])
    if cutoff:
        keep = ea <= 36.0
        C = np.where(keep, C, 0.0)
        grad = np.where(keep[None, :], grad, 0.0)
    return C, grad


def evaluate_np(patch, proposal, cutoff=True):
    """(chi2 [n_bands], dchi2 [n_bands, n_sources, 7], residual [npix]) in
    vectorised FP64 numpy; suitable for small patches.  ``proposal`` is a
    Source-struct array or a list of dicts with the same field names (FP64
    values)."""
    a = _arrays(patch)
    ns = int(patch.n_sources)
    nb = len(a["band_N"])
    chi2 = np.zeros(nb)
    dchi2 = np.zeros((nb, ns, 7))
    residual = np.zeros(len(a["data"]))
    for b in range(nb):
        npsf = int(a["n_psf_per_source"][b])
        for e in range(a["band_start"][b], a["band_start"][b] + a["band_N"][b]):
            sl = slice(a["exposure_start"][e], a["exposure_start"][e] + a["exposure_N"][e])
            xp = a["xpix"][sl].astype(np.float64)
            yp = a["ypix"][sl].astype(np.float64)
            d = a["data"][sl].astype(np.float64)
            ie = a["ierr"][sl].astype(np.float64)
            params, jacs = gaussians_np(patch, proposal, e)
            model = np.zeros_like(d)
            for par in params:
                model += pixel_values_np(par, xp, yp, cutoff)[0]
            r = d - model
            residual[sl] = r
            chi2[b] += np.sum((r * ie - d * ie) * (r * ie + d * ie))
            w = -2.0 * r * ie * ie
            for k, (par, jac) in enumerate(zip(params, jacs)):
                _, grad = pixel_values_np(par, xp, yp, cutoff)
                dchi2[b, k // npsf] += jac @ (grad @ w)
    return chi2, dchi2, residual


# ------------------------------------------------------ compiled reference

def reference_module():
    """The compiled reference CPU kernel module, or None when not built."""
    libs = glob.glob(os.path.join(REF_DIR, "compute_gaussians_kernel*.so"))
    if not libs:
        return None
    spec = importlib.util.spec_from_file_location("compute_gaussians_kernel", libs[0])
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


REF_PATCH_DTYPE = np.dtype([('data', np.uintp), ('ierr', np.uintp), ('xpix', np.uintp),
                            ('ypix', np.uintp), ('residual', np.uintp),
                            ('exposure_start', np.uintp), ('exposure_N', np.uintp),
                            ('band_start', np.uintp), ('band_N', np.uintp),
                            ('n_sources', np.int32), ('n_radii', np.int32),
                            ('rad2', np.uintp), ('D', np.uintp), ('crpix', np.uintp),
                            ('crval', np.uintp), ('CW', np.uintp), ('G', np.uintp),
                            ('n_psf_per_source', np.uintp), ('psfgauss', np.uintp),
                            ('psfgauss_start', np.uintp)], align=True)


class ReferenceKernel:
    """Drive the compiled reference kernel on a packed patch, band by band,
    as forcepho's CPUProposer does (forcepho/proposal.py:180-226)."""

    def __init__(self, patch, module=None):
        self.mod = module or reference_module()
        if self.mod is None:
            raise ImportError("reference kernel not built (make -C oracle ref)")
        self.a = _arrays(patch)
        self.residual = np.zeros(len(self.a["data"]), dtype=np.float32)
        self.nb = len(self.a["band_N"])
        self.ns = int(patch.n_sources)
        st = np.zeros(1, dtype=REF_PATCH_DTYPE)[0]
        for name in REF_PATCH_DTYPE.names:
            if name in ("n_sources", "n_radii"):
                st[name] = int(getattr(patch, name))
            elif name == "residual":
                st[name] = self.residual.ctypes.data
            else:
                st[name] = self.a[name].ctypes.data
        self.struct = np.array([st], dtype=REF_PATCH_DTYPE)
        self.ptr = int(self.struct.ctypes.data)

    def evaluate(self, proposal):
        """(chi2 [n_bands], dchi2 [n_bands, n_sources, 7]), bands in turn (the
        pybind11 kernel holds the GIL, so bands cannot run on threads)."""
        p = np.ascontiguousarray(proposal).copy()
        chi2 = np.zeros(self.nb)
        dchi2 = np.zeros((self.nb, self.ns, 7))
        rets = [self.mod.EvaluateProposal(b, self.ptr, p.ctypes.data) for b in range(self.nb)]
        for b, r in enumerate(rets):
            chi2[b] = r.chi2
            dchi2[b] = np.asarray(r.dchi2_dp).reshape(30, 7)[:self.ns]
        return chi2, dchi2


def reference_evaluate(patch, proposal):
    return ReferenceKernel(patch).evaluate(proposal)


# --------------------------------------------- parameter vector, all FP64

_SPLINES = None


def scipy_splines():
    """The reference's amplitude interpolants: one scipy RectBivariateSpline
    per mixture component over the table (forcepho/sources.py:623-645)."""
    global _SPLINES
    if _SPLINES is None:
        from scipy.interpolate import RectBivariateSpline
        table = os.path.join(os.path.dirname(HERE), "forcepho_amd", "data",
                             "sersic_splinedata_large.npz")
        with np.load(table) as f:
            n, r, A, radii = f["nsersic"], f["rh"], f["amplitudes"], f["radii"]
        ng, rg = np.unique(n), np.unique(r)
        sp = [RectBivariateSpline(ng, rg, A[:, i].reshape(len(ng), len(rg)), s=None)
              for i in range(A.shape[1])]
        _SPLINES = (sp, float(n.min()), np.asarray(radii, dtype=np.float64))
    return _SPLINES


def amplitudes_scipy(nsersic, rh):
    """(a, da/dnsersic, da/drh) with the branches of forcepho/sources.py:656-694."""
    sp, nmin, radii = scipy_splines()
    z = np.zeros(len(sp))
    if nsersic < nmin:
        a = z.copy()
        a[np.argmin(np.abs(rh - radii))] = 1.0
        return a, z.copy(), z.copy()
    if rh < 0:
        a = z.copy()
        a[0] = 1.0
        return a, z.copy(), z.copy()
    a = np.array([s(nsersic, rh)[0, 0] for s in sp])
    dn = np.array([s(nsersic, rh, dx=1)[0, 0] for s in sp])
    dr = np.array([s(nsersic, rh, dy=1)[0, 0] for s in sp])
    return a, dn, dr


def sources_from_theta(theta, n_bands):
    """FP64 source dicts from [flux_1..flux_NB, ra, dec, q, pa, n, rh] blocks."""
    theta = np.asarray(theta, dtype=np.float64).reshape(-1, n_bands + 6)
    out = []
    for t in theta:
        a, dn, dr = amplitudes_scipy(t[n_bands + 4], t[n_bands + 5])
        out.append({"fluxes": t[:n_bands], "ra": t[n_bands], "dec": t[n_bands + 1],
                    "q": t[n_bands + 2], "pa": t[n_bands + 3],
                    "mixture_amplitudes": a, "damplitude_dnsersic": dn,
                    "damplitude_drh": dr})
    return out


def stack_grad(dchi2):
    """[n_bands, n_sources, 7] -> per source [flux_b..., ra, dec, q, pa, n, rh]
    (forcepho/model.py:370-387)."""
    nb, ns, _ = dchi2.shape
    g = np.zeros((ns, nb + 6))
    for b in range(nb):
        g[:, b] += dchi2[b, :, 0]
        g[:, nb:] += dchi2[b, :, 1:]
    return g.reshape(-1)


def source_records64(sources, n_bands):
    """[n_sources, 66] FP64 records with the field order of the Source struct
    (ra, dec, q, pa, nsersic, rh, fluxes[30], amplitudes[10],
    damplitude_drh[10], damplitude_dnsersic[10])."""
    rec = np.zeros((len(sources), 66))
    for k, s in enumerate(sources):
        rec[k, 0], rec[k, 1], rec[k, 2], rec[k, 3] = s["ra"], s["dec"], s["q"], s["pa"]
        rec[k, 6:6 + n_bands] = s["fluxes"]
        na = len(s["mixture_amplitudes"])
        rec[k, 36:36 + na] = s["mixture_amplitudes"]
        rec[k, 46:46 + na] = s["damplitude_drh"]
        rec[k, 56:56 + na] = s["damplitude_dnsersic"]
    return rec


def evaluate64(patch, sources, cutoff=True, grad=True):
    """The C restatement fed FP64 source records (``sources_from_theta``
    dicts): chi2 [n_bands], dchi2 [n_bands, n_sources, 7] (None with
    ``grad=False``, which skips the gradient work)."""
    a = _arrays(patch)
    d = _desc(patch, a)
    nb, ns = len(a["band_N"]), int(patch.n_sources)
    rec = np.ascontiguousarray(source_records64(sources, nb))
    chi2 = np.zeros(nb)
    dchi2 = np.zeros((nb, ns, 7)) if grad else None
    rc = load().oracle_evaluate64(ctypes.byref(d), rec.ctypes.data_as(_dp), int(bool(cutoff)),
                                  chi2.ctypes.data_as(_dp),
                                  dchi2.ctypes.data_as(_dp) if grad else None, None, None)
    if rc != 0:
        raise RuntimeError("oracle_evaluate64 failed")
    return chi2, dchi2


def lnlike_theta(patch, theta, cutoff=True, grad=True):
    """lnlike = -chi2/2 and d lnlike / d theta, all FP64 (scipy splines, FP64
    source records, the C restatement); with ``grad=False`` only lnlike."""
    nb = len(np.asarray(patch.band_N))
    chi2, dchi2 = evaluate64(patch, sources_from_theta(theta, nb), cutoff=cutoff, grad=grad)
    if not grad:
        return -0.5 * chi2.sum()
    return -0.5 * chi2.sum(), -0.5 * stack_grad(dchi2)


def lnlike_theta_np(patch, theta, cutoff=True):
    """lnlike = -chi2/2 (summed over bands) and d lnlike / d theta in FP64."""
    nb = len(np.asarray(patch.band_N))
    chi2, dchi2, _ = evaluate_np(patch, sources_from_theta(theta, nb), cutoff=cutoff)
    return -0.5 * chi2.sum(), -0.5 * stack_grad(dchi2)
