"""Fitting drivers around the posterior: optimisers, NUTS and the linear
flux solve.

``run_opt`` / ``run_opt_bounded`` (scipy BFGS / L-BFGS-B) and ``run_lmc``
(NUTS) keep the interfaces of forcepho/fitting.py:255-499 and drive any
posterior object with ``lnprob_and_grad`` / ``nll`` and ``transform``, such
as :class:`forcepho_amd.model.FastPosterior` over the HIP proposer.

``optimize_fluxes`` is forcepho's conditional maximum-likelihood flux solve
(forcepho/fitting.py:573-627): given fixed source shapes, the model is linear
in the per-band fluxes, so the ML fluxes solve the weighted normal equations
(X W)(X W)^T f = (X W)(y W)^T per band, with X the per-source unit-flux
models (``design_matrix``, computed by the HIP kernels), y the data after
fixed sources are subtracted, and W = ierr.  ``(X W)(X W)^T`` is the flux
precision matrix.  The normal equations of every band are formed and solved
in FP64 on the GPU in one batched call.
"""

import time

import numpy as np
import torch

__all__ = ["Result", "run_opt", "run_opt_bounded", "run_lmc", "initial_cov",
           "optimize_one_band",
           "optimize_fluxes", "solve_normal_equations"]


class Result:
    """A namespace for the products of a fit (the in-memory part of
    forcepho/fitting.py:24-35; the HDF5 file round trip is not built)."""

    def __init__(self, **kwargs):
        for k, v in kwargs.items():
            setattr(self, k, v)


def _start(model, q):
    q = np.asarray(q, dtype=np.float64)
    return model.transform.inverse_transform(q) if model.transform is not None else q.copy()


def _optimize(model, q, method, jac, callback, bounds, opts):
    from scipy.optimize import minimize
    start = _start(model, q)
    t0 = time.time()
    fun = model.nll if jac else model.nll_nograd
    scires = minimize(fun, start.copy(), jac=True if jac else None, method=method,
                      options=opts, bounds=bounds, callback=callback)
    chain = np.atleast_2d(scires.x)
    if model.transform is not None:
        chain = model.transform.transform(chain)
    result = Result(ndim=len(q), starting_position=np.array(q, dtype=np.float64).copy(),
                    chain=chain, lnp=-0.5 * scires.fun, wall_time=time.time() - t0,
                    fun=scires.fun, jac=scires.jac, message=scires.message)
    return result, scires


def run_opt(model, q, jac=True, callback=None, **extras):
    """BFGS on the negative log posterior in the sampling space (the
    interface of forcepho/fitting.py:415-454): ``model`` provides ``nll``
    ((value, gradient) for ``jac``) or ``nll_nograd`` and ``transform``; ``q``
    is the start in the constrained space.  Returns (Result, scipy result);
    ``Result.chain`` holds the optimum in the constrained space."""
    return _optimize(model, q, "BFGS", jac, callback, None, dict(extras))


def run_opt_bounded(model, q, jac=True, callback=None, **extras):
    """L-BFGS-B (forcepho/fitting.py:457-499): ``bounds`` in ``extras``,
    replaced by (-200, 200) per parameter when the model samples in a
    transformed space."""
    bounds = extras.pop("bounds", None)
    opts = {"ftol": 1e-5, "gtol": 1e-5, "maxcor": 20}
    opts.update(extras)
    if model.transform is not None:
        bounds = [(-200.0, 200.0)] * len(q)
    return _optimize(model, q, "L-BFGS-B", jac, callback, bounds, opts)


class _ModelPosterior:
    """One chain over a host posterior object (``lnprob_and_grad`` of the
    sampling-space vector, as forcepho's Posterior), in the form the device
    samplers use: lnp [1] and gradient [ndim] tensors."""

    def __init__(self, model, ndim, device):
        self.model = model
        self.lower = torch.zeros(1, dtype=torch.float64, device=device)
        self.seg_idx = torch.arange(ndim, device=device)[None, :]
        self.seg_mask = torch.ones((1, ndim), dtype=torch.float64, device=device)
        self.param_patch = torch.zeros(ndim, dtype=torch.long, device=device)

    def segment_sum(self, v):
        return v.sum().reshape(1)

    def __call__(self, z):
        lnp, grad = self.model.lnprob_and_grad(z.detach().cpu().numpy())
        dev = self.lower.device
        return (torch.tensor([float(lnp)], dtype=torch.float64, device=dev),
                torch.as_tensor(np.asarray(grad, dtype=np.float64), device=dev))


def initial_cov(n_dim, init_cov=None, trace=None, regular_variance=1e-3):
    """The covariance ``get_pot`` starts the mass matrix from
    (forcepho/fitting.py:370-412): ``init_cov``, else the covariance of the
    samples in ``trace`` ([n_dim, n_samples]), else the identity; its
    diagonal is raised by ``regular_variance`` (up to 5 times) until it has
    a Cholesky factor."""
    if init_cov is not None:
        cov = np.array(init_cov, dtype=np.float64)
    elif trace is not None:
        cov = np.atleast_2d(np.cov(np.asarray(trace, dtype=np.float64), rowvar=True))
    else:
        cov = np.eye(n_dim)
    for _ in range(5):
        try:
            np.linalg.cholesky(cov)
            break
        except np.linalg.LinAlgError:
            cov[np.diag_indices_from(cov)] += regular_variance
    return cov


def run_lmc(model, q, n_draws, adapt=False, full=False, weight=10, warmup=[10], trace=None,
            z_cov=None, max_treedepth=10, discard_tuned_samples=True, progressbar=False,
            random_seed=0xDEADBEEF):
    """NUTS draws from ``model`` (the interface of forcepho/fitting.py:255-342,
    which drives littlemcmc): multinomial NUTS of :mod:`forcepho_amd.nuts` on
    ``model.lnprob_and_grad``, ``warmup[0]`` warm-up iterations adapting the
    step size and a diagonal (or with ``full`` a dense) inverse mass matrix,
    started from ``z_cov``, else from the covariance of ``trace`` ([ndim,
    n_samples]), else the identity, as ``get_pot`` chooses it
    (forcepho/fitting.py:370-412).  Warm-up states are never returned.
    ``adapt``, ``weight``, ``discard_tuned_samples`` and ``progressbar``
    belong to littlemcmc and are accepted for compatibility.  Returns
    (Result with ``chain`` [n_draws, ndim] in the constrained space, the
    sampler, per-draw statistics)."""
    from .nuts import NUTS
    start = _start(model, q)
    ndim = len(start)
    dev = torch.device("cuda") if torch.cuda.is_available() else torch.device("cpu")
    post = _ModelPosterior(model, ndim, dev)
    inv_mass = None
    cov = initial_cov(ndim, z_cov, trace)
    if z_cov is not None or trace is not None:
        inv_mass = cov[None] if full else np.diag(cov).copy()
    t0 = time.time()
    sampler = NUTS(post, start, step_size=0.1, max_treedepth=max_treedepth,
                   seed=int(random_seed) % (2 ** 31), inv_mass=inv_mass, dense=full)
    n_tune = int(warmup[0]) if len(warmup) else 0
    if n_tune > 0:
        sampler.adapt(n_tune)
    draws, stats = [], []
    for _ in range(int(n_draws)):
        sampler.trajectory()
        draws.append(sampler.z.cpu().numpy().copy())
        stats.append({"tree_depth": int(sampler.tree_depth[0]),
                      "n_steps": int(sampler.n_leapfrog[0]),
                      "diverging": bool(sampler.divergent[0]),
                      "mean_tree_accept": float(sampler.accept_prob[0]),
                      "step_size": float(sampler.eps_patch[0])})
    z = np.array(draws)
    chain = model.transform.transform(z) if model.transform is not None else z
    result = Result(ndim=len(q), starting_position=np.array(q, dtype=np.float64).copy(),
                    chain=chain, n_call=getattr(model, "ncall", None),
                    wall_time=time.time() - t0)
    return result, sampler, stats


def _device():
    if not torch.cuda.is_available():
        raise RuntimeError("optimize_fluxes needs the GPU (the HIP path has no CPU fallback)")
    return torch.device("cuda")


def solve_normal_equations(Xes, ws, ys, fixedXs=None):
    """Batched weighted least squares over bands on the device.

    Xes : list of [n_src, npix_b] unit-flux models; ws : list of [npix_b]
    weights (1/sigma); ys : list of [npix_b] data; fixedXs : optional list of
    [npix_b] (or [n_fixed, npix_b]) models of fixed sources to subtract.
    Returns (fluxes [n_bands, n_src], precisions [n_bands, n_src, n_src]) as
    FP64 tensors on the device.
    """
    dev = _device()
    ata, xty = [], []
    for b, (X, w, y) in enumerate(zip(Xes, ws, ys)):
        X = torch.as_tensor(np.asarray(X), dtype=torch.float64, device=dev)
        w = torch.as_tensor(np.asarray(w), dtype=torch.float64, device=dev)
        y = torch.as_tensor(np.asarray(y), dtype=torch.float64, device=dev)
        if fixedXs is not None:
            fx = torch.as_tensor(np.asarray(fixedXs[b]), dtype=torch.float64, device=dev)
            y = y - (fx.sum(dim=0) if fx.dim() == 2 else fx)
        Xp = X * w
        yp = y * w
        ata.append(Xp @ Xp.T)
        xty.append(Xp @ yp)
    A = torch.stack(ata)
    B = torch.stack(xty).unsqueeze(-1)
    flux = torch.linalg.solve(A, B).squeeze(-1)
    return flux, A


def optimize_one_band(X, w, y, fixedX=0):
    """Fluxes [n_src, 1] and precision matrix for one band
    (forcepho/fitting.py:573-595)."""
    fixed = None if np.isscalar(fixedX) and fixedX == 0 else [fixedX]
    flux, prec = solve_normal_equations([X], [w], [y], fixed)
    return flux[0].cpu().numpy()[:, None], prec[0].cpu().numpy()


def optimize_fluxes(patcher, active, return_all=False):
    """Maximum-likelihood fluxes of the ``active`` catalog, conditional on its
    shapes, for every band of ``patcher`` (forcepho/fitting.py:598-627).

    Returns (fluxes, precisions): lists over ``patcher.bandlist`` of
    [n_src] and [n_src, n_src] arrays; with ``return_all`` also the design
    matrices, data and weights.  A singular band raises
    ``numpy.linalg.LinAlgError``.
    """
    Xes, ys = patcher.design_matrix(active)
    ws = patcher.split_band("ierr")
    try:
        flux, prec = solve_normal_equations(Xes, ws, ys)
    except RuntimeError as err:   # torch reports a singular matrix this way
        raise np.linalg.LinAlgError(str(err)) from err
    fluxes = [np.atleast_1d(f) for f in flux.cpu().numpy()]
    precisions = list(prec.cpu().numpy())
    if return_all:
        return fluxes, precisions, Xes, ys, ws
    return fluxes, precisions
