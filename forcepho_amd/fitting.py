"""Linear flux optimisation on the device.

``optimize_fluxes`` is forcepho's conditional maximum-likelihood flux solve
(forcepho/fitting.py:573-627): given fixed source shapes, the model is linear
in the per-band fluxes, so the ML fluxes solve the weighted normal equations
(X W)(X W)^T f = (X W)(y W)^T per band, with X the per-source unit-flux
models (``design_matrix``, computed by the HIP kernels), y the data after
fixed sources are subtracted, and W = ierr.  ``(X W)(X W)^T`` is the flux
precision matrix.  The normal equations of every band are formed and solved
in FP64 on the GPU in one batched call.
"""

import numpy as np
import torch

__all__ = ["optimize_one_band", "optimize_fluxes", "solve_normal_equations"]


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
