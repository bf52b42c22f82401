"""Posteriors over a patch: ln-probability and its gradient in a sampling
space, for optimisers and samplers.

The public surface is forcepho's (forcepho/model.py): ``Posterior`` with
``lnprob``, ``lnprob_grad``, ``lnprob_and_grad``, ``nll``, ``nll_nograd``,
``residuals`` and ``check_grad`` (lines 42-257); ``FastPosterior`` over a
proposer and a patch (lines 260-405); the transforms ``Transform`` and
``BoundedTransform`` (lines 572-636); ``stack_grad`` (lines 370-387).  The
reference's Theano ops are replaced by :mod:`forcepho_amd.autograd`.

How it is organised here.  A posterior is a pipeline of three pieces:

* a *likelihood* in scene space, ``lnlike(q) -> (lnlike, d lnlike/dq)``,
  which each subclass supplies (``FastPosterior``: the proposer's chi2 and
  dchi2 of the packed patch);
* an optional prior ``lnprior(q) -> (lnprior, d lnprior/dq)``;
* a transform from the sampling space z to the scene space q; with
  ``sampling`` the log-determinant of its Jacobian is part of lnp.

The base class composes them once per new position and keeps the result of
the last position in an :class:`Evaluation` record, so that an optimiser
asking for the value and then the gradient at the same z pays for one
likelihood call.
"""

from dataclasses import dataclass

import numpy as np
from scipy.special import expit

from .limits import NPARAMS

__all__ = ["Posterior", "FastPosterior", "GPUPosterior", "Transform", "BoundedTransform",
           "Evaluation", "stack_grad"]


def stack_grad(chi2_derivs, n_sources, n_bands):
    """Per-band source gradients [n_bands, n_sources, 7] (d/d(flux, ra, dec,
    q, pa, nsersic, rh)) to the scene's parameter order: per source
    [flux_1..flux_NB, ra, dec, q, pa, nsersic, rh] (forcepho/model.py:
    370-387).  A source's flux in band b only enters band b; its shape
    enters every band, so those derivatives add up over bands."""
    d = np.asarray(chi2_derivs, dtype=np.float64).reshape(n_bands, n_sources, NPARAMS)
    out = np.empty((n_sources, n_bands + NPARAMS - 1), dtype=np.float64)
    out[:, :n_bands] = d[:, :, 0].T
    out[:, n_bands:] = d[:, :, 1:].sum(axis=0)
    return out.reshape(-1)


@dataclass
class Evaluation:
    """Everything computed at one sampling-space position."""
    z: np.ndarray
    q: np.ndarray
    lnlike: float
    lnlike_grad: np.ndarray
    lnprior: float
    lnprior_grad: object
    lnp: float
    lnp_grad: np.ndarray


class Posterior:
    """lnp(z) = lnlike(q(z)) + lnprior(q(z)) [+ ln|dq/dz| when sampling].

    Give either ``transform`` or both ``lower`` and ``upper`` (which make a
    :class:`BoundedTransform`); anything else raises ``ValueError``, as do
    keywords that are not understood."""

    sampling = True

    def __init__(self, **kwargs):
        transform = kwargs.pop("transform", None)
        lower = kwargs.pop("lower", None)
        upper = kwargs.pop("upper", None)
        if kwargs:
            raise ValueError(f"Posterior does not understand the argument(s) {sorted(kwargs)}")
        bounded = lower is not None and upper is not None
        if bounded == (transform is not None):
            raise ValueError('Must specify either "transform" or ("lower","upper")')
        self.transform = transform if transform is not None else BoundedTransform(lower, upper)
        self.ncall = 0
        self.ndof = 1.0
        self._lnpriorfn = None
        self._last = None

    # -- to be supplied by subclasses --------------------------------------
    def lnlike(self, q):
        """(lnlike, d lnlike / dq) in scene space."""
        raise NotImplementedError

    def residuals(self, z):
        raise NotImplementedError

    # -- composition --------------------------------------------------------
    def lnprior(self, q):
        """(lnprior, d lnprior / dq); zero unless a prior function is set."""
        if self._lnpriorfn is None:
            return 0.0, 0.0
        return self._lnpriorfn(q)

    def make_transform(self, z):
        """Scene-space parameters of the sampling-space position z."""
        if self.transform is None:
            return np.array(z, dtype=np.float64)
        return self.transform.transform(z)

    def evaluate(self, z):
        """Compose likelihood, prior and transform at z and remember it."""
        z = np.array(z, dtype=np.float64)
        q = self.make_transform(z)
        ll, ll_grad = self.lnlike(q)
        lpr, lpr_grad = self.lnprior(q)
        if self.transform is None:
            jac, ldj, ldj_grad = 1.0, 0.0, 0.0
        else:
            jac = self.transform.jacobian(z)
            ldj = self.transform.lndetjac(z) if self.sampling else 0.0
            ldj_grad = self.transform.lndetjac_grad(z) if self.sampling else 0.0
        lnp = ll + lpr + ldj
        lnp_grad = (ll_grad + lpr_grad) * jac + ldj_grad
        self.ncall += 1
        self._last = Evaluation(z, q, ll, ll_grad, lpr, lpr_grad, lnp, lnp_grad)
        return self._last

    def at(self, z):
        """The evaluation at z, reusing the last one when z is unchanged."""
        last = self._last
        z = np.asarray(z, dtype=np.float64)
        if last is None or last.z.shape != z.shape or not np.array_equal(last.z, z):
            last = self.evaluate(z)
        return last

    # forcepho's cached quantities of the last evaluation
    @property
    def _z(self):
        return None if self._last is None else self._last.z

    @property
    def _q(self):
        return None if self._last is None else self._last.q

    @property
    def _lnp(self):
        return None if self._last is None else self._last.lnp

    @property
    def _lnp_grad(self):
        return None if self._last is None else self._last.lnp_grad

    # -- the reference's accessors -------------------------------------------
    def lnprob(self, z):
        return self.at(z).lnp

    def lnprob_grad(self, z):
        return self.at(z).lnp_grad

    def lnprob_and_grad(self, z):
        e = self.at(z)
        return e.lnp, e.lnp_grad

    def nll(self, z):
        e = self.at(z)
        return -e.lnp / self.ndof, -e.lnp_grad / self.ndof

    def nll_nograd(self, z):
        return -self.at(z).lnp / self.ndof

    def check_grad(self, z, delta=1e-5):
        """(analytic gradient, central-difference gradient) of lnp at z with
        step ``delta`` (a scalar or one step per parameter)."""
        z = np.array(z, dtype=np.float64)
        analytic = np.array(self.lnprob_grad(z), dtype=np.float64)
        steps = np.broadcast_to(np.asarray(delta, dtype=np.float64), z.shape)
        numeric = np.empty_like(z)
        for i, h in enumerate(steps):
            e = np.zeros_like(z)
            e[i] = h
            numeric[i] = (self.lnprob(z + e) - self.lnprob(z - e)) / (2.0 * h)
        return analytic, numeric


class FastPosterior(Posterior):
    """lnlike from a proposer evaluating the packed patch: lnlike = -1/2
    sum_b chi2_b, gradient = -1/2 ``stack_grad(dchi2)`` (forcepho/model.py:
    302-387).  ``sampling`` adds the transform's log-Jacobian; ``logging``
    keeps the history of positions, lnlike and gradients; ``verbose`` prints
    every ``print_interval`` calls."""

    def __init__(self, proposer, patch, scene=None, lnprior=None, transform=None,
                 name="", print_interval=1000, verbose=False, debug=False,
                 logging=False, sampling=True, **kwargs):
        super().__init__(transform=transform, **kwargs)
        self.proposer = proposer
        self.patch = patch
        self.scene = scene if scene else patch.scene
        self._lnpriorfn = lnprior
        self.sampling = sampling
        self.name = name
        self.print_interval = print_interval
        self.verbose = verbose
        self.debug = debug
        self.logging = logging
        self._residuals = None
        self.pos_history, self.lnp_history, self.grad_history = [], [], []

    def _propose(self, q):
        self.scene.set_all_source_params(q)
        return self.scene.get_proposal()

    def lnlike(self, q):
        out = self.proposer.evaluate_proposal(self._propose(q), patch=self.patch)
        chi2, derivs = out[0], out[1]
        if len(out) > 2:
            self._residuals = out[2]
        ll = -0.5 * float(np.sum(chi2, dtype=np.float64))
        return ll, -0.5 * self.stack_grad(derivs)

    def evaluate(self, z):
        e = super().evaluate(z)
        if self.logging:
            self.pos_history.append(e.z)
            self.lnp_history.append(e.lnlike)
            self.grad_history.append(e.lnlike_grad)
        if self.verbose and self.ncall % self.print_interval == 0:
            print(f"-------\n {self.name} @ {self.ncall}\nq: {e.q}\nlnlike: {e.lnlike}")
        return e

    def stack_grad(self, chi2_derivs):
        return stack_grad(chi2_derivs, int(self.patch.n_sources), int(len(self.patch.band_N)))

    def residuals(self, q, unpack=True):
        """Residual images (data - model) of the scene parameters q."""
        keep = self.patch.return_residual
        self.patch.return_residual = True
        try:
            out = self.proposer.evaluate_proposal(self._propose(q), patch=self.patch,
                                                  unpack=unpack)
        finally:
            self.patch.return_residual = keep
        self._residuals = out[2]
        return self._residuals


GPUPosterior = FastPosterior


def sigmoid(x):
    return expit(x)


def logit(y):
    y = np.asarray(y, dtype=np.float64)
    return np.log(y) - np.log1p(-y)


def sigmoid_grad(z):
    y = expit(z)
    return y * (1.0 - y)


class Transform:
    """The identity between sampling and scene space (forcepho/model.py:
    572-593)."""

    def __init__(self, ndim):
        self.ndim = ndim

    def transform(self, z):
        return np.array(z, dtype=np.float64)

    def inverse_transform(self, Theta):
        return np.array(Theta, dtype=np.float64)

    def jacobian(self, z):
        """Diagonal of dq/dz."""
        return np.ones(len(z))

    def lndetjac(self, z):
        return float(np.sum(np.log(np.abs(self.jacobian(z)))))

    def lndetjac_grad(self, z):
        return np.zeros(len(z))


class BoundedTransform(Transform):
    """q = lower + (upper - lower) / (1 + exp(-z)), each parameter in its box
    (forcepho/model.py:596-636).  The log-determinant is summed in logs, so a
    long parameter vector does not underflow the Jacobian's product."""

    def __init__(self, lower, upper):
        self.lower = np.atleast_1d(np.asarray(lower, dtype=np.float64))
        self.upper = np.atleast_1d(np.asarray(upper, dtype=np.float64))
        if self.lower.shape != self.upper.shape:
            raise ValueError("lower and upper bounds differ in length")
        self.ndim = len(self.upper)

    @property
    def range(self):
        return self.upper - self.lower

    def transform(self, z):
        return self.lower + self.range * expit(z)

    def inverse_transform(self, Theta):
        return logit((np.asarray(Theta, dtype=np.float64) - self.lower) / self.range)

    def jacobian(self, z):
        return self.range * sigmoid_grad(z)

    def lndetjac(self, z):
        # log(y (1 - y)) = -|z| - 2 log(1 + exp(-|z|)), finite for any z
        a = np.abs(np.asarray(z, dtype=np.float64))
        return float(np.sum(np.log(np.abs(self.range))) - np.sum(a + 2.0 * np.log1p(np.exp(-a))))

    def lndetjac_grad(self, z):
        return 1.0 - 2.0 * expit(z)
