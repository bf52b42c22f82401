"""Posterior objects: ln-probability and gradient of a patch.

Mirrors forcepho/model.py: ``Posterior`` (caching, ``lnprob_and_grad``,
``nll``, central-difference ``check_grad``; lines 42-257), ``FastPosterior``
(transform -> proposal -> device -> lnlike/grad; lines 260-405) and the
parameter transforms ``Transform`` / ``BoundedTransform`` (lines 572-636).
The Theano ops of the reference are replaced by
:mod:`forcepho_amd.autograd` (a ``torch.autograd.Function``).
"""

import numpy as np
from scipy.special import expit

from .limits import NPARAMS

__all__ = ["Posterior", "FastPosterior", "GPUPosterior", "Transform", "BoundedTransform",
           "stack_grad"]


def stack_grad(chi2_derivs, n_sources, n_bands):
    """[n_bands, n_sources, 7] -> per source [flux_1..flux_NB, ra, dec, q, pa,
    nsersic, rh] (forcepho/model.py:370-387)."""
    grads = np.zeros([n_sources, n_bands + (NPARAMS - 1)], dtype=np.float64)
    for band, derivs in enumerate(chi2_derivs):
        grads[:, n_bands:] += derivs[:, 1:]
        grads[:, band] += derivs[:, 0]
    return grads.reshape(-1)


class Posterior:
    """Caching of lnp / grad and a numerical gradient check."""

    def __init__(self, **kwargs):
        self.ncall = 0
        self.ndof = 1.0
        has_bounds = (np.any(kwargs.get('lower')) and np.any(kwargs.get('upper')))
        if has_bounds == bool(kwargs.get('transform')):
            raise ValueError('Must specify either "transform" or ("lower","upper")')
        transform = kwargs.pop('transform', None)
        if transform is not None:
            self.transform = transform
        else:
            self.transform = BoundedTransform(kwargs.pop('lower'), kwargs.pop('upper'))
        if kwargs:
            raise ValueError(f'Passed argument(s) {list(kwargs.keys())} to Posterior '
                             'that were not understood')

    def evaluate(self, z):
        raise NotImplementedError

    def lnprob(self, z):
        if np.any(z != self._z):
            self.evaluate(z)
        return self._lnp

    def lnprob_grad(self, z):
        if np.any(z != self._z):
            self.evaluate(z)
        return self._lnp_grad

    def lnprior(self, q):
        if self._lnpriorfn is None:
            return 0.0, 0.0
        return self._lnpriorfn(q)

    def make_transform(self, z):
        if self.transform is not None:
            self._jacobian = self.transform.jacobian(z)
            self._lndetjac = self.transform.lndetjac(z)
            self._lndetjac_grad = self.transform.lndetjac_grad(z)
            return self.transform.transform(z)
        self._jacobian = 1.
        self._lndetjac = 0
        self._lndetjac_grad = 0
        return np.array(z)

    def nll(self, z):
        if np.any(z != self._z):
            self.evaluate(z)
        return -self._lnp / self.ndof, -self._lnp_grad / self.ndof

    def nll_nograd(self, z):
        if np.any(z != self._z):
            self.evaluate(z)
        return -self._lnp / self.ndof

    def lnprob_and_grad(self, z):
        if np.any(z != self._z):
            self.evaluate(z)
        return self._lnp, self._lnp_grad

    def residuals(self, z):
        raise NotImplementedError

    def check_grad(self, z, delta=1e-5):
        """(analytic, central-difference) gradients (forcepho/model.py:225-257)."""
        z0 = np.array(z, dtype=np.float64).copy()
        dlnp = self.lnprob_grad(z0)
        delta = np.zeros_like(dlnp) + delta
        dlnp_num = np.zeros(len(z0), dtype=np.float64)
        for i, dp in enumerate(delta):
            theta = z0.copy()
            theta[i] -= dp
            imlo = self.lnprob(theta.copy())
            theta[i] += 2 * dp
            imhi = self.lnprob(theta.copy())
            dlnp_num[i] = (imhi - imlo) / (2 * dp)
        return dlnp, dlnp_num


class FastPosterior(Posterior):
    """Posterior evaluated by a device proposer (forcepho/model.py:260-405)."""

    def __init__(self, proposer, patch, scene=None, lnprior=None, transform=None,
                 name="", print_interval=1000, verbose=False, debug=False,
                 logging=False, sampling=True, **kwargs):
        super().__init__(transform=transform, **kwargs)
        self.proposer = proposer
        self.patch = patch
        self.scene = scene if scene else self.patch.scene
        self._lnpriorfn = lnprior
        self.sampling = sampling
        self.ncall = 0
        self._z = -99
        self._q = -99
        self.mhalf = np.array(-0.5, dtype=np.float64)
        self.debug = debug
        self.logging = logging
        self.verbose = verbose
        self.print_interval = print_interval
        self.name = name
        if self.logging:
            self.pos_history, self.lnp_history, self.grad_history = [], [], []

    def evaluate(self, z):
        q = self.make_transform(z)
        self.scene.set_all_source_params(q)
        proposal = self.scene.get_proposal()
        lpr, lpr_grad = self.lnprior(q)
        ret = self.proposer.evaluate_proposal(proposal, patch=self.patch)
        if len(ret) == 3:
            chi2, chi2_derivs, self._residuals = ret
        else:
            chi2, chi2_derivs = ret
        ll = self.mhalf * np.array(chi2.astype(np.float64).sum(), dtype=np.float64)
        ll_grad = self.mhalf * self.stack_grad(chi2_derivs)
        self.ncall += 1
        if self.logging:
            self._lnlike, self._lnlike_grad = ll, ll_grad
            self._lnprior, self._lnprior_grad = lpr, lpr_grad
            self.pos_history.append(z)
            self.lnp_history.append(ll)
            self.grad_history.append(ll_grad)
        self._lnp = ll + lpr
        self._lnp_grad = (ll_grad + lpr_grad) * self._jacobian
        self._q = q
        self._z = z
        if self.sampling:
            self._lnp += self._lndetjac
            self._lnp_grad += self._lndetjac_grad
        if self.verbose and np.mod(self.ncall, self.print_interval) == 0:
            print(f"-------\n {self.name} @ {self.ncall}\nq: {q}\nlnlike: {ll}")

    def stack_grad(self, chi2_derivs):
        return stack_grad(chi2_derivs, int(self.patch.n_sources), int(len(self.patch.band_N)))

    def residuals(self, q, unpack=True):
        orig = self.patch.return_residual
        self.patch.return_residual = True
        self.scene.set_all_source_params(q)
        proposal = self.scene.get_proposal()
        chi2, chi2_derivs, self._residuals = self.proposer.evaluate_proposal(
            proposal, patch=self.patch, unpack=unpack)
        self.patch.return_residual = orig
        return self._residuals


GPUPosterior = FastPosterior


def sigmoid(x):
    return expit(x)


def logit(y):
    return -np.log(1. / y - 1.)


def sigmoid_grad(z):
    return (0.5 / np.cosh(0.5 * z)) ** 2


class Transform:
    """Identity transform (forcepho/model.py:572-593)."""

    def __init__(self, ndim):
        self.ndim = ndim

    def transform(self, z):
        return np.array(z)

    def jacobian(self, z):
        return np.ones(len(z))

    def lndetjac(self, z):
        return np.log(np.abs(np.prod(self.jacobian(z))))

    def lndetjac_grad(self, z):
        return np.zeros(len(z))

    def inverse_transform(self, Theta):
        return np.array(Theta)


class BoundedTransform(Transform):
    """q = lower + sigmoid(z) (upper - lower) (forcepho/model.py:596-636)."""

    def __init__(self, lower, upper):
        self.lower = np.atleast_1d(lower)
        self.upper = np.atleast_1d(upper)
        assert len(self.upper) == len(self.lower)
        self.ndim = len(self.upper)

    @property
    def range(self):
        return self.upper - self.lower

    def transform(self, z):
        return self.lower + sigmoid(z) * self.range

    def jacobian(self, z):
        return self.range * sigmoid_grad(z)

    def lndetjac_grad(self, z):
        return 1 - 2.0 * sigmoid(z)

    def inverse_transform(self, Theta):
        return logit((Theta - self.lower) / self.range)
