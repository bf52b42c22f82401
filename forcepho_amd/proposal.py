"""Proposers: evaluate chi2 and its gradient for a proposal on a patch.

``HIPProposer.evaluate_proposal`` is the drop-in for
``GPUProposer.evaluate_proposal`` (forcepho/proposal.py:96-152): it returns
``chi2`` (float64 [n_bands]) and ``chi2_derivs`` (float32 [n_bands,
n_sources, 7], order d/d(flux, ra, dec, q, pa, nsersic, rh)), plus the
residual images when ``patch.return_residual``.  ``chi2`` uses the kernel's
convention sum_pix (r^2 - d^2) ierr^2 (compute_gaussians_kernel.cu:277-286).
"""

import ctypes

import numpy as np

from . import _lib
from .limits import (MAXBANDS, MAXRADII, MAXSOURCES, NPARAMS, response_struct_dtype,
                     source_struct_dtype)

__all__ = ["ProposerBase", "HIPProposer", "GPUProposer", "Proposer",
           "source_struct_dtype", "response_struct_dtype"]


class ProposerBase:

    def __init__(self):
        pass

    def evaluate_proposal(self, proposal, **kwargs):
        raise NotImplementedError

    def unpack_residuals(self, residuals_flat, patch, reshape=False):
        """Split flat padded residuals per exposure (forcepho/proposal.py:38-52)."""
        residuals = np.split(residuals_flat, np.cumsum(patch.exposure_N)[:-1])
        if reshape:
            for e, residual in enumerate(residuals):
                residual = residual[:patch.original_sizes[e]]
                residuals[e] = residual.reshape(patch.original_shapes[e])
        return residuals


class HIPProposer(ProposerBase):
    """Synchronous evaluation through ``fpho_evaluate`` (host in, host out)."""

    def __init__(self, patch=None, chi_dtype=np.float64, **kwargs):
        self.patch = patch
        self.chi_dtype = chi_dtype

    def evaluate_proposal(self, proposal, patch=None, verbose=False, unpack=True):
        if patch is None:
            patch = self.patch
        n_bands, n_sources = int(len(patch.band_N)), int(patch.n_sources)
        proposal = np.ascontiguousarray(proposal, dtype=source_struct_dtype)
        if len(proposal) != n_sources:
            raise ValueError(f"proposal has {len(proposal)} sources, patch has {n_sources}")
        chi_out = np.empty(n_bands, dtype=np.float64)
        derivs = np.empty((n_bands, n_sources, NPARAMS), dtype=np.float32)
        _lib.check(_lib.lib().fpho_evaluate(
            patch.device_patch, ctypes.c_void_p(proposal.ctypes.data),
            _lib.fptr(chi_out, ctypes.c_double), _lib.fptr(derivs, ctypes.c_float)))
        chi_out = chi_out.astype(self.chi_dtype)
        if patch.return_residual:
            residuals = patch.retrieve_array("residual")
            if unpack:
                residuals = self.unpack_residuals(residuals, patch)
            return chi_out, derivs, residuals
        return chi_out, derivs

    def evaluate_response(self, proposal, patch=None):
        """The reference kernel's raw output layout: ``chi2`` (float64
        [n_bands]) and ``Response[n_bands]`` (``response_struct_dtype``,
        forcepho/proposal.py:242-244), i.e. what ``EvaluateProposal`` writes
        into ``cuda.Out(chi_derivs_out)`` before ``evaluate_proposal``
        reshapes it to [n_bands, MAXSOURCES, NPARAMS] (forcepho/proposal.py:
        119-141).  Entries past 7 n_sources of each band are zero."""
        if patch is None:
            patch = self.patch
        n_bands, n_sources = int(len(patch.band_N)), int(patch.n_sources)
        proposal = np.ascontiguousarray(proposal, dtype=source_struct_dtype)
        if len(proposal) != n_sources:
            raise ValueError(f"proposal has {len(proposal)} sources, patch has {n_sources}")
        chi_out = np.empty(n_bands, dtype=np.float64)
        resp = np.empty(n_bands, dtype=response_struct_dtype)
        _lib.check(_lib.lib().fpho_evaluate_response(
            patch.device_patch, ctypes.c_void_p(proposal.ctypes.data),
            _lib.fptr(chi_out, ctypes.c_double), ctypes.c_void_p(resp.ctypes.data)))
        return chi_out.astype(self.chi_dtype), resp


# names used by the reference for its GPU proposer
GPUProposer = HIPProposer
Proposer = HIPProposer
