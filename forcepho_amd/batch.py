"""Many patches in one set of launches, driven from torch on the device.

``PatchBatch`` wraps a ``fpho_batch``: the patches' proposals, chi2 and
gradient buffers are torch tensors in HBM, and every call is stream ordered
on torch's current stream (no host synchronisation, no allocation), so the
whole evaluation -- parameters -> proposals (FP64 splines) -> image Gaussians
-> persistent pixel kernel -> FP64 reduction -> stacked gradient -- can be
captured in a HIP graph.

There is no reference counterpart: forcepho evaluates one patch per call
(forcepho/proposal.py:96-152).  Per patch the results equal those of
``HIPProposer.evaluate_proposal`` (same kernels, same fixed reduction order).
"""

import ctypes

import numpy as np
import torch

from . import _lib
from .limits import NPARAMS, source_struct_dtype
from .spline import get_splines

__all__ = ["PatchBatch", "DeviceSplines"]


class DeviceSplines:
    """Sersic amplitude spline tables in HBM (fpho_spline)."""

    def __init__(self, splines=None, device=0):
        _lib.init(device)
        self.splines = splines or get_splines()
        sp = self.splines
        self._keep = [np.ascontiguousarray(sp.coeffs, dtype=np.float64),
                      np.ascontiguousarray(sp.tx, dtype=np.float64),
                      np.ascontiguousarray(sp.ty, dtype=np.float64),
                      np.ascontiguousarray(sp.radii, dtype=np.float64)]
        c, tx, ty, radii = self._keep
        h = ctypes.c_void_p()
        _lib.check(_lib.lib().fpho_spline_create(
            _lib.fptr(c, ctypes.c_double), int(sp.n_amp), _lib.fptr(tx, ctypes.c_double),
            len(tx), _lib.fptr(ty, ctypes.c_double), len(ty),
            _lib.fptr(radii, ctypes.c_double), float(sp.sersic_range[0]), ctypes.byref(h)))
        self.handle = h

    def free(self):
        if self.handle:
            _lib.lib().fpho_spline_free(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def _on_stream(theta, n, stream=None):
    """``theta`` as a contiguous float64 device vector of length ``n``, kept
    alive for ``stream``: a contiguous copy made here is freed when the
    caller drops it, and torch's allocator may hand its memory to work on the
    current stream while the library still reads it on ``stream``."""
    assert theta.dtype == torch.float64 and theta.is_cuda and theta.numel() == n
    theta = theta.contiguous()
    if stream is not None:
        theta.record_stream(stream)
    return theta


def _stream_ptr(stream=None):
    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(int(s.cuda_stream))


class PatchBatch:
    """A set of device patches (objects with ``device_patch`` handles, e.g.
    ``HIPPatch`` after ``send_to_device``) evaluated together."""

    def __init__(self, patches, device=0, precision="mixed", launch=None):
        _lib.init(device)
        self.patches = list(patches)
        self.device = torch.device("cuda", device)
        handles = (ctypes.c_void_p * len(self.patches))(
            *[p.device_patch.value if isinstance(p.device_patch, ctypes.c_void_p)
              else p.device_patch for p in self.patches])
        h = ctypes.c_void_p()
        _lib.check(_lib.lib().fpho_batch_create(handles, len(self.patches), ctypes.byref(h)))
        self.handle = h
        n = len(self.patches) + 1
        self.src_off = np.zeros(n, dtype=np.int64)
        self.chi2_off = np.zeros(n, dtype=np.int64)
        self.dchi2_off = np.zeros(n, dtype=np.int64)
        self.theta_off = np.zeros(n, dtype=np.int64)
        _lib.check(_lib.lib().fpho_batch_layout(
            self.handle, _lib.fptr(self.src_off, ctypes.c_int64),
            _lib.fptr(self.chi2_off, ctypes.c_int64), _lib.fptr(self.dchi2_off, ctypes.c_int64),
            _lib.fptr(self.theta_off, ctypes.c_int64)))
        self.n_sources_total = int(self.src_off[-1])
        self.n_theta = int(self.theta_off[-1])
        dev = self.device
        self.proposals = torch.zeros(self.n_sources_total * source_struct_dtype.itemsize,
                                     dtype=torch.uint8, device=dev)
        self.chi2 = torch.zeros(int(self.chi2_off[-1]), dtype=torch.float64, device=dev)
        self.dchi2 = torch.zeros(int(self.dchi2_off[-1]), dtype=torch.float64, device=dev)
        self.lnlike = torch.zeros(len(self.patches), dtype=torch.float64, device=dev)
        self.grad = torch.zeros(self.n_theta, dtype=torch.float64, device=dev)
        # parameter -> patch index (for the autograd backward)
        counts = np.diff(self.theta_off)
        self.param_patch = torch.as_tensor(np.repeat(np.arange(len(self.patches)), counts),
                                           device=dev)
        self.set_precision(precision)
        if launch is not None:
            self.set_launch(**launch)

    # -- configuration ------------------------------------------------------
    def set_precision(self, mode):
        self.precision = mode
        _lib.check(_lib.lib().fpho_batch_set_precision(self.handle, _lib.PRECISION[mode]))

    def set_launch(self, pix_per_lane=4, waves_per_group=4, persistent=1, groups_per_cu=8,
                   cull=1):
        l = _lib.Launch(pix_per_lane, waves_per_group, persistent, groups_per_cu, cull)
        _lib.check(_lib.lib().fpho_batch_set_launch(self.handle, ctypes.byref(l)))
        self.launch = dict(pix_per_lane=pix_per_lane, waves_per_group=waves_per_group,
                           persistent=persistent, groups_per_cu=groups_per_cu, cull=cull)

    def set_timing(self, on=True):
        _lib.check(_lib.lib().fpho_batch_set_timing(self.handle, int(bool(on))))

    def last_timing(self):
        """(convert, pixel pass, reduce) milliseconds of the last evaluation."""
        ms = np.zeros(3, dtype=np.float32)
        _lib.check(_lib.lib().fpho_batch_last_timing(self.handle, _lib.fptr(ms, ctypes.c_float)))
        return tuple(float(x) for x in ms)

    def status(self, clear=True):
        """The FPHO_STATUS_* bits the batch's kernels recorded since the last
        cleared read (waits for the device); ``_lib.describe_status`` names
        them."""
        w = ctypes.c_uint32()
        _lib.check(_lib.lib().fpho_batch_status(self.handle, ctypes.byref(w), int(bool(clear))))
        return w.value

    def info(self):
        v = [ctypes.c_int64() for _ in range(4)]
        _lib.check(_lib.lib().fpho_batch_info(self.handle, *[ctypes.byref(x) for x in v]))
        return dict(n_tiles=v[0].value, n_pairs=v[1].value, npix=v[2].value,
                    n_gauss=v[3].value)

    # -- data movement ------------------------------------------------------
    def set_proposals(self, proposals):
        """Copy host proposals (a list with one ``source_struct_dtype`` array
        per patch, or one concatenated array) to the device buffer."""
        if isinstance(proposals, (list, tuple)):
            proposals = np.concatenate([np.asarray(p, dtype=source_struct_dtype)
                                        for p in proposals])
        proposals = np.ascontiguousarray(proposals, dtype=source_struct_dtype)
        assert len(proposals) == self.n_sources_total
        host = torch.from_numpy(proposals.view(np.uint8).reshape(-1).copy())
        self.proposals.copy_(host)

    def chi2_host(self, i=None):
        c = self.chi2.cpu().numpy()
        if i is None:
            return [c[self.chi2_off[k]:self.chi2_off[k + 1]] for k in range(len(self.patches))]
        return c[self.chi2_off[i]:self.chi2_off[i + 1]]

    def dchi2_host(self, i=None):
        d = self.dchi2.cpu().numpy()
        out = []
        for k in range(len(self.patches)):
            nb = int(self.chi2_off[k + 1] - self.chi2_off[k])
            ns = int(self.src_off[k + 1] - self.src_off[k])
            out.append(d[self.dchi2_off[k]:self.dchi2_off[k + 1]].reshape(nb, ns, NPARAMS))
        return out if i is None else out[i]

    # -- stream-ordered evaluation -----------------------------------------
    def evaluate(self, proposals=None, stream=None):
        """chi2 and dchi2 of every patch for the proposals in ``proposals``
        (a uint8 device tensor of fpho_source records; default: the batch's
        own buffer)."""
        props = self.proposals if proposals is None else proposals
        _lib.check(_lib.lib().fpho_batch_evaluate_async(
            self.handle, _lib.dptr(props), _lib.dptr(self.chi2), _lib.dptr(self.dchi2),
            _stream_ptr(stream)))
        return self.chi2, self.dchi2

    def theta_to_proposals(self, theta, splines, stream=None):
        theta = _on_stream(theta, self.n_theta, stream)
        _lib.check(_lib.lib().fpho_batch_theta_to_proposals_async(
            self.handle, splines.handle, _lib.dptr(theta), _lib.dptr(self.proposals),
            _stream_ptr(stream)))
        return self.proposals

    def evaluate_theta(self, theta, splines, stream=None):
        """chi2 and dchi2 of every patch at the parameter vector ``theta``
        (float64 on the device): the image Gaussians are built from FP64
        source parameters (the FP32 records are written to ``proposals``
        too)."""
        theta = _on_stream(theta, self.n_theta, stream)
        _lib.check(_lib.lib().fpho_batch_evaluate_theta_async(
            self.handle, splines.handle, _lib.dptr(theta), _lib.dptr(self.proposals),
            _lib.dptr(self.chi2), _lib.dptr(self.dchi2), _stream_ptr(stream)))
        return self.chi2, self.dchi2

    def lnlike_and_grad(self, theta, splines, stream=None):
        """lnlike [n_patches] = -chi2/2 per patch and its gradient [n_theta]
        with respect to theta (per source [flux_1..flux_NB, ra, dec, q, pa,
        nsersic, rh]); all on the device."""
        self.evaluate_theta(theta, splines, stream)
        _lib.check(_lib.lib().fpho_batch_stack_grad_async(
            self.handle, _lib.dptr(self.chi2), _lib.dptr(self.dchi2), _lib.dptr(self.lnlike),
            _lib.dptr(self.grad), _stream_ptr(stream)))
        return self.lnlike, self.grad

    def free(self):
        if getattr(self, "handle", None):
            _lib.lib().fpho_batch_free(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
