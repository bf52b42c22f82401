"""ctypes binding of measure/libfpho_measure.so (measure/measure.h): the
measurement-only kernels bench.py uses.  Not part of the product; nothing in
forcepho_amd imports it.

* ``issue_calibration(kind)`` -- VALU issue rate of FP32 FMA / exp chains.
* ``count_pairs_in_cutoff(batch, proposals)`` -- (pixel, image Gaussian)
  pairs with exp argument <= 36, counted with the pixel pass's own FP32 test.
"""

import ctypes
import os

from forcepho_amd import _lib

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "libfpho_measure.so")

_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIBPATH):
            raise ImportError(f"{LIBPATH} is missing: build it with "
                              "`python -c 'import __graft_entry__ as g; g.build()'`")
        _lib.lib()   # the product library (and torch's HIP runtime) first
        h = ctypes.CDLL(LIBPATH)
        h.fpho_measure_last_error.argtypes = []
        h.fpho_measure_last_error.restype = ctypes.c_char_p
        h.fpho_measure_issue_calibration.argtypes = [
            ctypes.c_int, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double),
            ctypes.POINTER(ctypes.c_int)]
        h.fpho_measure_issue_calibration.restype = ctypes.c_int
        h.fpho_measure_batch_count_pairs.argtypes = [ctypes.c_void_p, ctypes.c_void_p,
                                                     ctypes.POINTER(ctypes.c_int64)]
        h.fpho_measure_batch_count_pairs.restype = ctypes.c_int
        _LIB = h
    return _LIB


def _check(rc):
    if rc != 0:
        raise _lib.FphoError(rc, lib().fpho_measure_last_error().decode())


def issue_calibration(kind):
    """(kernel ms, wave64 VALU instructions issued, shader clock kHz)."""
    ms, wi, clk = ctypes.c_double(), ctypes.c_double(), ctypes.c_int()
    _check(lib().fpho_measure_issue_calibration(int(kind), ctypes.byref(ms), ctypes.byref(wi),
                                                ctypes.byref(clk)))
    return ms.value, wi.value, clk.value


def count_pairs_in_cutoff(batch, proposals=None):
    """In-cutoff pairs of a PatchBatch for ``proposals`` (default: its own
    proposal buffer); synchronous."""
    n = ctypes.c_int64()
    props = batch.proposals if proposals is None else proposals
    _check(lib().fpho_measure_batch_count_pairs(batch.handle, _lib.dptr(props), ctypes.byref(n)))
    return n.value
