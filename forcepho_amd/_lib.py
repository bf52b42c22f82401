"""ctypes binding of libfpho_hip.so (declared in include/fpho.h).

Loading fails loudly when the library has not been built: there is no CPU
fallback anywhere in the product path.
"""

import ctypes
import os

import numpy as np

from ._build import LIBPATH

c_int = ctypes.c_int
c_i32 = ctypes.c_int32
c_i64 = ctypes.c_int64
c_void_p = ctypes.c_void_p
c_double_p = ctypes.POINTER(ctypes.c_double)
c_float_p = ctypes.POINTER(ctypes.c_float)
c_i16_p = ctypes.POINTER(ctypes.c_int16)
c_i32_p = ctypes.POINTER(ctypes.c_int32)
c_i64_p = ctypes.POINTER(ctypes.c_int64)

EXPORTED = [
    "fpho_last_error", "fpho_version", "fpho_init", "fpho_device_cu_count",
    "fpho_patch_create", "fpho_patch_set_meta", "fpho_patch_swap_data_residual",
    "fpho_patch_read", "fpho_patch_info", "fpho_patch_set_launch",
    "fpho_patch_set_precision", "fpho_patch_set_residual_writes", "fpho_patch_status",
    "fpho_patch_free",
    "fpho_evaluate", "fpho_evaluate_async", "fpho_evaluate_async_noalloc",
    "fpho_evaluate_response",
    "fpho_evaluate_response_async",
    "fpho_batch_create", "fpho_batch_layout", "fpho_batch_info",
    "fpho_batch_set_launch", "fpho_batch_set_precision",
    "fpho_batch_evaluate_async", "fpho_batch_set_timing", "fpho_batch_last_timing",
    "fpho_batch_status", "fpho_batch_free",
    "fpho_spline_create", "fpho_spline_free",
    "fpho_batch_theta_to_proposals_async", "fpho_batch_evaluate_theta_async",
    "fpho_batch_stack_grad_async",
]

ARRAY_IDS = {"data": 0, "ierr": 1, "xpix": 2, "ypix": 3, "residual": 4}
PRECISION = {"fp32": 0, "fp64_reduce": 1, "mixed": 2, "mixed_centre": 3}
# FPHO_STATUS_* bits of the device status word
STATUS_BITS = {0x1: "non-finite proposal value", 0x2: "axis ratio q <= 0",
               0x4: "image Gaussian covariance not positive definite",
               0x8: "non-finite image Gaussian parameter",
               0x10: "non-finite parameter vector entry"}


def describe_status(word):
    """Names of the FPHO_STATUS_* bits set in ``word``."""
    return [name for bit, name in STATUS_BITS.items() if word & bit]


class FphoError(RuntimeError):
    """An error reported by libfpho_hip (the message comes from the library)."""

    def __init__(self, code, msg):
        super().__init__(f"fpho error {code}: {msg}")
        self.code = code


class PatchDesc(ctypes.Structure):
    """== fpho_patch_desc (include/fpho.h)."""
    _fields_ = [("npix", c_i64),
                ("data", c_float_p), ("ierr", c_float_p),
                ("xpix", c_float_p), ("ypix", c_float_p),
                ("n_bands", c_i32), ("n_exp", c_i32),
                ("band_start", c_i16_p), ("band_N", c_i16_p),
                ("exposure_start", c_i32_p), ("exposure_N", c_i32_p),
                ("n_sources", c_i32), ("n_radii", c_i32),
                ("rad2", c_float_p), ("D", c_float_p), ("crpix", c_float_p),
                ("crval", c_float_p), ("CW", c_float_p), ("G", c_float_p),
                ("n_psf_per_source", c_i32_p),
                ("psfgauss", c_void_p), ("n_psfgauss", c_i64),
                ("psfgauss_start", c_i32_p)]


class Launch(ctypes.Structure):
    """== fpho_launch (include/fpho.h)."""
    _fields_ = [("pix_per_lane", c_i32), ("waves_per_group", c_i32),
                ("persistent", c_i32), ("groups_per_cu", c_i32), ("cull", c_i32)]


_LIB = None


def _declare(lib):
    P = c_void_p
    sig = {
        "fpho_last_error": ([], ctypes.c_char_p),
        "fpho_version": ([], c_int),
        "fpho_init": ([c_int], c_int),
        "fpho_device_cu_count": ([ctypes.POINTER(c_int)], c_int),
        "fpho_patch_create": ([ctypes.POINTER(PatchDesc), c_int, ctypes.POINTER(P)], c_int),
        "fpho_patch_set_meta": ([P, ctypes.POINTER(PatchDesc)], c_int),
        "fpho_patch_swap_data_residual": ([P], c_int),
        "fpho_patch_read": ([P, c_int, c_float_p], c_int),
        "fpho_patch_info": ([P, c_i32_p, c_i32_p, c_i64_p, c_i64_p], c_int),
        "fpho_patch_set_launch": ([P, ctypes.POINTER(Launch)], c_int),
        "fpho_patch_set_precision": ([P, c_int], c_int),
        "fpho_patch_set_residual_writes": ([P, c_int], c_int),
        "fpho_patch_status": ([P, ctypes.POINTER(ctypes.c_uint32), c_int], c_int),
        "fpho_patch_free": ([P], c_int),
        "fpho_evaluate": ([P, P, c_double_p, c_float_p], c_int),
        "fpho_evaluate_async": ([P, P, P, P, P], c_int),
        "fpho_evaluate_async_noalloc": ([P, P, P, P, P], c_int),
        "fpho_evaluate_response": ([P, P, c_double_p, P], c_int),
        "fpho_evaluate_response_async": ([P, P, P, P, P], c_int),
        "fpho_batch_create": ([ctypes.POINTER(P), c_int, ctypes.POINTER(P)], c_int),
        "fpho_batch_layout": ([P, c_i64_p, c_i64_p, c_i64_p, c_i64_p], c_int),
        "fpho_batch_info": ([P, c_i64_p, c_i64_p, c_i64_p, c_i64_p], c_int),
        "fpho_batch_set_launch": ([P, ctypes.POINTER(Launch)], c_int),
        "fpho_batch_set_precision": ([P, c_int], c_int),
        "fpho_batch_evaluate_async": ([P, P, P, P, P], c_int),
        "fpho_batch_set_timing": ([P, c_int], c_int),
        "fpho_batch_last_timing": ([P, c_float_p], c_int),
        "fpho_batch_status": ([P, ctypes.POINTER(ctypes.c_uint32), c_int], c_int),
        "fpho_batch_free": ([P], c_int),
        "fpho_spline_create": ([c_double_p, c_int, c_double_p, c_int, c_double_p, c_int,
                                c_double_p, ctypes.c_double, ctypes.POINTER(P)], c_int),
        "fpho_spline_free": ([P], c_int),
        "fpho_batch_theta_to_proposals_async": ([P, P, P, P, P], c_int),
        "fpho_batch_evaluate_theta_async": ([P, P, P, P, P, P, P], c_int),
        "fpho_batch_stack_grad_async": ([P, P, P, P, P, P], c_int),
    }
    for name, (args, res) in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res


def lib():
    """Load (once) and return the library; raises if it is not built."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIBPATH):
            raise ImportError(
                f"{LIBPATH} is missing: build it with "
                "`python -c 'import __graft_entry__ as g; g.build()'` "
                "(forcepho_amd has no CPU fallback)")
        # torch's wheel carries its own HIP runtime under the same soname
        # (libamdhip64.so.7). With torch loaded first, this library binds to
        # that copy and the process has one HIP runtime; loaded first, it
        # would bring the system copy in and torch's own copy, loaded later,
        # then finds no GPU.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        handle = ctypes.CDLL(LIBPATH)
        _declare(handle)
        _LIB = handle
    return _LIB


def check(code):
    """Raise FphoError for a nonzero return code."""
    if code != 0:
        msg = lib().fpho_last_error()
        raise FphoError(code, msg.decode() if msg else "")
    return code


def fptr(arr, ctype):
    """ctypes pointer to a contiguous numpy array (kept alive by the caller)."""
    assert arr.flags["C_CONTIGUOUS"], "array must be C-contiguous"
    return arr.ctypes.data_as(ctypes.POINTER(ctype))


def dptr(x):
    """Raw address of a torch tensor / numpy array / int as c_void_p."""
    if x is None:
        return None
    if isinstance(x, int):
        return c_void_p(x)
    if hasattr(x, "data_ptr"):
        return c_void_p(x.data_ptr())
    if isinstance(x, np.ndarray):
        return c_void_p(x.ctypes.data)
    raise TypeError(f"cannot take a device pointer of {type(x)}")


_INITIALISED = set()


def init(device=0):
    """Select the HIP device for this thread (idempotent)."""
    if device not in _INITIALISED:
        check(lib().fpho_init(int(device)))
        _INITIALISED.add(device)


def cu_count():
    n = c_int(0)
    check(lib().fpho_device_cu_count(ctypes.byref(n)))
    return n.value
