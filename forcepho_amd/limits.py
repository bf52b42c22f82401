"""Compile-time limits and the struct layouts shared with the HIP library.

The limits are those of forcepho/src/kernel_limits.h:5-26 (MAXBANDS,
MAXSOURCES, NPARAMS, MAX_EXP_ARG, MAXRADII).  MAXSOURCES only sizes the compat
``response_struct_dtype``; the native outputs are sized by the patch.

The numpy dtypes mirror ``fpho_source`` / ``fpho_psf_gauss`` of include/fpho.h,
which are the reference's ``Source`` (forcepho/src/proposal.cu:13-27,
forcepho/proposal.py:229-240) and ``PSFSourceGaussian``
(forcepho/src/patch.cc:103-119, forcepho/patches/storage.py:35-36).
"""

import numpy as np

MAXBANDS = 30
MAXSOURCES = 30
NPARAMS = 7
MAX_EXP_ARG = 36.0
MAXRADII = 10
MAX_NATIVE_SOURCES = 128

source_float_dt = np.float32
source_struct_dtype = np.dtype([('ra', source_float_dt),
                                ('dec', source_float_dt),
                                ('q', source_float_dt),
                                ('pa', source_float_dt),
                                ('nsersic', source_float_dt),
                                ('rh', source_float_dt),
                                ('fluxes', source_float_dt, MAXBANDS),
                                ('mixture_amplitudes', source_float_dt, MAXRADII),
                                ('damplitude_drh', source_float_dt, MAXRADII),
                                ('damplitude_dnsersic', source_float_dt, MAXRADII)],
                               align=True)

response_float_dt = source_float_dt
response_struct_dtype = np.dtype([('dchi2_dparam', response_float_dt,
                                   NPARAMS * MAXSOURCES)], align=True)

PSF_COLS = ["amp", "xcen", "ycen", "Cxx", "Cyy", "Cxy"]
psf_struct_dtype = np.dtype([(c, np.float32) for c in PSF_COLS] +
                            [("sersic_bin", np.int32)])

assert source_struct_dtype.itemsize == 264
assert response_struct_dtype.itemsize == 840
assert psf_struct_dtype.itemsize == 28
