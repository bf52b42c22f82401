"""forcepho_amd -- MI355X-native likelihood-and-gradient engine for forcepho
patches.

Host code (this package) keeps the reference's Patch / Proposer.
evaluate_proposal / Posterior surface; the arithmetic runs in hand-written
gfx950 HIP kernels of ``libfpho_hip.so`` reached through a C ABI
(include/fpho.h).
"""

from .limits import (MAXBANDS, MAXRADII, MAXSOURCES, NPARAMS, MAX_EXP_ARG,
                     source_struct_dtype, response_struct_dtype, psf_struct_dtype)

__version__ = "0.1.0"
