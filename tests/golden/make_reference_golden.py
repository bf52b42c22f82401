"""Write tests/golden/reference_c1.npz: chi2 and dchi2 of the reference's own
CPU kernel (forcepho/src/compute_gaussians_kernel.cc, compiled from its
sources into oracle/_ref by oracle/Makefile) on the seeded configuration-1
patch, at the truth and at 8 seeded proposals.  The patch is rebuilt from
the same seeds by the tests, so only the outputs are stored."""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", ".."))
sys.path.insert(0, os.path.join(HERE, ".."))

import oracle  # noqa: E402
from forcepho_amd import synth  # noqa: E402
from helpers import make_config_patch, proposal_for  # noqa: E402


def main():
    patch, scene, theta = make_config_patch(1)
    thetas = np.vstack([theta[None], synth.make_proposals(theta, 1, 8, config=1)])
    ref = oracle.ReferenceKernel(patch)
    chi2, dchi2 = [], []
    for t in thetas:
        c, d = ref.evaluate(proposal_for(scene, t))
        chi2.append(c)
        dchi2.append(d)
    out = os.path.join(HERE, "reference_c1.npz")
    np.savez(out, theta=thetas, chi2=np.array(chi2), dchi2=np.array(dchi2),
             data_checksum=np.array(float(np.sum(patch.data, dtype=np.float64))))
    print(out, np.array(chi2).ravel()[:3])


if __name__ == "__main__":
    main()
