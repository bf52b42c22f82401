"""Write tests/golden/spline_golden.npz: scipy RectBivariateSpline values and
derivatives (the reference's amplitude interpolants, forcepho/sources.py:
623-694, s=None) on the packaged amplitude table at seeded points, including
grid nodes and the table edges.  Run with the scipy of this image (1.15.3)."""
import os

import numpy as np
import scipy
from scipy.interpolate import RectBivariateSpline

HERE = os.path.dirname(os.path.abspath(__file__))
TABLE = os.path.join(HERE, "..", "..", "forcepho_amd", "data", "sersic_splinedata_large.npz")


def main():
    with np.load(TABLE) as f:
        n, r, A = f["nsersic"], f["rh"], f["amplitudes"]
    ng, rg = np.unique(n), np.unique(r)
    sp = [RectBivariateSpline(ng, rg, A[:, i].reshape(len(ng), len(rg)), s=None)
          for i in range(A.shape[1])]
    rng = np.random.default_rng(20240601)
    pts = [(rng.uniform(ng[0], ng[-1]), rng.uniform(rg[0], rg[-1])) for _ in range(64)]
    pts += [(rng.uniform(1, 5), rng.uniform(0.05, 0.3)) for _ in range(32)]
    pts += [(float(ng[i]), float(rg[j])) for i, j in [(0, 0), (10, 20), (54, 100), (3, 50)]]
    pts += [(float(ng[0]), 0.5), (float(ng[-1]), 0.5), (3.0, 0.0), (3.0, 1.0)]
    pts = np.array(pts)
    val = np.array([[s(a, b)[0, 0] for s in sp] for a, b in pts])
    dn = np.array([[s(a, b, dx=1)[0, 0] for s in sp] for a, b in pts])
    dr = np.array([[s(a, b, dy=1)[0, 0] for s in sp] for a, b in pts])
    out = os.path.join(HERE, "spline_golden.npz")
    np.savez(out, points=pts, value=val, d_nsersic=dn, d_rh=dr,
             scipy_version=np.array(scipy.__version__))
    print(out, pts.shape)


if __name__ == "__main__":
    main()
