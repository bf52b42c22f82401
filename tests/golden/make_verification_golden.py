"""Extract the inputs and outputs of forcepho's verification test into test
fixtures (run in a checkout that has the reference tree and h5dump; the
fixtures it writes are committed):

* ``reference_2021Nov30_f200w_sersic2.2_rhalf0.10.fits`` -- a copy of
  forcepho's ``tests/verification/data/reference-2021Nov30_f200w_sersic=2.2_
  rhalf=0.10.fits``: a 63 x 32 image of one Sersic source (n = 2.2,
  r_h = 0.10 arcsec, flux 1, q = 1, pa = 0) rendered by forcepho's pure-python
  path (``forcepho.slow``, make_reference.py), HDU 1-7 its derivative images
  dFLUX/d(param), the last HDU the one-row catalog.  Its header names the PSF
  file it was made with, ``psf_jwst_oct21_ng4m0_f200w.fits``.
* ``verification_f200w_inputs.npz`` -- what that rendering used besides the
  catalog:
  - ``psf_*``: the 4-Gaussian PSF table (amp, xcr, ycr, Cxx, Cyy, Cxy,
    sersic_bin) of ``tests/verification/data/psf_jwst_oct21_ng4m0_f200w.fits``
    (HDU 3);
  - ``amplitudes, nsersic, rh, radii``: the Sersic-mixture table of
    ``tests/data/sersic_mog_model.smooth=0.0150.h5`` (make_reference.py's
    default splinedata), dumped with h5dump as stored (float32 amplitudes,
    nsersic, rh; float64 radii), in the layout of
    ``forcepho_amd/data/sersic_splinedata_large.npz``.

Usage: python tests/golden/make_verification_golden.py [reference_root]
"""
import os
import shutil
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from forcepho_amd.patches import fits  # noqa: E402

IMAGE = "reference_2021Nov30_f200w_sersic2.2_rhalf0.10.fits"
INPUTS = "verification_f200w_inputs.npz"


def h5_dataset(h5dump, path, name, dtype, tmp):
    out = os.path.join(tmp, name + ".bin")
    subprocess.run([h5dump, "-b", "LE", "-d", "/" + name, "-o", out, path], check=True,
                   stdout=subprocess.DEVNULL)
    return np.fromfile(out, dtype)


def main(root="/root/reference"):
    tests = os.path.join(root, "tests")
    vdata = os.path.join(tests, "verification", "data")
    shutil.copyfile(os.path.join(vdata, "reference-2021Nov30_f200w_sersic=2.2_rhalf=0.10.fits"),
                    os.path.join(HERE, IMAGE))

    psf = fits.getdata(os.path.join(vdata, "psf_jwst_oct21_ng4m0_f200w.fits"), 3)
    out = {f"psf_{k}": np.asarray(psf[k]).astype(psf.dtype[k].newbyteorder("="))
           for k in psf.dtype.names}

    h5dump = os.environ.get("H5DUMP", "/opt/conda/bin/h5dump")
    mog = os.path.join(tests, "data", "sersic_mog_model.smooth=0.0150.h5")
    with tempfile.TemporaryDirectory() as tmp:
        amp = h5_dataset(h5dump, mog, "amplitudes", "<f4", tmp)
        out["nsersic"] = h5_dataset(h5dump, mog, "nsersic", "<f4", tmp)
        out["rh"] = h5_dataset(h5dump, mog, "rh", "<f4", tmp)
        out["radii"] = h5_dataset(h5dump, mog, "radii", "<f8", tmp)
    out["amplitudes"] = amp.reshape(out["nsersic"].size, -1)
    assert out["amplitudes"].shape[1] == out["radii"].size
    np.savez(os.path.join(HERE, INPUTS), **out)
    print(os.path.join(HERE, INPUTS), {k: v.shape for k, v in out.items()})


if __name__ == "__main__":
    main(*sys.argv[1:])
