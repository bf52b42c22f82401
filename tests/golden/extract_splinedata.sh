#!/bin/bash
# Convert the Sersic-mixture amplitude table shipped with the forcepho tests
# (tests/data/sersic_splinedata_large.h5) into the package's data file
# forcepho_amd/data/sersic_splinedata_large.npz.
# h5py is not available in this image, so the datasets are dumped as raw
# little-endian binary with h5dump and reassembled with numpy.
# Usage: extract_splinedata.sh [path/to/sersic_splinedata_large.h5]
set -euo pipefail
SRC=${1:-/root/reference/tests/data/sersic_splinedata_large.h5}
H5DUMP=${H5DUMP:-/opt/conda/bin/h5dump}
HERE=$(cd "$(dirname "$0")" && pwd)
TMP=$(mktemp -d)
trap 'rm -r "$TMP"' EXIT
for d in amplitudes nsersic radii rh; do
  "$H5DUMP" -b LE -d "/$d" -o "$TMP/$d.bin" "$SRC" > /dev/null
done
python3 - "$TMP" "$HERE/../../forcepho_amd/data/sersic_splinedata_large.npz" <<'PY'
import sys
import numpy as np
tmp, out = sys.argv[1], sys.argv[2]
amp = np.fromfile(f"{tmp}/amplitudes.bin", "<f8").reshape(-1, 9)
nsersic = np.fromfile(f"{tmp}/nsersic.bin", "<f4")   # stored as float32
rh = np.fromfile(f"{tmp}/rh.bin", "<f8")
radii = np.fromfile(f"{tmp}/radii.bin", "<f8")
assert amp.shape[0] == nsersic.size == rh.size
np.savez(out, amplitudes=amp, nsersic=nsersic, rh=rh, radii=radii)
print(out, amp.shape, nsersic.dtype, radii)
PY
