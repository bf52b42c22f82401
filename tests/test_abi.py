"""The C-ABI library: it loads, exports every entry point include/fpho.h
declares, its struct layouts match the ctypes/numpy mirrors, and input
validation fails with error codes before any device work (CPU only)."""
import ctypes
import os
import re
import subprocess
import sys

import numpy as np
import pytest

from forcepho_amd import _lib, limits

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "fpho.h")


def declared_functions():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(fpho_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = _lib.lib()
    names = declared_functions()
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(_lib.EXPORTED)


def test_struct_layouts_match_c(tmp_path):
    src = tmp_path / "layout.c"
    src.write_text(r'''
#include <stdio.h>
#include <stddef.h>
#include "fpho.h"
int main(void) {
  printf("%zu %zu %zu %zu\n", sizeof(fpho_source), sizeof(fpho_psf_gauss),
         sizeof(fpho_patch_desc), sizeof(fpho_launch));
  printf("%zu %zu %zu %zu\n", offsetof(fpho_patch_desc, n_bands),
         offsetof(fpho_patch_desc, n_sources), offsetof(fpho_patch_desc, psfgauss),
         offsetof(fpho_patch_desc, psfgauss_start));
  printf("%zu %zu\n", offsetof(fpho_source, fluxes), offsetof(fpho_source, damplitude_dnsersic));
  printf("%zu\n", sizeof(fpho_response));
  return 0;
}
''')
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), str(src), "-o",
                    str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()
    v = list(map(int, out))
    assert v[0] == limits.source_struct_dtype.itemsize == 264
    assert v[1] == limits.psf_struct_dtype.itemsize == 28
    assert v[2] == ctypes.sizeof(_lib.PatchDesc)
    assert v[3] == ctypes.sizeof(_lib.Launch)
    assert v[4] == _lib.PatchDesc.n_bands.offset
    assert v[5] == _lib.PatchDesc.n_sources.offset
    assert v[6] == _lib.PatchDesc.psfgauss.offset
    assert v[7] == _lib.PatchDesc.psfgauss_start.offset
    assert v[8] == limits.source_struct_dtype.fields["fluxes"][1]
    assert v[9] == limits.source_struct_dtype.fields["damplitude_dnsersic"][1]
    assert v[10] == limits.response_struct_dtype.itemsize == 840


def _tiny_desc(keep, **over):
    npix = 64
    arrs = dict(data=np.zeros(npix, np.float32), ierr=np.ones(npix, np.float32),
                xpix=np.zeros(npix, np.float32), ypix=np.zeros(npix, np.float32),
                band_start=np.zeros(1, np.int16), band_N=np.ones(1, np.int16),
                exposure_start=np.zeros(1, np.int32), exposure_N=np.full(1, npix, np.int32),
                rad2=np.ones(1, np.float32), D=np.eye(2, dtype=np.float32).ravel(),
                crpix=np.zeros(2, np.float32), crval=np.zeros(2, np.float32),
                CW=np.eye(2, dtype=np.float32).ravel(), G=np.ones(1, np.float32),
                n_psf_per_source=np.ones(1, np.int32), psfgauss_start=np.zeros(1, np.int32))
    psf = np.zeros(1, dtype=limits.psf_struct_dtype)
    psf["amp"], psf["Cxx"], psf["Cyy"] = 1.0, 1.0, 1.0
    arrs["psfgauss"] = psf
    arrs.update(over)
    keep.update(arrs)
    d = _lib.PatchDesc()
    d.npix = len(arrs["data"])
    for k in ("data", "ierr", "xpix", "ypix", "rad2", "D", "crpix", "crval", "CW", "G"):
        setattr(d, k, _lib.fptr(arrs[k], ctypes.c_float))
    for k in ("band_start", "band_N"):
        setattr(d, k, _lib.fptr(arrs[k], ctypes.c_int16))
    for k in ("exposure_start", "exposure_N", "n_psf_per_source", "psfgauss_start"):
        setattr(d, k, _lib.fptr(arrs[k], ctypes.c_int32))
    d.n_bands, d.n_exp, d.n_sources, d.n_radii = 1, 1, 1, 1
    d.psfgauss = ctypes.c_void_p(arrs["psfgauss"].ctypes.data)
    d.n_psfgauss = len(arrs["psfgauss"])
    return d


@pytest.mark.parametrize("case,code", [
    ("nan_pixel", 5), ("negative_ierr", 5), ("exposure_overrun", 1), ("psf_range", 1),
    ("psf_not_pd", 5), ("too_many_sources", 2), ("bad_bin", 1)])
def test_validation_errors_before_device_work(case, code):
    keep = {}
    if case == "nan_pixel":
        data = np.zeros(64, np.float32)
        data[5] = np.nan
        d = _tiny_desc(keep, data=data)
    elif case == "negative_ierr":
        ie = np.ones(64, np.float32)
        ie[3] = -1
        d = _tiny_desc(keep, ierr=ie)
    elif case == "exposure_overrun":
        d = _tiny_desc(keep, exposure_N=np.full(1, 65, np.int32))
    elif case == "psf_range":
        d = _tiny_desc(keep, psfgauss_start=np.full(1, 1, np.int32))
    elif case == "psf_not_pd":
        psf = np.zeros(1, dtype=limits.psf_struct_dtype)
        psf["Cxx"], psf["Cyy"], psf["Cxy"] = 1.0, 1.0, 2.0
        d = _tiny_desc(keep, psfgauss=psf)
    elif case == "too_many_sources":
        d = _tiny_desc(keep)
        d.n_sources = limits.MAX_NATIVE_SOURCES + 1
    elif case == "bad_bin":
        psf = np.zeros(1, dtype=limits.psf_struct_dtype)
        psf["Cxx"], psf["Cyy"], psf["sersic_bin"] = 1.0, 1.0, 3
        d = _tiny_desc(keep, psfgauss=psf)
    h = ctypes.c_void_p()
    rc = _lib.lib().fpho_patch_create(ctypes.byref(d), 0, ctypes.byref(h))
    assert rc == code, _lib.lib().fpho_last_error()
    assert not h.value
    assert len(_lib.lib().fpho_last_error()) > 0


def test_launch_validation():
    l = _lib.Launch(3, 4, 0, 8)
    rc = _lib.lib().fpho_patch_set_launch(ctypes.c_void_p(0), ctypes.byref(l))
    assert rc == 1  # null patch is rejected first


def test_one_hip_runtime_whatever_the_import_order():
    """torch's wheel carries its own HIP runtime under the system soname;
    loading the library before torch must not put a second runtime in the
    process (torch would then find no GPU)."""
    code = ("import sys; sys.path.insert(0, %r)\n"
            "from forcepho_amd import _lib; _lib.lib()\n"
            "import torch\n"
            "maps = open('/proc/self/maps').read().splitlines()\n"
            "print(len({l.split()[-1] for l in maps if 'libamdhip64' in l}))\n" % ROOT)
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True,
                         check=True).stdout.split()
    assert out[-1] == "1"


def test_new_entry_points_reject_null_arguments():
    """The compat Response output, residual-write switch and status words
    fail with FPHO_E_ARG on null handles, before any device work."""
    L = _lib.lib()
    z = ctypes.c_void_p(0)
    w = ctypes.c_uint32()
    assert L.fpho_evaluate_response(z, z, None, z) == 1
    assert L.fpho_evaluate_response_async(z, z, z, z, z) == 1
    assert L.fpho_patch_set_residual_writes(z, 1) == 1
    assert L.fpho_patch_status(z, ctypes.byref(w), 0) == 1
    assert L.fpho_batch_status(z, ctypes.byref(w), 0) == 1
    assert _lib.describe_status(0x12) == ["axis ratio q <= 0",
                                          "non-finite parameter vector entry"]


def test_measurement_library_is_separate():
    """Issue calibration and the in-cutoff pair count live in the bench-only
    measure/libfpho_measure.so, not in the product library or its header."""
    import measure
    m = measure.lib()
    for name in ("fpho_measure_issue_calibration", "fpho_measure_batch_count_pairs",
                 "fpho_measure_last_error"):
        assert hasattr(m, name)
    lib = _lib.lib()
    for name in ("fpho_issue_calibration", "fpho_batch_count_pairs"):
        assert not hasattr(lib, name)
    assert "calibrat" not in open(HEADER).read()


def test_c_host_compiles_links_and_validates_its_input(tmp_path):
    """The plain C host of the library (tests/c_host/evaluate_patch.c, used
    on the GPU by test_gpu_c_host.py) compiles against include/fpho.h with
    warnings as errors, resolves every library symbol it calls at load time
    (LD_BIND_NOW), and rejects a bad command line or input directory before
    it touches the device."""
    lib = os.path.dirname(_lib.LIBPATH)
    exe = tmp_path / "evaluate_patch"
    subprocess.run(["gcc", "-std=c99", "-O2", "-Wall", "-Wextra", "-Werror", "-I",
                    os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "c_host", "evaluate_patch.c"),
                    "-L", lib, "-lfpho_hip", "-Wl,-rpath," + lib, "-o", str(exe)], check=True)
    env = dict(os.environ, LD_BIND_NOW="1")
    usage = subprocess.run([str(exe)], capture_output=True, text=True, env=env)
    assert usage.returncode == 2 and "usage" in usage.stderr
    missing = subprocess.run([str(exe), str(tmp_path / "absent")], capture_output=True,
                             text=True, env=env)
    assert missing.returncode == 2 and "meta.txt" in missing.stderr


def test_precision_modes_match_the_header():
    """forcepho_amd._lib.PRECISION names every FPHO_PREC_* mode of
    include/fpho.h with its value; setting one on a null patch is an
    argument error before any device work."""
    text = open(HEADER).read()
    modes = {m.group(1).lower(): int(m.group(2))
             for m in re.finditer(r"#define FPHO_PREC_([A-Z0-9_]+) (\d+)", text)}
    assert modes == _lib.PRECISION
    z = ctypes.c_void_p(0)
    assert _lib.lib().fpho_patch_set_precision(z, 0) == 1   # null patch: FPHO_E_ARG
