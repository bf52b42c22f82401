"""Build the gfx950 shared library ``libfpho_hip.so`` in-tree with hipcc.

The library is written next to this file so that it travels with the source
tree; nothing is installed.  Packed FP32 instructions are disabled for all
device code (``-packed-fp32-ops``).  ``build_measure_library`` builds the
measurement-only ``measure/libfpho_measure.so`` that bench.py uses (issue
calibration, in-cutoff pair count); the product never loads it.
"""

import os
import shutil
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIBNAME = "libfpho_hip.so"
LIBPATH = os.path.join(HERE, LIBNAME)
SOURCES = ["kernels.hip", "runtime.hip"]
ARCH = "gfx950"


def hipcc_path():
    for cand in (os.environ.get("HIPCC"), shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found; the HIP library cannot be built")


def compile_flags():
    return ["-O3", "-std=c++17", f"--offload-arch={ARCH}",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-fPIC", "-Wall", "-Wno-unused-function"]


MEASURE_DIR = os.path.join(os.path.dirname(HERE), "measure")
MEASURE_LIBPATH = os.path.join(MEASURE_DIR, "libfpho_measure.so")


def _hipcc(cmd, out, verbose):
    tmp = out + ".tmp"
    cmd = cmd + ["-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    proc = subprocess.run(cmd, capture_output=True, text=True)
    # the host pass warns that it does not know the packed-fp32 feature
    msgs = [l for l in (proc.stdout + proc.stderr).splitlines()
            if "packed-fp32-ops" not in l and l.strip()]
    if proc.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + "\n".join(msgs))
    if verbose and msgs:
        print("\n".join(msgs))
    os.replace(tmp, out)
    return out


def build_library(verbose=False):
    """Compile every HIP source into ``libfpho_hip.so`` (always recompiles)."""
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    return _hipcc([hipcc_path()] + compile_flags() +
                  ["-shared", "-Wl,-soname,libfpho_hip.so"] + srcs, LIBPATH, verbose)


def build_measure_library(verbose=False):
    """Compile the measurement-only library against ``libfpho_hip.so``."""
    return _hipcc([hipcc_path()] + compile_flags() +
                  ["-shared", os.path.join(MEASURE_DIR, "measure.hip"), "-L" + HERE,
                   "-lfpho_hip", "-Wl,-rpath,$ORIGIN/../forcepho_amd"],
                  MEASURE_LIBPATH, verbose)


if __name__ == "__main__":
    print(build_library(verbose=True))
    print(build_measure_library(verbose=True))
