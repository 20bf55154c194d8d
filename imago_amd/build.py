"""Builds the HIP library (libimago_hip.so) in-tree for gfx950.

Each translation unit under csrc/ is compiled to an object in parallel,
then linked into imago_amd/libimago_hip.so against rocSOLVER/rocBLAS.
The integral file library (libimago_h5.so: host code over the HDF5 C
library of the image's conda tree) is built beside it when HDF5 is there.

Usage: python -m imago_amd.build [-v]   (or __graft_entry__.build())
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libimago_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    # packed FP32 instructions stay off on these parts
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", "-Wno-pass-failed",
]

SOURCES = ["imago_hip.hip", "prim_batch.hip", "asm_ol.hip", "asm_ke.hip", "asm_np.hip",
           "asm_ep.hip", "asm_epd.hip", "hetrd.hip", "elecstat.hip",
           "xcmesh.hip", "learned.hip", "tridiag_dc.hip", "dos.hip", "core_ortho.hip",
           "gemm_mfma.hip"]


# The MFMA kernels keep their accumulators in ordinary VGPRs: with the
# AGPR form the compiler copied every accumulator out of and back into the
# AGPRs on each K step, draining the matrix pipeline each time.
EXTRA_FLAGS = {
    "gemm_mfma.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form=1"],
    "tridiag_dc.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form=1"],
}


def _compile(src: str, verbose: bool) -> str:
    obj = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC] + HIP_FLAGS + EXTRA_FLAGS.get(src, []) + [
        "-I", os.path.join(ROOT, "include"), "-c", os.path.join(CSRC, src), "-o", obj]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stderr)
        raise subprocess.CalledProcessError(r.returncode, cmd)
    return obj


HDF5_ROOT = os.environ.get("IMAGO_HDF5_ROOT", "/opt/conda")
H5LIB = os.path.join(HERE, "libimago_h5.so")


def build_h5(verbose: bool = False):
    """libimago_h5.so from csrc/h5store.cpp, or None without HDF5."""
    if not os.path.exists(os.path.join(HDF5_ROOT, "include", "hdf5.h")):
        return None
    lib = os.path.join(HDF5_ROOT, "lib")
    cmd = ["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-I", os.path.join(ROOT, "include"),
           "-I", os.path.join(HDF5_ROOT, "include"), os.path.join(CSRC, "h5store.cpp"),
           "-o", H5LIB + ".tmp", "-L" + lib, "-lhdf5", "-Wl,-rpath," + lib]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(H5LIB + ".tmp", H5LIB)
    return H5LIB


FORTRAN = os.path.join(HERE, "fortran")
FBIN = os.path.join(FORTRAN, "_bin")
FLANG = "/opt/rocm/bin/amdflang"


def build_fortran(verbose: bool = False) -> str:
    """Compile the iso_c_binding module and the Fortran hosts that link
    against libimago_hip.so (imago_amd/fortran/_bin/hip_driver: a solve;
    scf_driver: the integrals and the SCF loop)."""
    os.makedirs(FBIN, exist_ok=True)
    mod = [FLANG, "-c", "-J" + FBIN, os.path.join(FORTRAN, "imago_hip.f90"), "-o",
           os.path.join(FBIN, "imago_hip.o")]
    drvs = [[FLANG, "-J" + FBIN, os.path.join(FORTRAN, name + ".f90"),
             os.path.join(FBIN, "imago_hip.o"), "-L" + HERE, "-limago_hip",
             "-Wl,-rpath,$ORIGIN/../..", "-o", os.path.join(FBIN, name)]
            for name in ("hip_driver", "scf_driver")]
    for cmd in [mod] + drvs:
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
    return os.path.join(FBIN, "hip_driver")


def build(verbose: bool = False) -> str:
    os.makedirs(OBJ, exist_ok=True)
    jobs = min(len(SOURCES), max(1, (os.cpu_count() or 2)))
    with cf.ThreadPoolExecutor(jobs) as ex:
        h5 = ex.submit(build_h5, verbose)
        objs = list(ex.map(lambda s: _compile(s, verbose), SOURCES))
        h5.result()
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC"] + objs + [
        "-o", LIB + ".tmp", "-L/opt/rocm/lib", "-lrocsolver", "-lrocblas",
        "-Wl,-rpath,/opt/rocm/lib"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(LIB + ".tmp", LIB)
    if os.path.exists(FLANG):
        build_fortran(verbose)
    return LIB


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print(LIB)
