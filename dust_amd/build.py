"""Build libdust_gpu.so for gfx950 in-tree with hipcc.

Every HIP translation unit under dust_amd/csrc is compiled to an object in
dust_amd/_build and linked into dust_amd/libdust_gpu.so.  Compiles run in
parallel; the build always recompiles (it is also the "does it build" check).
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libdust_gpu.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLANG = os.environ.get("FLANG", "/opt/rocm/llvm/bin/amdflang")
ROOT = os.path.dirname(HERE)
# Fortran binding (ISO_C_BINDING module), the modules with the reference's
# procedure names on top of it, and the Fortran test drivers
FORTRAN_MOD = os.path.join(HERE, "fortran", "mod_dust_gpu.f90")
FORTRAN_SHIMS = [os.path.join(HERE, "fortran", f) for f in ("mod_octree_gpu.f90",
                                                          "mod_linsys_gpu.f90")]
FORTRAN_DRIVER_SRC = os.path.join(ROOT, "tests", "fortran", "dust_gpu_driver.f90")
FORTRAN_DRIVER = os.path.join(ROOT, "tests", "fortran", "dust_gpu_driver")
SHIM_DRIVER_SRC = os.path.join(ROOT, "tests", "fortran", "dust_shim_driver.f90")
SHIM_DRIVER = os.path.join(ROOT, "tests", "fortran", "dust_shim_driver")
ARCH = "gfx950"
# HDF5 result files (libdust_h5.so, host code): the HDF5 C library of the image
HDF5_ROOT = os.environ.get("HDF5_ROOT", "/opt/conda")
H5_SRC = os.path.join(CSRC, "h5io.cpp")
H5_LIB = os.path.join(HERE, "libdust_h5.so")

SOURCES = [
    "fmm.hip",
    "fmm_prim.hip",
    "fmm_graph.hip",
    "fmm_tuned.hip",
    "aic.hip",
    "elem.hip",
    "wake.hip",
    "panwake.hip",
    "lu.hip",
    "lu_mixed.hip",
    "linsys.hip",
    "linsys_asm.hip",
    "pressure.hip",
    "wind.hip",
]

CFLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    # packed FP32 instructions are kept out of the generated code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def _compile(src):
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC] + CFLAGS + ["-c", os.path.join(CSRC, src), "-o", obj]
    p = subprocess.run(cmd, capture_output=True, text=True)
    out = "\n".join(l for l in (p.stdout + p.stderr).splitlines() if "packed-fp32-ops" not in l)
    return src, p.returncode, out, obj


def build(verbose=False, jobs=None):
    os.makedirs(BUILD, exist_ok=True)
    jobs = jobs or min(len(SOURCES), os.cpu_count() or 4, 8)
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        results = list(ex.map(_compile, SOURCES))
    objs = []
    for src, rc, out, obj in results:
        if rc != 0:
            raise RuntimeError(f"hipcc failed on {src}:\n{out}")
        if verbose and out.strip():
            print(out)
        objs.append(obj)
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError("link failed:\n" + p.stdout + p.stderr)
    os.replace(tmp, LIB)
    return LIB


def build_fortran():
    """Compile the Fortran binding module and the reference-named modules on
    it, and link the two Fortran test drivers against libdust_gpu.so (rpath
    relative to the drivers).  Returns the first driver's path, or None when
    no Fortran compiler is installed."""
    if not os.path.exists(FLANG):
        return None
    fdir = os.path.join(BUILD, "fortran")
    os.makedirs(fdir, exist_ok=True)

    def obj(src):
        return os.path.join(fdir, os.path.splitext(os.path.basename(src))[0] + ".o")

    mods = [FORTRAN_MOD] + FORTRAN_SHIMS
    steps = [[FLANG, "-O2", "-J", fdir, "-c", m, "-o", obj(m)] for m in mods]
    for src, exe, deps in ((FORTRAN_DRIVER_SRC, FORTRAN_DRIVER, [FORTRAN_MOD]),
                           (SHIM_DRIVER_SRC, SHIM_DRIVER, mods)):
        steps.append([FLANG, "-O2", "-J", fdir, "-c", src, "-o", obj(src)])
        steps.append([FLANG, obj(src)] + [obj(m) for m in deps] +
                     ["-L" + HERE, "-ldust_gpu", "-Wl,-rpath,$ORIGIN/../../dust_amd",
                      "-o", exe + ".tmp"])
    for cmd in steps:
        p = subprocess.run(cmd, capture_output=True, text=True)
        if p.returncode != 0:
            raise RuntimeError("flang failed: " + " ".join(cmd) + "\n" + p.stdout + p.stderr)
    for exe in (FORTRAN_DRIVER, SHIM_DRIVER):
        os.replace(exe + ".tmp", exe)
    return FORTRAN_DRIVER


def build_h5():
    """Compile libdust_h5.so (host C++ against the HDF5 C library, rpath to
    it).  Returns its path, or None when no HDF5 installation is found."""
    inc = os.path.join(HDF5_ROOT, "include")
    libdir = os.path.join(HDF5_ROOT, "lib")
    if not os.path.exists(os.path.join(inc, "hdf5.h")):
        return None
    tmp = H5_LIB + ".tmp"
    cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall", "-I" + inc, H5_SRC,
           "-L" + libdir, "-lhdf5", "-Wl,-rpath," + libdir, "-o", tmp]
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError("g++ failed on h5io.cpp:\n" + p.stdout + p.stderr)
    os.replace(tmp, H5_LIB)
    return H5_LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
    print(build_fortran())
    print(build_h5())
