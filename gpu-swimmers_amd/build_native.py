"""Build the native parts of lbswim for gfx950.

  lib/liblbswim_amd.so        HIP kernels + C ABI (include/lbswim.h), hipcc
  lbswim/_core.*.so           Cython bindings of the C ABI (the lbswim package)

Usage:  python gpu-swimmers_amd/build_native.py [--no-cython] [-j N]
Compiles every time it is called (no staleness check); objects go to _build/.
"""
import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
import sysconfig

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
INCLUDE = os.path.join(ROOT, "include")
BUILD = os.path.join(PKG, "_build")
LIBDIR = os.path.join(PKG, "lib")
LIB = os.path.join(LIBDIR, "liblbswim_amd.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

SOURCES = ["lattice_kernels.hip", "particle_kernels.hip", "particle_order.hip", "noalloc.hip", "capi.cpp"]

# -ffp-contract=off: the FP64 path reproduces the reference's unfused
#   arithmetic; FP32 kernels ask for fusion explicitly with fmaf.
# -packed-fp32-ops off: no v_pk_* FP32 instructions are generated.
HIP_FLAGS = [
    "-std=c++20", "-O3", "-fPIC", f"--offload-arch={ARCH}", "-ffp-contract=off",
    "-munsafe-fp-atomics", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-I" + INCLUDE, "-I" + CSRC,
]


def _run(cmd):
    res = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if res.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + res.stdout)
        raise SystemExit("build failed: %s" % os.path.basename(cmd[-1]))
    return res.stdout


def build_hip(jobs=4):
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    objs = []
    cmds = []
    for src in SOURCES:
        obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
        lang = ["-x", "hip"] if src.endswith(".hip") else []
        cmds.append([HIPCC] + HIP_FLAGS + ["-c", "-o", obj] + lang + [os.path.join(CSRC, src)])
        objs.append(obj)
    with cf.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        list(ex.map(_run, cmds))
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    return LIB


def build_cython():
    from Cython.Build import cythonize  # noqa: F401 (availability check)
    import numpy as np

    pyx = os.path.join(PKG, "lbswim", "_core.pyx")
    cpp = os.path.join(BUILD, "_core.cpp")
    os.makedirs(BUILD, exist_ok=True)
    _run([sys.executable, "-m", "cython", "-3", "--cplus", "-I", INCLUDE, "-o", cpp, pyx])
    suffix = sysconfig.get_config_var("EXT_SUFFIX")
    out = os.path.join(PKG, "lbswim", "_core" + suffix)
    tmp = out + ".tmp"
    _run(["g++", "-O2", "-fPIC", "-shared", "-std=c++17", "-w",
          "-I" + sysconfig.get_paths()["include"], "-I" + np.get_include(), "-I" + INCLUDE,
          "-DNPY_NO_DEPRECATED_API=NPY_1_7_API_VERSION",
          "-o", tmp, cpp, "-L" + LIBDIR, "-llbswim_amd", "-Wl,-rpath,$ORIGIN/../lib"])
    os.replace(tmp, out)
    return out


EXAMPLE = os.path.join(LIBDIR, "lbswim_readme_coupled")
DEVPTR = os.path.join(LIBDIR, "lbswim_device_pointers")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")


def build_example():
    """The C++ API programs used by the GPU tests (tests/cpp/): the README
    example, and the device-pointer check, which also reads device memory with
    the HIP runtime."""
    src = os.path.join(ROOT, "tests", "cpp", "readme_coupled.cpp")
    tmp = EXAMPLE + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-Wall", "-I" + INCLUDE, "-o", tmp, src, "-L" + LIBDIR, "-llbswim_amd",
          "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, EXAMPLE)
    src = os.path.join(ROOT, "tests", "cpp", "device_pointers.cpp")
    tmp = DEVPTR + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-Wall", "-D__HIP_PLATFORM_AMD__", "-I" + INCLUDE,
          "-I" + os.path.join(ROCM, "include"), "-o", tmp, src, "-L" + LIBDIR, "-llbswim_amd",
          "-L" + os.path.join(ROCM, "lib"), "-lamdhip64", "-Wl,-rpath,$ORIGIN",
          "-Wl,-rpath," + os.path.join(ROCM, "lib")])
    os.replace(tmp, DEVPTR)
    return EXAMPLE


def build_all(cython=True, jobs=4):
    lib = build_hip(jobs)
    outs = [lib, build_example(), DEVPTR]
    if cython:
        outs.append(build_cython())
    return outs


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--no-cython", action="store_true")
    ap.add_argument("-j", type=int, default=4)
    a = ap.parse_args()
    for o in build_all(cython=not a.no_cython, jobs=a.j):
        print("built", os.path.relpath(o, ROOT))
