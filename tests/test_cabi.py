"""CPU checks of the drop-in boundary: the C ABI library loads and exports
every symbol include/lbswim.h declares, arguments are validated before any
device work, the package imports, and the host-side constants and index
arithmetic the kernels use are right.  No compute call needs a GPU here."""
import ctypes
import json
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "lbswim.h")
LIB = os.path.join(ROOT, "gpu-swimmers_amd", "lib", "liblbswim_amd.so")
CSRC = os.path.join(ROOT, "gpu-swimmers_amd", "csrc")
GOLD = json.load(open(os.path.join(ROOT, "tests", "golden", "d3q15_tables.json")))["tables"]


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(lbswim_\w+)\s*\(", text)))


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(LIB):
        pytest.fail("liblbswim_amd.so is not built (run gpu-swimmers_amd/build_native.py)")
    return ctypes.CDLL(LIB)


def test_header_declares_the_reference_surface():
    names = declared_functions()
    for must in ("lbswim_lattice_create", "lbswim_lattice_step", "lbswim_lattice_calc_hydro",
                 "lbswim_lattice_init_from_hydro", "lbswim_lattice_zero_force", "lbswim_lattice_h2d",
                 "lbswim_lattice_d2h", "lbswim_swimmers_create", "lbswim_swimmers_add_forces",
                 "lbswim_swimmers_move", "lbswim_tracers_create", "lbswim_tracers_move", "lbswim_system_step"):
        assert must in names


def test_library_exports_every_declared_symbol(lib):
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_invalid_arguments_rejected_without_device(lib):
    lib.lbswim_last_error.restype = ctypes.c_char_p
    lib.lbswim_lattice_create.argtypes = [ctypes.c_int] * 3 + [ctypes.c_double] * 2 + [ctypes.c_int,
                                                                                      ctypes.c_void_p]
    out = ctypes.c_void_p()
    EINVAL = 1
    assert lib.lbswim_lattice_create(0, 4, 4, 0.25, 0.25, 0, ctypes.byref(out)) == EINVAL
    assert b"sizes" in lib.lbswim_last_error()
    assert lib.lbswim_lattice_create(4, 4, 4, -1.0, 0.25, 0, ctypes.byref(out)) == EINVAL
    assert lib.lbswim_lattice_create(4, 4, 4, 0.25, 0.25, 7, ctypes.byref(out)) == EINVAL
    assert lib.lbswim_lattice_create(2048, 2048, 1024, 0.25, 0.25, 0, ctypes.byref(out)) == EINVAL
    assert out.value is None
    assert lib.lbswim_lattice_step(None) == EINVAL
    assert lib.lbswim_lattice_sync_u(None) == EINVAL
    assert lib.lbswim_lattice_u_is_virtual(None, None) == EINVAL
    lib.lbswim_swimmers_create.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
    assert lib.lbswim_swimmers_create(-1, None, ctypes.byref(out)) == EINVAL
    lib.lbswim_tracers_create.argtypes = [ctypes.c_int, ctypes.c_void_p]
    assert lib.lbswim_tracers_create(-5, ctypes.byref(out)) == EINVAL


def test_package_imports_and_validates():
    import lbswim
    from lbswim import lb, swimmers, tracers, shared  # noqa: F401
    from lbswim.System import System  # noqa: F401

    with pytest.raises(ValueError):
        lb.Lattice(4, 4, 4, 0.25, 0.25, storage="f16")
    with pytest.raises(ValueError):
        lb.Lattice(-4, 4, 4, 0.25, 0.25)
    cp = swimmers.CommonParams()
    cp.P, cp.l, cp.alpha, cp.seed = 1e-3, 0.1, 0.01, 1234
    assert (cp.P, cp.l, cp.alpha, cp.seed) == (1e-3, 0.1, 0.01, 1234)
    import pickle
    cp2 = pickle.loads(pickle.dumps(cp))
    assert cp2.as_dict() == cp.as_dict()
    assert lbswim.LbswimError.__mro__[1] is RuntimeError


HOST_PROBE = r"""
#include <cstdio>
#include <cstdlib>
#include "d3q15.hpp"
using namespace lbs;
int main() {
  // exhaustive small divisors, sampled large numerators
  unsigned long long bad = 0;
  for (uint32_t d = 1; d <= 1200; ++d) {
    const FastDiv f = make_fastdiv(d);
    for (uint32_t x = 0; x < 70000; ++x) bad += fastdiv(x, f) != x / d;
    for (uint64_t x = 0; x < (1ull << 32); x += 0x10001ull * 977) bad += fastdiv(uint32_t(x), f) != uint32_t(x) / d;
    bad += fastdiv(0xFFFFFFFFu, f) != 0xFFFFFFFFu / d;
  }
  printf("bad %llu\n", bad);
  for (int m = 0; m < Q; ++m) for (int p = 0; p < Q; ++p) printf("mm %d %d %a\n", m, p, TAB.mm[m][p]);
  for (int p = 0; p < Q; ++p) for (int m = 0; m < Q; ++m) printf("mmi %d %d %a\n", p, m, TAB.mmi[p][m]);
  for (int p = 0; p < Q; ++p) printf("w %d %a\n", p, TAB.w[p]);
  for (int m = 0; m < Q; ++m) printf("norms %d %a\n", m, TAB.norms[m]);
  return 0;
}
"""


def test_kernel_constants_and_fastdiv_on_host(tmp_path):
    """The compile-time tables the kernels unroll against equal the reference's
    d3q15.c bitwise, and the site-index division is exact for all inputs tried."""
    src = tmp_path / "probe.cpp"
    src.write_text(HOST_PROBE)
    exe = tmp_path / "probe"
    subprocess.check_call(["/opt/rocm/bin/hipcc", "-std=c++20", "-O2", "-I" + CSRC, "-o", str(exe), str(src)],
                          stderr=subprocess.DEVNULL)
    out = subprocess.check_output([str(exe)], text=True).splitlines()
    assert out[0] == "bad 0"
    got = {"mm": np.zeros((15, 15)), "mmi": np.zeros((15, 15)), "w": np.zeros(15), "norms": np.zeros(15)}
    for line in out[1:]:
        parts = line.split()
        idx = tuple(int(i) for i in parts[1:-1])
        got[parts[0]][idx] = float.fromhex(parts[-1])
    for key, arr in got.items():
        np.testing.assert_array_equal(arr, np.array(GOLD[key]), err_msg=key)
