"""The C++ API (include/lbswim.hpp) on the README scenario against the oracle.

The program tests/cpp/readme_coupled.cpp is built in-tree by the build step
(gpu-swimmers_amd/build_native.py) and drives the lattice and particle
classes through their individual methods, like the reference's C++ users.
"""
import os
import subprocess
import sys

import numpy as np
import pytest

import lbswim_oracle as O
from cases import config1, run_oracle

pytestmark = pytest.mark.gpu

EXE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "gpu-swimmers_amd", "lib",
                   "lbswim_readme_coupled")


@pytest.mark.parametrize("alpha", [0.0, 0.01])
def test_cpp_readme_example_matches_oracle(tmp_path, alpha):
    assert os.path.exists(EXE), "C++ example not built (run gpu-swimmers_amd/build_native.py)"
    out = tmp_path / "state.bin"
    print("C++ README example, alpha=%g" % alpha, flush=True)
    sys.stdout.flush()
    res = subprocess.run([EXE, str(out), "1000", repr(alpha)], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stderr
    assert "done 1000 steps" in res.stdout
    n = 16 ** 3
    raw = np.fromfile(out, dtype=np.float64)
    assert raw.size == 15 * n + n + 3 * n + 3 + 3 + 3 * n
    cuts = np.cumsum([15 * n, n, 3 * n, 3, 3])
    fOld, rho, u, sw_r, sw_n, tr_r = np.split(raw, cuts)
    ref, _ = run_oracle(O, config1(1000, alpha=alpha))
    assert np.array_equal(fOld, ref["fOld"])
    assert np.array_equal(rho, ref["rho"])
    assert np.array_equal(u, ref["u"])
    assert np.array_equal(sw_r, ref["sw_r"].ravel())
    assert np.array_equal(sw_n, ref["sw_n"].ravel())
    assert np.array_equal(tr_r, ref["tr_r"].ravel())


def test_cpp_device_pointers_typed_and_fixed():
    """lat.data->u.device read as double*, particle device pointers fixed across
    the storage re-ordering after 64 moves, contents equal to D2H through
    slots(); FP32 lattices expose their float buffer as device_raw only."""
    exe = os.path.join(os.path.dirname(EXE), "lbswim_device_pointers")
    assert os.path.exists(exe), "device-pointer program not built (run gpu-swimmers_amd/build_native.py)"
    print("C++ device pointers, 70 steps", flush=True)
    res = subprocess.run([exe, "70"], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stdout + res.stderr
    assert res.stdout.strip().endswith("ok"), res.stdout
