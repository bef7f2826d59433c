"""The C ABI driven directly, as a binding in another language would drive it.

After coupled steps, fused or in two launches, u lives only in the particle
moves' interleaved copy (lbswim_lattice_u_is_virtual); a device pointer to u
obtained from lbswim_lattice_field must then hold u in the reference layout --
the values lbswim_lattice_d2h returns, the same from both forms.
"""
import ctypes
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "gpu-swimmers_amd", "lib", "liblbswim_amd.so")
LBSWIM_RHO, LBSWIM_U = 0, 1
LBSWIM_PART_R = 0
HIP_MEMCPY_DEVICE_TO_HOST = 2

P = ctypes.c_void_p


@pytest.fixture(scope="module")
def lib():
    lib = ctypes.CDLL(LIB)
    lib.lbswim_lattice_create.argtypes = [ctypes.c_int] * 3 + [ctypes.c_double] * 2 + [ctypes.c_int, P]
    lib.lbswim_lattice_field.argtypes = [P, ctypes.c_int, P, P, P]
    lib.lbswim_tracers_field.argtypes = [P, ctypes.c_int, P, P, P]
    lib.lbswim_system_step.argtypes = [P, P, ctypes.c_int, P, ctypes.c_int, ctypes.c_int]
    lib.lbswim_last_error.restype = ctypes.c_char_p
    return lib


def ok(lib, rc):
    assert rc == 0, lib.lbswim_last_error()


def host_view(ptr, count):
    return np.ctypeslib.as_array(ctypes.cast(ptr, ctypes.POINTER(ctypes.c_double)), shape=(count,))


def run(lib, fusion, nx=4, ny=8, nz=64, nt=300, steps=3):
    n = nx * ny * nz
    rng = np.random.default_rng(3)
    lat = P()
    ok(lib, lib.lbswim_lattice_create(nx, ny, nz, 0.3, 0.3, 0, ctypes.byref(lat)))
    host, size = P(), ctypes.c_size_t()
    for field, values in ((LBSWIM_RHO, 1 + 1e-2 * rng.standard_normal(n)), (LBSWIM_U, 1e-2 * rng.standard_normal(3 * n))):
        ok(lib, lib.lbswim_lattice_field(lat, field, ctypes.byref(host), None, ctypes.byref(size)))
        host_view(host, size.value)[:] = values
        ok(lib, lib.lbswim_lattice_h2d(lat, field))
    ok(lib, lib.lbswim_lattice_init_from_hydro(lat))
    ok(lib, lib.lbswim_lattice_set_fusion(lat, int(fusion), 0, 0))
    tr = P()
    ok(lib, lib.lbswim_tracers_create(nt, ctypes.byref(tr)))
    count = ctypes.c_size_t()
    ok(lib, lib.lbswim_tracers_field(tr, LBSWIM_PART_R, ctypes.byref(host), None, ctypes.byref(count)))
    host_view(host, count.value)[:] = rng.uniform(0, 1, 3 * nt) * np.repeat([nx, ny, nz], nt)
    ok(lib, lib.lbswim_tracers_h2d(tr, LBSWIM_PART_R))
    trs = (P * 1)(tr)
    ok(lib, lib.lbswim_system_step(lat, None, 0, trs, 1, steps))
    virt_before = ctypes.c_int(-1)
    ok(lib, lib.lbswim_lattice_u_is_virtual(lat, ctypes.byref(virt_before)))
    dev = P()
    ok(lib, lib.lbswim_lattice_field(lat, LBSWIM_U, None, ctypes.byref(dev), ctypes.byref(size)))
    virt_after = ctypes.c_int(-1)
    ok(lib, lib.lbswim_lattice_u_is_virtual(lat, ctypes.byref(virt_after)))
    ok(lib, lib.lbswim_synchronize())
    from_device = np.empty(3 * n)
    hip = ctypes.CDLL("libamdhip64.so")
    assert hip.hipMemcpy(P(from_device.ctypes.data), dev, ctypes.c_size_t(from_device.nbytes),
                         HIP_MEMCPY_DEVICE_TO_HOST) == 0
    ok(lib, lib.lbswim_lattice_d2h(lat, LBSWIM_U))
    ok(lib, lib.lbswim_lattice_field(lat, LBSWIM_U, ctypes.byref(host), None, ctypes.byref(size)))
    from_d2h = host_view(host, size.value).copy()
    ok(lib, lib.lbswim_tracers_d2h(tr, LBSWIM_PART_R))
    ok(lib, lib.lbswim_tracers_field(tr, LBSWIM_PART_R, ctypes.byref(host), None, ctypes.byref(count)))
    tr_r = host_view(host, count.value).copy()
    lib.lbswim_tracers_destroy(tr)
    lib.lbswim_lattice_destroy(lat)
    return virt_before.value, virt_after.value, from_device, from_d2h, tr_r


def test_device_pointer_to_u_after_fused_steps(lib):
    print("C ABI: 4x8x64, 300 tracers, 3 coupled steps, fused and two-launch", flush=True)
    fused = run(lib, True)
    plain = run(lib, False)
    # both forms of the coupled step write u only interleaved: virtual until
    # the device pointer was asked for
    assert fused[:2] == (1, 0)
    assert plain[:2] == (1, 0)
    assert np.array_equal(fused[2], fused[3])
    assert np.array_equal(fused[2], plain[2])
    assert np.array_equal(fused[4], plain[4])
