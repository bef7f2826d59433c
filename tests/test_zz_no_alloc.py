"""lbswim_system_step_noalloc (the coupled step for callers that cannot
allocate) against lbswim_system_step, through the C ABI.

Two systems built from identical inputs are stepped, one by each entry point;
every lattice and particle field must agree bit for bit (the FP64 form keeps
the reference's arithmetic, the FP32 form is deterministic too).  Each pair of
calls runs a system, then a smaller one, so the second call reuses the fixed
workspace the first left filled.  A system whose lists exceed the workspace is
refused before anything is enqueued.
"""
import ctypes
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "gpu-swimmers_amd", "lib", "liblbswim_amd.so")
P = ctypes.c_void_p
RHO, U, FORCE, FOLD = 0, 1, 2, 3
PART_R, PART_V, PART_N, PART_S, PART_PRNG = 0, 1, 2, 3, 4
EINVAL = 1


class Common(ctypes.Structure):
    _fields_ = [("P", ctypes.c_double), ("l", ctypes.c_double), ("alpha", ctypes.c_double),
                ("mobility", ctypes.c_double), ("translational_advection_off", ctypes.c_ubyte),
                ("rotational_advection_off", ctypes.c_ubyte), ("seed", ctypes.c_ulonglong)]


@pytest.fixture(scope="module")
def lib():
    lib = ctypes.CDLL(LIB)
    lib.lbswim_lattice_create.argtypes = [ctypes.c_int] * 3 + [ctypes.c_double] * 2 + [ctypes.c_int, P]
    lib.lbswim_lattice_field.argtypes = [P, ctypes.c_int, P, P, P]
    lib.lbswim_swimmers_create.argtypes = [ctypes.c_int, ctypes.POINTER(Common), P]
    lib.lbswim_swimmers_field.argtypes = [P, ctypes.c_int, P, P, P, P]
    lib.lbswim_tracers_field.argtypes = [P, ctypes.c_int, P, P, P]
    lib.lbswim_system_step.argtypes = [P, P, ctypes.c_int, P, ctypes.c_int, ctypes.c_int]
    lib.lbswim_system_step_noalloc.argtypes = [P, P, ctypes.c_int, P, ctypes.c_int, ctypes.c_int]
    lib.lbswim_last_error.restype = ctypes.c_char_p
    return lib


def ok(lib, rc):
    assert rc == 0, lib.lbswim_last_error()


def view(ptr, count, dtype=ctypes.c_double):
    return np.ctypeslib.as_array(ctypes.cast(ptr, ctypes.POINTER(dtype)), shape=(count,))


class System:
    """A lattice with one swimmer and one tracer array, built through the C ABI."""

    def __init__(self, lib, size, ns, nt, storage, seed):
        self.lib = lib
        nx, ny, nz = size
        n = nx * ny * nz
        rng = np.random.default_rng(seed)
        self.lat, self.sw, self.tr = P(), P(), P()
        ok(lib, lib.lbswim_lattice_create(nx, ny, nz, 0.3, 0.6, storage, ctypes.byref(self.lat)))
        host, cnt = P(), ctypes.c_size_t()
        for field, values in ((RHO, 1 + 1e-2 * rng.standard_normal(n)), (U, 1e-2 * rng.standard_normal(3 * n))):
            ok(lib, lib.lbswim_lattice_field(self.lat, field, ctypes.byref(host), None, ctypes.byref(cnt)))
            view(host, cnt.value)[:] = values
            ok(lib, lib.lbswim_lattice_h2d(self.lat, field))
        ok(lib, lib.lbswim_lattice_init_from_hydro(self.lat))
        cp = Common(1e-2, 0.5, 0.05, 2.0, 0, 0, seed)
        ok(lib, lib.lbswim_swimmers_create(ns, ctypes.byref(cp), ctypes.byref(self.sw)))
        L = np.repeat(np.array(size, float), ns)
        nv = rng.standard_normal((3, ns))
        for field, values in ((PART_R, rng.uniform(0, 1, 3 * ns) * L), (PART_N, (nv / np.linalg.norm(nv, axis=0)).ravel())):
            ok(lib, lib.lbswim_swimmers_field(self.sw, field, ctypes.byref(host), None, ctypes.byref(cnt), None))
            view(host, cnt.value)[:] = values
            ok(lib, lib.lbswim_swimmers_h2d(self.sw, field))
        ok(lib, lib.lbswim_tracers_create(nt, ctypes.byref(self.tr)))
        ok(lib, lib.lbswim_tracers_field(self.tr, PART_R, ctypes.byref(host), None, ctypes.byref(cnt)))
        view(host, cnt.value)[:] = rng.uniform(0, 1, 3 * nt) * np.repeat(np.array(size, float), nt)
        ok(lib, lib.lbswim_tracers_h2d(self.tr, PART_R))
        self.sws, self.trs = (P * 1)(self.sw), (P * 1)(self.tr)

    def step(self, steps, noalloc):
        fn = self.lib.lbswim_system_step_noalloc if noalloc else self.lib.lbswim_system_step
        return fn(self.lat, self.sws, 1, self.trs, 1, steps)

    def state(self):
        lib, out = self.lib, {}
        host, cnt = P(), ctypes.c_size_t()
        for name, field in (("fOld", FOLD), ("rho", RHO), ("u", U), ("force", FORCE)):
            ok(lib, lib.lbswim_lattice_d2h(self.lat, field))
            ok(lib, lib.lbswim_lattice_field(self.lat, field, ctypes.byref(host), None, ctypes.byref(cnt)))
            out[name] = view(host, cnt.value).copy()
        for name, field, dt in (("sw_r", PART_R, ctypes.c_double), ("sw_n", PART_N, ctypes.c_double),
                                ("sw_v", PART_V, ctypes.c_double), ("sw_prng", PART_PRNG, ctypes.c_uint64)):
            ok(lib, lib.lbswim_swimmers_d2h(self.sw, field))
            ok(lib, lib.lbswim_swimmers_field(self.sw, field, ctypes.byref(host), None, ctypes.byref(cnt), None))
            out[name] = view(host, cnt.value, dt).copy()
        for name, field in (("tr_r", PART_R), ("tr_s", PART_S), ("tr_v", PART_V)):
            ok(lib, lib.lbswim_tracers_d2h(self.tr, field))
            ok(lib, lib.lbswim_tracers_field(self.tr, field, ctypes.byref(host), None, ctypes.byref(cnt)))
            out[name] = view(host, cnt.value).copy()
        return out

    def close(self):
        self.lib.lbswim_tracers_destroy(self.tr)
        self.lib.lbswim_swimmers_destroy(self.sw)
        self.lib.lbswim_lattice_destroy(self.lat)


def compare(lib, size, ns, nt, storage, seed, steps, warm):
    """`warm` regular steps on both systems (the first allocates the fused
    step's interleaved u where the lattice takes it), then `steps` by each
    entry point."""
    print("noalloc vs system_step: %s ns=%d nt=%d storage=%d warm=%d steps=%d" % (size, ns, nt, storage, warm, steps),
          flush=True)
    a = System(lib, size, ns, nt, storage, seed)
    b = System(lib, size, ns, nt, storage, seed)
    if warm:
        ok(lib, a.step(warm, False))
        ok(lib, b.step(warm, False))
    ok(lib, a.step(steps, False))
    ok(lib, b.step(steps, True))
    sa, sb = a.state(), b.state()
    a.close()
    b.close()
    for k in sa:
        assert np.array_equal(sa[k], sb[k]), (size, storage, k)


@pytest.mark.parametrize("storage", [0, 1])
def test_noalloc_matches_system_step_twice(lib, storage):
    # a fused-step lattice (nz % 64 == 0), then a smaller two-launch one
    compare(lib, (24, 16, 64), 300, 600, storage, 11, steps=7, warm=1)
    compare(lib, (12, 10, 14), 40, 50, storage, 12, steps=9, warm=0)


def test_noalloc_refuses_lists_beyond_its_workspace(lib):
    print("noalloc: 60000 swimmers on 8^3 (beyond the workspace)", flush=True)
    s = System(lib, (8, 8, 8), 60000, 1, 0, 13)
    rc = s.step(1, True)
    msg = lib.lbswim_last_error().decode()
    ok(lib, lib.lbswim_synchronize())
    ts = ctypes.c_int(-1)
    ok(lib, lib.lbswim_lattice_get_time_step(s.lat, ctypes.byref(ts)))
    s.close()
    assert rc == EINVAL and "workspace" in msg, (rc, msg)
    assert ts.value == 0  # nothing was stepped
