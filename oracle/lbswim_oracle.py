"""ctypes front-end of the CPU oracle -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import
this module, and only as the checker or as the timed CPU baseline.  The
product (the HIP library and the lbswim package) never calls it.

Every function wraps one function of lbswim_oracle.c, which restates the
reference step by step (file:line citations live there).  Arrays use the
reference layout: lattice fields flat, element-slowest, (k, nx, ny, nz);
particle fields (3, N).
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liblbswim_oracle.so")

_D = ctypes.POINTER(ctypes.c_double)
_U64 = ctypes.POINTER(ctypes.c_uint64)
_U32 = ctypes.POINTER(ctypes.c_uint32)
_I = ctypes.c_int


class OrcParams(ctypes.Structure):
    """Mirror of the reference's LBParams (LBParams.h:4-30)."""

    _fields_ = [
        ("cs2", ctypes.c_double),
        ("w", ctypes.c_double * 15),
        ("xi", (ctypes.c_double * 3) * 15),
        ("ci", (ctypes.c_int * 3) * 15),
        ("Q", ((ctypes.c_double * 3) * 3) * 15),
        ("complement", ctypes.c_int * 15),
        ("norms", ctypes.c_double * 15),
        ("mm", (ctypes.c_double * 15) * 15),
        ("mmi", (ctypes.c_double * 15) * 15),
        ("tau_s", ctypes.c_double),
        ("tau_b", ctypes.c_double),
    ]


class OrcCommon(ctypes.Structure):
    """Mirror of the reference's CommonParams (SwimmerArray.h:8-16)."""

    _fields_ = [
        ("P", ctypes.c_double),
        ("l", ctypes.c_double),
        ("alpha", ctypes.c_double),
        ("mobility", ctypes.c_double),
        ("translational_advection_off", ctypes.c_ubyte),
        ("rotational_advection_off", ctypes.c_ubyte),
        ("seed", ctypes.c_ulonglong),
    ]


def build():
    """Compile liblbswim_oracle.so with oracle/Makefile."""
    subprocess.check_call(["make", "-s", "-C", HERE])


def build_native(path):
    """Compile the timing build (-march=native for the CPU that compiles it)
    to `path`; same arithmetic as the default build."""
    subprocess.check_call(["make", "-s", "-C", HERE, "native", "NATIVE_OUT=" + path])
    return path


def use_library(path):
    """Load the oracle from `path` instead of the default build (before any
    other call of this module)."""
    global LIB_PATH
    if _lib is not None and os.path.abspath(path) != os.path.abspath(LIB_PATH):
        raise RuntimeError("the oracle library is already loaded from %s" % LIB_PATH)
    LIB_PATH = path


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        L.orc_eigenset.argtypes = [ctypes.POINTER(OrcParams), ctypes.c_double, ctypes.c_double]
        L.orc_params_size.restype = ctypes.c_size_t
        L.orc_set_threads.argtypes = [_I]
        L.orc_get_max_threads.restype = _I
        L.orc_init_from_hydro.argtypes = [_I, _I, _I, _D, _D, _D]
        L.orc_zero_force.argtypes = [ctypes.c_long, _D]
        L.orc_step.argtypes = [_I, _I, _I, ctypes.c_double, ctypes.c_double, _D, _D, _D]
        L.orc_calc_hydro.argtypes = [_I, _I, _I, _D, _D, _D, _D]
        L.orc_peskin_delta.argtypes = [ctypes.c_double]
        L.orc_peskin_delta.restype = ctypes.c_double
        L.orc_interp.argtypes = [_I, _I, _I, _D, _D, _D]
        L.orc_swimmers_add_forces.argtypes = [_I, _I, _I, _I, ctypes.POINTER(OrcCommon), _D, _D, _D]
        L.orc_swimmers_move.argtypes = [_I, _I, _I, _I, ctypes.POINTER(OrcCommon), _D, _D, _D, _U64, _D]
        L.orc_swimmers_move.restype = ctypes.c_long
        L.orc_tracers_move.argtypes = [_I, _I, _I, _I, _D, _D, _D, _D]
        L.orc_philox4x32_10.argtypes = [_U32, _U32, _U32]
        L.orc_swimmer_uniform.argtypes = [ctypes.c_ulonglong, _I, ctypes.c_uint64]
        L.orc_swimmer_uniform.restype = ctypes.c_float
        L.orc_swimmer_normals.argtypes = [ctypes.c_ulonglong, _I, ctypes.c_uint64, _D]
        L.orc_bm_log.argtypes = [ctypes.c_double]
        L.orc_bm_log.restype = ctypes.c_double
        L.orc_bm_cossin_2pi.argtypes = [ctypes.c_double, _D, _D]
        L.orc_run.argtypes = [_I, _I, _I, ctypes.c_double, ctypes.c_double, _I,
                              _D, _D, _D, _D, _D, _I, ctypes.POINTER(OrcCommon),
                              _D, _D, _D, _U64, _I, _D, _D, _D]
        L.orc_run.restype = ctypes.c_long
        _lib = L
    return _lib


def set_threads(n):
    lib().orc_set_threads(int(n))


def max_threads():
    return lib().orc_get_max_threads()


def _p(a, ptr=_D):
    assert a.flags["C_CONTIGUOUS"], "oracle arrays must be C-contiguous"
    return a.ctypes.data_as(ptr)


def tables():
    """D3Q15 tables as numpy arrays, keyed like the reference's LBParams."""
    prm = OrcParams()
    lib().orc_eigenset(ctypes.byref(prm), 0.0, 0.0)
    out = {"cs2": prm.cs2}
    for name in ("w", "xi", "ci", "Q", "complement", "norms", "mm", "mmi"):
        out[name] = np.ctypeslib.as_array(getattr(prm, name)).copy()
    return out


def peskin_delta(x):
    return lib().orc_peskin_delta(float(x))


def philox4x32_10(ctr, key):
    c = np.ascontiguousarray(ctr, dtype=np.uint32)
    k = np.ascontiguousarray(key, dtype=np.uint32)
    out = np.zeros(4, dtype=np.uint32)
    lib().orc_philox4x32_10(_p(c, _U32), _p(k, _U32), _p(out, _U32))
    return out


def swimmer_uniform(seed, i, ctr):
    return lib().orc_swimmer_uniform(seed, i, ctr)


def swimmer_normals(seed, i, ctr):
    g = np.zeros(3)
    lib().orc_swimmer_normals(seed, i, ctr, _p(g))
    return g


def bm_log(x):
    """ln x as the Box-Muller transform evaluates it (series, same bits on the GPU)."""
    return lib().orc_bm_log(float(x))


def bm_cossin_2pi(u):
    """(cos 2 pi u, sin 2 pi u) as the Box-Muller transform evaluates them."""
    c, s = ctypes.c_double(), ctypes.c_double()
    lib().orc_bm_cossin_2pi(float(u), ctypes.byref(c), ctypes.byref(s))
    return c.value, s.value


class Lattice:
    """FP64 lattice state in the reference layout (lb.cu:11-17)."""

    def __init__(self, nx, ny, nz, tau_s, tau_b):
        self.size = (nx, ny, nz)
        self.n = nx * ny * nz
        self.tau_s, self.tau_b = float(tau_s), float(tau_b)
        self.rho = np.zeros(self.n)
        self.u = np.zeros(3 * self.n)
        self.force = np.zeros(3 * self.n)
        self.fOld = np.zeros(15 * self.n)
        self.fNew = np.zeros(15 * self.n)
        self.time_step = 0

    def InitFromHydro(self):
        lib().orc_init_from_hydro(*self.size, _p(self.rho), _p(self.u), _p(self.fOld))

    def ZeroForce(self):
        lib().orc_zero_force(self.n, _p(self.force))

    def Step(self):
        lib().orc_step(*self.size, self.tau_s, self.tau_b, _p(self.fOld), _p(self.force), _p(self.fNew))
        self.fOld, self.fNew = self.fNew, self.fOld
        self.time_step += 1

    def CalcHydro(self):
        lib().orc_calc_hydro(*self.size, _p(self.fOld), _p(self.force), _p(self.rho), _p(self.u))

    def interp(self, r):
        r = np.ascontiguousarray(r, dtype=np.float64)
        v = np.zeros(3)
        lib().orc_interp(*self.size, _p(self.u), _p(r), _p(v))
        return v


class Common:
    def __init__(self, P=0.0, l=0.0, alpha=0.0, mobility=0.0,
                 translational_advection_off=False, rotational_advection_off=False, seed=0):
        self.c = OrcCommon(P, l, alpha, mobility, int(translational_advection_off),
                           int(rotational_advection_off), seed)


class Swimmers:
    def __init__(self, n, common):
        self.num = n
        self.common = common
        self.r = np.zeros((3, n))
        self.v = np.zeros((3, n))
        self.n = np.zeros((3, n))
        self.prng = np.zeros(n, dtype=np.uint64)
        self.tumbles = 0

    def AddForces(self, lat):
        lib().orc_swimmers_add_forces(*lat.size, self.num, ctypes.byref(self.common.c),
                                      _p(self.r), _p(self.n), _p(lat.force))

    def Move(self, lat):
        self.tumbles += lib().orc_swimmers_move(*lat.size, self.num, ctypes.byref(self.common.c),
                                                _p(self.r), _p(self.v), _p(self.n),
                                                _p(self.prng, _U64), _p(lat.u))


class Tracers:
    def __init__(self, n):
        self.num = n
        self.r = np.zeros((3, n))
        self.s = np.zeros((3, n))
        self.v = np.zeros((3, n))

    def AddForces(self, lat):
        pass

    def Move(self, lat):
        lib().orc_tracers_move(*lat.size, self.num, _p(self.r), _p(self.s), _p(self.v), _p(lat.u))


def system_step(lat, things):
    """System.Step order (System.py:9-16)."""
    lat.ZeroForce()
    for th in things:
        th.AddForces(lat)
    lat.Step()
    lat.CalcHydro()
    for th in things:
        th.Move(lat)


def run(lat, steps, sw=None, tr=None):
    """`steps` coupled steps with one swimmer and one tracer array, in C."""
    empty = np.zeros((3, 1))
    e64 = np.zeros(1, dtype=np.uint64)
    ns = sw.num if sw is not None else 0
    nt = tr.num if tr is not None else 0
    cp = sw.common.c if sw is not None else OrcCommon()
    tumbles = lib().orc_run(
        *lat.size, lat.tau_s, lat.tau_b, int(steps), _p(lat.fOld), _p(lat.fNew), _p(lat.rho),
        _p(lat.u), _p(lat.force), ns, ctypes.byref(cp),
        _p(sw.r if ns else empty), _p(sw.v if ns else empty), _p(sw.n if ns else empty),
        _p(sw.prng if ns else e64, _U64), nt,
        _p(tr.r if nt else empty), _p(tr.s if nt else empty), _p(tr.v if nt else empty))
    lat.time_step += steps
    if sw is not None:
        sw.tumbles += tumbles
    return tumbles
