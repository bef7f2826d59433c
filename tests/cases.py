"""Seeded coupled-step cases and runners for the oracle and for the lbswim
package, so that parity tests feed both exactly the same inputs.

Synthetic inputs follow one draw order from numpy's PCG64: rho, u, swimmer r,
swimmer n, tracer r.  Arrays are in the reference layout.
"""
import dataclasses

import numpy as np

from util import isotropic, random_hydro, readme_mobility


@dataclasses.dataclass
class Case:
    size: tuple
    tau_s: float
    tau_b: float
    steps: int
    rho: np.ndarray
    u: np.ndarray
    sw_r: np.ndarray
    sw_n: np.ndarray
    common: dict
    tr_r: np.ndarray


def make_case(size, steps, tau_s=0.25, tau_b=None, ns=0, nt=0, alpha=0.0, seed=1234, drho=1e-3, du=1e-3,
              P=1e-3, l=0.1, grid_tracers=False, swimmer_at=None, swimmer_n=None):
    rng = np.random.default_rng(seed)
    n = size[0] * size[1] * size[2]
    if drho == 0 and du == 0:
        rho, u = np.ones(n), np.zeros(3 * n)
    else:
        rho, u = random_hydro(rng, n, drho, du)
    L = np.array(size, dtype=float)[:, None]
    sw_r = rng.uniform(0, 1, (3, ns)) * L if swimmer_at is None else np.array(swimmer_at, float).reshape(3, ns)
    sw_n = isotropic(rng, ns) if swimmer_n is None else np.array(swimmer_n, float).reshape(3, ns)
    if grid_tracers:
        tr_r = np.mgrid[: size[0], : size[1], : size[2]].reshape(3, -1).astype(float)
    else:
        tr_r = rng.uniform(0, 1, (3, nt)) * L
    common = dict(P=P, l=l, alpha=alpha, mobility=readme_mobility(tau_s), translational_advection_off=False,
                  rotational_advection_off=False, seed=seed)
    return Case(tuple(size), tau_s, tau_s if tau_b is None else tau_b, steps, rho, u, sw_r, sw_n, common, tr_r)


def config1(steps=1000, alpha=0.0):
    """16^3, tau 0.25, rho=1, u=0, one swimmer at (8,8,8) n=(0,0,1), P=1e-3,
    l=0.1, README mobility, seed 1234, 4096 tracers on the grid points."""
    return make_case((16, 16, 16), steps, 0.25, ns=1, alpha=alpha, drho=0.0, du=0.0, grid_tracers=True,
                     swimmer_at=[8.0, 8.0, 8.0], swimmer_n=[0.0, 0.0, 1.0])


def config2(steps=500, size=128, ns=4096, nt=100000):
    """rho = 1+1e-3 N(0,1), u = 1e-3 N(0,1); uniform swimmers with isotropic n,
    alpha = 0; uniform tracers."""
    return make_case((size, size, size), steps, 0.25, ns=ns, nt=nt, alpha=0.0)


OUT_KEYS = ("fOld", "rho", "u", "force", "sw_r", "sw_n", "sw_v", "sw_prng", "tr_r", "tr_s", "tr_v")


def run_oracle(O, case, steps=None):
    """Run `steps` (default case.steps) coupled steps with the oracle."""
    lat = O.Lattice(*case.size, case.tau_s, case.tau_b)
    lat.rho[:] = case.rho
    lat.u[:] = case.u
    lat.InitFromHydro()
    ns, nt = case.sw_r.shape[1], case.tr_r.shape[1]
    sw = tr = None
    if ns:
        sw = O.Swimmers(ns, O.Common(**case.common))
        sw.r[:] = case.sw_r
        sw.n[:] = case.sw_n
    if nt:
        tr = O.Tracers(nt)
        tr.r[:] = case.tr_r
    O.run(lat, case.steps if steps is None else steps, sw, tr)
    return _collect(lat.fOld, lat.rho, lat.u, lat.force, sw, tr, ns, nt), (sw.tumbles if sw else 0)


def _collect(f, rho, u, force, sw, tr, ns, nt):
    e3 = np.zeros((3, 0))
    return dict(fOld=np.array(f), rho=np.array(rho), u=np.array(u), force=np.array(force),
                sw_r=np.array(sw.r) if ns else e3, sw_n=np.array(sw.n) if ns else e3,
                sw_v=np.array(sw.v) if ns else e3,
                sw_prng=np.array(sw.prng) if ns else np.zeros(0, np.uint64),
                tr_r=np.array(tr.r) if nt else e3, tr_s=np.array(tr.s) if nt else e3,
                tr_v=np.array(tr.v) if nt else e3)


def build_lbswim(case, storage="f64", force_mode=None):
    """Lattice + arrays in the lbswim package, initialised from `case`."""
    from lbswim import lb, swimmers, tracers
    from lbswim.System import System

    lat = lb.Lattice(*case.size, case.tau_s, case.tau_b, storage=storage)
    if force_mode is not None:
        lat.set_force_mode(force_mode)
    lat.rho.data.ravel()[:] = case.rho
    lat.rho.H2D()
    lat.u.data.ravel()[:] = case.u
    lat.u.H2D()
    lat.ZeroForce()
    system = System(lat)
    ns, nt = case.sw_r.shape[1], case.tr_r.shape[1]
    sw = tr = None
    if ns:
        cp = swimmers.CommonParams()
        for k, v in case.common.items():
            setattr(cp, k, v)
        sw = swimmers.Array(ns, cp)
        sw.r[:] = case.sw_r
        sw.n[:] = case.sw_n
        sw.v[:] = 0.0
        sw.H2D()
        system.AddThing(sw)
    if nt:
        tr = tracers.Array(nt)
        tr.r[:] = case.tr_r
        tr.v[:] = 0.0
        tr.H2D()
        system.AddThing(tr)
    lat.InitFromHydro()
    return lat, system, sw, tr


def collect_lbswim(lat, sw, tr):
    for name in ("fOld", "rho", "u", "force"):
        getattr(lat, name).D2H()
    ns = sw.num if sw is not None else 0
    nt = tr.num if tr is not None else 0
    if sw is not None:
        sw.D2H()
    if tr is not None:
        tr.D2H()
    return _collect(lat.fOld.data.ravel(), lat.rho.data.ravel(), lat.u.data.ravel(), lat.force.data.ravel(),
                    sw, tr, ns, nt)


def run_lbswim(case, storage="f64", steps=None, force_mode=None):
    lat, system, sw, tr = build_lbswim(case, storage, force_mode)
    system.Step(case.steps if steps is None else steps)
    return collect_lbswim(lat, sw, tr), (sw.tumbles() if sw is not None else 0)
