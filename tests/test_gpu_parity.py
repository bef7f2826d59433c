"""GPU parity: the HIP path (through the lbswim package and the C ABI) against
the CPU oracle on the same seeded inputs.

Bar: the FP64 storage form keeps the reference's arithmetic, so lattice
fields must agree bit for bit on single operations and coupled steps, and
within 1e-12 norm-wise relative (||a-b||_inf/||b||_inf) after long coupled
runs (bit-identical in practice: the fused step spreads forces without
atomics).  The FP32 storage form must stay within 1e-5 relative of the FP64
run in u.  The swimmer force of the fused step is checked both stored (the
force array written and read back) and virtual (evaluated by the collide and
CalcHydro kernels from per-block point lists, written only on demand).
"""
import os
import pickle
import sys

import numpy as np
import pytest

import lbswim_oracle as O
from cases import build_lbswim, collect_lbswim, config1, config2, make_case, run_lbswim, run_oracle
from util import rel_inf

pytestmark = pytest.mark.gpu

PARITY = 1e-12  # FP64 coupled-run tolerance (norm-wise relative)
F32_U_TOL = 1e-5  # FP32-storage u error against FP64


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def oracle_threads():
    O.set_threads(min(16, os.cpu_count() or 1))


def compare(gpu, ref, keys, tol, label):
    worst = {}
    for k in keys:
        if ref[k].size == 0:
            continue
        worst[k] = rel_inf(gpu[k], ref[k])
    say("%s: %s" % (label, {k: "%.2e" % v for k, v in worst.items()}))
    bad = {k: v for k, v in worst.items() if not v <= tol}
    assert not bad, bad


# ------------------------------------------------------------- device math

def test_peskin_device_bitwise():
    from lbswim import _core
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.uniform(-2.5, 2.5, 200000), [0.0, -0.0, 1.0, -1.0, 2.0, -2.0, 0.5, 1.5, 1e-300,
                                                         np.nextafter(1.0, 0), np.nextafter(2.0, 0)]])
    say("peskin probe: %d points" % x.size)
    got = _core.diag_peskin(x)
    ref = np.array([O.peskin_delta(v) for v in x])
    assert np.array_equal(got, ref)


def test_interp_device_bitwise():
    from lbswim import _core
    size = (9, 7, 11)
    rng = np.random.default_rng(1)
    u = rng.standard_normal(3 * 9 * 7 * 11)
    r = np.vstack([rng.uniform(-1.5, s + 0.5, 3000) for s in size])
    say("interp probe: lattice %s, %d points" % (size, r.shape[1]))
    got = _core.diag_interp(size, u, r)
    lat = O.Lattice(*size, 0.25, 0.25)
    lat.u[:] = u
    ref = np.stack([lat.interp(r[:, i]) for i in range(r.shape[1])], axis=1)
    assert np.array_equal(got, ref)


# ------------------------------------------------------------- lattice ops

LATTICE_SIZES = [(4, 5, 6), (7, 5, 3), (1, 1, 1), (2, 1, 2), (16, 16, 16), (3, 70, 65), (5, 1, 130)]


@pytest.mark.parametrize("size", LATTICE_SIZES)
def test_lattice_ops_bitwise(size):
    from lbswim import lb
    say("lattice ops bitwise: size %s" % (size,))
    rng = np.random.default_rng(7)
    n = size[0] * size[1] * size[2]
    rho, u = 1.0 + 0.05 * rng.standard_normal(n), 0.05 * rng.standard_normal(3 * n)
    F = 1e-3 * rng.standard_normal(3 * n)
    ref = O.Lattice(*size, 0.3, 0.8)
    ref.rho[:], ref.u[:] = rho, u
    ref.InitFromHydro()
    gl = lb.Lattice(*size, 0.3, 0.8)
    gl.rho.data.ravel()[:] = rho
    gl.u.data.ravel()[:] = u
    gl.rho.H2D()
    gl.u.H2D()
    gl.InitFromHydro()
    gl.fOld.D2H()
    assert np.array_equal(gl.fOld.data.ravel(), ref.fOld), "InitFromHydro"
    for step in range(3):
        ref.force[:] = F * (step + 1)
        gl.force.data.ravel()[:] = F * (step + 1)
        gl.force.H2D()
        ref.Step()
        gl.Step()
        ref.CalcHydro()
        gl.CalcHydro()
        for name in ("fOld", "rho", "u"):
            getattr(gl, name).D2H()
            assert np.array_equal(getattr(gl, name).data.ravel(), getattr(ref, name)), (name, step)
    assert gl.time_step == 3
    gl.ZeroForce()
    gl.force.D2H()
    assert not gl.force.data.any()
    # fNew holds the previous distributions after the swap (lb.cu:242-244)
    gl.fNew.D2H()
    assert np.array_equal(gl.fNew.data.ravel(), ref.fNew)


@pytest.mark.parametrize("force_mode", ["stored", "virtual"])
def test_system_step_single_ops_bitwise(force_mode):
    """One coupled step with swimmers whose stencils cross the periodic faces."""
    case = make_case((12, 9, 10), 1, 0.3, 0.6, ns=40, nt=50, alpha=0.0, seed=5, drho=1e-2, du=1e-2, P=1e-2, l=0.7)
    case.sw_r[:, :4] = np.array([[0.1, 11.9, 0.0, 5.0], [8.95, 0.2, 0.0, 4.0], [9.99, 0.01, 0.0, 3.0]])
    say("system step bitwise: %s ns=40 nt=50 force %s" % (case.size, force_mode))
    ref, _ = run_oracle(O, case)
    got, _ = run_lbswim(case, force_mode=force_mode)
    # the fused step spreads forces as a per-site gather in swimmer order, so
    # even sites reached by several swimmers match the sequential oracle exactly
    for k in ref:
        assert np.array_equal(got[k], ref[k]), k


@pytest.mark.parametrize("force_mode", ["stored", "virtual"])
@pytest.mark.parametrize("size,ns,l", [((3, 2, 5), 7, 0.4), ((1, 6, 2), 3, 1.5), ((9, 33, 40), 300, 0.1),
                                       ((40, 20, 70), 2000, 2.5), ((5, 3, 300), 40, 0.8), ((6, 4, 257), 60, 1.2),
                                       ((20, 24, 30), 2600, 0.3)])
def test_spreading_exact_on_awkward_shapes(size, ns, l, force_mode):
    """Stencils wrapping several times onto lattices thinner than 4, partial
    blocks, z-runs crossing block boundaries and the periodic face, dense
    overlapping swimmers, long swimmers spanning blocks.  With the stored
    force, systems of up to 1024 swimmer ends (4096 on lattices thinner than
    4) walk all points per run, denser ones with every side >= 4 (2000 and
    2600 swimmers) use per-site segments, and the others sorted lists."""
    case = make_case(size, 3, 0.3, ns=ns, nt=0, alpha=0.0, seed=ns, drho=1e-2, du=1e-2, P=1e-2, l=l)
    say("spreading exact: %s ns=%d l=%g force %s" % (size, ns, l, force_mode))
    ref, _ = run_oracle(O, case)
    got, _ = run_lbswim(case, force_mode=force_mode)
    for k in ("force", "fOld", "sw_r", "sw_n"):
        assert np.array_equal(got[k], ref[k]), k


DENSE_SITE_CASES = [((24, 24, 24), 3700, 0.1, False), ((4, 5, 6), 2100, 0.3, False),
                    ((7, 33, 20), 2500, 2.5, False), ((16, 16, 16), 3000, 0.2, True)]


@pytest.mark.parametrize("size,ns,l,clustered", DENSE_SITE_CASES)
def test_dense_stored_force_by_site(size, ns, l, clustered):
    """Dense systems (more than 1024 swimmer ends, every side >= 4) form the
    stored force from per-site segments of the stencil entries, ordered by
    point id: the force, state and swimmers equal the oracle bit for bit over
    three steps -- also on a 4x5x6 lattice, where each stencil covers most of
    it, with swimmers longer than two sites, and with every swimmer inside one
    site (segments of about 6000 entries)."""
    case = make_case(size, 3, 0.3, ns=ns, nt=0, alpha=0.0, seed=ns + size[0], drho=1e-2, du=1e-2, P=1e-2, l=l)
    if clustered:
        rng = np.random.default_rng(9)
        case.sw_r[:] = 7.3 + rng.uniform(0.0, 1.0, case.sw_r.shape)
    say("dense stored force by site: %s ns=%d l=%g clustered=%s" % (size, ns, l, clustered))
    ref, _ = run_oracle(O, case)
    got, _ = run_lbswim(case, force_mode="stored")
    for k in ("force", "fOld", "rho", "u", "sw_r", "sw_n"):
        assert np.array_equal(got[k], ref[k]), k


@pytest.mark.parametrize("size,ns,l,clustered", DENSE_SITE_CASES[:2])
def test_dense_stored_force_by_site_fp32(size, ns, l, clustered):
    """The FP32 form stores the same double sums rounded once to float: its
    first step's force equals the oracle's rounded to FP32."""
    case = make_case(size, 1, 0.3, ns=ns, nt=0, alpha=0.0, seed=ns + size[0], drho=1e-2, du=1e-2, P=1e-2, l=l)
    say("dense stored force by site, FP32: %s ns=%d" % (size, ns))
    ref, _ = run_oracle(O, case)
    got, _ = run_lbswim(case, storage="f32", force_mode="stored")
    want = ref["force"].astype(np.float32).astype(np.float64)
    assert np.array_equal(got["force"], want)
    assert np.abs(want).max() > 0


@pytest.mark.parametrize("force_mode", ["stored", "virtual"])
def test_two_swimmer_arrays_in_order(force_mode):
    """Two swimmer arrays: contributions are added array by array, as System.Step does."""
    from lbswim import swimmers
    from lbswim.System import System
    size = (10, 9, 11)
    c1 = make_case(size, 0, 0.3, ns=20, nt=0, seed=1, drho=1e-2, du=1e-2, P=1e-2, l=0.5)
    c2 = make_case(size, 0, 0.3, ns=15, nt=0, seed=2, drho=1e-2, du=1e-2, P=2e-2, l=0.3)
    say("two swimmer arrays, force %s" % force_mode)
    lat, system, sw1, _ = build_lbswim(c1, force_mode=force_mode)
    cp = swimmers.CommonParams()
    for k, v in c2.common.items():
        setattr(cp, k, v)
    sw2 = swimmers.Array(15, cp)
    sw2.r[:] = c2.sw_r
    sw2.n[:] = c2.sw_n
    sw2.H2D()
    system.AddThing(sw2)
    system.Step(4)
    lat.force.D2H()
    lat.fOld.D2H()
    sw2.D2H()
    # oracle: same order by hand
    ol = O.Lattice(*size, 0.3, 0.3)
    ol.rho[:], ol.u[:] = c1.rho, c1.u
    ol.InitFromHydro()
    s1, s2 = O.Swimmers(20, O.Common(**c1.common)), O.Swimmers(15, O.Common(**c2.common))
    s1.r[:], s1.n[:] = c1.sw_r, c1.sw_n
    s2.r[:], s2.n[:] = c2.sw_r, c2.sw_n
    for _ in range(4):
        O.system_step(ol, [s1, s2])
    assert np.array_equal(lat.force.data.ravel(), ol.force)
    assert np.array_equal(lat.fOld.data.ravel(), ol.fOld)
    assert np.array_equal(sw2.r.data, s2.r)


def test_two_dense_swimmer_arrays_by_site():
    """Two swimmer arrays on a dense system (3000 swimmer ends on 16^3: the
    per-site segments): the first array's points come first in every site's
    sum, as System.Step adds them array by array."""
    from lbswim import swimmers
    from lbswim.System import System
    size = (16, 16, 16)
    c1 = make_case(size, 0, 0.3, ns=800, nt=0, seed=3, drho=1e-2, du=1e-2, P=1e-2, l=0.4)
    c2 = make_case(size, 0, 0.3, ns=700, nt=0, seed=4, drho=1e-2, du=1e-2, P=2e-2, l=0.9)
    say("two dense swimmer arrays by site")
    lat, system, sw1, _ = build_lbswim(c1)
    cp = swimmers.CommonParams()
    for k, v in c2.common.items():
        setattr(cp, k, v)
    sw2 = swimmers.Array(700, cp)
    sw2.r[:] = c2.sw_r
    sw2.n[:] = c2.sw_n
    sw2.H2D()
    system.AddThing(sw2)
    system.Step(3)
    assert lat.force_mode == "auto" and not lat.force_is_virtual
    lat.force.D2H()
    lat.fOld.D2H()
    sw1.D2H()
    sw2.D2H()
    ol = O.Lattice(*size, 0.3, 0.3)
    ol.rho[:], ol.u[:] = c1.rho, c1.u
    ol.InitFromHydro()
    s1, s2 = O.Swimmers(800, O.Common(**c1.common)), O.Swimmers(700, O.Common(**c2.common))
    s1.r[:], s1.n[:] = c1.sw_r, c1.sw_n
    s2.r[:], s2.n[:] = c2.sw_r, c2.sw_n
    for _ in range(3):
        O.system_step(ol, [s1, s2])
    assert np.array_equal(lat.force.data.ravel(), ol.force)
    assert np.array_equal(lat.fOld.data.ravel(), ol.fOld)
    assert np.array_equal(sw1.r.data, s1.r)
    assert np.array_equal(sw2.n.data, s2.n)


def test_virtual_force_written_on_demand():
    """After fused steps with a virtual force, every path that reads or
    replaces the force array sees what the stored form holds: D2H, a
    standalone Step/CalcHydro, a standalone AddForces on top, and an H2D."""
    case = make_case((14, 12, 20), 0, 0.3, ns=30, nt=20, seed=8, drho=1e-2, du=1e-2, P=1e-2, l=0.6)
    say("virtual force on demand")
    runs = {}
    for mode in ("stored", "virtual"):
        lat, system, sw, tr = build_lbswim(case, force_mode=mode)
        system.Step(3)
        out = {"after_fused": collect_lbswim(lat, sw, tr)}
        system.Step(2)
        lat.Step()                    # standalone Step reads the force of the last fused step
        lat.CalcHydro()
        out["standalone"] = collect_lbswim(lat, sw, tr)
        system.Step(1)
        sw.AddForces(lat)             # adds on top of the fused step's force
        lat.force.D2H()
        out["addforces"] = np.array(lat.force.data).ravel()
        system.Step(1)
        lat.force.data.ravel()[:] = 1e-4
        lat.force.H2D()               # replaces the virtual force
        lat.Step()
        out["h2d"] = collect_lbswim(lat, sw, tr)
        runs[mode] = out
    for stage in ("after_fused", "standalone", "h2d"):
        for k, v in runs["stored"][stage].items():
            assert np.array_equal(runs["virtual"][stage][k], v), (stage, k)
    # the standalone AddForces adds with hardware atomics: equal to rounding
    assert rel_inf(runs["virtual"]["addforces"], runs["stored"]["addforces"]) < 1e-14
    assert np.any(runs["stored"]["after_fused"]["force"] != 0)


def test_standalone_add_forces_atomic_path():
    """lbswim_swimmers_add_forces (hardware atomics, additive) against the oracle."""
    case = make_case((12, 14, 16), 0, 0.25, ns=500, nt=0, seed=3, drho=1e-2, du=1e-2, P=1e-2, l=0.6)
    say("standalone AddForces")
    lat, system, sw, _ = build_lbswim(case)
    lat.ZeroForce()
    sw.AddForces(lat)
    sw.AddForces(lat)
    lat.force.D2H()
    ol = O.Lattice(*case.size, 0.25, 0.25)
    osw = O.Swimmers(500, O.Common(**case.common))
    osw.r[:], osw.n[:] = case.sw_r, case.sw_n
    ol.ZeroForce()
    osw.AddForces(ol)
    osw.AddForces(ol)
    # atomics add in any order: equal to rounding
    assert rel_inf(lat.force.data.ravel(), ol.force) < 1e-14


# ------------------------------------------------------------- coupled runs

def test_config1_alpha0_1000_steps():
    case = config1(1000, alpha=0.0)
    say("config1 alpha=0: 16^3, 1 swimmer, 4096 tracers, 1000 steps")
    ref, _ = run_oracle(O, case)
    got, _ = run_lbswim(case)
    compare(got, ref, ("fOld", "rho", "u", "force", "sw_r", "sw_n", "sw_v", "tr_r", "tr_v", "tr_s"), PARITY,
            "config1 alpha=0 after 1000 steps")


def test_config1_alpha_tumbling_1000_steps():
    alpha = 0.01
    case = config1(1000, alpha=alpha)
    say("config1 alpha=0.01: 1000 steps")
    ref, t_ref = run_oracle(O, case)
    got, t_gpu = run_lbswim(case)
    # same counter-based stream on both sides: identical tumble decisions
    assert t_gpu == t_ref
    assert np.array_equal(got["sw_prng"], ref["sw_prng"])
    # tumble orientations use the same series log/cos/sin as the oracle: bit for bit
    for k in ref:
        assert np.array_equal(got[k], ref[k]), k


def test_tumble_rate_binomial_bound():
    from lbswim import lb, swimmers
    alpha, ns, steps = 0.01, 20000, 50
    say("tumble statistics: %d swimmers x %d steps, alpha=%g" % (ns, steps, alpha))
    lat = lb.Lattice(16, 16, 16, 0.25, 0.25)
    lat.rho.data[:] = 1.0
    lat.rho.H2D()
    lat.InitFromHydro()
    lat.CalcHydro()
    cp = swimmers.CommonParams()
    cp.P, cp.l, cp.alpha, cp.seed = 0.0, 0.1, alpha, 2024
    sw = swimmers.Array(ns, cp)
    rng = np.random.default_rng(3)
    sw.r[:] = rng.uniform(0, 16, (3, ns))
    sw.n[:] = np.array([0.0, 0.0, 1.0])[:, None]
    sw.H2D()
    for _ in range(steps):
        sw.Move(lat)
    N = ns * steps
    t = sw.tumbles()
    assert abs(t - alpha * N) < 4.0 * np.sqrt(N * alpha * (1 - alpha)), t
    sw.D2H()
    turned = sw.n.data[2] != 1.0
    nn = sw.n.data[:, turned]
    assert np.abs(nn.mean(axis=1)).max() < 4.0 / np.sqrt(3 * nn.shape[1])
    np.testing.assert_allclose(np.linalg.norm(sw.n.data, axis=0), 1.0, atol=1e-15)


def test_config2_reduced_500_steps():
    """The config2 recipe on a 32^3 lattice for the full 500 steps."""
    case = config2(steps=500, size=32, ns=64, nt=2000)
    say("config2 recipe at 32^3: 64 swimmers, 2000 tracers, 500 steps")
    ref, _ = run_oracle(O, case)
    got, _ = run_lbswim(case)
    compare(got, ref, ("fOld", "rho", "u", "sw_r", "sw_n", "sw_v", "tr_r", "tr_v"), PARITY,
            "config2@32^3 after 500 steps")
    for k in ref:  # deterministic spreading: in fact bit for bit
        assert np.array_equal(got[k], ref[k]), k


def test_config2_recipe_fused_500_steps():
    """The config2 recipe on a 64^3 lattice, where the coupled step takes the
    fused launch, the virtual force and the tracer moves on the second
    stream, for the full 500 steps: bit for bit with the oracle."""
    case = config2(steps=500, size=64, ns=256, nt=8000)
    say("config2 recipe at 64^3 (fused step): 256 swimmers, 8000 tracers, 500 steps")
    ref, _ = run_oracle(O, case)
    got, _ = run_lbswim(case)
    for k in ref:
        assert np.array_equal(got[k], ref[k]), k


# ------------------------------------------------------------- FP32 storage

@pytest.mark.parametrize("size,steps", [((32, 32, 32), 200), ((64, 48, 40), 100)])
def test_f32_storage_error_against_f64(size, steps):
    case = make_case(size, steps, 0.25, ns=64, nt=500, alpha=0.0, seed=99)
    say("f32 vs f64: %s, %d steps" % (size, steps))
    a, _ = run_lbswim(case, "f64")
    b, _ = run_lbswim(case, "f32")
    errs = {k: rel_inf(b[k], a[k]) for k in ("u", "rho", "sw_r", "tr_r", "fOld")}
    du = rel_inf(b["rho"] - 1.0, a["rho"] - 1.0)
    say("f32 errors: %s  drho %.2e" % ({k: "%.2e" % v for k, v in errs.items()}, du))
    assert errs["u"] <= F32_U_TOL
    assert errs["sw_r"] <= 1e-9 and errs["tr_r"] <= 1e-9


# ------------------------------------------------------------- properties

def test_conservation_at_256_cubed():
    """Total mass is conserved and total momentum unchanged (force dipoles sum
    to zero) at the full 256^3 size, FP64 storage."""
    case = make_case((256, 256, 256), 0, 0.25, ns=32768, nt=0, alpha=0.0, seed=11)
    say("conservation: 256^3, 32768 swimmers, 20 steps")
    lat, system, sw, tr = build_lbswim(case, "f64")
    t0 = lat.totals()
    system.Step(20)
    t1 = lat.totals()
    say("totals before %s after %s" % (t0, t1))
    assert abs(t1[0] - t0[0]) / t0[0] < 1e-13
    assert np.abs(t1[1:] - t0[1:]).max() < 1e-9


def test_pickle_continuation_identical():
    case = make_case((10, 12, 8), 0, 0.25, ns=5, nt=20, alpha=0.0, seed=4)
    say("pickle continuation")
    lat, system, sw, tr = build_lbswim(case, "f64")
    system.Step(7)
    blob = pickle.dumps((lat, sw, tr))
    lat2, sw2, tr2 = pickle.loads(blob)
    lat2.CalcHydro()
    from lbswim.System import System
    s2 = System(lat2)
    s2.AddThing(sw2)
    s2.AddThing(tr2)
    system.Step(5)
    s2.Step(5)
    a, b = collect_lbswim(lat, sw, tr), collect_lbswim(lat2, sw2, tr2)
    assert lat2.time_step == lat.time_step == 12
    for k in a:
        assert np.array_equal(a[k], b[k]), k


def test_empty_particle_arrays_and_python_driver():
    from lbswim import lb, swimmers, tracers
    from lbswim.System import System
    say("empty arrays")
    lat = lb.Lattice(8, 8, 8, 0.25, 0.25)
    lat.rho.data[:] = 1.0
    lat.rho.H2D()
    lat.InitFromHydro()
    cp = swimmers.CommonParams()
    sw = swimmers.Array(0, cp)
    tr = tracers.Array(0)

    class Passive:  # a non-native "thing" forces the Python driver
        def AddForces(self, lat):
            pass

        def Move(self, lat):
            pass

    s = System(lat)
    for th in (sw, tr, Passive()):
        s.AddThing(th)
    s.Step(3)
    assert lat.time_step == 3
    sw.D2H()
    tr.D2H()
    assert sw.r.data.shape == (3, 0)


@pytest.mark.parametrize("size,ns,virtual", [((128, 128, 128), 250, True), ((96, 96, 96), 300, True),
                                             ((96, 96, 96), 4000, False), ((48, 48, 48), 250, False),
                                             ((24, 24, 24), 2000, False)])
def test_auto_force_mode_density_rule(size, ns, virtual):
    """Force mode auto, a single system: the force stays virtual up to
    1 + n / 2^18 mean entries per 64-site list (about 17 per swimmer end)
    and is stored beyond; as an ensemble member the same system keeps the
    virtual force up to 8 entries."""
    from lbswim.ensemble import Ensemble
    case = make_case(size, 1, 0.3, ns=ns, nt=0, seed=ns)
    n = size[0] * size[1] * size[2]
    mean_entries = 2 * ns * 17 / (4 * ((n + 255) // 256))
    say("auto force mode: %s ns=%d (%.1f entries per list)" % (size, ns, mean_entries))
    lat, system, sw, tr = build_lbswim(case)
    system.Step(1)
    assert lat.force_is_virtual == virtual
    ens = Ensemble("f64")
    lat2, _, sw2, tr2 = build_lbswim(case)
    ens.add(lat2, [x for x in (sw2, tr2) if x is not None])
    ens.step(1)
    vsites, ssites = ens.sites()
    member_virtual = vsites + ens.fused_sites() == n
    assert member_virtual == (mean_entries <= 8), (vsites, ssites)
