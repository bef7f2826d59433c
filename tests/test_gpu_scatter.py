"""GPU tests of the scatter force of the FP32 form (LBSWIM_FORCE_SCATTER).

With the scatter force the coupled step builds no spreading lists: each
step's swimmer moves add the next step's ZeroForce + AddForces into a private
force buffer with FP32 atomics (the reference's own scatter,
SwimmerArray.cu:42-112), and the fused launch reads it.  Atomics add in a
varying order, so the results are compared with the deterministic stored
force of the same FP32 form within FP32 rounding, not bit for bit:

- the force array after a step (relative 2e-6: a few FP32 ulps of each sum)
  and the state after several steps (u, rho - 1 relative 1e-6, swimmer and
  tracer positions and orientations 1e-8 absolute);
- the FP64 oracle within the FP32 form's 1e-5 relative u error;
- the force buffer carried between calls is dropped when the swimmers change
  outside the step (H2D of r, a standalone Move, CommonParams.H2D) or a
  different set of arrays is stepped: every such case must match a run in
  which the stored force is rebuilt from the same state.
"""
import os
import sys

import numpy as np
import pytest

import lbswim_oracle as O
from cases import build_lbswim, collect_lbswim, make_case, run_oracle
from util import rel_inf

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def oracle_threads():
    O.set_threads(min(16, os.cpu_count() or 1))


def case_of(size, seed, ns=120, nt=400):
    case = make_case(size, 0, 0.3, 0.7, ns=ns, nt=nt, alpha=0.0, seed=seed, drho=1e-2, du=1e-2, P=1e-2, l=0.6)
    L = np.array(size, float)
    # swimmers across the periodic faces and the 64-site tile boundaries
    case.sw_r[:, 0] = [0.05, 0.1, 0.2]
    case.sw_r[:, 1] = L - 0.05
    case.sw_r[:, 2] = [L[0] / 2, L[1] - 0.5, 63.7 % L[2]]
    return case


def close(a, b, label):
    """FP32-rounding agreement of two FP32-form states."""
    checks = {"u": (rel_inf(a["u"], b["u"]), 1e-6), "rho": (rel_inf(a["rho"] - 1.0, b["rho"] - 1.0), 1e-6),
              "sw_r": (float(np.max(np.abs(a["sw_r"] - b["sw_r"]), initial=0.0)), 1e-8),
              "sw_n": (float(np.max(np.abs(a["sw_n"] - b["sw_n"]), initial=0.0)), 1e-8),
              "tr_r": (float(np.max(np.abs(a["tr_r"] - b["tr_r"]), initial=0.0)), 1e-8)}
    say("%s: %s" % (label, {k: "%.1e" % v for k, (v, _) in checks.items()}))
    for k, (v, tol) in checks.items():
        assert v <= tol, (label, k, v)


def stepped(case, mode, steps, rows=8, strips=0):
    lat, system, sw, tr = build_lbswim(case, "f32", mode)
    lat.set_fusion(True, rows, strips)
    assert lat.fusion[0]
    system.Step(steps)
    return lat, system, sw, tr


@pytest.mark.parametrize("size,rows", [((6, 16, 64), 8), ((5, 12, 128), 4), ((3, 8, 64), 2)])
def test_scatter_matches_stored_force(size, rows):
    case = case_of(size, seed=size[0] * 7 + rows)
    say("scatter vs stored force, FP32: %s rows %d" % (size, rows))
    lat, _, sw, tr = stepped(case, "scatter", 6, rows)
    got = collect_lbswim(lat, sw, tr)
    lat, _, sw, tr = stepped(case, "stored", 6, rows)
    want = collect_lbswim(lat, sw, tr)
    e = rel_inf(got["force"], want["force"])
    say("force after 6 steps: %.2e" % e)
    assert e <= 2e-6
    assert np.abs(want["force"]).max() > 0
    close(got, want, "state after 6 steps")


def test_scatter_against_the_fp64_oracle():
    case = case_of((8, 16, 64), seed=5)
    case.steps = 30
    say("scatter FP32 vs FP64 oracle: (8, 16, 64), 30 steps")
    ref, _ = run_oracle(O, case)
    lat, system, sw, tr = stepped(case, "scatter", case.steps)
    got = collect_lbswim(lat, sw, tr)
    e = rel_inf(got["u"], ref["u"])
    say("u error vs oracle %.2e" % e)
    assert e <= 1e-5


def test_scatter_buffer_dropped_when_swimmers_change():
    """Changes to the swimmers between calls (H2D of r, a standalone Move,
    new CommonParams) and a call with another set of arrays must reach the
    next step's force exactly as with the stored force."""
    case = case_of((4, 16, 64), seed=17, ns=80, nt=200)
    runs = {}
    for mode in ("scatter", "stored"):
        say("buffer invalidation, %s force" % mode)
        lat, system, sw, tr = stepped(case, mode, 3)
        # positions replaced from the host
        sw.D2H()
        sw.r[:] = (np.asarray(sw.r) + 0.37) % np.array(case.size, float)[:, None]
        sw.H2D()
        system.Step(2)
        # a move outside the coupled step
        sw.Move(lat)
        system.Step(2)
        # stronger swimmers
        sw.common.P = 3e-2
        system.Step(2)
        # a call without the swimmers, then with them again
        from lbswim import _core
        _core.system_step(lat, [], [tr], 1)
        system.Step(2)
        runs[mode] = collect_lbswim(lat, sw, tr)
    e = rel_inf(runs["scatter"]["force"], runs["stored"]["force"])
    say("force after the changes: %.2e" % e)
    assert e <= 2e-6
    close(runs["scatter"], runs["stored"], "state after the changes")


def test_scatter_calls_of_one_step_equal_one_call():
    case = case_of((4, 8, 64), seed=29)
    a = stepped(case, "scatter", 5)
    lat, system, sw, tr = stepped(case, "scatter", 0)
    for _ in range(5):
        system.Step(1)
    close(collect_lbswim(a[0], a[2], a[3]), collect_lbswim(lat, sw, tr), "five calls vs one")


def test_scatter_force_read_between_steps():
    """D2H of the force between scatter steps gives the force the last step
    used, and standalone Step/CalcHydro after it read that force."""
    case = case_of((4, 8, 64), seed=41)
    out = {}
    for mode in ("scatter", "stored"):
        lat, system, sw, tr = stepped(case, mode, 3)
        lat.force.D2H()
        f = np.array(lat.force.data, copy=True)
        lat.Step()
        lat.CalcHydro()
        system.Step(2)
        res = collect_lbswim(lat, sw, tr)
        res["f_mid"] = f
        out[mode] = res
    assert rel_inf(out["scatter"]["f_mid"], out["stored"]["f_mid"]) <= 2e-6
    close(out["scatter"], out["stored"], "after standalone Step/CalcHydro")


def test_scatter_mode_only_for_fp32():
    from lbswim import lb
    lat = lb.Lattice(4, 8, 64, 0.25, 0.25)
    with pytest.raises(Exception):
        lat.set_force_mode("scatter")
    lat32 = lb.Lattice(4, 8, 64, 0.25, 0.25, storage="f32")
    lat32.set_force_mode("scatter")
    assert lat32.force_mode == "scatter"


def test_autotune_may_choose_the_scatter_force():
    case = case_of((6, 16, 64), seed=3)
    lat, system, sw, tr = build_lbswim(case, "f32", "auto")
    rows, strips = system.Autotune(2)
    say("FP32 autotune: rows %d strips %d force %s" % (rows, strips, lat.force_mode))
    assert lat.force_mode in ("auto", "scatter")
    tuned = lat.time_step
    system.Step(3)
    got = collect_lbswim(lat, sw, tr)
    lat2, system2, sw2, tr2 = build_lbswim(case, "f32", "stored")
    system2.Step(tuned + 3)
    close(got, collect_lbswim(lat2, sw2, tr2), "after autotune")
