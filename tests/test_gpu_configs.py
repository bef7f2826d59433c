"""GPU parity at the workloads' own shapes, against the CPU oracle.

  C1  16^3, one swimmer, 4096 grid tracers, 1000 steps (FP32 error vs oracle;
      the FP64 runs are in test_gpu_parity.py)
  C2  128^3, 4096 swimmers, 1e5 tracers, alpha 0, 500 steps: FP64 bit for bit
      with the oracle, FP32 error against the oracle
  C3  256^3, 32768 swimmers (alpha 0.01), 1e6 tracers, 200 steps: FP64 bit
      for bit with the oracle, FP32 error against the oracle; FP32
      conservation totals
  C4  64 simulations of {16,24,32,48,64}^3 (the benchmark's ensemble_cases),
      1000 steps, graph replay and autotuner, both storage forms: every member
      bit-identical to its solo run, four FP64 members (the 64^3 one among
      them) bit-identical to the oracle, FP32 members within 1e-5 in u of
      their FP64 twins, FP64 mass conserved per member
  C5  twelve requests drawn from the serving ranges (cubes 24-96, 200-5000
      steps, 16-4096 swimmers, 0-1e5 tracers, FP32/FP64 50/50) served with
      continuous batching: every result equal to the request run alone, the
      three cheapest FP64 results bit-identical to the oracle

Error measures: ||a-b||_inf / ||b||_inf for lattice fields (rho as rho - 1),
minimum-image absolute displacement for positions.
"""
import os
import sys

import numpy as np
import pytest

import lbswim_oracle as O
from cases import Case, build_lbswim, collect_lbswim, config1, config2, make_case, run_lbswim, run_oracle
from util import rel_inf

pytestmark = pytest.mark.gpu

F32_U_TOL = 1e-5     # FP32-storage u (and rho - 1) error against FP64
F32_TRAJ_TOL = 1e-7  # FP32-storage particle trajectory error (lattice units)


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def oracle_threads():
    O.set_threads(min(16, os.cpu_count() or 1))


def traj_err(a, b, L):
    d = np.abs(np.asarray(a) - np.asarray(b))
    d = np.minimum(d, L - d)
    return float(d.max()) if d.size else 0.0


def f32_errors(got, ref, L):
    """FP32-storage run against the FP64 oracle of the same inputs."""
    e = {"u": rel_inf(got["u"], ref["u"]), "drho": rel_inf(got["rho"] - 1.0, ref["rho"] - 1.0)}
    for k in ("sw_r", "tr_r"):
        if ref[k].size:
            e[k] = traj_err(got[k], ref[k], L)
    if ref["sw_n"].size:
        e["sw_n"] = float(np.abs(got["sw_n"] - ref["sw_n"]).max())
    return e


def check_f32(e, label):
    say("%s FP32 vs FP64 oracle: %s" % (label, {k: "%.2e" % v for k, v in e.items()}))
    assert e["u"] <= F32_U_TOL and e["drho"] <= F32_U_TOL, e
    for k in ("sw_r", "tr_r", "sw_n"):
        if k in e:
            assert e[k] <= F32_TRAJ_TOL, (k, e[k])


def assert_bitwise(got, ref, keys, label):
    for k in keys:
        assert np.array_equal(got[k], ref[k]), (label, k)


ALL_KEYS = ("fOld", "rho", "u", "force", "sw_r", "sw_n", "sw_v", "sw_prng", "tr_r", "tr_v")


# ------------------------------------------------------------------ C1-C3

def test_c1_f32_error_against_oracle():
    case = config1(1000, alpha=0.01)
    say("C1: 16^3, 1 swimmer (alpha 0.01), 4096 tracers, 1000 steps, FP32 vs oracle")
    ref, t_ref = run_oracle(O, case)
    got, t_gpu = run_lbswim(case, "f32")
    assert t_gpu == t_ref  # the draws do not depend on the storage form
    check_f32(f32_errors(got, ref, 16.0), "C1")


@pytest.fixture(scope="module")
def c2_oracle():
    case = config2(steps=500)
    say("C2 oracle: 128^3, 4096 swimmers, 1e5 tracers, 500 steps")
    ref, _ = run_oracle(O, case)
    return case, ref


@pytest.mark.timeout(600)
def test_c2_full_size_500_steps_bitwise(c2_oracle):
    case, ref = c2_oracle
    say("C2: FP64 HIP, 500 steps")
    got, _ = run_lbswim(case)
    assert_bitwise(got, ref, ALL_KEYS, "C2")


@pytest.mark.timeout(600)
def test_c2_f32_error_against_oracle(c2_oracle):
    case, ref = c2_oracle
    say("C2: FP32 HIP, 500 steps")
    got, _ = run_lbswim(case, "f32")
    check_f32(f32_errors(got, ref, 128.0), "C2")


def c3_case():
    return make_case((256, 256, 256), 200, 0.25, ns=32768, nt=1_000_000, alpha=0.01, seed=1234)


@pytest.mark.timeout(900)
def test_c3_full_size_200_steps_against_oracle():
    """The precision configuration at full size: the FP64 form bit for bit
    with the oracle after 200 steps (tumbles included), the FP32 form within
    the stated bounds of it."""
    case = c3_case()
    say("C3 oracle: 256^3, 32768 swimmers (alpha 0.01), 1e6 tracers, 200 steps")
    ref, t_ref = run_oracle(O, case)
    say("C3: FP64 HIP")
    got, t_gpu = run_lbswim(case, "f64")
    assert t_gpu == t_ref
    assert_bitwise(got, ref, ("fOld", "rho", "u", "sw_r", "sw_n", "sw_prng", "tr_r"), "C3 f64")
    del got
    say("C3: FP32 HIP")
    got, t32 = run_lbswim(case, "f32")
    assert t32 == t_ref
    check_f32(f32_errors(got, ref, 256.0), "C3")


def test_f32_conservation_totals_256_cubed():
    """FP64 totals of the FP32-storage distributions over 200 coupled steps at
    256^3 with 32768 swimmers.  Rounding of the FP32 deviations is unbiased:
    about 1e-11 per site and step, a random walk of ~1e-6 in total mass over
    n = 1.7e7 sites and 200 steps (relative 4e-14), so the bound is relative
    mass drift <= 1e-10; momentum changes only by rounding as well (the
    force dipoles sum to zero), bound 1e-4 against |P| ~ 4."""
    case = make_case((256, 256, 256), 0, 0.25, ns=32768, nt=0, alpha=0.0, seed=11)
    say("FP32 conservation: 256^3, 32768 swimmers, 200 steps")
    lat, system, sw, tr = build_lbswim(case, "f32")
    t0 = np.array(lat.totals())
    system.Step(200)
    t1 = np.array(lat.totals())
    say("FP32 totals before %s after %s: mass drift %.2e (relative), momentum change %s"
        % (t0, t1, abs(t1[0] - t0[0]) / t0[0], t1[1:] - t0[1:]))
    assert abs(t1[0] - t0[0]) / t0[0] <= 1e-10
    assert np.abs(t1[1:] - t0[1:]).max() <= 1e-4


# ------------------------------------------------------------------ C4

def _ensemble_run(storage, cases, total):
    from lbswim.ensemble import Ensemble
    ens = Ensemble(storage)
    built = []
    for c in cases:
        lat, system, sw, tr = build_lbswim(c, storage)
        ens.add(lat, [sw])
        built.append((lat, sw, lat.totals() if storage == "f64" else None))
    choice = ens.autotune(3)
    done = built[0][0].time_step
    say("C4 %s: autotune chose %s after %d steps" % (storage, choice, done))
    assert 0 < done < total
    ens.step(total - done)
    out = []
    for lat, sw, t0 in built:
        assert lat.time_step == total
        out.append((collect_lbswim(lat, sw, None), t0, lat.totals() if storage == "f64" else None))
    del ens, built
    return out


@pytest.mark.timeout(900)
def test_c4_ensemble_at_its_shape():
    import bench
    from lbswim import _core
    total = 1000
    cases = bench.ensemble_cases(64)
    say("C4: %d sims, sizes %s, %d steps" % (len(cases), sorted({c.size[0] for c in cases}), total))
    res = {s: _ensemble_run(s, cases, total) for s in ("f64", "f32")}
    _core.synchronize()
    # every member exactly as alone, in both forms
    for storage in ("f64", "f32"):
        for i, c in enumerate(cases):
            got = res[storage][i][0]
            alone, _ = run_lbswim(c, storage, steps=total)
            assert_bitwise(got, alone, ("fOld", "rho", "u", "force", "sw_r", "sw_n", "sw_v", "sw_prng"),
                           "C4 %s member %d" % (storage, i))
    # FP64 members against the oracle: the first of each size 64, 16, 48, 32
    picks = []
    for L in (64, 16, 48, 32):
        picks.append(next(i for i, c in enumerate(cases) if c.size[0] == L))
    for i in picks:
        say("C4 member %d (%d^3, %d swimmers) vs oracle" % (i, cases[i].size[0], cases[i].sw_r.shape[1]))
        ref, _ = run_oracle(O, cases[i], steps=total)
        assert_bitwise(res["f64"][i][0], ref, ("fOld", "rho", "u", "force", "sw_r", "sw_n", "sw_prng"),
                       "C4 member %d vs oracle" % i)
    worst_u, worst_m = 0.0, 0.0
    for i in range(len(cases)):
        a, t0, t1 = res["f64"][i]
        b = res["f32"][i][0]
        worst_u = max(worst_u, rel_inf(b["u"], a["u"]))
        worst_m = max(worst_m, abs(t1[0] - t0[0]) / t0[0])
    say("C4: worst FP32 u error vs FP64 twin %.2e, worst FP64 mass drift %.2e" % (worst_u, worst_m))
    assert worst_u <= F32_U_TOL
    assert worst_m <= 1e-13


# ------------------------------------------------------------------ C5

def _oracle_case(req):
    from lbswim.serving import common_params, initial_state
    rho, u, sw_r, sw_n, tr_r = initial_state(req)
    return Case((req.size,) * 3, req.tau, req.tau, req.steps, rho, u, sw_r, sw_n, common_params(req), tr_r)


@pytest.mark.timeout(900)
def test_c5_serving_at_its_ranges():
    from lbswim.serving import Server, estimate_bytes, poisson_requests, run_alone
    reqs = poisson_requests(12, rate=16.0, seed=5)  # the serving ranges of the benchmark
    budget = 3 * max(estimate_bytes(r) for r in reqs)
    say("C5: %d requests %s, budget %d bytes"
        % (len(reqs), [(r.size, r.steps, r.swimmers, r.tracers, r.storage) for r in reqs], budget))
    srv = Server(hbm_budget=budget, max_members=64, chunk=100)
    peak = []
    stats = srv.run(reqs, realtime=False, on_chunk=lambda s: peak.append(len(s.active)))
    say("C5 stats %s, peak concurrent %d" % (stats, max(peak)))
    assert stats["completed"] == len(reqs) and stats["rejected"] == 0
    assert max(peak) >= 2
    for r in reqs:
        alone = run_alone(r)
        for k, v in alone.items():
            assert np.array_equal(r.result[k], v), (r.rid, k)
    # the three FP64 requests cheapest for the oracle, bit for bit
    f64 = sorted((r for r in reqs if r.storage == "f64"),
                 key=lambda r: r.steps * (r.sites + 64 * (r.tracers + 2 * r.swimmers)))[:3]
    for r in f64:
        say("C5 request %d (%d^3, %d steps, %d swimmers, %d tracers) vs oracle"
            % (r.rid, r.size, r.steps, r.swimmers, r.tracers))
        ref, _ = run_oracle(O, _oracle_case(r))
        for k in ("fOld", "rho", "u", "sw_r", "sw_n", "tr_r"):
            if k in r.result:
                assert np.array_equal(r.result[k], ref[k]), (r.rid, k)
