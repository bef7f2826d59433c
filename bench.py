"""Benchmark of the coupled lattice-Boltzmann / swimmer / tracer step on one GPU.

Workload (the 256^3 precision configuration): a 256^3 periodic D3Q15 MRT
lattice, tau_s = tau_b = 0.25, rho = 1 + 1e-3 N(0,1), u = 1e-3 N(0,1);
32768 swimmers (uniform positions, isotropic orientations, P = 1e-3,
l = 0.1, alpha = 0.01) and 1e6 uniform tracers, seed 1234.  One "step" is one
System.Step: ZeroForce, AddForces, collide+stream, CalcHydro, swimmer Move,
tracer Move.  All state is resident in HBM before the timed region.

`value` is MLUPS (lattice sites x steps / s / 1e6) of the FP64 drop-in form;
the FP32-storage form is timed on the same inputs and reported beside it.
Prints one JSON line.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--size 256] [--no-cpu]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [os.path.join(ROOT, "gpu-swimmers_amd"), os.path.join(ROOT, "tests"), os.path.join(ROOT, "oracle")]

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, vendor figure
ELEM = {"f64": 8, "f32": 4}


def collide_bytes(form, virtual):
    """Algorithmic bytes per site of one collide+stream launch: 15 f read and
    15 f written, plus 3 force components read unless the force is virtual
    (evaluated from the per-block point lists, which stay in cache)."""
    return (30 + (0 if virtual else 3)) * ELEM[form]


def fused_bytes(form, virtual):
    """Algorithmic bytes per site of one fused collide+stream+CalcHydro
    launch: 15 f read, 15 f written, rho and 3 u written (u component-
    interleaved, the layout the particle moves gather from), plus 3 force
    components read unless the force is virtual."""
    return (34 + (0 if virtual else 3)) * ELEM[form]


def particle_bytes(ns, nt):
    """Particle-array bytes of one coupled step: swimmers read r, n for the
    spreading points (48 B) and for the move (48 B), write r, v, n (72 B) and
    read and write the draw counter (16 B), plus 2 point records of 176 B
    written and read once; tracers read r, s (48 B) and write r, s, v (72 B)."""
    return ns * (48 + 48 + 72 + 16 + 2 * 2 * 176) + nt * (48 + 72)


def gathered_bytes(ns, nt, elem):
    """Stencil gathers of u (64 sites x 3 components) in the moves: two
    stencils per swimmer, one per tracer; mostly L2 hits."""
    return (2 * ns + nt) * 64 * 3 * elem


def step_bytes(form, virtual, fused):
    """Lattice bytes per site of the whole coupled step: collide+stream,
    CalcHydro (15 f read -- not with the fused launch, which takes them from
    the values it has just collided -- rho and 3 u written, 3 force read) and
    the force array written once; the force terms only with a stored force."""
    force = 0 if virtual else (3 if fused else 3 + 3) + 3
    return (30 + (0 if fused else 15) + 4 + force) * ELEM[form]


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def make_workload(size, seed=1234):
    from cases import make_case
    return make_case((size, size, size), 0, 0.25, ns=32768, nt=1_000_000, alpha=0.01, seed=seed)


def time_gpu(case, storage, steps, warmup, force_mode="auto", state_steps=200, fusion=True, tiles=None):
    from cases import build_lbswim
    import lbswim

    lat, system, sw, tr = build_lbswim(case, storage, force_mode)
    tuned = None
    if not fusion:
        lat.set_fusion(False)
    elif tiles is not None:
        # a fixed launch choice (profiling runs: counters distort the tuner's timings)
        lat.set_fusion(tiles[0] > 0, tiles[0], tiles[1])
        log("%s: fixed (rows, strips) = %s" % (storage, tiles))
    else:
        # the autotuner picks two launches or a fused tile shape (and, for
        # FP32, the list or the scatter force) on real steps
        tuned = system.Autotune(5)
        log("%s: autotuned (rows, strips) = %s, force %s" % (storage, tuned, lat.force_mode))
    fused = lat.fusion[0]
    main = "step_hydro" if fused else "collide_stream"  # the dominant lattice kernel
    system.Step(warmup)
    lbswim.synchronize()
    # timed region: HIP events around the dominant kernel's launches only (each
    # event pair costs a few microseconds of device time)
    lat.set_timing([main])
    lbswim.synchronize()
    t0 = time.perf_counter()
    system.Step(steps)
    lbswim.synchronize()
    t1 = time.perf_counter()
    main_ms, main_cnt = lat.kernel_times()[main]
    virtual = lat.force_is_virtual
    # per-kernel breakdown: a second pass with every kernel bracketed by events
    breakdown_steps = min(steps, 20)
    lat.set_timing(True)
    tb0 = time.perf_counter()
    system.Step(breakdown_steps)
    lbswim.synchronize()
    tb1 = time.perf_counter()
    times = lat.kernel_times()
    lat.set_timing(False)
    n = lat.addr.n
    ms_step = (t1 - t0) * 1e3 / steps
    mlups = n * steps / (t1 - t0) / 1e6
    res = {"ms_per_step": ms_step, "mlups": mlups, "kernels": {}, "virtual_force": virtual, "fused": fused,
           "fusion": list(lat.fusion), "breakdown_ms_per_step": (tb1 - tb0) * 1e3 / breakdown_steps}
    for name, (ms, cnt) in times.items():
        if cnt:
            res["kernels"][name] = {"avg_ms": ms / cnt, "launches": cnt}
    avg = main_ms / main_cnt
    # scatter force: the fused launch reads the force the moves scattered and
    # clears the other scatter buffer (3 components read, 3 written per site)
    scatter = lat.force_mode == "scatter" and fused
    res["force_mode"] = lat.force_mode
    if scatter:
        virtual = False
    res["main_kernel"] = "k_%s_%s%s" % (main, "scatter_" if scatter else ("vf_" if virtual else ""), storage)
    res["main_avg_ms"] = avg
    res["main_bytes"] = (fused_bytes if fused else collide_bytes)(storage, virtual) + (3 * ELEM[storage] if scatter else 0)
    res["step_bytes"] = step_bytes(storage, virtual, fused)
    res["main_gbs"] = res["main_bytes"] * n / (avg * 1e-3) / 1e9
    res["step_gbs"] = res["step_bytes"] * n / (ms_step * 1e-3) / 1e9
    res["tumbles"] = sw.tumbles()
    res["autotuned"] = list(tuned) if tuned is not None else None
    choice = (lat.fusion, lat.force_mode)
    del lat, system, sw, tr
    # final state for the FP32-vs-FP64 and oracle error reports: a fresh run of
    # exactly state_steps coupled steps (the configuration's own count) from
    # the same inputs with the launch choice the run above used (the tuner's
    # trial steps differ in number between the forms)
    res["state_steps"] = state_steps
    res["state"] = final_state(case, storage, choice, state_steps)
    return res


def final_state(case, storage, choice, steps):
    from cases import build_lbswim
    (active, rows, strips), force_mode = choice
    lat, system, sw, tr = build_lbswim(case, storage, force_mode)
    lat.set_fusion("always" if active else False, rows if active else 0, strips if active else 0)
    system.Step(steps)
    for name in ("rho", "u"):
        getattr(lat, name).D2H()
    sw.D2H()
    tr.D2H()
    out = {"rho": np.array(lat.rho.data).ravel(), "u": np.array(lat.u.data).ravel(),
           "sw_r": np.array(sw.r.data), "sw_n": np.array(sw.n.data), "tr_r": np.array(tr.r.data),
           "steps": lat.time_step}
    assert out["steps"] == steps, (out["steps"], steps)
    del lat, system, sw, tr
    return out


def precision_report(f64, f32, L):
    """Norm-wise relative errors ||a-b||_inf/||b||_inf of the FP32-storage run
    against the FP64 run of the same inputs and step count.  Trajectory
    errors use the minimum-image displacement (positions are periodic)."""
    def rel(a, b):
        return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))

    def traj(a, b, L):
        d = np.abs(a - b)
        d = np.minimum(d, L - d)
        return float(np.max(d))

    a, b = f32["state"], f64["state"]
    assert a["steps"] == b["steps"], (a["steps"], b["steps"])
    return {"u_rel": rel(a["u"], b["u"]), "drho_rel": rel(a["rho"] - 1.0, b["rho"] - 1.0),
            "swimmer_r_abs": traj(a["sw_r"], b["sw_r"], L), "swimmer_n_abs": float(np.max(np.abs(a["sw_n"] - b["sw_n"]))),
            "tracer_r_abs": traj(a["tr_r"], b["tr_r"], L),
            "note": "after %d coupled steps of both forms from identical inputs" % f64["state_steps"]}


def ensemble_cases(nsims=64, seed=1234):
    """64 independent sims: cubic sizes from {16,24,32,48,64}, tau ~ U[0.2,1],
    1-256 swimmers with alpha ~ U[0,0.05], per-sim seeds 1234+i."""
    from cases import make_case
    rng = np.random.default_rng(seed)
    out = []
    for i in range(nsims):
        L = int(rng.choice([16, 24, 32, 48, 64]))
        out.append(make_case((L, L, L), 0, float(rng.uniform(0.2, 1.0)), ns=int(rng.integers(1, 257)), nt=0,
                             alpha=float(rng.uniform(0.0, 0.05)), seed=seed + i))
    return out


def bench_ensemble(storage, steps, nsims=64, tune_steps=5, graph=True, sequential=True):
    """Batched (one launch per phase, graph replay, autotuned order) against
    the same sims run one after another with the plain per-system step."""
    from cases import build_lbswim
    import lbswim
    from lbswim.ensemble import Ensemble, ORDERS

    cases = ensemble_cases(nsims)
    sites = sum(c.size[0] * c.size[1] * c.size[2] for c in cases)
    # sequential plain runs
    t_seq = 0.0
    for c in (cases if sequential else []):
        lat, system, sw, tr = build_lbswim(c, storage)
        system.Step(2)
        lbswim.synchronize()
        t0 = time.perf_counter()
        system.Step(steps)
        lbswim.synchronize()
        t_seq += time.perf_counter() - t0
        del lat, system, sw, tr
    # batched
    ens = Ensemble(storage)
    ens.set_graph(graph)
    keep = []
    for c in cases:
        lat, system, sw, tr = build_lbswim(c, storage)
        ens.add(lat, [sw])
        keep.append((lat, sw))
    ens.autotune(tune_steps)
    order, per = ens.tuning
    ens.step(2)
    lbswim.synchronize()
    t0 = time.perf_counter()
    ens.step(steps)
    lbswim.synchronize()
    t_ens = time.perf_counter() - t0
    # roofline of the dominant lattice launch (the batched fused step of the
    # members with full 64-site rows, or the batched collide+stream of the
    # other virtual-force members, whichever covers more sites), each phase
    # timed with HIP events in an eager pass
    timed = 20
    ens.set_timing(True)
    ens.step(timed)
    kt = ens.kernel_times()
    ens.set_timing(False)
    vsites, ssites = ens.sites()
    fsites = ens.fused_sites()
    # device time of each phase per step (with lanes: summed over the lanes'
    # launches, which run side by side)
    phases = {k: round(ms / timed, 4) for k, (ms, cnt) in kt.items() if cnt}
    roof = None

    def gbs(byts, ms):
        return round(byts / (ms * 1e-3) / 1e9, 1)

    f_ms, f_n = kt["step_hydro"]
    c_ms, c_n = kt["collide_stream"]
    h_ms, h_n = kt["calc_hydro"]
    f_n, c_n, h_n = (timed if x else 0 for x in (f_n, c_n, h_n))  # per-step sums below
    if f_n and fsites >= vsites:
        avg = f_ms / f_n
        byts = fsites * fused_bytes(storage, True)
        roof = {"kernel": "kb_step_hydro<%s, virtual force>" % storage, "bound": "hbm", "achieved": gbs(byts, avg),
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(gbs(byts, avg) / HBM_PEAK_GBS, 4),
                "traffic": load_ensemble_traffic(storage, "kb_step_hydro<%s, virtual force>" % storage,
                                                 list(ens.fusion)) if sites == 6069760 else None,
                "algorithmic_bytes_per_step": byts, "launch_ms_per_step": round(avg, 4), "lanes": ens.lanes}
        if c_n:
            roof["collide_ms_per_step"] = round(c_ms / c_n, 4)
            roof["collide_gbs"] = gbs(vsites * collide_bytes(storage, True), c_ms / c_n)
    elif c_n:
        avg = c_ms / c_n
        byts = vsites * collide_bytes(storage, True)
        kname = "kb_collide<%s, virtual force>" % storage
        roof = {"kernel": kname, "bound": "hbm", "achieved": gbs(byts, avg), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(gbs(byts, avg) / HBM_PEAK_GBS, 4),
                "traffic": load_ensemble_traffic(storage, kname) if sites == 6069760 else None,
                "algorithmic_bytes_per_step": byts, "launch_ms_per_step": round(avg, 4), "lanes": ens.lanes,
                "calc_hydro_ms_per_step": round(h_ms / h_n, 4) if h_n else None,
                "calc_hydro_gbs": gbs(vsites * (19 * ELEM[storage]), h_ms / h_n) if h_n else None}
    # main-stream phases of one eager step (the stored-force chain runs beside
    # them on the side stream) against the replayed graphs' time per step
    main = sum(v for k, v in phases.items() if k != "stored_chain")
    return {"sims": nsims, "steps": steps, "sites": sites, "ms_batched": round(t_ens * 1e3, 2),
            "ms_sequential": round(t_seq * 1e3, 2) if t_seq else None,
            "sims_per_s": round(nsims / t_ens, 2),
            "sims_per_s_sequential": round(nsims / t_seq, 2) if t_seq else None,
            "speedup_vs_sequential": round(t_seq / t_ens, 2) if t_seq else None,
            "mlups": round(sites * steps / t_ens / 1e6, 1), "order": ORDERS[order], "blocks_per_workgroup": per,
            "lanes": ens.lanes,
            "fusion_rows_strip": list(ens.fusion), "sites_fused_virtual_stored": [fsites, vsites, ssites],
            "graph": graph, "roofline": roof, "phases_ms": phases,
            "graph_ms_per_step": round(t_ens * 1e3 / steps, 4), "main_phases_ms_per_step": round(main, 4)}


def bench_serving(nreq, rate, seed=1234):
    """Poisson stream of requests (cube size U[24,96], steps log-uniform in
    [200,5000], 16-4096 swimmers, 0-1e5 tracers, FP32/FP64 50/50) served with
    continuous batching under an HBM budget, then the same stream served one
    request at a time with the plain step."""
    from lbswim.serving import Server, poisson_requests, run_sequential
    reqs = poisson_requests(nreq, rate, seed)
    srv = Server(hbm_budget=64 << 30, max_members=64, chunk=100, keep_results=False)
    count = {"n": 0}

    def progress(label):
        def on_done(req):
            count["n"] += 1
            if count["n"] % 8 == 0 or count["n"] == nreq:
                log("serving %s: %d of %d requests done" % (label, count["n"], nreq))
        return on_done

    batched = srv.run(reqs, realtime=True, on_done=progress("batched"))
    srv.close()
    del srv
    count["n"] = 0
    seq = run_sequential(poisson_requests(nreq, rate, seed), realtime=True, on_done=progress("sequential"))
    mix = serving_mix_bound(poisson_requests(nreq, rate, seed))
    r = lambda d: {k: (round(v, 4) if isinstance(v, float) else v) for k, v in d.items()}  # noqa: E731
    return {"rate_per_s": rate, "batched": r(batched), "sequential": r(seq),
            "throughput_ratio": round(batched["mlups"] / seq["mlups"], 2),
            "p99_ratio": round(seq["p99_s"] / batched["p99_s"], 2),
            "mix_resident_mlups": round(mix, 1), "batched_vs_mix_resident": round(batched["mlups"] / mix, 3),
            # the stream's own floor: each request started at its arrival and run
            # as fast as it runs alone (sequential leg's service times)
            "makespan_floor_s": round(seq["floor_s"], 4),
            "batched_wall_vs_floor": round(seq["floor_s"] / batched["wall_s"], 3)}


def serving_mix_bound(reqs, steps=40):
    """Throughput of the same requests when all of them are resident at once
    and stepped together (two ensembles, graph replay): what the device gives
    this mix of lattice sizes and particle loads with no arrivals,
    admissions or collections around it."""
    from lbswim.ensemble import Ensemble
    from lbswim.serving import build
    import lbswim
    ens = {"f64": Ensemble("f64"), "f32": Ensemble("f32")}
    keep = []
    sites = 0
    for req in reqs:
        lat, things = build(req)
        ens[req.storage].add(lat, things)
        keep.append((lat, things))
        sites += req.sites
    for e in ens.values():
        e.step(2)
    lbswim.synchronize()
    t0 = time.perf_counter()
    for e in ens.values():
        e.step(steps)
    lbswim.synchronize()
    return sites * steps / (time.perf_counter() - t0) / 1e6


def oracle_errors(got, ref, L):
    """FP32-storage state against the FP64 oracle state of the same inputs and
    step count: ||a-b||_inf/||b||_inf for u and rho - 1, minimum-image
    absolute error for positions, absolute for orientations."""
    def rel(a, b):
        return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))

    def traj(a, b):
        d = np.abs(a - b)
        return float(np.max(np.minimum(d, L - d))) if d.size else 0.0

    out = {"u_rel": rel(got["u"], ref["u"]), "drho_rel": rel(got["rho"] - 1.0, ref["rho"] - 1.0)}
    if ref["sw_r"].size:
        out["swimmer_r_abs"] = traj(got["sw_r"], ref["sw_r"])
        out["swimmer_n_abs"] = float(np.max(np.abs(got["sw_n"] - ref["sw_n"])))
    if ref["tr_r"].size:
        out["tracer_r_abs"] = traj(got["tr_r"], ref["tr_r"])
    return out


def small_config_states(storage):
    """Final states of the 16^3 (1000 steps, alpha 0.01) and 128^3 (500
    steps) configurations run with the given storage form."""
    from cases import config1, config2, run_lbswim
    out = {}
    for name, case in (("C1", config1(1000, alpha=0.01)), ("C2", config2(500))):
        got, _ = run_lbswim(case, storage)
        out[name] = {k: got[k] for k in ("rho", "u", "sw_r", "sw_n", "tr_r")}
    return out


def _oracle_setup(O, case):
    lat = O.Lattice(*case.size, case.tau_s, case.tau_b)
    lat.rho[:] = case.rho
    lat.u[:] = case.u
    lat.InitFromHydro()
    sw = tr = None
    if case.sw_r.shape[1]:
        sw = O.Swimmers(case.sw_r.shape[1], O.Common(**case.common))
        sw.r[:] = case.sw_r
        sw.n[:] = case.sw_n
    if case.tr_r.shape[1]:
        tr = O.Tracers(case.tr_r.shape[1])
        tr.r[:] = case.tr_r
    return lat, sw, tr


def _oracle_state(lat, sw, tr):
    e3 = np.zeros((3, 0))
    return {"rho": lat.rho.copy(), "u": lat.u.copy(), "sw_r": sw.r.copy() if sw else e3,
            "sw_n": sw.n.copy() if sw else e3, "tr_r": tr.r.copy() if tr else e3}


def oracle_legs(case, f64_state, f32_state, f32_small, budget_s=240.0):
    """The FP64 oracle, built for this host (-O3 -march=native, OpenMP over
    its cores), on the benchmark workload and the two small configurations.

    - cpu_baseline: 256^3 coupled steps, 10 warm-up steps, then the median of
      3 timed runs of k steps (k chosen for about 5 s each);
    - the same run continued to the step count of the GPU states gives the
      FP64 reference for the FP32 error at 256^3 (and is compared bit for bit
      with the FP64 GPU state);
    - 16^3 x 1000 and 128^3 x 500 steps give the FP32 errors of those
      configurations.
    The 256^3 continuation is skipped if it would take more than budget_s."""
    import statistics
    import tempfile
    import lbswim_oracle as O
    build = "oracle built -O3 -march=native on this host"
    try:
        O.use_library(O.build_native(os.path.join(tempfile.mkdtemp(), "liblbswim_oracle_native.so")))
    except Exception as exc:  # no compiler on the host: the prebuilt oracle
        build = "prebuilt oracle (-O3; native build failed: %s)" % exc
    cores = int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count() or 1
    cores = min(cores, os.cpu_count() or cores)
    O.set_threads(cores)
    lat, sw, tr = _oracle_setup(O, case)
    t0 = time.perf_counter()
    O.run(lat, 10, sw, tr)
    per = (time.perf_counter() - t0) / 10
    k = int(max(1, min(60, round(5.0 / per))))
    runs = []
    for _ in range(3):
        t0 = time.perf_counter()
        O.run(lat, k, sw, tr)
        runs.append(time.perf_counter() - t0)
    med = statistics.median(runs)
    res = {"cpu_baseline": {
        "value": lat.n * k / med / 1e6, "unit": "MLUPS", "cores": cores, "kind": "port",
        "sample": "median of 3 runs of %d coupled steps (%.2f/%.2f/%.2f s) after 10 warm-up steps, 256^3 workload "
                  "(32768 swimmers, 1e6 tracers), FP64 oracle, %s, OpenMP over %d threads"
                  % (k, runs[0], runs[1], runs[2], build, cores)}}
    log("CPU baseline: %.1f MLUPS (%d threads)" % (res["cpu_baseline"]["value"], cores))
    if f32_small is None:
        return res
    errs = {}
    target = f64_state["steps"]
    assert f32_state["steps"] == target, (f32_state["steps"], target)
    left = target - (10 + 3 * k)
    if left >= 0 and left * med / k <= budget_s:
        log("oracle: continuing the 256^3 run to %d steps" % target)
        O.run(lat, left, sw, tr)
        ref = _oracle_state(lat, sw, tr)
        L = float(case.size[0])
        errs["C3"] = oracle_errors(f32_state, ref, L)
        res["f64_equals_oracle_256cubed"] = all(np.array_equal(f64_state[key], ref[key]) for key in ref)
    else:
        errs["C3"] = None
    del lat, sw, tr
    from cases import config1, config2
    for name, c, L in (("C1", config1(1000, alpha=0.01), 16.0), ("C2", config2(500), 128.0)):
        log("oracle: %s" % name)
        lat, sw, tr = _oracle_setup(O, c)
        O.run(lat, c.steps, sw, tr)
        errs[name] = oracle_errors(f32_small[name], _oracle_state(lat, sw, tr), L)
        del lat, sw, tr
    res["fp32_error_vs_fp64_oracle"] = {
        "C1": errs["C1"], "C2": errs["C2"], "C3": errs["C3"],
        "note": "FP32-storage state against the FP64 oracle after 1000 (16^3, alpha 0.01), 500 (128^3) and "
                "%d (256^3, alpha 0.01) coupled steps from identical inputs" % target}
    return res


def probe_dense(steps=40):
    """Coupled steps of dense small systems, the kind the serving mix holds
    (cube side, swimmers; FP64, no tracers, the serving leg's synthetic state):
    ms per step and of its swimmer-force phase (HIP events)."""
    from lbswim.serving import Request, build
    from lbswim.System import System
    import lbswim
    out = []
    for size, ns in ((24, 3710), (32, 3710), (58, 3547)):
        lat, things = build(Request(rid=0, size=size, steps=steps, swimmers=ns, tracers=0, storage="f64", seed=7))
        system = System(lat)
        for t in things:
            system.AddThing(t)
        system.Step(5)
        lbswim.synchronize()
        t0 = time.perf_counter()
        system.Step(steps)
        lbswim.synchronize()
        ms = (time.perf_counter() - t0) * 1e3 / steps
        lat.set_timing(["add_forces"])
        system.Step(10)
        lbswim.synchronize()
        fms, cnt = lat.kernel_times()["add_forces"]
        lat.set_timing(False)
        out.append({"side": size, "swimmers": ns, "force_mode": lat.force_mode, "ms_per_step": round(ms, 4),
                    "force_ms_per_step": round(fms / max(cnt, 1), 4)})
        log("dense %d^3 with %d swimmers: %.4f ms/step" % (size, ns, ms))
        del lat, things, system
    return out


def probe_mix(nreq=48, rate=16.0, steps=40):
    """The serving leg's requests (same seed) all resident in their two
    ensembles: ms per graph-replayed step of each ensemble, then each phase's
    device time per step in an eagerly launched, event-timed pass."""
    from lbswim.ensemble import Ensemble
    from lbswim.serving import build, poisson_requests
    import lbswim
    reqs = poisson_requests(nreq, rate, 1234)
    ens = {"f64": Ensemble("f64"), "f32": Ensemble("f32")}
    keep = []
    for req in reqs:
        lat, things = build(req)
        ens[req.storage].add(lat, things)
        keep.append((lat, things))
    out = {}
    for form, e in ens.items():
        e.step(2)
        lbswim.synchronize()
        t0 = time.perf_counter()
        e.step(steps)
        lbswim.synchronize()
        ms = (time.perf_counter() - t0) * 1e3 / steps
        e.set_timing(True)
        e.step(20)
        kt = e.kernel_times()
        e.set_timing(False)
        vs, ss = e.sites()
        out[form] = {"members": sum(1 for r in reqs if r.storage == form), "sites_virtual_stored": [vs, ss],
                     "fused_sites": e.fused_sites(), "graph_ms_per_step": round(ms, 4),
                     "phases_ms": {k: round(v / 20, 4) for k, (v, c) in kt.items() if c}}
        log("mix %s: %.4f ms per graph step" % (form, ms))
    return out


def load_pmc(name, kernel, tiles=None):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 --pmc
    summary profiles/<name>_pmc.json, or None when there is none or it was
    measured on another launch shape (tiles: the fused launch's rows and
    strips, or rows and strip planes in an ensemble)."""
    try:
        with open(os.path.join(ROOT, "profiles", name + "_pmc.json")) as fh:
            d = json.load(fh)
    except (OSError, ValueError):
        return None
    if d.get("kernel") != kernel:
        return None
    if tiles is not None and d.get("tiles") is not None and list(d["tiles"]) != list(tiles):
        return None
    return d.get("hbm_bytes_per_launch")


def load_ensemble_traffic(storage, kernel, tiles=None):
    """HBM bytes per launch of the ensemble's batched fused step
    (kb_step_hydro) or batched collide (kb_collide)."""
    return load_pmc("%s_%s_ensemble" % (kernel.split("<")[0], storage), kernel, tiles)


def load_traffic(kernel, tiles=None):
    """HBM bytes per launch of `kernel` (k_<name>_[vf_]<form>) at 256^3."""
    return load_pmc("%s_256cubed" % kernel[2:].replace("vf_", ""), kernel, tiles)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--size", type=int, default=256)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--forms", default="f64,f32")
    ap.add_argument("--force-mode", default="auto", choices=["auto", "stored", "virtual"],
                    help="swimmer force of the coupled step: virtual where sparse (auto), stored, virtual")
    ap.add_argument("--fusion", default="on", choices=["on", "off"],
                    help="collide/stream and CalcHydro as one launch where the lattice allows (on) or two launches")
    ap.add_argument("--tiles", default="auto",
                    help="fused tile shape 'rows,strips' of the coupled step (rows 0: two launches), or auto "
                         "(autotuned on real steps)")
    ap.add_argument("--probe", choices=["copy", "dense", "mix"],
                    help="run the streaming-copy bandwidth probe, time dense small systems' coupled steps, or time "
                         "the serving leg's requests all resident, instead")
    ap.add_argument("--ensemble-steps", type=int, default=1000, help="steps of the 64-sim ensemble leg (0: skip)")
    ap.add_argument("--serve-requests", type=int, default=48, help="requests of the serving leg (0: skip)")
    ap.add_argument("--ensemble-eager", action="store_true",
                    help="launch the ensemble's steps eagerly instead of replaying captured graphs")
    ap.add_argument("--ensemble-no-sequential", action="store_true",
                    help="skip the one-after-another baseline of the ensemble leg (profiling runs)")
    ap.add_argument("--serve-rate", type=float, default=16.0, help="Poisson arrival rate of the serving leg (1/s)")
    ap.add_argument("--legs", default="main,ensemble,serving",
                    help="legs to run: main (the 256^3 coupled step; required for the JSON line's value), "
                         "ensemble, serving")
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py measures one GPU only: there is no multi-GPU path (a lattice would be split into "
              "x-slabs exchanging one plane of 5 distributions per face each step; ensembles would assign "
              "whole simulations to GPUs); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    if a.probe == "copy":
        from lbswim import _core
        count = 1 << 28  # 2 GiB per buffer, 4 GiB of traffic per copy
        ms = _core.diag_copy(count, 20)
        gbs = 2 * 8 * count / (ms * 1e-3) / 1e9
        print(json.dumps({"probe": "copy", "doubles": count, "bytes_per_copy": 16 * count, "avg_ms": round(ms, 4),
                          "GBps": round(gbs, 1), "frac_of_peak": round(gbs / HBM_PEAK_GBS, 4)}), flush=True)
        return
    if a.probe == "dense":
        print(json.dumps({"probe": "dense", "systems": probe_dense()}), flush=True)
        return
    if a.probe == "mix":
        print(json.dumps({"probe": "mix", "ensembles": probe_mix(a.serve_requests, a.serve_rate)}), flush=True)
        return
    legs = {x for x in a.legs.split(",") if x}
    if "main" not in legs:  # partial runs (profiling): no headline value
        out = {"legs": sorted(legs)}
        if "ensemble" in legs and a.ensemble_steps > 0:
            out["ensemble"] = {}
            for form in [f for f in a.forms.split(",") if f]:
                log("ensemble %s: 64 sims x %d steps" % (form, a.ensemble_steps))
                out["ensemble"][form] = bench_ensemble(form, a.ensemble_steps, graph=not a.ensemble_eager,
                                                       sequential=not a.ensemble_no_sequential)
        if "serving" in legs and a.serve_requests > 0:
            out["serving"] = bench_serving(a.serve_requests, a.serve_rate)
        print(json.dumps(out), flush=True)
        return
    log("building workload %d^3" % a.size)
    case = make_workload(a.size)
    forms = [f for f in a.forms.split(",") if f]
    res = {}
    for form in forms:
        log("timing %s: %d warm-up + %d steps" % (form, a.warmup, a.steps))
        tiles = None if a.tiles == "auto" else tuple(int(x) for x in a.tiles.split(","))
        res[form] = time_gpu(case, form, a.steps, a.warmup, a.force_mode, fusion=a.fusion == "on", tiles=tiles)
        log("%s: %.3f ms/step, %.0f MLUPS, %s %.0f GB/s" % (form, res[form]["ms_per_step"], res[form]["mlups"],
                                                            res[form]["main_kernel"], res[form]["main_gbs"]))
    main_form = "f64" if "f64" in res else forms[0]
    m = res[main_form]
    n = a.size ** 3
    traffic = load_traffic(m["main_kernel"], m["fusion"][1:] if m["fused"] else None)
    out = {
        "metric": "MLUPS of the coupled step (ZeroForce+AddForces+collide/stream+CalcHydro+Move) at %d^3, "
                  "FP64 drop-in" % a.size,
        "value": round(m["mlups"], 2),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(m["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded PCG64 inputs of the configuration's shape)",
        "config": {"workload": "256^3 D3Q15 MRT, tau 0.25, 32768 swimmers (alpha 0.01), 1e6 tracers"
                   if a.size == 256 else "%d^3 reduced" % a.size,
                   "lattice": [a.size] * 3, "swimmers": 32768, "tracers": 1_000_000, "parallelism": "single GPU"},
        "roofline": {
            "kernel": m["main_kernel"],
            "bound": "hbm",
            "achieved": round(m["main_gbs"], 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(m["main_gbs"] / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "algorithmic_bytes_per_launch": m["main_bytes"] * n,
            "virtual_force": m["virtual_force"],
            "fusion": m["fusion"],
            "autotuned_rows_strips": m["autotuned"],
            "avg_launch_ms": round(m["main_avg_ms"], 4),
        },
        "step_hbm_frac": round(m["step_gbs"] / HBM_PEAK_GBS, 4),
        "step_bytes_per_site": m["step_bytes"],
        "particle_bytes_per_step": particle_bytes(32768, 1_000_000),
        "gathered_bytes_per_step": gathered_bytes(32768, 1_000_000, ELEM[main_form]),
        "kernels_ms": {k: round(v["avg_ms"], 4) for k, v in m["kernels"].items()},
        "kernels_note": "per-kernel averages from a second pass of min(steps, 20) steps with every kernel "
                        "bracketed by HIP events (%.4f ms/step with those events)" % m["breakdown_ms_per_step"],
    }
    if "f32" in res and "f64" in res:
        out["fp32_error_vs_fp64"] = precision_report(res["f64"], res["f32"], float(a.size))
    if "f32" in res and main_form != "f32":
        f = res["f32"]
        out["fp32_store"] = {"mlups": round(f["mlups"], 2), "ms_per_step": round(f["ms_per_step"], 4),
                             "speedup_vs_fp64": round(f["mlups"] / m["mlups"], 3), "force_mode": f["force_mode"],
                             "fusion": f["fusion"], "autotuned_rows_strips": f["autotuned"],
                             "kernel": f["main_kernel"], "kernel_gbs": round(f["main_gbs"], 1),
                             "kernel_frac": round(f["main_gbs"] / HBM_PEAK_GBS, 4),
                             "step_hbm_frac": round(f["step_gbs"] / HBM_PEAK_GBS, 4),
                             "kernels_ms": {k: round(v["avg_ms"], 4) for k, v in f["kernels"].items()}}
        if m["fused"] == f["fused"]:
            # the lattice pass alone (collide+stream+CalcHydro launch, timed in
            # the timed region of each form); the particle work around it is
            # FP64 in both forms
            out["fp32_store"]["lattice_pass_speedup_vs_fp64"] = round(m["main_avg_ms"] / f["main_avg_ms"], 3)
    small32 = None
    if "f32" in res and "f64" in res and not a.no_cpu and a.size == 256:
        log("FP32 runs of the 16^3 and 128^3 configurations")
        small32 = small_config_states("f32")
    if "ensemble" in legs and a.ensemble_steps > 0:
        out["ensemble"] = {}
        for form in forms:
            log("ensemble %s: 64 sims x %d steps, batched vs sequential" % (form, a.ensemble_steps))
            out["ensemble"][form] = bench_ensemble(form, a.ensemble_steps, graph=not a.ensemble_eager,
                                                       sequential=not a.ensemble_no_sequential)
            log("ensemble %s: %s" % (form, out["ensemble"][form]))
    if "serving" in legs and a.serve_requests > 0:
        log("serving: %d requests at %.1f/s, continuous batching vs sequential" % (a.serve_requests, a.serve_rate))
        out["serving"] = bench_serving(a.serve_requests, a.serve_rate)
        log("serving: %s" % out["serving"])
    if not a.no_cpu:
        log("CPU baseline and FP64 oracle references")
        if small32 is not None:
            out.update(oracle_legs(case, res["f64"]["state"], res["f32"]["state"], small32))
        else:
            out.update(oracle_legs(case, res[main_form]["state"], None, None))
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
