"""GPU tests of batched ensembles: every member must evolve exactly as it
does alone (lbswim_system_step), whatever the batch around it, the launch
order chosen by the autotuner, graph replay or eager launches."""
import sys

import numpy as np
import pytest

from cases import build_lbswim, collect_lbswim, make_case

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def uneven_cases(k, seed=0, sizes=(16, 24, 32, 48), steps=0):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(k):
        L = int(rng.choice(sizes))
        dims = (L, L, L) if i % 3 else (L, L // 2 + 3, L + 5)  # some non-cubic, odd-sized
        ns = int(rng.integers(1, 65))
        nt = int(rng.integers(0, 300))
        tau = float(rng.uniform(0.2, 1.0))
        alpha = float(rng.uniform(0.0, 0.05))
        out.append(make_case(dims, steps, tau, ns=ns, nt=nt, alpha=alpha, seed=1234 + i + 100 * seed))
    return out


def build_member(case, storage):
    lat, system, sw, tr = build_lbswim(case, storage)
    things = [t for t in (sw, tr) if t is not None]
    return lat, system, sw, tr, things


def run_alone(case, storage, steps):
    lat, system, sw, tr, _ = build_member(case, storage)
    system.Step(steps)
    return collect_lbswim(lat, sw, tr)


def assert_same(a, b, label):
    for k in a:
        assert np.array_equal(a[k], b[k]), (label, k)


@pytest.mark.parametrize("storage,graph,steps,force_mode", [("f64", True, 9, "auto"), ("f64", False, 4, "auto"),
                                                            ("f32", True, 8, "auto"), ("f64", True, 7, "virtual"),
                                                            ("f64", False, 3, "stored"), ("f32", True, 5, "virtual")])
def test_members_match_alone(storage, graph, steps, force_mode):
    from lbswim.ensemble import Ensemble
    cases = uneven_cases(8, seed=1)
    say("ensemble vs alone: %s graph=%s steps=%d force %s sizes=%s" % (storage, graph, steps, force_mode,
                                                                        [c.size for c in cases]))
    ens = Ensemble(storage)
    ens.set_graph(graph)
    ens.set_force_mode(force_mode)
    built = []
    for c in cases:
        lat, system, sw, tr, things = build_member(c, storage)
        ens.add(lat, things)
        built.append((lat, sw, tr))
    ens.step(steps)
    for i, (c, (lat, sw, tr)) in enumerate(zip(cases, built)):
        assert lat.time_step == steps
        assert_same(collect_lbswim(lat, sw, tr), run_alone(c, storage, steps), "member %d" % i)


def test_membership_changes_and_direct_use():
    from lbswim.ensemble import Ensemble
    cases = uneven_cases(5, seed=2)
    say("ensemble membership changes")
    ens = Ensemble("f64")
    built = []
    for c in cases[:4]:
        lat, system, sw, tr, things = build_member(c, "f64")
        built.append((ens.add(lat, things), lat, system, sw, tr))
    ens.step(3)
    ens.remove(built[1][0])           # member 1 leaves after 3 steps
    lat, system, sw, tr, things = build_member(cases[4], "f64")
    mid = ens.add(lat, things)        # a newcomer takes the free slot
    assert mid == built[1][0]
    ens.step(4)
    built[0][2].Step(2)               # member 0 also stepped on its own
    ens.step(1)
    expect = {0: 3 + 4 + 2 + 1, 2: 8, 3: 8}
    for j, steps in expect.items():
        _, L, _, S, T = built[j]
        assert L.time_step == steps
        assert_same(collect_lbswim(L, S, T), run_alone(cases[j], "f64", steps), "member %d" % j)
    assert lat.time_step == 5
    assert_same(collect_lbswim(lat, sw, tr), run_alone(cases[4], "f64", 5), "newcomer")
    _, L1, _, S1, T1 = built[1]
    assert L1.time_step == 3
    assert_same(collect_lbswim(L1, S1, T1), run_alone(cases[1], "f64", 3), "removed member")


def test_autotune_keeps_results():
    from lbswim.ensemble import Ensemble
    cases = uneven_cases(6, seed=3)
    say("ensemble autotune")
    ens = Ensemble("f64")
    built = []
    for c in cases:
        lat, system, sw, tr, things = build_member(c, "f64")
        ens.add(lat, things)
        built.append((lat, sw, tr))
    # 3 orders x 3 work-item sizes, then 1, 2 and 4 lanes, x (2 + 3) real
    # steps (no member of these sizes takes the fused step: no tile trials)
    choice = ens.autotune(3)
    assert choice in (0, 1, 2)
    order, per = ens.tuning
    assert order == choice and per in (1, 2, 4)
    assert built[0][0].time_step == (9 + 3) * (2 + 3)
    ens.step(5)
    total = (9 + 3) * (2 + 3) + 5
    for c, (lat, sw, tr) in zip(cases, built):
        assert lat.time_step == total
        assert_same(collect_lbswim(lat, sw, tr), run_alone(c, "f64", total), "after autotune")


@pytest.mark.parametrize("per", [3, 8])
def test_work_items_of_several_blocks(per):
    """Batched launches whose workgroups each take `per` consecutive lattice
    blocks of a member (the autotuner's work-item size) give the same
    results as running each member alone."""
    from lbswim.ensemble import Ensemble
    cases = uneven_cases(6, seed=4)
    say("ensemble with %d blocks per work item" % per)
    ens = Ensemble("f64")
    ens.set_work(per)
    built = []
    for c in cases:
        lat, system, sw, tr, things = build_member(c, "f64")
        ens.add(lat, things)
        built.append((lat, sw, tr))
    ens.step(5)
    assert ens.tuning[1] == per
    for c, (lat, sw, tr) in zip(cases, built):
        assert_same(collect_lbswim(lat, sw, tr), run_alone(c, "f64", 5), "per %d" % per)


# Members the batched fused step takes (virtual force, nz % 64 == 0, rows
# dividing ny) beside members it does not (nz not a multiple of 64, a dense
# stored-force member, ny not a multiple of 8).
# four members the batched fused step advances (virtual force: at most 8
# mean list entries), a virtual-force member of the two launches (16^3) and a
# stored-force one
FUSED_CASES = [((8, 16, 64), 30, 200), ((5, 8, 128), 16, 0), ((6, 24, 64), 10, 100), ((4, 12, 64), 10, 50),
               ((16, 16, 16), 3, 60), ((12, 12, 24), 200, 0)]


@pytest.mark.parametrize("storage,graph,fusion", [("f64", True, (8, 32)), ("f64", False, (8, 3)),
                                                  ("f64", True, (4, 5)), ("f32", True, (8, 32)),
                                                  ("f32", False, (4, 64)), ("f64", True, (0, 32))])
def test_fused_members_match_alone(storage, graph, fusion):
    from lbswim.ensemble import Ensemble
    cases = [make_case(size, 0, 0.3 + 0.1 * i, ns=ns, nt=nt, alpha=0.02, seed=77 + i)
             for i, (size, ns, nt) in enumerate(FUSED_CASES)]
    say("ensemble with fused members: %s graph=%s fusion=%s" % (storage, graph, fusion))
    ens = Ensemble(storage)
    ens.set_graph(graph)
    ens.set_fusion(*fusion)
    built = []
    for c in cases:
        lat, system, sw, tr, things = build_member(c, storage)
        ens.add(lat, things)
        built.append((lat, system, sw, tr))
    steps = 7
    ens.step(steps)
    assert ens.fusion == fusion
    fused = ens.fused_sites()
    want = sum(c.size[0] * c.size[1] * c.size[2] for c in cases[:4] if fusion[0] and c.size[1] % fusion[0] == 0)
    assert fused == want, (fused, want)
    for i, (c, (lat, system, sw, tr)) in enumerate(zip(cases, built)):
        assert lat.time_step == steps
        assert_same(collect_lbswim(lat, sw, tr), run_alone(c, storage, steps), "member %d" % i)
    # the members go on alone from the ensemble's state (u written interleaved
    # by the fused launch, the force array as the call left it)
    lat, system, sw, tr = built[0]
    system.Step(2)
    assert_same(collect_lbswim(lat, sw, tr), run_alone(cases[0], storage, steps + 2), "member 0 continued")


@pytest.mark.parametrize("storage,graph,lanes", [("f64", True, 2), ("f64", False, 3), ("f32", True, 2),
                                                 ("f64", True, 8)])
def test_lanes_match_alone(storage, graph, lanes):
    """Members split over lanes (groups stepped on streams of their own) --
    the fused, two-launch and stored-force members of FUSED_CASES together
    with uneven ones -- evolve exactly as alone; a membership change and a
    change of the lane count between calls keep them so."""
    from lbswim.ensemble import Ensemble
    cases = [make_case(size, 0, 0.3 + 0.1 * i, ns=ns, nt=nt, alpha=0.02, seed=177 + i)
             for i, (size, ns, nt) in enumerate(FUSED_CASES)] + uneven_cases(4, seed=5)
    say("ensemble lanes: %s graph=%s lanes=%d, %d members" % (storage, graph, lanes, len(cases)))
    ens = Ensemble(storage)
    ens.set_graph(graph)
    ens.set_lanes(lanes)
    built = []
    for c in cases:
        lat, system, sw, tr, things = build_member(c, storage)
        built.append((lat, sw, tr, ens.add(lat, things)))
    ens.step(5)
    assert ens.lanes == lanes
    # drop one member, step the rest with one lane fewer
    ens.remove(built[2][3])
    ens.set_lanes(max(1, lanes - 1))
    ens.step(4)
    for i, (c, (lat, sw, tr, _)) in enumerate(zip(cases, built)):
        steps = 5 if i == 2 else 9
        assert lat.time_step == steps
        assert_same(collect_lbswim(lat, sw, tr), run_alone(c, storage, steps), "member %d" % i)


def test_autotune_chooses_lanes_and_keeps_results():
    from lbswim.ensemble import Ensemble
    cases = uneven_cases(6, seed=9)
    ens = Ensemble("f64")
    built = []
    for c in cases:
        lat, system, sw, tr, things = build_member(c, "f64")
        ens.add(lat, things)
        built.append((lat, sw, tr))
    ens.autotune(2)
    say("ensemble autotune: order/work %s fusion %s lanes %d" % (ens.tuning, ens.fusion, ens.lanes))
    assert ens.lanes in (1, 2, 4)
    ens.step(3)
    total = built[0][0].time_step
    for i, (c, (lat, sw, tr)) in enumerate(zip(cases, built)):
        assert lat.time_step == total
        assert_same(collect_lbswim(lat, sw, tr), run_alone(c, "f64", total), "member %d" % i)


@pytest.mark.parametrize("storage,graph,force_mode", [("f64", True, "auto"), ("f32", False, "auto"),
                                                      ("f64", False, "stored")])
def test_dense_members_by_site_match_alone_and_oracle(storage, graph, force_mode):
    """Dense members (more than 1024 swimmer ends on lattices with every side
    >= 4) write their force from per-site segments numbered across the batch,
    beside sparse members with lists and a dense member of a thin lattice
    (side 3) that keeps the all-points or list path: every member equals its
    run alone, and in FP64 the dense ones equal the oracle bit for bit."""
    import os

    import lbswim_oracle as O
    from cases import run_oracle
    from lbswim.ensemble import Ensemble
    O.set_threads(min(16, os.cpu_count() or 1))
    cases = [make_case((24, 24, 24), 4, 0.3, ns=2000, nt=100, seed=11, drho=1e-2, du=1e-2, P=1e-2, l=0.3),
             make_case((20, 26, 17), 4, 0.4, ns=700, nt=0, seed=12, drho=1e-2, du=1e-2, P=1e-2, l=1.2),
             make_case((32, 32, 32), 4, 0.5, ns=20, nt=200, seed=13),
             make_case((3, 30, 28), 4, 0.3, ns=2500, nt=0, seed=14, drho=1e-2, du=1e-2, P=1e-2, l=0.2),
             make_case((16, 16, 16), 4, 0.3, ns=1500, nt=50, seed=15, drho=1e-2, du=1e-2, P=1e-2, l=0.5)]
    steps = 4
    say("dense members by site: %s graph=%s force %s" % (storage, graph, force_mode))
    ens = Ensemble(storage)
    ens.set_graph(graph)
    ens.set_force_mode(force_mode)
    built = []
    for c in cases:
        lat, system, sw, tr, things = build_member(c, storage)
        ens.add(lat, things)
        built.append((lat, sw, tr))
    ens.step(steps)
    for i, (c, (lat, sw, tr)) in enumerate(zip(cases, built)):
        got = collect_lbswim(lat, sw, tr)
        assert_same(got, run_alone(c, storage, steps), "member %d" % i)
        if storage == "f64" and i in (0, 1, 4):
            ref, _ = run_oracle(O, c, steps)
            for k in ("force", "fOld", "rho", "u", "sw_r", "sw_n", "tr_r"):
                assert np.array_equal(got[k], ref[k]), ("oracle", i, k)


def test_dense_members_by_site_with_lanes_and_membership_changes():
    """Site segments numbered across members survive two lanes (each lane
    numbers its own members' segments) and membership changes (a dense
    member leaves, another joins, so every later member's first site moves):
    every member equals its run alone, step counts included."""
    from lbswim.ensemble import Ensemble
    mk = lambda size, ns, seed, nt=0: make_case(size, 0, 0.3, ns=ns, nt=nt, seed=seed, drho=1e-2, du=1e-2,  # noqa: E731
                                                P=1e-2, l=0.3)
    cases = [mk((24, 24, 24), 1800, 21, 50), mk((16, 18, 20), 900, 22), mk((40, 40, 40), 30, 23, 100),
             mk((20, 20, 20), 1200, 24), mk((12, 16, 28), 700, 25)]
    say("dense members by site: lanes 2, membership changes")
    ens = Ensemble("f64")
    ens.set_lanes(2)
    built = []
    for c in cases[:4]:
        lat, system, sw, tr, things = build_member(c, "f64")
        built.append([ens.add(lat, things), lat, sw, tr, 0])
    ens.step(2)
    for b in built:
        b[4] += 2
    ens.remove(built[0][0])            # the first dense member leaves
    lat, system, sw, tr, things = build_member(cases[4], "f64")
    built.append([ens.add(lat, things), lat, sw, tr, 0])
    ens.step(3)
    for b in built[1:]:
        b[4] += 3
    for i, (c, b) in enumerate(zip(cases, built)):
        _, lat, sw, tr, steps = b
        assert lat.time_step == steps, (i, lat.time_step, steps)
        assert_same(collect_lbswim(lat, sw, tr), run_alone(c, "f64", steps), "member %d" % i)
