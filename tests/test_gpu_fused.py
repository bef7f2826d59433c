"""GPU parity of the fused coupled step: Step + CalcHydro in one launch.

Where a lattice qualifies (by default nz % 64 == 0 and ny % rows == 0; with
fusion "always" any lattice with ny >= rows) lbswim_system_step collides and
computes rho/u in one sweep over tiles of rows with a ring of neighbour
sites recomputed.  Its distributions, rho, u, force and particle
state must equal the two-launch path bit for bit -- for every tile shape
(rows, strips), lattices thinner than the halo (nx = 1, ny = rows), both
storage forms and both force modes -- and the FP64 form must equal the CPU
oracle bit for bit.
"""
import os
import sys

import numpy as np
import pytest

import lbswim_oracle as O
from cases import build_lbswim, collect_lbswim, make_case, run_oracle

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def oracle_threads():
    O.set_threads(min(16, os.cpu_count() or 1))


def run(case, storage, force_mode, fusion, steps, mode=True):
    lat, system, sw, tr = build_lbswim(case, storage, force_mode)
    if fusion is None:
        lat.set_fusion(False)
    else:
        lat.set_fusion(mode, *fusion)
    active, rows, _ = lat.fusion
    assert active == (fusion is not None), lat.fusion
    if fusion is not None and fusion[0]:
        assert rows == fusion[0]
    system.Step(steps)
    return collect_lbswim(lat, sw, tr)


def swimmer_case(size, seed, ns=60, nt=300):
    case = make_case(size, 0, 0.3, 0.7, ns=ns, nt=nt, alpha=0.0, seed=seed, drho=1e-2, du=1e-2, P=1e-2, l=0.6)
    # swimmers straddling the periodic faces and the 64-site tile boundaries
    L = np.array(size, float)
    case.sw_r[:, 0] = [0.05, 0.1, 0.2]
    case.sw_r[:, 1] = L - 0.05
    case.sw_r[:, 2] = [L[0] / 2, L[1] - 0.5, 63.7 % L[2]]
    case.sw_r[:, 3] = [0.5, L[1] / 2, 64.2 % L[2]]
    return case


SHAPES = [
    # (size, rows, strips)
    ((4, 8, 64), 4, 0),
    ((4, 8, 64), 8, 3),
    ((3, 4, 128), 2, 1),
    ((3, 4, 128), 4, 2),
    ((1, 2, 64), 2, 0),     # one plane; the halo rows are the tile's own rows
    ((2, 6, 64), 2, 2),
    ((7, 16, 192), 8, 0),
    ((5, 12, 64), 4, 5),    # one plane per strip
]


@pytest.mark.parametrize("storage", ["f64", "f32"])
@pytest.mark.parametrize("force_mode", ["stored", "virtual"])
@pytest.mark.parametrize("size,rows,strips", SHAPES)
def test_fused_equals_two_launches(size, rows, strips, force_mode, storage):
    case = swimmer_case(size, seed=size[0] * 100 + rows)
    say("fused vs two launches: %s rows=%d strips=%d force %s %s" % (size, rows, strips, force_mode, storage))
    a = run(case, storage, force_mode, None, 4)
    b = run(case, storage, force_mode, (rows, strips), 4)
    for k in a:
        assert np.array_equal(a[k], b[k]), k


# Lattices whose tiles are not full: nz below 64 (idle lanes), nz above 64
# but not a multiple of it and ny not a multiple of the rows (the last tile
# overlaps its neighbour), a single z site.  Fusion "always".
ANY_SHAPES = [
    ((4, 8, 16), 8, 0),
    ((5, 12, 24), 4, 0),
    ((3, 13, 48), 8, 2),
    ((4, 8, 96), 8, 0),
    ((3, 9, 70), 4, 2),
    ((2, 6, 1), 2, 0),
    ((6, 10, 33), 8, 0),
]


@pytest.mark.parametrize("storage", ["f64", "f32"])
@pytest.mark.parametrize("force_mode", ["stored", "virtual"])
@pytest.mark.parametrize("size,rows,strips", ANY_SHAPES)
def test_fused_any_shape_equals_two_launches(size, rows, strips, force_mode, storage):
    case = swimmer_case(size, seed=size[2] * 10 + rows)
    say("fused (any shape) vs two launches: %s rows=%d strips=%d force %s %s"
        % (size, rows, strips, force_mode, storage))
    a = run(case, storage, force_mode, None, 4)
    b = run(case, storage, force_mode, (rows, strips), 4, mode="always")
    for k in a:
        assert np.array_equal(a[k], b[k]), k


def test_fused_any_shape_equals_oracle():
    case = make_case((7, 11, 48), 6, 0.25, ns=50, nt=400, alpha=0.02, seed=31, drho=1e-2, du=1e-2, P=1e-2, l=0.5)
    say("fused (any shape) vs oracle: (7, 11, 48), 6 steps")
    ref, _ = run_oracle(O, case)
    got = run(case, "f64", "auto", (0, 0), case.steps, mode="always")
    for k in ref:
        assert np.array_equal(got[k], ref[k]), k


@pytest.mark.parametrize("force_mode", ["stored", "virtual"])
def test_fused_equals_oracle(force_mode):
    case = make_case((6, 8, 128), 5, 0.25, ns=40, nt=500, alpha=0.0, seed=21, drho=1e-2, du=1e-2, P=1e-2, l=0.5)
    say("fused vs oracle: (6, 8, 128), 5 steps, force %s" % force_mode)
    ref, _ = run_oracle(O, case)
    got = run(case, "f64", force_mode, (0, 0), case.steps)
    for k in ref:
        assert np.array_equal(got[k], ref[k]), k


def test_fusion_skipped_where_unsupported():
    """nz not a multiple of 64, or ny not a multiple of the requested rows:
    the coupled step runs the two launches."""
    from lbswim import lb
    lat = lb.Lattice(8, 8, 48, 0.25, 0.25)
    assert lat.fusion[0] is False
    lat = lb.Lattice(8, 6, 64, 0.25, 0.25)
    lat.set_fusion(True, 4, 0)
    assert lat.fusion[0] is False
    lat.set_fusion(True, 2, 0)
    assert lat.fusion == (True, 2, 0)
    lat.set_fusion("always", 4, 0)   # tiles of 4 rows overlap on 6 rows
    assert lat.fusion == (True, 4, 0)
    lat.set_fusion("always", 8, 0)   # more rows than the lattice has
    assert lat.fusion[0] is False
    del lat
    # 15 n sizeof(double) above 4 GiB: the fused launch's 32-bit offsets
    # cannot reach the whole buffer
    lat = lb.Lattice(280, 256, 512, 0.25, 0.25)
    assert lat.fusion[0] is False
    del lat
    lat = lb.Lattice(280, 256, 512, 0.25, 0.25, storage="f32")
    assert lat.fusion[0] is True


def test_moves_use_u_written_after_fused_steps():
    """The fused step also writes u component-interleaved for the particle
    moves; a u replaced afterwards (H2D, standalone CalcHydro) must be the one
    the next moves interpolate -- as with the two launches."""
    case = swimmer_case((4, 8, 64), seed=77)
    out = []
    for fusion in (None, (0, 0)):
        lat, system, sw, tr = build_lbswim(case, "f64", "virtual")
        lat.set_fusion(fusion is not None)
        system.Step(2)
        lat.u.D2H()
        lat.u.data[...] = 1.5 * lat.u.data + 1e-3
        lat.u.H2D()
        sw.Move(lat)
        tr.Move(lat)
        lat.CalcHydro()  # two-launch CalcHydro rewrites u in the reference layout
        sw.Move(lat)
        tr.Move(lat)
        out.append(collect_lbswim(lat, sw, tr))
    for k in out[0]:
        assert np.array_equal(out[0][k], out[1][k]), k


def test_u_written_on_demand_after_fused_steps():
    """Coupled steps, fused or in two launches, leave u only in the moves'
    interleaved copy; InitFromHydro, D2H and sync_u must see the same u
    from both."""
    case = swimmer_case((4, 8, 64), seed=91)
    out = []
    for fusion in (None, (0, 0)):
        lat, system, sw, tr = build_lbswim(case, "f64", "virtual")
        lat.set_fusion(fusion is not None)
        system.Step(3)
        assert lat.u_is_virtual
        lat.InitFromHydro()  # rebuilds f from rho and u
        assert not lat.u_is_virtual
        system.Step(2)
        lat.sync_u()
        assert not lat.u_is_virtual
        lat.sync_u()  # idempotent
        system.Step(1)
        out.append(collect_lbswim(lat, sw, tr))
        assert not lat.u_is_virtual
    for k in out[0]:
        assert np.array_equal(out[0][k], out[1][k]), k


def test_tracer_moves_overlapping_the_next_step():
    """With the fused step the tracer moves run on a second stream beside the
    swimmer moves and the next step's spreading, and the next fused launch
    waits for them.  Two tracer arrays, calls of odd and even step counts, and
    a standalone move between calls must give the two launches' results bit
    for bit."""
    from lbswim import tracers
    case = swimmer_case((4, 8, 64), seed=123, nt=500)
    out = []
    for fusion in (None, (0, 0)):
        lat, system, sw, tr = build_lbswim(case, "f64", "virtual")
        lat.set_fusion(fusion is not None)
        tr2 = tracers.Array(200)
        tr2.r[:] = case.tr_r[:, ::-1][:, :200] * 0.5 + 0.25
        tr2.v[:] = 0.0
        tr2.H2D()
        system.AddThing(tr2)
        system.Step(3)
        system.Step(1)
        tr2.Move(lat)  # reads the current buffer after the call has returned
        system.Step(4)
        res = collect_lbswim(lat, sw, tr)
        tr2.D2H()
        res["tr2_r"] = np.array(tr2.r).copy()
        res["tr2_v"] = np.array(tr2.v).copy()
        out.append(res)
    for k in out[0]:
        assert np.array_equal(out[0][k], out[1][k]), k


def test_interleaved_systems_and_tracers_without_swimmers():
    """Two systems stepped in alternation share the library's second stream;
    a system with tracers and no swimmers (zero force) still moves them
    beside the next step.  Each must give its two-launch result alone, bit
    for bit."""
    from lbswim import lb, tracers
    from lbswim.System import System

    def tracer_only(case, fusion):
        lat = lb.Lattice(*case.size, case.tau_s, case.tau_b)
        lat.rho.data.ravel()[:] = case.rho
        lat.rho.H2D()
        lat.u.data.ravel()[:] = case.u
        lat.u.H2D()
        lat.ZeroForce()
        lat.InitFromHydro()
        lat.set_fusion(fusion)
        system = System(lat)
        tr = tracers.Array(case.tr_r.shape[1])
        tr.r[:] = case.tr_r
        tr.v[:] = 0.0
        tr.H2D()
        system.AddThing(tr)
        return lat, system, tr

    case_a = swimmer_case((4, 8, 64), seed=5)
    case_b = make_case((5, 4, 128), 0, 0.3, ns=0, nt=700, alpha=0.0, seed=6, drho=1e-2, du=1e-2)
    lat, system, sw, tr = build_lbswim(case_a, "f64", "virtual")
    lat.set_fusion(False)
    system.Step(5)
    solo = [collect_lbswim(lat, sw, tr)]
    lat, system, tr = tracer_only(case_b, False)
    system.Step(5)
    solo.append(collect_lbswim(lat, None, tr))
    lat_a, sys_a, sw_a, tr_a = build_lbswim(case_a, "f64", "virtual")
    lat_b, sys_b, tr_b = tracer_only(case_b, True)
    assert lat_a.fusion[0] and lat_b.fusion[0]
    for n in (2, 1, 2):
        sys_a.Step(n)
        sys_b.Step(n)
    got = [collect_lbswim(lat_a, sw_a, tr_a), collect_lbswim(lat_b, None, tr_b)]
    for want, have in zip(solo, got):
        for k in want:
            assert np.array_equal(want[k], have[k]), k


@pytest.mark.parametrize("storage,size", [("f64", (272, 256, 512)), ("f32", (544, 256, 512))])
def test_fused_at_the_32bit_offset_limit(storage, size):
    """The fused step addresses each distribution buffer with 32-bit byte
    offsets (15 n sizeof(T) < 4 GiB).  A lattice at 99.6 % of that limit
    must still give the two launches' distributions, rho and u bit for bit."""
    from lbswim import _core
    n = size[0] * size[1] * size[2]
    elem = 8 if storage == "f64" else 4
    assert 0.99 * 2 ** 32 < 15 * n * elem < 2 ** 32
    case = make_case(size, 0, 0.3, ns=256, nt=2000, alpha=0.0, seed=9, drho=1e-2, du=1e-2)
    say("fused vs two launches at the offset limit: %s %s" % (size, storage))
    out = []
    for fusion in (None, (0, 0)):
        lat, system, sw, tr = build_lbswim(case, storage, "virtual")
        lat.set_fusion(fusion is not None)
        assert lat.fusion[0] == (fusion is not None)
        system.Step(2)
        _core.synchronize()
        res = {}
        for name in ("fOld", "rho", "u"):
            arr = getattr(lat, name)
            arr.D2H()
            res[name] = np.array(arr.data, copy=True)
        tr.D2H()
        res["tr_r"] = np.array(tr.r, copy=True)
        out.append(res)
        del lat, system, sw, tr
    for k in out[0]:
        assert np.array_equal(out[0][k], out[1][k]), k


@pytest.mark.parametrize("size", [(4, 8, 64), (5, 12, 48)])
def test_autotune_keeps_results(size):
    """System.Autotune times the two launches and the fused tile shapes on
    real coupled steps and keeps the fastest; the system's state after the
    tuning and further steps equals plain stepping for the same count."""
    case = swimmer_case(size, seed=size[2] + 3)
    say("coupled-step autotune: %s" % (size,))
    lat, system, sw, tr = build_lbswim(case, "f64", "auto")
    rows, strips = system.Autotune(2)
    assert rows in (0, 2, 4, 8) and strips >= 0
    assert lat.fusion[0] == (rows > 0)
    tuned = lat.time_step
    assert tuned >= 4 * (2 + 2)  # two launches and at least three tile shapes
    system.Step(3)
    got = collect_lbswim(lat, sw, tr)
    lat2, system2, sw2, tr2 = build_lbswim(case, "f64", "auto")
    system2.Step(tuned + 3)
    want = collect_lbswim(lat2, sw2, tr2)
    for k in want:
        assert np.array_equal(got[k], want[k]), k
