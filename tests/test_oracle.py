"""Pins the CPU oracle (oracle/lbswim_oracle.c) without the reference.

The reference ships no tests or recorded outputs, so the oracle is pinned by
(a) the golden D3Q15 tables extracted from the reference's d3q15.c,
(b) known answers at special points (rest state, omega = 1 on a 1x1x1
lattice, grid-point stencil weights, Philox known-answer vectors),
(c) invariants and closed forms (mass and momentum balance, uniform flow,
Peskin moment conditions, linear-field interpolation, dipole moments,
shear-wave decay at nu = tau_s/3), and
(d) an independent vectorised numpy formulation of the step on tiny lattices.
Each test names the mistake it is meant to catch.
"""
import json
import os

import numpy as np
import pytest

import lbswim_oracle as O
from util import rel_inf, random_hydro, readme_mobility

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = json.load(open(os.path.join(HERE, "golden", "d3q15_tables.json")))["tables"]


@pytest.fixture(scope="module")
def tab():
    return O.tables()


# ---------------------------------------------------------------- tables (a1)

def test_tables_equal_reference_source_bitwise(tab):
    # golden values are d3q15.c:2-186 evaluated as C folds its literals
    for key, ref in GOLD.items():
        ours = np.asarray(tab[key])
        np.testing.assert_array_equal(ours, np.asarray(ref, dtype=ours.dtype), err_msg=key)


def test_tables_closed_forms(tab):
    mm, mmi, w, xi = tab["mm"], tab["mmi"], tab["w"], tab["xi"]
    np.testing.assert_allclose(mm @ mmi, np.eye(15), atol=4e-16)
    np.testing.assert_allclose(mmi, w[:, None] * mm.T * tab["norms"][None, :], rtol=2e-16, atol=0)
    np.testing.assert_array_equal(xi[tab["complement"]], -xi)
    np.testing.assert_allclose(tab["Q"], xi[:, :, None] * xi[:, None, :] - np.eye(3) / 3.0, atol=1e-16)
    np.testing.assert_allclose(1.0 / tab["norms"], (w[None, :] * mm**2).sum(axis=1), rtol=1e-15)
    assert abs(w.sum() - 1.0) < 1e-15 and np.allclose((w[:, None, None] * xi[:, :, None] * xi[:, None, :]).sum(0),
                                                      np.eye(3) / 3.0)


# ---------------------------------------------------------------- Peskin (a8)

def test_peskin_special_values():
    assert O.peskin_delta(0.0) == 0.5
    assert O.peskin_delta(1.0) == 0.25 and O.peskin_delta(-1.0) == 0.25
    assert O.peskin_delta(2.0) == 0.0 and O.peskin_delta(-2.0) == 0.0 and O.peskin_delta(3.7) == 0.0


def test_peskin_moment_conditions():
    rng = np.random.default_rng(5)
    for r in rng.uniform(-3.0, 40.0, 200):
        x0 = np.ceil(r - 2.0)
        d = r - (x0 + np.arange(4))
        phi = np.array([O.peskin_delta(x) for x in d])
        assert abs(phi.sum() - 1.0) < 2e-16 * 8
        assert abs((d * phi).sum()) < 4e-15
        assert abs((phi**2).sum() - 3.0 / 8.0) < 1e-15


# ------------------------------------------------------- lattice known answers

def _lat(size, tau=0.25, taub=None):
    return O.Lattice(*size, tau, tau if taub is None else taub)


def test_rest_state_is_fixed_point(tab):
    # f = w exactly after InitFromHydro; the step keeps it to rounding only
    # (sum_p w_p = 1 - 3.3e-16 in double, so it is not a bitwise fixed point)
    lat = _lat((4, 5, 6))
    lat.rho[:] = 1.0
    lat.InitFromHydro()
    f = lat.fOld.reshape(15, -1).copy()
    assert np.array_equal(f, np.repeat(tab["w"][:, None], lat.n, axis=1))
    for _ in range(10):
        lat.ZeroForce()
        lat.Step()
    np.testing.assert_allclose(lat.fOld.reshape(15, -1), f, rtol=0, atol=1e-15)
    lat.CalcHydro()
    np.testing.assert_allclose(lat.u, 0.0, atol=1e-16)


def test_init_from_hydro_known_answer_with_sxz_quirk(tab):
    # lb.cu:341-369: m = (rho, rho u, rho ux^2, rho ux uy, rho ux uy [sic], rho uy^2, rho uy uz, rho uz^2, 0...)
    lat = _lat((2, 3, 4))
    rng = np.random.default_rng(1)
    lat.rho[:], lat.u[:] = random_hydro(rng, lat.n, 0.1, 0.1)
    lat.InitFromHydro()
    r, (ux, uy, uz) = lat.rho, lat.u.reshape(3, -1)
    m = np.zeros((15, lat.n))
    m[0], m[1], m[2], m[3] = r, r * ux, r * uy, r * uz
    m[4], m[5], m[6], m[7], m[8], m[9] = r * ux * ux, r * ux * uy, r * ux * uy, r * uy * uy, r * uy * uz, r * uz * uz
    np.testing.assert_allclose(lat.fOld.reshape(15, -1), tab["mmi"] @ m, rtol=0, atol=2e-16)
    # the quirk is visible: the SXZ mode of f is rho ux uy, not rho ux uz
    sxz = tab["mm"][6] @ lat.fOld.reshape(15, -1)
    np.testing.assert_allclose(sxz, r * ux * uy, atol=1e-16)
    assert np.max(np.abs(sxz - r * ux * uz)) > 1e-3


def test_omega_one_single_site_known_answer(tab):
    # On a 1x1x1 periodic lattice streaming is the identity, so Step returns the
    # post-collision distributions.  With tau = 1/2 (omega = 1) lb.cu:163-178
    # set the stress to its (quirky) forced equilibrium:
    #   TrS' = rho usq + uF;  S'_ab - TrS'/3 d_ab = rho (u_a u_b - usq d_ab) + (u_a F_b + F_a u_b - 2 uF d_ab)
    lat = _lat((1, 1, 1), 0.5, 0.5)
    rho, u0 = 1.2, np.array([0.1, 0.2, 0.3])
    F = np.array([0.01, -0.02, 0.03])
    lat.rho[:] = rho
    lat.u[:] = u0
    lat.InitFromHydro()
    f0 = lat.fOld.copy()
    lat.force[:] = F
    lat.Step()
    m0 = tab["mm"] @ f0
    u = (m0[1:4] + 0.5 * F) / m0[0]
    usq, uF = u @ u, u @ F
    j = m0[1:4] + F
    tr = m0[0] * usq + uF
    S = m0[0] * (np.outer(u, u) - usq * np.eye(3)) + (np.outer(u, F) + np.outer(F, u) - 2 * uF * np.eye(3))
    S += tr / 3.0 * np.eye(3)
    m = np.zeros(15)
    m[0], m[1:4] = m0[0], j
    m[4:10] = [S[0, 0], S[0, 1], S[0, 2], S[1, 1], S[1, 2], S[2, 2]]
    np.testing.assert_allclose(lat.fOld, tab["mmi"] @ m, rtol=0, atol=4e-16)


def _numpy_step(tab, f, F, size, tau_s, tau_b):
    """Vectorised restatement: modes by matrix product, relaxation on arrays,
    push streaming by np.roll (fNew[p](x + c_p) = f*_p(x))."""
    n = f.shape[1]
    m = tab["mm"] @ f
    om_s, om_b = 1.0 / (tau_s + 0.5), 1.0 / (tau_b + 0.5)
    u = (m[1:4] + 0.5 * F) / m[0]
    m[1:4] += F
    usq, uF = (u * u).sum(0), (u * F).sum(0)
    idx = {(0, 0): 4, (0, 1): 5, (0, 2): 6, (1, 1): 7, (1, 2): 8, (2, 2): 9}
    S = np.empty((3, 3, n))
    for (a, b), k in idx.items():
        S[a, b] = S[b, a] = m[k]
    tr = S[0, 0] + S[1, 1] + S[2, 2]
    for a in range(3):
        S[a, a] -= tr / 3
    tr = tr - om_b * (tr - m[0] * usq) + 2 * om_b * tau_b * uF
    d = np.eye(3)
    for a in range(3):
        for b in range(3):
            S[a, b] = S[a, b] - om_s * (S[a, b] - m[0] * (u[a] * u[b] - usq * d[a, b]))
            S[a, b] = S[a, b] + 2 * om_s * tau_s * (u[a] * F[b] + F[a] * u[b] - 2 * uF * d[a, b])
        S[a, a] += tr / 3
    for (a, b), k in idx.items():
        m[k] = S[a, b]
    m[10:] = 0.0
    fpc = tab["mmi"] @ m
    out = np.empty_like(fpc)
    for p in range(15):
        out[p] = np.roll(fpc[p].reshape(size), shift=tuple(tab["ci"][p]), axis=(0, 1, 2)).ravel()
    return out


@pytest.mark.parametrize("size", [(4, 5, 6), (7, 5, 3), (1, 2, 3)])
def test_step_matches_vectorised_restatement(tab, size):
    lat = _lat(size, 0.3, 0.7)
    rng = np.random.default_rng(11)
    lat.rho[:], lat.u[:] = random_hydro(rng, lat.n, 0.05, 0.05)
    lat.InitFromHydro()
    lat.force[:] = 1e-3 * rng.standard_normal(3 * lat.n)
    f, F = lat.fOld.reshape(15, -1).copy(), lat.force.reshape(3, -1).copy()
    lat.Step()
    ref = _numpy_step(tab, f, F, size, 0.3, 0.7)
    assert rel_inf(lat.fOld.reshape(15, -1), ref) < 1e-14


def test_uniform_flow_preserved():
    lat = _lat((3, 4, 5))
    lat.rho[:] = 1.01
    lat.u.reshape(3, -1)[:] = np.array([0.01, -0.02, 0.005])[:, None]
    lat.InitFromHydro()
    for _ in range(20):
        lat.ZeroForce()
        lat.Step()
    lat.CalcHydro()
    np.testing.assert_allclose(lat.rho, 1.01, rtol=1e-14)
    np.testing.assert_allclose(lat.u.reshape(3, -1), np.array([0.01, -0.02, 0.005])[:, None] * np.ones((3, lat.n)),
                               rtol=0, atol=1e-15)


def test_mass_and_momentum_balance(tab):
    lat = _lat((6, 5, 4), 0.4, 0.9)
    rng = np.random.default_rng(3)
    lat.rho[:], lat.u[:] = random_hydro(rng, lat.n, 0.02, 0.02)
    lat.InitFromHydro()
    xi = tab["xi"]
    for _ in range(5):
        f = lat.fOld.reshape(15, -1)
        mass, mom = f.sum(), xi.T @ f.sum(1)
        lat.force[:] = 1e-3 * rng.standard_normal(3 * lat.n)
        Ftot = lat.force.reshape(3, -1).sum(1)
        lat.Step()
        f = lat.fOld.reshape(15, -1)
        assert abs(f.sum() - mass) < 1e-12
        np.testing.assert_allclose(xi.T @ f.sum(1), mom + Ftot, rtol=0, atol=1e-13)


def test_init_then_hydro_round_trip():
    lat = _lat((5, 4, 3))
    rng = np.random.default_rng(4)
    rho, u = random_hydro(rng, lat.n, 0.05, 0.05)
    lat.rho[:], lat.u[:] = rho, u
    lat.InitFromHydro()
    lat.rho[:] = 0
    lat.u[:] = 0
    lat.CalcHydro()
    np.testing.assert_allclose(lat.rho, rho, rtol=1e-15)
    np.testing.assert_allclose(lat.u, u, rtol=0, atol=1e-16)


def test_shear_wave_decays_at_lattice_viscosity():
    # u_x(z) = A sin(k z); amplitude decays as exp(-nu k^2 t), nu = tau_s/3 (lb.pyx:129-131)
    L, tau, A = 32, 0.25, 1e-6
    lat = _lat((1, 1, L), tau)
    z = np.arange(L)
    lat.rho[:] = 1.0
    lat.u.reshape(3, -1)[0] = A * np.sin(2 * np.pi * z / L)
    lat.InitFromHydro()

    def amp():
        lat.CalcHydro()
        return 2.0 / L * (lat.u.reshape(3, -1)[0] * np.sin(2 * np.pi * z / L)).sum()

    for _ in range(50):
        lat.Step()
    a0 = amp()
    for _ in range(200):
        lat.Step()
    a1 = amp()
    rate = -np.log(a1 / a0) / 200
    k = 2 * np.pi / L
    assert abs(rate / (tau / 3 * k * k) - 1.0) < 0.02


# ------------------------------------------------------- interpolation (a9)

def test_interp_constant_and_linear_fields():
    lat = _lat((10, 11, 12))
    x, y, z = np.meshgrid(np.arange(10), np.arange(11), np.arange(12), indexing="ij")
    u = lat.u.reshape(3, 10, 11, 12)
    u[0] = 0.3
    u[1] = 0.1 + 0.01 * x - 0.02 * y + 0.005 * z
    u[2] = -0.2 + 0.03 * z
    rng = np.random.default_rng(8)
    for r in rng.uniform(3.0, 7.0, (50, 3)):
        v = lat.interp(r)
        assert abs(v[0] - 0.3) < 1e-15
        assert abs(v[1] - (0.1 + 0.01 * r[0] - 0.02 * r[1] + 0.005 * r[2])) < 1e-15
        assert abs(v[2] - (-0.2 + 0.03 * r[2])) < 1e-15


def test_interp_grid_point_weights():
    lat = _lat((8, 8, 8))
    u = lat.u.reshape(3, 8, 8, 8)
    u[0, 3, 4, 5] = 1.0
    assert lat.interp([3.0, 4.0, 5.0])[0] == 0.125          # 1/2 * 1/2 * 1/2
    assert lat.interp([4.0, 4.0, 5.0])[0] == 0.0625         # 1/4 * 1/2 * 1/2
    assert lat.interp([3.0, 4.0, 5.0 + 8.0 - 1e-300])[0] == 0.125  # periodic image


# ------------------------------------------------------- swimmers / tracers

def _swimmers(n, **kw):
    return O.Swimmers(n, O.Common(**kw))


def test_force_dipole_zero_sum_and_moment():
    lat = _lat((16, 16, 16))
    P, l = 1e-3, 5.0
    sw = _swimmers(1, P=P, l=l)
    sw.r[:, 0] = [8.3, 7.6, 10.2]
    sw.n[:, 0] = [0.0, 0.0, 1.0]
    lat.ZeroForce()
    sw.AddForces(lat)
    F = lat.force.reshape(3, 16, 16, 16)
    assert np.abs(F.sum(axis=(1, 2, 3))).max() < 1e-18
    z = np.arange(16)[None, None, :]
    # head +P n at r, tail -P n at r - l n: sum z F_z = P*r_z - P*(r_z - l) = P l
    assert abs((F[2] * z).sum() - P * l) < 1e-15
    assert abs(F[2][F[2] > 0].sum() - P) < 1e-17  # head spreads weight 1 (stencils disjoint)


def test_swimmer_in_quiescent_fluid_swims_straight():
    lat = _lat((12, 12, 12))
    mob = readme_mobility(0.25)
    sw = _swimmers(3, P=1e-3, l=0.1, mobility=mob)
    rng = np.random.default_rng(2)
    sw.r[:] = rng.uniform(0, 12, (3, 3))
    n = rng.standard_normal((3, 3))
    sw.n[:] = n / np.linalg.norm(n, axis=0)
    r0, n0 = sw.r.copy(), sw.n.copy()
    sw.Move(lat)
    np.testing.assert_allclose(sw.r, np.mod(r0 + mob * 1e-3 * n0, 12), rtol=0, atol=1e-14)
    np.testing.assert_allclose(sw.v, mob * 1e-3 * n0, rtol=1e-15)
    np.testing.assert_allclose(sw.n, n0, rtol=0, atol=2e-16)


def test_swimmer_rotates_in_linear_shear():
    # u_x = g z: v(r) - v(r - l n) = g l n_z e_x exactly (linear reproduction),
    # so n' = normalise(n + g n_z e_x)
    L, g = 16, 1e-3
    lat = _lat((L, L, L))
    lat.u.reshape(3, L, L, L)[0] = g * np.arange(L)[None, None, :]
    sw = _swimmers(1, P=0.0, l=0.5, translational_advection_off=True)
    sw.r[:, 0] = [8.0, 8.0, 7.5]
    n = np.array([0.6, 0.0, 0.8])
    sw.n[:, 0] = n
    sw.Move(lat)
    expect = n + np.array([g * n[2], 0, 0])
    np.testing.assert_allclose(sw.n[:, 0], expect / np.linalg.norm(expect), rtol=0, atol=1e-15)


def test_tracer_uniform_flow():
    lat = _lat((6, 7, 8))
    u = np.array([0.01, -0.03, 0.02])
    lat.u.reshape(3, -1)[:] = u[:, None]
    tr = O.Tracers(5)
    rng = np.random.default_rng(9)
    tr.r[:] = rng.uniform(0, 6, (3, 5))
    r0 = tr.r.copy()
    tr.Move(lat)
    np.testing.assert_allclose(tr.v, u[:, None] * np.ones((3, 5)), rtol=0, atol=4e-17)
    np.testing.assert_allclose(tr.r, np.mod(r0 + u[:, None] + np.array([6, 7, 8])[:, None], [[6], [7], [8]]),
                               atol=1e-15)
    np.testing.assert_allclose(tr.s, u[:, None] * np.ones((3, 5)), atol=4e-17)


# ------------------------------------------------------- random streams

def test_philox_known_answer_vectors():
    # Random123 kat_vectors for philox4x32_10
    assert list(O.philox4x32_10([0, 0, 0, 0], [0, 0])) == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]
    assert list(O.philox4x32_10([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2)) == [0x408F276D, 0x41C83B0E, 0xA20BC7C6,
                                                                         0x6D5451FD]
    assert list(O.philox4x32_10([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0])) == [
        0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]


def test_box_muller_log_and_cossin_accuracy():
    """The series log and cos/sin(2 pi u) of the Box-Muller transform against
    numpy (a wrong or missing series term shows as an error far above 1e-15;
    a wrong quadrant as an error of order 1)."""
    rng = np.random.default_rng(5)
    x = np.concatenate([rng.uniform(0, 1, 3000), 2.0 ** -rng.uniform(0, 54, 3000),
                        [2.0 ** -54, 0.5, 0.70710678118654752, 0.7071067811865476, 1 - 2 ** -53]])
    got = np.array([O.bm_log(v) for v in x])
    assert np.max(np.abs(got - np.log(x)) / np.maximum(1.0, np.abs(np.log(x)))) < 4e-16
    u = np.concatenate([rng.uniform(0, 1, 4000), np.arange(8) / 8.0, [2 ** -53, 0.125 - 2 ** -55, 0.375 + 2 ** -54]])
    cs = np.array([O.bm_cossin_2pi(v) for v in u])
    # reference in extended precision: 2 pi u rounded to double alone is off by up to 4e-16
    ang = np.longdouble("6.283185307179586476925286766559005768") * u.astype(np.longdouble)
    assert np.max(np.abs(cs[:, 0] - np.cos(ang))) < 3e-16
    assert np.max(np.abs(cs[:, 1] - np.sin(ang))) < 3e-16
    # exact at the quadrant points
    assert O.bm_cossin_2pi(0.0) == (1.0, 0.0) and O.bm_cossin_2pi(0.25) == (0.0, 1.0)
    assert O.bm_cossin_2pi(0.5) == (-1.0, 0.0) and O.bm_cossin_2pi(0.75) == (0.0, -1.0)


def test_uniforms_and_normals_statistics():
    U = np.array([O.swimmer_uniform(1234, i, c) for i in range(200) for c in range(100)])
    assert U.min() > 0.0 and U.max() <= 1.0
    assert abs(U.mean() - 0.5) < 4 * np.sqrt(1 / 12 / U.size)
    G = np.array([O.swimmer_normals(99, i, c) for i in range(100) for c in range(100)]).ravel()
    assert abs(G.mean()) < 4 / np.sqrt(G.size)
    assert abs(G.var() - 1.0) < 4 * np.sqrt(2.0 / G.size)


def test_tumble_rate_within_binomial_bound():
    lat = _lat((8, 8, 8))
    alpha, ns, steps = 0.05, 400, 50
    sw = _swimmers(ns, P=0.0, l=0.1, alpha=alpha, seed=77)
    rng = np.random.default_rng(6)
    sw.r[:] = rng.uniform(0, 8, (3, ns))
    sw.n[:] = np.array([0.0, 0.0, 1.0])[:, None]
    nz = []
    for _ in range(steps):
        before = sw.n.copy()
        sw.Move(lat)
        changed = np.any(sw.n != before, axis=0)
        nz.extend(sw.n[:, changed].T.tolist())
    N = ns * steps
    assert abs(sw.tumbles - alpha * N) < 4 * np.sqrt(N * alpha * (1 - alpha))
    nz = np.array(nz)
    assert np.abs(nz.mean(axis=0)).max() < 4 / np.sqrt(3 * len(nz))  # isotropic new orientations
    np.testing.assert_allclose(np.linalg.norm(sw.n, axis=0), 1.0, atol=1e-15)


# ------------------------------------------------------- regression fixture

def test_coupled_fixture_regression():
    """Oracle output for a small coupled run, committed under tests/golden/."""
    from golden_cases import coupled_small
    ref = np.load(os.path.join(HERE, "golden", "coupled_small.npz"))
    out = coupled_small(O)
    for k in ref.files:
        np.testing.assert_array_equal(out[k], ref[k], err_msg=k)


def test_native_timing_build_same_results(tmp_path):
    """The CPU baseline times a build compiled with -march=native on its host;
    it must compute what the pinned build computes (a vectorised reduction
    or a contracted multiply-add would change the bits)."""
    import subprocess
    import sys
    lib = O.build_native(str(tmp_path / "liblbswim_oracle_native.so"))
    out = tmp_path / "native.npz"
    code = ("import sys; sys.path[:0] = %r; import numpy as np, lbswim_oracle as O; O.use_library(%r); "
            "from golden_cases import coupled_small; np.savez(%r, **coupled_small(O))"
            % ([HERE, os.path.join(os.path.dirname(HERE), "oracle")], lib, str(out)))
    subprocess.check_call([sys.executable, "-c", code])
    ref = np.load(os.path.join(HERE, "golden", "coupled_small.npz"))
    got = np.load(out)
    for k in ref.files:
        np.testing.assert_array_equal(got[k], ref[k], err_msg=k)
