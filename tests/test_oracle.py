"""Pins the oracle (the CPU restatement of the reference path) against what
the mathematics fixes: closed forms, conservation laws, textbook solutions.
The reference ships no golden data (its regression data is generated
locally and git-ignored), so these checks are what makes the oracle
trustworthy; a dropped term, a wrong sign or a transposed index in any of
the restated formulas fails one of them."""
import json
import math
import os

import numpy as np
import pytest

from life_amd import config as cfg

CX = np.array([0, 1, -1, 0, 0, 1, -1, 1, -1])
CY = np.array([0, 0, 0, 1, -1, 1, -1, -1, 1])
CS2 = 1.0 / 3.0


# ------------------------------------------------------------------ delta

def test_dirac_special_values(oracle):
    d = oracle.dirac_delta
    assert d(0.0) == pytest.approx(2.0 / 3.0, abs=1e-16)
    assert d(0.5) == pytest.approx(0.5, abs=1e-16)
    assert d(0.5 + 1e-12) == pytest.approx(0.5, abs=1e-10)  # both branches meet at 1/2
    assert d(1.0) == pytest.approx(1.0 / 6.0, abs=1e-16)
    assert d(1.5) == 0.0 and d(1.6) == 0.0 and d(-1.6) == 0.0
    assert d(-0.3) == d(0.3)


def test_dirac_moment_conditions(oracle):
    """Roma et al. (1999): sum_i d(x-i) = 1, sum_i (x-i) d(x-i) = 0,
    sum_i d(x-i)^2 = 1/2 for every x."""
    rng = np.random.default_rng(1)
    for x in rng.uniform(-3, 3, 200):
        i = np.arange(math.floor(x) - 3, math.floor(x) + 4)
        w = np.array([oracle.dirac_delta(x - k) for k in i])
        assert w.sum() == pytest.approx(1.0, abs=1e-14)
        assert ((x - i) * w).sum() == pytest.approx(0.0, abs=1e-14)
        assert (w * w).sum() == pytest.approx(0.5, abs=1e-14)


def test_find_support_brute_force(oracle):
    """findSupport equals the brute-force set |x-i| < 1.5, |y-j| < 1.5 in
    the grid, i-major, with tensor-product weights."""
    rng = np.random.default_rng(2)
    Dx, Nx, Ny = 0.7, 20, 12
    for _ in range(200):
        x, y = rng.uniform(-1.0, Nx * Dx + 1.0), rng.uniform(-1.0, Ny * Dx + 1.0)
        ij, d = oracle.find_support(x, y, Dx, Nx, Ny)
        want = [(i, j) for i in range(Nx) for j in range(Ny) if abs(x / Dx - i) < 1.5 and abs(y / Dx - j) < 1.5]
        assert [tuple(r) for r in ij] == want
        for (i, j), w in zip(want, d):
            assert w == oracle.dirac_delta(x / Dx - i) * oracle.dirac_delta(y / Dx - j)
        if len(want) == 9:
            # interior marker: weights sum to 1 and reproduce a linear field exactly
            assert d.sum() == pytest.approx(1.0, abs=1e-14)
            assert (d * ij[:, 0]).sum() == pytest.approx(x / Dx, abs=1e-12)
            assert (d * ij[:, 1]).sum() == pytest.approx(y / Dx, abs=1e-12)


# ------------------------------------------------------------------ equilibria / collision

def moments(f):
    return f.sum(), (CX * f).sum(), (CY * f).sum()


@pytest.mark.parametrize("cm", [0, 1])
def test_equilibrium_moments(oracle, cm):
    rng = np.random.default_rng(3)
    for _ in range(50):
        rho, ux, uy = 1 + 0.1 * rng.standard_normal(), 0.1 * rng.standard_normal(), 0.1 * rng.standard_normal()
        f = oracle.equilibrium(cm, rho, ux, uy)
        r, jx, jy = moments(f)
        assert r == pytest.approx(rho, rel=1e-14)
        assert jx == pytest.approx(rho * ux, abs=1e-15)
        assert jy == pytest.approx(rho * uy, abs=1e-15)
        # second moments rho (cs^2 I + u u)
        assert (CX * CX * f).sum() == pytest.approx(rho * (CS2 + ux * ux), rel=1e-13)
        assert (CY * CY * f).sum() == pytest.approx(rho * (CS2 + uy * uy), rel=1e-13)
        assert (CX * CY * f).sum() == pytest.approx(rho * ux * uy, abs=1e-15)
        if cm:
            # product form (4th-order Hermite): 1D moments factorise
            m = lambda a, u: [1.0, u, CS2 + u * u][a]
            for a in range(3):
                for b in range(3):
                    assert (CX ** a * CY ** b * f).sum() == pytest.approx(rho * m(a, ux) * m(b, uy), abs=1e-14)


def test_bgk_collision_conserves(oracle):
    """BGK + Guo: mass conserved, momentum increases by exactly F."""
    rng = np.random.default_rng(4)
    for _ in range(50):
        f = oracle.equilibrium(0, 1.0, 0.05, -0.02) * (1 + 0.02 * rng.standard_normal(9))
        F = 1e-4 * rng.standard_normal(2)
        r, jx, jy = moments(f)
        ux, uy = (jx + 0.5 * F[0]) / r, (jy + 0.5 * F[1]) / r
        out = oracle.collide_site(0, 1.7, f, r, ux, uy, F[0], F[1])
        r2, jx2, jy2 = moments(out)
        assert r2 == pytest.approx(r, rel=1e-14)
        assert jx2 == pytest.approx(jx + F[0], abs=1e-15)
        assert jy2 == pytest.approx(jy + F[1], abs=1e-15)


def central_moments(f, ux, uy):
    dx, dy = CX - ux, CY - uy
    return np.array([f.sum(), (dx * f).sum(), (dy * f).sum(), ((dx ** 2 + dy ** 2) * f).sum(),
                     ((dx ** 2 - dy ** 2) * f).sum(), (dx * dy * f).sum(), (dx ** 2 * dy * f).sum(),
                     (dx * dy ** 2 * f).sum(), (dx ** 2 * dy ** 2 * f).sum()])


def test_cm_collision_post_moments(oracle):
    """Central moments (De Rosis 2019): the post-collision central moments
    are exactly k0..k8 -- this checks every coefficient of the
    back-transformation table."""
    rng = np.random.default_rng(5)
    for _ in range(50):
        f = oracle.equilibrium(1, 1.0, 0.08, 0.03) * (1 + 0.03 * rng.standard_normal(9))
        F = 1e-3 * rng.standard_normal(2)
        omega = rng.uniform(0.5, 1.9)
        r, jx, jy = moments(f)
        ux, uy = (jx + 0.5 * F[0]) / r, (jy + 0.5 * F[1]) / r
        pre = central_moments(f, ux, uy)
        out = oracle.collide_site(1, omega, f, r, ux, uy, F[0], F[1])
        post = central_moments(out, ux, uy)
        want = [r, 0.5 * F[0], 0.5 * F[1], 2 * r * CS2, (1 - omega) * pre[4], (1 - omega) * pre[5],
                0.5 * F[1] * CS2, 0.5 * F[0] * CS2, r * CS2 * CS2]
        assert np.allclose(post, want, rtol=1e-12, atol=1e-15)


def test_cm_omega_one_gives_equilibrium(oracle):
    rng = np.random.default_rng(6)
    f = oracle.equilibrium(1, 1.0, 0.04, -0.07) * (1 + 0.03 * rng.standard_normal(9))
    r, jx, jy = moments(f)
    out = oracle.collide_site(1, 1.0, f, r, jx / r, jy / r)
    assert np.allclose(out, oracle.equilibrium(1, r, jx / r, jy / r), rtol=1e-13, atol=1e-16)


# ------------------------------------------------------------------ lattice

def test_periodic_box_conservation(oracle):
    """All sides periodic (eFluid): mass is conserved exactly and momentum
    grows by N*F per step."""
    for cm in (0, 1):
        p = cfg.lattice_case(16, 12, 0.8, wall_left=cfg.FLUID, wall_right=cfg.FLUID, wall_bottom=cfg.FLUID,
                             wall_top=cfg.FLUID, profile=cfg.UNIFORM, uxInlet_p=0.0, dpdx=2e-5, central_moments=cm)
        c = oracle.OracleCase(p, "")
        from tests._util import random_state
        f, rho, u = random_state(16, 12, seed=9, cm=cm)
        jx0 = (f * CX).sum()
        c.set_grid(f=f, rho=f.sum(-1), u=np.stack([((f * CX).sum(-1) + 0.5 * 2e-5) / f.sum(-1),
                                                      (f * CY).sum(-1) / f.sum(-1)], -1))
        c.step(30)
        f2, _, _, _ = c.grid()
        assert f2.sum() == pytest.approx(f.sum(), rel=1e-13)
        assert (f2 * CX).sum() - jx0 == pytest.approx(30 * 16 * 12 * 2e-5, rel=1e-9)


@pytest.mark.parametrize("cm", [0, 1])
def test_poiseuille(oracle, cm):
    """Body-force-driven channel with regularised no-slip walls on the
    first/last rows: steady profile u = F y (H - y) / (2 nu), H = Ny - 1."""
    Ny, tau, F = 17, 0.8, 1e-6
    p = cfg.lattice_case(4, Ny, tau, wall_left=cfg.FLUID, wall_right=cfg.FLUID, wall_bottom=cfg.WALL,
                         wall_top=cfg.WALL, profile=cfg.UNIFORM, uxInlet_p=0.0, dpdx=F, central_moments=cm)
    c = oracle.OracleCase(p, "")
    c.step(6000)
    _, _, u, _ = c.grid()
    nu = cfg.lattice_viscosity(1.0 / tau)
    y = np.arange(Ny, dtype=float)
    exact = F * y * (Ny - 1 - y) / (2 * nu)
    assert np.max(np.abs(u[2, :, 0] - exact)) / exact.max() < 2e-3
    assert np.max(np.abs(u[:, :, 1])) < 1e-12


@pytest.mark.parametrize("cm", [0, 1])
def test_womersley_channel(oracle, cm):
    """WOMERSLEY forcing (Grid.cpp:52-61): after the start-up transient has
    decayed (rate nu pi^2 / H^2), the channel velocity follows the planar
    Womersley solution through a whole period; a wrong period, sign or
    amplitude of the oscillating force fails it."""
    Ny, tau, Wo, G = 33, 0.8, 4.0, 1e-5
    p = cfg.womersley_channel(Ny=Ny, tau=tau, Wo=Wo, G=G, central_moments=cm)
    nu = cfg.lattice_viscosity(1.0 / tau)
    H = Ny - 1.0
    period = H * H * np.pi / (2.0 * Wo * Wo * nu)
    c = oracle.OracleCase(p, "")
    t = 9000
    c.step(t)
    y = np.arange(Ny, dtype=float)
    amp = G / (2.0 * np.pi / period)  # the inviscid amplitude G / omega
    worst = 0.0
    for _ in range(8):
        c.step(int(period / 8))
        t += int(period / 8)
        _, _, u, _ = c.grid()
        exact = np.array(cfg.womersley_profile(y, float(t), H, nu, Wo, G))
        worst = max(worst, np.max(np.abs(u[1, :, 0] - exact)) / amp)
        assert np.max(np.abs(u[:, :, 1])) < 1e-6 * amp
    print("womersley worst", worst)
    assert worst < 1e-2
    c.close()


def test_regularised_wall_velocity(oracle):
    """After a step, the rebuilt wall populations carry the imposed velocity
    exactly (regularised BC, Latt et al. 2008)."""
    p, g = cfg.lid_driven_cavity(N=20, tau=0.7, u_lid=0.08)
    c = oracle.OracleCase(p, g)
    c.step(5)
    _, _, u, _ = c.grid()
    assert np.max(np.abs(u[0, 1:-1])) < 1e-15 and np.max(np.abs(u[-1, 1:-1])) < 1e-15
    assert np.max(np.abs(u[1:-1, 0])) < 1e-15
    assert np.allclose(u[1:-1, -1, 0], 0.08, atol=1e-15) and np.max(np.abs(u[1:-1, -1, 1])) < 1e-15


def test_pressure_outlet_density(oracle):
    from tests._util import bc_mix_params
    p, g = bc_mix_params("channel_pressure")
    c = oracle.OracleCase(p, g)
    c.step(10)
    _, rho, _, _ = c.grid()
    assert np.allclose(rho[-1, 1:-1], 1.0, atol=1e-14)


def test_convective_outlet_uniform_flow(oracle):
    """A uniform flow leaves the convective outlet untouched (delU = 0)."""
    p = cfg.lattice_case(24, 10, 0.7, wall_left=cfg.VELOCITY, wall_right=cfg.CONVECTIVE, wall_bottom=cfg.FREESLIP,
                         wall_top=cfg.FREESLIP, profile=cfg.UNIFORM, uxInlet_p=0.05, ux0_p=0.05)
    c = oracle.OracleCase(p, "")
    c.step(20)
    _, rho, u, _ = c.grid()
    assert np.allclose(u[..., 0], 0.05, atol=1e-13) and np.max(np.abs(u[..., 1])) < 1e-13
    assert np.allclose(rho, 1.0, atol=1e-13)


# Ghia, Ghia & Shin (1982), Table I, Re = 100: u/U on the vertical centreline
_GHIA = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                                    "ghia1982_re100_u_centreline.json")))
GHIA_Y, GHIA_U = _GHIA["y"], _GHIA["u"]


def test_lid_driven_cavity_ghia(oracle):
    """Cavity at Re = 100 against the Ghia et al. centreline velocities."""
    oracle.set_threads(4)
    N, U = 65, 0.1
    nu = U * (N - 1) / 100.0
    p, g = cfg.lid_driven_cavity(N=N, tau=3 * nu + 0.5, u_lid=U)
    c = oracle.OracleCase(p, g)
    c.step(16000)
    oracle.set_threads(1)
    _, _, u, _ = c.grid()
    prof = u[(N - 1) // 2, :, 0] / U
    y = np.arange(N) / (N - 1)
    got = np.interp(GHIA_Y, y, prof)
    assert np.max(np.abs(got - np.array(GHIA_U))) < 0.03


# ------------------------------------------------------------------ IBM / epsilon

def test_single_marker_epsilon(oracle):
    """One isolated marker: A = ds * sum_s delta_s^2 = ds / 4, so eps = 4 / ds."""
    p = cfg.lattice_case(30, 30, 0.8, wall_left=cfg.WALL, wall_right=cfg.WALL, profile=cfg.UNIFORM, uxInlet_p=0.0)
    g = "CIRCLE 1 0 12.3 14.6 0.0 0.0 0.2\n"  # floor(2 pi 0.2) = 1 marker
    c = oracle.OracleCase(p, g)
    m = c.markers()
    assert c.n_markers == 1 and m["ds"][0] == 10.0
    assert m["eps"][0] == pytest.approx(4.0 / m["ds"][0], rel=1e-13)


def test_overlap_removal(oracle):
    """Markers of a rigid body closer than Dx/2 to a flexible body are
    removed (Objects.cpp:867-938)."""
    p, g = cfg.flexible_fsi(Nx=200, Ny=62, D=12.0, centre=(30.0, 30.0), L=42.0, h=2.4, n_el=6)
    c = oracle.OracleCase(p, g)
    circle, fil = c.body_info(0), c.body_info(1)
    assert circle["n"] == int(math.floor(2 * math.pi * 6.0)) - 1
    assert fil["n"] == 43 and fil["first"] == circle["n"]


def test_lu_solve(oracle):
    rng = np.random.default_rng(7)
    for n, bc in [(8, 0), (30, 3), (50, 2)]:
        A = rng.standard_normal((n, n)) + n * np.eye(n) * rng.uniform(0.1, 1)
        b = rng.standard_normal(n)
        x = oracle.lu_solve(A, b, bc)
        assert np.all(x[:bc] == 0.0)
        ref = np.linalg.solve(A[bc:, bc:], b[bc:])
        assert np.allclose(x[bc:], ref, rtol=1e-10, atol=1e-12)


# ------------------------------------------------------------------ FEM

def test_cantilever_static(oracle):
    """Tip load P on a clamped beam: tip deflection P L^3 / (3 E I)."""
    L, h, E, P = 100.0, 2.0, 1e4, 1e-4
    I = 1.0 * h ** 3 / 12
    d = oracle.fem_cantilever(L, h, E, 10.0, 20, 1.0, P)
    assert d == pytest.approx(P * L ** 3 / (3 * E * I), rel=1e-6)


def test_cantilever_period(oracle):
    """Free vibration after release: first bending period
    2 pi / (1.8751^2 sqrt(E I / (rho A L^4))), Newmark (1/4, 1/2) keeps the
    amplitude."""
    L, h, E, rho, P = 100.0, 2.0, 1e4, 10.0, 1e-4
    I, A = h ** 3 / 12, h
    T = 2 * np.pi / (1.8751040687 ** 2 * np.sqrt(E * I / (rho * A * L ** 4)))
    dt = T / 400
    tip = oracle.fem_cantilever(L, h, E, rho, 20, dt, P, nsteps=1600)
    z = np.where(np.diff(np.sign(tip)) != 0)[0]
    tz = [(k + tip[k] / (tip[k] - tip[k + 1])) * dt for k in z]
    assert 2 * np.mean(np.diff(tz)) == pytest.approx(T, rel=1e-2)
    d0 = P * L ** 3 / (3 * E * I)
    assert np.max(tip) <= d0 * 1.001 and np.min(tip) >= -d0 * 1.001


def test_initial_deflection_reaches_target(oracle):
    """INITIAL_DEFLECT (FEMBody.cpp:352-445): the static feedback loop stops
    when the tip's body-frame deflection is -fraction * L0 within its 1e-10
    tolerance; the markers follow the deflected beam."""
    p, g = cfg.inverted_flag(Nx=256, Ny=128, n_el=20, initial_deflect=0.01)
    oc = oracle.OracleCase(p, g)
    U = oc.fem(0)[0]
    nd = len(U)
    a = math.pi  # plate angle 180 degrees
    local_y = -math.sin(a) * U[nd - 3] + math.cos(a) * U[nd - 2]
    assert abs(local_y / p.ref_L + 0.01) <= 1.0001e-10
    pos = oc.markers()["pos"]
    # free end (upstream tip) displaced by +0.01 L in y, clamped end fixed
    assert abs(pos[-1, 1] - (0.5 * (p.Ny - 1) + 0.01 * p.ref_L)) < 1e-3 * p.ref_L
    assert pos[0, 1] == 0.5 * (p.Ny - 1)
    oc.close()
