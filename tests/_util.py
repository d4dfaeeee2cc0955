"""Shared helpers for the parity tests (small cases the oracle runs in seconds)."""
import numpy as np

from life_amd import config as cfg


def rel_err(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.max(np.abs(b))
    return float(np.max(np.abs(a - b)) / (den if den > 0 else 1.0))


def bc_mix_params(kind, Nx=40, Ny=24, cm=0, tau=0.7):
    """Small channels exercising each boundary type of Grid.cpp:302-474."""
    if kind == "ldc":
        p, g = cfg.lid_driven_cavity(N=Ny, tau=tau, u_lid=0.08, central_moments=cm)
        return p.replace(Nx=Nx), g
    if kind == "channel_pressure":
        return cfg.lattice_case(Nx, Ny, tau, wall_left=cfg.VELOCITY, wall_right=cfg.PRESSURE, wall_bottom=cfg.WALL,
                                wall_top=cfg.WALL, profile=cfg.PARABOLIC, uxInlet_p=0.05, central_moments=cm,
                                inlet_ramp=30.0), ""
    if kind == "channel_convective":
        return cfg.lattice_case(Nx, Ny, tau, wall_left=cfg.VELOCITY, wall_right=cfg.CONVECTIVE,
                                wall_bottom=cfg.WALL, wall_top=cfg.FREESLIP, profile=cfg.UNIFORM,
                                uxInlet_p=0.05, central_moments=cm), ""
    if kind == "periodic_force":
        return cfg.lattice_case(Nx, Ny, tau, wall_left=cfg.FLUID, wall_right=cfg.FLUID, wall_bottom=cfg.WALL,
                                wall_top=cfg.FREESLIP, profile=cfg.UNIFORM, uxInlet_p=0.0, dpdx=1e-5,
                                central_moments=cm), ""
    if kind == "shear_velocity":
        return cfg.lattice_case(Nx, Ny, tau, wall_left=cfg.VELOCITY, wall_right=cfg.CONVECTIVE,
                                wall_bottom=cfg.VELOCITY, wall_top=cfg.VELOCITY, profile=cfg.SHEAR,
                                uxInlet_p=0.04, uyInlet_p=0.01, ux0_p=0.02, central_moments=cm), ""
    if kind == "boundary_layer":
        return cfg.lattice_case(Nx, Ny, tau, wall_left=cfg.VELOCITY, wall_right=cfg.PRESSURE, wall_bottom=cfg.WALL,
                                wall_top=cfg.FREESLIP, profile=cfg.BOUNDARY_LAYER, uxInlet_p=0.05,
                                central_moments=cm), ""
    raise KeyError(kind)


BC_KINDS = ["ldc", "channel_pressure", "channel_convective", "periodic_force", "shear_velocity", "boundary_layer"]


def random_state(Nx, Ny, seed, cm=0, amp=0.02):
    """Near-equilibrium random populations with consistent rho/u (AoS)."""
    from oracle import oracle as O
    rng = np.random.default_rng(seed)
    rho = 1.0 + amp * rng.standard_normal((Nx, Ny))
    u = 0.05 * rng.standard_normal((Nx, Ny, 2))
    f = np.zeros((Nx, Ny, 9))
    for i in range(Nx):
        for j in range(Ny):
            f[i, j] = O.equilibrium(cm, rho[i, j], u[i, j, 0], u[i, j, 1])
    f *= 1.0 + 0.01 * rng.standard_normal(f.shape)
    return f, rho, u
