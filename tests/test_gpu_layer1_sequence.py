"""The layer-1 drop-in sequence: a case stepped only through the C-ABI
calls that the patched reference makes (life_grid_create, life_ibm_set_markers,
life_ibm_find_support, life_ibm_epsilon_matrix + life_solve_lapack +
life_ibm_set_epsilon at set-up; per step life_lbm_kernel(g, t),
life_ibm_interpolate, life_ibm_spread, and for moving markers
life_ibm_update_markers + life_ibm_find_support(g, 1) + the epsilon calls
before the interpolation).  The marker vectors are the ones the reference's
own body construction makes (taken from the oracle's restatement of it at
t = 0), as they would stay in the reference's IBMNode vectors.

- Rigid bodies (Objects.cpp:26-60 with hasFlex false: lbmKernel, then
  interpolation, then spreading; main.cpp:70 steps t from 1): against the
  oracle's full steps at 1e-12, with an inlet ramp and with WOMERSLEY forcing
  so that the t passed to life_lbm_kernel matters, and with two bodies as
  two epsilon systems or one (UNI_EPSILON).
- Moving markers (the recomputeObjectVals path, Objects.cpp:152-232, with a
  prescribed filament motion standing in for the FEM update): against the
  oracle's unit steps (lattice kernel, marker update, supports and spacing,
  epsilon, interpolation, spreading) in the same order, at 1e-12.
"""
import numpy as np
import pytest

from life_amd import config as cfg
from life_amd import lib as L
from life_amd.case import Grid
from tests._util import rel_err

pytestmark = pytest.mark.gpu

TOL = 1e-12


def two_cylinders(uni):
    p, _ = cfg.cylinder_channel(Nx=200, Ny=64, D=12.0, centre=(40.0, 31.5), u_max=0.06, Re=40.0, ramp_steps=0.0)
    g = ("CIRCLE 1 0 40.0 24.0 0.0 0.0 6.0\n"
         "CIRCLE 1 1 90.0 38.5 0.0 0.0 5.0\n")
    return p.replace(wall_right=cfg.PRESSURE, uni_epsilon=uni), g


RIGID = {
    "cylinder_ramp": lambda: cfg.cylinder_channel(Nx=160, Ny=64, D=16.0, centre=(40.0, 31.5), u_max=0.06, Re=40.0,
                                                  ramp_steps=20.0),
    "womersley_cylinder": lambda: cfg.womersley_cylinder(),
    "two_cylinders": lambda: two_cylinders(0),
    "two_cylinders_uni": lambda: two_cylinders(1),
}


def systems(p, bodies, n):
    return [(0, n)] if p.uni_epsilon else [(b["first"], b["n"]) for b in bodies]


def layer1_setup(p, oc):
    """initialiseObjects (Objects.cpp:941-961) through the C-ABI."""
    mk = oc.markers()
    bodies = [oc.body_info(b) for b in range(oc.n_bodies)]
    gr = Grid(p, precision=L.LIFE_FP64)
    gr.set_markers(mk["pos"], mk["vel"], mk["body"], [b["flex"] for b in bodies])
    gr.find_support(0)
    layer1_epsilon(gr, p, bodies, len(mk["body"]))
    return gr, bodies


def layer1_epsilon(gr, p, bodies, n):
    """computeEpsilon (Objects.cpp:235-321): one system per body, or one
    over all markers with UNI_EPSILON."""
    eps = np.zeros(n)
    for first, m in systems(p, bodies, n):
        A = gr.epsilon_matrix(first, m)
        eps[first:first + m] = L.solve_lapack(A, np.ones(m))
    gr.set_epsilon(eps)


def compare(gr, oc, force_last):
    f, rho, u, fibm = gr.download()
    fo, rhoo, uo, fibmo = oc.grid()
    mg, mo = gr.markers(), oc.markers()
    errs = {"f": rel_err(f, fo), "rho": rel_err(rho, rhoo), "u": rel_err(u, uo), "fibm": rel_err(fibm, fibmo),
            "pos": rel_err(mg["pos"], mo["pos"]), "force": rel_err(mg["force"], mo["force"]),
            "eps": rel_err(mg["eps"], mo["eps"]), "ds": rel_err(mg["ds"], mo["ds"]),
            "returned_force": rel_err(force_last, mo["force"])}
    for k, e in errs.items():
        assert e <= TOL, (k, e)
    return errs


@pytest.mark.parametrize("name", sorted(RIGID))
def test_rigid_bodies_through_layer1(oracle, gpu_lib, name):
    p, g = RIGID[name]()
    steps = 40
    oc = oracle.OracleCase(p, g)
    gr, bodies = layer1_setup(p, oc)
    if name.startswith("two"):
        assert len(bodies) == 2 and all(b["flex"] == 1 for b in bodies)
    for t in range(1, steps + 1):
        gr.lbm_kernel(t)
        force = gr.interpolate()
        gr.spread()
    oc.step(steps)
    errs = compare(gr, oc, force)
    print("%s: layer-1 sequence vs oracle after %d steps: %s" % (
        name, steps, " ".join("%s %.1g" % kv for kv in errs.items())), flush=True)
    assert np.abs(force).max() > 0
    gr.close()


def filament_motion(pos0, root, L_fil, t, amp=1.5, period=100.0):
    """A prescribed bending of a vertical filament clamped at `root`:
    x displacement amp sin(2 pi t / period) (s / L)^2 with s the height above
    the root, and its time derivative."""
    s = (pos0[:, 1] - root[1]) / L_fil
    w = 2.0 * np.pi / period
    pos = pos0.copy()
    vel = np.zeros_like(pos0)
    pos[:, 0] += amp * np.sin(w * t) * s * s
    vel[:, 0] = amp * w * np.cos(w * t) * s * s
    return pos, vel


def test_moving_markers_through_layer1(oracle, gpu_lib):
    p, g = cfg.canopy(Nx=240, Ny=80, n_filaments=2, L=30.0, n_el=8, min_spacing=40.0, x_range=(60, 160),
                      ramp_steps=0.0, u_in=0.05, kb_range=(1e-2, 1e-1))
    p = p.replace(ux0_p=0.05)
    oc = oracle.OracleCase(p, g)
    gr, bodies = layer1_setup(p, oc)
    assert all(b["flex"] == 0 for b in bodies)
    pos0 = oc.markers()["pos"].copy()
    n = len(pos0)
    roots = {b["first"]: pos0[b["first"]] for b in bodies}
    steps = 30
    for t in range(1, steps + 1):
        pos, vel = pos0.copy(), np.zeros_like(pos0)
        for b in bodies:
            sl = slice(b["first"], b["first"] + b["n"])
            pos[sl], vel[sl] = filament_motion(pos0[sl], roots[b["first"]], 30.0, t)
        # the library: the calls of the patched recomputeObjectVals and objectKernel
        gr.lbm_kernel(t)
        gr.update_markers(pos, vel)
        gr.find_support(1)
        layer1_epsilon(gr, p, bodies, n)
        force = gr.interpolate()
        gr.spread()
        # the oracle: the same restated steps in the same order
        oc.lbm_kernel()
        oc.set_markers(pos=pos, vel=vel)
        oc.ibm_supports()
        eps = np.zeros(n)
        for k, (first, m) in enumerate(systems(p, bodies, n)):
            eps[first:first + m] = oracle.lu_solve(oc.epsilon_matrix(k), np.ones(m))
        oc.set_markers(eps=eps)
        oc.ibm_interpolate()
        oc.ibm_spread()
    errs = compare(gr, oc, force)
    print("moving filaments: layer-1 sequence vs oracle after %d steps: %s" % (
        steps, " ".join("%s %.1g" % kv for kv in errs.items())), flush=True)
    # the markers really moved and the moving-wall velocity entered the forcing
    assert np.abs(gr.markers()["pos"] - pos0).max() > 0.5
    gr.close()
