"""Force coefficients of the Schaefer-Turek cylinder cases
(:func:`life_amd.config.schafer_turek`) measured on the device path: the
coefficients are the reference's own (ObjectsClass::writeTotalForces,
Objects.cpp:517-534: C = F / (0.5 rho_ref U_ref^2 L_ref) with F the sum of
-force * epsilon * ds over the markers), read through life_case_total_force."""
import numpy as np

from . import config as cfg
from . import lib as L
from .case import Case


def lift_strouhal(cl, dt, ref_L, ref_U):
    """Strouhal number f L / U from the upward zero crossings of the lift
    about its mean (linear interpolation between samples dt seconds apart);
    None with fewer than three crossings."""
    z = cl - cl.mean()
    up = [k - z[k] / (z[k + 1] - z[k]) for k in range(len(z) - 1) if z[k] < 0.0 <= z[k + 1]]
    if len(up) < 3:
        return None
    period = (up[-1] - up[0]) / (len(up) - 1) * dt
    return ref_L / (ref_U * period)


def schafer_turek_run(test, res_factor, seconds=20.0, precision=L.LIFE_FP64, sample_fraction=0.25):
    """Run one Schaefer-Turek case for `seconds` of physical time.

    2D-1: the drag and lift coefficients at the end and their change over
    the last physical second (steadiness).  2D-2: the maximum drag and lift
    coefficients and the Strouhal number over the last `sample_fraction` of
    the run, sampled every r^2 steps (about 200 samples per shedding period).
    Also returns the steps, lattice size and wall time."""
    import time
    p, g = cfg.schafer_turek(test, res_factor)
    steps = int(round(seconds / p.tStep))
    c = Case(p, g, precision=precision)
    t0 = time.perf_counter()
    out = {"test": test, "res_factor": int(res_factor), "D_lattice": 10 * int(res_factor), "Nx": p.Nx, "Ny": p.Ny,
           "steps": steps, "seconds": seconds}
    try:
        if test == "2D-1":
            last = int(round(1.0 / p.tStep))
            c.step(steps - last)
            a = c.total_force()
            c.step(last)
            b = c.total_force()
            out.update(cd=float(b[2]), cl=float(b[3]), cd_change_last_second=float(abs(b[2] - a[2])))
        else:
            every = int(res_factor) ** 2
            window = int(steps * sample_fraction) // every
            c.step(steps - window * every)
            cd, cl = np.empty(window), np.empty(window)
            for k in range(window):
                c.step(every)
                f = c.total_force()
                cd[k], cl[k] = f[2], f[3]
            out.update(cd_max=float(cd.max()), cl_max=float(cl.max()),
                       strouhal=lift_strouhal(cl, every * p.tStep, p.ref_L, p.ref_U))
    finally:
        c.close()
    out["wall_s"] = time.perf_counter() - t0
    return out


# Published FSI2 results (Turek & Hron 2006, "Proposal for numerical
# benchmarking of fluid-structure interaction between an elastic object and
# laminar incompressible flow"): displacement of point A (the trailing edge
# of the flag) in mm as mean +- amplitude, and the frequency of u_y in Hz.
TUREK_HRON_FSI2 = {"ux_mean": -14.58, "ux_amp": 12.44, "uy_mean": 1.23, "uy_amp": 80.6, "f_y": 2.0}


def turek_hron_run(res_factor, seconds=12.0, window_s=3.0, precision=L.LIFE_FP64, sample_s=0.005):
    """The reference's TurekHron example (examples/TurekHron/params.h) at
    resolution factor r (Nx = 250 r + 1, Ny = 41 r + 1, tStep = 0.001 / r^2):
    displacement of the flag's last marker (point A, at (0.6, 0.2)) sampled
    every `sample_s` seconds; mean and amplitude of u_x and u_y in mm over
    the last `window_s` seconds and the frequencies of u_y and u_x from their
    upward crossings of the window mean."""
    import time
    p, g = cfg.reference_example("TurekHron")
    r = int(res_factor)
    p = p.replace(Nx=r * 250 + 1, Ny=r * 41 + 1, tStep=0.001 / (r * r))
    c = Case(p, g, precision=precision)
    steps = int(round(seconds / p.tStep))
    every = int(round(sample_s / p.tStep))
    t0 = time.perf_counter()
    try:
        b = c.body_info(1)
        tip = b["first"] + b["n"] - 1
        a0 = c.markers()["pos"][tip].copy()
        ts, ux, uy = [], [], []
        for k in range(steps // every):
            c.step(every)
            q = c.markers()["pos"][tip]
            ts.append((k + 1) * every * p.tStep)
            ux.append(q[0] - a0[0])
            uy.append(q[1] - a0[1])
    finally:
        c.close()
    ts, ux, uy = np.array(ts), 1e3 * np.array(ux), 1e3 * np.array(uy)
    w = ts > ts[-1] - window_s

    def freq(v):
        z = v[w] - v[w].mean()
        tw = ts[w]
        up = [tw[k] - z[k] * (tw[k + 1] - tw[k]) / (z[k + 1] - z[k]) for k in range(len(z) - 1) if z[k] < 0.0 <= z[k + 1]]
        return (len(up) - 1) / (up[-1] - up[0]) if len(up) > 2 else None

    return {"res_factor": r, "steps": steps, "point_A": a0.tolist(),
            "ux_mean": float((ux[w].max() + ux[w].min()) / 2), "ux_amp": float((ux[w].max() - ux[w].min()) / 2),
            "uy_mean": float((uy[w].max() + uy[w].min()) / 2), "uy_amp": float((uy[w].max() - uy[w].min()) / 2),
            "f_y": freq(uy), "f_x": freq(ux), "wall_s": time.perf_counter() - t0}
