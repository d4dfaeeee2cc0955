"""Throughput benchmark of the LIFE fluid + coupling hot path on one MI355X.

Workload (default): the canopy case -- D2Q9 BGK on an 8192 x 4096 lattice
with FP32-shifted populations, uniform inlet, convective outlet, no-slip
bottom, free-slip top, and 64 clamped flexible filaments (length 200, 201
markers and 50 corotational beam elements each) coupled through the
implicit immersed boundary, with the FP64 FEM on the device (one workgroup
per filament; LIFE_DEVICE_FEM=0 selects the host FEM, which the FP64
populations use by default).  A step is one full time step (lattice step +
coupling subiterations + spread).

Prints one JSON line: MLUPS of the whole step (lattice sites x steps /
wall time), the roofline of the fused stream-collide kernel (72 algorithmic
bytes per site update over its HIP-event-timed average launch), the FP32 vs
FP64 velocity error on the same run, and a CPU baseline (the oracle, FP64,
OpenMP, on a bounded number of steps of the same workload).

--workload serve measures the serving engine instead: a seeded Poisson
stream of independent inverted-flag cases (arrivals, sizes, Re, stiffness
and element counts drawn as described in life_amd.config.serving_stream)
run by the persistent request loop with FP32-shifted populations, reporting
cases/hour, p50/p99 time-to-solution (arrival to finish) and aggregate
MLUPS; beside it, a prefix of the same stream is run both by the engine and
by the plain FP64 path one case at a time in arrival order, and the FP32
results are compared with the FP64 ones (tip position, force).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
BYTES_PER_LU = {"f32": 72.0, "f64": 144.0}


def workload(name):
    from life_amd import config as cfg
    if name == "canopy8192":
        p, g = cfg.canopy()
        return p, g, dict(workload="canopy_8192x4096_bgk", Nx=p.Nx, Ny=p.Ny, filaments=64, filament_length=200,
                          fem_elements=50, markers_per_filament=201, filament_thickness=4.0, density_ratio=50.0,
                          tau=0.6, u_in=0.05, seed=20261202)
    if name == "canopy8192_rigidfree":
        p, g = cfg.canopy(n_filaments=0)
        return p, g, dict(workload="channel_8192x4096_bgk_no_bodies", Nx=p.Nx, Ny=p.Ny, tau=0.6, u_in=0.05)
    raise SystemExit("unknown workload %s" % name)


def read_traffic(kernel, wl):
    """HBM bytes per launch measured by rocprofv3 --pmc (FETCH_SIZE +
    WRITE_SIZE, scaled by a copy kernel of known size), recorded in
    profiles/traffic.txt as 'kernel workload bytes'."""
    path = os.path.join(ROOT, "profiles", "traffic.txt")
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        for line in fh:
            parts = line.split()
            if len(parts) == 3 and parts[0] == kernel and parts[1] == wl:
                return float(parts[2])
    return None


def cpu_baseline(p, g, steps, threads):
    from oracle import oracle as O
    O.set_threads(threads)
    c = O.OracleCase(p, g)
    c.step(1)  # first step uses the initial arrays; time the steady path
    t0 = time.perf_counter()
    c.step(steps)
    dt = time.perf_counter() - t0
    c.close()
    return p.Nx * p.Ny * steps / dt / 1e6


def cpu_model():
    """The host CPU as lscpu names it (model name, sockets x cores x
    threads), falling back to /proc/cpuinfo."""
    import subprocess
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        kv = dict(line.split(":", 1) for line in out.splitlines() if ":" in line)
        kv = {k.strip(): v.strip() for k, v in kv.items()}
        if "Model name" in kv:
            return "%s (%s CPUs online; %s socket(s) x %s core(s) x %s thread(s))" % (
                kv["Model name"], kv.get("CPU(s)", "?"), kv.get("Socket(s)", "?"),
                kv.get("Core(s) per socket", "?"), kv.get("Thread(s) per core", "?"))
    except (OSError, subprocess.SubprocessError, ValueError):
        pass
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_threads():
    """Threads for the CPU baseline: the host's CPU share for this GPU
    (OMP_NUM_THREADS where the job sets it), else every online CPU."""
    v = os.environ.get("OMP_NUM_THREADS")
    if v and v.isdigit() and int(v) > 0:
        return int(v)
    return os.cpu_count() or 1


def calibrate_copy(nbytes, reps):
    """Device-to-device copies of a known size, for scaling the rocprofv3
    FETCH_SIZE / WRITE_SIZE counters (run under rocprofv3 --pmc)."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(C.byref(b), C.c_size_t(nbytes)) == 0
    assert hip.hipMemset(a, 1, C.c_size_t(nbytes)) == 0
    for _ in range(reps):
        assert hip.hipMemcpy(b, a, C.c_size_t(nbytes), 3) == 0  # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    print("copied %d bytes x %d" % (nbytes, reps))


def serve_summary(srv, ids, wall):
    import numpy as np
    from life_amd import lib
    infos = [srv.info(i) for i in ids]
    done = [x for x in infos if x["status"] == lib.LIFE_REQ_DONE]
    tts = np.array([x["t_finish"] - x["t_arrival"] for x in done]) if done else np.zeros(1)
    st = srv.stats()
    return {"cases": len(infos), "done": len(done), "wall_s": wall,
            "cases_per_hour": len(done) / wall * 3600.0,
            "p50_s": float(np.percentile(tts, 50)), "p99_s": float(np.percentile(tts, 99)),
            "mean_s": float(tts.mean()), "mlups": st["site_updates"] / wall / 1e6,
            "mlups_busy": st["site_updates"] / max(st["busy_s"], 1e-9) / 1e6,
            "engine_steps": st["steps"], "ms_per_engine_step": st["busy_s"] * 1e3 / max(st["steps"], 1.0),
            "graph_captures": st["graph_captures"], "peak_running": st["peak_running"],
            "peak_bytes": st["peak_bytes"]}, infos


def run_one_at_a_time(stream, prec):
    """Plain path, one case at a time in arrival order: case k starts at
    max(its arrival, the previous finish)."""
    import numpy as np
    from life_amd.case import Case
    t0 = time.perf_counter()
    out = []
    for r in stream:
        now = time.perf_counter() - t0
        if now < r["arrival"]:
            time.sleep(r["arrival"] - now)
        c = Case(r["params"], r["geometry"], precision=prec)
        c.step(r["steps"])
        m = c.markers()["pos"].reshape(-1, 2)
        out.append({"t_finish": time.perf_counter() - t0, "t_arrival": r["arrival"], "tip": m[-1].copy(),
                    "force": c.total_force().copy(), "L": r["params"].ref_L})
        c.close()
        print("bench: one-at-a-time case %d/%d done" % (len(out), len(stream)), file=sys.stderr, flush=True)
    wall = time.perf_counter() - t0
    tts = np.array([x["t_finish"] - x["t_arrival"] for x in out])
    return {"cases": len(out), "wall_s": wall, "cases_per_hour": len(out) / wall * 3600.0,
            "p50_s": float(np.percentile(tts, 50)), "p99_s": float(np.percentile(tts, 99)),
            "mlups": sum(r["Nx"] * r["Ny"] * r["steps"] for r in stream) / wall / 1e6}, out


def _tip_force_err(a_tip, a_C, b_tip, b_C, L):
    import numpy as np
    return (float(np.linalg.norm(np.asarray(a_tip[:2]) - np.asarray(b_tip[:2])) / L),
            float(np.abs(np.asarray(a_C) - np.asarray(b_C)).max()))


# Accuracy bars of the FP32-shifted served results against plain FP64 at the
# short horizon (before the flapping flags' trajectories decorrelate): tip
# position error over the plate length, and force-coefficient error.
SERVE_TIP_BAR = 1e-6
SERVE_DC_BAR = 1e-3


def serve_accuracy(cases, pinfo, pout, n_short, short_steps, slots, budget):
    """FP32-shifted (served) against plain FP64 (one at a time) per case.

    Short horizon (short_steps, the first n_short cases): held to the bars
    SERVE_TIP_BAR and SERVE_DC_BAR.  Full horizon: the flapping flags are
    chaotic, so the FP32 errors are given beside plain FP64 against itself
    with the inflow perturbed by 1e-12 (relative) on the same cases -- the
    sensitivity of the exact method to a rounding-level change."""
    from life_amd import lib
    from life_amd.case import Case
    from life_amd.serve import Server
    full = [_tip_force_err(a["tip"], a["total_force"][2:], b["tip"], b["force"][2:], b["L"])
            for a, b in zip(pinfo, pout)]
    m = min(n_short, len(cases))
    pert = []
    for r, b in zip(cases[:m], pout[:m]):
        p = r["params"].replace(uxInlet_p=r["params"].uxInlet_p * (1.0 + 1e-12))
        c = Case(p, r["geometry"], precision=lib.LIFE_FP64)
        c.step(r["steps"])
        tip = c.markers()["pos"].reshape(-1, 2)[-1].copy()
        pert.append(_tip_force_err(tip, c.total_force()[2:], b["tip"], b["force"][2:], b["L"]))
        c.close()
    srv = Server(precision=lib.LIFE_FP32_SHIFTED, max_cases=slots, hbm_budget=budget)
    ids = [srv.submit(r["params"], r["geometry"], short_steps) for r in cases[:m]]
    srv.drain()
    short = []
    for rid, r in zip(ids, cases[:m]):
        a = srv.info(rid)
        c = Case(r["params"], r["geometry"], precision=lib.LIFE_FP64)
        c.step(short_steps)
        tip = c.markers()["pos"].reshape(-1, 2)[-1].copy()
        short.append(_tip_force_err(a["tip"], a["total_force"][2:], tip, c.total_force()[2:], r["params"].ref_L))
        c.close()
    srv.close()
    mx = lambda v, i: max(x[i] for x in v)
    return {"short_horizon": {"steps": short_steps, "cases": m, "max_tip_err_over_L": mx(short, 0),
                              "max_abs_dC": mx(short, 1), "bars": {"tip_err_over_L": SERVE_TIP_BAR,
                                                                   "abs_dC": SERVE_DC_BAR},
                              "within_bars": mx(short, 0) <= SERVE_TIP_BAR and mx(short, 1) <= SERVE_DC_BAR},
            "full_horizon": {"steps": cases[0]["steps"], "cases": len(full),
                             "fp32_vs_fp64_max_tip_err_over_L": mx(full, 0), "fp32_vs_fp64_max_abs_dC": mx(full, 1),
                             "perturbed_cases": m,
                             "fp32_vs_fp64_first_cases_max_tip_err_over_L": mx(full[:m], 0),
                             "fp32_vs_fp64_first_cases_max_abs_dC": mx(full[:m], 1),
                             "fp64_perturbed_1e-12_vs_fp64_max_tip_err_over_L": mx(pert, 0),
                             "fp64_perturbed_1e-12_vs_fp64_max_abs_dC": mx(pert, 1)}}


def serve_run(cases, slots, budget, prec=None):
    from life_amd import lib
    from life_amd.serve import Server
    srv = Server(precision=lib.LIFE_FP32_SHIFTED if prec is None else prec, max_cases=slots, hbm_budget=budget)
    t0 = time.perf_counter()
    ids = srv.run(cases)
    wall = time.perf_counter() - t0
    summary, infos = serve_summary(srv, ids, wall)
    srv.close()
    return summary, infos


def serving_comparison(cases, slots, budget, n_short=8, short_steps=300, capacity=True, accuracy=True):
    """The serving engine against the plain FP64 path one case at a time,
    on the same cases: the engine on the timed stream (arrivals as drawn),
    the engine with every case queued at once (its own throughput, not
    bounded by arrivals), and the plain path in arrival order (a case
    starts at max(its arrival, the previous finish))."""
    from life_amd import lib
    print("bench: serving %d cases x %d steps (arrival stream)" % (len(cases), cases[0]["steps"]),
          file=sys.stderr, flush=True)
    served, sinfo = serve_run(cases, slots, budget)
    cap = None
    if capacity:
        print("bench: serving the same cases queued at once", file=sys.stderr, flush=True)
        cap, _ = serve_run([dict(r, arrival=0.0) for r in cases], slots, budget)
    print("bench: plain FP64 one case at a time", file=sys.stderr, flush=True)
    plain, pout = run_one_at_a_time(cases, lib.LIFE_FP64)
    out = {"cases": len(cases), "steps_per_case": cases[0]["steps"], "served_fp32": served,
           "plain_fp64_one_at_a_time": plain,
           "throughput_ratio": served["cases_per_hour"] / plain["cases_per_hour"],
           "p99_ratio": plain["p99_s"] / served["p99_s"]}
    if cap is not None:
        out["served_fp32_all_queued_at_once"] = cap
        out["capacity_throughput_ratio"] = cap["cases_per_hour"] / plain["cases_per_hour"]
    if accuracy:
        print("bench: serving accuracy", file=sys.stderr, flush=True)
        out["accuracy"] = serve_accuracy(cases, sinfo, pout, n_short, short_steps, slots, budget)
    return out


def bench_serve(args):
    from life_amd import config as cfg

    stream = cfg.serving_stream(n=args.serve_cases, steps=args.serve_steps)
    wl = serve_workload(args, len(stream))
    if args.capacity_only:
        # the engine alone with every case queued at once (for profiling)
        cap, _ = serve_run([dict(r, arrival=0.0) for r in stream], args.serve_slots, args.serve_budget)
        print(json.dumps({"metric": "FSI cases/hour (serving engine, all cases queued at once)",
                          "value": cap["cases_per_hour"], "unit": "cases/hour", "n_gpus": 1,
                          "steps": args.serve_steps, "warmup": 0, "ms_per_step": cap["ms_per_engine_step"],
                          "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
                          "data": "synthetic", "config": wl, "serving": {"served_fp32_all_queued_at_once": cap}}))
        return
    base = args.baseline_cases if args.baseline_cases > 0 else len(stream)
    cmp_ = serving_comparison(stream[:base], args.serve_slots, args.serve_budget, args.accuracy_cases,
                              args.accuracy_steps)
    served = cmp_["served_fp32"]
    out = {
        "metric": "FSI cases/hour (serving engine, continuous batching)",
        "value": served["cases_per_hour"],
        "unit": "cases/hour",
        "n_gpus": 1,
        "steps": args.serve_steps,
        "warmup": 0,
        "ms_per_step": served["ms_per_engine_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": dict(wl, cases=base),
        "serving": cmp_,
    }
    print(json.dumps(out))


def serve_workload(args, n):
    return {"workload": "serve_inverted_flag_stream", "cases": n, "steps_per_case": args.serve_steps,
            "arrival_rate_per_s": 4.0, "widths": [512, 1024, 2048], "heights": [128, 256, 512],
            "plate_length": "0.25*height", "fem_elements": [20, 60], "Re": [50, 400], "K_B": [0.1, 1.0],
            "collision": "central_moments", "seed": 20261203, "slots": args.serve_slots,
            "hbm_budget_bytes": args.serve_budget, "precision": "fp32_shifted populations, fp64 FEM/coupling"}


CONFIGS = {
    # name: (builder, default steps, oracle parity steps, description)
    "ldc256": ("lid_driven_cavity", 20000, 200, "lid-driven cavity 256x256, BGK, tau 0.6, lid 0.05"),
    "cylinder1024": ("cylinder_channel", 50000, 100,
                     "rigid cylinder D=40 in a 1024x256 channel, central moments, Re 100, parabolic inlet"),
    "fsi2048": ("flexible_fsi", 4000, 30,
                "cylinder D=80 + flexible plate L=280 h=16 (40 elements) in 2048x410, central moments, Re 100"),
}


def coefficient_summary(series, p):
    """Force coefficients over the second half of a run: mean drag, lift
    amplitude and rms, and the shedding Strouhal number from the lift's
    upward zero crossings (St = f D / U with the case's reference length and
    velocity, time in units of Dt)."""
    import numpy as np
    t, cd, cl = series[:, 0], series[:, 3], series[:, 4]
    h = len(t) // 2
    t, cd, cl = t[h:], cd[h:], cl[h:]
    out = {"samples": int(len(t)), "from_step": int(t[0]), "to_step": int(t[-1]), "cd_mean": float(cd.mean()),
           "cl_mean": float(cl.mean()), "cl_rms": float(np.sqrt(np.mean((cl - cl.mean()) ** 2))),
           "cl_amplitude": float((cl.max() - cl.min()) / 2.0)}
    z = cl - cl.mean()
    up = [t[k] - z[k] * (t[k + 1] - t[k]) / (z[k + 1] - z[k]) for k in range(len(z) - 1) if z[k] < 0.0 <= z[k + 1]]
    if len(up) >= 3:
        period = (up[-1] - up[0]) / (len(up) - 1)  # in steps
        out["strouhal"] = float(p.ref_L / (p.ref_U * period * time_step(p)))
        out["shedding_periods"] = len(up) - 1
    return out


def time_step(p):
    """Dt = Dx^2 nu_lattice / nu_p (Grid.cpp:1255-1258), omega as params.h sets it."""
    from life_amd import config as cfg
    Dx = p.height_p / (p.Ny - 1)
    omega = p.omega
    if omega <= 0:
        cs = 1.0 / 3.0 ** 0.5
        omega = 1.0 / (p.nu_p * p.tStep / ((cs * cs) * (Dx * Dx)) + 0.5)
    return Dx * Dx * cfg.lattice_viscosity(omega) / p.nu_p


def bench_config(args):
    """One of the reference-style validation configurations on its full
    lattice: FP64 (plain) and FP32-shifted runs of the same steps, MLUPS of
    each and the FP32 results against the FP64 ones.  (FP64 parity with the
    oracle on these full lattices is a GPU test, tests/test_gpu_configs.py.)"""
    import numpy as np
    from life_amd import config as cfg
    from life_amd import lib
    from life_amd.case import Case

    builder, dsteps, _, desc = CONFIGS[args.workload]
    steps = args.steps if args.steps_given else dsteps
    p, g = getattr(cfg, builder)()
    N = p.Nx * p.Ny
    runs = {}
    # bodies: the force coefficients are sampled about 1000 times over the
    # run (one host synchronisation per sample)
    chunk = max(1, steps // 1000)
    for name, prec in (("f64", lib.LIFE_FP64), ("f32", lib.LIFE_FP32_SHIFTED)):
        c = Case(p, g, precision=prec)
        series = []
        t0 = time.perf_counter()
        if c.n_markers:
            done = 0
            while done < steps:
                k = min(chunk, steps - done)
                c.step(k)
                done += k
                series.append([done] + list(c.total_force()))
        else:
            c.step(steps)
        wall = time.perf_counter() - t0
        _, rho, u, _ = c.fields()
        r = {"mlups": N * steps / wall / 1e6, "wall_s": wall, "rho": rho, "u": u, "series": np.array(series)}
        if c.n_markers:
            r["force"] = c.total_force()
            r["tip"] = c.markers()["pos"].reshape(-1, 2)[-1].copy()
        runs[name] = r
        c.close()
        print("bench: %s %s %d steps in %.2f s" % (args.workload, name, steps, wall), file=sys.stderr, flush=True)
    a, b = runs["f32"], runs["f64"]
    res = {"u_l2_rel_f32_vs_f64": float(np.linalg.norm(a["u"] - b["u"]) / np.linalg.norm(b["u"])),
           "rho_max_abs_f32_vs_f64": float(np.abs(a["rho"] - b["rho"]).max())}
    if args.workload == "ldc256":
        m = p.Nx // 2
        res["centreline_u_f64"] = b["u"][m, :, 0][::16].tolist()
        res["centreline_v_f64"] = b["u"][:, m, 1][::16].tolist()
        res["centreline_max_abs_f32_vs_f64"] = float(max(np.abs(a["u"][m, :, 0] - b["u"][m, :, 0]).max(),
                                                       np.abs(a["u"][:, m, 1] - b["u"][:, m, 1]).max()))
    if "force" in b:
        res["coefficients_f64"] = coefficient_summary(b["series"], p)
        res["coefficients_f32"] = coefficient_summary(a["series"], p)
        res["force_f64"] = b["force"].tolist()
        res["force_rel_err_f32"] = float(np.linalg.norm(a["force"][:2] - b["force"][:2]) /
                                         np.linalg.norm(b["force"][:2]))
        res["tip_f64"] = b["tip"].tolist()
        res["tip_err_f32_over_L"] = float(np.linalg.norm(a["tip"] - b["tip"]) / p.ref_L)
    out = {"metric": "D2Q9 MLUPS (plain FP64, full step incl. coupling)", "value": runs["f64"]["mlups"],
           "unit": "MLUPS", "n_gpus": 1, "steps": steps, "warmup": 0,
           "ms_per_step": runs["f64"]["wall_s"] * 1e3 / steps, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": args.workload, "description": desc, "Nx": p.Nx, "Ny": p.Ny},
           "mlups_f32_shifted": runs["f32"]["mlups"], "results": res}
    print(json.dumps(out))


def bench_schafer_turek(args):
    """The Schaefer-Turek cylinder benchmarks on the reference's Cylinder
    example by resolution factor (D = 10 r lattice units): 2D-1 (Re 20,
    30 s) at r = 2 and 4, 2D-2 (Re 100, 20 s) at r = 2, 4 and 8 on the plain
    FP64 path, and 2D-2 at r = 4 on the FP32-shifted path; each coefficient
    beside its published bounds.  The headline number is the MLUPS of the
    largest FP64 run (1761 x 329, 1.28 M steps)."""
    from life_amd import lib
    from life_amd.config import SCHAFER_TUREK_BOUNDS
    from life_amd.validation import schafer_turek_run

    runs = []
    for test, r, secs, prec in (("2D-1", 2, 30.0, lib.LIFE_FP64), ("2D-1", 4, 30.0, lib.LIFE_FP64),
                                ("2D-2", 2, 20.0, lib.LIFE_FP64), ("2D-2", 4, 20.0, lib.LIFE_FP64),
                                ("2D-2", 8, 20.0, lib.LIFE_FP64), ("2D-2", 4, 20.0, lib.LIFE_FP32_SHIFTED)):
        out = schafer_turek_run(test, r, seconds=secs, precision=prec)
        out["precision"] = "f64" if prec == lib.LIFE_FP64 else "f32_shifted"
        out["mlups"] = out["Nx"] * out["Ny"] * out["steps"] / out["wall_s"] / 1e6
        runs.append(out)
        print("bench: schafer_turek %s" % json.dumps(out), file=sys.stderr, flush=True)
    big = runs[4]
    out = {"metric": "D2Q9 MLUPS (plain FP64, full step incl. coupling)", "value": big["mlups"], "unit": "MLUPS",
           "n_gpus": 1, "steps": big["steps"], "warmup": 0, "ms_per_step": big["wall_s"] * 1e3 / big["steps"],
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": "schafer_turek", "Nx": big["Nx"], "Ny": big["Ny"],
                      "description": "Schaefer-Turek 2D-1/2D-2 on the reference's Cylinder example, r = 2, 4, 8"},
           "results": {"published_bounds": SCHAFER_TUREK_BOUNDS, "runs": runs}}
    print(json.dumps(out))


def main():
    if "--calibrate-copy" in sys.argv:
        calibrate_copy(1 << 30, 5)
        return
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", default="canopy8192")
    ap.add_argument("--precision", default="f32", choices=["f32", "f64"])
    ap.add_argument("--spinup", type=int, default=-1,
                    help="steps run before the warmup (default 5000: past the canopy's 2000-step inlet ramp and "
                         "into the loaded flow, where every step takes two coupling subiterations)")
    ap.add_argument("--horizon", type=int, default=10000,
                    help="step at which the FP32 vs FP64 velocity error is also reported (0: off)")
    ap.add_argument("--no-accuracy", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=30)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-serve", action="store_true", help="canopy: leave out the serving leg")
    ap.add_argument("--capacity-only", action="store_true",
                    help="serve: the engine alone, every case queued at once (no plain path, no accuracy)")
    ap.add_argument("--serve-cases", type=int, default=256)
    ap.add_argument("--serve-steps", type=int, default=5000)
    ap.add_argument("--serve-slots", type=int, default=128)
    ap.add_argument("--serve-budget", type=float, default=200e9)
    ap.add_argument("--bench-serve-cases", type=int, default=32,
                    help="canopy: cases of the serving stream in the serving leg")
    ap.add_argument("--baseline-cases", type=int, default=0,
                    help="serve: cases run by both the engine and the plain path (0: all)")
    ap.add_argument("--accuracy-cases", type=int, default=8)
    ap.add_argument("--accuracy-steps", type=int, default=300)
    args = ap.parse_args()
    args.steps_given = any(a == "--steps" or a.startswith("--steps=") for a in sys.argv[1:])
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: the library runs each case on a single MI355X "
              "(multi-GPU sharding is a design note, not an implementation).", file=sys.stderr)
        sys.exit(2)
    # the CPU baseline's OpenMP threads stay on neighbouring cores (read by
    # the OpenMP runtime when the first library that uses it loads)
    os.environ.setdefault("OMP_PROC_BIND", "close")

    if args.workload == "serve":
        bench_serve(args)
        return
    if args.workload in CONFIGS:
        bench_config(args)
        return
    if args.workload == "schafer_turek":
        bench_schafer_turek(args)
        return

    import numpy as np

    from life_amd import lib
    from life_amd.case import Case

    p, geom, wcfg = workload(args.workload)
    prec = lib.LIFE_FP32_SHIFTED if args.precision == "f32" else lib.LIFE_FP64
    # the flow reaches the filaments after the inlet ramp: from about step
    # 4000 the coupling takes a second subiteration, from 4500 at every step
    spin = args.spinup if args.spinup >= 0 else (5000 if p.inlet_ramp > 0 else 0)
    N = p.Nx * p.Ny
    print("bench: building %s (%s)" % (wcfg["workload"], args.precision), file=sys.stderr, flush=True)
    case = Case(p, geom, precision=prec)
    # past the inlet ramp and into the loaded flow: the timed steps carry
    # two coupling subiterations each, whatever --warmup says
    case.step(spin)
    case.step(args.warmup)
    print("bench: spin-up %d + warmup %d steps done" % (spin, args.warmup), file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    ms_dev, ms_kernel = case.step_timed(args.steps)
    wall = time.perf_counter() - t0
    t_end = case.t
    mlups = N * args.steps / wall / 1e6
    k_avg_s = ms_kernel / 1e3 / args.steps
    algo_bytes = BYTES_PER_LU[args.precision] * N
    achieved = algo_bytes / k_avg_s / 1e9
    kernel = "k_push32" if args.precision == "f32" else "k_push64"
    traffic = read_traffic(kernel, wcfg["workload"])
    subits = case.coupling()["subIt"]
    print("bench: %d steps in %.3f s wall (%.3f ms device), K1 avg %.3f ms, %d subiterations in the last step" %
          (args.steps, wall, ms_dev, k_avg_s * 1e3, subits), file=sys.stderr, flush=True)
    u32 = None
    if not args.no_accuracy and args.precision == "f32":
        _, _, u32, _ = case.fields()
    # four more timed regions of the same length: the spread of the
    # full-step rate and of the kernel average, reported beside the
    # contract's single region (short runs only)
    reps = [(mlups, k_avg_s * 1e3)]
    for _ in range(4 if args.steps <= 500 else 0):
        t1 = time.perf_counter()
        _, mk = case.step_timed(args.steps)
        reps.append((N * args.steps / (time.perf_counter() - t1) / 1e6, mk / args.steps))
    # the same kernel with nothing beside it: a few steps without the
    # overlap of the next step's stream-collide and the coupling
    alone_steps = min(args.steps, 20)
    saved = os.environ.get("LIFE_OVERLAP")
    os.environ["LIFE_OVERLAP"] = "0"
    _, ms_alone = case.step_timed(alone_steps)
    if saved is None:
        del os.environ["LIFE_OVERLAP"]
    else:
        os.environ["LIFE_OVERLAP"] = saved
    k_alone_s = ms_alone / 1e3 / alone_steps

    accuracy = None
    fp64 = None
    if u32 is not None:
        # the same case with FP64 populations (the plain method, equal to the
        # oracle): velocity error at the end of the timed window, the FP64
        # stream-collide timed over the same window, and the error again at
        # the horizon
        ref = Case(p, geom, precision=lib.LIFE_FP64)
        ref.step(t_end - args.steps)
        t1 = time.perf_counter()
        _, k64 = ref.step_timed(args.steps)
        w64 = time.perf_counter() - t1
        _, _, u64, _ = ref.fields()
        accuracy = {"u_l2_rel_vs_fp64": float(np.linalg.norm(u32 - u64) / np.linalg.norm(u64)), "steps": t_end}
        # the yardstick: FP64 against itself with the inflow perturbed by
        # 1e-12 (relative), at the same steps (once the flow reaches the
        # filaments, their motion amplifies any difference)
        print("bench: FP64 with the inflow perturbed by 1e-12", file=sys.stderr, flush=True)
        per = Case(p.replace(uxInlet_p=p.uxInlet_p * (1.0 + 1e-12)), geom, precision=lib.LIFE_FP64)
        per.step(t_end)
        _, _, up, _ = per.fields()
        accuracy["fp64_perturbed_1e-12_vs_fp64"] = float(np.linalg.norm(up - u64) / np.linalg.norm(u64))
        del u32, u64, up
        k64_s = k64 / 1e3 / args.steps
        fp64 = {"mlups": N * args.steps / w64 / 1e6, "ms_per_step": w64 * 1e3 / args.steps,
                "roofline": {"bound": "hbm", "kernel": "k_push64", "kernel_ms": k64_s * 1e3,
                             "achieved": BYTES_PER_LU["f64"] * N / k64_s / 1e9,
                             "frac": BYTES_PER_LU["f64"] * N / k64_s / 1e9 / HBM_PEAK_GBS}}
        if args.horizon > case.t:
            print("bench: FP32 and FP64 to step %d" % args.horizon, file=sys.stderr, flush=True)
            case.step(args.horizon - case.t)
            ref.step(args.horizon - ref.t)
            per.step(args.horizon - per.t)
            _, _, a, _ = case.fields()
            _, _, b, _ = ref.fields()
            _, _, c, _ = per.fields()
            accuracy["horizon"] = {"steps": args.horizon,
                                   "u_l2_rel_vs_fp64": float(np.linalg.norm(a - b) / np.linalg.norm(b)),
                                   "fp64_perturbed_1e-12_vs_fp64": float(np.linalg.norm(c - b) / np.linalg.norm(b)),
                                   "subiterations_last_step": case.coupling()["subIt"]}
            del a, b, c
        per.close()
        ref.close()
    case.close()

    cpu = None
    if not args.no_cpu:
        threads = cpu_threads()
        print("bench: CPU baseline (oracle, FP64, %d threads, %d steps)" % (threads, args.cpu_steps),
              file=sys.stderr, flush=True)
        v = cpu_baseline(p, geom, args.cpu_steps, threads)
        cpu = {"value": v, "unit": "MLUPS", "cores": threads, "kind": "port",
               "sample": "%d full steps (lattice + coupling, FP64) of the same workload from its start, timed "
                         "after the first" % args.cpu_steps, "cpu_model": cpu_model()}

    serving = None
    if not args.no_serve and args.bench_serve_cases > 0:
        from life_amd import config as cfg
        stream = cfg.serving_stream(n=args.bench_serve_cases, steps=args.serve_steps)
        serving = dict(serve_workload(args, len(stream)),
                       **serving_comparison(stream, args.serve_slots, args.serve_budget, args.accuracy_cases,
                                            args.accuracy_steps))

    out = {
        "metric": "D2Q9 MLUPS (full step incl. IBM/FEM coupling)",
        "value": mlups,
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": args.precision,
        "data": "synthetic",
        "config": dict(wcfg, precision=("fp32_shifted" if args.precision == "f32" else "fp64"),
                       spinup_steps=spin, timed_from_step=t_end - args.steps, subiterations_last_step=subits),
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "kernel": kernel,
                     "kernel_ms": k_avg_s * 1e3, "kernel_mlups": N / k_avg_s / 1e6,
                     "live": "beside the coupling chain; with the FEM on the device the bulk kernel runs on "
                             "XCDs 2-7 (192 of 256 CUs) and the chain on XCDs 0-1 (LIFE_COUPLING_XCDS)",
                     "alone": {"steps": alone_steps, "kernel_ms": k_alone_s * 1e3,
                               "achieved": algo_bytes / k_alone_s / 1e9,
                               "frac": algo_bytes / k_alone_s / 1e9 / HBM_PEAK_GBS}},
        "cpu_baseline": cpu,
        "accuracy": accuracy,
        "fp64_same_window": fp64,
        "repeats": {"n": len(reps), "mlups": [r[0] for r in reps], "kernel_ms": [r[1] for r in reps],
                    "median_mlups": float(np.median([r[0] for r in reps])),
                    "median_kernel_ms": float(np.median([r[1] for r in reps]))},
        "serving": serving,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
