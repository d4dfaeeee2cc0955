"""Throughput benchmark of the LIFE fluid + coupling hot path on one MI355X.

Workload (default): the canopy case -- D2Q9 BGK on an 8192 x 4096 lattice
with FP32-shifted populations, uniform inlet, convective outlet, no-slip
bottom, free-slip top, and 64 clamped flexible filaments (length 200, 201
markers and 50 corotational beam elements each) coupled through the
implicit immersed boundary with FP64 FEM on the host.  A step is one full
time step (lattice step + coupling subiterations + spread).

Prints one JSON line: MLUPS of the whole step (lattice sites x steps /
wall time), the roofline of the fused stream-collide kernel (72 algorithmic
bytes per site update over its HIP-event-timed average launch), the FP32 vs
FP64 velocity error on the same run, and a CPU baseline (the oracle, FP64,
OpenMP, on a bounded number of steps of the same workload).
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
                          fem_elements=50, markers_per_filament=201, tau=0.6, u_in=0.05, seed=20261202)
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
    ap.add_argument("--no-accuracy", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=6)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: this build runs each case on a single MI355X "
              "(multi-GPU sharding is a design note, not an implementation).", file=sys.stderr)
        sys.exit(2)

    import numpy as np

    from life_amd import lib
    from life_amd.case import Case

    p, geom, wcfg = workload(args.workload)
    prec = lib.LIFE_FP32_SHIFTED if args.precision == "f32" else lib.LIFE_FP64
    print("bench: building %s (%s)" % (wcfg["workload"], args.precision), file=sys.stderr, flush=True)
    case = Case(p, geom, precision=prec)
    case.step(args.warmup)
    print("bench: warmup done (%d steps)" % args.warmup, file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    ms_dev, ms_kernel = case.step_timed(args.steps)
    wall = time.perf_counter() - t0
    N = p.Nx * p.Ny
    mlups = N * args.steps / wall / 1e6
    k_avg_s = ms_kernel / 1e3 / args.steps
    algo_bytes = BYTES_PER_LU[args.precision] * N
    achieved = algo_bytes / k_avg_s / 1e9
    kernel = "k_push32" if args.precision == "f32" else "k_push64"
    traffic = read_traffic(kernel, wcfg["workload"])
    print("bench: %d steps in %.3f s wall (%.3f ms device), K1 avg %.3f ms" %
          (args.steps, wall, ms_dev, k_avg_s * 1e3), file=sys.stderr, flush=True)

    accuracy = None
    if not args.no_accuracy and args.precision == "f32":
        _, _, u32, _ = case.fields()
        case.close()
        ref = Case(p, geom, precision=lib.LIFE_FP64)
        ref.step(args.warmup + args.steps)
        _, _, u64, _ = ref.fields()
        ref.close()
        accuracy = {"u_l2_rel_vs_fp64": float(np.linalg.norm(u32 - u64) / np.linalg.norm(u64)),
                    "steps": args.warmup + args.steps}
        del u32, u64
    else:
        case.close()

    cpu = None
    if not args.no_cpu:
        threads = min(16, os.cpu_count() or 1)
        print("bench: CPU baseline (oracle, FP64, %d threads, %d steps)" % (threads, args.cpu_steps),
              file=sys.stderr, flush=True)
        v = cpu_baseline(p, geom, args.cpu_steps, threads)
        cpu = {"value": v, "unit": "MLUPS", "cores": threads, "kind": "port",
               "sample": "%d full steps (lattice + coupling) of the same workload in FP64 after 1 step" %
                         args.cpu_steps}

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
        "config": dict(wcfg, precision=("fp32_shifted" if args.precision == "f32" else "fp64")),
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "kernel": kernel,
                     "kernel_ms": k_avg_s * 1e3, "kernel_mlups": N / k_avg_s / 1e6},
        "cpu_baseline": cpu,
        "accuracy": accuracy,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
