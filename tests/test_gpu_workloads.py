"""The benchmark workloads at their own sizes, against the oracle.

* Canopy, 8192 x 4096 with 64 flexible filaments (the bandwidth-roofline
  workload): the plain FP64 method equals the oracle at 1e-12 on the full
  lattice (inflow at full speed from the first step, so the coupling
  subiterates), and the FP32-shifted populations stay within the stated
  velocity bar of FP64 past the 2000-step inlet ramp.
* Serving stream: cases drawn from the serving workload's own size
  distribution (512-2048 x 128-512, 20-60 elements, Re 50-400) served
  together in FP64 equal the oracle's run of each case at 1e-12.
* FP32-shifted populations with flexible bodies against the FP64 oracle:
  a reduced cylinder + plate case and a reduced canopy over a few thousand
  steps, with stated bars on the velocity field, the tip position and the
  force coefficients.
"""
import numpy as np
import pytest

from life_amd import config as cfg
from life_amd import lib as L
from life_amd.case import Case
from life_amd.serve import Server
from oracle import oracle as O
from tests._util import rel_err

pytestmark = pytest.mark.gpu

TOL = 1e-12


def u_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def test_canopy_full_size_fp64_matches_oracle(oracle, gpu_lib):
    p, g = cfg.canopy(ramp_steps=0.0)
    p = p.replace(ux0_p=0.05)
    steps = 3
    O.set_threads(16)
    oc = O.OracleCase(p, g)
    oc.step(steps)
    gc = Case(p, g, precision=L.LIFE_FP64)
    gc.step(steps)
    _, rg, ug, _ = gc.fields()
    _, ro, uo, _ = oc.grid()
    errs = dict(rho=rel_err(rg, ro), u=rel_err(ug, uo), force=rel_err(gc.total_force(), oc.total_force()),
                pos=rel_err(gc.markers()["pos"], oc.markers()["pos"]))
    del rg, ug, ro, uo
    for b in range(gc.n_bodies):
        errs["U%d" % b] = rel_err(gc.fem(b)[0], oc.fem(b)[0])
    print("canopy 8192x4096 FP64, %d steps, subIt %d: max err %.3g" % (steps, gc.coupling()["subIt"],
                                                                        max(errs.values())), flush=True)
    assert gc.coupling()["subIt"] == oc.coupling()["subIt"] >= 2
    for k, e in errs.items():
        assert e <= TOL, (k, e)
    oc.close()
    gc.close()


def test_canopy_full_size_fp32_accuracy(gpu_lib):
    """FP32-shifted against FP64 on the canopy after the inlet ramp (2000
    steps) and 500 steps of loaded flow: velocity L2 <= 1e-5 (the stated
    bar of the shifted-FP32 form); the FP64 run is the oracle's (previous
    test)."""
    p, g = cfg.canopy()
    steps = 2500
    out = {}
    for prec in (L.LIFE_FP64, L.LIFE_FP32_SHIFTED):
        c = Case(p, g, precision=prec)
        c.step(steps)
        out[prec] = (c.fields()[2], c.markers()["pos"], c.total_force())
        c.close()
    (u64, pos64, F64), (u32, pos32, F32) = out[L.LIFE_FP64], out[L.LIFE_FP32_SHIFTED]
    l2 = u_l2(u32, u64)
    tip = float(np.max(np.abs(pos32 - pos64))) / p.ref_L
    print("canopy FP32 vs FP64 after %d steps: u L2 %.3g, max marker displacement error / L %.3g, C %s vs %s"
          % (steps, l2, tip, F32[2:], F64[2:]), flush=True)
    assert l2 <= 1e-5
    assert tip <= 1e-5


def test_serving_stream_sizes_fp64_match_oracle(oracle, gpu_lib):
    """One case of every grid size of the serving stream (3 widths x 3
    heights, the first of each in the stream's own seeded order), served
    together (continuous batching, FP64), each equal to the oracle's run of
    it."""
    stream, seen = [], set()
    for r in cfg.serving_stream(n=128):
        if (r["Nx"], r["Ny"]) not in seen:
            seen.add((r["Nx"], r["Ny"]))
            stream.append(r)
    assert len(stream) == 9
    steps = 40
    srv = Server(precision=L.LIFE_FP64, max_cases=9)
    ids = [srv.submit(r["params"], r["geometry"], steps, keep_fields=True) for r in stream]
    srv.drain()
    O.set_threads(16)
    for rid, r in zip(ids, stream):
        info = srv.info(rid)
        assert info["status"] == L.LIFE_REQ_DONE, srv.error(rid)
        rho, u, pos = srv.fields(rid)
        oc = O.OracleCase(r["params"], r["geometry"])
        oc.step(steps)
        _, ro, uo, _ = oc.grid()
        errs = dict(rho=rel_err(rho, ro), u=rel_err(u, uo), pos=rel_err(pos, oc.markers()["pos"]),
                    force=rel_err(np.array(info["total_force"]), oc.total_force()))
        print("served %dx%d (%d elements, Re %.0f): max err %.3g" % (r["Nx"], r["Ny"], r["n_el"], r["Re"],
                                                                     max(errs.values())), flush=True)
        for k, e in errs.items():
            assert e <= TOL, (k, e)
        oc.close()
    srv.close()


# FP32-shifted with flexible bodies: (builder, steps, bars)
FP32_CASES = {
    "cylinder_plate": (lambda: cfg.flexible_fsi(Nx=512, Ny=103, D=20.0, centre=(50.0, 51.0), L=70.0, h=4.0,
                                                n_el=10, ramp_steps=1000.0), 3000),
    "canopy": (lambda: cfg.canopy(Nx=640, Ny=192, n_filaments=4, L=60.0, n_el=15, min_spacing=60.0,
                                  x_range=(100, 400), ramp_steps=500.0), 3000),
}
# velocity L2, marker position / L, |dC| over |C|
FP32_BARS = (1e-5, 1e-5, 1e-4)


@pytest.mark.parametrize("name", sorted(FP32_CASES))
def test_fp32_flexible_against_oracle(oracle, gpu_lib, name):
    build, steps = FP32_CASES[name]
    p, g = build()
    O.set_threads(16)
    oc = O.OracleCase(p, g)
    oc.step(steps)
    gc = Case(p, g, precision=L.LIFE_FP32_SHIFTED)
    gc.step(steps)
    uo, ug = oc.grid()[2], gc.fields()[2]
    l2 = u_l2(ug, uo)
    pos = float(np.max(np.abs(gc.markers()["pos"] - oc.markers()["pos"]))) / p.ref_L
    Co, Cg = oc.total_force()[2:], gc.total_force()[2:]
    dC = float(np.max(np.abs(Cg - Co)) / np.max(np.abs(Co)))
    print("%s FP32 vs oracle after %d steps: u L2 %.3g, marker pos / L %.3g, |dC|/|C| %.3g (C %s)"
          % (name, steps, l2, pos, dC, Co), flush=True)
    assert l2 <= FP32_BARS[0]
    assert pos <= FP32_BARS[1]
    assert dC <= FP32_BARS[2]
    oc.close()
    gc.close()
