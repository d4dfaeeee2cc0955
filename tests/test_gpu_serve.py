"""Serving engine (life_server_*): every case's results are the same as a
life_case run of the same inputs (the single-case path, itself checked
against the oracle), whether it runs alone or beside other cases of other
sizes, collision operators and body counts, admitted at other times, with
or without the hipGraph lattice phase.  The bar is bitwise equality: the
batched kernels do each case's arithmetic exactly as the single-case
kernels do."""
import numpy as np
import pytest

from life_amd import config as cfg
from life_amd import lib as L
from life_amd.case import Case
from life_amd.serve import Server

pytestmark = pytest.mark.gpu


def flag(Nx=160, Ny=64, n_el=8, Re=120.0, kb=0.4, cm=1):
    return cfg.inverted_flag(Nx=Nx, Ny=Ny, length_frac=0.25, n_el=n_el, Re=Re, kb=kb, central_moments=cm)


def cylinder():
    return cfg.cylinder_channel(Nx=160, Ny=64, D=16.0, centre=(40.0, 31.5), u_max=0.06, Re=40.0, ramp_steps=0.0)


def cavity():
    return cfg.lid_driven_cavity(N=96, tau=0.6, u_lid=0.05)


def canopy():
    p, g = cfg.canopy(Nx=320, Ny=96, n_filaments=3, L=40.0, n_el=10, min_spacing=30.0, x_range=(60, 220),
                      ramp_steps=0.0, u_in=0.05, kb_range=(1e-2, 1e-1))
    return p.replace(ux0_p=0.05), g


def run_case(pg, steps, prec):
    p, g = pg
    c = Case(p, g, precision=prec)
    c.step(steps)
    _, rho, u, _ = c.fields()
    mk = c.markers()["pos"] if c.n_markers else np.zeros((0, 2))
    out = dict(rho=rho, u=u, pos=mk.reshape(-1, 2), force=c.total_force())
    c.close()
    return out


def served(srv, rid):
    info = srv.info(rid)
    assert info["status"] == L.LIFE_REQ_DONE, srv.error(rid)
    rho, u, pos = srv.fields(rid)
    return dict(rho=rho, u=u, pos=pos, force=np.array(info["total_force"]))


def assert_same(a, b):
    for k in ("rho", "u", "pos", "force"):
        assert a[k].shape == b[k].shape, k
        assert np.array_equal(a[k], b[k]), k


@pytest.mark.parametrize("prec", [L.LIFE_FP64, L.LIFE_FP32_SHIFTED])
def test_server_matches_case(prec):
    """Each of four different cases run together in one server equals its
    own life_case run (FP64: the plain method, so also the oracle's)."""
    steps = 12
    cases = [flag(), cylinder(), cavity(), canopy()]
    ref = [run_case(pg, steps, prec) for pg in cases]
    srv = Server(precision=prec, max_cases=8)
    ids = [srv.submit(p, g, steps, keep_fields=True) for p, g in cases]
    srv.drain()
    for r, rid in zip(ref, ids):
        assert_same(served(srv, rid), r)
    st = srv.stats()
    assert st["steps"] == steps and st["peak_running"] == 4
    srv.close()


@pytest.mark.parametrize("graph", [True, False])
def test_alone_equals_mix(graph):
    """A case alone and the same case admitted mid-stream beside others
    (different sizes, BGK and central moments, rigid and flexible bodies,
    different step counts): identical results."""
    prec = L.LIFE_FP32_SHIFTED
    x = flag(Nx=192, Ny=96, n_el=10, Re=200.0, kb=0.3, cm=1)
    steps = 16
    alone = Server(precision=prec, max_cases=4, use_graph=graph)
    ra = alone.submit(*x, steps, keep_fields=True)
    alone.drain()
    a = served(alone, ra)
    alone.close()

    mix = Server(precision=prec, max_cases=8, use_graph=graph)
    mix.submit(*flag(Nx=128, Ny=64, n_el=6, Re=80.0, kb=0.8, cm=0), 9)
    mix.submit(*cylinder(), 30)
    for _ in range(3):
        mix.step()
    rx = mix.submit(*x, steps, keep_fields=True)
    mix.submit(*canopy(), 5)
    mix.step()
    mix.submit(*cavity(), 20)
    mix.drain()
    assert_same(served(mix, rx), a)
    mix.close()


@pytest.mark.parametrize("prec", [L.LIFE_FP64, L.LIFE_FP32_SHIFTED])
def test_womersley_served_matches_case(prec):
    """An oscillating-force case (WOMERSLEY) admitted after the engine has
    already stepped: the force follows the case's own time step, not the
    engine's, so it equals its own life_case run."""
    pw, gw = cfg.womersley_cylinder()
    pw = pw.replace(womersley=8.0)
    steps = 15
    ref = run_case((pw, gw), steps, prec)
    srv = Server(precision=prec, max_cases=4)
    srv.submit(*cylinder(), 30)
    for _ in range(4):
        srv.step()
    rid = srv.submit(pw, gw, steps, keep_fields=True)
    srv.drain()
    assert_same(served(srv, rid), ref)
    srv.close()


@pytest.mark.parametrize("prec,graph", [(L.LIFE_FP64, True), (L.LIFE_FP32_SHIFTED, True),
                                        (L.LIFE_FP32_SHIFTED, False)])
def test_early_bulk_matches_plain(monkeypatch, prec, graph):
    """The next step's bulk stream-collide launched beside the coupling
    (while the set of cases stays fixed), as soon as the lattice phase is
    queued (mode 1, the default) or when the FEM starts (mode 2), gives
    the same results, bit for bit, as the plain order, through admissions
    and retirements."""
    pw, gw = cfg.womersley_cylinder()
    work = [(flag(), 10), (cylinder(), 14), (cavity(), 7), (canopy(), 9), ((pw.replace(womersley=8.0), gw), 12)]
    out = {}
    for early in ("0", "1", "2"):
        monkeypatch.setenv("LIFE_SERVE_EARLY", early)
        srv = Server(precision=prec, max_cases=3, use_graph=graph)
        ids = [srv.submit(p, g, n, keep_fields=True) for (p, g), n in work]
        srv.drain()
        out[early] = [served(srv, rid) for rid in ids]
        st = srv.stats()
        print("early", early, "early steps", st["early_steps"], "of", st["steps"], flush=True)
        assert (st["early_steps"] > 0) == (early != "0")
        srv.close()
    for mode in ("1", "2"):
        for a, b in zip(out["0"], out[mode]):
            assert_same(a, b)


def test_admission_by_budget():
    """With room for one case at a time the others wait in arrival order;
    every request still finishes with the single-case result."""
    prec = L.LIFE_FP32_SHIFTED
    cases = [flag(Nx=128, Ny=64, n_el=6), flag(Nx=160, Ny=64, n_el=8), cavity()]
    steps = 6
    ref = [run_case(pg, steps, prec) for pg in cases]
    # room for the largest single footprint (about 120 B/site + 1 MiB), not two
    srv = Server(precision=prec, max_cases=8, hbm_budget=float(160 * 64 * 130 + (1 << 20)))
    ids = [srv.submit(p, g, steps, keep_fields=True) for p, g in cases]
    run, queued = srv.step()
    assert run == 1 and queued == 2
    srv.drain()
    infos = [srv.info(i) for i in ids]
    for a, b in zip(infos, infos[1:]):
        assert b["t_admit"] >= a["t_finish"]
    for r, rid in zip(ref, ids):
        assert_same(served(srv, rid), r)
    srv.close()


def test_timed_stream_runs_to_completion():
    """The persistent request loop admits a small seeded Poisson stream and
    finishes every request; time-to-solution is arrival to finish."""
    stream = cfg.serving_stream(n=6, rate=200.0, steps=8, widths=(128, 192), heights=(64, 96), n_el_range=(6, 10))
    srv = Server(precision=L.LIFE_FP32_SHIFTED, max_cases=4)
    ids = srv.run(stream)
    for rid in ids:
        info = srv.info(rid)
        assert info["status"] == L.LIFE_REQ_DONE, srv.error(rid)
        assert info["steps_done"] == 8
        assert info["t_finish"] > info["t_admit"] >= info["t_arrival"] - 1e-9
    st = srv.stats()
    assert st["site_updates"] == sum(r["Nx"] * r["Ny"] * 8 for r in stream)
    srv.close()


def test_womersley_gravity_request_fails_alone():
    """A WOMERSLEY case with gravity (force_xy per site) is refused by the
    engine with a message; the request beside it finishes as usual."""
    pw, gw = cfg.womersley_cylinder()
    srv = Server(precision=L.LIFE_FP64, max_cases=4)
    bad = srv.submit(pw.replace(womersley=8.0, gravityY=-1e-4), gw, 5)
    good = srv.submit(*cylinder(), 5, keep_fields=True)
    srv.drain()
    assert srv.info(bad)["status"] == L.LIFE_REQ_FAILED
    assert "single-case path" in srv.error(bad)
    assert srv.info(good)["status"] == L.LIFE_REQ_DONE, srv.error(good)
    srv.close()
