"""The FEM on the device (fem_dev.hip), the default with FP32-shifted
populations: the same case stepped with the device FEM and with the host FEM
(LIFE_DEVICE_FEM=0) gives the same trajectory.  The device FEM keeps the
host's operations and order except for the linear solve of each Newton
iteration (block cyclic reduction over the beam's 3 x 3 node blocks in
place of the banded LU: the same solution to rounding) and the device's
atan2/sin/cos, which may round differently in the last bit, so the bar is
1e-9 relative after tens of coupled steps (the FP32 populations themselves
sit ~1e-6 from FP64), with identical subiteration counts at every step.
The cases cover both root conditions (CLAMPED: 3 constrained DOFs,
SUPPORTED: 2) and chains of 3 to 51 nodes, so the reduction meets node
counts of 2^k + 1, odd and even.  Also: a restart written with the device
FEM continues bit for bit, and the outputs that read the FEM state (FEM
displacements, total force, restart files) see the device's state."""
import numpy as np
import pytest

from life_amd import config as cfg
from life_amd import lib as L
from life_amd.case import Case
from tests._util import rel_err

pytestmark = pytest.mark.gpu

BAR = 1e-9


def fsi():
    return cfg.flexible_fsi(Nx=200, Ny=62, D=12.0, centre=(30.0, 30.0), L=42.0, h=2.4, n_el=6, u_mean=0.05,
                            Re=60.0, ramp_steps=0.0, uni_epsilon=0, central_moments=1, stiffness=200.0)


def canopy():
    p, g = cfg.canopy(Nx=320, Ny=96, n_filaments=3, L=40.0, n_el=10, min_spacing=30.0, x_range=(60, 220),
                      ramp_steps=0.0, u_in=0.05, kb_range=(1e-2, 1e-1))
    return p.replace(ux0_p=0.05), g


def flag():
    return cfg.inverted_flag(Nx=192, Ny=96, n_el=12, Re=150.0, kb=0.3, initial_deflect=0.01)


def long_filament():
    p, g = cfg.canopy(Nx=200, Ny=340, n_filaments=1, L=300.0, n_el=30, min_spacing=30.0, x_range=(60, 120),
                      ramp_steps=0.0, u_in=0.05, kb_range=(5e-2, 5e-2))
    return p.replace(ux0_p=0.05), g


def supported():
    """Filaments on pinned roots (SUPPORTED) of 16 and 2 elements."""
    p, g = cfg.canopy(Nx=320, Ny=96, n_filaments=2, L=40.0, n_el=16, min_spacing=60.0, x_range=(60, 220),
                      ramp_steps=0.0, u_in=0.05, kb_range=(5e-2, 1e-1))
    lines = g.replace("CLAMPED", "SUPPORTED").splitlines()
    lines[1] = lines[1].replace("FLEXIBLE 16", "FLEXIBLE 2")
    return p.replace(ux0_p=0.05), "\n".join(lines) + "\n"


CASES = {"fsi": (fsi, 30), "canopy": (canopy, 30), "flag": (flag, 30), "long_filament": (long_filament, 12),
         "supported": (supported, 30)}


def run(pg, steps, monkeypatch, device):
    monkeypatch.setenv("LIFE_DEVICE_FEM", "1" if device else "0")
    p, g = pg
    c = Case(p, g, precision=L.LIFE_FP32_SHIFTED)
    subits = []
    for _ in range(steps):
        c.step(1)
        subits.append(c.coupling()["subIt"])
    flex = [k for k in range(c.n_bodies) if c.body_info(k)["ndof"]]
    out = dict(u=c.fields()[2], pos=c.markers()["pos"], force=c.total_force(),
               U=np.concatenate([c.fem(k)[0] for k in flex]), subits=subits)
    c.close()
    return out


@pytest.mark.parametrize("name", sorted(CASES))
def test_device_fem_matches_host_fem(monkeypatch, gpu_lib, name):
    build, steps = CASES[name]
    pg = build()
    h = run(pg, steps, monkeypatch, False)
    d = run(pg, steps, monkeypatch, True)
    errs = {k: rel_err(d[k], h[k]) for k in ("u", "pos", "force", "U")}
    print("%s: %d steps, subiterations %s, device vs host FEM: %s" % (
        name, steps, d["subits"][-5:], " ".join("%s %.2g" % kv for kv in errs.items())), flush=True)
    assert d["subits"] == h["subits"]
    for k, e in errs.items():
        assert e <= BAR, (k, e)


def test_device_fem_is_default(monkeypatch, gpu_lib):
    """Without LIFE_DEVICE_FEM set, FP32-shifted cases use the device FEM:
    the flag case, whose device and host FEM trajectories differ in the last
    bits after 30 steps, follows the device FEM's bit for bit."""
    pg = flag()
    monkeypatch.delenv("LIFE_DEVICE_FEM", raising=False)
    p, g = pg
    c = Case(p, g, precision=L.LIFE_FP32_SHIFTED)
    c.step(30)
    u_default = c.fields()[2]
    c.close()
    d = run(pg, 30, monkeypatch, True)
    h = run(pg, 30, monkeypatch, False)
    assert not np.array_equal(d["u"], h["u"])  # the two are distinguishable here
    assert np.array_equal(u_default, d["u"])


def test_device_fem_restart(monkeypatch, gpu_lib, tmp_path):
    """The device FEM state goes into the restart files and comes back: a
    case read from them holds the writer's FEM state and markers exactly,
    and stepping on from there with the device FEM matches the host FEM
    stepping on from the same files."""
    p, g = fsi()
    monkeypatch.setenv("LIFE_DEVICE_FEM", "1")
    a = Case(p, g, precision=L.LIFE_FP32_SHIFTED)
    a.step(9)
    a.write_restart(str(tmp_path))
    Ua, pa = [x.copy() for x in a.fem(1)], a.markers()["pos"]
    a.close()
    out = {}
    for dev in (True, False):
        monkeypatch.setenv("LIFE_DEVICE_FEM", "1" if dev else "0")
        c = Case(p, g, precision=L.LIFE_FP32_SHIFTED)
        c.read_restart(str(tmp_path))
        for x, y in zip(c.fem(1), Ua):
            assert np.array_equal(x, y)
        assert np.array_equal(c.markers()["pos"], pa)
        c.step(7)
        out[dev] = (c.fields()[2], c.markers()["pos"], c.fem(1)[0], c.coupling()["subIt"])
        c.close()
    (ud, pd, Ud, sd), (uh, ph, Uh, sh) = out[True], out[False]
    assert sd == sh
    assert rel_err(ud, uh) <= BAR and rel_err(pd, ph) <= BAR and rel_err(Ud, Uh) <= BAR


def test_served_device_fem_equals_case(monkeypatch, gpu_lib):
    """The serving engine's batched device FEM (every flexible body of every
    case in one launch, per-case Aitken factors) gives each case exactly its
    single-case device-FEM result, beside cases of other sizes and kinds."""
    from life_amd.serve import Server
    monkeypatch.setenv("LIFE_DEVICE_FEM", "1")
    cases = [fsi(), canopy(), flag(),
             cfg.cylinder_channel(Nx=160, Ny=64, D=16.0, centre=(40.0, 31.5), u_max=0.06, Re=40.0, ramp_steps=0.0)]
    steps = 14
    ref = []
    for p, g in cases:
        c = Case(p, g, precision=L.LIFE_FP32_SHIFTED)
        c.step(steps)
        ref.append((c.fields()[1], c.fields()[2], c.markers()["pos"], c.total_force()))
        c.close()
    srv = Server(precision=L.LIFE_FP32_SHIFTED, max_cases=8)
    ids = [srv.submit(p, g, steps, keep_fields=True) for p, g in cases]
    srv.drain()
    for rid, (rho, u, pos, force) in zip(ids, ref):
        info = srv.info(rid)
        assert info["status"] == L.LIFE_REQ_DONE, srv.error(rid)
        r2, u2, p2 = srv.fields(rid)
        assert np.array_equal(r2, rho) and np.array_equal(u2, u) and np.array_equal(p2, pos)
        assert np.array_equal(np.array(info["total_force"]), force)
    srv.close()


@pytest.mark.parametrize("name", ["canopy", "flag", "fsi"])
def test_spread_behind_fem_is_neutral(monkeypatch, gpu_lib, name):
    """The final spread queued behind the FEM of a subiteration predicted to
    be the step's last (the default) and the spread queued after the host has
    read the residual (LIFE_SPEC_SPREAD=0) give the same trajectory bit for
    bit.  Where a step takes more subiterations than the one before, the
    queued spread is redone by the next subiteration's."""
    build, steps = CASES[name]
    pg = build()
    monkeypatch.setenv("LIFE_SPEC_SPREAD", "0")
    a = run(pg, steps, monkeypatch, True)
    monkeypatch.setenv("LIFE_SPEC_SPREAD", "1")
    b = run(pg, steps, monkeypatch, True)
    grew = sum(1 for k in range(1, steps) if b["subits"][k] > b["subits"][k - 1])
    print("%s: subiterations %s (%d steps with more than the step before)" % (name, b["subits"], grew), flush=True)
    assert a["subits"] == b["subits"]
    for k in ("u", "pos", "force", "U"):
        assert np.array_equal(a[k], b[k]), k
