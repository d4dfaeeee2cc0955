"""The device FEM (fem_dev.hip) against the oracle directly.  With
LIFE_DEVICE_FEM=1 the plain method (FP64 populations; lattice, supports,
spacing and epsilon in the oracle's operation order) runs its FEM on the
device, so the only differences from the oracle are the device FEM's own:
its linear solve (block cyclic reduction in place of the banded
elimination) and the device library's atan2/sin/cos, which round differently
in the last bits.  After tens of coupled steps the FEM displacements, marker
positions and forces, the velocity field and the total force agree with the
oracle to 1e-10 relative (1e-12 is the bar of the bit-compatible host FEM),
with the same subiteration counts; and the run really used the device FEM
(its FEM state is not the host FEM's bit for bit)."""
import numpy as np
import pytest

from life_amd import config as cfg
from life_amd.case import Case
from tests._util import rel_err

pytestmark = pytest.mark.gpu

BAR = 1e-10


def fsi():
    return cfg.flexible_fsi(Nx=200, Ny=62, D=12.0, centre=(30.0, 30.0), L=42.0, h=2.4, n_el=6, u_mean=0.05,
                            Re=60.0, ramp_steps=0.0, uni_epsilon=0, central_moments=1, stiffness=200.0)


def canopy():
    p, g = cfg.canopy(Nx=320, Ny=96, n_filaments=3, L=40.0, n_el=10, min_spacing=30.0, x_range=(60, 220),
                      ramp_steps=0.0, u_in=0.05, kb_range=(1e-2, 1e-1))
    return p.replace(ux0_p=0.05), g


def flag():
    return cfg.inverted_flag(Nx=192, Ny=96, n_el=12, Re=150.0, kb=0.3, initial_deflect=0.01)


CASES = {"fsi": (fsi, 3, 8), "canopy": (canopy, 3, 8), "flag": (flag, 3, 8)}


def flex_fem(c):
    return np.concatenate([c.fem(k)[0] for k in range(c.n_bodies) if c.body_info(k)["ndof"]])


@pytest.mark.parametrize("name", sorted(CASES))
def test_device_fem_matches_oracle(oracle, gpu_lib, monkeypatch, name):
    build, rounds, steps = CASES[name]
    p, g = build()
    oc = oracle.OracleCase(p, g)
    monkeypatch.setenv("LIFE_DEVICE_FEM", "1")
    dc = Case(p, g, precision=0)
    monkeypatch.setenv("LIFE_DEVICE_FEM", "0")
    hc = Case(p, g, precision=0)
    worst = {}
    for _ in range(rounds):
        oc.step(steps)
        dc.step(steps)
        hc.step(steps)
        assert dc.coupling()["subIt"] == oc.coupling()["subIt"]
        mo, md = oc.markers(), dc.markers()
        fo, fd = oc.grid(), dc.fields()
        errs = {"U": rel_err(flex_fem(dc), flex_fem(oc)), "pos": rel_err(md["pos"], mo["pos"]),
                "force": rel_err(md["force"], mo["force"]), "eps": rel_err(md["eps"], mo["eps"]),
                "u": rel_err(fd[2], fo[2]), "total": rel_err(dc.total_force(), oc.total_force())}
        for k, e in errs.items():
            worst[k] = max(worst.get(k, 0.0), e)
            assert e <= BAR, (k, e)
    print("%s: device FEM vs oracle after %d steps: %s" % (
        name, rounds * steps, " ".join("%s %.2g" % kv for kv in worst.items())), flush=True)
    # the host FEM run is the oracle's to 1e-12; the device FEM's is not bit for bit the host's
    assert rel_err(flex_fem(hc), flex_fem(oc)) <= 1e-12
    assert not np.array_equal(flex_fem(dc), flex_fem(hc))
    dc.close()
    hc.close()
