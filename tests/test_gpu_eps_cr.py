"""Banded epsilon systems of FP32-shifted cases by block cyclic reduction
(ibm.hip: k_eps_band_cr, k_eps_band_cr_batch and the fused serving kernel).

- Against the pivoted elimination on the same case (LIFE_EPS_CR_MAX_BW=0
  sends every system there): the reduction solves the same systems to
  rounding, so after tens of coupled steps the trajectories agree to 1e-9
  relative with identical subiteration counts, while not being bitwise equal
  (the two paths really differ).
- The fallback inside the reduction kernels (a band wider than the limit is
  solved there by the pivoted elimination): with the limit lowered to 2, the
  filament at 45 degrees (markers one spacing apart along the diagonal:
  half-bandwidth 3 or 4) takes the fallback and the vertical ones
  (half-bandwidth 2) the reduction.  The served case equals the single case
  bit for bit under each limit, so the fused serving kernel and the
  single-case kernels make the same choice and the same arithmetic.
"""
import numpy as np
import pytest

from life_amd import config as cfg
from life_amd import lib as L
from life_amd.case import Case
from life_amd.serve import Server
from tests._util import rel_err

pytestmark = pytest.mark.gpu

BAR = 1e-9
STEPS = 24


def mixed():
    """Two vertical filaments and one at 45 degrees in a uniform stream."""
    p, g = cfg.canopy(Nx=320, Ny=128, n_filaments=2, L=40.0, n_el=10, min_spacing=60.0, x_range=(60, 200),
                      ramp_steps=0.0, u_in=0.05, kb_range=(5e-2, 1e-1))
    lines = g.strip().splitlines()
    f = lines[0].split()
    # FILAMENT count ID x y ... L h angle FLEXIBLE nEl BC rho E: a third body,
    # clamped at (240, 20) and leaning downstream at 45 degrees
    f[2], f[3], f[4], f[9] = "2", "240", "20.0", "45.0"
    lines.append(" ".join(f))
    return p.replace(ux0_p=0.05), "\n".join(lines) + "\n"


def run(pg, steps, monkeypatch, max_bw):
    monkeypatch.setenv("LIFE_EPS_CR_MAX_BW", str(max_bw))
    p, g = pg
    c = Case(p, g, precision=L.LIFE_FP32_SHIFTED)
    subits = []
    for _ in range(steps):
        c.step(1)
        subits.append(c.coupling()["subIt"])
    mk = c.markers()
    out = dict(u=c.fields()[2], pos=mk["pos"], eps=mk["eps"], force=c.total_force(), subits=subits)
    c.close()
    return out


def test_mixed_geometry_has_three_flexible_bodies(gpu_lib):
    p, g = mixed()
    c = Case(p, g, precision=L.LIFE_FP32_SHIFTED)
    assert c.n_bodies == 3 and all(c.body_info(k)["ndof"] for k in range(3))
    pos = c.markers()["pos"]
    first, n = c.body_info(2)["first"], c.body_info(2)["n"]
    d = np.abs(np.diff(pos[first:first + n], axis=0))
    # the third body's markers step along the diagonal
    assert np.allclose(d[:, 0], d[:, 1], rtol=1e-6)
    c.close()


def test_reduction_matches_pivoted_elimination(monkeypatch, gpu_lib):
    pg = mixed()
    cr = run(pg, STEPS, monkeypatch, 4)
    pv = run(pg, STEPS, monkeypatch, 0)
    errs = {k: rel_err(cr[k], pv[k]) for k in ("u", "pos", "eps", "force")}
    print("reduction vs pivoted after %d steps: %s" % (STEPS, " ".join("%s %.2g" % kv for kv in errs.items())))
    assert cr["subits"] == pv["subits"]
    for k, e in errs.items():
        assert e <= BAR, (k, e)
    assert not np.array_equal(cr["eps"], pv["eps"])


@pytest.mark.parametrize("max_bw", [2, 4])
def test_served_equals_case_with_fallback(monkeypatch, gpu_lib, max_bw):
    pg = mixed()
    ref = run(pg, STEPS, monkeypatch, max_bw)
    p, g = pg
    srv = Server(precision=L.LIFE_FP32_SHIFTED, max_cases=4)
    rid = srv.submit(p, g, STEPS, keep_fields=True)
    srv.drain()
    info = srv.info(rid)
    assert info["status"] == L.LIFE_REQ_DONE, srv.error(rid)
    _, u, pos = srv.fields(rid)
    assert np.array_equal(u, ref["u"]) and np.array_equal(pos, ref["pos"])
    assert np.array_equal(np.array(info["total_force"]), ref["force"])
    srv.close()


def test_lower_limit_changes_the_path(monkeypatch, gpu_lib):
    """With the limit at 2 the diagonal filament leaves the reduction: the
    trajectory differs from the default's in the last bits only."""
    pg = mixed()
    a = run(pg, STEPS, monkeypatch, 4)
    b = run(pg, STEPS, monkeypatch, 2)
    assert not np.array_equal(a["eps"], b["eps"])
    assert a["subits"] == b["subits"]
    for k in ("u", "pos", "eps"):
        assert rel_err(a[k], b[k]) <= BAR, k
