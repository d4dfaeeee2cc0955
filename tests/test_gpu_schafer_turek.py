"""Physics anchor: the Schaefer-Turek cylinder benchmarks computed by the
plain FP64 path on the reference's own Cylinder example
(examples/Cylinder/params.h: BGK, tau 0.53, pressure outlet) at resolution
factors r = 2 and 4 (D = 20 and 40 lattice units), against the published
bounds (Schaefer & Turek 1996).

- 2D-2 (Re 100): the Strouhal number lies in the published band at both
  resolutions; the maximum lift lies within 3 % of its band; the maximum
  drag lies above its band and falls towards it (3.41 at r = 2, 3.30 at
  r = 4 and 3.26 at r = 8: the diffuse interface's error).
- 2D-1 (Re 20): the flow is steady after 30 s; drag and lift fall with
  resolution towards the band (drag 5.87 at r = 2 and 5.71 at r = 4, within
  2.5 % of the band's upper bound).
A misplaced cylinder, a wrong viscosity or time step, a lost or doubled
force term or a mis-scaled coefficient moves these well outside the bars."""
import pytest

from life_amd.config import SCHAFER_TUREK_BOUNDS as B
from life_amd.validation import schafer_turek_run

pytestmark = pytest.mark.gpu


def within(x, band, rel):
    return band[0] * (1.0 - rel) <= x <= band[1] * (1.0 + rel)


def test_2d2_periodic_shedding(gpu_lib):
    r2 = schafer_turek_run("2D-2", 2)
    r4 = schafer_turek_run("2D-2", 4)
    for r in (r2, r4):
        print("2D-2 r=%d: CDmax %.4f CLmax %.4f St %.4f (%d steps, %.1f s)" % (
            r["res_factor"], r["cd_max"], r["cl_max"], r["strouhal"], r["steps"], r["wall_s"]), flush=True)
        assert within(r["strouhal"], B["2D-2"]["strouhal"], 0.0), r
        assert within(r["cl_max"], B["2D-2"]["cl_max"], 0.03), r
        assert B["2D-2"]["cd_max"][0] <= r["cd_max"] <= B["2D-2"]["cd_max"][1] * 1.06, r
    assert r4["cd_max"] < r2["cd_max"]
    assert within(r4["cd_max"], B["2D-2"]["cd_max"], 0.025)


def test_2d1_steady(gpu_lib):
    r2 = schafer_turek_run("2D-1", 2, seconds=30.0)
    r4 = schafer_turek_run("2D-1", 4, seconds=30.0)
    for r in (r2, r4):
        print("2D-1 r=%d: CD %.5f CL %.5f (change over the last second %.1e; %d steps, %.1f s)" % (
            r["res_factor"], r["cd"], r["cl"], r["cd_change_last_second"], r["steps"], r["wall_s"]), flush=True)
        assert r["cd_change_last_second"] <= 1e-3, r
        assert B["2D-1"]["cd"][0] <= r["cd"] <= B["2D-1"]["cd"][1] * 1.06, r
        assert within(r["cl"], B["2D-1"]["cl"], 0.5), r
    assert r4["cd"] < r2["cd"] and r4["cl"] < r2["cl"]
    assert within(r4["cd"], B["2D-1"]["cd"], 0.025)
