"""Plain-method (FP64) parity with the oracle on the full lattices of the
reference-style validation configurations, for a bounded number of steps:
lid-driven cavity 256x256 (BGK), rigid cylinder in a 1024x256 channel
(central moments, parabolic inlet, convective outlet), and the cylinder +
flexible plate FSI case on 2048x410 (uniform epsilon system, FP64 FEM,
Aitken coupling), with the inlet at full speed from the first step so that
the coupling subiterates; and the reference's PELskin example at its own
size (735x91, physical units, ten flexible filaments in the oscillating
Womersley channel).  Fields, forces and structure at 1e-12 relative."""
import pytest

from life_amd import config as cfg
from life_amd import lib as L
from life_amd.case import Case
from oracle import oracle as O
from tests._util import rel_err

pytestmark = pytest.mark.gpu

TOL = 1e-12

CASES = {
    "ldc256": (lambda: cfg.lid_driven_cavity(), 200),
    "cylinder1024": (lambda: cfg.cylinder_channel(ramp_steps=0.0), 60),
    "fsi2048": (lambda: cfg.flexible_fsi(ramp_steps=0.0), 15),
    "pelskin735": (lambda: cfg.pelskin(), 80),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_full_lattice_parity(name, oracle):
    build, steps = CASES[name]
    p, g = build()
    O.set_threads(16)
    oc = O.OracleCase(p, g)
    oc.step(steps)
    _, orho, ou, _ = oc.grid()
    gc = Case(p, g, precision=L.LIFE_FP64)
    gc.step(steps)
    _, grho, gu, _ = gc.fields()
    print("%s: %d steps, u rel %.3g, rho rel %.3g" % (name, steps, rel_err(gu, ou), rel_err(grho, orho)), flush=True)
    assert rel_err(grho, orho) <= TOL
    assert rel_err(gu, ou) <= TOL
    if gc.n_markers:
        assert rel_err(gc.total_force(), oc.total_force()) <= TOL
        mo, mg = oc.markers(), gc.markers()
        assert rel_err(mg["pos"], mo["pos"]) <= TOL
        for b in range(gc.n_bodies):
            if gc.body_info(b)["ndof"]:
                Uo = oc.fem(b)[0]
                Ug = gc.fem(b)[0]
                assert rel_err(Ug, Uo) <= TOL
        assert gc.coupling()["subIt"] == oc.coupling()["subIt"]
    oc.close()
    gc.close()
