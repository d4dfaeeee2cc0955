"""The reference's own regression cases (testing/run-tests.sh:34-110): every
example under examples/ at its own params.h values and geometry.config,
500 steps (store-ref-data.sh:26 sets nSteps = 500), plain FP64 method on
the GPU against the oracle at 1e-12 relative on rho, u, the populations,
the force coefficients, the marker positions, the FEM displacements and the
coupling state.  TurekHron runs a second time from the restart files the
first run wrote, as run-tests.sh:70-72 does: 500 steps, restart, 500 more,
against 1000 uninterrupted oracle steps."""
import numpy as np
import pytest

from life_amd import config as cfg
from life_amd import lib as L
from life_amd.case import Case
from oracle import oracle as O
from tests._util import rel_err

pytestmark = pytest.mark.gpu

TOL = 1e-12
STEPS = 500


def compare(name, oc, gc):
    fo, ro, uo, _ = oc.grid()
    fg, rg, ug, _ = gc.fields()
    errs = {"rho": rel_err(rg, ro), "u": rel_err(ug, uo), "f": rel_err(fg, fo)}
    if gc.n_markers:
        errs["force"] = rel_err(gc.total_force(), oc.total_force())
        errs["pos"] = rel_err(gc.markers()["pos"], oc.markers()["pos"])
        for b in range(gc.n_bodies):
            if gc.body_info(b)["ndof"]:
                errs["U%d" % b] = rel_err(gc.fem(b)[0], oc.fem(b)[0])
    print("%s: t=%d %s" % (name, gc.t, " ".join("%s %.2g" % kv for kv in errs.items())), flush=True)
    for k, e in errs.items():
        assert e <= TOL, (k, e)
    if gc.n_markers:
        co, cg = oc.coupling(), gc.coupling()
        assert cg["subIt"] == co["subIt"]
        assert cg["relax"] == pytest.approx(co["relax"], rel=1e-12, abs=1e-300)


@pytest.mark.parametrize("name", [n for n in cfg.REFERENCE_EXAMPLES if n != "TurekHron"])
def test_reference_example(name, oracle):
    p, g = cfg.reference_example(name)
    O.set_threads(16)
    oc = O.OracleCase(p, g)
    gc = Case(p, g, precision=L.LIFE_FP64)
    oc.step(STEPS)
    gc.step(STEPS)
    compare(name, oc, gc)
    oc.close()
    gc.close()


def test_reference_example_turek_hron_restart(oracle, tmp_path):
    p, g = cfg.reference_example("TurekHron")
    O.set_threads(16)
    gc = Case(p, g, precision=L.LIFE_FP64)
    gc.step(STEPS)
    results = str(tmp_path / "Results")
    gc.write_restart(results)
    gc.close()
    gc = Case(p, g, precision=L.LIFE_FP64)
    gc.read_restart(results)
    assert gc.t == STEPS
    gc.step(STEPS)
    oc = O.OracleCase(p, g)
    oc.step(2 * STEPS)
    compare("TurekHron (restarted)", oc, gc)
    oc.close()
    gc.close()
