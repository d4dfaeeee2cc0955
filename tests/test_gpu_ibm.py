"""GPU parity of the immersed-boundary coupling (ObjectsClass::objectKernel,
Objects.cpp:26-321) for rigid and flexible bodies, plain method (FP64):
supports, spacing, epsilon, marker forces, FEM state and lattice fields
against the oracle at 1e-12 relative."""
import numpy as np
import pytest

from life_amd import config as cfg
from tests._util import rel_err

pytestmark = pytest.mark.gpu

TOL = 1e-12


def small_cylinder():
    return cfg.cylinder_channel(Nx=160, Ny=64, D=16.0, centre=(40.0, 31.5), u_max=0.06, Re=40.0, ramp_steps=0.0)


def small_fsi(uni=1, cm=1):
    return cfg.flexible_fsi(Nx=200, Ny=62, D=12.0, centre=(30.0, 30.0), L=42.0, h=2.4, n_el=6, u_mean=0.05,
                            Re=60.0, ramp_steps=0.0, uni_epsilon=uni, central_moments=cm, stiffness=200.0)


def small_canopy(cm=0):
    p, g = cfg.canopy(Nx=320, Ny=96, n_filaments=3, L=40.0, n_el=10, min_spacing=30.0, x_range=(60, 220),
                      ramp_steps=0.0, u_in=0.05, kb_range=(1e-2, 1e-1))
    return p.replace(central_moments=cm, ux0_p=0.05), g


def compare_markers(oc, gc):
    mo = oc.markers()
    mg = gc.markers()
    cnt_o, ij_o, d_o = oc.supports()
    assert np.array_equal(mg["supp_count"], cnt_o)
    for k in range(len(cnt_o)):
        n = cnt_o[k]
        assert np.array_equal(mg["supp_ij"][k, :n], ij_o[k, :n])
        assert np.array_equal(mg["supp_delta"][k, :n], d_o[k, :n])
    assert rel_err(mg["ds"], mo["ds"]) <= TOL
    assert rel_err(mg["eps"], mo["eps"]) <= TOL
    assert rel_err(mg["pos"], mo["pos"]) <= TOL
    assert rel_err(mg["force"], mo["force"]) <= TOL


def compare_fields(oc, gc):
    fo, ro, uo, bo = oc.grid()
    fg, rg, ug, bg = gc.fields()
    assert rel_err(bg, bo) <= TOL
    assert rel_err(rg, ro) <= TOL
    assert rel_err(ug, uo) <= TOL
    assert rel_err(fg, fo) <= TOL


def build(oracle, p, g, precision=0):
    from life_amd.case import Case
    oc = oracle.OracleCase(p, g)
    gc = Case(p, g, precision=precision)
    return oc, gc


def test_initial_objects(oracle, gpu_lib):
    """initialiseObjects: supports, ds and epsilon at t = 0 (Objects.cpp:941-961)."""
    p, g = small_cylinder()
    oc, gc = build(oracle, p, g)
    assert gc.n_markers == oc.n_markers
    compare_markers(oc, gc)


def test_rigid_cylinder(oracle, gpu_lib):
    p, g = small_cylinder()
    oc, gc = build(oracle, p, g)
    print("rigid cylinder markers", oc.n_markers, flush=True)
    for _ in range(3):
        oc.step(15)
        gc.step(15)
        compare_markers(oc, gc)
        compare_fields(oc, gc)
        assert rel_err(gc.total_force(), oc.total_force()) <= TOL


@pytest.mark.parametrize("uni", [1, 0])
def test_flexible_fsi(oracle, gpu_lib, uni):
    p, g = small_fsi(uni=uni)
    oc, gc = build(oracle, p, g)
    print("fsi markers", oc.n_markers, "uni", uni, flush=True)
    for _ in range(3):
        oc.step(8)
        gc.step(8)
        co, cg = oc.coupling(), gc.coupling()
        assert co["subIt"] == cg["subIt"]
        assert abs(co["relax"] - cg["relax"]) <= TOL * max(1.0, abs(co["relax"]))
        U_o = oc.fem(1)[0]
        U_g = gc.fem(1)[0]
        assert rel_err(U_g, U_o) <= TOL
        compare_markers(oc, gc)
        compare_fields(oc, gc)
        assert rel_err(gc.total_force(), oc.total_force()) <= TOL


def test_canopy_small(oracle, gpu_lib):
    p, g = small_canopy()
    oc, gc = build(oracle, p, g)
    print("canopy markers", oc.n_markers, flush=True)
    oc.step(10)
    gc.step(10)
    for b in range(oc.n_bodies):
        assert rel_err(gc.fem(b)[0], oc.fem(b)[0]) <= TOL
    compare_markers(oc, gc)
    compare_fields(oc, gc)


def _state(gc):
    f, rho, u, b = gc.fields()
    mk = gc.markers()
    flex = [k for k in range(gc.n_bodies) if gc.body_info(k)["ndof"]]
    return dict(f=f, rho=rho, u=u, force_ibm=b, pos=mk["pos"], eps=mk["eps"], force=mk["force"],
                fem=np.concatenate([gc.fem(k)[0] for k in flex]) if flex else np.zeros(0), total=gc.total_force(),
                subit=gc.coupling()["subIt"], flex=np.array(flex))


@pytest.mark.parametrize("mode", ["1", "2"])
@pytest.mark.parametrize("which", ["canopy", "fsi", "womersley"])
def test_overlapped_step_matches_plain_order(oracle, gpu_lib, monkeypatch, mode, which):
    """The overlapped schedule (next step's stream-collide on a second
    stream beside the coupling, step tail behind it, spread beside the FEM,
    deferred out-of-band check) gives bit-identical results to the plain
    order, and both equal the oracle, on small cases where it is not the
    default: filaments (banded epsilon), cylinder + plate with
    subiterations (uniform epsilon), and a cylinder in the oscillating
    Womersley channel (the early stream-collide takes the next step's
    force)."""
    from life_amd.case import Case
    if which == "womersley":
        p, g = cfg.womersley_cylinder()
        p = p.replace(womersley=8.0)  # a short period: the force changes visibly from step to step
        flex_ok = False
    else:
        p, g = small_canopy() if which == "canopy" else small_fsi(uni=1)
        flex_ok = True
    steps = 10 if which == "canopy" else 12
    runs = {}
    for m in ("0", mode):
        monkeypatch.setenv("LIFE_OVERLAP", m)
        gc = Case(p, g, precision=0)
        gc.step(steps // 2)
        gc.step(steps - steps // 2)
        runs[m] = _state(gc)
        gc.close()
    a, b = runs["0"], runs[mode]
    for k in a:
        assert np.array_equal(np.asarray(a[k]), np.asarray(b[k])), k
    oc = oracle.OracleCase(p, g)
    oc.step(steps)
    assert rel_err(b["u"], oc.grid()[2]) <= TOL
    assert rel_err(b["force"], oc.markers()["force"]) <= TOL
    if flex_ok:
        assert rel_err(b["fem"], np.concatenate([oc.fem(k)[0] for k in b["flex"]])) <= TOL
    oc.close()


def test_long_filament(oracle, gpu_lib):
    """A 301-marker filament: more markers than the fused per-body kernel
    and the LDS-staged solve take, so supports, spacing and the banded
    epsilon solve run as separate kernels with the band in global memory."""
    p, g = cfg.canopy(Nx=200, Ny=340, n_filaments=1, L=300.0, n_el=30, min_spacing=30.0, x_range=(60, 120),
                      ramp_steps=0.0, u_in=0.05, kb_range=(5e-2, 5e-2))
    p = p.replace(ux0_p=0.05)
    oc, gc = build(oracle, p, g)
    assert oc.n_markers == 301
    oc.step(6)
    gc.step(6)
    assert rel_err(gc.fem(0)[0], oc.fem(0)[0]) <= TOL
    compare_markers(oc, gc)
    compare_fields(oc, gc)


def test_inverted_flag_initial_deflection(oracle, gpu_lib):
    """INITIAL_DEFLECT (the reference's InvertedFlag example): static FEM at
    creation, then the coupled steps, against the oracle."""
    p, g = cfg.inverted_flag(Nx=192, Ny=96, n_el=12, Re=150.0, kb=0.3, initial_deflect=0.01)
    oc, gc = build(oracle, p, g)
    assert rel_err(gc.fem(0)[0], oc.fem(0)[0]) <= TOL
    compare_markers(oc, gc)
    oc.step(12)
    gc.step(12)
    assert rel_err(gc.fem(0)[0], oc.fem(0)[0]) <= TOL
    compare_markers(oc, gc)
    compare_fields(oc, gc)
    assert rel_err(gc.total_force(), oc.total_force()) <= TOL


def test_layer1_epsilon_matrix(oracle, gpu_lib):
    """life_ibm_epsilon_matrix + life_solve_lapack reproduce epsilon."""
    from life_amd import lib
    p, g = small_fsi(uni=1)
    oc, gc = build(oracle, p, g)
    n = gc.n_markers
    gc.grid.n_markers = n
    A = gc.grid.epsilon_matrix(0, n)
    eps = lib.solve_lapack(A, np.ones(n))
    assert rel_err(eps, oc.markers()["eps"]) <= TOL
    # epsilon * A = 1 (Objects.cpp:303-306)
    assert np.max(np.abs(A @ eps - 1.0)) < 1e-10


@pytest.mark.parametrize("cm", [1, 0])
def test_gravity_flexible_matches_oracle(oracle, gpu_lib, cm):
    """Gravity: the uniform body force on the fluid (Grid.cpp:1042-1043) and
    the weight of the beam elements in the FEM load (FEMElement.cpp:35),
    with a flexible plate, FP64 on the GPU against the oracle; the weight
    visibly moves the plate against the same case without gravity."""
    p0, g = small_fsi(uni=1, cm=cm)
    p = p0.replace(gravityX=1e-6, gravityY=-6e-6)
    oc, gc = build(oracle, p, g)
    steps = 20
    oc.step(steps)
    gc.step(steps)
    co, cg = oc.coupling(), gc.coupling()
    assert co["subIt"] == cg["subIt"]
    U_o, U_g = oc.fem(1)[0], gc.fem(1)[0]
    assert rel_err(U_g, U_o) <= TOL
    compare_markers(oc, gc)
    compare_fields(oc, gc)
    assert rel_err(gc.total_force(), oc.total_force()) <= TOL
    o0 = oracle.OracleCase(p0, g)
    o0.step(steps)
    dU = float(np.max(np.abs(U_o - o0.fem(1)[0])) / np.max(np.abs(U_o)))
    print("gravity cm %d: subIt %d, FEM displacement change from gravity %.3g" % (cm, cg["subIt"], dU), flush=True)
    assert dU > 1e-3
    o0.close()
    oc.close()
    gc.close()


def test_womersley_gravity_flexible_matches_oracle(oracle, gpu_lib):
    """The reference's PELskin example (WOMERSLEY, ten clamped flexible
    filaments, physical units, 735 x 91) with gravity added: force_xy per
    site from each site's density and the weight of the beam elements in
    the FEM load, FP64 against the oracle; the weight roughly doubles the
    filaments' displacement over these steps."""
    p0, g = cfg.pelskin()
    p = p0.replace(gravityX=4.0, gravityY=-9.81)
    oracle.set_threads(16)
    oc, gc = build(oracle, p, g)
    steps = 40
    oc.step(steps)
    gc.step(steps)
    co, cg = oc.coupling(), gc.coupling()
    assert co["subIt"] == cg["subIt"]
    for b in range(oc.n_bodies):
        assert rel_err(gc.fem(b)[0], oc.fem(b)[0]) <= TOL
    compare_markers(oc, gc)
    compare_fields(oc, gc)
    assert rel_err(gc.total_force(), oc.total_force()) <= TOL
    o0 = oracle.OracleCase(p0, g)
    o0.step(steps)
    change = float(np.max(np.abs(oc.fem(0)[0] - o0.fem(0)[0])) / np.max(np.abs(o0.fem(0)[0])))
    print("PELskin with gravity: subIt %d, FEM displacement change from gravity %.3g" % (cg["subIt"], change),
          flush=True)
    assert change > 0.1
    o0.close()
    oc.close()
    gc.close()
