"""Runtime case description: the reference's compile-time params.h as data.

LIFE bakes each case into ``inc/params.h`` (params.h:26-110) and reads bodies
from ``input/geometry.config`` (Objects.cpp:324-488).  Serving mixes grid
sizes, so here both are runtime values: :class:`CaseParams` carries the
params.h constants with the same names and meanings, and the geometry stays
the reference's text format (one ``CIRCLE``/``FILAMENT`` line per body group).

The named workloads below are the benchmark and parity cases.  They are all
set up in lattice units (``Dx = 1``, ``Dt = 1``): ``height_p = Ny - 1``,
``rho_p = 1`` and ``nu_p`` equal to the lattice viscosity that the solver
derives from ``omega`` (Grid.cpp:1255-1258), so positions, lengths and
velocities below are in lattice units.
"""
import dataclasses
import math
import random

# eLatType (inc/defs.h:52)
FLUID, WALL, VELOCITY, FREESLIP, PRESSURE, CONVECTIVE = range(6)
# eProfileType (inc/defs.h:53); UNIFORM = PROFILE undefined
UNIFORM, PARABOLIC, SHEAR, BOUNDARY_LAYER = -1, 0, 1, 2


@dataclasses.dataclass
class CaseParams:
    Nx: int
    Ny: int
    height_p: float
    rho_p: float
    nu_p: float
    ux0_p: float = 0.0
    uy0_p: float = 0.0
    gravityX: float = 0.0
    gravityY: float = 0.0
    dpdx: float = 0.0
    dpdy: float = 0.0
    wall_left: int = VELOCITY
    wall_right: int = PRESSURE
    wall_bottom: int = WALL
    wall_top: int = WALL
    profile: int = PARABOLIC
    uxInlet_p: float = 1.0
    uyInlet_p: float = 0.0
    alpha: float = 0.25
    delta: float = 0.5
    relaxMax: float = 1.0
    subTol: float = 1e-8
    omega: float = 0.0          # > 0: set directly (params.h:85)
    tStep: float = 0.0          # used when omega <= 0 (params.h:89-90)
    central_moments: int = 0
    inlet_ramp: float = 0.0     # INLET_RAMP, <= 0: undefined
    uni_epsilon: int = 0
    initial_deflect: float = 0.0
    ref_nu: float = 0.0
    ref_rho: float = 1.0
    ref_P: float = 0.0
    ref_L: float = 1.0
    ref_U: float = 1.0
    womersley: float = 0.0      # WOMERSLEY, 0: undefined

    def replace(self, **kw):
        return dataclasses.replace(self, **kw)


def lattice_viscosity(omega):
    """nu = (1/omega - 1/2) c_s^2 with the solver's own rounding (Grid.cpp:1247, 1255-1256)."""
    cs = 1.0 / math.sqrt(3.0)
    tau = 1.0 / omega
    return (tau - 0.5) * (cs * cs)


def lattice_case(Nx, Ny, tau, **kw):
    """CaseParams in lattice units: Dx = 1 and Dt = 1 exactly."""
    omega = 1.0 / tau
    nu = lattice_viscosity(omega)
    base = dict(Nx=Nx, Ny=Ny, height_p=float(Ny - 1), rho_p=1.0, nu_p=nu, omega=omega, ref_nu=nu)
    base.update(kw)
    return CaseParams(**base)


# ----------------------------------------------------------------------------
# Named workloads
# ----------------------------------------------------------------------------

def lid_driven_cavity(N=256, tau=0.6, u_lid=0.05, central_moments=0):
    """Lid-driven cavity, N x N, BGK, regularised no-slip walls, uniform lid
    velocity on the top wall, f = f_eq(rho=1, u=0).  Re = u_lid (N-1) / nu."""
    p = lattice_case(N, N, tau, wall_left=WALL, wall_right=WALL, wall_bottom=WALL, wall_top=VELOCITY,
                     profile=UNIFORM, uxInlet_p=u_lid, uyInlet_p=0.0, central_moments=central_moments,
                     ref_L=float(N - 1), ref_U=u_lid)
    return p, ""


def womersley_channel(Nx=4, Ny=33, tau=0.8, Wo=4.0, G=1e-5, central_moments=0):
    """Oscillating pressure-gradient channel (WOMERSLEY, Grid.cpp:52-61):
    periodic in x, regularised no-slip walls on the first and last rows,
    body force G cos(2 pi t / T) with period T = H^2 pi / (2 Wo^2 nu),
    H = Ny - 1, so Wo = (H / 2) sqrt(omega / nu)."""
    return lattice_case(Nx, Ny, tau, wall_left=FLUID, wall_right=FLUID, wall_bottom=WALL, wall_top=WALL,
                        profile=UNIFORM, uxInlet_p=0.0, dpdx=G, womersley=Wo, central_moments=central_moments,
                        ref_L=float(Ny - 1), ref_U=G)


def womersley_cylinder(Nx=96, Ny=49, D=10.0, tau=0.8, Wo=4.0, G=2e-5, central_moments=1):
    """The Womersley channel with a rigid CIRCLE on the centre line (the
    oscillating force acting beside the immersed-boundary force)."""
    p = womersley_channel(Nx=Nx, Ny=Ny, tau=tau, Wo=Wo, G=G, central_moments=central_moments)
    geom = "CIRCLE 1 0 %.17g %.17g 0.0 0.0 %.17g\n" % (Nx / 3.0, (Ny - 1) / 2.0, D / 2.0)
    return p.replace(ref_L=D), geom


def pelskin(resFactor=3):
    """The reference's PELskin example (examples/PELskin/params.h and
    input/geometry.config) in its own physical units: a periodic channel of
    height 0.06 m driven by the oscillating pressure gradient dpdx = 4550
    Pa/m at Womersley number 7.52, BGK, with ten clamped flexible filaments
    (20 elements each) standing on the bottom wall, uniform epsilon system,
    subTol 1e-4.  Nx = 245 resFactor, Ny = 30 resFactor + 1, tStep =
    (0.001 / 3) / resFactor^2 (omega follows from it)."""
    p = CaseParams(Nx=resFactor * (5 * 9 + 10 * 20), Ny=resFactor * (10 * 3) + 1, height_p=0.06, rho_p=1200.0,
                   nu_p=0.0001, dpdx=4550.0, wall_left=FLUID, wall_right=FLUID, wall_bottom=WALL, wall_top=WALL,
                   profile=UNIFORM, uxInlet_p=0.0, subTol=1e-4, tStep=(0.001 / 3.0) / (resFactor * resFactor),
                   uni_epsilon=1, womersley=7.52, ref_nu=0.0001, ref_rho=1200.0, ref_L=0.02, ref_U=0.06)
    geom = "FILAMENT 10 0 0.2 0.0 0.01 0.0 0.02 0.001 90.0 FLEXIBLE 20 CLAMPED 1200.0 1.2e6\n"
    return p, geom


def reference_example(name):
    """The reference's own example cases (examples/<name>/params.h and
    input/geometry.config, run for 500 steps by testing/run-tests.sh) in
    their own physical units; omega follows from tStep as params.h:89-90
    derives it.  PELskin is :func:`pelskin`."""
    if name == "ChannelFlow":
        # BGK, uniform inlet, pressure outlet, no bodies
        p = CaseParams(Nx=500 + 1, Ny=50 + 1, height_p=1.0, rho_p=1.0, nu_p=0.01, wall_left=VELOCITY,
                       wall_right=PRESSURE, wall_bottom=WALL, wall_top=WALL, profile=UNIFORM, uxInlet_p=1.0,
                       tStep=0.0005, ref_nu=0.01, ref_rho=1.0, ref_L=1.0, ref_U=1.0)
        return p, ""
    if name == "Cylinder":
        return schafer_turek("2D-2", 2)
    if name == "Honami":
        # 128 clamped flexible filaments on the bottom wall, boundary-layer
        # inlet ramped over 250 s, free-slip top
        p = CaseParams(Nx=1200 + 127 * 15 + 1, Ny=90 + 1, height_p=3.0, rho_p=1.0, nu_p=0.0125, wall_left=VELOCITY,
                       wall_right=PRESSURE, wall_bottom=WALL, wall_top=FREESLIP, profile=BOUNDARY_LAYER,
                       uxInlet_p=1.0, tStep=0.00625 / (3.0 * 3.0), inlet_ramp=250.0, ref_nu=0.0125, ref_rho=1.0,
                       ref_L=1.0, ref_U=1.0)
        return p, "FILAMENT 128 0 10.0 0.0 0.5 0.0 1.0 0.02 90.0 FLEXIBLE 20 CLAMPED 50.000000 269536.024785\n"
    if name == "InvertedFlag":
        # one plate clamped at its trailing edge with an initial deflection,
        # uniform flow, convective outlet
        r = 3
        p = CaseParams(Nx=r * 10 * 21 + 1, Ny=r * 10 * 15 + 1, height_p=0.855, rho_p=1.2047, nu_p=1.5111e-2,
                       ux0_p=8.0, wall_left=VELOCITY, wall_right=CONVECTIVE, wall_bottom=VELOCITY,
                       wall_top=VELOCITY, profile=UNIFORM, uxInlet_p=8.0, subTol=1e-5, tStep=0.00005 / (r * r),
                       initial_deflect=0.01, ref_nu=1.5111e-2, ref_rho=1.2047, ref_L=0.057, ref_U=8.0)
        return p, "FILAMENT 1 0 0.342 0.4275 0.0 0.0 0.057 0.000205 180.0 FLEXIBLE 20 CLAMPED 1780.0 3.0e9\n"
    if name == "LidDrivenCavity":
        # central moments, moving top wall
        p = CaseParams(Nx=100 + 1, Ny=100 + 1, height_p=1.0, rho_p=1.0, nu_p=0.0002, wall_left=WALL,
                       wall_right=WALL, wall_bottom=WALL, wall_top=VELOCITY, profile=UNIFORM, uxInlet_p=1.0,
                       tStep=0.001, central_moments=1, ref_nu=0.0002, ref_rho=1.0, ref_L=1.0, ref_U=1.0)
        return p, ""
    if name == "TurekHron":
        # FSI2-type benchmark: rigid cylinder + flexible plate, one epsilon
        # system over all markers, parabolic inlet ramped over 2 s
        r = 2
        p = CaseParams(Nx=r * 250 + 1, Ny=r * 41 + 1, height_p=0.41, rho_p=1000.0, nu_p=0.001, wall_left=VELOCITY,
                       wall_right=PRESSURE, wall_bottom=WALL, wall_top=WALL, profile=PARABOLIC, uxInlet_p=1.0,
                       tStep=0.001 / (r * r), inlet_ramp=2.0, uni_epsilon=1, ref_nu=0.001, ref_rho=1000.0,
                       ref_L=0.1, ref_U=1.0)
        return p, ("CIRCLE 1 0 0.2 0.2 0.0 0.0 0.05\n"
                   "FILAMENT 1 1 0.25 0.2 0.0 0.0 0.35 0.02 0.0 FLEXIBLE 20 CLAMPED 10000.0 1.4e6\n")
    if name == "PELskin":
        return pelskin()
    raise KeyError(name)


REFERENCE_EXAMPLES = ("ChannelFlow", "Cylinder", "Honami", "InvertedFlag", "LidDrivenCavity", "PELskin", "TurekHron")


# Published bounds of the Schaefer-Turek benchmark (Schaefer & Turek 1996,
# "Benchmark computations of laminar flow around a cylinder"): drag and lift
# coefficients of the steady case 2D-1 (Re 20), maximum drag and lift and the
# Strouhal number of the periodic case 2D-2 (Re 100).
SCHAFER_TUREK_BOUNDS = {
    "2D-1": {"cd": (5.57, 5.59), "cl": (0.0104, 0.0110)},
    "2D-2": {"cd_max": (3.22, 3.24), "cl_max": (0.99, 1.01), "strouhal": (0.295, 0.305)},
}


def schafer_turek(test="2D-2", res_factor=2):
    """The reference's Cylinder example (examples/Cylinder/params.h and
    input/geometry.config) at resolution factor r = res_factor: the
    Schaefer-Turek channel of height 0.41 m and length 2.2 m, a rigid
    cylinder of diameter 0.1 m centred at (0.2, 0.2), nu = 0.001 m^2/s, BGK,
    parabolic inlet of mean uxInlet_p ramped over 2 s, pressure outlet,
    no-slip walls.  Nx = 220 r + 1, Ny = 41 r + 1 (D = 10 r lattice units);
    tStep = 0.001 / r^2, so the lattice viscosity stays 0.01 (tau 0.53) and
    the lattice inlet speed is 0.1 uxInlet_p / r (diffusive scaling: the
    compressibility error falls with r).  "2D-2" is the example as shipped
    (mean inlet 1 m/s, Re 100, r = 2); "2D-1" has mean inlet 0.2 m/s, Re 20."""
    U = {"2D-1": 0.2, "2D-2": 1.0}[test]
    r = int(res_factor)
    p = CaseParams(Nx=r * 220 + 1, Ny=r * 41 + 1, height_p=0.41, rho_p=1.0, nu_p=0.001, wall_left=VELOCITY,
                   wall_right=PRESSURE, wall_bottom=WALL, wall_top=WALL, profile=PARABOLIC, uxInlet_p=U,
                   tStep=0.001 / (r * r), inlet_ramp=2.0, ref_nu=0.001, ref_rho=1.0, ref_L=0.1, ref_U=U)
    return p, "CIRCLE 1 0 0.2 0.2 0.0 0.0 0.05\n"


def womersley_profile(y, t, H, nu, Wo, G):
    """Planar Womersley flow between walls at y = 0 and y = H driven by the
    body force G cos(omega t): u = Re{G / (i omega) (1 - cosh(k (y - H/2)) /
    cosh(k H / 2)) e^(i omega t)}, k = sqrt(i omega / nu)."""
    import cmath
    omega = 4.0 * Wo * Wo * nu / (H * H)
    k = cmath.sqrt(1j * omega / nu)
    out = []
    for yy in y:
        U = G / (1j * omega) * (1.0 - cmath.cosh(k * (yy - H / 2.0)) / cmath.cosh(k * H / 2.0))
        out.append((U * cmath.exp(1j * omega * t)).real)
    return out


def cylinder_channel(Nx=1024, Ny=256, D=40.0, centre=(200.0, 128.0), u_max=0.04, Re=100.0,
                     ramp_steps=1000.0):
    """Rigid cylinder in a channel: central moments, parabolic inlet (peak
    u_max, so mean = u_max / 1.5), convective outlet, no-slip walls, one
    rigid CIRCLE, per-body epsilon.  Re = mean * D / nu.

    The convective outlet (Grid.cpp:468-495) sets the outgoing populations
    from a velocity extrapolation and does not hold the density, so against
    the fixed-velocity inlet the mean density rises (to about 1.2 after
    40000 steps in a Schaefer-Turek-sized channel at Re 20) and force
    coefficients, which are normalised by the reference density, rise with
    it.  :func:`schafer_turek` (pressure outlet) is the case whose
    coefficients compare with published values."""
    u_mean = u_max / 1.5
    nu = u_mean * D / Re
    tau = 3.0 * nu + 0.5
    p = lattice_case(Nx, Ny, tau, wall_left=VELOCITY, wall_right=CONVECTIVE, wall_bottom=WALL, wall_top=WALL,
                     profile=PARABOLIC, uxInlet_p=u_mean, central_moments=1, inlet_ramp=ramp_steps,
                     ref_L=D, ref_U=u_mean)
    geom = "CIRCLE 1 0 %.17g %.17g 0.0 0.0 %.17g\n" % (centre[0], centre[1], D / 2.0)
    return p, geom


def flexible_fsi(Nx=2048, Ny=410, D=80.0, centre=(200.0, 200.0), L=280.0, h=16.0, n_el=40,
                 density_ratio=10.0, stiffness=1.4e3, u_mean=0.04, Re=100.0, ramp_steps=2000.0,
                 uni_epsilon=1, central_moments=1):
    """Turek-Hron-type FSI: rigid cylinder with an attached clamped flexible
    plate (corotational beam), Newmark alpha = 0.25, block Gauss-Seidel with
    Aitken relaxation, subTol = 1e-8.  E = stiffness * rho_f * u_mean^2."""
    nu = u_mean * D / Re
    tau = 3.0 * nu + 0.5
    p = lattice_case(Nx, Ny, tau, wall_left=VELOCITY, wall_right=PRESSURE, wall_bottom=WALL, wall_top=WALL,
                     profile=PARABOLIC, uxInlet_p=u_mean, central_moments=central_moments,
                     inlet_ramp=ramp_steps, uni_epsilon=uni_epsilon, ref_L=D, ref_U=u_mean)
    E = stiffness * 1.0 * u_mean * u_mean
    x0 = centre[0] + D / 2.0
    geom = ("CIRCLE 1 0 %.17g %.17g 0.0 0.0 %.17g\n" % (centre[0], centre[1], D / 2.0) +
            "FILAMENT 1 1 %.17g %.17g 0.0 0.0 %.17g %.17g 0.0 FLEXIBLE %d CLAMPED %.17g %.17g\n"
            % (x0, centre[1], L, h, n_el, density_ratio, E))
    return p, geom


def canopy(Nx=8192, Ny=4096, n_filaments=64, L=200.0, n_el=50, min_spacing=60.0, tau=0.6, u_in=0.05,
           h=4.0, density_ratio=50.0, kb_range=(1e-3, 1e-1), seed=20261202, x_range=None,
           ramp_steps=2000.0):
    """Bandwidth-roofline canopy: BGK, uniform inlet, convective outlet,
    no-slip bottom, free-slip top (as the reference's Honami case), and
    n_filaments clamped vertical flexible filaments rooted on the bottom wall
    at seeded random x positions with a minimum spacing.  Bending stiffness
    K_B = E I / (rho_f U^2 L^3), log-uniform; I = Dx h^3 / 12.  Thickness
    and density follow the Honami filaments (examples/Honami: density 50
    against fluid 1, thickness 0.02 of the length): mass ratio
    rho_s h / (rho_f L) = 1.  (With h = 1 and density 10, mass ratio 0.05,
    the coupling of the most flexible filaments stops converging within 20
    subiterations near step 9 370.)"""
    rng = random.Random(seed)
    lo, hi = x_range if x_range else (0.1 * Nx, 0.6 * Nx)
    # uniform random positions with a minimum gap: sorted uniform samples on
    # the range shortened by the gaps, then the k-th shifted by k gaps
    free = (hi - lo) - (n_filaments - 1) * min_spacing
    if n_filaments and free < 0:
        raise ValueError("cannot place %d filaments with spacing %g" % (n_filaments, min_spacing))
    base = sorted(rng.uniform(0.0, free) for _ in range(n_filaments))
    xs = [float(math.floor(lo + b + k * min_spacing)) for k, b in enumerate(base)]
    I = 1.0 * h ** 3 / 12.0
    lines = []
    for k, x in enumerate(xs):
        kb = math.exp(rng.uniform(math.log(kb_range[0]), math.log(kb_range[1])))
        E = kb * 1.0 * u_in * u_in * L ** 3 / I
        lines.append("FILAMENT 1 %d %.17g 0.0 0.0 0.0 %.17g %.17g 90.0 FLEXIBLE %d CLAMPED %.17g %.17g"
                     % (k, x, L, h, n_el, density_ratio, E))
    p = lattice_case(Nx, Ny, tau, wall_left=VELOCITY, wall_right=CONVECTIVE, wall_bottom=WALL,
                     wall_top=FREESLIP, profile=UNIFORM, uxInlet_p=u_in, central_moments=0,
                     inlet_ramp=ramp_steps, uni_epsilon=0, ref_L=L, ref_U=u_in)
    return p, "\n".join(lines) + ("\n" if lines else "")


def inverted_flag(Nx=512, Ny=256, length_frac=0.25, n_el=20, Re=200.0, kb=0.5, u_in=0.04,
                  density_ratio=10.0, h=1.0, central_moments=0, initial_deflect=0.0):
    """Inverted flag (the reference's InvertedFlag case): one flexible
    plate clamped at its trailing edge (angle 180 degrees), uniform inflow,
    convective outlet, velocity top/bottom.  Plate length = length_frac * Ny.
    initial_deflect: INITIAL_DEFLECT (fraction of L; the reference's
    InvertedFlag example uses 0.01)."""
    L = float(int(length_frac * Ny))
    nu = u_in * L / Re
    tau = 3.0 * nu + 0.5
    I = 1.0 * h ** 3 / 12.0
    E = kb * 1.0 * u_in * u_in * L ** 3 / I
    x0 = 0.4 * Nx + L
    y0 = 0.5 * (Ny - 1)
    p = lattice_case(Nx, Ny, tau, wall_left=VELOCITY, wall_right=CONVECTIVE, wall_bottom=VELOCITY,
                     wall_top=VELOCITY, profile=UNIFORM, uxInlet_p=u_in, ux0_p=u_in,
                     central_moments=central_moments, subTol=1e-5, ref_L=L, ref_U=u_in,
                     initial_deflect=initial_deflect)
    geom = ("FILAMENT 1 0 %.17g %.17g 0.0 0.0 %.17g %.17g 180.0 FLEXIBLE %d CLAMPED %.17g %.17g\n"
            % (x0, y0, L, h, n_el, density_ratio, E))
    return p, geom


def serving_stream(n=256, rate=4.0, steps=5000, seed=20261203, widths=(512, 1024, 2048),
                   heights=(128, 256, 512), n_el_range=(20, 60), Re_range=(50.0, 400.0),
                   kb_range=(0.1, 1.0), central_moments=1):
    """Stream of independent inverted-flag cases with seeded Poisson arrivals
    (mean `rate` per second): width and height uniform over the given sets,
    plate length 0.25 * height, element count uniform in n_el_range, Re
    uniform, K_B log-uniform.  Central-moments collision (the reference's
    more stable operator at the small relaxation times of Re = 400 on a
    32-site plate)."""
    rng = random.Random(seed)
    t = 0.0
    out = []
    for k in range(n):
        t += rng.expovariate(rate)
        Nx = rng.choice(widths)
        Ny = rng.choice(heights)
        n_el = rng.randint(n_el_range[0], n_el_range[1])
        Re = rng.uniform(Re_range[0], Re_range[1])
        kb = math.exp(rng.uniform(math.log(kb_range[0]), math.log(kb_range[1])))
        p, g = inverted_flag(Nx, Ny, 0.25, n_el, Re, kb, central_moments=central_moments)
        out.append({"arrival": t, "params": p, "geometry": g, "steps": steps, "Nx": Nx, "Ny": Ny,
                    "n_el": n_el, "Re": Re, "kb": kb})
    return out
