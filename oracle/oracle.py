"""ctypes front-end to the oracle (test infrastructure only).

The oracle is the plain-C FP64 restatement of the LIFE reference path in
``life_oracle.c``.  Only tests, ``__graft_entry__.smoke()`` and the
CPU-baseline leg of ``bench.py`` import this module; the product never does.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liblife_oracle.so")

FLUID, WALL, VELOCITY, FREESLIP, PRESSURE, CONVECTIVE = range(6)


class OrcParams(C.Structure):
    _fields_ = [
        ("Nx", C.c_int), ("Ny", C.c_int),
        ("height_p", C.c_double), ("rho_p", C.c_double), ("nu_p", C.c_double),
        ("ux0_p", C.c_double), ("uy0_p", C.c_double),
        ("gravityX", C.c_double), ("gravityY", C.c_double), ("dpdx", C.c_double), ("dpdy", C.c_double),
        ("wall_left", C.c_int), ("wall_right", C.c_int), ("wall_bottom", C.c_int), ("wall_top", C.c_int),
        ("profile", C.c_int),
        ("uxInlet_p", C.c_double), ("uyInlet_p", C.c_double),
        ("alpha", C.c_double), ("delta", C.c_double),
        ("relaxMax", C.c_double), ("subTol", C.c_double),
        ("omega", C.c_double), ("tStep", C.c_double),
        ("central_moments", C.c_int),
        ("inlet_ramp", C.c_double),
        ("uni_epsilon", C.c_int),
        ("initial_deflect", C.c_double),
        ("ref_nu", C.c_double), ("ref_rho", C.c_double), ("ref_P", C.c_double),
        ("ref_L", C.c_double), ("ref_U", C.c_double),
        ("womersley", C.c_double),
    ]


def build():
    """Compile the oracle shared library in place (gcc, no GPU)."""
    subprocess.check_call(["make", "-s", "-C", _HERE], stdout=subprocess.DEVNULL)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = C.CDLL(_LIB)
        dp = C.POINTER(C.c_double)
        ip = C.POINTER(C.c_int)
        vp = C.c_void_p
        L.orc_last_error.restype = C.c_char_p
        L.orc_create.restype = vp
        L.orc_create.argtypes = [C.POINTER(OrcParams), C.c_char_p]
        L.orc_destroy.argtypes = [vp]
        L.orc_scalars.argtypes = [vp, dp]
        L.orc_time.argtypes = [vp]
        L.orc_step.argtypes = [vp, C.c_int]
        L.orc_lbm_kernel.argtypes = [vp]
        L.orc_get_grid.argtypes = [vp, dp, dp, dp, dp]
        L.orc_set_grid.argtypes = [vp, dp, dp, dp, dp]
        L.orc_get_types.argtypes = [vp, ip]
        L.orc_num_markers.argtypes = [vp]
        L.orc_num_bodies.argtypes = [vp]
        L.orc_get_markers.argtypes = [vp, dp, dp, dp, dp, dp, ip]
        L.orc_get_supports.argtypes = [vp, ip, ip, dp]
        L.orc_body_info.argtypes = [vp, C.c_int, ip]
        L.orc_get_fem.argtypes = [vp, C.c_int, dp, dp, dp]
        L.orc_body_nodes.argtypes = [vp, C.c_int, ip]
        L.orc_get_fem_state.argtypes = [vp, C.c_int, ip, dp, dp, dp, dp, dp]
        L.orc_coupling_info.argtypes = [vp, dp]
        L.orc_total_force.argtypes = [vp, dp]
        L.orc_dirac_delta.restype = C.c_double
        L.orc_dirac_delta.argtypes = [C.c_double]
        L.orc_find_support.argtypes = [C.c_double, C.c_double, C.c_double, C.c_int, C.c_int, ip, dp]
        L.orc_lu_solve.argtypes = [C.c_int, C.c_int, dp, dp, dp]
        L.orc_collide_site.argtypes = [C.c_int, C.c_double, dp, C.c_double, C.c_double, C.c_double,
                                       C.c_double, C.c_double, dp]
        L.orc_equilibrium.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, dp]
        L.orc_set_threads.argtypes = [C.c_int]
        L.orc_fem_cantilever.argtypes = [C.c_double, C.c_double, C.c_double, C.c_double, C.c_int, C.c_double,
                                         C.c_double, C.c_int, dp]
        L.orc_ibm_interpolate.argtypes = [vp]
        L.orc_ibm_spread.argtypes = [vp]
        L.orc_set_markers.argtypes = [vp, dp, dp, dp, dp, dp]
        L.orc_ibm_supports.argtypes = [vp]
        L.orc_epsilon_matrix.argtypes = [vp, C.c_int, dp]
        L.orc_get_interp.argtypes = [vp, dp, dp]
        L.orc_normal_vector.argtypes = [vp, C.c_int, C.c_int, ip, ip]
        L.orc_get_delu.argtypes = [vp, dp]
        L.orc_get_inlet.argtypes = [vp, dp]
        L.orc_fem_load.argtypes = [vp, C.c_int, dp]
        L.orc_aitken_relax.restype = C.c_double
        L.orc_aitken_relax.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, C.c_double]
        L.orc_sub_residual.argtypes = [C.c_int, dp, dp, dp, dp, dp]
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def set_threads(n):
    lib().orc_set_threads(int(n))


def make_params(p):
    """Fill an OrcParams from any object with the same attribute names."""
    s = OrcParams()
    for name, _ in OrcParams._fields_:
        setattr(s, name, getattr(p, name))
    return s


class OracleCase:
    """One LIFE case run by the oracle: grid + bodies + FEM, FP64, CPU."""

    def __init__(self, params, geometry=""):
        L = lib()
        self._p = make_params(params)
        h = L.orc_create(C.byref(self._p), geometry.encode() if geometry else None)
        if not h:
            raise RuntimeError("oracle: " + L.orc_last_error().decode())
        self.h = h
        self.Nx, self.Ny = params.Nx, params.Ny
        self.N = self.Nx * self.Ny

    def close(self):
        if self.h:
            lib().orc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def scalars(self):
        out = np.zeros(10)
        lib().orc_scalars(self.h, _dp(out))
        keys = ["omega", "tau", "nu", "Dx", "Dt", "Dm", "Drho", "cs", "cs2", "cs4"]
        return dict(zip(keys, out.tolist()))

    @property
    def t(self):
        return lib().orc_time(self.h)

    def step(self, n=1):
        if lib().orc_step(self.h, int(n)) != 0:
            raise RuntimeError("oracle: " + lib().orc_last_error().decode())

    def lbm_kernel(self):
        if lib().orc_lbm_kernel(self.h) != 0:
            raise RuntimeError("oracle: " + lib().orc_last_error().decode())

    def grid(self):
        f = np.zeros(self.N * 9)
        rho = np.zeros(self.N)
        u = np.zeros(self.N * 2)
        fibm = np.zeros(self.N * 2)
        lib().orc_get_grid(self.h, _dp(f), _dp(rho), _dp(u), _dp(fibm))
        return (f.reshape(self.Nx, self.Ny, 9), rho.reshape(self.Nx, self.Ny),
                u.reshape(self.Nx, self.Ny, 2), fibm.reshape(self.Nx, self.Ny, 2))

    def set_grid(self, f=None, rho=None, u=None, fibm=None):
        def arg(a):
            if a is None:
                return None, None
            a = np.ascontiguousarray(a, dtype=np.float64).ravel()
            return a, _dp(a)
        fa, fp = arg(f)
        ra, rp = arg(rho)
        ua, up = arg(u)
        ba, bp = arg(fibm)
        lib().orc_set_grid(self.h, fp, rp, up, bp)

    def types(self):
        t = np.zeros(self.N, dtype=np.int32)
        lib().orc_get_types(self.h, _ip(t))
        return t.reshape(self.Nx, self.Ny)

    @property
    def n_markers(self):
        return lib().orc_num_markers(self.h)

    @property
    def n_bodies(self):
        return lib().orc_num_bodies(self.h)

    def markers(self):
        n = self.n_markers
        pos, vel, force = np.zeros((n, 2)), np.zeros((n, 2)), np.zeros((n, 2))
        eps, ds = np.zeros(n), np.zeros(n)
        body = np.zeros(n, dtype=np.int32)
        lib().orc_get_markers(self.h, _dp(pos), _dp(vel), _dp(force), _dp(eps), _dp(ds), _ip(body))
        return dict(pos=pos, vel=vel, force=force, eps=eps, ds=ds, body=body)

    def supports(self):
        n = self.n_markers
        cnt = np.zeros(n, dtype=np.int32)
        ij = np.zeros((n, 9, 2), dtype=np.int32)
        d = np.zeros((n, 9))
        lib().orc_get_supports(self.h, _ip(cnt), _ip(ij), _dp(d))
        return cnt, ij, d

    def body_info(self, b):
        info = np.zeros(5, dtype=np.int32)
        lib().orc_body_info(self.h, int(b), _ip(info))
        return dict(ID=int(info[0]), flex=int(info[1]), n=int(info[2]), first=int(info[3]), ndof=int(info[4]))

    def body_nodes(self, b):
        """(shape, marker indices): shape 0 circle, 1 filament; the body's
        markers in its own node order."""
        idx = np.zeros(self.body_info(b)["n"], np.int32)
        kind = lib().orc_body_nodes(self.h, int(b), _ip(idx))
        return int(kind), idx

    def fem(self, b):
        nd = self.body_info(b)["ndof"]
        U, Ud, Udd = np.zeros(nd), np.zeros(nd), np.zeros(nd)
        lib().orc_get_fem(self.h, int(b), _dp(U), _dp(Ud), _dp(Udd))
        return U, Ud, Udd

    def fem_state(self, b):
        """U_n, U_nm1, R_k, per-element (L, angle) and per-node (pos0, pos,
        angle0, angle) of flexible body b."""
        nd = self.body_info(b)["ndof"]
        sizes = np.zeros(2, np.int32)
        lib().orc_get_fem_state(self.h, int(b), _ip(sizes), None, None, None, None, None)
        nel, nnode = int(sizes[0]), int(sizes[1])
        Un, Unm1, Rk = np.zeros(nd), np.zeros(nd), np.zeros(nd)
        el, nodes = np.zeros((nel, 2)), np.zeros((nnode, 6))
        lib().orc_get_fem_state(self.h, int(b), _ip(sizes), _dp(Un), _dp(Unm1), _dp(Rk), _dp(el), _dp(nodes))
        return dict(U_n=Un, U_nm1=Unm1, R_k=Rk, elements=el, nodes=nodes)

    def coupling(self):
        out = np.zeros(3)
        lib().orc_coupling_info(self.h, _dp(out))
        return dict(subIt=int(out[0]), subRes=out[1], relax=out[2])

    def total_force(self):
        out = np.zeros(4)
        lib().orc_total_force(self.h, _dp(out))
        return out

    # -- unit-level access to single restated steps --------------------------

    def ibm_interpolate(self):
        lib().orc_ibm_interpolate(self.h)

    def ibm_spread(self):
        lib().orc_ibm_spread(self.h)

    def set_markers(self, pos=None, vel=None, force=None, eps=None, ds=None):
        keep = []

        def arg(a):
            if a is None:
                return None
            a = np.ascontiguousarray(a, dtype=np.float64).ravel()
            keep.append(a)
            return _dp(a)
        lib().orc_set_markers(self.h, arg(pos), arg(vel), arg(force), arg(eps), arg(ds))

    def ibm_supports(self):
        if lib().orc_ibm_supports(self.h) != 0:
            raise RuntimeError("oracle: " + lib().orc_last_error().decode())

    def epsilon_matrix(self, sys=0):
        n = lib().orc_epsilon_matrix(self.h, int(sys), None)
        A = np.zeros((n, n))
        lib().orc_epsilon_matrix(self.h, int(sys), _dp(A))
        return A

    def interp(self):
        n = self.n_markers
        rho, mom = np.zeros(n), np.zeros((n, 2))
        lib().orc_get_interp(self.h, _dp(rho), _dp(mom))
        return rho, mom

    def normal(self, i, j):
        n = np.zeros(2, dtype=np.int32)
        nd = C.c_int(0)
        if lib().orc_normal_vector(self.h, int(i), int(j), _ip(n), C.byref(nd)) != 0:
            raise RuntimeError("oracle: " + lib().orc_last_error().decode())
        return (int(n[0]), int(n[1])), int(nd.value)

    def delu(self):
        d = np.zeros((self.Ny, 2))
        lib().orc_get_delu(self.h, _dp(d))
        return d

    def inlet(self):
        d = np.zeros((self.Ny, 2))
        lib().orc_get_inlet(self.h, _dp(d))
        return d

    def fem_load(self, b):
        nd = self.body_info(b)["ndof"]
        R = np.zeros(nd)
        lib().orc_fem_load(self.h, int(b), _dp(R))
        return R


def dirac_delta(x):
    return lib().orc_dirac_delta(float(x))


def find_support(x, y, Dx, Nx, Ny):
    ij = np.zeros(18, dtype=np.int32)
    d = np.zeros(9)
    n = lib().orc_find_support(float(x), float(y), float(Dx), int(Nx), int(Ny), _ip(ij), _dp(d))
    return ij[: 2 * n].reshape(n, 2), d[:n]


def lu_solve(A, b, bc=0):
    A = np.ascontiguousarray(A, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    x = np.zeros_like(b)
    lib().orc_lu_solve(A.shape[0], int(bc), _dp(A), _dp(b), _dp(x))
    return x


def collide_site(cm, omega, fsite, rho, ux, uy, Fx=0.0, Fy=0.0):
    fs = np.ascontiguousarray(fsite, dtype=np.float64)
    out = np.zeros(9)
    lib().orc_collide_site(int(cm), float(omega), _dp(fs), float(rho), float(ux), float(uy),
                           float(Fx), float(Fy), _dp(out))
    return out


def fem_cantilever(L, h, E, rho, nel, Dt, P, nsteps=0):
    """Static tip deflection (nsteps=0) or tip history after release."""
    out = np.zeros(max(1, nsteps))
    lib().orc_fem_cantilever(float(L), float(h), float(E), float(rho), int(nel), float(Dt), float(P),
                             int(nsteps), _dp(out))
    return out[0] if nsteps == 0 else out


def aitken_relax(sub_it, relax, relax_max, num, den):
    return lib().orc_aitken_relax(int(sub_it), float(relax), float(relax_max), float(num), float(den))


def sub_residual(U, Ukm1, Rkm1):
    """(R_k, |R_k|^2, numerator, denominator) of the subiteration residual."""
    U = np.ascontiguousarray(U, dtype=np.float64)
    Ukm1 = np.ascontiguousarray(Ukm1, dtype=np.float64)
    Rkm1 = np.ascontiguousarray(Rkm1, dtype=np.float64)
    Rk = np.zeros_like(U)
    out = np.zeros(3)
    lib().orc_sub_residual(len(U), _dp(U), _dp(Ukm1), _dp(Rk), _dp(Rkm1), _dp(out))
    return Rk, out[0], out[1], out[2]


def equilibrium(cm, rho, ux, uy):
    out = np.zeros(9)
    lib().orc_equilibrium(int(cm), float(rho), float(ux), float(uy), _dp(out))
    return out
