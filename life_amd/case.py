"""Python handles over the C ABI: :class:`Grid` (layer 1, the device
lattice that replaces GridClass's arrays and lbmKernel) and :class:`Case`
(layer 2, a whole case: lattice + bodies + FP64 FEM coupling)."""
import ctypes as C

import numpy as np

from . import lib as L


class Grid:
    """Device lattice of one case (GridClass state, Grid.h:76-92)."""

    def __init__(self, params, precision=L.LIFE_FP64, device=0, _handle=None):
        self.params = params
        self.Nx, self.Ny = params.Nx, params.Ny
        self.N = self.Nx * self.Ny
        self._owner = _handle is None
        if _handle is None:
            h = C.c_void_p()
            self._p = L.make_params(params)
            L.check(L.load().life_grid_create(C.byref(self._p), int(precision), int(device), C.byref(h)))
            self.h = h
        else:
            self.h = C.c_void_p(_handle)

    def close(self):
        if self._owner and self.h:
            L.load().life_grid_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def scalars(self):
        out = np.zeros(10)
        L.check(L.load().life_grid_scalars(self.h, L.dptr(out)))
        keys = ["omega", "tau", "nu", "Dx", "Dt", "Dm", "Drho", "cs", "cs2", "cs4"]
        return dict(zip(keys, out.tolist()))

    def lbm_kernel(self, t):
        L.check(L.load().life_lbm_kernel(self.h, int(t)))

    def sync(self):
        L.check(L.load().life_grid_sync(self.h))

    def blowup_check(self):
        """(i, j) of the first site with a NaN velocity, or None."""
        i, j = C.c_int(-1), C.c_int(-1)
        L.check(L.load().life_grid_blowup_check(self.h, C.byref(i), C.byref(j)))
        return None if i.value < 0 else (i.value, j.value)

    def download(self):
        f = np.zeros(self.N * 9)
        rho = np.zeros(self.N)
        u = np.zeros(self.N * 2)
        fibm = np.zeros(self.N * 2)
        L.check(L.load().life_grid_download(self.h, L.dptr(f), L.dptr(rho), L.dptr(u), L.dptr(fibm)))
        return (f.reshape(self.Nx, self.Ny, 9), rho.reshape(self.Nx, self.Ny),
                u.reshape(self.Nx, self.Ny, 2), fibm.reshape(self.Nx, self.Ny, 2))

    def upload(self, f, rho, u, fibm=None):
        f = np.ascontiguousarray(f, dtype=np.float64).ravel()
        rho = np.ascontiguousarray(rho, dtype=np.float64).ravel()
        u = np.ascontiguousarray(u, dtype=np.float64).ravel()
        fb = None if fibm is None else np.ascontiguousarray(fibm, dtype=np.float64).ravel()
        L.check(L.load().life_grid_upload(self.h, L.dptr(f), L.dptr(rho), L.dptr(u), L.dptr(fb)))

    # -- markers (layer 1 IBM entry points)
    def set_markers(self, pos, vel, body, body_flex):
        pos = np.ascontiguousarray(pos, dtype=np.float64)
        vel = np.ascontiguousarray(vel, dtype=np.float64)
        body = np.ascontiguousarray(body, dtype=np.int32)
        bf = np.ascontiguousarray(body_flex, dtype=np.int32)
        self.n_markers = len(body)
        L.check(L.load().life_ibm_set_markers(self.h, len(body), L.dptr(pos), L.dptr(vel), L.iptr(body),
                                             len(bf), L.iptr(bf)))

    def update_markers(self, pos=None, vel=None):
        pos = None if pos is None else np.ascontiguousarray(pos, dtype=np.float64)
        vel = None if vel is None else np.ascontiguousarray(vel, dtype=np.float64)
        L.check(L.load().life_ibm_update_markers(self.h, L.dptr(pos), L.dptr(vel)))

    def find_support(self, flexible_only=0):
        L.check(L.load().life_ibm_find_support(self.h, int(flexible_only)))

    def epsilon_matrix(self, first, n):
        A = np.zeros((n, n))
        L.check(L.load().life_ibm_epsilon_matrix(self.h, int(first), int(n), L.dptr(A)))
        return A

    def epsilon_matrix_noalloc(self, first, n):
        """The same matrix through the allocation-free entry point."""
        A = np.zeros((n, n))
        L.check(L.load().life_ibm_epsilon_matrix_noalloc(self.h, int(first), int(n), L.dptr(A)))
        return A

    def set_epsilon(self, eps):
        eps = np.ascontiguousarray(eps, dtype=np.float64)
        L.check(L.load().life_ibm_set_epsilon(self.h, L.dptr(eps)))

    def interpolate(self):
        force = np.zeros((self.n_markers, 2))
        L.check(L.load().life_ibm_interpolate(self.h, L.dptr(force)))
        return force

    def spread(self):
        L.check(L.load().life_ibm_spread(self.h))

    def markers(self, n=None):
        n = self.n_markers if n is None else n
        pos, vel, force = np.zeros((n, 2)), np.zeros((n, 2)), np.zeros((n, 2))
        eps, ds = np.zeros(n), np.zeros(n)
        cnt = np.zeros(n, dtype=np.int32)
        ij = np.zeros((n, 9, 2), dtype=np.int32)
        d = np.zeros((n, 9))
        L.check(L.load().life_ibm_download(self.h, L.dptr(pos), L.dptr(vel), L.dptr(force), L.dptr(eps),
                                          L.dptr(ds), L.iptr(cnt), L.iptr(ij), L.dptr(d)))
        return dict(pos=pos, vel=vel, force=force, eps=eps, ds=ds, supp_count=cnt, supp_ij=ij, supp_delta=d)


class Case:
    """One LIFE case on the GPU: lattice, bodies, FP64 FEM and coupling."""

    def __init__(self, params, geometry="", precision=L.LIFE_FP64, device=0):
        self.params = params
        self.Nx, self.Ny = params.Nx, params.Ny
        self.N = self.Nx * self.Ny
        self._p = L.make_params(params)
        h = C.c_void_p()
        L.check(L.load().life_case_create(C.byref(self._p), geometry.encode() if geometry else None,
                                          int(precision), int(device), C.byref(h)))
        self.h = h
        self.grid = Grid(params, _handle=L.load().life_case_grid(self.h))

    def close(self):
        if self.h:
            L.load().life_case_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def t(self):
        t = C.c_int(0)
        L.check(L.load().life_case_time(self.h, C.byref(t)))
        return t.value

    def step(self, n=1):
        L.check(L.load().life_case_step(self.h, int(n)))

    def set_check_interval(self, every):
        """Blow-up test cadence in steps (0: off); see life_case_set_check_interval."""
        L.check(L.load().life_case_set_check_interval(self.h, int(every)))

    def step_timed(self, n):
        tot, ker = C.c_double(0), C.c_double(0)
        L.check(L.load().life_case_step_timed(self.h, int(n), C.byref(tot), C.byref(ker)))
        return tot.value, ker.value

    @property
    def n_markers(self):
        n = C.c_int(0)
        L.check(L.load().life_case_num_markers(self.h, C.byref(n)))
        return n.value

    @property
    def n_bodies(self):
        n = C.c_int(0)
        L.check(L.load().life_case_num_bodies(self.h, C.byref(n)))
        return n.value

    def body_info(self, b):
        info = np.zeros(5, dtype=np.int32)
        L.check(L.load().life_case_body_info(self.h, int(b), L.iptr(info)))
        return dict(ID=int(info[0]), flex=int(info[1]), n=int(info[2]), first=int(info[3]), ndof=int(info[4]))

    def fem(self, b):
        nd = self.body_info(b)["ndof"]
        U, Ud, Udd = np.zeros(nd), np.zeros(nd), np.zeros(nd)
        L.check(L.load().life_case_get_fem(self.h, int(b), L.dptr(U), L.dptr(Ud), L.dptr(Udd)))
        return U, Ud, Udd

    # result files in the reference's formats (Results/ layout)
    def write_restart(self, results):
        L.check(L.load().life_case_write_restart(self.h, str(results).encode()))

    def read_restart(self, results):
        L.check(L.load().life_case_read_restart(self.h, str(results).encode()))

    def write_vtk(self, results, fem=False):
        L.check(L.load().life_case_write_vtk(self.h, str(results).encode(), 1 if fem else 0))

    def write_forces(self, results):
        L.check(L.load().life_case_write_forces(self.h, str(results).encode()))

    def write_tips(self, results):
        L.check(L.load().life_case_write_tips(self.h, str(results).encode()))

    def coupling(self):
        out = np.zeros(3)
        L.check(L.load().life_case_coupling_info(self.h, L.dptr(out)))
        return dict(subIt=int(out[0]), subRes=out[1], relax=out[2])

    def total_force(self):
        out = np.zeros(4)
        L.check(L.load().life_case_total_force(self.h, L.dptr(out)))
        return out

    def fields(self):
        return self.grid.download()

    def markers(self):
        self.grid.n_markers = self.n_markers
        return self.grid.markers()
