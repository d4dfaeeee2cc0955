"""ctypes binding of liblife_hip.so (the C ABI in include/life_hip.h).

Loading fails loudly when the library has not been built: the package has
no CPU fallback for any entry point.
"""
import atexit
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liblife_hip.so")

LIFE_FP64 = 0
LIFE_FP32_SHIFTED = 1


class LifeParams(C.Structure):
    """struct life_params (include/life_hip.h), the runtime params.h."""
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


def make_params(p):
    s = LifeParams()
    for name, _ in LifeParams._fields_:
        setattr(s, name, getattr(p, name))
    return s


class ServerConfig(C.Structure):
    """struct life_server_config (include/life_hip.h)."""
    _fields_ = [("device", C.c_int), ("precision", C.c_int), ("max_cases", C.c_int),
                ("hbm_budget", C.c_double), ("use_graph", C.c_int), ("check_every", C.c_int)]


LIFE_REQ_QUEUED, LIFE_REQ_RUNNING, LIFE_REQ_DONE, LIFE_REQ_FAILED = 0, 1, 2, 3


class RequestInfo(C.Structure):
    """struct life_request_info (include/life_hip.h)."""
    _fields_ = [("status", C.c_int), ("steps_done", C.c_int), ("nsteps", C.c_int), ("subits", C.c_int), ("markers", C.c_int),
                ("t_arrival", C.c_double), ("t_admit", C.c_double), ("t_finish", C.c_double),
                ("sites", C.c_double), ("bytes", C.c_double),
                ("total_force", C.c_double * 4), ("tip", C.c_double * 4)]


# name -> (restype, argtypes)
_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_vp = C.c_void_p
SIGNATURES = {
    "life_last_error": (C.c_char_p, []),
    "life_version": (C.c_char_p, []),
    "life_device_count": (C.c_int, [_ip]),
    "life_release_cached": (C.c_int, []),
    "life_grid_create": (C.c_int, [C.POINTER(LifeParams), C.c_int, C.c_int, C.POINTER(_vp)]),
    "life_grid_destroy": (C.c_int, [_vp]),
    "life_grid_scalars": (C.c_int, [_vp, _dp]),
    "life_lbm_kernel": (C.c_int, [_vp, C.c_int]),
    "life_grid_download": (C.c_int, [_vp, _dp, _dp, _dp, _dp]),
    "life_grid_upload": (C.c_int, [_vp, _dp, _dp, _dp, _dp]),
    "life_grid_sync": (C.c_int, [_vp]),
    "life_grid_blowup_check": (C.c_int, [_vp, _ip, _ip]),
    "life_ibm_set_markers": (C.c_int, [_vp, C.c_int, _dp, _dp, _ip, C.c_int, _ip]),
    "life_ibm_update_markers": (C.c_int, [_vp, _dp, _dp]),
    "life_ibm_find_support": (C.c_int, [_vp, C.c_int]),
    "life_ibm_epsilon_matrix": (C.c_int, [_vp, C.c_int, C.c_int, _dp]),
    "life_ibm_epsilon_matrix_noalloc": (C.c_int, [_vp, C.c_int, C.c_int, _dp]),
    "life_ibm_set_epsilon": (C.c_int, [_vp, _dp]),
    "life_ibm_interpolate": (C.c_int, [_vp, _dp]),
    "life_ibm_spread": (C.c_int, [_vp]),
    "life_ibm_download": (C.c_int, [_vp, _dp, _dp, _dp, _dp, _dp, _ip, _ip, _dp]),
    "life_solve_lapack": (C.c_int, [C.c_int, C.c_int, _dp, _dp, _dp]),
    "life_case_create": (C.c_int, [C.POINTER(LifeParams), C.c_char_p, C.c_int, C.c_int, C.POINTER(_vp)]),
    "life_case_destroy": (C.c_int, [_vp]),
    "life_case_step": (C.c_int, [_vp, C.c_int]),
    "life_case_step_timed": (C.c_int, [_vp, C.c_int, _dp, _dp]),
    "life_case_time": (C.c_int, [_vp, _ip]),
    "life_case_set_check_interval": (C.c_int, [_vp, C.c_int]),
    "life_case_grid": (_vp, [_vp]),
    "life_case_num_markers": (C.c_int, [_vp, _ip]),
    "life_case_num_bodies": (C.c_int, [_vp, _ip]),
    "life_case_body_info": (C.c_int, [_vp, C.c_int, _ip]),
    "life_case_get_fem": (C.c_int, [_vp, C.c_int, _dp, _dp, _dp]),
    "life_case_coupling_info": (C.c_int, [_vp, _dp]),
    "life_case_total_force": (C.c_int, [_vp, _dp]),
    "life_case_write_restart": (C.c_int, [_vp, C.c_char_p]),
    "life_case_read_restart": (C.c_int, [_vp, C.c_char_p]),
    "life_case_write_vtk": (C.c_int, [_vp, C.c_char_p, C.c_int]),
    "life_case_write_forces": (C.c_int, [_vp, C.c_char_p]),
    "life_case_write_tips": (C.c_int, [_vp, C.c_char_p]),
    "life_server_create": (C.c_int, [C.POINTER(ServerConfig), C.POINTER(_vp)]),
    "life_server_destroy": (C.c_int, [_vp]),
    "life_server_submit": (C.c_int, [_vp, C.POINTER(LifeParams), C.c_char_p, C.c_int, C.c_int, _ip]),
    "life_server_step": (C.c_int, [_vp, _ip, _ip]),
    "life_server_run": (C.c_int, [_vp, C.c_int, C.POINTER(LifeParams), C.POINTER(C.c_char_p), _ip, _dp, _ip]),
    "life_server_info": (C.c_int, [_vp, C.c_int, C.POINTER(RequestInfo)]),
    "life_server_request_error": (C.c_char_p, [_vp, C.c_int]),
    "life_server_fields": (C.c_int, [_vp, C.c_int, _dp, _dp, _dp]),
    "life_server_release": (C.c_int, [_vp, C.c_int]),
    "life_server_stats": (C.c_int, [_vp, _dp]),
}

_lib = None


class LifeError(RuntimeError):
    pass


def load():
    """Load liblife_hip.so (build it first with life_amd.build())."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise LifeError("liblife_hip.so is not built (%s); run life_amd.build()" % LIB_PATH)
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
        # pooled streams and blocks go back to HIP before the interpreter
        # (and with it the HIP runtime) shuts down
        atexit.register(L.life_release_cached)
    return _lib


def check(rc):
    if rc != 0:
        raise LifeError(load().life_last_error().decode())


def device_count():
    n = C.c_int(0)
    check(load().life_device_count(C.byref(n)))
    return n.value


def dptr(a):
    return None if a is None else a.ctypes.data_as(_dp)


def iptr(a):
    return None if a is None else a.ctypes.data_as(_ip)


def solve_lapack(A, b, bc=0):
    """Utils::solveLAPACK replacement (host FP64 canonical elimination)."""
    import numpy as np
    A = np.ascontiguousarray(A, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    x = np.zeros_like(b)
    check(load().life_solve_lapack(A.shape[0], int(bc), dptr(A), dptr(b), dptr(x)))
    return x
