"""Python handle over the serving engine (life_server_* in
include/life_hip.h): continuous batching of independent cases on one
device."""
import ctypes as C

import numpy as np

from . import lib as L


class Server:
    def __init__(self, precision=L.LIFE_FP32_SHIFTED, max_cases=64, hbm_budget=200e9, use_graph=True, device=0,
                 check_every=0):
        self.cfg = L.ServerConfig(int(device), int(precision), int(max_cases), float(hbm_budget),
                                  1 if use_graph else 0, int(check_every))
        h = C.c_void_p()
        L.check(L.load().life_server_create(C.byref(self.cfg), C.byref(h)))
        self.h = h
        self._meta = {}

    def close(self):
        if self.h:
            L.load().life_server_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, params, geometry, nsteps, keep_fields=False):
        p = L.make_params(params)
        rid = C.c_int(-1)
        L.check(L.load().life_server_submit(self.h, C.byref(p), geometry.encode(), int(nsteps),
                                            1 if keep_fields else 0, C.byref(rid)))
        self._meta[rid.value] = (params.Nx, params.Ny)
        return rid.value

    def step(self):
        run, q = C.c_int(0), C.c_int(0)
        L.check(L.load().life_server_step(self.h, C.byref(run), C.byref(q)))
        return run.value, q.value

    def drain(self, max_steps=10 ** 9):
        """Step until nothing is running or queued."""
        for _ in range(max_steps):
            run, q = self.step()
            if run == 0 and q == 0:
                return
        raise RuntimeError("server did not drain")

    def run(self, stream):
        """Persistent loop over a timed stream (list of dicts with arrival,
        params, geometry, steps); returns the request ids."""
        n = len(stream)
        ps = (L.LifeParams * n)(*[L.make_params(r["params"]) for r in stream])
        geo = (C.c_char_p * n)(*[r["geometry"].encode() for r in stream])
        steps = np.ascontiguousarray([r["steps"] for r in stream], dtype=np.int32)
        arr = np.ascontiguousarray([r["arrival"] for r in stream], dtype=np.float64)
        ids = np.zeros(n, dtype=np.int32)
        L.check(L.load().life_server_run(self.h, n, ps, geo, L.iptr(steps), L.dptr(arr), L.iptr(ids)))
        for r, i in zip(stream, ids):
            self._meta[int(i)] = (r["params"].Nx, r["params"].Ny)
        return [int(i) for i in ids]

    def info(self, rid):
        r = L.RequestInfo()
        L.check(L.load().life_server_info(self.h, int(rid), C.byref(r)))
        d = {name: getattr(r, name) for name, _ in L.RequestInfo._fields_}
        d["total_force"] = list(r.total_force)
        d["tip"] = list(r.tip)
        return d

    def error(self, rid):
        return L.load().life_server_request_error(self.h, int(rid)).decode()

    def fields(self, rid):
        Nx, Ny = self._meta[rid]
        nm = self.info(rid)["markers"]
        rho = np.zeros(Nx * Ny)
        u = np.zeros(2 * Nx * Ny)
        mp = np.zeros(max(2 * nm, 1))
        L.check(L.load().life_server_fields(self.h, int(rid), L.dptr(rho), L.dptr(u), L.dptr(mp)))
        return rho.reshape(Nx, Ny), u.reshape(Nx, Ny, 2), mp[:2 * nm].reshape(nm, 2)

    def release(self, rid):
        L.check(L.load().life_server_release(self.h, int(rid)))

    def stats(self):
        out = np.zeros(7)
        L.check(L.load().life_server_stats(self.h, L.dptr(out)))
        keys = ["site_updates", "steps", "busy_s", "graph_captures", "peak_running", "peak_bytes", "early_steps"]
        return dict(zip(keys, out.tolist()))
