"""Result files against a second, independent encoder run on the oracle.

The library writes its restart and VTK files from device state.  Here the
same files are encoded a second time, in Python, from the oracle's state
after the same steps, following the reference's writers field by field:

* Fluid.restart, GridClass::writeRestart (Grid.cpp:1163-1229);
* IBM.restart and FEM.restart, ObjectsClass::writeRestart
  (Objects.cpp:1205-1355);
* Fluid.<t>.vti, GridClass::writeVTK (Grid.cpp:789-900);
* IBM.<t>.vtp and FEM.<t>.vtp, ObjectsClass::writeVTK (Objects.cpp:709-855).

An encoder is a list of segments: raw bytes (headers, XML text, integer
counts and ids, connectivity), compared byte for byte, and float64 arrays,
compared at 1e-12 relative to the largest magnitude of their field.  The
library's file must have exactly the encoder's length.  Neither side reads
the other's files, so a layout or unit mistake in the library's writer (an
order, a count, a missing field, a wrong conversion) fails here even when
its own reader agrees with it."""
import os
import struct

import numpy as np
import pytest

from life_amd import config as cfg
from life_amd import lib as L
from life_amd.case import Case
from oracle import oracle as O

pytestmark = pytest.mark.gpu

TOL = 1e-12


class Enc:
    """Segments of an expected file: ('raw', bytes) or ('f8', array)."""

    def __init__(self):
        self.seg = []

    def raw(self, b):
        if self.seg and self.seg[-1][0] == "raw":
            self.seg[-1] = ("raw", self.seg[-1][1] + b)
        else:
            self.seg.append(("raw", b))

    def text(self, s):
        self.raw(s.encode())

    def i4(self, *v):
        self.raw(struct.pack("<%di" % len(v), *v))

    def u8(self, *v):
        self.raw(struct.pack("<%dQ" % len(v), *v))

    def i8(self, a):
        self.raw(np.asarray(a, dtype="<i8").tobytes())

    def f8(self, a):
        self.seg.append(("f8", np.ascontiguousarray(a, dtype=np.float64).reshape(-1)))

    def length(self):
        return sum(len(x) if k == "raw" else 8 * x.size for k, x in self.seg)


def check_file(path, enc, what):
    data = open(path, "rb").read()
    assert len(data) == enc.length(), (what, len(data), enc.length())
    pos = 0
    worst = 0.0
    for k, x in enc.seg:
        if k == "raw":
            got = data[pos:pos + len(x)]
            assert got == x, (what, pos, got[:80], x[:80])
            pos += len(x)
        else:
            got = np.frombuffer(data[pos:pos + 8 * x.size], dtype="<f8")
            scale = max(float(np.max(np.abs(x))) if x.size else 0.0, 1e-300)
            err = float(np.max(np.abs(got - x))) / scale if x.size else 0.0
            worst = max(worst, err)
            assert err <= TOL, (what, pos, err)
            pos += 8 * x.size
    return worst


# ------------------------------------------------------------------ encoders

def enc_fluid_restart(oc):
    """Grid.cpp:1163-1229: t, Nx, Ny (int), omega, Dx, Dt, Dm (double), then
    for i, for j: i, j (int), rho, ux, uy, Fx_ibm, Fy_ibm, f_0..f_8."""
    s = oc.scalars()
    f, rho, u, fibm = oc.grid()
    e = Enc()
    e.i4(oc.t, oc.Nx, oc.Ny)
    e.f8([s["omega"], s["Dx"], s["Dt"], s["Dm"]])
    for i in range(oc.Nx):
        # one column of sites: (i, j) ints then 14 doubles per site
        for j in range(oc.Ny):
            e.i4(i, j)
            e.f8(np.concatenate(([rho[i, j]], u[i, j], fibm[i, j], f[i, j])))
    return e


def enc_ibm_restart(oc):
    """Objects.cpp:1207-1262: body count (size_t), per body ID (int), node
    count (size_t), per node pos x, y, vel x, y, force x, y."""
    mk = oc.markers()
    e = Enc()
    e.u8(oc.n_bodies)
    for b in range(oc.n_bodies):
        info = oc.body_info(b)
        _, idx = oc.body_nodes(b)
        e.i4(info["ID"])
        e.u8(len(idx))
        e.f8(np.concatenate((mk["pos"][idx], mk["vel"][idx], mk["force"][idx]), axis=1))
    return e


def enc_fem_restart(oc):
    """Objects.cpp:1265-1353: flexible body count (int), relax (double),
    per flexible body ID (int), element and node counts (size_t), U, Udot,
    Udotdot, U_n, U_nm1, R_k, then (L, angle) per element and (pos0 x, y,
    pos x, y, angle0, angle) per node."""
    flex = [b for b in range(oc.n_bodies) if oc.body_info(b)["ndof"]]
    e = Enc()
    e.i4(len(flex))
    e.f8([oc.coupling()["relax"]])
    for b in flex:
        U, Ud, Udd = oc.fem(b)
        st = oc.fem_state(b)
        e.i4(oc.body_info(b)["ID"])
        e.u8(st["elements"].shape[0], st["nodes"].shape[0])
        e.f8(np.concatenate((U, Ud, Udd, st["U_n"], st["U_nm1"], st["R_k"], st["elements"].reshape(-1),
                             st["nodes"].reshape(-1))))
    return e


def g(x):
    """std::ostream's default rendering of a double (6 significant digits)."""
    return "%g" % x


def enc_fluid_vtk(oc, p):
    """Grid.cpp:789-900: ImageData with Density, Pressure and Velocity
    appended raw, each behind a UInt64 byte count, j outer and i inner;
    physical units."""
    s = oc.scalars()
    _, rho, u, _ = oc.grid()
    Nx, Ny = oc.Nx, oc.Ny
    nb = Nx * Ny * 8
    e = Enc()
    e.text('<?xml version="1.0"?>\n<VTKFile type="ImageData" version="1.0" byte_order="LittleEndian" '
           'header_type="UInt64">\n')
    e.text('\t<ImageData WholeExtent="0 %d 0 %d 0 0" Origin="0 0 0" Spacing="%s %s %s">\n'
           % (Nx - 1, Ny - 1, g(s["Dx"]), g(s["Dx"]), g(s["Dx"])))
    e.text('\t\t<Piece Extent="0 %d 0 %d 0 0">\n\t\t\t<PointData>\n' % (Nx - 1, Ny - 1))
    e.text('\t\t\t\t<DataArray type="Float64" Name="Density" format="appended" offset="0"/>\n')
    e.text('\t\t\t\t<DataArray type="Float64" Name="Pressure" format="appended" offset="%d"/>\n' % (nb + 8))
    e.text('\t\t\t\t<DataArray type="Float64" Name="Velocity" NumberOfComponents="3" format="appended" '
           'offset="%d"/>\n' % (2 * (nb + 8)))
    e.text('\t\t\t</PointData>\n\t\t</Piece>\n\t</ImageData>\n\t<AppendedData encoding="raw">\n\t\t_')
    r = rho.T  # [j, i]
    e.u8(nb)
    e.f8(r * s["Drho"])
    e.u8(nb)
    e.f8(p.ref_P + (r - p.rho_p / s["Drho"]) * s["cs2"] * s["Dm"] / (s["Dx"] * (s["Dt"] * s["Dt"])))
    e.u8(3 * nb)
    vel = np.zeros((Ny, Nx, 3))
    vel[..., 0] = u[..., 0].T * (s["Dx"] / s["Dt"])
    vel[..., 1] = u[..., 1].T * (s["Dx"] / s["Dt"])
    e.f8(vel)
    e.text("\n\t</AppendedData>\n</VTKFile>\n")
    return e


def enc_body_vtk(oc, fem):
    """Objects.cpp:709-855: PolyData, one piece per body (FEM: flexible
    bodies only, their beam nodes), points, then line connectivity and
    offsets; closed polylines for circles."""
    mk = oc.markers()
    pieces = []
    for b in range(oc.n_bodies):
        info = oc.body_info(b)
        kind, idx = oc.body_nodes(b)
        if fem:
            if not info["ndof"]:
                continue
            pts = oc.fem_state(b)["nodes"][:, 2:4]
        else:
            pts = mk["pos"][idx]
        nn = pts.shape[0]
        pieces.append((pts, nn, nn if kind == 0 else nn - 1))
    e = Enc()
    e.text('<?xml version="1.0"?>\n<VTKFile type="PolyData" version="1.0" byte_order="LittleEndian" '
           'header_type="UInt64">\n\t<PolyData>\n')
    off = 0
    for pts, nn, nl in pieces:
        e.text('\t\t<Piece NumberOfPoints="%d" NumberOfVerts="0" NumberOfLines="%d" NumberOfStrips="0" '
               'NumberOfPolys="0">\n' % (nn, nl))
        e.text('\t\t\t<Points>\n\t\t\t\t<DataArray type="Float64" Name="Points" NumberOfComponents="3" '
               'format="appended" offset="%d"/>\n\t\t\t</Points>\n' % off)
        off += 24 * nn + 8
        e.text('\t\t\t<Lines>\n\t\t\t\t<DataArray type="Int64" Name="connectivity" format="appended" '
               'offset="%d"/>\n' % off)
        off += 16 * nl + 8
        e.text('\t\t\t\t<DataArray type="Int64" Name="offsets" format="appended" offset="%d"/>\n'
               '\t\t\t</Lines>\n\t\t</Piece>\n' % off)
        off += 8 * nl + 8
    e.text('\t</PolyData>\n\t<AppendedData encoding="raw">\n\t\t_')
    for pts, nn, nl in pieces:
        e.u8(24 * nn)
        xyz = np.zeros((nn, 3))
        xyz[:, :2] = pts
        e.f8(xyz)
        n = np.arange(nl)
        e.u8(16 * nl)
        e.i8(np.stack((n, (n + 1) % nn), axis=1))
        e.u8(8 * nl)
        e.i8(2 * (n + 1))
    e.text("\n\t</AppendedData>\n</VTKFile>\n")
    return e


# ------------------------------------------------------------------ cases

def fsi():
    """Rigid circle (closed polyline) + flexible plate, central moments."""
    return cfg.flexible_fsi(Nx=200, Ny=62, D=12.0, centre=(30.0, 30.0), L=42.0, h=2.4, n_el=6, u_mean=0.05,
                            Re=60.0, ramp_steps=0.0, uni_epsilon=1, central_moments=1, stiffness=200.0)


def canopy():
    """Three flexible filaments (open polylines), BGK, convective outlet."""
    p, g_ = cfg.canopy(Nx=320, Ny=96, n_filaments=3, L=40.0, n_el=10, min_spacing=30.0, x_range=(60, 220),
                       ramp_steps=40.0, u_in=0.05, kb_range=(1e-2, 1e-1))
    return p.replace(ux0_p=0.05), g_


CASES = {"fsi": (fsi, 13), "canopy": (canopy, 9)}


@pytest.mark.parametrize("name", sorted(CASES))
def test_files_equal_independent_encoding_of_oracle(oracle, gpu_lib, tmp_path, name):
    build, steps = CASES[name]
    p, geom = build()
    O.set_threads(8)
    oc = O.OracleCase(p, geom)
    oc.step(steps)
    gc = Case(p, geom, precision=L.LIFE_FP64)
    gc.step(steps)
    res = str(tmp_path / "Results")
    gc.write_restart(res)
    gc.write_vtk(res, fem=True)
    gc.close()
    worst = {}
    worst["Fluid.restart"] = check_file(os.path.join(res, "Restart", "Fluid.restart"), enc_fluid_restart(oc),
                                        "Fluid.restart")
    worst["IBM.restart"] = check_file(os.path.join(res, "Restart", "IBM.restart"), enc_ibm_restart(oc),
                                      "IBM.restart")
    worst["FEM.restart"] = check_file(os.path.join(res, "Restart", "FEM.restart"), enc_fem_restart(oc),
                                      "FEM.restart")
    vtk = os.path.join(res, "VTK")
    worst["Fluid.vti"] = check_file(os.path.join(vtk, "Fluid.%d.vti" % steps), enc_fluid_vtk(oc, p), "Fluid.vti")
    worst["IBM.vtp"] = check_file(os.path.join(vtk, "IBM.%d.vtp" % steps), enc_body_vtk(oc, False), "IBM.vtp")
    worst["FEM.vtp"] = check_file(os.path.join(vtk, "FEM.%d.vtp" % steps), enc_body_vtk(oc, True), "FEM.vtp")
    print("%s after %d steps, worst relative difference per file: %s" % (
        name, steps, " ".join("%s %.2g" % kv for kv in worst.items())), flush=True)
    oc.close()
