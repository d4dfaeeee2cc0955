"""Result files in the reference's formats (restart, VTK, force and tip
logs): layouts parsed back and compared with the case's own state, and a
restart that continues a run bit for bit (FP64), as GridClass/ObjectsClass
readRestart + writeRestart do (Grid.cpp:1072-1229, Objects.cpp:979-1386).
The reference ships no result files to compare against, so the layouts are
checked against the format as its writers define it."""
import os

import numpy as np
import pytest

from life_amd import config as cfg
from life_amd import lib as L
from life_amd.case import Case

pytestmark = pytest.mark.gpu


def fsi():
    return cfg.flexible_fsi(Nx=200, Ny=62, D=12.0, centre=(30.0, 30.0), L=42.0, h=2.4, n_el=6, u_mean=0.05,
                            Re=60.0, ramp_steps=0.0, uni_epsilon=1, central_moments=1, stiffness=200.0)


def canopy():
    p, g = cfg.canopy(Nx=320, Ny=96, n_filaments=3, L=40.0, n_el=10, min_spacing=30.0, x_range=(60, 220),
                      ramp_steps=40.0, u_in=0.05, kb_range=(1e-2, 1e-1))
    return p.replace(ux0_p=0.05), g


def state(c):
    f, rho, u, b = c.fields()
    mk = c.markers()
    out = dict(f=f, rho=rho, u=u, force_ibm=b, pos=mk["pos"], vel=mk["vel"], force=mk["force"], eps=mk["eps"],
               total=c.total_force())
    for k in range(c.n_bodies):
        if c.body_info(k)["ndof"]:
            U, Ud, Udd = c.fem(k)
            out["U%d" % k], out["Ud%d" % k], out["Udd%d" % k] = U, Ud, Udd
    return out


def womersley():
    p, g = cfg.womersley_cylinder()
    return p.replace(womersley=8.0), g


def womersley_gravity():
    """Per-site force_xy (WOMERSLEY with gravity): the restarted step takes
    its force from the restart's density."""
    p, g = womersley()
    return p.replace(gravityY=-1e-4), g


@pytest.mark.parametrize("which", ["fsi", "canopy", "womersley", "womersley_gravity"])
def test_restart_continues_bitwise(tmp_path, which):
    """Run A for n1 + n2 steps; run B for n1 steps, write the restart, and
    run C from that restart for n2 steps: C equals A bit for bit (for the
    oscillating force too: it resumes at the restart's time step)."""
    p, g = {"fsi": fsi, "canopy": canopy, "womersley": womersley, "womersley_gravity": womersley_gravity}[which]()
    n1, n2 = 9, 7
    a = Case(p, g, precision=L.LIFE_FP64)
    a.step(n1 + n2)
    sa = state(a)
    a.close()
    b = Case(p, g, precision=L.LIFE_FP64)
    b.step(n1)
    b.write_restart(tmp_path)
    b.close()
    c = Case(p, g, precision=L.LIFE_FP64)
    c.read_restart(tmp_path)
    assert c.t == n1
    c.step(n2)
    sc = state(c)
    c.close()
    for k in sa:
        assert np.array_equal(sa[k], sc[k]), k


def test_restart_layout(tmp_path):
    """Fluid.restart: t, Nx, Ny (int32), omega, Dx, Dt, Dm (float64), then
    per site in i-major order i, j (int32), rho, ux, uy, Fx_ibm, Fy_ibm and
    the nine populations (float64).  IBM.restart: body count (size_t), per
    body ID (int32), marker count (size_t) and x, y, vx, vy, Fx, Fy per
    marker.  FEM.restart: flexible body count (int32), relaxation factor,
    per body ID, element and node counts, U, Udot, Udotdot, U_n, U_nm1, R_k,
    (L, angle) per element and (pos0, pos, angle0, angle) per node."""
    p, g = fsi()
    c = Case(p, g, precision=L.LIFE_FP64)
    c.step(11)
    c.write_restart(tmp_path)
    f, rho, u, b = c.fields()
    sc = c.grid.scalars()
    raw = open(os.path.join(tmp_path, "Restart", "Fluid.restart"), "rb").read()
    hdr = np.frombuffer(raw[:12], dtype="<i4")
    assert list(hdr) == [11, p.Nx, p.Ny]
    h2 = np.frombuffer(raw[12:44], dtype="<f8")
    assert h2[0] == sc["omega"] and h2[1] == sc["Dx"] and h2[2] == sc["Dt"] and h2[3] == sc["Dm"]
    rec = np.dtype([("i", "<i4"), ("j", "<i4"), ("rho", "<f8"), ("u", "<f8", 2), ("fibm", "<f8", 2),
                    ("f", "<f8", 9)])
    sites = np.frombuffer(raw[44:], dtype=rec)
    N = p.Nx * p.Ny
    assert sites.size == N
    ids = np.arange(N)
    assert np.array_equal(sites["i"], ids // p.Ny) and np.array_equal(sites["j"], ids % p.Ny)
    assert np.array_equal(sites["rho"], rho.reshape(-1))
    assert np.array_equal(sites["u"], u.reshape(N, 2))
    assert np.array_equal(sites["fibm"], b.reshape(N, 2))
    assert np.array_equal(sites["f"], f.reshape(N, 9))

    mk = c.markers()
    raw = open(os.path.join(tmp_path, "Restart", "IBM.restart"), "rb").read()
    nb = int(np.frombuffer(raw[:8], dtype="<u8")[0])
    assert nb == c.n_bodies
    off = 8
    for k in range(nb):
        info = c.body_info(k)
        bid = int(np.frombuffer(raw[off:off + 4], dtype="<i4")[0])
        nn = int(np.frombuffer(raw[off + 4:off + 12], dtype="<u8")[0])
        off += 12
        assert bid == info["ID"] and nn == info["n"]
        vals = np.frombuffer(raw[off:off + 48 * nn], dtype="<f8").reshape(nn, 6)
        off += 48 * nn
        sl = slice(info["first"], info["first"] + nn)
        assert np.array_equal(vals[:, 0:2], mk["pos"][sl])
        assert np.array_equal(vals[:, 2:4], mk["vel"][sl])
        assert np.array_equal(vals[:, 4:6], mk["force"][sl])
    assert off == len(raw)

    raw = open(os.path.join(tmp_path, "Restart", "FEM.restart"), "rb").read()
    nf = int(np.frombuffer(raw[:4], dtype="<i4")[0])
    assert nf == 1
    relax = float(np.frombuffer(raw[4:12], dtype="<f8")[0])
    assert relax == c.coupling()["relax"]
    off = 12
    k = [q for q in range(c.n_bodies) if c.body_info(q)["ndof"]][0]
    info = c.body_info(k)
    bid = int(np.frombuffer(raw[off:off + 4], dtype="<i4")[0])
    nel, nnode = (int(v) for v in np.frombuffer(raw[off + 4:off + 20], dtype="<u8"))
    off += 20
    assert bid == info["ID"] and nnode == nel + 1 and 3 * nnode == info["ndof"]
    U, Ud, Udd = c.fem(k)
    vecs = np.frombuffer(raw[off:off + 6 * 8 * info["ndof"]], dtype="<f8").reshape(6, info["ndof"])
    off += 6 * 8 * info["ndof"]
    assert np.array_equal(vecs[0], U) and np.array_equal(vecs[1], Ud) and np.array_equal(vecs[2], Udd)
    off += 16 * nel + 48 * nnode
    assert off == len(raw)
    c.close()


def test_vtk(tmp_path):
    """Fluid.<t>.vti (ImageData, appended raw, UInt64 headers): density,
    pressure and velocity in physical units, j-major; IBM.<t>.vtp and
    FEM.<t>.vtp (PolyData): one piece per body with points and lines."""
    p, g = fsi()
    c = Case(p, g, precision=L.LIFE_FP64)
    c.step(5)
    c.write_vtk(tmp_path, fem=True)
    f, rho, u, b = c.fields()
    sc = c.grid.scalars()
    Nx, Ny = p.Nx, p.Ny
    raw = open(os.path.join(tmp_path, "VTK", "Fluid.5.vti"), "rb").read()
    head, data = raw.split(b'<AppendedData encoding="raw">\n\t\t_', 1)
    text = head.decode()
    assert text.startswith('<?xml version="1.0"?>\n<VTKFile type="ImageData" version="1.0" '
                           'byte_order="LittleEndian" header_type="UInt64">\n')
    assert '\t<ImageData WholeExtent="0 %d 0 %d 0 0" Origin="0 0 0" Spacing="%s %s %s">\n' % (
        Nx - 1, Ny - 1, "%g" % sc["Dx"], "%g" % sc["Dx"], "%g" % sc["Dx"]) in text
    nb = Nx * Ny * 8
    assert 'Name="Pressure" format="appended" offset="%d"' % (nb + 8) in text
    assert 'Name="Velocity" NumberOfComponents="3" format="appended" offset="%d"' % (2 * (nb + 8)) in text
    pos = 0
    arrays = []
    for comps in (1, 1, 3):
        size = int(np.frombuffer(data[pos:pos + 8], dtype="<u8")[0])
        assert size == comps * nb
        arrays.append(np.frombuffer(data[pos + 8:pos + 8 + size], dtype="<f8"))
        pos += 8 + size
    assert data[pos:] == b"\n\t</AppendedData>\n</VTKFile>\n"
    rho_ji = rho.reshape(Nx, Ny).T.reshape(-1)
    assert np.array_equal(arrays[0], rho_ji * sc["Drho"])
    press = p.ref_P + (rho_ji - p.rho_p / sc["Drho"]) * sc["cs2"] * sc["Dm"] / (sc["Dx"] * (sc["Dt"] * sc["Dt"]))
    assert np.array_equal(arrays[1], press)
    vel = arrays[2].reshape(Ny, Nx, 3)
    uu = u.reshape(Nx, Ny, 2)
    assert np.array_equal(vel[:, :, 0], uu[:, :, 0].T * (sc["Dx"] / sc["Dt"]))
    assert np.array_equal(vel[:, :, 1], uu[:, :, 1].T * (sc["Dx"] / sc["Dt"]))
    assert not vel[:, :, 2].any()

    mk = c.markers()
    for name, nbodies in (("IBM.5.vtp", c.n_bodies), ("FEM.5.vtp", 1)):
        raw = open(os.path.join(tmp_path, "VTK", name), "rb").read()
        head, data = raw.split(b'<AppendedData encoding="raw">\n\t\t_', 1)
        assert head.decode().count("<Piece ") == nbodies
        pos = 0
        for k in range(c.n_bodies):
            info = c.body_info(k)
            if name.startswith("FEM") and not info["ndof"]:
                continue
            nn = info["n"] if name.startswith("IBM") else info["ndof"] // 3
            size = int(np.frombuffer(data[pos:pos + 8], dtype="<u8")[0])
            assert size == 24 * nn
            pts = np.frombuffer(data[pos + 8:pos + 8 + size], dtype="<f8").reshape(nn, 3)
            pos += 8 + size
            if name.startswith("IBM"):
                assert np.array_equal(pts[:, :2], mk["pos"][info["first"]:info["first"] + nn])
            nl = nn if (name.startswith("IBM") and info["flex"] == 1) else nn - 1
            size = int(np.frombuffer(data[pos:pos + 8], dtype="<u8")[0])
            assert size == 16 * nl
            conn = np.frombuffer(data[pos + 8:pos + 8 + size], dtype="<i8").reshape(nl, 2)
            assert np.array_equal(conn[:, 0], np.arange(nl)) and np.array_equal(conn[:, 1], (np.arange(nl) + 1) % nn)
            pos += 8 + size
            size = int(np.frombuffer(data[pos:pos + 8], dtype="<u8")[0])
            offs = np.frombuffer(data[pos + 8:pos + 8 + size], dtype="<i8")
            assert np.array_equal(offs, 2 * (np.arange(nl) + 1))
            pos += 8 + size
        assert data[pos:] == b"\n\t</AppendedData>\n</VTKFile>\n"
    c.close()


def test_logs_and_restart_cleanup(tmp_path):
    """TotalForces.out / TipPositions.out / TipVelocities.out: tab-separated,
    header first, one line per call with 10 significant digits; reading a
    restart drops the lines of steps at or after the restart step
    (Objects.cpp:1512-1586)."""
    p, g = fsi()
    c = Case(p, g, precision=L.LIFE_FP64)
    dt = c.grid.scalars()["Dt"]
    expect = []
    for t in range(1, 5):
        c.step(1)
        if t == 2:
            c.write_restart(tmp_path)
        c.write_forces(tmp_path)
        c.write_tips(tmp_path)
        F = c.total_force()
        expect.append(("%d\t%.10g\t%.10g\t%.10g\t%.10g\t%.10g" % ((t, dt * t) + tuple(F)), t))
    mk = c.markers()
    c.close()
    lines = open(os.path.join(tmp_path, "TotalForces.out")).read().split("\n")
    assert lines[0] == "Timestep\tTime (s)\tFx (N)\tFy (N)\tCx\tCy"
    assert lines[1:] == [e for e, _ in expect]
    tips = open(os.path.join(tmp_path, "TipPositions.out")).read().split("\n")
    assert tips[0] == "Timestep\tTime (s)\tTipX\tTipY" and len(tips) == 5
    last = mk["pos"][-1]
    assert tips[4] == "%d\t%.10g\t%.10g\t%.10g" % (4, dt * 4, last[0], last[1])
    c2 = Case(p, g, precision=L.LIFE_FP64)
    c2.read_restart(tmp_path)
    c2.close()
    lines = open(os.path.join(tmp_path, "TotalForces.out")).read().split("\n")
    assert lines[1:] == [e for e, t in expect if t < 2]
    assert len(open(os.path.join(tmp_path, "TipVelocities.out")).read().split("\n")) == 2
