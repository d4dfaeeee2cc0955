"""CPU-side checks of the C ABI library: it loads, exports every entry point
include/life_hip.h declares, and its host-only solver (the replacement for
Utils::solveLAPACK) reproduces the oracle's canonical elimination bit for
bit.  No GPU work is launched here."""
import os
import re
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "life_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(life_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_header():
    from life_amd import lib
    L = lib.load()
    names = declared_symbols()
    assert len(names) >= 25
    for n in names:
        assert hasattr(L, n), "missing export %s" % n
    # the ctypes table covers the whole header
    assert set(names) == set(lib.SIGNATURES)


def test_version_and_device_count():
    from life_amd import lib
    assert b"gfx950" in lib.load().life_version()
    assert lib.device_count() >= 0


def test_solve_lapack_matches_oracle(oracle):
    from life_amd import lib
    rng = np.random.default_rng(11)
    for n, bc, density in [(12, 0, 1.0), (40, 3, 0.2), (81, 2, 0.08), (150, 3, 0.05)]:
        A = rng.standard_normal((n, n)) * (rng.random((n, n)) < density)
        A += np.diag(rng.uniform(0.5, 2.0, n)) * rng.choice([-1.0, 1.0], n)
        b = rng.standard_normal(n)
        x = lib.solve_lapack(A, b, bc)
        y = oracle.lu_solve(A, b, bc)
        assert np.array_equal(x, y)


def test_solve_lapack_banded_matches_oracle(oracle):
    """Band matrices with pivoting (the shape of the FEM and filament
    epsilon systems)."""
    from life_amd import lib
    rng = np.random.default_rng(12)
    n = 123
    A = np.zeros((n, n))
    for k in range(-5, 6):
        A += np.diag(rng.standard_normal(n - abs(k)), k)
    b = rng.standard_normal(n)
    assert np.array_equal(lib.solve_lapack(A, b, 3), oracle.lu_solve(A, b, 3))


def test_solve_lapack_reports_singular():
    from life_amd import lib
    with pytest.raises(lib.LifeError):
        lib.solve_lapack(np.zeros((4, 4)), np.ones(4))


def test_lattice_units(oracle):
    """The named workloads are in lattice units: Dx = Dt = 1 exactly."""
    from life_amd import config as cfg
    for p, g in [cfg.lid_driven_cavity(N=32), cfg.cylinder_channel(Nx=64, Ny=32, D=8.0, centre=(16, 16))]:
        s = oracle.OracleCase(p, g).scalars()
        assert s["Dx"] == 1.0 and s["Dt"] == 1.0


def test_bench_refuses_multi_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2 and "one GPU" in r.stderr
