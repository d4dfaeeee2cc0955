"""The allocation-free epsilon-matrix entry point
(life_ibm_epsilon_matrix_noalloc) against life_ibm_epsilon_matrix, both
through the C ABI on the same grid: equal bit for bit on two calls in a row,
the second with a smaller system (the shared device array still holds the
first call's larger matrix), and a clear error past its size limit."""
import numpy as np
import pytest

from life_amd import config as cfg
from life_amd import lib as L
from life_amd.case import Case

pytestmark = pytest.mark.gpu


def test_noalloc_epsilon_matrix_equals_allocating_form(gpu_lib):
    # cylinder + flexible plate, markers moved by a few coupled steps
    p, g = cfg.flexible_fsi(Nx=200, Ny=62, D=12.0, centre=(30.0, 30.0), L=42.0, h=2.4, n_el=6, u_mean=0.05,
                            Re=60.0, ramp_steps=0.0, uni_epsilon=0, central_moments=1, stiffness=200.0)
    c = Case(p, g, precision=L.LIFE_FP64)
    c.step(4)
    gr = c.grid
    n = c.n_markers
    gr.n_markers = n
    calls = [(0, n), (5, n // 3)]  # the second smaller, overlapping the first's rows
    for first, m in calls:
        A = gr.epsilon_matrix(first, m)
        B = gr.epsilon_matrix_noalloc(first, m)
        print("epsilon system [%d, %d): %d non-zeros, max |A| %.3g" % (first, first + m, np.count_nonzero(A),
                                                                      np.abs(A).max()), flush=True)
        assert np.count_nonzero(A) >= m
        assert np.array_equal(A, B)
    c.close()


def test_noalloc_epsilon_matrix_limit(gpu_lib):
    """Six filaments of 201 markers: the whole set (1206) is past the
    1024-marker limit and is refused; one filament's system is served."""
    p, g = cfg.canopy(Nx=1024, Ny=320, n_filaments=6, L=200.0, n_el=50, min_spacing=60.0, x_range=(100, 700),
                      ramp_steps=0.0)
    c = Case(p, g, precision=L.LIFE_FP64)
    gr = c.grid
    n = c.n_markers
    assert n > 1024
    gr.n_markers = n
    with pytest.raises(L.LifeError, match="allocation-free limit of 1024"):
        gr.epsilon_matrix_noalloc(0, n)
    first = c.body_info(2)["first"]
    m = c.body_info(2)["n"]
    assert np.array_equal(gr.epsilon_matrix(first, m), gr.epsilon_matrix_noalloc(first, m))
    c.close()
