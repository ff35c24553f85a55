"""The fourth-difference sensor sum (artificial.jl:67-92) on grids whose
extents are not multiples of the device tile (64 x 16 columns, 16 planes), so
partial tiles, wrapped halos (periodic) and clamped halos (closed, indices held
to [1, n]) all occur, and a collapsed dimension: the artificial coefficients it
feeds against the oracle on the same inputs (1e-9, the sensors' amplification
of rounding), dQ and the primitives 1e-12."""
import numpy as np
import pytest

import cases
from test_gpu_solver import _compare_rhs

pytestmark = pytest.mark.gpu
TWO_PI = 2 * np.pi


@pytest.fixture(scope="module")
def ctx():
    import compactles_hip as ch
    return ch.Context(0)


@pytest.mark.parametrize("n, bcs", [
    ((70, 19, 21), ((1, 1), (0, 0), (1, 1))),     # x, z walls; y periodic
    ((33, 40, 18), ((1, 1), (0, 0), (0, 0))),     # x walls; 18 planes: one full z block and 2 planes
    ((65, 17, 1), ((1, 1), (1, 1), (0, 0))),      # 2-D, z collapsed
    ((130, 1, 35), ((0, 0), (0, 0), (1, 1))),     # y collapsed, x periodic
])
def test_sensor_on_partial_tiles(ctx, n, bcs):
    periodic = tuple(bcs[k] == (0, 0) for k in range(3))
    L = (TWO_PI, 0.8 * TWO_PI, 0.7 * TWO_PI)
    X, Y, Z = cases.grid(n, L, periodic=periodic)
    q = cases.conserved(cases.smooth_prim(X, Y, Z))
    os_ = cases.oracle_solver(n, L, q, mu0=1e-3, art=True, bcs=bcs)
    gs = cases.gpu_solver(ctx, n, L, q, mu0=1e-3, art=True, bcs=bcs)
    os_.apply_bcs()
    gs.apply_bcs()
    os_.compute_rhs(False)
    gs.compute_rhs(False)
    assert np.abs(os_.interior(os_.field("mu_art"))).max() > 0
    _compare_rhs(gs, os_, n)
