"""The solver path at a size where launches write more than the 256 MB
Infinity Cache holds, so the kernels take their non-temporal store variants
(the fluxes: 15 fields; the z sweeps' four-field gradient launches): one RHS
and one step at 210 x 208 x 206 (9.0 M points) against the oracle on the same
inputs (dQ and the primitives 1e-12, the sensor-fed coefficients 1e-9, the
stepped state 1e-12)."""
import numpy as np
import pytest

import cases
import helpers as Hh
from test_gpu_batch import _state_err
from test_gpu_solver import _compare_rhs

pytestmark = pytest.mark.gpu
TWO_PI = 2 * np.pi


@pytest.fixture(scope="module")
def ctx():
    import compactles_hip as ch
    return ch.Context(0)


def test_rhs_and_step_above_cache_size(ctx):
    n, L = (210, 208, 206), (TWO_PI,) * 3
    N = n[0] * n[1] * n[2]
    assert N * 8 * 15 > 256 * 2 ** 20 and N * 8 * 4 > 256 * 2 ** 20   # fluxes and z-sweep gradients
    X, Y, Z = cases.grid(n, L)
    q = cases.conserved(cases.smooth_prim(X, Y, Z))
    del X, Y, Z
    os_ = cases.oracle_solver(n, L, q, mu0=1e-3, art=True)
    gs = cases.gpu_solver(ctx, n, L, q, mu0=1e-3, art=True)
    del q
    os_.compute_rhs(False)
    gs.compute_rhs(False)
    _compare_rhs(gs, os_, n)
    r1, _ = os_.max_rate()
    r2, _ = gs.max_rate()
    assert abs(r1 - r2) <= 1e-13 * r1
    dt = 0.5 / r1
    os_.step(dt, True)
    gs.step(dt, True)
    Qg = gs.allocate_state()
    gs.download(Qg)
    err = _state_err([Hh.interior(Qg[..., c], n) for c in range(os_.n_cons)],
                     [os_.interior(os_.Q[..., c]) for c in range(os_.n_cons)])
    assert err <= 1e-12
