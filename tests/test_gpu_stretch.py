"""Stretched Cartesian grids on the GPU (cles_solver_create_ex) against the
oracle on the same inputs: one RHS (dQ 1e-12; sensor-fed coefficients 1e-9),
one step and filter (state 1e-12), 20 run! steps (dt history 1e-12, kinetic
energy 1e-12, enstrophy 1e-11), with the sine clustering along x and along y;
exact freestream; the FP32-mixed form against the FP64 oracle; NSCBC faces
on a stretched dimension (1e-12 / 1e-10)."""
import numpy as np
import pytest

import cases
import helpers as Hh
from test_gpu_solver import _compare_rhs

pytestmark = pytest.mark.gpu
A, XC = 0.4, 0.35


@pytest.fixture(scope="module")
def ctx():
    import compactles_hip as ch
    return ch.Context(0)


def _map(n):
    xi = np.arange(n) / (n - 1)
    x = xi - (A / (2 * np.pi)) * (np.sin(2 * np.pi * (xi - XC)) + np.sin(2 * np.pi * XC))
    return x, 1.0 * (1 - A * np.cos(2 * np.pi * (xi - XC)))


def _case(n, dim, prim_fn):
    """Stretch and slip walls along `dim`, periodic elsewhere; domain [0, 1]^3."""
    coords, jac = [], [None, None, None]
    for d in range(3):
        if d == dim:
            x, m = _map(n[d])
            jac[d] = m
        else:
            x = np.arange(n[d]) / n[d]
        coords.append(x)
    X, Y, Z = np.meshgrid(*coords, indexing="ij")
    q = cases.conserved(prim_fn(X, Y, Z))
    bcs = tuple((1, 1) if d == dim else (0, 0) for d in range(3))
    return q, jac, bcs


def _flow(X, Y, Z):
    k = 2 * np.pi
    return (0.2 * np.sin(k * Y) * np.sin(np.pi * X), 0.1 * np.cos(k * Z) * np.sin(np.pi * Y),
            0.05 * np.sin(k * X + k * Y), 1.0 + 0.1 * np.cos(k * Z) * np.cos(np.pi * X),
            1.0 + 0.05 * np.sin(k * Y + k * Z))


def _pair(ctx, n, dim, prim_fn=_flow, precision="fp64", **kw):
    import compactles_hip as ch
    q, jac, bcs = _case(n, dim, prim_fn)
    st = [None, None, None]
    st[dim] = ch.sine_cluster(0.0, 1.0, XC, A)
    L = (1.0, 1.0, 1.0)
    os_ = cases.oracle_solver(n, L, q, jac=jac, bcs=bcs, **kw)
    gs = cases.gpu_solver(ctx, n, L, q, bcs=bcs, stretch=tuple(st), precision=precision, **kw)
    return os_, gs


@pytest.mark.parametrize("dim", [0, 1])
def test_rhs_step_filter_match_oracle(ctx, dim):
    n = (28, 24, 16)
    os_, gs = _pair(ctx, n, dim, mu0=1e-3, art=True)
    os_.apply_bcs()
    gs.apply_bcs()
    os_.compute_rhs(False)
    gs.compute_rhs(False)
    _compare_rhs(gs, os_, n)
    r1, _ = os_.max_rate()
    r2, _ = gs.max_rate()
    assert abs(r1 - r2) <= 1e-13 * r1
    dt = 0.5 / r1
    os_.step(dt, True)
    gs.step(dt, True)
    os_.filter_state()
    gs.filter_state()
    Qg = gs.allocate_state()
    gs.download(Qg)
    for c in range(5):
        assert Hh.relerr(Hh.interior(Qg[..., c], n), os_.interior(os_.Q[..., c])) <= 1e-12, (dim, c)


def test_run_matches_oracle(ctx):
    n = (28, 24, 16)
    os_, gs = _pair(ctx, n, 0, mu0=1e-3, art=True)
    d1 = os_.run(1e9, 20, record=20)
    d2 = gs.run(1e9, 20, record=20)
    assert np.abs(d1 - d2).max() <= 1e-12 * d1.max()
    ke1, en1 = os_.diagnostics()
    ke2, en2 = gs.diagnostics()
    assert abs(ke1 - ke2) <= 1e-12 * ke1
    assert abs(en1 - en2) <= 1e-11 * en1


def test_freestream_on_gpu(ctx):
    n = (24, 12, 12)
    _, gs = _pair(ctx, n, 0, prim_fn=lambda X, Y, Z: (0 * X, 0 * X, 0 * X, 1.0 + 0 * X, 1.0 + 0 * X), art=False)
    gs.apply_bcs()
    gs.compute_rhs(False)
    for c in range(5):
        assert np.abs(gs.field(f"dQ[{c + 1}]")).max() < 1e-12


def test_fp32_mixed_stretched(ctx):
    n = (28, 24, 16)
    os_, gs = _pair(ctx, n, 0, precision="fp32", mu0=1e-3, art=True)
    os_.run(1e9, 10)
    gs.run(1e9, 10)
    Qg = gs.allocate_state()
    gs.download(Qg)
    refs = [os_.interior(os_.Q[..., c]) for c in range(5)]
    mom = max(np.abs(refs[c]).max() for c in (1, 2, 3))
    err = max(np.abs(Hh.interior(Qg[..., c], n) - refs[c]).max() / (mom if c in (1, 2, 3) else np.abs(refs[c]).max())
              for c in range(5))
    print(f"\nFP32-mixed stretched 28x24x16, 10 steps vs FP64 oracle: Q {err:.2e}")
    assert err <= 1e-5


def _through(X, Y, Z):
    k = 2 * np.pi
    return (0.3 + 0.05 * np.sin(np.pi * X) * np.cos(k * Y), 0.04 * np.cos(k * Z) * np.sin(np.pi * X),
            0.03 * np.sin(k * Y + k * Z), 1.0 + 0.05 * np.cos(np.pi * X) * np.sin(k * Y),
            1.0 + 0.05 * np.sin(k * Y + k * Z))


@pytest.mark.parametrize("kind", ["outflow", "inflow_outflow"])
def test_nscbc_on_stretched_grid_matches_oracle(ctx, kind):
    """nscbc.jl:108-130, :257: the face derivatives take inv_h of the stretch."""
    import compactles_hip as ch
    n = (28, 16, 12)
    q, jac, bcs = _case(n, 0, _through)
    out = {"pinf": 0.97, "sigma": 0.3, "beta_t": -1.0}
    faces = {(0, 1): out}
    if kind == "inflow_outflow":
        faces[(0, 0)] = {"u": (0.32, 0.01, 0.0), "T": 1.0, "Y": (1.0,), "eta_u": 0.28, "eta_T": 0.35,
                         "eta_t": 0.22, "eta_Y": 0.3}
    L = (1.0, 1.0, 1.0)
    os_ = cases.oracle_solver(n, L, q, jac=jac, bcs=cases.nscbc_codes(bcs, faces), nscbc=faces, mu0=1e-3)
    gs = cases.gpu_solver(ctx, n, L, q, bcs=bcs, nscbc=faces, mu0=1e-3,
                          stretch=(ch.sine_cluster(0.0, 1.0, XC, A), None, None))
    os_.compute_rhs(False)
    gs.compute_rhs(False)
    _compare_rhs(gs, os_, n)
    d1 = os_.run(1e9, 15, record=15)
    d2 = gs.run(1e9, 15, record=15)
    assert np.abs(d1 - d2).max() <= 1e-12 * d1.max()
    Qg = gs.allocate_state()
    gs.download(Qg)
    for c in range(5):
        assert Hh.relerr(Hh.interior(Qg[..., c], n), os_.interior(os_.Q[..., c])) <= 1e-10, c
