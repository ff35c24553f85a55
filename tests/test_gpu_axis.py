"""The cylindrical axis (AxisBC) on the GPU against the oracle on the same
inputs: the reference's fold known-answer tests on the device (d/dr of an odd
r exp(-4 r^2) < 5e-6 and of an even exp(-4 r^2) < 2e-5, C6 and C10,
runtests.jl:171-189, 231-249) and equal to the oracle's to 1e-12; an
axisymmetric swirling jet (RHS, grad_u, step and filter 1e-12; 20 run! steps:
dt 1e-12, state, kinetic energy and enstrophy 1e-10); freestream; the
pressure-pulse smoke case; the FP32-mixed form; refusals (an odd theta count,
the wrong face or metric)."""
import numpy as np
import pytest

import cases
import helpers as Hh
import oracle as O
from test_gpu_batch import _state_err
from test_gpu_solver import _compare_rhs

pytestmark = pytest.mark.gpu
AX_OR = ((O.BC_AXIS, O.BC_SLIPWALL), (0, 0), (0, 0))


@pytest.fixture(scope="module")
def ctx():
    import compactles_hip as ch
    return ch.Context(0)


def _grid(n, L):
    h = L[0] / (n[0] - 0.5)
    r = (np.arange(n[0]) + 0.5) * h
    z = np.arange(n[2]) * L[2] / n[2] if n[2] > 1 else np.zeros(1)
    return np.meshgrid(r, np.zeros(1), z, indexing="ij")


def _gpu(ctx, n, L, q, deriv="lele_d1_6", precision="fp64", **kw):
    import compactles_hip as ch
    prob = ch.Problem(domain=((0.0, L[0]), (0.0, L[1]), (0.0, L[2])), metric=ch.CylindricalMetric(),
                      bcs=((ch.AxisBC(), ch.SlipWallBC()), (ch.PeriodicBC(), ch.PeriodicBC()),
                           (ch.PeriodicBC(), ch.PeriodicBC())),
                      ic=lambda X, Y, Z: None, transport=ch.Transport(mu0=kw.pop("mu0", 0.0)))
    sch = {"lele_d1_6": ch.lele_d1_6, "lele_d1_10": ch.lele_d1_10}[deriv]()
    num = ch.Numerics(n_global=n, deriv=sch, art=ch.ArtParams(enabled=kw.pop("art", True)), precision=precision)
    s = ch.Solver(ctx, prob, num)
    Q = s.allocate_state()
    sl = tuple(slice(s.halo[k], s.halo[k] + n[k]) for k in range(3))
    for c, qc in enumerate(q):
        Q[sl + (c,)] = qc
    s.upload(Q)
    return s


def _pair(ctx, n, L, prim_fn, deriv="lele_d1_6", precision="fp64", art=True, mu0=0.0):
    X, Y, Z = _grid(n, L)
    q = cases.conserved(prim_fn(X, Y, Z))
    os_ = cases.oracle_solver(n, L, q, bcs=AX_OR, metric="cylindrical", deriv=deriv, art=art, mu0=mu0)
    gs = _gpu(ctx, n, L, q, deriv=deriv, precision=precision, art=art, mu0=mu0)
    return os_, gs, (X, Y, Z)


def _jet(L):
    def prim(X, Y, Z):
        k = 2 * np.pi / L[2]
        g = np.exp(-6 * X * X)
        return (0.08 * X * g * np.cos(k * Z), 0.3 * X * g, 0.2 * g + 0.05 * np.sin(k * Z) * (1 - X * X),
                1.0 + 0.05 * g * np.cos(k * Z), 1.0 + 0.03 * g * np.sin(k * Z))
    return prim


def _q_err(gs, os_, n):
    Qg = gs.allocate_state()
    gs.download(Qg)
    return _state_err([Hh.interior(Qg[..., c], n) for c in range(os_.n_cons)],
                      [os_.interior(os_.Q[..., c]) for c in range(os_.n_cons)])


@pytest.mark.parametrize("deriv", ["lele_d1_6", "lele_d1_10"])
def test_fold_known_answers_on_gpu(ctx, deriv):
    n, L = (64, 1, 12), (1.0, 1.0, 0.5)

    def prim(X, Y, Z):
        return X * np.exp(-4 * X * X), 0 * X, np.exp(-4 * X * X), 1.0 + 0 * X, 1.0 + 0 * X
    os_, gs, (X, Y, Z) = _pair(ctx, n, L, prim, deriv=deriv, art=False)
    os_.compute_rhs(False)
    gs.compute_rhs(False)
    odd, even = gs.field("grad_u[11]"), gs.field("grad_u[13]")
    assert np.abs(odd - (1 - 8 * X * X) * np.exp(-4 * X * X)).max() < 5e-6
    assert np.abs(even - (-8 * X * np.exp(-4 * X * X))).max() < 2e-5
    for nm in ("grad_u[11]", "grad_u[13]", "grad_u[22]"):
        assert Hh.relerr(gs.field(nm), os_.interior(os_.field(nm))) <= 1e-12, nm


def test_swirling_jet_rhs_step_filter(ctx):
    n, L = (40, 1, 16), (1.0, 1.0, 0.5)
    os_, gs, _ = _pair(ctx, n, L, _jet(L), mu0=2e-3)
    os_.apply_bcs()
    gs.apply_bcs()
    os_.compute_rhs(False)
    gs.compute_rhs(False)
    _compare_rhs(gs, os_, n)
    for a in (1, 2, 3):
        for b in (1, 2, 3):
            ref = os_.interior(os_.field(f"grad_u[{a}{b}]"))
            if np.abs(ref).max() > 0:
                assert Hh.relerr(gs.field(f"grad_u[{a}{b}]"), ref) <= 1e-12, (a, b)
    r1, _ = os_.max_rate()
    r2, _ = gs.max_rate()
    assert abs(r1 - r2) <= 1e-13 * r1
    dt = 0.5 / r1
    os_.step(dt, True)
    gs.step(dt, True)
    os_.filter_state()
    gs.filter_state()
    assert _q_err(gs, os_, n) <= 1e-12


def test_swirling_jet_run(ctx):
    n, L = (40, 1, 16), (1.0, 1.0, 0.5)
    os_, gs, _ = _pair(ctx, n, L, _jet(L), mu0=2e-3)
    d1 = os_.run(1e9, 20, record=20)
    d2 = gs.run(1e9, 20, record=20)
    assert np.abs(d1 - d2).max() <= 1e-12 * d1.max()
    assert _q_err(gs, os_, n) <= 1e-10
    ke1, en1 = os_.diagnostics()
    ke2, en2 = gs.diagnostics()
    assert abs(ke1 - ke2) <= 1e-10 * ke1
    assert abs(en1 - en2) <= 1e-10 * en1


def test_freestream_and_volume_weights(ctx):
    n, L = (64, 1, 12), (1.0, 1.0, 1.0)
    os_, gs, _ = _pair(ctx, n, L, lambda X, Y, Z: (0 * X, 0 * X, 1.0 + 0 * X, 1.0 + 0 * X, 1.0 + 0 * X), art=False)
    gs.apply_bcs()
    gs.compute_rhs(False)
    for c in range(5):
        assert np.abs(gs.field(f"dQ[{c + 1}]")).max() < 1e-12, c
    # kinetic energy of u_z = 1, rho = 1 over the unit cylinder per radian: 1/4 up to
    # the outer wall's O(h^2) edge term (the folded edge takes a full weight)
    ke_g, _ = gs.diagnostics()
    ke_o, _ = os_.diagnostics()
    h = 1.0 / (64 - 0.5)
    assert abs(ke_g - ke_o) <= 1e-13 * ke_o
    assert 0 <= ke_g - 0.25 < 0.5 * h * h


def test_pressure_pulse_smoke(ctx):
    n, L = (48, 1, 1), (1.0, 1.0, 1.0)
    os_, gs, _ = _pair(ctx, n, L, lambda X, Y, Z: (0 * X, 0 * X, 0 * X, 1 + np.exp(-40 * (X - 0.4) ** 2),
                                                  1.0 + 0 * X))
    d1 = os_.run(1e9, 3, record=3)
    d2 = gs.run(1e9, 3, record=3)
    assert np.isfinite(d2).all()
    assert np.abs(d1 - d2).max() <= 1e-12 * d1.max()
    assert _q_err(gs, os_, n) <= 1e-12


def test_fp32_mixed_axis(ctx):
    n, L = (40, 1, 16), (1.0, 1.0, 0.5)
    os_, gs, _ = _pair(ctx, n, L, _jet(L), precision="fp32", mu0=2e-3)
    os_.run(1e9, 10)
    gs.run(1e9, 10)
    err = _q_err(gs, os_, n)
    print(f"\nFP32-mixed axisymmetric jet 40x1x16, 10 steps vs FP64 oracle: Q {err:.2e}")
    assert err <= 1e-5


def test_axis_refusals(ctx):
    import compactles_hip as ch
    n = (32, 15, 12)
    q = cases.conserved(tuple(np.zeros(n) + v for v in (0, 0, 0, 1, 1)))
    with pytest.raises(ch.ClesError, match="even theta point count"):
        _gpu(ctx, n, (1.0, 2 * np.pi, 0.5), q)
    prob = ch.Problem(domain=((0.0, 1.0),) * 3, bcs=((ch.AxisBC(), ch.SlipWallBC()), (ch.PeriodicBC(),) * 2,
                                                     (ch.PeriodicBC(),) * 2), ic=lambda x, y, z: None)
    with pytest.raises(ch.ClesError, match="CylindricalMetric"):
        ch.Solver(ctx, prob, ch.Numerics(n_global=(32, 1, 12)))
    prob = ch.Problem(domain=((0.0, 1.0),) * 3, bcs=((ch.SlipWallBC(), ch.AxisBC()), (ch.PeriodicBC(),) * 2,
                                                     (ch.PeriodicBC(),) * 2), ic=lambda x, y, z: None,
                      metric=ch.CylindricalMetric())
    with pytest.raises(ch.ClesError, match="low end"):
        ch.Solver(ctx, prob, ch.Numerics(n_global=(32, 1, 12)))
