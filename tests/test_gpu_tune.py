"""The line-tile autotuner on the GPU and parity of the line kernels its
choices select, against the oracle at the FP64 bar: long x lines (more than
64 chunks, the block x-sweep kernel) through the operator plugin, a tuned
solver's run (kinetic energy and enstrophy), and pinned short and long chunk
lengths through one RHS, step and filter on grids with closed and periodic x
lines."""
import numpy as np
import pytest

import cases
import helpers as Hh
from test_gpu_operators import SCHEMES, _gpu_apply, _oracle_apply
from test_gpu_solver import _compare_rhs

pytestmark = pytest.mark.gpu
TWO_PI = 2 * np.pi


@pytest.fixture(scope="module")
def ctx():
    import compactles_hip as ch
    return ch.Context(0)


@pytest.fixture(autouse=True)
def _empty_table():
    import compactles_hip as ch
    ch.tune.clear()
    yield
    ch.tune.clear()


@pytest.mark.parametrize("name", SCHEMES)
@pytest.mark.parametrize("periodic", [True, False])
def test_long_x_lines_match_oracle(ctx, name, periodic):
    n = (1100, 6, 5)   # 128 chunks per line: the block x sweep with its LDS reduced solve
    f = Hh.random_field(n, seed=5)
    h = 0.37 / n[0]
    err = Hh.relerr(_gpu_apply(ctx, name, 1, n, periodic, h, f), _oracle_apply(name, 1, n, periodic, h, f))
    assert err <= 1e-12, (name, periodic, err)


def _tgv_spec(ch, n):
    prob = ch.Problem(domain=((0.0, TWO_PI),) * 3, bcs=((ch.PeriodicBC(), ch.PeriodicBC()),) * 3,
                      ic=lambda X, Y, Z: None)
    return ch.SolverSpec(prob, ch.Numerics(n_global=n))


def test_tuned_solver_matches_oracle(ctx):
    import compactles_hip as ch
    n, L = (32, 32, 32), (TWO_PI,) * 3
    spec = _tgv_spec(ch, n)
    chunks, ms = ch.tune.tune(ctx, spec, reps=3)
    print(f"\ntuned line tiles for TGV 32^3: {chunks}, ms per repetition {ms}")
    assert chunks[0] in (16, 8, 4) and chunks[1] in (32, 16, 8, 4) and chunks[2] in (32, 16, 8, 4)
    assert all(t > 0 for t in ms)
    assert ch.tune.lookup(spec) == (chunks, (True, True, True))
    X, Y, Z = cases.grid(n, L)
    q = cases.conserved(cases.tgv_prim(X, Y, Z))
    os_ = cases.oracle_solver(n, L, q, art=True)
    gs = cases.gpu_solver(ctx, n, L, q, art=True)
    d1 = os_.run(1e9, 20, record=20)
    d2 = gs.run(1e9, 20, record=20)
    assert np.abs(d1 - d2).max() <= 1e-12 * d1.max()
    ke1, en1 = os_.diagnostics()
    ke2, en2 = gs.diagnostics()
    assert abs(ke1 - ke2) <= 1e-12 * ke1
    assert abs(en1 - en2) <= 1e-11 * en1


# the last two: x lines of more than 64 chunks of at most 8 points (the block
# x sweep at capacity 8), closed and periodic
@pytest.mark.parametrize("n,per,chunks", [((40, 24, 36), (False, True, True), (4, 4, 4)),
                                          ((40, 24, 36), (False, True, True), (8, 16, 8)),
                                          ((40, 24, 36), (False, True, True), (16, 32, 32)),
                                          ((1030, 12, 10), (False, True, True), (8, 8, 8)),
                                          ((1024, 12, 10), (True, True, True), (8, 8, 8))])
def test_pinned_chunks_match_oracle(ctx, n, per, chunks):
    # x spacing 0.025 on every grid: a field resolved much more finely has
    # fourth differences near rounding level, and the artificial-property
    # sensors then amplify last-bit differences beyond the FP64 bar
    L = (0.025 * n[0], 0.7, 0.9)
    X, Y, Z = cases.grid(n, L, periodic=per)
    ky, kz = 2 * np.pi / L[1], 2 * np.pi / L[2]
    prim = (0.2 * np.sin(3 * X) * np.cos(ky * Y) + 0.1 * np.sin(2 * X) ** 2, 0.05 * np.cos(2 * X + ky * Y),
            0.04 * np.sin(X + kz * Z), 1.0 + 0.1 * np.cos(X) * np.cos(kz * Z), 1.0 + 0.05 * np.sin(ky * Y + X))
    q = cases.conserved(prim)
    bcs = ((0, 0) if per[0] else (1, 1), (0, 0), (0, 0))
    os_ = cases.oracle_solver(n, L, q, mu0=1e-3, art=True, bcs=bcs)
    gs = cases.gpu_solver(ctx, n, L, q, mu0=1e-3, art=True, bcs=bcs, chunks=chunks)
    os_.apply_bcs()
    gs.apply_bcs()
    os_.compute_rhs(False)
    gs.compute_rhs(False)
    _compare_rhs(gs, os_, n)
    r1, _ = os_.max_rate()
    gs.max_rate()
    dt = 0.5 / r1
    os_.step(dt, True)
    gs.step(dt, True)
    os_.filter_state()
    gs.filter_state()
    Qg = gs.allocate_state()
    gs.download(Qg)
    for c in range(5):
        assert Hh.relerr(Hh.interior(Qg[..., c], n), os_.interior(os_.Q[..., c])) <= 1e-12, (chunks, c)
