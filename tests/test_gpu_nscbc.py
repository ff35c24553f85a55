"""GPU parity of the NSCBC boundary corrections (nscbc.jl:59-306) against the
oracle on the same inputs, through the C ABI.  Bar: FP64 normwise relative
error <= 1e-12 on dQ for one RHS; <= 1e-10 on the state, kinetic energy and
the dt sequence after a short run of the shock / interface case."""
import numpy as np
import pytest

import cases
import helpers as Hh
from test_oracle_nscbc import TWO_SPECIES, _smooth_state

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    import compactles_hip as ch
    return ch.Context(0)


def _inflow(u, T, Y):
    return {"u": u, "T": T, "Y": Y, "eta_u": 0.28, "eta_T": 0.35, "eta_t": 0.22, "eta_Y": 0.3}


OUT = {"pinf": 0.97, "sigma": 0.3, "beta_t": -1.0}


def _permute(q, dim):
    """Move the closed (x) axis of a state built along x to `dim`; the
    momentum components follow the axes."""
    if dim == 0:
        return q
    perm = [0, 1, 2]
    perm[0], perm[dim] = perm[dim], perm[0]
    Ns = len(q) - 4
    out = [np.ascontiguousarray(np.transpose(a, perm)) for a in q]
    mom = out[Ns:Ns + 3]
    mom[0], mom[dim] = mom[dim], mom[0]
    return out[:Ns] + mom + out[Ns + 3:]


@pytest.mark.parametrize("case", ["2d_two_species_x", "3d_x", "3d_y_faces", "3d_z_beta_t"])
def test_rhs_nscbc_parity(ctx, case):
    if case == "2d_two_species_x":
        n, L, species, dim = (40, 12, 1), (1.3, 0.5, 0.4), TWO_SPECIES, 0
        faces = {(0, 0): _inflow((0.32, 0.01, 0.0), 0.95, (0.9, 0.1)), (0, 1): OUT}
    elif case == "3d_x":
        n, L, species, dim = (32, 10, 10), (1.3, 0.5, 0.4), ((1.0, 1.4),), 0
        faces = {(0, 0): _inflow((0.32, 0.01, -0.02), 0.95, (1.0,)), (0, 1): OUT}
    elif case == "3d_y_faces":
        n, L, species, dim = (10, 32, 12), (0.5, 1.3, 0.4), ((1.0, 1.4),), 1
        faces = {(1, 0): OUT, (1, 1): _inflow((0.01, -0.3, 0.02), 1.05, (1.0,))}
    else:
        n, L, species, dim = (12, 10, 32), (0.5, 0.4, 1.3), TWO_SPECIES, 2
        faces = {(2, 0): {"pinf": 1.02, "sigma": 0.2, "beta_t": 0.5}, (2, 1): dict(OUT, Lref=0.7)}
    nx = [n[dim], n[1] if dim != 1 else n[0], n[2] if dim != 2 else n[0]]
    Lx = [L[dim], L[1] if dim != 1 else L[0], L[2] if dim != 2 else L[0]]
    q = _permute(_smooth_state(tuple(nx), tuple(Lx), species, u0=0.3), dim)
    bcs = [(0, 0)] * 3
    bcs[dim] = (1, 1)
    codes = cases.nscbc_codes(tuple(bcs), faces)
    os_ = cases.oracle_solver(n, L, q, species=species, art=True, bcs=codes, nscbc=faces)
    gs = cases.gpu_solver(ctx, n, L, q, species=species, art=True, bcs=tuple(bcs), nscbc=faces)
    os_.compute_rhs(False)
    gs.compute_rhs(False)
    for c in range(os_.n_cons):
        ref = os_.interior(os_.dQ[..., c])
        got = gs.field(f"dQ[{c + 1}]")
        assert Hh.relerr(got, ref) <= 1e-12, (case, c, Hh.relerr(got, ref))
    # the faces really were corrected: the plane differs from a plain closed line
    plain = cases.oracle_solver(n, L, q, species=species, art=True, bcs=tuple(bcs))
    plain.compute_rhs(False)
    face = [slice(None)] * 3
    face[dim] = 0
    d = np.abs(os_.interior(os_.dQ)[tuple(face)] - plain.interior(plain.dQ)[tuple(face)]).max()
    assert d > 1e-6


def test_run_shock_interface_nscbc(ctx):
    """Short run of the two-dimensional shock / interface case (two species,
    NSCBC inflow and outflow, art on, cfl 0.4, filter every step)."""
    n = (96, 32, 1)
    q, L, faces = cases.shock_interface(n)
    bcs = ((1, 1), (0, 0), (0, 0))
    codes = cases.nscbc_codes(bcs, faces)
    os_ = cases.oracle_solver(n, L, q, species=cases.HE_CO2, art=True, cfl=0.4, bcs=codes, nscbc=faces)
    gs = cases.gpu_solver(ctx, n, L, q, species=cases.HE_CO2, art=True, cfl=0.4, bcs=bcs, nscbc=faces)
    nsteps = 40
    d1 = os_.run(1e9, nsteps, record=nsteps)
    d2 = gs.run(1e9, nsteps, record=nsteps)
    assert np.abs(d1 - d2).max() <= 1e-10 * d1.max()
    for c in range(os_.n_cons):
        ref = os_.interior(os_.Q[..., c])
        got = gs.field(f"Q[{c + 1}]")
        assert Hh.relerr(got, ref) <= 1e-10, (c, Hh.relerr(got, ref))
    ke1, _ = os_.diagnostics()
    ke2, _ = gs.diagnostics()
    assert abs(ke1 - ke2) <= 1e-10 * ke1
