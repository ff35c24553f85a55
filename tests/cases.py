"""Shared case builders: the same physical setup for the GPU solver (through
the compactles_hip frontend) and for the oracle (padded-layout restatement)."""
import numpy as np

import oracle as O
from compactles_hip.workloads import (HE_CO2, conserved, grid, nscbc_bc, shock_interface, smooth_prim,  # noqa: F401
                                      tgv_prim, tube_batch, tube_specs, tube_state)

PER = (0, 0)


def oracle_solver(n, L, q, jac=None, x=None, **kw):
    d = O.make_desc(n, L, **kw)
    s = O.Solver(d, jac, x)
    Q = s.Q
    sl = tuple(slice(s.halo[k], s.halo[k] + n[k]) for k in range(3))
    for c, qc in enumerate(q):
        Q[sl + (c,)] = qc
    return s


def nasa9_table(eos):
    """Oracle make_desc(nasa9=...) table of a frontend Nasa9Mixture."""
    return (eos.T_guess, [(s.R, [(iv.Tmin, iv.Tmax, iv.a, iv.b1) for iv in s.intervals]) for s in eos.sp])


def nasa9_oracle_kw(eos):
    """make_desc keywords of a Nasa9Mixture (species carry R; gamma unused)."""
    return {"nasa9": nasa9_table(eos), "species": tuple((s.R, 0.0) for s in eos.sp)}


def nscbc_codes(bcs, nscbc):
    """BC code table with the NSCBC faces filled in (oracle codes)."""
    out = [list(b) for b in (bcs or ((0, 0),) * 3)]
    for (k, s), par in (nscbc or {}).items():
        out[k][s] = O.BC_NSCBC_OUT if "pinf" in par else O.BC_NSCBC_IN
    return tuple(tuple(b) for b in out)


def gpu_solver(ctx, n, L, q, species=((1.0, 1.4),), mu0=0.0, art=True, cfl=0.5, bcs=None, deriv="lele_d1_6",
               precision="fp64", beta_sensor="strain", bc_funs=None, filter_interval=1, origin=(0.0, 0.0, 0.0),
               nscbc=None, chunks=None, control=None, stretch=None, eos=None, metric="cartesian"):
    """chunks: line-tile lengths (x, y, z) for this solver's plans, set in the
    table for its creation only (None: the table / rule as it stands).
    eos: a frontend EOS object (None: the IdealMixture of `species`)."""
    import compactles_hip as ch
    if eos is None:
        eos = ch.IdealMixture([ch.IdealSpecies(f"s{k}", R, g) for k, (R, g) in enumerate(species)])
    code = {0: ch.PeriodicBC, 1: ch.SlipWallBC}
    bcs = bcs or ((0, 0),) * 3
    bco = []
    for k in range(3):
        pair = []
        for s in range(2):
            if bc_funs and (k, s) in bc_funs:
                pair.append(ch.DirichletBC(bc_funs[(k, s)]))
            elif nscbc and (k, s) in nscbc:
                pair.append(nscbc_bc(ch, nscbc[(k, s)]))
            else:
                pair.append(code[bcs[k][s]]())
        bco.append(tuple(pair))
    prob = ch.Problem(domain=tuple((origin[k], origin[k] + L[k]) for k in range(3)), bcs=tuple(bco),
                      ic=lambda X, Y, Z: None, eos=eos, transport=ch.Transport(mu0=mu0), metric=metric)
    sch = {"lele_d1_6": ch.lele_d1_6, "lele_d1_10": ch.lele_d1_10, "pade_d1_4": ch.pade_d1_4}[deriv]()
    num = ch.Numerics(n_global=n, deriv=sch, art=ch.ArtParams(enabled=art, beta_sensor=beta_sensor), cfl=cfl,
                      precision=precision, filter_interval=filter_interval, control=control or ch.StepControl(),
                      stretch=stretch)
    if chunks is None:
        s = ch.Solver(ctx, prob, num)
    else:
        spec = ch.SolverSpec(prob, num)
        ch.tune.pin(spec, chunks)
        try:
            s = ch.Solver(ctx, prob, num)
        finally:
            ch.tune.unpin(spec)
    Q = s.allocate_state()
    sl = tuple(slice(s.halo[k], s.halo[k] + n[k]) for k in range(3))
    for c, qc in enumerate(q):
        Q[sl + (c,)] = qc
    s.upload(Q)
    return s


def tube_oracle(spec, precision_art=True, cfl=0.4, eos=None):
    q, (qL, qR), L, h = tube_state(spec, eos)
    n = spec["n"]
    kw = {"species": HE_CO2} if eos is None else nasa9_oracle_kw(eos)
    return oracle_solver((n, 1, 1), (L, h, h), q, art=precision_art, cfl=cfl,
                         bcs=((O.BC_DIRICHLET, O.BC_DIRICHLET), (0, 0), (0, 0)), bc_state={(0, 0): qL, (0, 1): qR},
                         **kw)


def tube_solver(ctx, spec, precision="fp64", art=True, cfl=0.4):
    """One tube through the plain single-domain solver (Dirichlet ends
    holding the same end states as the batch's instance)."""
    from compactles_hip import workloads
    return workloads.tube_solver(ctx, spec, precision, art, cfl)

