"""Shared case builders: the same physical setup for the GPU solver (through
the compactles_hip frontend) and for the oracle (padded-layout restatement)."""
import numpy as np

import oracle as O

PER = (0, 0)


def tgv_prim(X, Y, Z, p0=71.43, amp=1.0):
    """README.md:33-44 quick-start Taylor-Green initial state."""
    u = amp * np.sin(X) * np.cos(Y)
    v = -amp * np.cos(X) * np.sin(Y)
    return u, v, 0 * X, np.full_like(X, p0), np.ones_like(X)


def smooth_prim(X, Y, Z):
    """A smooth fully three-dimensional periodic state (density, velocity and
    pressure all vary), used by the RHS parity checks."""
    u = 0.3 * np.sin(X) * np.cos(Y) + 0.1 * np.cos(Z)
    v = -0.3 * np.cos(X) * np.sin(Y) + 0.05 * np.sin(2 * Z)
    w = 0.2 * np.sin(X + Z) * np.cos(Y)
    p = 1.0 + 0.1 * np.cos(X) * np.cos(Z) + 0.05 * np.sin(Y)
    rho = 1.0 + 0.2 * np.sin(Y) * np.cos(X + Z)
    return u, v, w, p, rho


def conserved(prim, species=((1.0, 1.4),), Y=None):
    """problem.jl:76-90 conserved_from_prim on arrays (rho, p given)."""
    u, v, w, p, rho = prim
    Y = Y if Y is not None else [np.ones_like(u)]
    Rm = sum(y * R for y, (R, g) in zip(Y, species))
    cvm = sum(y * (R / (g - 1)) for y, (R, g) in zip(Y, species))
    T = p / (rho * Rm)
    ke = 0.5 * (u * u + v * v + w * w)
    return [rho * y for y in Y] + [rho * u, rho * v, rho * w, rho * (cvm * T + ke)]


def oracle_solver(n, L, q, **kw):
    d = O.make_desc(n, L, **kw)
    s = O.Solver(d)
    Q = s.Q
    sl = tuple(slice(s.halo[k], s.halo[k] + n[k]) for k in range(3))
    for c, qc in enumerate(q):
        Q[sl + (c,)] = qc
    return s


def gpu_solver(ctx, n, L, q, species=((1.0, 1.4),), mu0=0.0, art=True, cfl=0.5, bcs=None, deriv="lele_d1_6",
               precision="fp64", beta_sensor="strain", bc_funs=None, filter_interval=1, origin=(0.0, 0.0, 0.0)):
    import compactles_hip as ch
    eos = ch.IdealMixture([ch.IdealSpecies(f"s{k}", R, g) for k, (R, g) in enumerate(species)])
    code = {0: ch.PeriodicBC, 1: ch.SlipWallBC}
    bcs = bcs or ((0, 0),) * 3
    bco = []
    for k in range(3):
        pair = []
        for s in range(2):
            if bc_funs and (k, s) in bc_funs:
                pair.append(ch.DirichletBC(bc_funs[(k, s)]))
            else:
                pair.append(code[bcs[k][s]]())
        bco.append(tuple(pair))
    prob = ch.Problem(domain=tuple((origin[k], origin[k] + L[k]) for k in range(3)), bcs=tuple(bco),
                      ic=lambda X, Y, Z: None, eos=eos, transport=ch.Transport(mu0=mu0))
    sch = {"lele_d1_6": ch.lele_d1_6, "lele_d1_10": ch.lele_d1_10, "pade_d1_4": ch.pade_d1_4}[deriv]()
    num = ch.Numerics(n_global=n, deriv=sch, art=ch.ArtParams(enabled=art, beta_sensor=beta_sensor), cfl=cfl,
                      precision=precision, filter_interval=filter_interval)
    s = ch.Solver(ctx, prob, num)
    Q = s.allocate_state()
    sl = tuple(slice(s.halo[k], s.halo[k] + n[k]) for k in range(3))
    for c, qc in enumerate(q):
        Q[sl + (c,)] = qc
    s.upload(Q)
    return s


def grid(n, L, periodic=(True, True, True), origin=(0.0, 0.0, 0.0)):
    xs = []
    for k in range(3):
        if n[k] == 1:
            xs.append(np.array([origin[k]]))
        else:
            h = L[k] / (n[k] if periodic[k] else n[k] - 1)
            xs.append(origin[k] + np.arange(n[k]) * h)
    return np.meshgrid(*xs, indexing="ij")
