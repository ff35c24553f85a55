"""Workloads of the measurements and examples: the Taylor-Green initial
state of the reference quick start, the two-dimensional He/CO2 shock /
interface case with NSCBC faces, and seeded batches of one-dimensional
two-gas shock tubes, as conserved arrays and frontend objects.  Nothing here
touches the device until a solver or batch is made from it."""
import numpy as np

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


def grid(n, L, periodic=(True, True, True), origin=(0.0, 0.0, 0.0)):
    xs = []
    for k in range(3):
        if n[k] == 1:
            xs.append(np.array([origin[k]]))
        else:
            h = L[k] / (n[k] if periodic[k] else n[k] - 1)
            xs.append(origin[k] + np.arange(n[k]) * h)
    return np.meshgrid(*xs, indexing="ij")


# --------------------------------------------------------------------------
# Two-dimensional shock / perturbed-interface case with NSCBC faces (the shape
# of docs/case_studies/shock_tube_2d.jl:37-131): He (light) and CO2 (heavy)
# ideal gases, R_k = R_u / M_k with R_u = 8.314462618 J/(mol K) and M in
# g/mol (kJ/(kg K)); gamma 5/3 and 1.289.  A Mach-1.5 shock in He from the
# normal-shock relations, frame velocity U, inflow at low x holding the shocked
# state, characteristic outflow at high x relaxing to the undisturbed pressure,
# periodic y.
# --------------------------------------------------------------------------
HE_CO2 = ((8.314462618 / 4.002602, 5.0 / 3.0), (8.314462618 / 44.0095, 1.289))


def shock_interface(n, Lx=2.0, Ly=1.0, MS=1.5, U=0.3):
    (RL, g), (RH, _) = HE_CO2
    p1, T1 = 1.0, 1.0
    rho1 = p1 / (RL * T1)
    c1 = np.sqrt(g * p1 / rho1)
    p2 = p1 * (2 * g * MS ** 2 - (g - 1)) / (g + 1)
    rho2 = rho1 * ((g + 1) * MS ** 2) / ((g - 1) * MS ** 2 + 2)
    u2 = c1 * 2 / (g + 1) * (MS ** 2 - 1) / MS
    T2 = p2 / (rho2 * RL)
    L = (Lx, Ly, Ly / n[1])
    X, Yg, Z = grid(n, L, periodic=(False, True, True))
    hx = Lx / (n[0] - 1)
    xs, xi0, amp = 0.15 * Lx, 0.25 * Lx, 0.05 * Ly
    s = 0.5 * (1 + np.tanh((X - xs) / (2 * hx)))                       # 0 shocked, 1 undisturbed
    th = 0.5 * (1 + np.tanh((X - (xi0 + amp * np.cos(2 * np.pi * Yg / Ly))) / (3 * hx)))   # 0 He, 1 CO2
    Ys = [1 - th, th]
    u = (1 - s) * u2 + U
    p = (1 - s) * p2 + s * p1
    T = (1 - s) * T2 + s * T1
    Rm = Ys[0] * RL + Ys[1] * RH
    rho = p / (Rm * T)
    q = conserved((u, 0 * X, 0 * X, p, rho), species=HE_CO2, Y=Ys)
    nscbc = {(0, 0): {"u": (u2 + U, 0.0, 0.0), "T": T2, "Y": (1.0, 0.0), "eta_u": 0.28, "eta_T": 0.28,
                      "eta_t": 0.28, "eta_Y": 0.28},
             (0, 1): {"pinf": p1, "sigma": 0.25, "beta_t": -1.0}}
    return q, L, nscbc


# --------------------------------------------------------------------------
# Batches of one-dimensional two-gas shock tubes (the shape of
# test/cases.jl:84-107): He at (p_L, rho_L) on the left, CO2 at (1, 1) on the
# right, diaphragm at x = 0.5 smeared over 2h, Dirichlet ends holding the
# initial end states; N_i ~ U{nmin..nmax}, p_L ~ U[2, 20], rho_L ~ U[0.1, 10]
# from one seeded generator.
# --------------------------------------------------------------------------


def tube_specs(count, seed=20261202, nmin=128, nmax=2048):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        n = int(rng.integers(nmin, nmax + 1))
        pl = float(rng.uniform(2.0, 20.0))
        rl = float(rng.uniform(0.1, 10.0))
        out.append({"n": n, "pL": pl, "rhoL": rl, "pR": 1.0, "rhoR": 1.0})
    return out


def tube_state(spec, eos=None):
    """(q arrays over the nodes, (left, right) Dirichlet conserved states, L, h).
    eos: a frontend Nasa9Mixture of two species (None: the HE_CO2 ideal gases)."""
    n = spec["n"]
    L = 1.0
    h = L / (n - 1)
    x = np.arange(n) * h
    th = 0.5 * (1 + np.tanh((x - 0.5) / (2 * h)))
    Y = [1 - th, th]
    p = (1 - th) * spec["pL"] + th * spec["pR"]
    rho = (1 - th) * spec["rhoL"] + th * spec["rhoR"]
    z = 0 * x
    if eos is not None:
        import compactles_hip as ch
        q = ch.conserved_from_prim(eos, ch.Prim(u=(z, z, z), p=p, rho=rho, Y=tuple(Y)))
        qL = [float(v) for v in ch.conserved_from_prim(eos, ch.Prim(p=spec["pL"], rho=spec["rhoL"], Y=(1.0, 0.0)))]
        qR = [float(v) for v in ch.conserved_from_prim(eos, ch.Prim(p=spec["pR"], rho=spec["rhoR"], Y=(0.0, 1.0)))]
        return [np.asarray(a).reshape(n, 1, 1) for a in q], (qL, qR), L, h
    q = conserved((z, z, z, p, rho), species=HE_CO2, Y=Y)
    one = np.ones(1)
    qL = [float(v[0]) for v in conserved((0 * one, 0 * one, 0 * one, spec["pL"] * one, spec["rhoL"] * one),
                                          species=HE_CO2, Y=[one, 0 * one])]
    qR = [float(v[0]) for v in conserved((0 * one, 0 * one, 0 * one, spec["pR"] * one, spec["rhoR"] * one),
                                          species=HE_CO2, Y=[0 * one, one])]
    return [a.reshape(n, 1, 1) for a in q], (qL, qR), L, h


def nscbc_bc(ch, par):
    """Frontend NSCBC object from the oracle's parameter dict."""
    if "pinf" in par:
        return ch.NSCBCOutflowBC(pinf=par["pinf"], sigma=par.get("sigma", 0.25), Lref=par.get("Lref", 0.0),
                                 beta_t=par.get("beta_t", -1.0))
    return ch.NSCBCInflowBC(u=tuple(par["u"]), T_ion=par["T"], Y=tuple(par.get("Y", (1.0,))),
                            eta_u=par.get("eta_u", 0.28), eta_T=par.get("eta_T", 0.28),
                            eta_t=par.get("eta_t", 0.28), eta_Y=par.get("eta_Y", 0.28), Lref=par.get("Lref", 0.0))


def tube_batch(ctx, specs, precision="fp64", tfinal=None, nmax=2 ** 62, art=True, cfl=0.4, eos=None):
    """A GPU Batch of the tubes, uploaded with the same initial arrays."""
    import compactles_hip as ch
    gas = eos
    if eos is None:
        eos = ch.IdealMixture([ch.IdealSpecies(f"s{k}", R, g) for k, (R, g) in enumerate(HE_CO2)])
    dummy = ch.DirichletBC(lambda x, y, z, t: None)
    prob = ch.Problem(domain=((0.0, 1.0),) * 3, bcs=((dummy, dummy), (ch.PeriodicBC(), ch.PeriodicBC()),
                                                   (ch.PeriodicBC(), ch.PeriodicBC())),
                      ic=lambda x, y, z: None, eos=eos, transport=ch.Transport(mu0=0.0))
    num = ch.Numerics(n_global=(1, 1, 1), art=ch.ArtParams(enabled=art), cfl=cfl, filter_interval=1,
                      precision=precision)
    insts, states = [], []
    for k, sp in enumerate(specs):
        q, (qL, qR), L, h = tube_state(sp, gas)
        tf = tfinal[k] if isinstance(tfinal, (list, tuple)) else (tfinal if tfinal is not None else float("inf"))
        nm = nmax[k] if isinstance(nmax, (list, tuple)) else nmax
        insts.append(ch.Instance(n=sp["n"], domain=(0.0, L), ic=None, tfinal=tf, nmax=nm,
                                 lo=ch.Prim(rho=sp["rhoL"], p=sp["pL"], Y=(1.0, 0.0)),
                                 hi=ch.Prim(rho=sp["rhoR"], p=sp["pR"], Y=(0.0, 1.0))))
        states.append(q)
    b = ch.Batch(ctx, prob, num, insts)
    for k, q in enumerate(states):
        n = specs[k]["n"]
        Q = np.zeros((n + 8, 1, 1, len(q)), order="F")
        for c in range(len(q)):
            Q[4:4 + n, 0, 0, c] = q[c][:, 0, 0]
        b.upload(k, Q)
    return b


def _he_co2():
    import compactles_hip as ch
    return ch.IdealMixture([ch.IdealSpecies(f"s{k}", R, g) for k, (R, g) in enumerate(HE_CO2)])


def tube_solver(ctx, spec, precision="fp64", art=True, cfl=0.4):
    """One tube through the plain single-domain solver, with Dirichlet ends
    holding the same end states as the batch's instance."""
    import compactles_hip as ch
    q, _, L, h = tube_state(spec)
    n = spec["n"]

    def end(rho, p, y0):
        return ch.DirichletBC(lambda X, Y, Z, t: ch.Prim(rho=rho + 0 * X, p=p + 0 * X,
                                                         Y=(y0 + 0 * X, 1.0 - y0 + 0 * X)))
    per = (ch.PeriodicBC(), ch.PeriodicBC())
    prob = ch.Problem(domain=((0.0, L), (0.0, h), (0.0, h)),
                      bcs=((end(spec["rhoL"], spec["pL"], 1.0), end(spec["rhoR"], spec["pR"], 0.0)), per, per),
                      ic=lambda X, Y, Z: None, eos=_he_co2(), transport=ch.Transport(mu0=0.0))
    num = ch.Numerics(n_global=(n, 1, 1), art=ch.ArtParams(enabled=art), cfl=cfl, precision=precision)
    s = ch.Solver(ctx, prob, num)
    Q = s.allocate_state()
    for c, qc in enumerate(q):
        Q[s.halo[0]:s.halo[0] + n, 0, 0, c] = qc[:, 0, 0]
    s.upload(Q)
    return s


def tgv_problem(ch, n, precision="fp64"):
    """The reference quick start's Taylor-Green vortex (README.md:33-44) on an
    n^3 periodic grid: C6, C8 filter every step, Cook properties on, cfl 0.5."""
    def ic(X, Y, Z):
        return ch.Prim(u=(np.sin(X) * np.cos(Y), -np.cos(X) * np.sin(Y), 0.0 * X), p=71.43, rho=1.0)

    prob = ch.Problem(domain=((0.0, 2 * np.pi),) * 3, bcs=((ch.PeriodicBC(), ch.PeriodicBC()),) * 3, ic=ic,
                      name="taylor-green")
    num = ch.Numerics(n_global=(n, n, n), precision=precision, cfl=0.5)
    return prob, num


def shock_problem(ch, n, precision="fp64"):
    """The two-dimensional He|CO2 shock / interface case of shock_interface as
    (Problem, Numerics, padded initial state)."""
    q, L, faces = shock_interface(n)
    bcs = ((nscbc_bc(ch, faces[(0, 0)]), nscbc_bc(ch, faces[(0, 1)])),
           (ch.PeriodicBC(), ch.PeriodicBC()), (ch.PeriodicBC(), ch.PeriodicBC()))
    prob = ch.Problem(domain=tuple((0.0, L[k]) for k in range(3)), bcs=bcs, ic=lambda X, Y, Z: None,
                      eos=_he_co2(), transport=ch.Transport(mu0=0.0))
    num = ch.Numerics(n_global=n, cfl=0.4, precision=precision)
    spec = ch.SolverSpec(prob, num)
    Q = spec.allocate_state()
    sl = tuple(slice(spec.halo[k], spec.halo[k] + n[k]) for k in range(3))
    for c, qc in enumerate(q):
        Q[sl + (c,)] = qc
    return prob, num, Q


def fourier_field(n, kmax, seed, slope=-5.0 / 3.0):
    """Periodic field sum_{1<=|k|<=kmax} |k|^slope cos(k.x + phi_k) on [0,2pi)^3,
    phases from numpy.random.default_rng(seed) drawn in lexicographic (kx,ky,kz)
    order over the half space kx>0 | (kx==0 & ky>0) | (kx==ky==0 & kz>0);
    synthesised with an inverse real FFT and normalised to unit RMS."""
    rng = np.random.default_rng(seed)
    F = np.zeros((n[0], n[1], n[2] // 2 + 1), dtype=complex)
    for kx in range(0, kmax + 1):
        for ky in range(-kmax, kmax + 1):
            for kz in range(-kmax, kmax + 1):
                if not (kx > 0 or (kx == 0 and ky > 0) or (kx == 0 and ky == 0 and kz > 0)):
                    continue
                kk = kx * kx + ky * ky + kz * kz
                if kk == 0 or kk > kmax * kmax:
                    continue
                amp = kk ** (slope / 2)
                ph = rng.uniform(0, 2 * np.pi)
                c = 0.5 * amp * np.exp(1j * ph)
                # place the pair (k, -k) in the rfft half-spectrum (z is the halved axis)
                if kz > 0:
                    F[kx % n[0], ky % n[1], kz] += c
                elif kz < 0:
                    F[(-kx) % n[0], (-ky) % n[1], -kz] += np.conj(c)
                else:
                    F[kx % n[0], ky % n[1], 0] += c
                    F[(-kx) % n[0], (-ky) % n[1], 0] += np.conj(c)
    f = np.fft.irfftn(F, s=n, axes=(0, 1, 2)) * (n[0] * n[1] * n[2])
    f /= np.sqrt(np.mean(f * f))
    return np.asfortranarray(f)
