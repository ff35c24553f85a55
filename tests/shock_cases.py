"""The reference's one-dimensional shock-capturing validation cases
(test/cases.jl: `tube`, `lax`, `shu_osher`, `woodward`, `noh_case` with
nu = 1) and their measurements (test/validation.jl, test/references.jl), for
the oracle and, through the frontend, the GPU solver.

Each case is one-dimensional along x with y and z collapsed; its initial state
is the reference's smeared diaphragm (tanh_blend over delta = 2h unless
given).  A case is a dict: n, L, origin, q (conserved arrays over the nodes),
bcs (oracle codes), bc_state (conserved Dirichlet states), gamma, cfl, tfin, x.
"""
import numpy as np

import oracle as O


def _conserved(rho, u, p, gamma):
    z = 0 * rho
    return [rho, rho * u, z, z, p / (gamma - 1) + 0.5 * rho * u * u]


def tanh_blend(x, x0, delta):
    """problem.jl:172."""
    return 0.5 * (1 + np.tanh((x - x0) / delta))


def tube(left, right, N, tfin, L=1.0, x0=0.5, xlo=0.0, gamma=1.4, cfl=0.4, rhofun=None, delta=None):
    """test/cases.jl:75-104: Dirichlet ends holding the end states, art on,
    filter every step, mu0 = 0."""
    h = L / (N - 1)
    d = 2 * h if delta is None else delta
    x = xlo + np.arange(N) * h
    rL, uL, pL = left
    rR, uR, pR = right
    rat = (lambda xx: rR + 0 * xx) if rhofun is None else rhofun
    th = tanh_blend(x, x0, d)
    rho = (1 - th) * rL + th * rat(x)
    u = (1 - th) * uL + th * uR
    p = (1 - th) * pL + th * pR
    qL = [float(v[0]) for v in _conserved(np.array([rL]), np.array([uL]), np.array([pL]), gamma)]
    xe = np.array([xlo + L])
    qR = [float(v[0]) for v in _conserved(rat(xe), np.array([uR]), np.array([pR]), gamma)]
    return {"n": (N, 1, 1), "L": (L, h, h), "origin": (xlo, 0.0, 0.0), "x": x,
            "q": [a.reshape(N, 1, 1) for a in _conserved(rho, u, p, gamma)],
            "bcs": ((O.BC_DIRICHLET, O.BC_DIRICHLET), (0, 0), (0, 0)), "bc_state": {(0, 0): qL, (0, 1): qR},
            "gamma": gamma, "cfl": cfl, "tfin": tfin}


# test/cases.jl:110-115
LAX_L, LAX_R, LAX_T, LAX_N = (0.445, 0.698, 3.528), (0.5, 0.0, 0.571), 0.14, 400


def lax(N=LAX_N):
    return tube(LAX_L, LAX_R, N, LAX_T)


# test/cases.jl:120-127
SO_L, SO_T, SO_N = (3.857143, 2.629369, 10.333333333333333), 1.8, 800


def shu_osher(N=SO_N):
    return tube(SO_L, (1.0, 0.0, 1.0), N, SO_T, L=10.0, xlo=-5.0, x0=-4.0,
                rhofun=lambda x: 1 + 0.2 * np.sin(5 * x))


# test/cases.jl:132-152
WC_T, WC_N = 0.038, 800


def woodward(N=WC_N):
    h = 1.0 / (N - 1)
    d = 2 * h
    x = np.arange(N) * h
    p = 1000 * (1 - tanh_blend(x, 0.1, d)) + 0.01 * (tanh_blend(x, 0.1, d) - tanh_blend(x, 0.9, d)) + \
        100 * tanh_blend(x, 0.9, d)
    rho = np.ones(N)
    return {"n": (N, 1, 1), "L": (1.0, h, h), "origin": (0.0, 0.0, 0.0), "x": x,
            "q": [a.reshape(N, 1, 1) for a in _conserved(rho, 0 * rho, p, 1.4)],
            "bcs": ((1, 1), (0, 0), (0, 0)), "bc_state": None, "gamma": 1.4, "cfl": 0.3, "tfin": WC_T}


# test/cases.jl:203-212, nu = 1 (planar: slip wall at x = 0, the exact inflow at x = 1)
NOH_G, NOH_T, NOH_P0, NOH_CFL, NOH_N = 5.0 / 3.0, 0.6, 1e-4, 0.15, 400


def noh_exact(r, t, nu, g, rho0=1.0):
    """test/references.jl:106-114."""
    S = (g - 1) / 2
    if r <= S * t:
        rho = rho0 * ((g + 1) / (g - 1)) ** nu
        return rho, 0.0, (g - 1) * rho * 0.5
    return rho0 * (1 + t / r) ** (nu - 1), -1.0, 0.0


def noh1(N=NOH_N):
    """Cold gas at u = -1 onto a slip wall; the inflow face holds the exact
    pre-shock state (rho 1, u -1, p 1e-4), constant in time for nu = 1."""
    R = 1.0
    h = R / (N - 1)
    x = np.arange(N) * h
    one = np.ones(N)
    qin = [float(v[0]) for v in _conserved(np.array([1.0]), np.array([-1.0]), np.array([NOH_P0]), NOH_G)]
    return {"n": (N, 1, 1), "L": (R, R / N, R / N), "origin": (0.0, 0.0, 0.0), "x": x,
            "q": [a.reshape(N, 1, 1) for a in _conserved(one, -one, NOH_P0 * one, NOH_G)],
            "bcs": ((1, O.BC_DIRICHLET), (0, 0), (0, 0)), "bc_state": {(0, 1): qin},
            "gamma": NOH_G, "cfl": NOH_CFL, "tfin": NOH_T}


# --- running ------------------------------------------------------------------
def oracle_run(case, nmax=2_000_000):
    """The oracle's run!(...; tfinal) from the case's state; returns (x, rho,
    u, p, completed) with the primitives of the final state."""
    import cases
    s = cases.oracle_solver(case["n"], case["L"], case["q"], origin=case["origin"],
                            species=((1.0, case["gamma"]),), art=True, cfl=case["cfl"], bcs=case["bcs"],
                            bc_state=case["bc_state"])
    s.run(case["tfin"], nmax)
    t = s.time[0]
    s.max_rate()
    prof = [s.interior(s.field(f))[:, 0, 0].copy() for f in ("rho", "u", "p")]
    return (case["x"], *prof, bool(np.isfinite(t) and t >= case["tfin"] * (1 - 1e-9)))


def gpu_solver(ctx, case, precision="fp64"):
    import cases
    import compactles_hip as ch

    funs = {}
    for (k, sd), q in (case["bc_state"] or {}).items():
        rho, m, E = q[0], q[1], q[4]
        u = m / rho
        p = (case["gamma"] - 1) * (E - 0.5 * rho * u * u)
        funs[(k, sd)] = (lambda r, uu, pp: lambda X, Y, Z, t: ch.Prim(rho=r + 0 * X, u=(uu + 0 * X, 0 * X, 0 * X),
                                                                      p=pp + 0 * X))(rho, u, p)
    bcs = tuple(tuple(1 if b == O.BC_DIRICHLET else b for b in pair) for pair in case["bcs"])
    return cases.gpu_solver(ctx, case["n"], case["L"], case["q"], species=((1.0, case["gamma"]),), art=True,
                            cfl=case["cfl"], bcs=bcs, bc_funs=funs, origin=case["origin"], precision=precision)


def gpu_run(ctx, case, precision="fp64", nmax=2_000_000):
    s = gpu_solver(ctx, case, precision)
    s.run(case["tfin"], nmax)
    t = s.t
    s.max_rate()
    prof = [s.field(f)[:, 0, 0].copy() for f in ("rho", "u", "p")]
    s.close()
    return (case["x"], *prof, bool(np.isfinite(t) and t >= case["tfin"] * (1 - 1e-9)))


# --- measurements (test/references.jl, test/cases.jl) ---------------------------
def l1(a, b):
    """test/references.jl:171."""
    return float(np.abs(np.asarray(a) - np.asarray(b)).mean())


def front_position(xs, f, level):
    """test/references.jl:160-168: the outermost crossing of `level`."""
    for i in range(len(f) - 1, 0, -1):
        lo, hi = f[i - 1], f[i]
        if (lo - level) * (hi - level) <= 0 and lo != hi:
            return xs[i - 1] + (xs[i] - xs[i - 1]) * (level - lo) / (hi - lo)
    return float("nan")


def noh_metrics(xs, rho, nu):
    """test/cases.jl:324-333: plateau, wall deficit, shock position, pre-shock L1."""
    exact = 4.0 ** nu
    Rs = (NOH_G - 1) / 2 * NOH_T
    core = (xs >= 0.3 * Rs) & (xs <= 0.7 * Rs)
    pre = xs > Rs + 0.05
    ex = [noh_exact(x, NOH_T, nu, NOH_G)[0] for x in xs[pre]]
    return (float(rho[core].mean()), 1 - rho[0] / exact, front_position(xs, rho, 0.5 * exact),
            l1(rho[pre], ex))


def so_band(xs):
    """test/cases.jl:336: the wave-train window."""
    return (xs >= 0.5) & (xs <= 2.2)


def read_ref(path):
    """test/validation.jl:27-38: a stored profile as (x, rho, u, p)."""
    a = np.loadtxt(path, delimiter=",", comments="#")
    return a[:, 0], a[:, 1], a[:, 2], a[:, 3]


def interp1(xs, ys, xq):
    """test/cases.jl:53-60 (np.interp clamps outside the range the same way)."""
    return np.interp(xq, xs, ys)
