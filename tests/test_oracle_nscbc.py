"""Pins the oracle's NSCBC boundary corrections (CPU, no GPU needed).

  * the reference's own known-answer tests: a matched uniform stream receives
    no correction, a mismatched one a non-trivial correction
    (test/runtests.jl:423-437 outflow, :485-512 inflow);
  * the LODI mathematics, independently of the oracle's code: the change the
    correction makes to dQ on the face, mapped through the exact Jacobian of
    the conserved-from-primitive map (complex-step derivative of
    Q(rho, u, p, Y) = [rho Y_k, rho u, p cv_m/R_m + rho |u|^2 / 2]), must equal
    the primitive time-derivative deltas of the LODI relations of Poinsot &
    Lele (1992) with the Yoo & Im (2007) transverse term (outflow), built from
    one-sided derivatives computed here by a dense solve of the closed-line
    operator and spectrally exact periodic transverse derivatives.
"""
import numpy as np
import pytest

import cases
import oracle as O
from test_oracle_operators import SCHEMES, dense_operator, multiplier

BC_PER, BC_WALL, BC_DIR, BC_OUT, BC_IN = 0, 1, 2, 3, 4
TWO_SPECIES = ((1.0, 1.4), (0.25, 1.6))


def _solver(n, L, q, bcs, species=((1.0, 1.4),), nscbc=None, bc_state=None, art=False):
    d = O.make_desc(n, L, species=species, art=art, bcs=bcs, nscbc=nscbc, bc_state=bc_state)
    s = O.Solver(d)
    sl = tuple(slice(s.halo[k], s.halo[k] + n[k]) for k in range(3))
    for c, qc in enumerate(q):
        s.Q[sl + (c,)] = qc
    return s


def _uniform(n, L, u, p, rho, Y=None, species=((1.0, 1.4),)):
    X, Yg, Z = cases.grid(n, L, periodic=(False, True, True))
    one = np.ones_like(X)
    Ys = None if Y is None else [y * one for y in Y]
    return cases.conserved((u[0] * one, u[1] * one, u[2] * one, p * one, rho * one), species=species, Y=Ys)


def test_outflow_matched_stream_no_correction():
    """test/runtests.jl:423-437."""
    n, L = (32, 12, 12), (1.0, 0.4, 0.4)
    q = _uniform(n, L, (0.3, 0, 0), 1.0, 1.0)
    probe = _solver(n, L, q, ((BC_PER, BC_PER),) * 3)
    state = probe.conserved_from_prim(rho=1.0, u=(0.3, 0, 0), p=1.0)
    s = _solver(n, L, q, ((BC_DIR, BC_OUT), (0, 0), (0, 0)), nscbc={(0, 1): {"pinf": 1.0, "sigma": 0.25,
                                                                               "beta_t": -1.0}},
                bc_state={(0, 0): state})
    s.apply_bcs()
    s.compute_rhs(False)
    nx = n[0]
    m = np.abs(s.interior(s.dQ)[nx - 1]).max()
    assert m < 1e-8


def _inflow_desc(u, T, Y=(1.0,)):
    return {"u": u, "T": T, "Y": Y, "eta_u": 0.28, "eta_T": 0.28, "eta_t": 0.28, "eta_Y": 0.28}


def test_inflow_matched_and_mismatched_stream():
    """test/runtests.jl:485-512: T_ion = 1, R = 1, so rho = p = 1."""
    n, L = (32, 12, 12), (1.0, 0.4, 0.4)
    uin = (0.3, 0.0, 0.0)
    nscbc = {(0, 0): _inflow_desc(uin, 1.0), (0, 1): {"pinf": 1.0, "sigma": 0.25, "beta_t": -1.0}}
    bcs = ((BC_IN, BC_OUT), (0, 0), (0, 0))
    s = _solver(n, L, _uniform(n, L, uin, 1.0, 1.0), bcs, nscbc=nscbc)
    s.apply_bcs()
    s.compute_rhs(False)
    assert np.abs(s.interior(s.dQ)[0]).max() < 1e-8
    s2 = _solver(n, L, _uniform(n, L, (0.15, 0, 0), 1.0, 1.0), bcs, nscbc=nscbc)
    s2.apply_bcs()
    s2.compute_rhs(False)
    assert np.abs(s2.interior(s2.dQ)[0]).max() > 1e-3


# ---- LODI mathematics -----------------------------------------------------
def _q_of_prim(rho, u, v, w, p, Y, species):
    Rm = sum(y * R for y, (R, g) in zip(Y, species))
    cvm = sum(y * R / (g - 1) for y, (R, g) in zip(Y, species))
    return [rho * y for y in Y] + [rho * u, rho * v, rho * w, p * cvm / Rm + 0.5 * rho * (u * u + v * v + w * w)]


def _jvp(prim, dprim, species, h=1e-30):
    """Exact directional derivative of Q(prim) along dprim (complex step)."""
    rho, u, v, w, p, Y = prim
    drho, du, dv, dw, dp, dY = dprim
    q = _q_of_prim(rho + 1j * h * drho, u + 1j * h * du, v + 1j * h * dv, w + 1j * h * dw, p + 1j * h * dp,
                   [y + 1j * h * dy for y, dy in zip(Y, dY)], species)
    return [np.imag(x) / h for x in q]


def _dn(f, h, axis):
    """C6 derivative along a closed axis: dense solve of the assembled line."""
    q, lhs, a0, coeffs, sym, closures = SCHEMES["lele_d1_6"]
    A, B = dense_operator(q, lhs, a0, coeffs, sym, closures, f.shape[axis], h)
    M = np.linalg.solve(A, B)
    return np.moveaxis(np.tensordot(M, np.moveaxis(f, axis, 0), axes=1), 0, axis)


def _dt(f, h, axis):
    """C6 derivative along a periodic axis through its exact modified wavenumber."""
    n = f.shape[axis]
    w = 2 * np.pi * np.fft.fftfreq(n)          # k h for the integer modes of the period n h
    shp = [1] * f.ndim
    shp[axis] = -1
    mult = (multiplier("lele_d1_6", w) / h).reshape(shp)
    return np.real(np.fft.ifft(np.fft.fft(f, axis=axis) * mult, axis=axis))


def _smooth_state(n, L, species, u0):
    X, Yg, Z = cases.grid(n, L, periodic=(False, True, True))
    ky, kz = 2 * np.pi / L[1], 2 * np.pi / L[2]
    u = u0 + 0.05 * np.sin(3 * X) * np.cos(ky * Yg) + 0.02 * np.cos(kz * Z)
    v = 0.04 * np.cos(2 * X + ky * Yg) + (0.03 * np.sin(kz * Z) if n[2] > 1 else 0)
    w = (0.03 * np.cos(X) * np.sin(ky * Yg + kz * Z)) if n[2] > 1 else 0 * X
    p = 1.0 + 0.08 * np.cos(2.5 * X) * np.sin(ky * Yg) + 0.03 * np.cos(kz * Z)
    rho = 1.1 + 0.1 * np.sin(1.7 * X + ky * Yg) + 0.02 * np.sin(kz * Z)
    if len(species) == 1:
        Y = None
    else:
        th = 0.5 * (1 + np.tanh((X - 0.5 * L[0]) / (0.15 * L[0]) + 0.3 * np.sin(ky * Yg)))
        Y = [1 - th, th]
    return cases.conserved((u, v, w, p, rho), species=species, Y=Y)


def _face_delta(n, L, species, bc_face, params, side):
    """(solver with the NSCBC face, face index, dQ_nscbc - dQ_plain on the face)."""
    q = _smooth_state(n, L, species, u0=0.35 if side == 1 else 0.3)
    other = BC_WALL
    bcs_nscbc = [[0, 0], [0, 0], [0, 0]]
    bcs_plain = [[0, 0], [0, 0], [0, 0]]
    bcs_nscbc[0] = [other, other]
    bcs_plain[0] = [other, other]
    bcs_nscbc[0][side] = bc_face
    s = _solver(n, L, q, tuple(map(tuple, bcs_nscbc)), species=species, nscbc={(0, side): params})
    r = _solver(n, L, q, tuple(map(tuple, bcs_plain)), species=species)
    s.compute_rhs(False)
    r.compute_rhs(False)
    i = 0 if side == 0 else n[0] - 1
    dq = s.interior(s.dQ)[i] - r.interior(r.dQ)[i]
    return s, i, dq


def _prims(s, i, Ns):
    g = lambda nm: s.interior(s.field(nm))
    P = {nm: g(nm) for nm in ("rho", "u", "v", "w", "p", "c", "T", "cp")}
    P["Y"] = [g(f"Y[{k + 1}]") for k in range(Ns)]
    return P


@pytest.mark.parametrize("n,species", [((40, 12, 1), TWO_SPECIES), ((32, 10, 10), ((1.0, 1.4),))])
@pytest.mark.parametrize("side", [0, 1])
def test_outflow_lodi_deltas(n, species, side):
    L = (1.3, 0.5, 0.4)
    pinf, sigma = 0.97, 0.3
    s, i, dq = _face_delta(n, L, species, BC_OUT, {"pinf": pinf, "sigma": sigma, "beta_t": -1.0}, side)
    Ns = len(species)
    P = _prims(s, i, Ns)
    hx = L[0] / (n[0] - 1)
    hy, hz = L[1] / n[1], L[2] / n[2]
    f = lambda a: a[i]
    rho, un, c, p = f(P["rho"]), f(P["u"]), f(P["c"]), f(P["p"])
    dpn = f(_dn(P["p"], hx, 0))
    dun = f(_dn(P["u"], hx, 0))
    # computed incoming amplitude: L1 at the high face, L5 at the low face
    Lc = (un - c) * (dpn - rho * c * dun) if side == 1 else (un + c) * (dpn + rho * c * dun)
    sgn = -1.0 if side == 1 else 1.0
    tin = 0.0
    for ax, ut_name, h in ((1, "v", hy), (2, "w", hz)):
        if n[ax] == 1:
            continue
        ut = f(P[ut_name])
        tin = tin + ut * f(_dt(P["p"], h, ax)) + rho * c * c * f(_dt(P[ut_name], h, ax)) \
            + sgn * rho * c * ut * f(_dt(P["u"], h, ax))
    Ma = np.abs(un) / c
    dL = sigma * (1 - Ma ** 2) * c / L[0] * (p - pinf) - Ma * tin - Lc
    # LODI: rho_t gains -dL/2c^2, p_t -dL/2, u_n,t -+dL/(2 rho c); Y and u_t untouched
    zero = 0 * rho
    dprim = (-dL / (2 * c * c), -sgn * dL / (2 * rho * c), zero, zero, -dL / 2, [zero] * Ns)
    prim = (rho, un, f(P["v"]), f(P["w"]), p, [f(y) for y in P["Y"]])
    expect = _jvp(prim, dprim, species)
    for cc in range(Ns + 4):
        scale = max(np.abs(expect[cc]).max(), 1e-300)
        assert np.abs(dq[..., cc] - expect[cc]).max() <= 1e-9 * max(scale, np.abs(dq).max() * 1e-3), cc
    assert np.abs(dq).max() > 1e-4          # the correction is not trivially zero


@pytest.mark.parametrize("n,species,Yt", [((40, 12, 1), TWO_SPECIES, (0.9, 0.1)),
                                          ((32, 10, 10), ((1.0, 1.4),), (1.0,))])
@pytest.mark.parametrize("side", [0, 1])
def test_inflow_lodi_deltas(n, species, Yt, side):
    L = (1.3, 0.5, 0.4)
    ut, Tt = (0.32, 0.01, -0.02), 0.95
    eta = dict(eta_u=0.28, eta_T=0.35, eta_t=0.22, eta_Y=0.3)
    par = {"u": ut, "T": Tt, "Y": Yt, **eta}
    s, i, dq = _face_delta(n, L, species, BC_IN, par, side)
    Ns = len(species)
    P = _prims(s, i, Ns)
    hx = L[0] / (n[0] - 1)
    f = lambda a: a[i]
    rho, un, c, p, T = f(P["rho"]), f(P["u"]), f(P["c"]), f(P["p"]), f(P["T"])
    vt, wt = f(P["v"]), f(P["w"])
    dpn, drn, dun = f(_dn(P["p"], hx, 0)), f(_dn(P["rho"], hx, 0)), f(_dn(P["u"], hx, 0))
    dvn, dwn = f(_dn(P["v"], hx, 0)), f(_dn(P["w"], hx, 0))
    Lr = L[0]
    Ma = np.abs(un) / c
    K = c / Lr
    L1c = (un - c) * (dpn - rho * c * dun)
    L5c = (un + c) * (dpn + rho * c * dun)
    L2c = un * (c * c * drn - dpn)
    rel = eta["eta_u"] * rho * c * c * (1 - Ma ** 2) / Lr * (un - ut[0])
    dL1 = -rel - L1c if side == 1 else 0 * rho
    dL5 = rel - L5c if side == 0 else 0 * rho
    dL2 = eta["eta_T"] * rho * c ** 3 * (Tt - T) / (Lr * T) - L2c
    dLt1 = eta["eta_t"] * K * (vt - ut[1]) - un * dvn
    dLt2 = eta["eta_t"] * K * (wt - ut[2]) - un * dwn
    dLs = [eta["eta_Y"] * K * (f(P["Y"][k]) - Yt[k]) - un * (f(_dn(P["Y"][k], hx, 0)) if Ns > 1 else 0)
           for k in range(Ns)]
    # LODI: d1 = [L2 + (L5 + L1)/2]/c^2, d2 = (L5 + L1)/2, d3 = (L5 - L1)/(2 rho c)
    dprim = (-(dL2 + 0.5 * (dL5 + dL1)) / (c * c), -(dL5 - dL1) / (2 * rho * c), -dLt1, -dLt2,
             -0.5 * (dL5 + dL1), [-x for x in dLs])
    prim = (rho, un, vt, wt, p, [f(y) for y in P["Y"]])
    expect = _jvp(prim, dprim, species)
    big = np.abs(dq).max()
    for cc in range(Ns + 4):
        assert np.abs(dq[..., cc] - expect[cc]).max() <= 1e-9 * big, cc
    assert big > 1e-4
