"""Exact Riemann solution of the ideal-gas Euler equations (Toro,
"Riemann Solvers and Numerical Methods for Fluid Dynamics", ch. 4): Newton
iteration on the pressure function for the star state, then sampling of the
self-similar wave fan.  Used by the Sod/Lax checks (the reference's own Sod
test, test/runtests.jl:1903-1951, compares against the same solution)."""
import math

import numpy as np


def _fk(p, rk, pk, ck, g):
    if p > pk:
        A = 2.0 / ((g + 1) * rk)
        B = (g - 1) / (g + 1) * pk
        return (p - pk) * math.sqrt(A / (p + B)), math.sqrt(A / (p + B)) * (1 - (p - pk) / (2 * (B + p)))
    e = (g - 1) / (2 * g)
    return 2 * ck / (g - 1) * ((p / pk) ** e - 1), (1 / (rk * ck)) * (p / pk) ** (-(g + 1) / (2 * g))


def star(left, right, g):
    rL, uL, pL = left
    rR, uR, pR = right
    cL, cR = math.sqrt(g * pL / rL), math.sqrt(g * pR / rR)
    p = max(0.5 * (pL + pR), 1e-8)
    for _ in range(200):
        fL, dL = _fk(p, rL, pL, cL, g)
        fR, dR = _fk(p, rR, pR, cR, g)
        pn = max(p - (fL + fR + uR - uL) / (dL + dR), 1e-12)
        if abs(pn - p) <= 1e-14 * 0.5 * (p + pn):
            p = pn
            break
        p = pn
    fL, _ = _fk(p, rL, pL, cL, g)
    fR, _ = _fk(p, rR, pR, cR, g)
    return p, 0.5 * (uL + uR) + 0.5 * (fR - fL), cL, cR


def sample(S, left, right, g, ps, us, cL, cR):
    rL, uL, pL = left
    rR, uR, pR = right
    G1, G6, G7 = (g - 1) / (2 * g), (g - 1) / (g + 1), (g - 1) / 2
    if S <= us:
        if ps > pL:
            if S <= uL - cL * math.sqrt((g + 1) / (2 * g) * ps / pL + G1):
                return rL, uL, pL
            return rL * ((ps / pL + G6) / (G6 * ps / pL + 1)), us, ps
        if S <= uL - cL:
            return rL, uL, pL
        if S >= us - cL * (ps / pL) ** G1:
            return rL * (ps / pL) ** (1 / g), us, ps
        u = 2 / (g + 1) * (cL + G7 * uL + S)
        c = 2 / (g + 1) * (cL + G7 * (uL - S))
        return rL * (c / cL) ** (2 / (g - 1)), u, pL * (c / cL) ** (2 * g / (g - 1))
    if ps > pR:
        if S >= uR + cR * math.sqrt((g + 1) / (2 * g) * ps / pR + G1):
            return rR, uR, pR
        return rR * ((ps / pR + G6) / (G6 * ps / pR + 1)), us, ps
    if S >= uR + cR:
        return rR, uR, pR
    if S <= us + cR * (ps / pR) ** G1:
        return rR * (ps / pR) ** (1 / g), us, ps
    u = 2 / (g + 1) * (-cR + G7 * uR + S)
    c = 2 / (g + 1) * (cR - G7 * (uR - S))
    return rR * (c / cR) ** (2 / (g - 1)), u, pR * (c / cR) ** (2 * g / (g - 1))


def profile(x, t, x0, left, right, g):
    ps, us, cL, cR = star(left, right, g)
    out = np.array([sample((xi - x0) / t, left, right, g, ps, us, cL, cR) for xi in x])
    return out[:, 0], out[:, 1], out[:, 2]
