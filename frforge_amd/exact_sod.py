"""Exact solution of the 1D Euler Riemann problem for an ideal gas (Toro, "Riemann Solvers and
Numerical Methods for Fluid Dynamics", ch. 4): Newton iteration on the star pressure, then
sampling of the self-similar wave pattern.  The reference evaluates the Sod tube with it
(ExactSod.jl:1-168) for the L1 error and the excess-dissipation mask of run_sod.
"""
from __future__ import annotations

import math
from dataclasses import dataclass


@dataclass(frozen=True)
class RiemannProblem1D:
    rho_l: float = 1.0
    u_l: float = 0.0
    p_l: float = 1.0
    rho_r: float = 0.125
    u_r: float = 0.0
    p_r: float = 0.1
    gamma: float = 1.4
    x0: float = 0.5


SOD = RiemannProblem1D()


def _f(p, rho, pk, ck, g):
    """Toro (4.6)-(4.7): the velocity jump across one wave, and its derivative."""
    if p > pk:   # shock
        a = 2.0 / ((g + 1.0) * rho)
        b = (g - 1.0) / (g + 1.0) * pk
        q = math.sqrt(a / (p + b))
        return (p - pk) * q, q * (1.0 - 0.5 * (p - pk) / (p + b))
    r = p / pk   # rarefaction
    return (2.0 * ck / (g - 1.0)) * (r ** ((g - 1.0) / (2.0 * g)) - 1.0), (1.0 / (rho * ck)) * r ** (-(g + 1.0) / (2.0 * g))


def star_state(pr: RiemannProblem1D):
    """(p*, u*) from Newton on f_L(p) + f_R(p) + u_R - u_L = 0 to full double precision."""
    g = pr.gamma
    cl = math.sqrt(g * pr.p_l / pr.rho_l)
    cr = math.sqrt(g * pr.p_r / pr.rho_r)
    p = max(0.5 * (pr.p_l + pr.p_r), 1e-8)
    for _ in range(100):
        fl, dl = _f(p, pr.rho_l, pr.p_l, cl, g)
        fr, dr = _f(p, pr.rho_r, pr.p_r, cr, g)
        step = (fl + fr + pr.u_r - pr.u_l) / (dl + dr)
        p_new = max(p - step, 1e-14)
        if abs(p_new - p) <= 1e-16 * p:
            p = p_new
            break
        p = p_new
    fl, _ = _f(p, pr.rho_l, pr.p_l, cl, g)
    fr, _ = _f(p, pr.rho_r, pr.p_r, cr, g)
    return p, 0.5 * (pr.u_l + pr.u_r) + 0.5 * (fr - fl)


def sample(pr: RiemannProblem1D, x: float, t: float):
    """Primitives (rho, u, p) at (x, t) (Toro section 4.5)."""
    if t <= 0.0:
        return (pr.rho_l, pr.u_l, pr.p_l) if x < pr.x0 else (pr.rho_r, pr.u_r, pr.p_r)
    g = pr.gamma
    ps, us = star_state(pr)
    s = (x - pr.x0) / t
    if s < us:   # left of the contact
        rho, u, p = pr.rho_l, pr.u_l, pr.p_l
        c = math.sqrt(g * p / rho)
        if ps > p:   # left shock
            sl = u - c * math.sqrt((g + 1.0) / (2.0 * g) * ps / p + (g - 1.0) / (2.0 * g))
            if s < sl:
                return rho, u, p
            rs = rho * (ps / p + (g - 1.0) / (g + 1.0)) / ((g - 1.0) / (g + 1.0) * ps / p + 1.0)
            return rs, us, ps
        head = u - c
        tail = us - c * (ps / p) ** ((g - 1.0) / (2.0 * g))
        if s < head:
            return rho, u, p
        if s < tail:   # inside the left fan
            uf = 2.0 / (g + 1.0) * (c + (g - 1.0) / 2.0 * u + s)
            cf = 2.0 / (g + 1.0) * (c + (g - 1.0) / 2.0 * (u - s))
            rf = rho * (cf / c) ** (2.0 / (g - 1.0))
            return rf, uf, p * (rf / rho) ** g
        return rho * (ps / p) ** (1.0 / g), us, ps
    rho, u, p = pr.rho_r, pr.u_r, pr.p_r
    c = math.sqrt(g * p / rho)
    if ps > p:   # right shock
        sr = u + c * math.sqrt((g + 1.0) / (2.0 * g) * ps / p + (g - 1.0) / (2.0 * g))
        if s > sr:
            return rho, u, p
        rs = rho * (ps / p + (g - 1.0) / (g + 1.0)) / ((g - 1.0) / (g + 1.0) * ps / p + 1.0)
        return rs, us, ps
    head = u + c
    tail = us + c * (ps / p) ** ((g - 1.0) / (2.0 * g))
    if s > head:
        return rho, u, p
    if s > tail:   # inside the right fan
        uf = 2.0 / (g + 1.0) * (-c + (g - 1.0) / 2.0 * u + s)
        cf = 2.0 / (g + 1.0) * (c - (g - 1.0) / 2.0 * (u - s))
        rf = rho * (cf / c) ** (2.0 / (g - 1.0))
        return rf, uf, p * (rf / rho) ** g
    return rho * (ps / p) ** (1.0 / g), us, ps
