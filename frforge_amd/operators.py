"""Reference-element operators for FR on Gauss-Legendre (GL) or Gauss-Lobatto (GLL) points.

Host-side set-up, run once per degree; the arrays are uploaded to the device as they are.
Follows FRForge `src/fr/Points.jl`, `src/fr/Correction.jl` and `src/fr/Operators.jl`.
All arrays are float64 numpy arrays; matrices are indexed ``D[i, j]`` as in the reference.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

_EPS = np.finfo(np.float64).eps


def gauss_legendre_nodes_weights(n: int):
    """Gauss-Legendre nodes and weights on [-1, 1] by Golub-Welsch (Points.jl:10-33).

    The eigen-decomposition uses LAPACK through numpy; the reference uses LAPACK through
    Julia.  Node bits can differ in the last ulp between the two LAPACK routines.
    """
    if n < 1:
        raise ValueError(f"n must be >= 1, got {n}")
    if n == 1:
        return np.array([0.0]), np.array([2.0])
    k = np.arange(1, n, dtype=np.float64)
    beta = k / np.sqrt(4.0 * k * k - 1.0)
    J = np.zeros((n, n))
    for i in range(n - 1):
        J[i, i + 1] = beta[i]
        J[i + 1, i] = beta[i]
    vals, vecs = np.linalg.eigh(J)
    perm = np.argsort(vals, kind="stable")
    xi = vals[perm].copy()
    V = vecs[:, perm]
    w = np.array([2.0 * V[0, i] ** 2 for i in range(n)])
    s = 0.0
    for v in w:                      # sequential sum, as Julia's sum for n < 16
        s += v
    w *= 2.0 / s
    return xi, w


def legendre_P(k: int, x):
    """P_k(x) by the three-term recurrence (Correction.jl:8-24)."""
    if k < 0:
        raise ValueError("k must be >= 0")
    x = np.asarray(x, dtype=np.float64)
    if k == 0:
        return np.ones_like(x)
    if k == 1:
        return x.copy()
    pm2, pm1 = np.ones_like(x), x.copy()
    pk = pm1
    for m in range(2, k + 1):
        pk = ((float(2 * m - 1) * x * pm1) - (float(m - 1) * pm2)) / float(m)
        pm2, pm1 = pm1, pk
    return pk


def legendre_P_prime(k: int, x: float) -> float:
    """P_k'(x) with the reference's endpoint branches (Correction.jl:32-45)."""
    if k < 0:
        raise ValueError("k must be >= 0")
    if k == 0:
        return 0.0
    if x >= 1.0 - 1000 * _EPS:
        return float(k * (k + 1)) / 2.0
    if x <= -1.0 + 1000 * _EPS:
        return (1.0 if k % 2 else -1.0) * float(k * (k + 1)) / 2.0
    pk = float(legendre_P(k, x))
    pkm1 = float(legendre_P(k - 1, x))
    return (float(k) * pkm1 - float(k) * x * pk) / (1.0 - x * x)


def gauss_lobatto_legendre_nodes_weights(n: int):
    """GLL nodes/weights by Newton on P_p' (Operators.jl:10-55)."""
    if n < 1:
        raise ValueError(f"n must be >= 1, got {n}")
    if n == 1:
        return np.array([0.0]), np.array([2.0])
    p = n - 1
    xi = np.empty(n)
    xi[0], xi[-1] = -1.0, 1.0
    for j in range(2, n):
        x = float(np.cos(np.pi * float(j - 1) / float(p)))
        for _ in range(80):
            Pp = float(legendre_P(p, x))
            dPp = legendre_P_prime(p, x)
            one_m_x2 = 1.0 - x * x
            if abs(one_m_x2) < 1000 * _EPS:
                break
            d2Pp = (2.0 * x * dPp - float(p * (p + 1)) * Pp) / one_m_x2
            if abs(d2Pp) < _EPS:
                break
            xnew = x - dPp / d2Pp
            xnew = min(max(xnew, -1.0 + 10 * _EPS), 1.0 - 10 * _EPS)
            if abs(xnew - x) < 10 * _EPS:
                x = xnew
                break
            x = xnew
        xi[j - 1] = x
    xi.sort()
    xi[0], xi[-1] = -1.0, 1.0
    w = np.empty(n)
    for i in range(n):
        Pp = float(legendre_P(p, xi[i]))
        w[i] = 2.0 / max(float(p) * float(p + 1) * Pp * Pp, _EPS)
    s = 0.0
    for v in w:
        s += v
    w *= 2.0 / s
    return xi, w


def lagrange_at(xi: np.ndarray, x: float) -> np.ndarray:
    """Lagrange basis values l_j(x) on nodes xi (Points.jl:65-76)."""
    n = len(xi)
    ell = np.ones(n)
    for j in range(n):
        for k in range(n):
            if k != j:
                ell[j] *= (x - xi[k]) / (xi[j] - xi[k])
    return ell


def differentiation_matrix(xi: np.ndarray) -> np.ndarray:
    """D[i, j] = l_j'(xi_i) (Points.jl:83-104); diagonal from sum_j D[i, j] = 0."""
    n = len(xi)
    D = np.zeros((n, n))
    for i in range(n):
        for j in range(n):
            if i != j:
                num = 1.0
                den = xi[j] - xi[i]
                for k in range(n):
                    if k != i and k != j:
                        num *= (xi[i] - xi[k]) / (xi[j] - xi[k])
                D[i, j] = num / den
        s = None
        for k in range(n):
            if k != i:
                s = D[i, k] if s is None else s + D[i, k]
        D[i, i] = -(s if s is not None else 0.0)
    return D


def g_DG_values_and_derivs(p: int, xi: np.ndarray):
    """g_L, g_R (Radau-based DG correction functions) and their derivatives (Correction.jl:57-86)."""
    if p < 0:
        raise ValueError("p must be >= 0")
    rRm1 = float(legendre_P(p + 1, -1.0)) - float(legendre_P(p, -1.0))
    rLp1 = float(legendre_P(p + 1, 1.0)) + float(legendre_P(p, 1.0))
    if abs(rRm1) < 100 * _EPS or abs(rLp1) < 100 * _EPS:
        raise ValueError(f"correction normaliser vanished for p={p}")
    n = len(xi)
    gL, gR, gLxi, gRxi = (np.zeros(n) for _ in range(4))
    for i, x in enumerate(xi):
        x = float(x)
        Pp, Pp1 = float(legendre_P(p, x)), float(legendre_P(p + 1, x))
        gL[i] = (Pp1 - Pp) / rRm1
        gR[i] = (Pp1 + Pp) / rLp1
        dPp, dPp1 = legendre_P_prime(p, x), legendre_P_prime(p + 1, x)
        gLxi[i] = (dPp1 - dPp) / rRm1
        gRxi[i] = (dPp1 + dPp) / rLp1
    return gL, gR, gLxi, gRxi


def legendre_vandermonde(xi: np.ndarray) -> np.ndarray:
    """V[j, k] = P_k(xi_j) (Correction.jl:93-100)."""
    n = len(xi)
    V = np.zeros((n, n))
    for j in range(n):
        for k in range(n):
            V[j, k] = float(legendre_P(k, xi[j]))
    return V


def g_DG_endpoints(p: int):
    """g_L, g_R at xi = -1, +1 (Correction.jl:105-116)."""
    gL, gR, _, _ = g_DG_values_and_derivs(p, np.array([-1.0, 1.0]))
    return {"gL_m1": gL[0], "gL_p1": gL[1], "gR_m1": gR[0], "gR_p1": gR[1]}


@dataclass(frozen=True)
class FROperators:
    """Reference-element operators (Operators.jl:62-75)."""

    p: int
    points: str
    xi: np.ndarray
    w: np.ndarray
    D: np.ndarray
    ell_L: np.ndarray
    ell_R: np.ndarray
    gL_xi: np.ndarray
    gR_xi: np.ndarray
    gL: np.ndarray
    gR: np.ndarray
    V_legendre: np.ndarray

    @property
    def n_points(self) -> int:
        return len(self.xi)


def build_operators(p: int, points: str = "gl") -> FROperators:
    """Build nodes, weights, D, l(+-1), g_DG derivatives and the Legendre Vandermonde
    (Operators.jl:85-100).  ``points`` is "gl" (default) or "gll"."""
    if p < 0:
        raise ValueError("p must be >= 0")
    if points not in ("gl", "gll"):
        raise ValueError(f"points must be 'gl' or 'gll', got {points}")
    n = p + 1
    if points == "gl":
        xi, w = gauss_legendre_nodes_weights(n)
    else:
        xi, w = gauss_lobatto_legendre_nodes_weights(n)
    D = differentiation_matrix(xi)
    lL = lagrange_at(xi, -1.0)
    lR = lagrange_at(xi, 1.0)
    gL, gR, gLxi, gRxi = g_DG_values_and_derivs(p, xi)
    V = legendre_vandermonde(xi)
    return FROperators(p, points, xi, w, D, lL, lR, gLxi, gRxi, gL, gR, V)


def n_points(ops: FROperators) -> int:
    return len(ops.xi)
