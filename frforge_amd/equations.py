"""2D compressible Euler host helpers (FRForge `src/equations/Euler2D.jl`).

These run on the host for initial conditions and diagnostics; the device kernels carry their
own copy of the point physics.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

_EPS = np.finfo(np.float64).eps


@dataclass(frozen=True)
class Euler2D:
    """Conserved U = (rho, rho u, rho v, E) with ideal-gas EOS (Euler2D.jl:8-12)."""

    gamma: float = 1.4

    @property
    def n_vars(self) -> int:
        return 4


def pressure(eq: Euler2D, U) -> np.ndarray:
    """(gamma-1)(E - rho (u^2+v^2)/2) with the rho floor (Euler2D.jl:14-19).  U[..., 4]."""
    U = np.asarray(U, dtype=np.float64)
    rho = np.maximum(U[..., 0], _EPS)
    u = U[..., 1] / rho
    v = U[..., 2] / rho
    return (eq.gamma - 1.0) * (U[..., 3] - 0.5 * rho * (u * u + v * v))


def primitives_to_conserved(eq: Euler2D, rho, u, v, p) -> np.ndarray:
    """(Euler2D.jl:21-25)."""
    rho, u, v, p = (np.asarray(a, dtype=np.float64) for a in (rho, u, v, p))
    E = p / (eq.gamma - 1.0) + 0.5 * rho * (u * u + v * v)
    return np.stack(np.broadcast_arrays(rho, rho * u, rho * v, E), axis=-1)


def conserved_to_primitives(eq: Euler2D, U):
    """(Euler2D.jl:27-33)."""
    U = np.asarray(U, dtype=np.float64)
    rho = U[..., 0]
    u = U[..., 1] / np.maximum(rho, _EPS)
    v = U[..., 2] / np.maximum(rho, _EPS)
    return rho, u, v, pressure(eq, U)


def sound_speed(eq: Euler2D, U) -> np.ndarray:
    """(Euler2D.jl:35-39)."""
    U = np.asarray(U, dtype=np.float64)
    rho = np.maximum(U[..., 0], _EPS)
    p = np.maximum(pressure(eq, U), _EPS)
    return np.sqrt(eq.gamma * p / rho)
