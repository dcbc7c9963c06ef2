"""Initial conditions and seeded inputs for the 2D Euler cases (host side).

Follows FRForge `src/verification/Cases2D_Riemann.jl`.  All callables are vectorised:
they take arrays of x, y and return (..., 4) conserved states.
"""
from __future__ import annotations

import math

import numpy as np

from .equations import Euler2D, primitives_to_conserved

RIEMANN2D_CFG6 = {   # Lax-Liu configuration 6 (Cases2D_Riemann.jl:10-15), states (rho, u, v, p)
    "ne": (1.0, 0.75, -0.5, 1.0),
    "nw": (2.0, 0.75, 0.5, 1.0),
    "sw": (1.0, -0.75, 0.5, 1.0),
    "se": (3.0, -0.75, -0.5, 1.0),
}
RIEMANN2D_CFG3 = {   # Lax-Liu configuration 3 (Cases2D_Riemann.jl:20-25)
    "ne": (1.5, 0.0, 0.0, 1.5),
    "nw": (0.5323, 1.206, 0.0, 0.3),
    "sw": (0.138, 1.206, 1.206, 0.029),
    "se": (0.5323, 0.0, 1.206, 0.3),
}


def riemann2d_ic(eq: Euler2D, x, y, config="cfg6", x0: float = 0.5, y0: float = 0.5):
    """Four-quadrant piecewise-constant state (Cases2D_Riemann.jl:33-51)."""
    if isinstance(config, str):
        states = {"cfg6": RIEMANN2D_CFG6, "cfg3": RIEMANN2D_CFG3}[config]
    else:
        states = config
    x = np.asarray(x, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)
    ne = (x >= x0) & (y >= y0)
    nw = (x < x0) & (y >= y0)
    sw = (x < x0) & (y < y0)
    out = np.empty(x.shape + (4,))
    for key, mask in (("se", ~(ne | nw | sw)), ("ne", ne), ("nw", nw), ("sw", sw)):
        out[mask] = primitives_to_conserved(eq, *states[key])
    return out


def isentropic_vortex_primitives(x, y, t, gamma=1.4, beta=5.0, x0=5.0, y0=5.0, u_inf=1.0, v_inf=1.0):
    """Yee isentropic vortex primitives (Cases2D_Riemann.jl:163-186)."""
    x = np.asarray(x, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)
    xd = x - x0 - u_inf * t
    yd = y - y0 - v_inf * t
    r2 = xd * xd + yd * yd
    expf = np.exp(0.5 * (1.0 - r2))
    two_pi = 2.0 * math.pi
    du = -(beta / two_pi) * expf * yd
    dv = (beta / two_pi) * expf * xd
    T = 1.0 - (gamma - 1.0) * beta ** 2 / (8.0 * gamma * (math.pi * math.pi)) * np.exp(1.0 - r2)
    T = np.maximum(T, 1e-14)
    rho = T ** (1.0 / (gamma - 1.0))
    p = rho ** gamma
    return rho, u_inf + du, v_inf + dv, p


def vortex_ic(eq: Euler2D, t=0.0, **kw):
    """Vectorised IC / exact-solution callable for the isentropic vortex."""
    def f(x, y):
        return primitives_to_conserved(eq, *isentropic_vortex_primitives(x, y, t, gamma=eq.gamma, **kw))
    return f


class SplitMix64:
    """SplitMix64 generator (Steele, Lea, Flood 2014), used for every seeded input so the
    same seed gives the same states in any implementation."""

    MASK = (1 << 64) - 1

    def __init__(self, seed: int):
        self.state = seed & self.MASK

    def next_u64(self) -> int:
        self.state = (self.state + 0x9E3779B97F4A7C15) & self.MASK
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & self.MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & self.MASK
        return z ^ (z >> 31)

    def uniform(self, lo: float, hi: float) -> float:
        """53-bit uniform in [lo, hi)."""
        return lo + (hi - lo) * ((self.next_u64() >> 11) * (1.0 / (1 << 53)))


def seeded_quadrant_states(seed: int) -> dict:
    """Four quadrant states rho~U[0.5,2], u,v~U[-1,1], p~U[0.1,1.5], drawn in the order
    ne, nw, sw, se and (rho, u, v, p) within each quadrant."""
    g = SplitMix64(seed)
    out = {}
    for key in ("ne", "nw", "sw", "se"):
        rho = g.uniform(0.5, 2.0)
        u = g.uniform(-1.0, 1.0)
        v = g.uniform(-1.0, 1.0)
        p = g.uniform(0.1, 1.5)
        out[key] = (rho, u, v, p)
    return out


ENSEMBLE_MESHES = (64, 128, 256, 512)
ENSEMBLE_DEGREES = (2, 3, 4, 5)
ENSEMBLE_FLUXES = ("rusanov", "hllc")


def ensemble_instances(seed: int, n: int = 64, meshes=ENSEMBLE_MESHES, degrees=ENSEMBLE_DEGREES):
    """The seeded ensemble of independent 2D instances: instance k uses seed + k; even k is an
    isentropic vortex (periodic [-5,5]^2, NullCapturing), odd k a four-quadrant Riemann problem
    with quadrant states from seeded_quadrant_states(seed + k) (transmissive [0,1]^2, Persson +
    BR0).  Mesh size, degree and flux are drawn from SplitMix64(seed + k) in that order.
    Returns a list of dicts {k, seed, kind, n, p, flux}."""
    out = []
    for k in range(n):
        s = seed + k
        g = SplitMix64(s ^ 0x5DEECE66D)
        m = meshes[g.next_u64() % len(meshes)]
        p = degrees[g.next_u64() % len(degrees)]
        flux = ENSEMBLE_FLUXES[g.next_u64() % len(ENSEMBLE_FLUXES)]
        out.append({"k": k, "seed": s, "kind": "vortex" if k % 2 == 0 else "riemann", "n": m, "p": p,
                    "flux": flux})
    return out
