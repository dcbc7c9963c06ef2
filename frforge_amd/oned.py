"""1D Euler host side: the reference's Euler1D, Mesh1D and SolutionState for the Sod tube
configuration, mirroring Euler.jl, Mesh1D.jl and SolutionState.jl (src/equations,
src/mesh, src/solvestate).  Layout u[j, e, c] with shape (Np, Nel, 3), Fortran order, as the
reference's Array{T,3}.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Any, Optional

import numpy as np

from .mesh import DirichletBC, PeriodicBC, TransmissiveBC
from .operators import FROperators
from .state import SchemeConfig

EPS = np.finfo(np.float64).eps


@dataclass(frozen=True)
class Euler1D:
    """Conserved (rho, rho u, E), ideal gas (Euler.jl:8-13)."""
    gamma: float = 1.4
    neq: int = 3

    def primitives_to_conserved(self, rho, u, p):
        """E = p/(gamma-1) + 0.5 rho u^2 (Euler.jl:46-52)."""
        rho = np.asarray(rho, dtype=np.float64)
        u = np.asarray(u, dtype=np.float64)
        p = np.asarray(p, dtype=np.float64)
        E = p / (self.gamma - 1.0) + 0.5 * rho * u * u
        return np.stack([rho, rho * u, E], axis=-1)

    def pressure(self, U):
        """Floored-density pressure (Euler.jl:22-27)."""
        U = np.asarray(U, dtype=np.float64)
        rs = np.maximum(U[..., 0], EPS)
        u = U[..., 1] / rs
        return (self.gamma - 1.0) * (U[..., 2] - 0.5 * rs * u * u)


@dataclass
class Mesh1D:
    """Uniform mesh on [x_left, x_right] (Mesh1D.jl:35-48): vertices k (x_right - x_left)/n,
    dx = diff(vertices), J = dx / 2."""
    x_left: float
    x_right: float
    n_elements: int
    left_bc: Any = field(default_factory=PeriodicBC)
    right_bc: Any = field(default_factory=PeriodicBC)

    def __post_init__(self):
        if self.n_elements < 1:
            raise ValueError("n_elements must be >= 1")
        if not self.x_right > self.x_left:
            raise ValueError("x_right must be > x_left")
        if isinstance(self.left_bc, PeriodicBC) != isinstance(self.right_bc, PeriodicBC):
            raise ValueError("Periodic BC must be set on both ends")
        n = self.n_elements
        a, b = float(self.x_left), float(self.x_right)
        # correctly rounded a + k (b - a) / n, as the reference's range(a, b; length = n + 1)
        k = np.arange(n + 1, dtype=np.float64)
        self.x_vertices = a + (b - a) * k / n if a != 0.0 else (b * k) / n
        self.x_vertices[0], self.x_vertices[-1] = a, b
        self.dx = np.diff(self.x_vertices)
        self.J = self.dx / 2.0

    def physical_coords(self, ops: FROperators) -> np.ndarray:
        """(Np, Nel) solution-point coordinates: (xL + xR)/2 + J xi (Mesh1D.jl:51-58)."""
        xL, xR = self.x_vertices[:-1], self.x_vertices[1:]
        return (xL + xR)[None, :] / 2.0 + self.J[None, :] * ops.xi[:, None]

    def bc_names(self):
        """Device kinds of the (left, right) ends: periodic, transmissive, or a constant Dirichlet
        state (``DirichletBC(state=U)``, Residual.jl:75-79,100-104)."""
        def one(bc):
            if isinstance(bc, PeriodicBC):
                return "periodic"
            if isinstance(bc, TransmissiveBC):
                return "transmissive"
            if isinstance(bc, DirichletBC):
                if bc.state is None:
                    raise NotImplementedError("the 1D device path takes constant Dirichlet states "
                                              "(DirichletBC(state=U))")
                return "dirichlet"
            raise NotImplementedError("the 1D device path implements periodic, transmissive and "
                                      "Dirichlet ends")
        return one(self.left_bc), one(self.right_bc)


@dataclass
class SolutionState1D:
    u: np.ndarray
    t: float
    p: int
    mesh: Mesh1D
    ops: FROperators
    scheme: SchemeConfig
    device: Optional[Any] = field(default=None, repr=False)


def allocate_state_1d(mesh: Mesh1D, ops: FROperators, neq: int = 3,
                      scheme: Optional[SchemeConfig] = None) -> SolutionState1D:
    """Zero state (SolutionState.jl:25-39)."""
    u = np.zeros((ops.p + 1, mesh.n_elements, neq), order="F")
    return SolutionState1D(u, 0.0, ops.p, mesh, ops, scheme or SchemeConfig())


def set_initial_condition_1d(state: SolutionState1D, f) -> None:
    """u[j, e, :] = f(x_je) (SolutionState.jl:59-75); f maps an array of x to (..., Neq)."""
    X = state.mesh.physical_coords(state.ops)
    state.u[...] = np.asarray(f(X), dtype=np.float64)


def sod_ic(eq: Euler1D, x, x0: float = 0.5):
    """rho, u, p = (1, 0, 1) left of x0, (0.125, 0, 0.1) right (Cases.jl:940-946)."""
    x = np.asarray(x)
    left = x < x0
    return eq.primitives_to_conserved(np.where(left, 1.0, 0.125), 0.0 * x, np.where(left, 1.0, 0.1))


def discrete_mass_1d(state: SolutionState1D, c: int = 0) -> float:
    """sum_e sum_j w_j J_e u[j, e, c] in the reference's loop order (SolutionState.jl:80-90)."""
    w, J = state.ops.w, state.mesh.J
    m = 0.0
    for e in range(state.u.shape[1]):
        for j in range(state.u.shape[0]):
            m += w[j] * J[e] * state.u[j, e, c]
    return m


def positivity_ok_1d(eq: Euler1D, u: np.ndarray, atol: float = 0.0) -> bool:
    """rho > atol, p > atol and both finite at every solution point (Euler.jl:62-77)."""
    rho = u[:, :, 0]
    p = eq.pressure(u)
    return bool(np.all((rho > atol) & (p > atol) & np.isfinite(rho) & np.isfinite(p)))


def solution_extrema(state: SolutionState1D, c: int = 0):
    """(min, max) of component c over all solution points (Metrics.jl:64-73)."""
    v = state.u[:, :, c]
    return float(np.min(v)), float(np.max(v))


def sample_solution_1d(state: SolutionState1D, component: int = 0):
    """All solution-point samples sorted by x (Metrics.jl:95-121); component index."""
    X = state.mesh.physical_coords(state.ops)
    xs = X.reshape(-1, order="F")
    vs = state.u[:, :, component].reshape(-1, order="F")
    perm = np.argsort(xs, kind="stable")
    return xs[perm], vs[perm]
