"""Structured 2D quad mesh, boundary-condition types and affine mesh metrics (host side).

Follows FRForge `src/mesh/Mesh2D.jl`, `src/mesh/BoundaryConditions.jl` and the affine branch
of `src/mesh/MeshMetrics2D.jl`.  Element index ``e = jy * nx + jx`` (0-based here; the
reference is 1-based, Mesh2D.jl:100).
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Optional

import numpy as np

from .operators import FROperators

_EPS = np.finfo(np.float64).eps


class AbstractBC:
    """Boundary-condition tag (BoundaryConditions.jl:6)."""


class PeriodicBC(AbstractBC):
    """Periodic domain end; both ends of an axis must be periodic (BoundaryConditions.jl:9)."""

    def __repr__(self):
        return "PeriodicBC()"


class TransmissiveBC(AbstractBC):
    """Zero-order extrapolation: ghost state = interior trace (BoundaryConditions.jl:12)."""

    def __repr__(self):
        return "TransmissiveBC()"


class ReflectingBC(AbstractBC):
    """Slip wall (BoundaryConditions.jl:27): the ghost state mirrors the velocity through the
    outward normal (reflect_conserved, :43-60)."""

    def __repr__(self):
        return "ReflectingBC()"


class DirichletBC(AbstractBC):
    """Prescribed state (BoundaryConditions.jl:22).  The device path takes a constant conserved
    state: ``DirichletBC(state=U)``; a ``u_func(x, y, t)`` of the reference's form is accepted
    on the host side only."""

    def __init__(self, u_func=None, state=None):
        if u_func is None and state is None:
            raise ValueError("DirichletBC needs a state or a u_func")
        self.u_func = u_func
        self.state = None if state is None else tuple(float(v) for v in state)

    def __repr__(self):
        return f"DirichletBC(state={self.state})" if self.state is not None else "DirichletBC(u_func)"


class GhostStateBC(AbstractBC):
    """User ghost state (BoundaryConditions.jl:31-33,112-130).  The device path runs the form the
    reduced double-Mach reflection uses (Cases2D_Optional.jl:72-84), built by ``split``: at a face
    point with coordinate s along the side (x on the bottom and top, y on the left and right) and
    time t, the ghost is ``U_lo`` where s < a + (b + c t) / d, and otherwise a reflecting wall
    (``hi="reflecting"``) or the state ``U_hi`` (``hi="dirichlet"``).  A general
    ``ghost(u_int, nx, ny, x, y, t)`` callable is accepted on the host side only."""

    def __init__(self, ghost=None):
        self.ghost = ghost
        self.coef = None
        self.U_lo = self.U_hi = None
        self.hi = None

    @classmethod
    def split(cls, a, b, c, d, U_lo, hi="reflecting", U_hi=None):
        if hi not in ("reflecting", "dirichlet"):
            raise ValueError("hi must be 'reflecting' or 'dirichlet'")
        if hi == "dirichlet" and U_hi is None:
            raise ValueError("a Dirichlet hi side needs U_hi")
        if d == 0:
            raise ValueError("d must be nonzero")
        bc = cls()
        bc.coef = (float(a), float(b), float(c), float(d))
        bc.U_lo = tuple(float(v) for v in U_lo)
        bc.hi = hi
        bc.U_hi = None if U_hi is None else tuple(float(v) for v in U_hi)
        return bc

    def key(self):
        return (self.coef, self.U_lo, self.hi, self.U_hi)

    def __repr__(self):
        if self.coef is None:
            return "GhostStateBC(ghost)"
        return f"GhostStateBC.split(coef={self.coef}, hi={self.hi!r})"


def _linspace(a: float, b: float, n: int) -> np.ndarray:
    """Vertices a + k (b - a) / n, k = 0..n; correctly rounded whenever the reference's
    ``range(a, b; length=n+1)`` is exact (all meshes used here: power-of-two spacings)."""
    k = np.arange(n + 1, dtype=np.float64)
    v = a + (k * (b - a)) / float(n)
    v[0], v[-1] = a, b
    return v


@dataclass
class Mesh2D:
    """``nx x ny`` affine quad mesh (Mesh2D.jl:11-91)."""

    x_left: float
    x_right: float
    y_bottom: float
    y_top: float
    nx: int
    ny: int
    left_bc: AbstractBC = field(default_factory=PeriodicBC)
    right_bc: AbstractBC = field(default_factory=PeriodicBC)
    bottom_bc: AbstractBC = field(default_factory=PeriodicBC)
    top_bc: AbstractBC = field(default_factory=PeriodicBC)
    # optional solid-element mask, True = inactive / wall interior (Mesh2D.jl:31-32,65-69)
    solid: Optional[np.ndarray] = None

    def __post_init__(self):
        if self.nx < 1 or self.ny < 1:
            raise ValueError("nx, ny must be >= 1")
        if not self.x_right > self.x_left:
            raise ValueError("x_right > x_left required")
        if not self.y_top > self.y_bottom:
            raise ValueError("y_top > y_bottom required")
        if isinstance(self.left_bc, PeriodicBC) != isinstance(self.right_bc, PeriodicBC):
            raise ValueError("x-periodic requires both left and right PeriodicBC")
        if isinstance(self.bottom_bc, PeriodicBC) != isinstance(self.top_bc, PeriodicBC):
            raise ValueError("y-periodic requires both bottom and top PeriodicBC")
        if self.solid is not None:
            m = np.asarray(self.solid, dtype=bool).ravel()
            if m.size != self.nx * self.ny:
                raise ValueError(f"solid mask length must be n_elements={self.nx * self.ny}")
            self.solid = m
        self.x_vertices = _linspace(float(self.x_left), float(self.x_right), self.nx)
        self.y_vertices = _linspace(float(self.y_bottom), float(self.y_top), self.ny)
        self.dx = np.diff(self.x_vertices)
        self.dy = np.diff(self.y_vertices)
        self.Jx = self.dx / 2.0
        self.Jy = self.dy / 2.0

    @property
    def n_elements(self) -> int:
        return self.nx * self.ny

    def element_index(self, jx: int, jy: int) -> int:
        return jy * self.nx + jx

    def is_solid(self, e: int) -> bool:
        """True if element ``e`` is marked solid (Mesh2D.jl:93-97)."""
        return self.solid is not None and bool(self.solid[e])

    def element_coords(self, e: int):
        jy, jx = divmod(e, self.nx)
        return jx, jy

    def sp_coordinates(self, ops: FROperators):
        """Physical (x, y) of every solution point, arrays of shape (Np, Np, Nel) in
        Fortran order (physical_xy, MeshMetrics2D.jl:49-58)."""
        Np, nx = ops.n_points, self.nx
        xc = (self.x_vertices[:-1] + self.x_vertices[1:]) / 2.0     # (nx,)
        yc = (self.y_vertices[:-1] + self.y_vertices[1:]) / 2.0     # (ny,)
        X1 = xc[None, :] + self.Jx[None, :] * ops.xi[:, None]        # X1[i, jx]
        Y1 = yc[None, :] + self.Jy[None, :] * ops.xi[:, None]        # Y1[j, jy]
        e = np.arange(self.n_elements)
        jx, jy = e % nx, e // nx
        X = np.broadcast_to(X1[:, None, jx], (Np, Np, len(e)))      # X[i, j, e] = X1[i, jx]
        Y = np.broadcast_to(Y1[None, :, jy], (Np, Np, len(e)))      # Y[i, j, e] = Y1[j, jy]
        return np.asfortranarray(X), np.asfortranarray(Y)


def is_periodic(bc: AbstractBC) -> bool:
    return isinstance(bc, PeriodicBC)


@dataclass
class MeshMetrics2D:
    """Metric arrays at solution points and faces (MeshMetrics2D.jl:15-39), affine branch."""

    x: np.ndarray
    y: np.ndarray
    x_xi: np.ndarray
    x_eta: np.ndarray
    y_xi: np.ndarray
    y_eta: np.ndarray
    J: np.ndarray
    h_char: np.ndarray


def build_mesh_metrics(mesh: Mesh2D, ops: FROperators) -> MeshMetrics2D:
    """Volume metrics by applying D to the SP coordinates, in the reference's summation
    order (MeshMetrics2D.jl:82-122), and affine h_char = min(dx, dy) (:206-209).

    Face normals are not stored on the host: the device library rebuilds them from the
    vertices and operators it is given.
    """
    Np = ops.n_points
    X, Y = mesh.sp_coordinates(ops)
    D = ops.D
    shape = X.shape
    xxi = np.zeros(shape, order="F")
    yxi = np.zeros(shape, order="F")
    xeta = np.zeros(shape, order="F")
    yeta = np.zeros(shape, order="F")
    for i in range(Np):
        for k in range(Np):
            xxi[i] = xxi[i] + D[i, k] * X[k]
            yxi[i] = yxi[i] + D[i, k] * Y[k]
    for j in range(Np):
        for k in range(Np):
            xeta[:, j] = xeta[:, j] + D[j, k] * X[:, k]
            yeta[:, j] = yeta[:, j] + D[j, k] * Y[:, k]
    J = xxi * yeta - xeta * yxi
    jx = np.arange(mesh.n_elements) % mesh.nx
    jy = np.arange(mesh.n_elements) // mesh.nx
    h_char = np.minimum(mesh.dx[jx], mesh.dy[jy])
    return MeshMetrics2D(X, Y, xxi, xeta, yxi, yeta, J, h_char)
