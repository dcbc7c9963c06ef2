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
    # amplitude of the analytic wavy warp (Mesh2D.jl:30; make_wavy_mesh2d), or None
    wavy_amp: Optional[float] = None

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
    """Metric arrays at solution points and faces (MeshMetrics2D.jl:15-39).  The affine branch
    keeps the volume arrays only (the device rebuilds its affine tables from the vertices); curved
    metrics (``is_curved``) carry the face normals and scalings too, and the device takes them as
    they are (frf_set_curved_metrics)."""

    x: np.ndarray
    y: np.ndarray
    x_xi: np.ndarray
    x_eta: np.ndarray
    y_xi: np.ndarray
    y_eta: np.ndarray
    J: np.ndarray
    h_char: np.ndarray
    nx_W: Optional[np.ndarray] = None
    ny_W: Optional[np.ndarray] = None
    sJ_W: Optional[np.ndarray] = None
    nx_E: Optional[np.ndarray] = None
    ny_E: Optional[np.ndarray] = None
    sJ_E: Optional[np.ndarray] = None
    nx_S: Optional[np.ndarray] = None
    ny_S: Optional[np.ndarray] = None
    sJ_S: Optional[np.ndarray] = None
    nx_N: Optional[np.ndarray] = None
    ny_N: Optional[np.ndarray] = None
    sJ_N: Optional[np.ndarray] = None
    is_curved: bool = False


def build_mesh_metrics(mesh: Mesh2D, ops: FROperators) -> MeshMetrics2D:
    """Metrics by applying D to the SP coordinates, in the reference's summation order
    (MeshMetrics2D.jl:82-218): volume terms, face normals and scalings by Lagrange extrapolation
    of the volume terms, and h_char = min(dx, dy) on affine meshes.  A wavy mesh (``wavy_amp``)
    takes the warped coordinates (physical_xy, :49-57) and the curved h_char
    min(sqrt(sum w_i w_j |J|), 2 min sJ) (:191-208)."""
    Np = ops.n_points
    X, Y = mesh.sp_coordinates(ops)
    curved = mesh.wavy_amp is not None
    if curved:
        X, Y = wavy_physical(X, Y, mesh.wavy_amp)
        X, Y = np.asfortranarray(X), np.asfortranarray(Y)
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
    lL, lR = ops.ell_L, ops.ell_R
    eps = np.finfo(np.float64).eps
    fz = lambda: np.zeros((Np, mesh.n_elements))
    xW, yW, xE, yE, xS, yS, xN, yN = (fz() for _ in range(8))
    for i in range(Np):                 # extrapolation to xi = -1, +1 (row j = first index)
        xW = xW + lL[i] * xeta[i]
        yW = yW + lL[i] * yeta[i]
        xE = xE + lR[i] * xeta[i]
        yE = yE + lR[i] * yeta[i]
    for j in range(Np):                 # to eta = -1, +1 (column i)
        xS = xS + lL[j] * xxi[:, j]
        yS = yS + lL[j] * yxi[:, j]
        xN = xN + lR[j] * xxi[:, j]
        yN = yN + lR[j] * yxi[:, j]
    F = np.asfortranarray
    sE = np.maximum(np.sqrt(yE * yE + xE * xE), eps)
    sW = np.maximum(np.sqrt(yW * yW + xW * xW), eps)
    sN = np.maximum(np.sqrt(yN * yN + xN * xN), eps)
    sS = np.maximum(np.sqrt(yS * yS + xS * xS), eps)
    faces = dict(nx_E=F(yE / sE), ny_E=F(-xE / sE), sJ_E=F(sE), nx_W=F(-yW / sW), ny_W=F(xW / sW),
                 sJ_W=F(sW), nx_N=F(-yN / sN), ny_N=F(xN / sN), sJ_N=F(sN), nx_S=F(yS / sS),
                 ny_S=F(-xS / sS), sJ_S=F(sS))
    jx = np.arange(mesh.n_elements) % mesh.nx
    jy = np.arange(mesh.n_elements) // mesh.nx
    if curved:
        area = np.zeros(mesh.n_elements)
        for j in range(Np):
            for i in range(Np):
                area = area + ops.w[i] * ops.w[j] * np.abs(J[i, j])
        h_area = np.sqrt(np.maximum(area, eps))
        h_face = np.minimum.reduce([sW.min(axis=0), sE.min(axis=0), sS.min(axis=0), sN.min(axis=0)])
        h_char = np.minimum(h_area, 2.0 * h_face)
    else:
        h_char = np.minimum(mesh.dx[jx], mesh.dy[jy])
    return MeshMetrics2D(X, Y, xxi, xeta, yxi, yeta, J, h_char, is_curved=curved, **faces)


# ---------------------------------------------------------------- wavy (curved) meshes

_TWO_PI = 2.0 * np.pi   # T(2pi) in Float64
_PI = np.pi


def wavy_physical(X, Y, amp):
    """Global wavy map from computational (X, Y) to physical (x, y) (MeshMetrics2D.jl:294-303)."""
    x = X + amp * np.sin(_TWO_PI * Y) * np.sin(_PI * X)
    y = Y + amp * np.sin(_TWO_PI * X) * np.sin(_PI * Y)
    return x, y


def wavy_partials(X, Y, amp):
    """Analytic Jacobian (x_X, x_Y, y_X, y_Y) of the wavy map (MeshMetrics2D.jl:305-318), in the
    reference's operation order."""
    x_X = 1.0 + amp * np.sin(_TWO_PI * Y) * _PI * np.cos(_PI * X)
    x_Y = amp * _TWO_PI * np.cos(_TWO_PI * Y) * np.sin(_PI * X)
    y_X = amp * _TWO_PI * np.cos(_TWO_PI * X) * np.sin(_PI * Y)
    y_Y = 1.0 + amp * np.sin(_TWO_PI * X) * _PI * np.cos(_PI * Y)
    return x_X, x_Y, y_X, y_Y


def make_wavy_mesh2d(nx: int, ny: int, ops: FROperators, amp: float = 0.05, x_left=0.0,
                     x_right=1.0, y_bottom=0.0, y_top=1.0, left_bc=None, right_bc=None,
                     bottom_bc=None, top_bc=None) -> Mesh2D:
    """Structured mesh with the sinusoidal warp (MeshMetrics2D.jl:253-292); periodic by default.
    The warp lives in ``wavy_amp``: allocate_state builds the analytic metrics from it, as the
    reference's curved cases do (``allocate_state(...; wavy_amp)``)."""
    mesh = Mesh2D(x_left, x_right, y_bottom, y_top, nx, ny,
                  left_bc=left_bc or PeriodicBC(), right_bc=right_bc or PeriodicBC(),
                  bottom_bc=bottom_bc or PeriodicBC(), top_bc=top_bc or PeriodicBC())
    mesh.wavy_amp = float(amp)
    return mesh


def build_mesh_metrics_analytic_wavy(mesh: Mesh2D, ops: FROperators, amp: float) -> MeshMetrics2D:
    """Analytic wavy metrics by the chain rule through the affine computational coordinates
    (MeshMetrics2D.jl:320-442): volume terms at the solution points, face normals and scalings
    from the map evaluated on the faces (continuous across elements), affine h_char."""
    Np, nx, Nel = ops.n_points, mesh.nx, mesh.n_elements
    xi = ops.xi
    e = np.arange(Nel)
    jx, jy = e % nx, e // nx
    Jx, Jy = mesh.Jx[jx], mesh.Jy[jy]                       # (Nel,)
    xc = (mesh.x_vertices[jx] + mesh.x_vertices[jx + 1]) / 2.0
    yc = (mesh.y_vertices[jy] + mesh.y_vertices[jy + 1]) / 2.0
    X = xc[None, None, :] + Jx[None, None, :] * xi[:, None, None]   # X[i, ., e]
    Y = yc[None, None, :] + Jy[None, None, :] * xi[None, :, None]   # Y[., j, e]
    X, Y = np.broadcast_arrays(X, Y)
    x, y = wavy_physical(X, Y, amp)
    x_X, x_Y, y_X, y_Y = wavy_partials(X, Y, amp)
    x_xi = x_X * Jx
    x_eta = x_Y * Jy
    y_xi = y_X * Jx
    y_eta = y_Y * Jy
    J = x_xi * y_eta - x_eta * y_xi
    f = lambda a: np.asfortranarray(a)
    out = {}
    eps = np.finfo(np.float64).eps
    Yf = yc[None, :] + Jy[None, :] * xi[:, None]            # (Np, Nel): W/E face points
    Xf = xc[None, :] + Jx[None, :] * xi[:, None]            # S/N face points
    for side, sgn in (("W", -1.0), ("E", 1.0)):
        Xs = xc + Jx * sgn
        _, x_Yf, _, y_Yf = wavy_partials(np.broadcast_to(Xs, Yf.shape), Yf, amp)
        xeta = x_Yf * Jy
        yeta = y_Yf * Jy
        sf = np.maximum(np.sqrt(yeta * yeta + xeta * xeta), eps)
        out["sJ_" + side] = f(sf)
        out["nx_" + side] = f((-yeta if side == "W" else yeta) / sf)
        out["ny_" + side] = f((xeta if side == "W" else -xeta) / sf)
    for side, sgn in (("S", -1.0), ("N", 1.0)):
        Ys = yc + Jy * sgn
        x_Xf, _, y_Xf, _ = wavy_partials(Xf, np.broadcast_to(Ys, Xf.shape), amp)
        xxi = x_Xf * Jx
        yxi = y_Xf * Jx
        sf = np.maximum(np.sqrt(yxi * yxi + xxi * xxi), eps)
        out["sJ_" + side] = f(sf)
        out["nx_" + side] = f((yxi if side == "S" else -yxi) / sf)
        out["ny_" + side] = f((-xxi if side == "S" else xxi) / sf)
    h_char = np.minimum(mesh.dx[jx], mesh.dy[jy])
    return MeshMetrics2D(f(x), f(y), f(x_xi), f(x_eta), f(y_xi), f(y_eta), f(J), h_char,
                         is_curved=True, **out)
