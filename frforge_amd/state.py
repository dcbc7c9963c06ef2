"""2D FR solution state and host-side diagnostics (FRForge `src/solvestate/SolutionState2D.jl`).

``state.u`` is a float64 numpy array of shape (Np, Np, Nel, 4) in Fortran order, i.e. the
reference's column-major ``u[i, j, e, c]``; its raw memory is exactly what the device library
uploads and downloads.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Any, Callable, Optional

import numpy as np

from .equations import Euler2D, pressure
from .mesh import Mesh2D, MeshMetrics2D, build_mesh_metrics, build_mesh_metrics_analytic_wavy
from .operators import FROperators


@dataclass(frozen=True)
class SchemeConfig:
    """Scheme axes (SchemeConfig.jl:15-34): solution points, interface flux, time scheme."""

    points: str = "gl"
    flux: str = "rusanov"
    time: str = "ssp_rk3"

    def __post_init__(self):
        if self.points not in ("gl", "gll"):
            raise ValueError(f"points must be 'gl' or 'gll', got {self.points}")
        if self.flux not in ("rusanov", "hllc"):
            raise ValueError(f"flux must be 'rusanov' or 'hllc', got {self.flux}")
        if self.time not in ("ssp_rk3", "ssp_rk2"):
            raise ValueError(f"time must be 'ssp_rk3' or 'ssp_rk2', got {self.time}")


DEFAULT_SCHEME = SchemeConfig()


@dataclass
class SolutionState2D:
    """(SolutionState2D.jl:9-18).  ``precision`` selects the device arithmetic: "fp64" (plain)
    or "fp32_mixed"; the state itself is always float64."""

    u: np.ndarray
    t: float
    p: int
    mesh: Mesh2D
    ops: FROperators
    scheme: SchemeConfig
    metrics: MeshMetrics2D
    precision: str = "fp64"
    device: Any = field(default=None, repr=False)   # lazily created device context

    @property
    def n_vars(self) -> int:
        return self.u.shape[3]


def allocate_state(mesh: Mesh2D, ops: FROperators, neq: int = 4,
                   scheme: Optional[SchemeConfig] = None, precision: str = "fp64",
                   metrics=None, wavy_amp: Optional[float] = None) -> SolutionState2D:
    """(SolutionState2D.jl:20-43): the given metrics, else the analytic wavy ones when
    ``wavy_amp`` is set, else build_mesh_metrics (affine, or the discrete metrics of a wavy mesh).
    Only Neq = 4 (2D Euler) runs on the device; curved metrics run in the FP64 form."""
    if precision not in ("fp64", "fp32_mixed"):
        raise ValueError(f"precision must be 'fp64' or 'fp32_mixed', got {precision}")
    Np = ops.n_points
    u = np.zeros((Np, Np, mesh.n_elements, neq), order="F")
    sch = scheme if scheme is not None else SchemeConfig(points=ops.points)
    if metrics is not None:
        met = metrics
    elif wavy_amp is not None:
        met = build_mesh_metrics_analytic_wavy(mesh, ops, float(wavy_amp))
    else:
        met = build_mesh_metrics(mesh, ops)
    return SolutionState2D(u, 0.0, ops.p, mesh, ops, sch, met, precision)


def set_initial_condition(state: SolutionState2D, f: Callable) -> SolutionState2D:
    """Sample ``f(x, y)`` at every solution point (SolutionState2D.jl:63-83).

    ``f`` is called once with arrays X, Y of shape (Np, Np, Nel) and returns an array of
    shape (Np, Np, Nel, Neq) (or broadcastable to it)."""
    X, Y = state.metrics.x, state.metrics.y
    val = np.asarray(f(X, Y), dtype=np.float64)
    state.u[...] = np.broadcast_to(val, state.u.shape)
    state.t = 0.0
    if state.device is not None:
        state.device.mark_host_dirty()
    return state


def _seq_sum(terms: np.ndarray) -> float:
    """Left-to-right sum over Fortran order, as the reference's scalar accumulation loops."""
    flat = np.asarray(terms, dtype=np.float64).ravel(order="F")
    if flat.size == 0:
        return 0.0
    return float(np.cumsum(flat)[-1])


def discrete_mass(state: SolutionState2D, c: int = 0) -> float:
    """sum w_i w_j |J| u_c (SolutionState2D.jl:85-96); ``c`` is 0-based."""
    w = state.ops.w
    W = w[:, None, None] * w[None, :, None]
    return _seq_sum(W * np.abs(state.metrics.J) * state.u[:, :, :, c])


def l2_error(state: SolutionState2D, uexact: Callable, c: int = 0) -> float:
    """sqrt(sum w_i w_j |J| (u_c - exact_c)^2) (SolutionState2D.jl:98-110); ``uexact`` is
    vectorised like the initial-condition callables; ``c`` is 0-based."""
    ex = np.asarray(uexact(state.metrics.x, state.metrics.y), dtype=np.float64)
    exc = ex if ex.ndim == 3 else ex[..., c]
    err = state.u[:, :, :, c] - exc
    w = state.ops.w
    W = w[:, None, None] * w[None, :, None]
    acc = _seq_sum(W * np.abs(state.metrics.J) * err * err)
    return float(np.sqrt(acc))


def positivity_ok(eq: Euler2D, state: SolutionState2D, atol: float = 0.0) -> bool:
    """rho > atol and p > atol, both finite, at every SP (Euler2D.jl:153-163)."""
    U = np.stack([state.u[:, :, :, c] for c in range(4)], axis=-1)
    p = pressure(eq, U)
    rho = U[..., 0]
    ok = (rho > atol) & (p > atol) & np.isfinite(rho) & np.isfinite(p)
    return bool(np.all(ok))


def has_nonfinite(a: np.ndarray) -> bool:
    """any(!isfinite) (Residual.jl:219-226)."""
    return not bool(np.all(np.isfinite(a)))
