"""Optional 2D wall benchmarks: the reduced forward-facing step on a solid-element mask
(Cases2D_Optional.jl:142-267).

The case runner builds the reference's problem (mesh, sides, mask, inflow state, method) and
advances it through the device path (frforge_amd.solver.integrate); the report carries the
reference's keys and metrics.
"""
from __future__ import annotations

import math
import time
from typing import Optional

import numpy as np

from .capturing import AbstractCapturingMethod, PerssonAVMethod, method_params
from .equations import Euler2D, pressure, primitives_to_conserved
from .mesh import DirichletBC, Mesh2D, ReflectingBC, TransmissiveBC
from .operators import build_operators
from .report import case_report_dict
from .state import allocate_state, has_nonfinite, set_initial_condition


def make_ffs_solid_mask(nx: int, ny: int, Lx: float, Ly: float, step_x: float = 0.6,
                        step_h: float = 0.2) -> np.ndarray:
    """Elements whose centres lie inside the step x >= step_x, y <= step_h
    (Cases2D_Optional.jl:142-167); element e = jy * nx + jx."""
    dx, dy = Lx / nx, Ly / ny
    xc = (np.arange(nx) + 0.5) * dx
    yc = (np.arange(ny) + 0.5) * dy
    return ((xc[None, :] >= step_x) & (yc[:, None] <= step_h)).ravel()


def fluid_positivity_ok(state, eq: Euler2D) -> bool:
    """Density and pressure positive and finite at every point of every fluid element
    (Cases2D_Optional.jl:212-227)."""
    u = state.u
    fluid = np.ones(u.shape[2], dtype=bool) if state.mesh.solid is None else ~state.mesh.solid
    rho = u[:, :, fluid, 0]
    pr = pressure(eq, u[:, :, fluid, :])
    return bool(np.all(rho > 0) and np.all(pr > 0) and np.all(np.isfinite(rho)) and np.all(np.isfinite(pr)))


def run_forward_facing_step(p: int = 1, nx: int = 24, ny: int = 8, t_final: float = 0.05,
                            cfl: float = 0.04, gamma: float = 1.4, Lx: float = 1.5, Ly: float = 0.5,
                            step_x: float = 0.3, step_h: float = 0.1, M_in: float = 3.0,
                            method: Optional[AbstractCapturingMethod] = None,
                            method_name: str = "persson_av", require_positivity: bool = False,
                            precision: str = "fp64", integrator=None):
    """Reduced forward-facing step (Cases2D_Optional.jl:169-267): a Mach M_in inflow (constant
    Dirichlet state) on the left, transmissive right, reflecting bottom and top, and the step
    as solid elements.  Pass: not diverged and finite (positivity of the fluid elements only
    when require_positivity).  Returns (case, state, eq)."""
    from .suite2d import device_integrator_2d
    t0 = time.time()
    method = method or PerssonAVMethod.create(c_av=0.15)
    integrator = integrator or device_integrator_2d
    eq = Euler2D(gamma)
    ops = build_operators(p)
    c0 = math.sqrt(gamma * 1.0 / 1.4)
    U_in = primitives_to_conserved(eq, 1.4, M_in * c0, 0.0, 1.0)
    solid = make_ffs_solid_mask(nx, ny, Lx, Ly, step_x=step_x, step_h=step_h)
    mesh = Mesh2D(0.0, Lx, 0.0, Ly, nx, ny, left_bc=DirichletBC(state=U_in),
                  right_bc=TransmissiveBC(), bottom_bc=ReflectingBC(), top_bc=ReflectingBC(),
                  solid=solid)
    state = allocate_state(mesh, ops, precision=precision)
    set_initial_condition(state, lambda x, y: U_in)
    result = integrator(state, eq, method, t_final, cfl)
    diverged = result["status"] != "ok"
    nan_det = diverged or has_nonfinite(state.u)
    pos = (not nan_det) and fluid_positivity_ok(state, eq)
    passed = (not diverged) and (not nan_det) and ((not require_positivity) or pos)
    case = case_report_dict(
        name=f"forward_facing_step_p{p}_{method_name}", case_type="discontinuous",
        equation="euler2d", p=p, capturing_method=method_name, passed=passed, diverged=diverged,
        nan_detected=nan_det, conservation_residual=0.0, conservation_pass=True,
        conservation_metric="none", positivity_ok=pos, wall_time_sec=time.time() - t0,
        n_elements=nx * ny, t_final=t_final, shock_thickness_unit="sp_spacings", overshoot=0.0,
        metrics={"nx": nx, "ny": ny, "Lx": Lx, "Ly": Ly, "step_x": step_x, "step_h": step_h,
                 "n_solid": int(solid.sum()), "n_steps": int(result["n_steps"]), "M_in": M_in,
                 "method_params": method_params(method), "ci_tier": "full_nightly",
                 "note": "optional reduced FFS via solid mask (full/nightly tier)"})
    return case, state, eq
