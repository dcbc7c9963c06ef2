"""Optional 2D wall benchmarks (Cases2D_Optional.jl): the reduced double-Mach reflection on
moving ghost states, the reduced forward-facing step on a solid-element mask, and the suite that
runs both.

Each case runner builds the reference's problem (mesh, sides, mask, states, method) and advances
it through the device path (frforge_amd.solver.integrate); the reports carry the reference's keys
and metrics.
"""
from __future__ import annotations

import math
import time
from typing import Optional

import numpy as np

from .capturing import AbstractCapturingMethod, PerssonAVMethod, method_params
from .equations import Euler2D, pressure, primitives_to_conserved
from .mesh import DirichletBC, GhostStateBC, Mesh2D, ReflectingBC, TransmissiveBC
from .operators import build_operators
from .report import case_report_dict
from .state import allocate_state, has_nonfinite, positivity_ok, set_initial_condition


def double_mach_states(eq: Euler2D, strength: str = "reduced"):
    """(U_post, U_pre) of the inclined shock (Cases2D_Optional.jl:7-26): the Woodward-Colella
    Mach 10 states for "classic", a milder post-shock state for "reduced"."""
    U_pre = primitives_to_conserved(eq, 1.4, 0.0, 0.0, 1.0)
    if strength == "classic":
        U_post = primitives_to_conserved(eq, 8.0, 8.25 * math.cos(math.pi / 6),
                                         -8.25 * math.sin(math.pi / 6), 116.5)
    else:
        U_post = primitives_to_conserved(eq, 3.5, 1.6 * math.cos(math.pi / 6),
                                         -1.6 * math.sin(math.pi / 6), 4.5)
    return U_post, U_pre


def double_mach_ic(eq: Euler2D, x_shock0: float = 1.0 / 6.0, strength: str = "reduced"):
    """Post-shock state left of the 60-degree front x < x_shock0 + y / sqrt(3), pre-shock right
    of it (Cases2D_Optional.jl:28-37), as a set_initial_condition function."""
    U_post, U_pre = double_mach_states(eq, strength)

    def f(x, y):
        xs = x_shock0 + y / math.sqrt(3.0)
        return np.where((x < xs)[..., None], U_post, U_pre)
    return f


def run_double_mach_reflection(p: int = 1, nx: int = 24, ny: int = 6, t_final: float = 0.04,
                               cfl: float = 0.04, gamma: float = 1.4, x_shock0: float = 1.0 / 6.0,
                               Lx: float = 2.0, Ly: float = 0.5, strength: str = "reduced",
                               method: Optional[AbstractCapturingMethod] = None,
                               method_name: str = "persson_av", require_positivity: bool = False,
                               precision: str = "fp64", integrator=None):
    """Reduced double-Mach reflection (Cases2D_Optional.jl:39-140): post-shock inflow on the left,
    transmissive right, a bottom that is post-shock for x < x_shock0 and a reflecting wall from
    there on, and a top whose post/pre-shock switch moves with the shock foot,
    x_s(t) = x_shock0 + (Ly + vs t) / sqrt(3).  Both ghost sides are GhostStateBC.split forms of
    the reference's closures.  Returns (case, state, eq)."""
    from .suite2d import device_integrator_2d
    t0 = time.time()
    method = method or PerssonAVMethod.create(c_av=0.15)
    integrator = integrator or device_integrator_2d
    eq = Euler2D(gamma)
    ops = build_operators(p)
    U_post, U_pre = double_mach_states(eq, strength)
    vs = 20.0 if strength == "classic" else 2.5
    bottom = GhostStateBC.split(x_shock0, 0.0, 0.0, 1.0, U_post, hi="reflecting")
    top = GhostStateBC.split(x_shock0, Ly, vs, math.sqrt(3.0), U_post, hi="dirichlet", U_hi=U_pre)
    mesh = Mesh2D(0.0, Lx, 0.0, Ly, nx, ny, left_bc=DirichletBC(state=U_post),
                  right_bc=TransmissiveBC(), bottom_bc=bottom, top_bc=top)
    state = allocate_state(mesh, ops, precision=precision)
    set_initial_condition(state, double_mach_ic(eq, x_shock0, strength))
    result = integrator(state, eq, method, t_final, cfl)
    diverged = result["status"] != "ok"
    nan_det = diverged or has_nonfinite(state.u)
    pos = (not nan_det) and positivity_ok(eq, state)
    passed = (not diverged) and (not nan_det) and ((not require_positivity) or pos)
    rho = state.u[:, :, :, 0]
    case = case_report_dict(
        name=f"double_mach_{strength}_p{p}_{method_name}", case_type="discontinuous",
        equation="euler2d", p=p, capturing_method=method_name, passed=passed, diverged=diverged,
        nan_detected=nan_det, conservation_residual=0.0, conservation_pass=True,
        conservation_metric="none", positivity_ok=pos, wall_time_sec=time.time() - t0,
        n_elements=nx * ny, t_final=t_final, shock_thickness_unit="sp_spacings", overshoot=0.0,
        metrics={"nx": nx, "ny": ny, "Lx": Lx, "Ly": Ly, "strength": strength,
                 "n_steps": int(result["n_steps"]),
                 "ρ_min": float("nan") if nan_det else float(rho.min()),
                 "ρ_max": float("nan") if nan_det else float(rho.max()),
                 "method_params": method_params(method), "ci_tier": "full_nightly",
                 "note": "optional reduced Double Mach-like (full/nightly tier)"})
    return case, state, eq


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


def run_p33b_optional_suite(precision: str = "fp64", integrator=None):
    """The reduced double-Mach reflection and forward-facing step at their ultra-light settings
    (Cases2D_Optional.jl:269-330); a case that diverges or turns non-finite is a hard failure.
    Returns (cases, overall, hard_fails)."""
    cases, hard_fails = [], []
    c_dmr, _, _ = run_double_mach_reflection(
        p=1, nx=12, ny=4, t_final=0.03, cfl=0.04, Lx=1.2, Ly=0.4, strength="reduced",
        method=PerssonAVMethod.create(c_av=0.2), method_name="persson_av",
        require_positivity=False, precision=precision, integrator=integrator)
    cases.append(c_dmr)
    if c_dmr["diverged"] or c_dmr["nan_detected"]:
        hard_fails.append("reduced Double Mach diverged/NaN")
        c_dmr["pass"] = False
    else:
        c_dmr["pass"] = True
    c_ffs, _, _ = run_forward_facing_step(
        p=1, nx=12, ny=4, t_final=0.02, cfl=0.025, Lx=1.0, Ly=0.4, step_x=0.3, step_h=0.1,
        M_in=2.0, method=PerssonAVMethod.create(c_av=0.25), method_name="persson_av",
        require_positivity=False, precision=precision, integrator=integrator)
    cases.append(c_ffs)
    if c_ffs["diverged"] or c_ffs["nan_detected"]:
        hard_fails.append("reduced FFS diverged/NaN")
        c_ffs["pass"] = False
    else:
        c_ffs["pass"] = True
    return cases, not hard_fails, hard_fails
