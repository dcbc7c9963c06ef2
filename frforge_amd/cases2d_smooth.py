"""2D Euler verification cases on Cartesian meshes (Cases2D_Smooth.jl:89-345, 627-676): the
periodic density-wave order study, the density jump, and the M8 and P3.1 suites that run them.

The runners build the reference's problems and advance them through the device path
(frforge_amd.solver.integrate); the reports carry the reference's keys.  The M8 suite's
scalar-advection order study (``run_advection2d_smooth_order``, Neq = 1) has no device path: the
2D device library implements the Euler equations only.
"""
from __future__ import annotations

import math
import time
from typing import Optional

import numpy as np

from .capturing import AbstractCapturingMethod, NullCapturing, PerssonAVMethod, method_params
from .equations import Euler2D, primitives_to_conserved
from .mesh import Mesh2D, TransmissiveBC
from .metrics import (DEFAULT_ORDER_TOLERANCE, conservation_residual_relative, observed_orders,
                      order_pass, overshoot_metric)
from .operators import build_operators
from .report import case_report_dict
from .state import (allocate_state, discrete_mass, has_nonfinite, l2_error, positivity_ok,
                    set_initial_condition)

_TWO_PI = 2.0 * math.pi


def _density_wave(eq, t):
    """rho = 1 + 0.2 sin(2 pi (x - t)) sin(2 pi (y - t)), u = v = p = 1."""
    def f(x, y):
        rho = 1.0 + 0.2 * np.sin(_TWO_PI * (x - t)) * np.sin(_TWO_PI * (y - t))
        return primitives_to_conserved(eq, rho, 1.0, 1.0, 1.0)
    return f


def run_euler2d_smooth_order(p: int = 2, n_list=(8, 16, 32), t_final: float = 0.25, cfl: float = 0.06,
                             gamma: float = 1.4, order_tol: float = DEFAULT_ORDER_TOLERANCE,
                             method: Optional[AbstractCapturingMethod] = None,
                             method_name: str = "null", precision: str = "fp64", integrator=None):
    """Periodic density wave on n x n meshes of the unit square (Cases2D_Smooth.jl:89-200):
    density L2 errors at a fixed dt set by the finest mesh, observed orders against p + 1, and a
    CFL run on the finest mesh for the periodic mass-conservation gate."""
    from .suite2d import device_integrator_2d
    t0 = time.time()
    method = method or NullCapturing()
    integrator = integrator or device_integrator_2d
    eq = Euler2D(gamma)
    ops = build_operators(p)
    formal = p + 1
    lam = 1.0 + math.sqrt(gamma)
    h = 1.0 / max(n_list)
    dt_fixed = (cfl / (p + 1)) * h / ((2 * p + 1) * lam)
    exact = _density_wave(eq, t_final)
    mesh_sizes, errors = [], []
    diverged, pos_ok = False, True
    for n in n_list:
        state = allocate_state(Mesh2D(0.0, 1.0, 0.0, 1.0, n, n), ops, precision=precision)
        set_initial_condition(state, _density_wave(eq, 0.0))
        r = integrator(state, eq, method, t_final, cfl, dt=dt_fixed)
        mesh_sizes.append(1.0 / n)
        if r["status"] != "ok":
            diverged = True
            errors.append(float("nan"))
            continue
        pos_ok = pos_ok and positivity_ok(eq, state)
        errors.append(l2_error(state, exact, 0))
    obs = observed_orders(mesh_sizes, errors)
    opass = (not diverged) and order_pass(obs, formal, order_tol)
    n = n_list[-1]
    state = allocate_state(Mesh2D(0.0, 1.0, 0.0, 1.0, n, n), ops, precision=precision)
    set_initial_condition(state, _density_wave(eq, 0.0))
    M0 = discrete_mass(state, 0)
    r = integrator(state, eq, method, t_final, cfl)
    MT = discrete_mass(state, 0)
    cres = conservation_residual_relative(M0, MT)
    cpass = r["status"] == "ok" and (cres <= 1e-8 or abs(MT - M0) <= 1e-10)
    return case_report_dict(
        name=f"euler2d_smooth_order_p{p}_{method_name}", case_type="smooth_order",
        equation="euler2d", p=p, capturing_method=method_name,
        passed=opass and cpass and not diverged and pos_ok, diverged=diverged, nan_detected=diverged,
        conservation_residual=cres, conservation_pass=cpass,
        conservation_metric="periodic_mass_change", positivity_ok=pos_ok,
        wall_time_sec=time.time() - t0,
        metrics={"n_list": list(n_list), "method_params": method_params(method)},
        extra={"mesh_sizes": mesh_sizes, "l2_errors": errors, "observed_orders": obs,
               "formal_order": formal, "order_pass": opass, "order_tolerance": order_tol})


def run_euler2d_discontinuous(p: int = 1, nx: int = 20, ny: int = 20, t_final: float = 0.1,
                              cfl: float = 0.1, gamma: float = 1.4,
                              method: Optional[AbstractCapturingMethod] = None,
                              method_name: str = "null", precision: str = "fp64", integrator=None):
    """Density jump at x = 0.5 on the transmissive unit square (Cases2D_Smooth.jl:202-278); the
    overshoot is that of the density against [0.125, 1].  Returns (case, state, eq)."""
    from .suite2d import device_integrator_2d
    t0 = time.time()
    method = method or NullCapturing()
    integrator = integrator or device_integrator_2d
    eq = Euler2D(gamma)
    ops = build_operators(p)
    B = TransmissiveBC
    mesh = Mesh2D(0.0, 1.0, 0.0, 1.0, nx, ny, left_bc=B(), right_bc=B(), bottom_bc=B(), top_bc=B())
    state = allocate_state(mesh, ops, precision=precision)
    UL = primitives_to_conserved(eq, 1.0, 0.0, 0.0, 1.0)
    UR = primitives_to_conserved(eq, 0.125, 0.0, 0.0, 0.1)
    set_initial_condition(state, lambda x, y: np.where((x < 0.5)[..., None], UL, UR))
    rho0 = state.u[:, :, :, 0]
    rmin0, rmax0 = float(rho0.min()), float(rho0.max())
    r = integrator(state, eq, method, t_final, cfl)
    diverged = r["status"] != "ok"
    nan_det = diverged or has_nonfinite(state.u)
    pos = (not nan_det) and positivity_ok(eq, state)
    rho = state.u[:, :, :, 0]
    _, _, eta = overshoot_metric(float(rho.min()), float(rho.max()), 0.125, 1.0)
    case = case_report_dict(
        name=f"euler2d_density_jump_p{p}_{method_name}", case_type="discontinuous",
        equation="euler2d", p=p, capturing_method=method_name,
        passed=(not diverged) and pos and (not nan_det), diverged=diverged, nan_detected=nan_det,
        conservation_residual=0.0, conservation_pass=True, conservation_metric="none",
        positivity_ok=pos, wall_time_sec=time.time() - t0, n_elements=nx * ny, t_final=t_final,
        shock_thickness_unit="sp_spacings", overshoot=eta,
        metrics={"nx": nx, "ny": ny, "n_steps": int(r["n_steps"]), "ρ_min0": rmin0, "ρ_max0": rmax0,
                 "method_params": method_params(method)})
    return case, state, eq


def run_m8_2d_suite(precision: str = "fp64", integrator=None):
    """Baseline 2D suite (Cases2D_Smooth.jl:280-345) without its scalar-advection order study:
    Euler order, the density jump with NullCapturing and Persson AV, and the smooth wave with
    Persson AV gated on stability and conservation.  Returns (cases, overall, hard_fails,
    state_disc, eq_disc)."""
    kw = dict(precision=precision, integrator=integrator)
    cases, hard = [], []
    c2 = run_euler2d_smooth_order(p=2, n_list=(16, 32), t_final=0.25, cfl=0.06, **kw)
    cases.append(c2)
    if not c2["order_pass"]:
        hard.append(f"2D Euler order failed: {c2['observed_orders']}")
    c3, state_disc, eq_disc = run_euler2d_discontinuous(p=1, nx=16, ny=16, t_final=0.05, **kw)
    cases.append(c3)
    if not c3["pass"]:
        hard.append("2D discontinuous Euler (null) failed")
    pers = PerssonAVMethod.create(c_av=0.1)
    c4, _, _ = run_euler2d_discontinuous(p=1, nx=12, ny=12, t_final=0.05, cfl=0.08, method=pers,
                                         method_name="persson_av", **kw)
    cases.append(c4)
    if not c4["pass"]:
        hard.append("2D discontinuous Euler (persson_av) failed")
    c5 = run_euler2d_smooth_order(p=2, n_list=(8, 16), t_final=0.15, cfl=0.05, method=pers,
                                  method_name="persson_av", **kw)
    cases.append(c5)
    if c5["diverged"] or not c5["positivity_ok"] or not c5["conservation_pass"]:
        hard.append("2D Euler smooth + persson_av unstable or non-conservative")
        c5["pass"] = False
    else:
        c5["pass"] = True
        c5["metrics"]["note"] = "capturing-on smooth: stability/conservation gated; order informative"
    overall = all(c["pass"] is True for c in cases) and not hard
    return cases, overall, hard, state_disc, eq_disc


def run_p31_2d_capturing_suite(precision: str = "fp64", integrator=None):
    """Focused 2D capturing suite (Cases2D_Smooth.jl:625-676).  Returns (cases, overall,
    hard_fails)."""
    kw = dict(precision=precision, integrator=integrator)
    cases, hard = [], []
    pers = PerssonAVMethod.create(c_av=0.1)
    c_null, _, _ = run_euler2d_discontinuous(p=1, nx=12, ny=12, t_final=0.05, method_name="null", **kw)
    cases.append(c_null)
    c_av, _, _ = run_euler2d_discontinuous(p=1, nx=12, ny=12, t_final=0.05, cfl=0.08, method=pers,
                                           method_name="persson_av", **kw)
    cases.append(c_av)
    if not c_av["pass"]:
        hard.append("persson_av 2D jump failed")
    if c_av["pass"] and c_null["pass"] and c_av["overshoot"] > c_null["overshoot"] + 0.5:
        hard.append("persson_av overshoot much larger than null (unexpected)")
    c_sm = run_euler2d_smooth_order(p=2, n_list=(8, 16), t_final=0.15, cfl=0.05, method=pers,
                                    method_name="persson_av", **kw)
    if c_sm["diverged"] or not c_sm["positivity_ok"] or not c_sm["conservation_pass"]:
        c_sm["pass"] = False
        hard.append("persson_av 2D smooth unstable")
    else:
        c_sm["pass"] = True
    cases.append(c_sm)
    return cases, (not hard) and all(c["pass"] is True for c in cases), hard
