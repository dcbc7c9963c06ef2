"""The reference's 2D verification cases and their CI-light suite (Cases2D_Riemann.jl:69-373):
the isentropic-vortex order study, the four-quadrant Riemann problem, and the suite that scores
them into a schema v1 report.  Every time loop runs on the device through
frforge_amd.solver.integrate (or a given ``integrator(state, eq, method, t_final, cfl, dt)``);
``precision`` selects the plain FP64 or the FP32-mixed device form, so the same report can be
written for both and their scores compared."""
from __future__ import annotations

import math
import time
from typing import Any, Dict, Optional

import numpy as np

from .capturing import AbstractCapturingMethod, NullCapturing, PerssonAVMethod, method_params
from .cases import riemann2d_ic, vortex_ic
from .equations import Euler2D
from .mesh import Mesh2D, TransmissiveBC
from .metrics import (DEFAULT_ORDER_TOLERANCE, conservation_residual_relative, observed_orders,
                      order_pass, overshoot_metric)
from .operators import build_operators
from .report import case_report_dict, report_skeleton
from .scoring import collect_hard_gate_failures
from .state import allocate_state, discrete_mass, has_nonfinite, l2_error, positivity_ok, set_initial_condition


def device_integrator_2d(state, eq, method, t_final, cfl, dt=None):
    from .solver import integrate
    return integrate(state, eq, method, t_final, cfl=cfl, dt=dt)


def run_isentropic_vortex_order(p: int = 2, n_list=(8, 16), t_final: float = 0.5, cfl: float = 0.08,
                                L: float = 10.0, gamma: float = 1.4, beta: float = 5.0,
                                x0: float = 5.0, y0: float = 5.0, u_inf: float = 1.0,
                                v_inf: float = 1.0, order_tol: float = DEFAULT_ORDER_TOLERANCE,
                                method: Optional[AbstractCapturingMethod] = None,
                                method_name: str = "null", precision: str = "fp64",
                                integrator=None) -> Dict[str, Any]:
    """Density L2 errors at a fixed dt on periodic n x n meshes of [0, L]^2, observed orders
    against p + 1 with tolerance order_tol + 0.15, and a CFL run on the finest mesh for the
    mass-conservation gate (Cases2D_Riemann.jl:194-313)."""
    t0 = time.time()
    method = method or NullCapturing()
    integrator = integrator or device_integrator_2d
    eq = Euler2D(gamma)
    ops = build_operators(p)
    formal = p + 1
    lam = abs(u_inf) + abs(v_inf) + beta / (2.0 * math.pi) + math.sqrt(gamma) + 1.0
    h = L / max(n_list)
    dt_fixed = (cfl / (p + 1)) * h / ((2 * p + 1) * lam)
    kw = dict(beta=beta, x0=x0, y0=y0, u_inf=u_inf, v_inf=v_inf)

    def fresh(n):
        st = allocate_state(Mesh2D(0.0, L, 0.0, L, n, n), ops, precision=precision)
        set_initial_condition(st, vortex_ic(eq, t=0.0, **kw))
        return st

    mesh_sizes, l2 = [], []
    diverged, pos = False, True
    for n in n_list:
        st = fresh(n)
        r = integrator(st, eq, method, t_final, 0.2, dt_fixed)
        mesh_sizes.append(L / n)
        if r["status"] != "ok":
            diverged = True
            l2.append(float("nan"))
            continue
        pos = pos and positivity_ok(eq, st)
        l2.append(l2_error(st, vortex_ic(eq, t=t_final, **kw), 0))
    obs = observed_orders(mesh_sizes, l2)
    opass = (not diverged) and order_pass(obs, formal, order_tol + 0.15)

    st = fresh(n_list[-1])
    m0 = discrete_mass(st, 0)
    r = integrator(st, eq, method, t_final, cfl, None)
    mt = discrete_mass(st, 0)
    cres = conservation_residual_relative(m0, mt)
    cpass = r["status"] == "ok" and (cres <= 1e-8 or abs(mt - m0) <= 1e-10)
    pos = pos and positivity_ok(eq, st)
    return case_report_dict(
        name=f"isentropic_vortex_order_p{p}_{method_name}", case_type="smooth_order",
        equation="euler2d", p=p, capturing_method=method_name,
        passed=opass and cpass and not diverged and pos, diverged=diverged, nan_detected=diverged,
        conservation_residual=cres, conservation_pass=cpass,
        conservation_metric="periodic_mass_change", positivity_ok=pos,
        wall_time_sec=time.time() - t0,
        metrics={"n_list": list(n_list), "L": L, "β": beta, "t_final": t_final,
                 "method_params": method_params(method), "ci_tier": "required_ci_light",
                 "note": "isentropic vortex = exact-solution smooth Euler (MMS-style) gate",
                 "precision": precision},
        extra={"mesh_sizes": mesh_sizes, "l2_errors": l2, "observed_orders": obs,
               "formal_order": formal, "order_pass": opass, "order_tolerance": order_tol + 0.15})


def run_euler2d_riemann(p: int = 1, nx: int = 16, ny: int = 16, t_final: float = 0.08,
                        cfl: float = 0.08, gamma: float = 1.4, x0: float = 0.5, y0: float = 0.5,
                        config="cfg6", method: Optional[AbstractCapturingMethod] = None,
                        method_name: str = "persson_av", require_positivity: bool = True,
                        precision: str = "fp64", flux: str = "rusanov", integrator=None):
    """Four-quadrant Riemann problem on [0,1]^2 with transmissive sides: density overshoot
    against the initial bounds widened to [0.5, 3] and the finite/positive gate
    (Cases2D_Riemann.jl:69-155).  ``config`` is "cfg6", "cfg3" or a dict of quadrant states
    (the seeded four-quadrant problems).  Returns (case, state, eq)."""
    from .state import SchemeConfig
    t0 = time.time()
    method = method or PerssonAVMethod.create(c_av=0.1)
    integrator = integrator or device_integrator_2d
    eq = Euler2D(gamma)
    ops = build_operators(p)
    T = TransmissiveBC
    mesh = Mesh2D(0.0, 1.0, 0.0, 1.0, nx, ny, left_bc=T(), right_bc=T(), bottom_bc=T(), top_bc=T())
    state = allocate_state(mesh, ops, scheme=SchemeConfig(flux=flux), precision=precision)
    set_initial_condition(state, lambda x, y: riemann2d_ic(eq, x, y, config=config, x0=x0, y0=y0))
    rmin0, rmax0 = float(np.min(state.u[:, :, :, 0])), float(np.max(state.u[:, :, :, 0]))
    result = integrator(state, eq, method, t_final, cfl, None)
    diverged = result["status"] != "ok"
    nan = diverged or has_nonfinite(state.u)
    pos = (not nan) and positivity_ok(eq, state)
    rho = state.u[:, :, :, 0]
    rmin, rmax = (float("nan"), float("nan")) if nan else (float(np.min(rho)), float(np.max(rho)))
    _, _, eta = overshoot_metric(rmin, rmax, min(rmin0, 0.5), max(rmax0, 3.0))
    cfg = config if isinstance(config, str) else "seeded"
    case = case_report_dict(
        name=f"euler2d_riemann_{cfg}_p{p}_{method_name}", case_type="discontinuous",
        equation="euler2d", p=p, capturing_method=method_name,
        passed=(not diverged) and (not nan) and ((not require_positivity) or pos),
        diverged=diverged, nan_detected=nan, conservation_residual=0.0, conservation_pass=True,
        conservation_metric="none", positivity_ok=pos, wall_time_sec=time.time() - t0,
        n_elements=nx * ny, t_final=t_final, shock_thickness_unit="sp_spacings", overshoot=eta,
        metrics={"config": cfg, "nx": nx, "ny": ny, "n_steps": result["n_steps"],
                 "ρ_min0": rmin0, "ρ_max0": rmax0, "ρ_min": rmin, "ρ_max": rmax,
                 "method_params": method_params(method),
                 "ci_tier": "required_ci_light" if cfg == "cfg6" else "full_nightly",
                 "require_positivity": require_positivity, "precision": precision})
    return case, state, eq


def run_benchmark_suite_2d(precision: str = "fp64", integrator=None):
    """The 2D CI-light suite: vortex order p=2 on 8^2/16^2 to t=0.5, cfg6 Riemann p=1 16^2 with
    Persson AV to t=0.08, and a NullCapturing cfg6 smoke on 12^2 to t=0.04
    (Cases2D_Riemann.jl:321-373).  Returns (cases, overall_pass, hard_gate_failures)."""
    kw = dict(precision=precision, integrator=integrator)
    cases, hard = [], []
    cv = run_isentropic_vortex_order(p=2, n_list=[8, 16], t_final=0.5, cfl=0.08, L=10.0, **kw)
    cases.append(cv)
    if not cv["order_pass"] or cv["diverged"] or not cv["conservation_pass"]:
        hard.append(f"isentropic vortex order failed: orders={cv['observed_orders']} L2={cv['l2_errors']}")
    cr, _, _ = run_euler2d_riemann(p=1, nx=16, ny=16, t_final=0.08, cfl=0.08, config="cfg6",
                                   method=PerssonAVMethod.create(c_av=0.1), method_name="persson_av", **kw)
    cases.append(cr)
    if not cr["pass"]:
        hard.append("2D Riemann cfg6 + persson_av failed")
    crn, _, _ = run_euler2d_riemann(p=1, nx=12, ny=12, t_final=0.04, cfl=0.06, config="cfg6",
                                    method=NullCapturing(), method_name="null", **kw)
    cases.append(crn)
    if not crn["pass"]:
        hard.append("2D Riemann cfg6 null failed")
    overall = not hard and all(c["pass"] is True for c in cases)
    return cases, overall, hard


def benchmark_report_2d(precision: str = "fp64", integrator=None) -> Dict[str, Any]:
    """The 2D suite as a schema v1 report (command "test", suite "benchmarks")."""
    t0 = time.time()
    cases, overall, hard = run_benchmark_suite_2d(precision, integrator)
    return report_skeleton(command="test", suite="benchmarks", method_name="persson_av",
                           method_params=method_params(PerssonAVMethod.create()),
                           overall_pass=overall, diverged=any(c["diverged"] for c in cases),
                           nan_detected=any(c["nan_detected"] for c in cases),
                           wall_time_sec=time.time() - t0, hard_gate_failures=hard, cases=cases)
