"""Curved-element verification cases (Cases2D_Smooth.jl:345-620): free-stream preservation on
affine and wavy meshes, a density jump on a wavy mesh, and the CI-light suite that runs them.

The wavy cases use the reference's analytic wavy metrics (``allocate_state(...; wavy_amp)``),
which the device takes through frf_set_curved_metrics (FP64 form).  The suite's scalar-advection
order study (``run_advection2d_curved_order``, Neq = 1) has no device path: the 2D device library
implements the Euler equations only.
"""
from __future__ import annotations

import time
from typing import Optional

import numpy as np

from .capturing import AbstractCapturingMethod, NullCapturing, PerssonAVMethod, method_params
from .equations import Euler2D, primitives_to_conserved
from .mesh import Mesh2D, TransmissiveBC, make_wavy_mesh2d
from .operators import build_operators
from .report import case_report_dict
from .state import allocate_state, has_nonfinite, positivity_ok, set_initial_condition


def run_freestream_preservation_2d(p: int = 2, nx: int = 4, ny: int = 4, curved: bool = False,
                                   amp: float = 0.05, t_final: float = 0.1, cfl: float = 0.1,
                                   gamma: float = 1.4, tol: Optional[float] = None,
                                   integrator=None, residual_fn=None):
    """Constant free stream (rho, u, v, p) = (1, 0.5, 0.25, 1) on the periodic unit square,
    affine or wavy with the analytic metrics (Cases2D_Smooth.jl:345-418): the instantaneous
    residual and the deviation after t_final must stay below 1e-10 / 1e-8 (affine) or 5e-4 /
    5e-5 (curved, spectral GCL error)."""
    from .solver import residual as device_residual
    from .suite2d import device_integrator_2d
    t0 = time.time()
    integrator = integrator or device_integrator_2d
    residual_fn = residual_fn or device_residual
    eq = Euler2D(gamma)
    ops = build_operators(p)
    mesh = make_wavy_mesh2d(nx, ny, ops, amp=amp) if curved else Mesh2D(0.0, 1.0, 0.0, 1.0, nx, ny)
    state = allocate_state(mesh, ops, wavy_amp=amp if curved else None)
    U0 = primitives_to_conserved(eq, 1.0, 0.5, 0.25, 1.0)
    set_initial_condition(state, lambda x, y: U0)
    du = np.zeros(state.u.shape, order="F")
    residual_fn(du, state, eq, NullCapturing())
    res_max = float(np.max(np.abs(du)))
    result = integrator(state, eq, NullCapturing(), t_final, cfl)
    err = max(float(np.max(np.abs(state.u[:, :, :, c] - U0[c]))) for c in range(4))
    pos = positivity_ok(eq, state)
    tol_res = tol if tol is not None else (5e-4 if curved else 1e-10)
    tol_err = 5e-5 if curved else 1e-8
    passed = (result["status"] == "ok" and res_max < tol_res and err < tol_err and pos
              and bool(np.all(state.metrics.J > 0)))
    return case_report_dict(
        name=f"freestream_2d_p{p}_" + ("curved" if curved else "cartesian"),
        case_type="conservation", equation="euler2d", p=p, capturing_method="null", passed=passed,
        diverged=result["status"] != "ok", nan_detected=has_nonfinite(state.u),
        conservation_residual=err, conservation_pass=err < tol_err,
        conservation_metric="freestream_max_dev", positivity_ok=pos,
        wall_time_sec=time.time() - t0,
        metrics={"residual_max": res_max, "freestream_error": err, "curved": curved, "amp": amp,
                 "tol_res": tol_res, "tol_err": tol_err, "n_steps": int(result["n_steps"])})


def run_euler2d_curved_discontinuous(p: int = 1, nx: int = 12, ny: int = 12, t_final: float = 0.03,
                                     cfl: float = 0.04, amp: float = 0.01, gamma: float = 1.4,
                                     method: Optional[AbstractCapturingMethod] = None,
                                     method_name: str = "null", integrator=None):
    """Sod-like density jump at x = 0.5 on a transmissive wavy mesh (Cases2D_Smooth.jl:495-561).
    Pass: not diverged, finite and positive."""
    from .suite2d import device_integrator_2d
    t0 = time.time()
    method = method or NullCapturing()
    integrator = integrator or device_integrator_2d
    eq = Euler2D(gamma)
    ops = build_operators(p)
    mesh = make_wavy_mesh2d(nx, ny, ops, amp=amp, left_bc=TransmissiveBC(), right_bc=TransmissiveBC(),
                            bottom_bc=TransmissiveBC(), top_bc=TransmissiveBC())
    state = allocate_state(mesh, ops, wavy_amp=amp)
    UL = primitives_to_conserved(eq, 1.0, 0.0, 0.0, 1.0)
    UR = primitives_to_conserved(eq, 0.125, 0.0, 0.0, 0.1)
    set_initial_condition(state, lambda x, y: np.where((x < 0.5)[..., None], UL, UR))
    result = integrator(state, eq, method, t_final, cfl)
    diverged = result["status"] != "ok"
    nan_det = diverged or has_nonfinite(state.u)
    pos = (not nan_det) and positivity_ok(eq, state)
    return case_report_dict(
        name=f"euler2d_curved_jump_p{p}_{method_name}", case_type="discontinuous",
        equation="euler2d", p=p, capturing_method=method_name,
        passed=(not diverged) and pos and (not nan_det), diverged=diverged, nan_detected=nan_det,
        conservation_residual=0.0, conservation_pass=True, conservation_metric="none",
        positivity_ok=pos, wall_time_sec=time.time() - t0, n_elements=nx * ny, t_final=t_final,
        shock_thickness_unit="sp_spacings", overshoot=0.0,
        metrics={"nx": nx, "ny": ny, "amp": amp, "n_steps": int(result["n_steps"]),
                 "method_params": method_params(method)})


def run_p32_curved_suite(integrator=None, residual_fn=None):
    """The CI-light curved suite (Cases2D_Smooth.jl:563-620) without its scalar-advection order
    study: Cartesian and curved free stream, and the curved jump with NullCapturing and Persson
    AV.  Returns (cases, overall, hard_fails)."""
    cases, hard = [], []
    c1 = run_freestream_preservation_2d(p=2, nx=4, ny=4, curved=False, integrator=integrator,
                                        residual_fn=residual_fn)
    cases.append(c1)
    if not c1["pass"]:
        hard.append("cartesian freestream failed")
    c2 = run_freestream_preservation_2d(p=3, nx=4, ny=4, curved=True, amp=0.05, t_final=0.05,
                                        integrator=integrator, residual_fn=residual_fn)
    cases.append(c2)
    if not c2["pass"]:
        hard.append(f"curved freestream failed res={c2['metrics']['residual_max']}")
    kw = dict(p=1, nx=12, ny=12, t_final=0.03, amp=0.01, cfl=0.04, integrator=integrator)
    c4 = run_euler2d_curved_discontinuous(method=NullCapturing(), method_name="null", **kw)
    cases.append(c4)
    if not c4["pass"]:
        hard.append("curved jump null failed")
    c5 = run_euler2d_curved_discontinuous(method=PerssonAVMethod.create(c_av=0.1),
                                          method_name="persson_av", **kw)
    cases.append(c5)
    if not c5["pass"]:
        hard.append("curved jump persson_av failed")
    return cases, (not hard) and all(c["pass"] is True for c in cases), hard
