"""The reference's quantitative verification suite for 1D Euler (run_m5_quant_suite,
Cases.jl:1188-1273) and the invent scoring step built on it: the smooth density-wave order
studies (Cases.jl:396-536), the Sod tube (frforge_amd.sod) and the Shu-Osher problem
(Cases.jl:925-1183), each run through an integrator that defaults to the device time loop,
gathered into a schema v1 report with absolute scores, and a candidate method classified
against a baseline report (Candidate.jl:136-183)."""
from __future__ import annotations

import math
import time
from typing import Any, Dict, List, Optional, Tuple

import numpy as np

from .capturing import (AbstractCapturingMethod, NullCapturing, PerssonAVMethod,
                        get_capturing_method, method_params)
from .mesh import PeriodicBC, TransmissiveBC
from .metrics import (DEFAULT_ORDER_TOLERANCE, conservation_residual_absolute,
                      conservation_residual_relative, excess_dissipation, interp_1d,
                      observed_orders, order_pass, overshoot_metric, shock_thickness_sp,
                      smooth_region_mask)
from .oned import (Euler1D, Mesh1D, allocate_state_1d, discrete_mass_1d, positivity_ok_1d,
                   sample_solution_1d, set_initial_condition_1d, solution_extrema)
from .operators import build_operators
from .report import case_report_dict, report_skeleton
from .scoring import classify_candidate, collect_hard_gate_failures
from .sod import device_integrator, run_sod
from .state import SchemeConfig

TWO_PI = 2.0 * math.pi     # Julia's 2π as a Float64
_vsin = np.vectorize(math.sin, otypes=[np.float64])


def euler_density_wave_conserved(eq: Euler1D, x, t, A: float = 0.2, u0: float = 1.0, p0: float = 1.0):
    """rho = 1 + A sin(2 pi (x - u0 t)), u = u0, p = p0 (Cases.jl:404-414)."""
    x = np.asarray(x, dtype=np.float64)
    rho = 1.0 + A * _vsin(TWO_PI * (x - u0 * t))
    return eq.primitives_to_conserved(rho, u0 + 0.0 * x, p0 + 0.0 * x)


def shu_osher_ic(eq: Euler1D, x, x0: float = 1.0):
    """Left state (3.857143, 2.629369, 10.33333) for x < x0, else rho = 1 + 0.2 sin(5 (x - 5)),
    u = 0, p = 1, on [0, 10] (Cases.jl:930-938)."""
    x = np.asarray(x, dtype=np.float64)
    left = x < x0
    rho = np.where(left, 3.857143, 1.0 + 0.2 * _vsin(5.0 * (x - 5.0)))
    return eq.primitives_to_conserved(rho, np.where(left, 2.629369, 0.0), np.where(left, 10.33333, 1.0))


def _euler_order_study_dt(p: int, ne_fine: int, lam_max: float, cfl: float = 0.2) -> float:
    """(cfl / (p+1)) h / ((2p+1) lambda_max) with h = 1/ne_fine (Cases.jl:416-421)."""
    h = 1.0 / ne_fine
    cfl_eff = cfl / (p + 1)
    return cfl_eff * h / ((2 * p + 1) * max(lam_max, np.finfo(np.float64).eps))


def l2_error_1d(state, uexact, c: int = 0) -> float:
    """sqrt(sum_e sum_j w_j J_e (u - u_exact)^2) in the reference's loop order
    (SolutionState.jl:93-107); uexact maps an array of x to (..., Neq)."""
    X = state.mesh.physical_coords(state.ops)
    ex = np.asarray(uexact(X), dtype=np.float64)[..., c]
    w, J = state.ops.w, state.mesh.J
    acc = 0.0
    for e in range(state.u.shape[1]):
        for j in range(state.u.shape[0]):
            err = state.u[j, e, c] - ex[j, e]
            acc += w[j] * J[e] * err * err
    return math.sqrt(acc)


def run_euler_smooth_order(p: int = 3, n_elements_list=(8, 16, 32), t_final: float = 1.0,
                           cfl: float = 0.2, order_tol: float = DEFAULT_ORDER_TOLERANCE,
                           gamma: float = 1.4, method: Optional[AbstractCapturingMethod] = None,
                           scheme: Optional[SchemeConfig] = None, integrator=None) -> Dict[str, Any]:
    """Periodic density-wave order study on [0, 1] at a fixed dt, then a CFL run on the finest
    mesh for the conservation gate (Cases.jl:433-536)."""
    t0 = time.time()
    method = method or NullCapturing()
    scheme = scheme or SchemeConfig()
    integrator = integrator or device_integrator
    eq = Euler1D(gamma)
    ops = build_operators(p)
    formal = p + 1
    dt_fixed = _euler_order_study_dt(p, max(n_elements_list), 2.5, cfl=cfl)
    mesh_sizes: List[float] = []
    l2: List[float] = []
    diverged = nan = False
    pos = True
    uexact = lambda x: euler_density_wave_conserved(eq, x, t_final)

    def fresh(ne):
        mesh = Mesh1D(0.0, 1.0, ne, left_bc=PeriodicBC(), right_bc=PeriodicBC())
        st = allocate_state_1d(mesh, ops, scheme=scheme)
        set_initial_condition_1d(st, lambda x: euler_density_wave_conserved(eq, x, 0.0))
        return st

    for ne in n_elements_list:
        st = fresh(ne)
        r = integrator(st, eq, method, t_final, 0.2, dt_fixed)
        mesh_sizes.append(1.0 / ne)
        if r["status"] != "ok":
            diverged = nan = True
            l2.append(float("nan"))
            continue
        l2.append(l2_error_1d(st, uexact, 0))
        pos = pos and positivity_ok_1d(eq, st.u)
        if not np.all(np.isfinite(st.u)):
            diverged = nan = True
    obs = observed_orders(mesh_sizes, l2)
    opass = (not diverged) and order_pass(obs, formal, order_tol)

    st = fresh(n_elements_list[-1])
    m0 = [discrete_mass_1d(st, c) for c in range(3)]
    r = integrator(st, eq, method, t_final, cfl, None)
    mt = [discrete_mass_1d(st, c) for c in range(3)]
    cres_comp = [conservation_residual_relative(m0[c], mt[c]) for c in range(3)]
    abs_comp = [conservation_residual_absolute(m0[c], mt[c]) for c in range(3)]
    cpass = r["status"] == "ok" and all(cres_comp[c] <= 1e-10 or abs_comp[c] <= 1e-11 for c in range(3))
    pos = pos and positivity_ok_1d(eq, st.u)
    return case_report_dict(
        name=f"euler_smooth_order_p{p}", case_type="smooth_order", equation="euler1d", p=p,
        capturing_method="null", passed=opass and cpass and not diverged and not nan and pos,
        diverged=diverged, nan_detected=nan, conservation_residual=max(cres_comp),
        conservation_pass=cpass, conservation_metric="periodic_mass_change", positivity_ok=pos,
        wall_time_sec=time.time() - t0,
        metrics={"γ": gamma, "t_final": t_final, "cfl": cfl, "dt_fixed": dt_fixed,
                 "conservation_residual_components": cres_comp,
                 "mass_abs_change_components": abs_comp,
                 "n_elements_list": list(n_elements_list)},
        extra={"mesh_sizes": mesh_sizes, "l2_errors": l2, "observed_orders": obs,
               "formal_order": formal, "order_pass": opass, "order_tolerance": order_tol})


def load_shu_osher_reference(path: str) -> Tuple[np.ndarray, np.ndarray]:
    """x, rho columns of a CSV, skipping comments, blanks and short lines (Cases.jl:1169-1184)."""
    xs, rs = [], []
    with open(path) as f:
        for line in f:
            if line.startswith("#") or not line.strip():
                continue
            parts = line.rstrip("\n").split(",")
            if len(parts) < 2:
                continue
            xs.append(float(parts[0]))
            rs.append(float(parts[1]))
    return np.array(xs), np.array(rs)


def run_shu_osher(p: int = 2, n_elements: int = 80, t_final: float = 1.8, cfl: float = 0.15,
                  gamma: float = 1.4, method: Optional[AbstractCapturingMethod] = None,
                  method_name: str = "persson_av", scheme: Optional[SchemeConfig] = None,
                  integrator=None, reference_csv: Optional[str] = None) -> Dict[str, Any]:
    """Shu-Osher on [0, 10], transmissive ends: overshoot against the initial bounds, excess
    dissipation against NullCapturing on the same mesh, shock thickness, and L1 against the
    frozen reference density when `reference_csv` is given (Cases.jl:1068-1167)."""
    t0 = time.time()
    method = method or PerssonAVMethod.create()
    scheme = scheme or SchemeConfig()
    integrator = integrator or device_integrator
    eq = Euler1D(gamma)
    ops = build_operators(p)
    mesh = Mesh1D(0.0, 10.0, n_elements, left_bc=TransmissiveBC(), right_bc=TransmissiveBC())

    def fresh():
        st = allocate_state_1d(mesh, ops, scheme=scheme)
        set_initial_condition_1d(st, lambda x: shu_osher_ic(eq, x))
        return st

    state = fresh()
    umin0, umax0 = solution_extrema(state, 0)
    m0 = discrete_mass_1d(state, 0)
    result = integrator(state, eq, method, t_final, cfl, None)
    diverged = result["status"] != "ok"
    nan = diverged or not np.all(np.isfinite(state.u))
    pos = (not nan) and positivity_ok_1d(eq, state.u)
    mt = discrete_mass_1d(state, 0)
    x, rho = sample_solution_1d(state, 0)
    umin, umax = solution_extrema(state, 0)
    _, _, eta = overshoot_metric(umin, umax, umin0, max(umax0, 4.5))

    state_n = fresh()
    rn = integrator(state_n, eq, NullCapturing(), t_final, cfl, None)
    dex = None
    if rn["status"] == "ok":
        xn, rho_n = sample_solution_1d(state_n, 0)
        mask = smooth_region_mask(x, rho_n, grad_frac=0.12, exclude_windows=[(1.5, 3.0)])
        dex = excess_dissipation(rho, interp_1d(xn, rho_n, x), mask)
    delta = shock_thickness_sp(x, rho)
    l1_ref = None
    if reference_csv is not None:
        xr, rr = load_shu_osher_reference(reference_csv)
        ri = interp_1d(xr, rr, x)
        l1_ref = float(np.sum(np.abs(rho - ri)) / (np.sum(np.abs(ri)) + 1e-30))
    return case_report_dict(
        name=f"shu_osher_p{p}_ne{n_elements}_{method_name}", case_type="discontinuous",
        equation="euler1d", p=p, capturing_method=method_name,
        passed=(not diverged) and (not nan) and pos, diverged=diverged, nan_detected=nan,
        conservation_residual=abs(mt - m0), conservation_pass=True, conservation_metric="none",
        positivity_ok=pos, wall_time_sec=time.time() - t0, n_elements=n_elements,
        t_final=t_final, excess_dissipation=dex, shock_thickness=delta,
        shock_thickness_unit="sp_spacings", overshoot=eta,
        metrics={"mass_change": mt - m0, "n_steps": result["n_steps"], "u_min": umin, "u_max": umax,
                 "reference_file": "test/data/shu_osher_ref.csv" if reference_csv else None},
        extra={"l1_error_vs_reference": l1_ref, "method_params": method_params(method)})


def run_quant_suite(method_name: str = "persson_av", scheme: Optional[SchemeConfig] = None,
                    light: bool = False, integrator=None, shu_osher_reference: Optional[str] = None):
    """Smooth order p=2 and p=3, Sod p=2 Ne=64, Shu-Osher p=1 Ne=100 (c_av raised to at least 0.3
    for a plain Persson method); `light` is the reduced suite.  Returns (cases, overall_pass,
    hard_gate_failures) (Cases.jl:1192-1273)."""
    scheme = scheme or SchemeConfig()
    method = get_capturing_method(method_name)
    kw = dict(scheme=scheme, integrator=integrator)

    def shu_method():
        if isinstance(method, PerssonAVMethod):
            return PerssonAVMethod.create(kappa=method.sensor.kappa, s0_factor=method.sensor.s0_factor,
                                          c_av=max(method.dissip.c_av, 0.3), av_form=method.dissip.av_form)
        return method

    cases = []
    if light:
        cases.append(run_euler_smooth_order(p=2, n_elements_list=[16, 32], cfl=0.1, **kw))
        cases.append(run_sod(p=2, n_elements=32, t_final=0.15, method=method, method_name=method_name, **kw))
        cases.append(run_shu_osher(p=1, n_elements=40, t_final=0.4, cfl=0.1, method=shu_method(),
                                   method_name=method_name, reference_csv=shu_osher_reference, **kw))
    else:
        for p in (2, 3):
            nlist = [16, 32, 64] if p == 2 else [8, 16, 32]
            cases.append(run_euler_smooth_order(p=p, n_elements_list=nlist, cfl=0.1, **kw))
        cases.append(run_sod(p=2, n_elements=64, t_final=0.2, method=method, method_name=method_name, **kw))
        cases.append(run_shu_osher(p=1, n_elements=100, t_final=1.8, cfl=0.1, method=shu_method(),
                                   method_name=method_name, reference_csv=shu_osher_reference, **kw))
    hard = collect_hard_gate_failures(cases)
    overall = all(c["pass"] is True for c in cases) and not hard
    return cases, overall, hard


def quant_report(method_name: str = "persson_av", command: str = "invent", light: bool = False,
                 baseline_name=None, integrator=None, shu_osher_reference: Optional[str] = None,
                 scheme: Optional[SchemeConfig] = None) -> Dict[str, Any]:
    """The quant suite of one method as a schema v1 report with absolute scores."""
    t0 = time.time()
    cases, overall, hard = run_quant_suite(method_name, scheme=scheme, light=light,
                                           integrator=integrator,
                                           shu_osher_reference=shu_osher_reference)
    return report_skeleton(
        command=command, suite="quant", method_name=method_name,
        method_params=method_params(get_capturing_method(method_name)), baseline_name=baseline_name,
        overall_pass=overall, diverged=any(c["diverged"] for c in cases),
        nan_detected=any(c["nan_detected"] for c in cases), wall_time_sec=time.time() - t0,
        hard_gate_failures=hard, cases=cases, scheme=scheme)


def score_candidate(method_name: str, baseline_name: str = "persson_av", delta: float = 0.02,
                    light: bool = False, integrator=None, shu_osher_reference: Optional[str] = None,
                    baseline_report: Optional[Dict[str, Any]] = None):
    """Run the quant suite for a candidate method (and for the baseline unless its report is
    given) and classify the candidate; returns (method_report, baseline_report, comparison)."""
    kw = dict(light=light, integrator=integrator, shu_osher_reference=shu_osher_reference)
    base = baseline_report or quant_report(baseline_name, **kw)
    rep = quant_report(method_name, baseline_name=baseline_name, **kw)
    cmp = classify_candidate(rep, base, delta=delta)
    rep["summary"]["tradeoff_notes"] = cmp["tradeoff_notes"]
    for k in ("candidate_status", "composite_margin", "tradeoff_ok", "vtk_produced"):
        rep[k] = cmp[k]
    rep["baseline_composite"] = cmp["baseline_composite"]
    rep["score_margin_threshold"] = cmp["score_margin_threshold"]
    return rep, base, cmp
