"""1D Euler boundary checks (Cases.jl:590-700): a uniform state between transmissive ends, and
between constant Dirichlet ends, must stay uniform."""
from __future__ import annotations

import time

import numpy as np

from .capturing import NullCapturing
from .mesh import DirichletBC, TransmissiveBC
from .oned import Euler1D, Mesh1D, allocate_state_1d, positivity_ok_1d, set_initial_condition_1d
from .operators import build_operators
from .report import case_report_dict


def _freestream_case(bc_name, U0, p, n_elements, t_final, cfl, gamma, integrator, make_bc):
    from .sod import device_integrator
    t0 = time.time()
    integrator = integrator or device_integrator
    eq = Euler1D(gamma)
    ops = build_operators(p)
    mesh = Mesh1D(0.0, 1.0, n_elements, left_bc=make_bc(), right_bc=make_bc())
    state = allocate_state_1d(mesh, ops)
    set_initial_condition_1d(state, lambda x: np.broadcast_to(U0, np.shape(x) + (3,)))
    r = integrator(state, eq, NullCapturing(), t_final, cfl)
    dev = float(max(np.max(np.abs(state.u[:, :, c] - U0[c])) for c in range(3)))
    pos = positivity_ok_1d(eq, state.u)
    ok = r["status"] == "ok" and dev < 1e-10 and pos
    return case_report_dict(
        name=f"bc_{bc_name}_freestream_p{p}", case_type="other", equation="euler1d", p=p,
        capturing_method="null", passed=ok, diverged=r["status"] != "ok",
        nan_detected=not bool(np.all(np.isfinite(state.u))), conservation_residual=dev,
        conservation_pass=dev < 1e-10, conservation_metric="none", positivity_ok=pos,
        wall_time_sec=time.time() - t0,
        metrics={"max_deviation_from_uniform": dev, "n_steps": int(r["n_steps"]), "bc": bc_name})


def run_bc_transmissive_test(p: int = 3, n_elements: int = 16, t_final: float = 0.2, cfl: float = 0.2,
                             gamma: float = 1.4, integrator=None):
    """Uniform (rho, u, p) = (1, 0.5, 1) between transmissive ends (Cases.jl:590-646)."""
    U0 = Euler1D(gamma).primitives_to_conserved(1.0, 0.5, 1.0)
    return _freestream_case("transmissive", U0, p, n_elements, t_final, cfl, gamma, integrator,
                            TransmissiveBC)


def run_bc_dirichlet_test(p: int = 3, n_elements: int = 16, t_final: float = 0.2, cfl: float = 0.2,
                          gamma: float = 1.4, integrator=None):
    """Uniform (rho, u, p) = (1, 0.3, 1) between Dirichlet ends holding the same state
    (Cases.jl:648-700)."""
    U0 = Euler1D(gamma).primitives_to_conserved(1.0, 0.3, 1.0)
    return _freestream_case("dirichlet", U0, p, n_elements, t_final, cfl, gamma, integrator,
                            lambda: DirichletBC(state=U0))


def run_euler_conservation(p: int = 3, n_elements: int = 32, t_final: float = 1.0, cfl: float = 0.2,
                           gamma: float = 1.4, method=None, integrator=None):
    """Periodic density wave: the discrete mass of every component is conserved to roundoff
    (Cases.jl:535-588)."""
    from .metrics import conservation_residual_absolute, conservation_residual_relative
    from .oned import discrete_mass_1d
    from .mesh import PeriodicBC
    from .quant import euler_density_wave_conserved
    from .sod import device_integrator
    t0 = time.time()
    method = method or NullCapturing()
    integrator = integrator or device_integrator
    eq = Euler1D(gamma)
    ops = build_operators(p)
    state = allocate_state_1d(Mesh1D(0.0, 1.0, n_elements, left_bc=PeriodicBC(), right_bc=PeriodicBC()), ops)
    set_initial_condition_1d(state, lambda x: euler_density_wave_conserved(eq, x, 0.0))
    M0 = [discrete_mass_1d(state, c) for c in range(3)]
    r = integrator(state, eq, method, t_final, cfl)
    MT = [discrete_mass_1d(state, c) for c in range(3)]
    rel = [conservation_residual_relative(M0[c], MT[c]) for c in range(3)]
    ab = [conservation_residual_absolute(M0[c], MT[c]) for c in range(3)]
    cpass = r["status"] == "ok" and all(rel[c] <= 1e-11 or ab[c] <= 1e-12 for c in range(3))
    diverged = r["status"] != "ok"
    pos = positivity_ok_1d(eq, state.u)
    return case_report_dict(
        name=f"euler_conservation_p{p}", case_type="other", equation="euler1d", p=p,
        capturing_method="null", passed=cpass and not diverged and pos, diverged=diverged,
        nan_detected=diverged or not bool(np.all(np.isfinite(state.u))), conservation_residual=max(rel),
        conservation_pass=cpass, conservation_metric="periodic_mass_change", positivity_ok=pos,
        wall_time_sec=time.time() - t0,
        metrics={"mass_initial": M0, "mass_final": MT, "mass_abs_change_components": ab,
                 "conservation_residual_components": rel, "n_elements": n_elements,
                 "n_steps": int(r["n_steps"])})


def run_m3_euler_suite(integrator=None):
    """1D Euler suite (Cases.jl:702-745): order studies at p = 2, 3, 4 on meshes in their
    asymptotic range, conservation, and the two boundary freestream checks.  Returns (cases,
    overall, hard_fails)."""
    from .quant import run_euler_smooth_order
    cases, hard = [], []
    for p in (2, 3, 4):
        nlist = (16, 32, 64) if p == 2 else ((8, 16, 32) if p == 3 else (16, 32, 64))
        c = run_euler_smooth_order(p=p, n_elements_list=nlist, cfl=0.1, integrator=integrator)
        cases.append(c)
        if not c["order_pass"]:
            hard.append(f"euler order failed p={p} observed={c['observed_orders']}")
        if not c["conservation_pass"]:
            hard.append(f"euler order-case conservation failed p={p}")
        if not c["positivity_ok"]:
            hard.append(f"euler positivity failed p={p}")
        cc = run_euler_conservation(p=p, n_elements=32, integrator=integrator)
        cases.append(cc)
        if not cc["conservation_pass"]:
            hard.append(f"euler conservation failed p={p} res={cc['conservation_residual']}")
    bt = run_bc_transmissive_test(integrator=integrator)
    cases.append(bt)
    if not bt["pass"]:
        hard.append("transmissive BC freestream test failed")
    bd = run_bc_dirichlet_test(integrator=integrator)
    cases.append(bd)
    if not bd["pass"]:
        hard.append("dirichlet BC freestream test failed")
    return cases, all(c["pass"] is True for c in cases) and not hard, hard
