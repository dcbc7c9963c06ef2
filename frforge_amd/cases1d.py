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
