"""Split ghost sides (GhostStateBC with a moving switch point) on the GPU against the oracle
(whose split sides tests/test_oracle_walls.py pins): per-stage residuals at several boundary times
within 1e-12 (FP64) and 1e-4 (FP32-mixed), the reduced double-Mach reflection run step for step,
and the C ABI's checks."""
import math

import numpy as np
import pytest

from frforge_amd import (Mesh2D, SchemeConfig, TransmissiveBC, allocate_state, build_operators,
                         residual, set_initial_condition)
from frforge_amd.cases2d_optional import (double_mach_ic, double_mach_states,
                                          run_double_mach_reflection, run_p33b_optional_suite)
from frforge_amd.device import DeviceContext, FrfError
from frforge_amd.mesh import DirichletBC, GhostStateBC

from helpers import EQ, method_for, oracle_for_state, oracle_integrator_2d, rel_err_per_component

pytestmark = pytest.mark.gpu

U_POST, U_PRE = double_mach_states(EQ)


def _case(p, capturing, flux, precision, nx=30, ny=13, hi_bottom="reflecting"):
    """The double-Mach sides on [0, 2] x [0, 1/2]: split bottom and top, Dirichlet left."""
    Ly = 0.5
    bottom = GhostStateBC.split(1.0 / 6.0, 0.0, 0.0, 1.0, U_POST, hi=hi_bottom,
                                U_hi=U_PRE if hi_bottom == "dirichlet" else None)
    top = GhostStateBC.split(1.0 / 6.0, Ly, 2.5, math.sqrt(3.0), U_POST, hi="dirichlet", U_hi=U_PRE)
    mesh = Mesh2D(0.0, 2.0, 0.0, Ly, nx, ny, left_bc=DirichletBC(state=U_POST),
                  right_bc=TransmissiveBC(), bottom_bc=bottom, top_bc=top)
    st = allocate_state(mesh, build_operators(p), scheme=SchemeConfig(flux=flux), precision=precision)
    f = double_mach_ic(EQ)
    set_initial_condition(st, lambda x, y: f(x, y) * (1.0 + 0.02 * np.sin(5 * x + 3 * y))[..., None])
    m = method_for(capturing)
    return st, m, oracle_for_state(st, m)


@pytest.mark.parametrize("t", [0.0, 0.13, 0.4])
@pytest.mark.parametrize("capturing", ["null", "persson_br0", "persson_dd"])
@pytest.mark.parametrize("flux", ["rusanov", "hllc"])
@pytest.mark.parametrize("p", [1, 3])
def test_split_residual_fp64(p, flux, capturing, t):
    st, m, orc = _case(p, capturing, flux, "fp64")
    st.t = t
    du = np.zeros(st.u.shape, order="F")
    residual(du, st, EQ, m)
    err = rel_err_per_component(du, orc.residual(st.u, t))
    print(f"split fp64 p={p} {flux} {capturing} t={t}: {err:.3e}", flush=True)
    assert err <= 1e-12


@pytest.mark.parametrize("hi_bottom", ["reflecting", "dirichlet"])
@pytest.mark.parametrize("t", [0.0, 0.4])
@pytest.mark.parametrize("p", [2, 3, 4])
def test_split_residual_mixed(p, t, hi_bottom):
    st, m, orc = _case(p, "persson_br0", "rusanov", "fp32_mixed", hi_bottom=hi_bottom)
    st.t = t
    du = np.zeros(st.u.shape, order="F")
    residual(du, st, EQ, m)
    err = rel_err_per_component(du, orc.residual(st.u, t))
    print(f"split mixed p={p} t={t} {hi_bottom}: {err:.3e}", flush=True)
    assert err <= 1e-4


@pytest.mark.parametrize("p,nx,ny", [(1, 12, 4), (1, 24, 6), (2, 24, 6)])
def test_double_mach_matches_oracle(p, nx, ny):
    """The reduced double-Mach reflection (Cases2D_Optional.jl:39-140) on the device and through
    the oracle's time loop: same status and steps, states within 1e-11.  The top's switch point
    moves with the stage times of every step."""
    kw = dict(p=p, nx=nx, ny=ny, t_final=0.03, cfl=0.04, Lx=1.2, Ly=0.4)
    case_d, st_d, _ = run_double_mach_reflection(**kw)
    case_o, st_o, _ = run_double_mach_reflection(**kw, integrator=oracle_integrator_2d)
    err = rel_err_per_component(st_d.u, st_o.u)
    print(f"dmr p={p} {nx}x{ny}: steps {case_d['metrics']['n_steps']} "
          f"(oracle {case_o['metrics']['n_steps']}) pass {case_d['pass']} rel err {err:.3e}", flush=True)
    assert case_d["metrics"]["n_steps"] == case_o["metrics"]["n_steps"]
    assert (case_d["pass"], case_d["diverged"], case_d["positivity_ok"]) == \
        (case_o["pass"], case_o["diverged"], case_o["positivity_ok"])
    assert case_d["pass"]
    assert err <= 1e-11


def test_double_mach_mixed_follows_fp64():
    kw = dict(p=2, nx=24, ny=6, t_final=0.03, cfl=0.04, Lx=1.2, Ly=0.4)
    case_d, st_d, _ = run_double_mach_reflection(**kw)
    case_m, st_m, _ = run_double_mach_reflection(**kw, precision="fp32_mixed")
    err = rel_err_per_component(st_m.u, st_d.u)
    print(f"dmr mixed vs fp64: rel err {err:.3e}", flush=True)
    assert case_m["metrics"]["n_steps"] == case_d["metrics"]["n_steps"]
    assert case_m["pass"] == case_d["pass"]
    assert err <= 1e-4


def test_optional_suite_on_the_device():
    """run_p33b_optional_suite (Cases2D_Optional.jl:269-330): the reference's smoke tests ask
    that both reduced cases finish (test/test_2d.jl:263-301); the device reports match the
    oracle's."""
    cases, overall, hard = run_p33b_optional_suite()
    cases_o, overall_o, hard_o = run_p33b_optional_suite(integrator=oracle_integrator_2d)
    assert overall and overall_o and hard == hard_o == []
    for c, o in zip(cases, cases_o):
        assert c["name"] == o["name"]
        assert c["metrics"]["n_steps"] == o["metrics"]["n_steps"]
        assert (c["pass"], c["positivity_ok"]) == (o["pass"], o["positivity_ok"])


def test_split_side_checks():
    kw = dict(p=1, nx=8, ny=4, bc_left="dirichlet", bc_right="transmissive")
    dev = DeviceContext(**kw, bc_bottom="ghost_split", bc_top="reflecting")
    try:
        coef = (0.5, 0.0, 0.0, 1.0)
        with pytest.raises(FrfError, match="not a FRF_BC_GHOST_SPLIT side"):
            dev.set_boundary_split("top", coef, U_POST, "reflecting")
        with pytest.raises(FrfError, match="reflecting or Dirichlet"):
            dev.set_boundary_split("bottom", coef, U_POST, "transmissive")
        with pytest.raises(FrfError, match="divisor"):
            dev.set_boundary_split("bottom", (0.5, 0.0, 0.0, 0.0), U_POST, "reflecting")
        with pytest.raises(FrfError, match="no Dirichlet state"):
            dev.set_boundary_split("bottom", coef, U_POST, "dirichlet")
        dev.set_boundary_split("bottom", coef, U_POST, "dirichlet", U_PRE)
    finally:
        dev.close()
