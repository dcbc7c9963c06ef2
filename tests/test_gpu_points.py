"""Gauss-Lobatto-Legendre solution points (SchemeConfig(points = :gll), Operators.jl:85-100) on the
GPU: the device takes whatever operators the host builds, so the GLL residual and step must match
the oracle given the same GLL operators, within the usual bars (FP64 1e-12, FP32-mixed 1e-4)."""
import numpy as np
import pytest

from frforge_amd import (Mesh2D, PeriodicBC, SchemeConfig, TransmissiveBC, allocate_state,
                         build_operators, residual, set_initial_condition, ssp_rk3)
from frforge_amd.cases import riemann2d_ic, vortex_ic

from helpers import EQ, method_for, oracle_for_state, rel_err_per_component

pytestmark = pytest.mark.gpu


def _case(p, capturing, precision, bc="periodic", ic="vortex"):
    ops = build_operators(p, points="gll")
    B = PeriodicBC if bc == "periodic" else TransmissiveBC
    dom = (-5.0, 5.0, -5.0, 5.0) if ic == "vortex" else (0.0, 1.0, 0.0, 1.0)
    mesh = Mesh2D(*dom, 19, 11, left_bc=B(), right_bc=B(), bottom_bc=B(), top_bc=B())
    st = allocate_state(mesh, ops, scheme=SchemeConfig(points="gll"), precision=precision)
    if ic == "vortex":
        set_initial_condition(st, vortex_ic(EQ, x0=0.0, y0=0.0))
    else:
        set_initial_condition(st, lambda x, y: riemann2d_ic(EQ, x, y, config="cfg6"))
    m = method_for(capturing)
    return st, m, oracle_for_state(st, m)


@pytest.mark.parametrize("capturing", ["null", "persson_br0", "persson_dd"])
@pytest.mark.parametrize("ic,bc", [("vortex", "periodic"), ("cfg6", "transmissive")])
@pytest.mark.parametrize("p", [1, 3, 5])
def test_gll_residual_fp64(p, ic, bc, capturing):
    st, m, orc = _case(p, capturing, "fp64", bc, ic)
    assert st.ops.points == "gll" and st.ops.xi[0] == -1.0
    du = np.zeros(st.u.shape, order="F")
    residual(du, st, EQ, m)
    err = rel_err_per_component(du, orc.residual(st.u))
    print(f"gll fp64 p={p} {ic} {capturing}: {err:.3e}", flush=True)
    assert err <= 1e-12


@pytest.mark.parametrize("p", [2, 3, 4])
def test_gll_residual_mixed(p):
    st, m, orc = _case(p, "persson_br0", "fp32_mixed", "transmissive", "cfg6")
    du = np.zeros(st.u.shape, order="F")
    residual(du, st, EQ, m)
    err = rel_err_per_component(du, orc.residual(st.u))
    print(f"gll mixed p={p}: {err:.3e}", flush=True)
    assert err <= 1e-4


def test_gll_run_matches_oracle():
    st, m, orc = _case(3, "persson_br0", "fp64", "transmissive", "cfg6")
    st_o, n_o, t_o, u_o = orc.run(st.u, 0.0, 0.02, cfl=0.08)
    r = ssp_rk3(st, EQ, m, 0.02, cfl=0.08)
    err = rel_err_per_component(st.u, u_o)
    print(f"gll run: n={r['n_steps']} (oracle {n_o}) rel err {err:.3e}", flush=True)
    assert (r["status"], r["n_steps"]) == (st_o, n_o) and r["t"] == t_o
    assert err <= 1e-11
