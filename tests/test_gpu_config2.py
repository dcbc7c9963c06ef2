"""The isentropic-vortex configuration at its full size on the GPU: periodic [-5,5]^2, 256 x 256
quads, p = 4, vortex beta = 5 at the origin, free stream (1, 1, 1, 1), Rusanov, NullCapturing.

* per-stage residual of the FP64 plain form against the oracle on identical states (the initial
  state and the three stage states of a step): <= 1e-12 normwise per conserved component;
* the whole run to t = 1 at CFL 0.08 in FP64 and in FP32-mixed: density L2 error against the
  exact (translated) vortex stated, mixed against FP64 density L1 <= 1e-6.
"""
import math

import numpy as np
import pytest

from frforge_amd import (Mesh2D, NullCapturing, allocate_state, build_operators, l2_error, residual,
                         set_initial_condition, ssp_rk3)
from frforge_amd.cases import vortex_ic
from oracle.oracle import Oracle

from helpers import EQ, oracle_ops, rel_err_per_component

pytestmark = pytest.mark.gpu

N, P = 256, 4


def _state(precision="fp64"):
    st = allocate_state(Mesh2D(-5.0, 5.0, -5.0, 5.0, N, N), build_operators(P), precision=precision)
    set_initial_condition(st, vortex_ic(EQ, x0=0.0, y0=0.0))
    return st


@pytest.mark.timeout(300)
def test_per_stage_residual_256_p4_matches_oracle():
    st = _state()
    m = st.mesh
    orc = Oracle(P, N, N, m.x_vertices, m.y_vertices, oracle_ops(st.ops))
    dt = 0.5 * orc.compute_dt(st.u, 0.08)
    u0 = st.u.copy(order="F")
    # stage states of one SSP-RK3 step, formed from the device's own residuals
    states = [u0]
    du = np.zeros(st.u.shape, order="F")
    worst = 0.0
    for k in range(3):
        st.u[...] = states[-1]
        residual(du, st, EQ, NullCapturing())
        ref = orc.residual(states[-1])
        err = rel_err_per_component(du, ref)
        worst = max(worst, err)
        print(f"256^2 p=4 vortex stage {k + 1} residual rel err {err:.3e}", flush=True)
        if k == 0:
            states.append(u0 + dt * du)
        elif k == 1:
            states.append(0.75 * u0 + 0.25 * (states[-1] + dt * du))
    assert worst <= 1e-12


@pytest.mark.timeout(300)
def test_run_to_t1_fp64_and_mixed():
    out = {}
    for prec in ("fp64", "fp32_mixed"):
        st = _state(prec)
        r = ssp_rk3(st, EQ, NullCapturing(), 1.0, cfl=0.08)
        assert r["status"] == "ok" and r["t"] == 1.0
        l2 = l2_error(st, vortex_ic(EQ, t=1.0, x0=0.0, y0=0.0), 0)
        out[prec] = (r["n_steps"], l2, st.u[:, :, :, 0].copy())
        print(f"{prec}: {r['n_steps']} steps to t=1, density L2 error vs exact {l2:.6e}", flush=True)
        st.device.close()
        st.device = None
    (n64, e64, r64), (n32, e32, r32) = out["fp64"], out["fp32_mixed"]
    l1 = float(np.sum(np.abs(r32 - r64)) / np.sum(np.abs(r64)))
    print(f"mixed vs fp64 density L1 {l1:.3e}, L2-error difference {abs(e32 - e64) / e64:.3e}", flush=True)
    assert abs(n32 - n64) <= 1 and l1 <= 1e-6
    assert e64 < 1e-3
