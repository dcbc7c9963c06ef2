"""Curved metrics on the GPU against the oracle given the same metric arrays (whose curved path
tests/test_oracle_curved.py pins): per-stage residuals on wavy meshes within 1e-12 (FP64) and 1e-4
(FP32-mixed), with periodic, transmissive and reflecting sides and a solid block; the curved cases
run step for step; and the checks of frf_set_curved_metrics."""
import numpy as np
import pytest

from frforge_amd import SchemeConfig, allocate_state, build_operators, residual, set_initial_condition, ssp_rk3
from frforge_amd.cases import vortex_ic
from frforge_amd.cases2d_curved import (run_euler2d_curved_discontinuous, run_freestream_preservation_2d,
                                        run_p32_curved_suite)
from frforge_amd.device import FrfError
from frforge_amd.equations import primitives_to_conserved
from frforge_amd.mesh import ReflectingBC, TransmissiveBC, make_wavy_mesh2d

from helpers import (EQ, method_for, oracle_for_state, oracle_integrator_2d, oracle_residual,
                     rel_err_per_component)

pytestmark = pytest.mark.gpu


def _smooth(x, y):
    rho = 1.0 + 0.2 * np.sin(2 * np.pi * x) * np.cos(2 * np.pi * y)
    u = 0.4 + 0.2 * np.cos(2 * np.pi * y)
    v = -0.3 + 0.1 * np.sin(2 * np.pi * x)
    p = 1.0 + 0.1 * np.cos(2 * np.pi * (x + y))
    return primitives_to_conserved(EQ, rho, u, v, p)


def _case(p, capturing, flux, sides="periodic", nx=13, ny=9, amp=0.05, solid=None, noise=0.0):
    ops = build_operators(p)
    if sides == "periodic":
        mesh = make_wavy_mesh2d(nx, ny, ops, amp=amp)
    else:
        mesh = make_wavy_mesh2d(nx, ny, ops, amp=amp, left_bc=TransmissiveBC(), right_bc=TransmissiveBC(),
                                bottom_bc=ReflectingBC(), top_bc=ReflectingBC())
        mesh.solid = solid
    st = allocate_state(mesh, ops, scheme=SchemeConfig(flux=flux), wavy_amp=amp)
    set_initial_condition(st, _smooth)
    if noise:
        st.u[:, :, :, 0] *= 1.0 + noise * np.random.default_rng(3).standard_normal(st.u.shape[:3])
    m = method_for(capturing)
    return st, m, oracle_for_state(st, m)


@pytest.mark.parametrize("capturing", ["null", "persson_br0", "persson_dd"])
@pytest.mark.parametrize("flux", ["rusanov", "hllc"])
@pytest.mark.parametrize("sides", ["periodic", "walls"])
@pytest.mark.parametrize("p", [1, 3, 5])
def test_curved_residual_fp64(p, sides, flux, capturing):
    st, m, orc = _case(p, capturing, flux, sides, noise=0.05)
    du = np.zeros(st.u.shape, order="F")
    residual(du, st, EQ, m)
    err = rel_err_per_component(du, orc.residual(st.u))
    print(f"curved fp64 p={p} {sides} {flux} {capturing}: {err:.3e}", flush=True)
    assert err <= 1e-12


@pytest.mark.parametrize("p", [2, 4])
def test_curved_residual_with_a_solid_block(p):
    mask = np.zeros((9, 13), dtype=bool)
    mask[3:6, 5:8] = True
    st, m, orc = _case(p, "persson_br0", "hllc", "walls", solid=mask.ravel(), noise=0.05)
    du = np.zeros(st.u.shape, order="F")
    residual(du, st, EQ, m)
    err = rel_err_per_component(du, orc.residual(st.u))
    print(f"curved solid p={p}: {err:.3e}", flush=True)
    assert err <= 1e-12 and not np.any(du[:, :, mask.ravel(), :])


def test_curved_run_matches_oracle():
    st, m, orc = _case(3, "persson_br0", "rusanov", "periodic")
    st_o, n_o, t_o, u_o = orc.run(st.u, 0.0, 0.03, cfl=0.1)
    r = ssp_rk3(st, EQ, m, 0.03, cfl=0.1)
    err = rel_err_per_component(st.u, u_o)
    print(f"curved run: n={r['n_steps']} (oracle {n_o}) rel err {err:.3e}", flush=True)
    assert (r["status"], r["n_steps"]) == (st_o, n_o) and r["t"] == t_o
    assert err <= 1e-11


@pytest.mark.parametrize("method_name", ["null", "persson_av"])
def test_curved_jump_matches_oracle(method_name):
    from frforge_amd import NullCapturing, PerssonAVMethod
    m = NullCapturing() if method_name == "null" else PerssonAVMethod.create(c_av=0.1)
    kw = dict(p=1, nx=12, ny=12, t_final=0.03, amp=0.01, cfl=0.04, method=m, method_name=method_name)
    c_d = run_euler2d_curved_discontinuous(**kw)
    c_o = run_euler2d_curved_discontinuous(**kw, integrator=oracle_integrator_2d)
    print(f"curved jump {method_name}: steps {c_d['metrics']['n_steps']} (oracle {c_o['metrics']['n_steps']})",
          flush=True)
    assert c_d["metrics"]["n_steps"] == c_o["metrics"]["n_steps"]
    assert c_d["pass"] and c_o["pass"]


def test_curved_freestream_and_suite_on_the_device():
    """The reference's free-stream bars (test/test_2d.jl:97-131) and the curved suite without its
    scalar-advection case, on the device; the residual maxima equal the oracle's to 1e-12."""
    c1 = run_freestream_preservation_2d(p=3, nx=4, ny=4, curved=True, amp=0.05, t_final=0.05)
    o1 = run_freestream_preservation_2d(p=3, nx=4, ny=4, curved=True, amp=0.05, t_final=0.05,
                                        integrator=oracle_integrator_2d, residual_fn=oracle_residual)
    print(f"freestream curved: res {c1['metrics']['residual_max']:.3e} "
          f"(oracle {o1['metrics']['residual_max']:.3e}) err {c1['metrics']['freestream_error']:.3e}", flush=True)
    assert c1["pass"] and c1["metrics"]["residual_max"] < 5e-4
    assert abs(c1["metrics"]["residual_max"] - o1["metrics"]["residual_max"]) <= 1e-12 * max(
        1.0, o1["metrics"]["residual_max"])
    cases, overall, hard = run_p32_curved_suite()
    assert overall and hard == []


@pytest.mark.parametrize("capturing", ["null", "persson_br0", "persson_dd"])
@pytest.mark.parametrize("sides", ["periodic", "walls"])
@pytest.mark.parametrize("p", [2, 3, 4])
def test_curved_residual_mixed(p, sides, capturing):
    """The FP32-mixed form on curved metrics (the general-metric kernel in FP32, FP64 state):
    within 1e-4 of the oracle normwise."""
    ops = build_operators(p)
    st_ref, m, orc = _case(p, capturing, "rusanov", sides, noise=0.05)
    st = allocate_state(st_ref.mesh, ops, precision="fp32_mixed", wavy_amp=0.05)
    st.u[...] = st_ref.u
    du = np.zeros(st.u.shape, order="F")
    residual(du, st, EQ, m)
    err = rel_err_per_component(du, orc.residual(st.u))
    print(f"curved mixed p={p} {sides} {capturing}: {err:.3e}", flush=True)
    assert err <= 1e-4


def test_curved_mixed_run_follows_fp64():
    st, m, _ = _case(3, "persson_br0", "rusanov", "periodic")
    st_m = allocate_state(st.mesh, st.ops, precision="fp32_mixed", wavy_amp=0.05)
    st_m.u[...] = st.u
    r = ssp_rk3(st, EQ, m, 0.03, cfl=0.1)
    r_m = ssp_rk3(st_m, EQ, m, 0.03, cfl=0.1)
    err = rel_err_per_component(st_m.u, st.u)
    print(f"curved mixed run: n={r_m['n_steps']} (fp64 {r['n_steps']}) rel err {err:.3e}", flush=True)
    assert (r_m["status"], r_m["n_steps"]) == (r["status"], r["n_steps"])
    assert err <= 1e-5


def test_curved_metric_checks():
    from frforge_amd.mesh import build_mesh_metrics
    st, m, _ = _case(2, "null", "rusanov")
    # discrete metrics of a wavy mesh carry a curved h_char the device does not keep
    st_d = allocate_state(st.mesh, st.ops, metrics=build_mesh_metrics(st.mesh, st.ops))
    set_initial_condition(st_d, _smooth)
    with pytest.raises(NotImplementedError, match="h_char"):
        residual(np.zeros(st_d.u.shape, order="F"), st_d, EQ, m)
    # the learned sensor lives in the affine mixed kernel
    from frforge_amd.device import DeviceContext
    dev = DeviceContext(p=3, nx=4, ny=4, bc_left="periodic", bc_right="periodic", bc_bottom="periodic",
                        bc_top="periodic", precision="fp32_mixed", capturing="persson_br0")
    try:
        ops = build_operators(3)
        dev.set_operators(ops.xi, ops.w, ops.D, ops.ell_L, ops.ell_R, ops.gL_xi, ops.gR_xi, ops.V_legendre)
        mesh = make_wavy_mesh2d(4, 4, ops, amp=0.05)
        dev.set_geometry(mesh.x_vertices, mesh.y_vertices)
        st_w = allocate_state(mesh, ops, wavy_amp=0.05)
        dev.set_curved_metrics(st_w.metrics)
        from frforge_amd.learned import MLPWeights
        with pytest.raises(FrfError, match="affine FP32-mixed kernel"):
            dev.set_learned_sensor(MLPWeights.init(seed=0).pack())
    finally:
        dev.close()
