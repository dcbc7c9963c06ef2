"""Solid-element masks on the GPU against the oracle (whose solid faces tests/test_oracle_walls.py
pins): per-stage residuals of the FP64 form within 1e-12 and of the FP32-mixed form within 1e-4
(normwise per component), solid elements' du exactly zero, the reduced forward-facing step run
step for step, and the C ABI's rejection of masks it does not implement."""
import numpy as np
import pytest

from frforge_amd import (Mesh2D, SchemeConfig, TransmissiveBC, allocate_state, build_operators,
                         residual, set_initial_condition, ssp_rk3)
from frforge_amd.cases2d_optional import make_ffs_solid_mask, run_forward_facing_step
from frforge_amd.device import DeviceContext, FrfError
from frforge_amd.equations import primitives_to_conserved
from frforge_amd.mesh import DirichletBC, ReflectingBC

from helpers import EQ, method_for, oracle_for_state, oracle_integrator_2d, rel_err_per_component

pytestmark = pytest.mark.gpu

NX, NY = 37, 21
U_IN = primitives_to_conserved(EQ, 1.4, 0.9, 0.05, 1.1)


def _ic(x, y):
    rho = np.where(x < 0.3 + 0.2 * y, 2.0, 1.0) + 0.05 * np.sin(3 * x + 2 * y)
    u = 0.3 * np.cos(np.pi * y) + 0.1
    v = -0.2 * np.sin(np.pi * x)
    p = np.where(x < 0.3 + 0.2 * y, 2.0, 1.0) * (1.0 + 0.05 * np.cos(2 * (x + y)))
    return primitives_to_conserved(EQ, rho, u, v, p)


def _blocks_mask():
    """Solids across the tile edges of every tile shape (x multiples of 8 and 16, y multiples of
    4 and 8), isolated elements, and elements on the domain sides and corners."""
    m = np.zeros((NY, NX), dtype=bool)
    m[3:7, 8] = True           # column across an x tile edge
    m[8, 14:21] = True         # bar along a y tile edge
    m[16:18, 15:17] = True     # 2x2 block across both
    m[15, 30] = True           # isolated
    m[20, 0] = m[0, 36] = m[0, 20] = m[10, 36] = True   # corners and sides
    return m.ravel()


MASKS = {"step": make_ffs_solid_mask(NX, NY, 1.0, 1.0, step_x=0.55, step_h=0.3),
         "blocks": _blocks_mask()}


def _case(p, mask, capturing, flux, precision):
    mesh = Mesh2D(0.0, 1.0, 0.0, 1.0, NX, NY, left_bc=DirichletBC(state=U_IN),
                  right_bc=TransmissiveBC(), bottom_bc=ReflectingBC(), top_bc=ReflectingBC(),
                  solid=MASKS[mask])
    st = allocate_state(mesh, build_operators(p), scheme=SchemeConfig(flux=flux), precision=precision)
    set_initial_condition(st, _ic)
    m = method_for(capturing)
    return st, m, oracle_for_state(st, m)


@pytest.mark.parametrize("capturing", ["null", "persson_br0", "persson_dd"])
@pytest.mark.parametrize("flux", ["rusanov", "hllc"])
@pytest.mark.parametrize("mask", ["step", "blocks"])
@pytest.mark.parametrize("p", [1, 3, 5])
def test_solid_residual_fp64(p, mask, flux, capturing):
    st, m, orc = _case(p, mask, capturing, flux, "fp64")
    du = np.zeros(st.u.shape, order="F")
    residual(du, st, EQ, m)
    ref = orc.residual(st.u)
    err = rel_err_per_component(du, ref)
    print(f"solid fp64 p={p} {mask} {flux} {capturing}: {err:.3e}", flush=True)
    assert err <= 1e-12
    assert not np.any(du[:, :, MASKS[mask], :])


@pytest.mark.parametrize("capturing", ["null", "persson_br0"])
@pytest.mark.parametrize("mask", ["step", "blocks"])
@pytest.mark.parametrize("p", [2, 3, 4])
def test_solid_residual_mixed(p, mask, capturing):
    st, m, orc = _case(p, mask, capturing, "rusanov", "fp32_mixed")
    du = np.zeros(st.u.shape, order="F")
    residual(du, st, EQ, m)
    err = rel_err_per_component(du, orc.residual(st.u))
    print(f"solid mixed p={p} {mask} {capturing}: {err:.3e}", flush=True)
    assert err <= 1e-4
    assert not np.any(du[:, :, MASKS[mask], :])


# the reduced FFS of the reference's smoke test (test/test_2d.jl:282-301).  At the case's own
# defaults (24 x 8, M_in = 3, c_av = 0.15), and at p >= 3 in the smoke set-up, the run blows up
# in the oracle as on the device, so the runs compare p = 1 and 2.
FFS_SMOKE = dict(t_final=0.02, cfl=0.025, Lx=1.0, Ly=0.4, step_x=0.3, step_h=0.1, M_in=2.0)


def _ffs(p, nx=12, ny=4, **kw):
    from frforge_amd import PerssonAVMethod
    return run_forward_facing_step(p=p, nx=nx, ny=ny, method=PerssonAVMethod.create(c_av=0.25),
                                   **FFS_SMOKE, **kw)


@pytest.mark.parametrize("p,nx,ny", [(1, 12, 4), (2, 12, 4), (2, 24, 8)])
def test_forward_facing_step_matches_oracle(p, nx, ny):
    """The reduced FFS case (Cases2D_Optional.jl:169-267) on the device and through the oracle's
    time loop: same status and step count, states within 1e-11, solid elements untouched; the
    reference's smoke test asks for a run that does not diverge."""
    case_d, st_d, _ = _ffs(p, nx, ny)
    case_o, st_o, _ = _ffs(p, nx, ny, integrator=oracle_integrator_2d)
    err = rel_err_per_component(st_d.u, st_o.u)
    print(f"ffs p={p} {nx}x{ny}: steps {case_d['metrics']['n_steps']} (oracle {case_o['metrics']['n_steps']}) "
          f"pass {case_d['pass']} rel err {err:.3e}", flush=True)
    assert case_d["metrics"]["n_steps"] == case_o["metrics"]["n_steps"]
    assert (case_d["pass"], case_d["diverged"], case_d["positivity_ok"]) == \
        (case_o["pass"], case_o["diverged"], case_o["positivity_ok"])
    assert err <= 1e-11
    assert case_d["pass"] and not case_d["nan_detected"] and case_d["metrics"]["n_solid"] > 0
    solid = st_d.mesh.solid
    U_in = st_d.u[0, 0, solid, :][0]
    assert np.array_equal(st_d.u[:, :, solid, :], np.broadcast_to(U_in, st_d.u[:, :, solid, :].shape))


def test_forward_facing_step_mixed_follows_fp64():
    case_d, st_d, _ = _ffs(2, 24, 8)
    case_m, st_m, _ = _ffs(2, 24, 8, precision="fp32_mixed")
    err = rel_err_per_component(st_m.u, st_d.u)
    print(f"ffs mixed vs fp64: rel err {err:.3e}", flush=True)
    assert case_m["metrics"]["n_steps"] == case_d["metrics"]["n_steps"]
    assert case_m["pass"] == case_d["pass"]
    assert err <= 1e-4


def test_solid_run_fp64_matches_oracle():
    st, m, orc = _case(3, "blocks", "persson_br0", "hllc", "fp64")
    st_o, n_o, t_o, u_o = orc.run(st.u, 0.0, 0.02, cfl=0.08)
    r = ssp_rk3(st, EQ, m, 0.02, cfl=0.08)
    err = rel_err_per_component(st.u, u_o)
    print(f"solid run: n={r['n_steps']} (oracle {n_o}) rel err {err:.3e}", flush=True)
    assert (r["status"], r["n_steps"]) == (st_o, n_o) and r["t"] == t_o
    assert err <= 1e-11


def test_solid_mask_rejections():
    dev = DeviceContext(p=2, nx=8, ny=4, bc_left="periodic", bc_right="periodic",
                        bc_bottom="transmissive", bc_top="transmissive")
    try:
        with pytest.raises(FrfError, match="non-periodic"):
            dev.set_solid_mask(np.zeros(32, dtype=bool))
    finally:
        dev.close()
    dev = DeviceContext(p=2, nx=8, ny=4, bc_left="transmissive", bc_right="transmissive",
                        bc_bottom="reflecting", bc_top="reflecting")
    try:
        with pytest.raises(FrfError, match="n_elements=32"):
            dev.set_solid_mask(np.zeros(31, dtype=bool))
        dev.set_solid_mask(np.zeros(32, dtype=bool))
        dev.set_solid_mask(None)
    finally:
        dev.close()


@pytest.mark.parametrize("precision", ["fp64", "fp32_mixed"])
def test_reflecting_wall_keeps_freestream(precision):
    """The reference's freestream wall test (test/test_2d.jl:242-261): uniform flow along a
    reflecting bottom, Dirichlet inflow, transmissive right and top, p = 1 on 4 x 2 elements;
    after t = 0.05 the density stays within 0.05 of 1 and the state is positive."""
    from frforge_amd import NullCapturing, integrate
    from frforge_amd.state import positivity_ok
    U0 = primitives_to_conserved(EQ, 1.0, 0.5, 0.0, 1.0)
    mesh = Mesh2D(0.0, 1.0, 0.0, 0.5, 4, 2, left_bc=DirichletBC(state=U0), right_bc=TransmissiveBC(),
                  bottom_bc=ReflectingBC(), top_bc=TransmissiveBC())
    st = allocate_state(mesh, build_operators(1), precision=precision)
    set_initial_condition(st, lambda x, y: U0)
    r = integrate(st, EQ, NullCapturing(), 0.05, cfl=0.1)
    err = np.max(np.abs(st.u[:, :, :, 0] - 1.0))
    print(f"freestream wall {precision}: {r['n_steps']} steps, max |rho - 1| {err:.3e}", flush=True)
    assert r["status"] == "ok" and positivity_ok(EQ, st)
    assert err < 0.05
