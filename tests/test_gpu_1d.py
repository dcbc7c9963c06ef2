"""1D Euler path on the GPU (frf1d_*) against the 1D oracle and the reference's Sod fixture."""
import math

import numpy as np
import pytest

from frforge_amd import NullCapturing, PerssonAVMethod, build_operators
from frforge_amd.mesh import PeriodicBC, TransmissiveBC
from frforge_amd.oned import Euler1D, Mesh1D, allocate_state_1d, set_initial_condition_1d, sod_ic
from frforge_amd.sod import run_sod
from frforge_amd.solver1d import compute_dt_1d, residual_1d, ssp_rk3_1d, ssp_rk3_step_1d

from test_oracle1d import FIXTURE, _oracle_for

pytestmark = pytest.mark.gpu

EQ = Euler1D()


def _method(capt):
    if capt == "null":
        return NullCapturing()
    return PerssonAVMethod.create(av_form="conservative_br0" if capt == "persson_br0" else "element_local_DD")


def _state(p, n, bc, kind="sod"):
    ops = build_operators(p)
    B = PeriodicBC if bc == "periodic" else TransmissiveBC
    st = allocate_state_1d(Mesh1D(0.0, 1.0, n, left_bc=B(), right_bc=B()), ops)
    if kind == "sod":
        set_initial_condition_1d(st, lambda x: sod_ic(EQ, x))
    else:
        set_initial_condition_1d(st, lambda x: EQ.primitives_to_conserved(
            1.0 + 0.3 * np.sin(2 * np.pi * x) + 0.4 * (x > 0.5), 0.2 + 0.1 * np.cos(2 * np.pi * x), 1.0 + 0 * x))
    return st


def _rel(a, b):
    return max(np.max(np.abs(a[..., c] - b[..., c])) / max(np.max(np.abs(b[..., c])), 1e-300) for c in range(3))


@pytest.mark.parametrize("capt", ["null", "persson_br0", "persson_dd"])
@pytest.mark.parametrize("bc,kind", [("transmissive", "sod"), ("periodic", "wave")])
@pytest.mark.parametrize("p", [1, 2, 3, 5])
def test_residual_1d_matches_oracle(p, bc, kind, capt):
    st = _state(p, 37, bc, kind)
    m = _method(capt)
    du = np.zeros_like(st.u, order="F")
    residual_1d(du, st, EQ, m)
    ref = _oracle_for(st, EQ, m).residual(st.u)
    err = _rel(du, ref)
    print(f"1D p={p} {bc} {capt}: rel err {err:.3e}", flush=True)
    assert err <= 1e-12


@pytest.mark.parametrize("capt", ["null", "persson_br0"])
def test_run_1d_matches_oracle(capt):
    st = _state(3, 50, "transmissive")
    m = _method(capt)
    st_o, n_o, t_o, u_o = _oracle_for(st, EQ, m).run(st.u, 0.0, 0.05, cfl=0.15)
    r = ssp_rk3_1d(st, EQ, m, 0.05, cfl=0.15)
    assert (r["status"], r["n_steps"]) == (st_o, n_o) and r["t"] == t_o
    assert _rel(st.u, u_o) <= 1e-11


def test_sod_fixture_on_gpu():
    """The reference's snapshot case sod_p2_ne64_persson_av, integrated on the GPU."""
    r = run_sod(p=2, n_elements=64, t_final=0.2)
    m = r["metrics"]
    print({k: m[k] for k in ("n_steps", "u_min", "u_max", "l1_density_vs_exact")}, flush=True)
    assert m["n_steps"] == FIXTURE["n_steps"]
    assert math.isclose(m["u_max"], FIXTURE["u_max"], rel_tol=1e-9)
    assert math.isclose(m["u_min"], FIXTURE["u_min"], rel_tol=1e-9)
    assert math.isclose(m["l1_density_vs_exact"], FIXTURE["l1_density_vs_exact"], rel_tol=1e-9)
    assert math.isclose(r["overshoot"], FIXTURE["overshoot"], rel_tol=1e-9)
    assert math.isclose(r["shock_thickness"], FIXTURE["shock_thickness"], rel_tol=1e-9)
    assert math.isclose(m["excess_dissipation_vs_null"], FIXTURE["excess_dissipation_vs_null"], rel_tol=1e-8)


def test_config1_sod_200_elements_p3():
    """The configuration this path is quoted on: 200 elements, p = 3, Persson + BR0, CFL 0.1 to
    t = 0.2, against the oracle."""
    st = _state(3, 200, "transmissive")
    m = _method("persson_br0")
    st_o, n_o, t_o, u_o = _oracle_for(st, EQ, m).run(st.u, 0.0, 0.2, cfl=0.1)
    r = ssp_rk3_1d(st, EQ, m, 0.2, cfl=0.1)
    print(f"config 1: n={r['n_steps']} (oracle {n_o}) rel err {_rel(st.u, u_o):.3e}", flush=True)
    assert (r["status"], r["n_steps"]) == (st_o, n_o) == ("ok", n_o)
    assert _rel(st.u, u_o) <= 1e-10


def test_step_dt_and_divergence_1d():
    st = _state(2, 10, "transmissive")
    orc = _oracle_for(st, EQ, NullCapturing())
    assert compute_dt_1d(st, EQ, 0.1) == orc.compute_dt(st.u, 0.1)
    st.u[1, 3, 0] = np.nan
    u_before = st.u.copy(order="F")
    assert ssp_rk3_step_1d(st, EQ, NullCapturing(), 1e-3) == "diverged"
    assert st.t == 0.0
    np.testing.assert_array_equal(np.isnan(st.u), np.isnan(u_before))


@pytest.mark.parametrize("capt", ["null", "br0", "dd"])
@pytest.mark.parametrize("p", [1, 3, 5])
def test_residual_1d_dirichlet_matches_oracle(p, capt):
    """Constant Dirichlet ends (Residual.jl:75-79,100-104) on the device against the oracle."""
    from frforge_amd.mesh import DirichletBC
    from helpers import oracle_for_state_1d
    eq = Euler1D(1.4)
    UL = np.asarray(eq.primitives_to_conserved(1.2, 0.3, 1.4))
    UR = np.asarray(eq.primitives_to_conserved(0.2, -0.1, 0.15))
    st = allocate_state_1d(Mesh1D(0.0, 1.0, 23, left_bc=DirichletBC(state=UL), right_bc=DirichletBC(state=UR)),
                           build_operators(p))
    set_initial_condition_1d(st, lambda x: sod_ic(eq, x, x0=0.45))
    m = {"null": NullCapturing(), "br0": PerssonAVMethod.create(c_av=0.2),
         "dd": PerssonAVMethod.create(c_av=0.2, av_form="element_local_DD")}[capt]
    du = np.zeros(st.u.shape, order="F")
    residual_1d(du, st, eq, m)
    ref = oracle_for_state_1d(st, m, eq).residual(st.u)
    err = max(np.max(np.abs(du[:, :, c] - ref[:, :, c])) / max(np.max(np.abs(ref[:, :, c])), 1e-300) for c in range(3))
    print(f"1d dirichlet p={p} {capt}: {err:.3e}", flush=True)
    assert err <= 1e-12


def test_bc_freestream_cases_on_the_device():
    """test_euler.jl:55-63 on the device: the uniform state stays uniform to 1e-12 between
    transmissive and between Dirichlet ends, with the oracle's step counts."""
    from frforge_amd.cases1d import run_bc_dirichlet_test, run_bc_transmissive_test
    from helpers import oracle_integrator_1d
    for run in (run_bc_transmissive_test, run_bc_dirichlet_test):
        c = run()
        o = run(integrator=oracle_integrator_1d)
        print(f"{c['name']}: dev {c['metrics']['max_deviation_from_uniform']:.3e} "
              f"steps {c['metrics']['n_steps']} (oracle {o['metrics']['n_steps']})", flush=True)
        assert c["pass"] and c["metrics"]["max_deviation_from_uniform"] < 1e-12
        assert c["metrics"]["n_steps"] == o["metrics"]["n_steps"]


def test_m3_euler_suite_on_the_device():
    """run_m3_euler_suite (Cases.jl:702-745) on the device: every case passes (test_euler.jl:29-63
    gates), and the order studies' L2 errors equal the oracle's to 1e-10 relative."""
    from frforge_amd.cases1d import run_m3_euler_suite
    from helpers import oracle_integrator_1d
    cases, overall, hard = run_m3_euler_suite()
    cases_o, overall_o, _ = run_m3_euler_suite(integrator=oracle_integrator_1d)
    assert overall and overall_o and hard == []
    for c, o in zip(cases, cases_o):
        assert c["name"] == o["name"] and c["pass"] == o["pass"]
        if "l2_errors" in c:
            print(f"{c['name']}: orders {c['observed_orders']}", flush=True)
            assert np.allclose(c["l2_errors"], o["l2_errors"], rtol=1e-10, atol=0)
            assert all(q >= c["formal_order"] - 0.3 for q in c["observed_orders"])
