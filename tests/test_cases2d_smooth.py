"""2D Euler verification cases on Cartesian meshes through the oracle's time loop (no GPU): the
reference's own tests of them (test/test_2d.jl:38-95) and the P3.1 capturing suite
(Cases2D_Smooth.jl:625-676)."""
import numpy as np

from frforge_amd import PerssonAVMethod
from frforge_amd.cases2d_smooth import (run_euler2d_discontinuous, run_euler2d_smooth_order,
                                        run_p31_2d_capturing_suite)

from helpers import oracle_integrator_2d


def test_discontinuous_euler_runs():
    """test_2d.jl:38-42 and :68-81: the density jump with NullCapturing and with Persson AV."""
    c, st, _ = run_euler2d_discontinuous(p=1, nx=12, ny=12, t_final=0.05, integrator=oracle_integrator_2d)
    assert c["pass"] and c["positivity_ok"]
    assert c["name"] == "euler2d_density_jump_p1_null" and c["metrics"]["ρ_min0"] == 0.125
    c, _, _ = run_euler2d_discontinuous(p=1, nx=12, ny=12, t_final=0.05, cfl=0.08,
                                        method=PerssonAVMethod.create(c_av=0.1), method_name="persson_av",
                                        integrator=oracle_integrator_2d)
    assert c["pass"] and c["positivity_ok"] and not c["diverged"]
    assert 0.0 <= c["overshoot"] < 0.1


def test_persson_smooth_stability():
    """test_2d.jl:83-95: the smooth density wave with Persson AV stays stable, positive and
    conservative."""
    c = run_euler2d_smooth_order(p=2, n_list=(8, 16), t_final=0.15, cfl=0.05,
                                 method=PerssonAVMethod.create(c_av=0.1), method_name="persson_av",
                                 integrator=oracle_integrator_2d)
    assert not c["diverged"] and c["positivity_ok"] and c["conservation_pass"]
    assert np.all(np.isfinite(c["l2_errors"])) and c["observed_orders"][0] > 2.0


def test_p31_capturing_suite():
    cases, overall, hard = run_p31_2d_capturing_suite(integrator=oracle_integrator_2d)
    assert overall and hard == []
    assert [c["name"] for c in cases] == ["euler2d_density_jump_p1_null", "euler2d_density_jump_p1_persson_av",
                                          "euler2d_smooth_order_p2_persson_av"]
