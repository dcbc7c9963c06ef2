"""The batched ensemble at its full size on the GPU: 64 seeded instances (meshes 64^2..512^2,
p 2..5, Rusanov or HLLC, vortex or four-quadrant Riemann) stepped together by the batched
FP32-mixed form with autotuned tiles.  Every instance's status and step count equal those of the
same fixed-dt steps in the FP64 plain form, and every instance that completes its steps stays
within density L1 1e-6 (relative, quadrature-weighted) of FP64."""
import pytest

from frforge_amd.ensemble import run_ensemble

pytestmark = pytest.mark.gpu


@pytest.mark.timeout(600)
def test_ensemble_matches_fp64():
    r = run_ensemble(8)
    e = r["density_rel_l1_vs_fp64"]
    print(f"ensemble: {r['instances']} instances, {r['dof'] / 1e6:.0f} M DOF, completed {r['completed']}, "
          f"diverged {r['diverged']}, status and steps equal to FP64: {r['status_and_steps_match_fp64']}, "
          f"density L1 vs FP64 max {e['max']:.3e} median {e['median']:.3e}, "
          f"{r['dof_stage_per_s']:.3e} DOF-stage/s", flush=True)
    assert r["status_and_steps_match_fp64"] == r["instances"] == 64
    assert r["completed"] + r["diverged"] == 64 and e["instances"] == r["completed"] > 0
    assert e["max"] <= 1e-6
