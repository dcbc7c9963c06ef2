"""The throughput workload at its full size on the GPU: the seeded four-quadrant Riemann problem
(seed 2026) on [0,1]^2, 1024 x 1024 quads, p = 3 GL, Rusanov, Persson AV (conservative BR0,
c_av = 0.1), transmissive sides; density perturbed by 1 % noise so that every element carries
structure and the sensor fires across the whole grid.

* per-stage residual against the oracle on identical states (the perturbed state and the two
  intermediate states of one SSP-RK3 step, formed from the device's own FP64 residuals): FP64
  plain <= 1e-12 and FP32-mixed <= 1e-4 normwise per conserved component;
* the learned sensor on the same mesh: residual within 1e-3 of the oracle's residual with the
  viscosity of the host network's sigma (the bar of the small learned-sensor test);
* properties independent of size, on the same mesh with periodic sides: a shift of the state by
  whole elements shifts the residual (to 1e-10 in FP64 and 1e-5 in FP32-mixed: the metrics come
  from D applied to the node coordinates, as in the reference, so J varies by roundoff, 7.5e-12
  relative at this size, from element to element), and the weighted sum
  sum_e sum_ij w_i w_j J du_c vanishes to roundoff (1e-12 of the sum of magnitudes in FP64,
  1e-7 in FP32-mixed), since the common face fluxes telescope;
* a uniform state has a residual of exactly zero in the FP32-mixed form and the oracle's
  free-stream roundoff in the FP64 form.
"""
import numpy as np
import pytest

from frforge_amd import allocate_state, residual

from helpers import EQ, make_case, rel_err_per_component

pytestmark = pytest.mark.gpu

N, P, SEED = 1024, 3, 2026


def _free(st):
    if st.device is not None:
        st.device.close()
        st.device = None


def _mixed_twin(st):
    tw = allocate_state(st.mesh, st.ops, scheme=st.scheme, precision="fp32_mixed")
    tw.u[...] = st.u
    return tw


@pytest.fixture(scope="module")
def transmissive_case():
    st, m, orc = make_case(P, N, N, bc="transmissive", capturing="persson_br0", ic="riemann", seed=SEED,
                           noise=0.01)
    yield st, m, orc
    _free(st)


@pytest.mark.timeout(600)
def test_per_stage_residual_matches_oracle(transmissive_case):
    st, m, orc = transmissive_case
    mixed = _mixed_twin(st)
    dt = 0.5 * orc.compute_dt(st.u, 0.08)
    u0 = st.u.copy(order="F")
    states = [u0]
    du = np.zeros(st.u.shape, order="F")
    du32 = np.zeros(st.u.shape, order="F")
    worst64 = worst32 = 0.0
    for k in range(3):
        st.u[...] = states[-1]
        mixed.u[...] = states[-1]
        residual(du, st, EQ, m)
        residual(du32, mixed, EQ, m)
        ref = orc.residual(states[-1])
        e64, e32 = rel_err_per_component(du, ref), rel_err_per_component(du32, ref)
        worst64, worst32 = max(worst64, e64), max(worst32, e32)
        print(f"{N}^2 p={P} stage {k + 1}: fp64 rel err {e64:.3e}, fp32_mixed {e32:.3e}", flush=True)
        if k == 0:
            states.append(u0 + dt * du)
        elif k == 1:
            states.append(0.75 * u0 + 0.25 * (states[-1] + dt * du))
    st.u[...] = u0
    _free(mixed)
    assert worst64 <= 1e-12 and worst32 <= 1e-4


@pytest.mark.timeout(600)
def test_learned_residual_matches_oracle(transmissive_case):
    from frforge_amd.learned import LearnedSensorMethod, element_modes, features
    from oracle.oracle import Oracle
    from helpers import oracle_ops
    st, _, _ = transmissive_case
    mixed = _mixed_twin(st)
    m = LearnedSensorMethod.create()
    du = np.zeros(st.u.shape, order="F")
    residual(du, mixed, EQ, m)
    _free(mixed)
    sigma = m.weights().predict_bf16(features(element_modes(st.u, st.ops.V_legendre))).astype(np.float64)
    me = st.mesh
    orc = Oracle(P, N, N, me.x_vertices, me.y_vertices, oracle_ops(st.ops), bc_x="transmissive",
                 bc_y="transmissive", capturing="null", c_av=0.1)
    ref = orc.apply_br0(st.u, orc.element_viscosities(st.u, sigma), orc.residual(st.u))
    err = rel_err_per_component(du, ref)
    print(f"{N}^2 p={P} learned residual vs oracle with host-network viscosity: {err:.3e}, "
          f"sigma > 0 in {np.mean(sigma > 0):.3f} of elements", flush=True)
    assert err <= 1e-3


def _roll_elements(a, sx, sy):
    s = a.shape
    b = np.reshape(a, (s[0], s[1], N, N, s[3]))  # element e = jy * nx + jx
    return np.asfortranarray(np.reshape(np.roll(b, (sy, sx), axis=(2, 3)), s))


@pytest.mark.timeout(600)
@pytest.mark.parametrize("precision", ["fp64", "fp32_mixed"])
def test_periodic_shift_and_conservation(precision):
    st, m, _ = make_case(P, N, N, bc="periodic", capturing="persson_br0", ic="riemann", seed=SEED,
                         noise=0.01, precision=precision)
    du = np.zeros(st.u.shape, order="F")
    residual(du, st, EQ, m)
    sx, sy = 37, 411
    st.u[...] = _roll_elements(st.u, sx, sy)
    du_s = np.zeros(st.u.shape, order="F")
    residual(du_s, st, EQ, m)
    _free(st)
    shift = rel_err_per_component(du_s, _roll_elements(du, sx, sy))
    ww = np.outer(st.ops.w, st.ops.w)[:, :, None]
    rel = []
    for c in range(4):
        t = ww * st.metrics.J * du[:, :, :, c]
        rel.append(abs(float(np.sum(t))) / float(np.sum(np.abs(t))))
    print(f"{N}^2 p={P} periodic {precision}: shifted residual rel err {shift:.3e}; "
          f"weighted residual sums / sums of magnitudes {[f'{r:.2e}' for r in rel]}", flush=True)
    assert shift <= (1e-10 if precision == "fp64" else 1e-5)
    assert max(rel) <= (1e-12 if precision == "fp64" else 1e-7)


@pytest.mark.timeout(300)
@pytest.mark.parametrize("precision", ["fp64", "fp32_mixed"])
def test_uniform_state_residual(precision):
    """FP32-mixed: exactly zero (traces and derivatives act on differences from a line's
    first node, so every difference of a constant line is zero).  FP64 plain: the reference's
    own free-stream roundoff, which grows as the inverse square of the element size because the
    metrics come from D applied to the node coordinates (9e-8 in E at this size, 4e-12 at
    8 x 8); equal to the oracle's to 1e-12 of the residual scale 2 N max|U| of an O(1) state on
    this mesh."""
    st, m, orc = make_case(P, N, N, bc="transmissive", capturing="persson_br0", ic="riemann", seed=SEED,
                           precision=precision)
    U = np.array([1.3, 0.4, -0.7, 2.9])
    st.u[...] = U
    du = np.full(st.u.shape, np.nan, order="F")
    residual(du, st, EQ, m)
    _free(st)
    print(f"{N}^2 p={P} uniform {precision}: max |du| {np.max(np.abs(du)):.3e}", flush=True)
    if precision == "fp32_mixed":
        assert not np.any(du)
    else:
        ref = orc.residual(st.u)
        d = float(np.max(np.abs(du - ref)))
        print(f"oracle max |du| {np.max(np.abs(ref)):.3e}, max |device - oracle| {d:.3e}", flush=True)
        assert np.all(np.isfinite(du)) and d <= 1e-12 * 2 * N * np.max(np.abs(U))
