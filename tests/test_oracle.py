"""Pins of the CPU oracle (the checker the GPU path is held to).

Each test is a known answer from the reference's own test suite, or a closed form / invariant
chosen so that a plausible slip in the restatement (a swapped xi/eta index, a wrong Jacobian
factor, a dropped correction term, a sign error on a boundary normal) fails it.
"""
import math

import numpy as np
import pytest

from frforge_amd import (Euler2D, Mesh2D, PeriodicBC, TransmissiveBC, allocate_state,
                         build_operators, l2_error, primitives_to_conserved, set_initial_condition)
from frforge_amd.cases import vortex_ic
from oracle import oracle as O

from helpers import EQ, make_case, oracle_ops

G = 1.4


def _flux_phys(U, nx, ny, g=G):
    rho, mu, mv, E = U
    u, v = mu / rho, mv / rho
    p = (g - 1) * (E - 0.5 * rho * (u * u + v * v))
    F = np.array([mu, mu * u + p, mv * u, (E + p) * u])
    Gy = np.array([mv, mu * v, mv * v + p, (E + p) * v])
    return F * nx + Gy * ny


# ------------------------------------------------------------------ known answers

@pytest.mark.parametrize("ops_from", ["host", "oracle"])
def test_vortex_order_known_answer(ops_from):
    """test_2d.jl:316-321 (run_isentropic_vortex_order p=2, n=[8,16], t=0.5, cfl=0.08)."""
    p, L = 2, 10.0
    ops = build_operators(p)
    d = oracle_ops(ops)
    if ops_from == "oracle":
        oo = O.build_operators(p)
        d = {k: oo[k] for k in ("xi", "w", "D", "lL", "lR", "gLxi", "gRxi", "V")}
    lam = 1.0 + 1.0 + 5.0 / (2 * math.pi) + math.sqrt(1.4) + 1.0
    dt = (0.08 / (p + 1)) * (L / 16) / ((2 * p + 1) * lam)
    errs = []
    for n in (8, 16):
        mesh = Mesh2D(0.0, L, 0.0, L, n, n)
        st = allocate_state(mesh, ops)
        set_initial_condition(st, vortex_ic(EQ))
        orc = O.Oracle(p, n, n, mesh.x_vertices, mesh.y_vertices, d)
        status, nsteps, t, u = orc.run(st.u, 0.0, 0.5, dt=dt)
        assert status == "ok" and nsteps == 747 and t == 0.5
        st.u[...] = u
        errs.append(l2_error(st, vortex_ic(EQ, t=0.5)))
    order = math.log(errs[0] / errs[1]) / math.log(2.0)
    assert math.isclose(errs[0], 0.02069688015968434, rel_tol=1e-12, abs_tol=1e-14)
    assert math.isclose(errs[1], 0.0031146731348864697, rel_tol=1e-12, abs_tol=1e-14)
    assert math.isclose(order, 2.7322606381318906, rel_tol=1e-10, abs_tol=1e-12)


@pytest.mark.parametrize("flux", ["rusanov", "hllc"])
@pytest.mark.parametrize("bc", ["periodic", "transmissive"])
@pytest.mark.parametrize("capturing", ["null", "persson_br0", "persson_dd"])
def test_freestream(flux, bc, capturing):
    """test_2d.jl:304-315: uniform (1, 0.5, 0.25, 1) on 4x4, p=2 gives max|du| < 1e-12."""
    st, m, orc = make_case(2, 4, 4, bc=bc, flux=flux, capturing=capturing, ic="vortex",
                           domain=(0.0, 1.0, 0.0, 1.0))
    st.u[...] = primitives_to_conserved(EQ, 1.0, 0.5, 0.25, 1.0)
    du = orc.residual(st.u)
    assert np.max(np.abs(du)) < 1e-12


def test_viscous_mass_on_constants():
    """test_2d.jl:47 (viscous_mass_residual_scale_2d < 1e-10): BR0 with c_av = 1, sigma = 1,
    lambda = 1 on a constant field (PerssonAV2D.jl:401-423)."""
    st, m, orc = make_case(2, 4, 4, bc="periodic", domain=(0.0, 1.0, 0.0, 1.0))
    st.u[...] = 1.0
    h = np.minimum(st.mesh.dx[0], st.mesh.dy[0])
    eps = np.full(st.mesh.n_elements, 1.0 * 1.0 * (h / 2) * 1.0)
    du = orc.apply_br0(st.u, eps)
    w = st.ops.w
    mass = np.sum(w[:, None, None] * w[None, :, None] * st.mesh.Jx[0] * st.mesh.Jy[0] * du[:, :, :, 0])
    assert abs(mass) < 1e-10
    assert np.max(np.abs(du)) < 1e-12


def test_sensor_fires_on_jump():
    """test_2d.jl:51-61: diagonal jump on 4x4, p=2: sigma in [0, 1], max sigma > 0.01."""
    st, m, orc = make_case(2, 4, 4, bc="periodic", domain=(0.0, 1.0, 0.0, 1.0))
    UL = primitives_to_conserved(EQ, 1.0, 0.0, 0.0, 1.0)
    UR = primitives_to_conserved(EQ, 0.125, 0.0, 0.0, 0.1)
    X, Y = st.metrics.x, st.metrics.y
    st.u[...] = np.where((X + Y < 0.7)[..., None], UL, UR)
    s = orc.sense(st.u)
    assert np.all((s >= 0) & (s <= 1)) and s.max() > 0.01


def test_hllc_stationary_contact():
    """test_schemes.jl:97-102, in 2D along both axes: momentum flux = p, mass/energy flux 0."""
    for nx, ny in ((1.0, 0.0), (0.0, 1.0), (0.6, 0.8)):
        uL = primitives_to_conserved(EQ, 1.0, 0.0, 0.0, 1.0)
        uR = primitives_to_conserved(EQ, 0.1, 0.0, 0.0, 1.0)
        f = O.flux_n("hllc", G, uL, uR, nx, ny)
        np.testing.assert_allclose(f, [0.0, nx, ny, 0.0], atol=1e-10)


# ------------------------------------------------------------------ closed forms and invariants

@pytest.mark.parametrize("kind", ["rusanov", "hllc"])
def test_flux_consistency_and_symmetry(kind):
    rng = np.random.default_rng(0)
    for _ in range(50):
        prim = [rng.uniform(0.2, 2), rng.uniform(-1, 1), rng.uniform(-1, 1), rng.uniform(0.1, 2)]
        U = primitives_to_conserved(EQ, *prim)
        th = rng.uniform(0, 2 * np.pi)
        n = (np.cos(th), np.sin(th))
        np.testing.assert_allclose(O.flux_n(kind, G, U, U, *n), _flux_phys(U, *n), rtol=1e-13, atol=1e-13)
        V = primitives_to_conserved(EQ, rng.uniform(0.2, 2), rng.uniform(-1, 1), rng.uniform(-1, 1),
                                    rng.uniform(0.1, 2))
        a = O.flux_n(kind, G, U, V, *n)
        b = O.flux_n(kind, G, V, U, -n[0], -n[1])
        np.testing.assert_allclose(a, -b, rtol=1e-12, atol=1e-12)


def test_rusanov_closed_form():
    uL = primitives_to_conserved(EQ, 1.0, 0.3, -0.2, 1.0)
    uR = primitives_to_conserved(EQ, 0.5, -0.1, 0.4, 0.4)
    n = (0.0, 1.0)
    cl, cr = math.sqrt(1.4 * 1.0 / 1.0), math.sqrt(1.4 * 0.4 / 0.5)
    lam = max(abs(-0.2) + cl, abs(0.4) + cr)
    ref = 0.5 * (_flux_phys(uL, *n) + _flux_phys(uR, *n)) - 0.5 * lam * (uR - uL)
    np.testing.assert_allclose(O.flux_n("rusanov", G, uL, uR, *n), ref, rtol=1e-14)


def test_hllc_supersonic_upwinds():
    """All waves to the right: HLLC returns F(uL).n exactly (HLLC.jl:272-273)."""
    uL = primitives_to_conserved(EQ, 1.0, 3.0, 0.5, 1.0)
    uR = primitives_to_conserved(EQ, 0.7, 2.5, 0.1, 0.6)
    np.testing.assert_allclose(O.flux_n("hllc", G, uL, uR, 1.0, 0.0), _flux_phys(uL, 1.0, 0.0), rtol=1e-15)
    np.testing.assert_allclose(O.flux_n("hllc", G, uR, uL, -1.0, 0.0), _flux_phys(uL, -1.0, 0.0), rtol=1e-15)


@pytest.mark.parametrize("p", [1, 2, 3, 4, 5])
def test_sensor_closed_forms(p):
    """Modal energies of prescribed Legendre content: a constant has no top-mode energy;
    2 + a P_p(xi) + b P_p(eta) + c P_1(xi)P_1(eta) has e_high = a^2 + b^2 exactly."""
    st, m, orc = make_case(p, 3, 2, bc="periodic", capturing="persson_br0",
                           domain=(0.0, 3.0, 0.0, 2.0))
    xi = st.ops.xi
    P = np.polynomial.legendre.Legendre
    Pp, P1 = P.basis(p)(xi), P.basis(1)(xi)
    s0 = -4.0 * math.log10(max(p, 1))
    sig = lambda eh, et: 1.0 / (1.0 + math.exp(-4.0 * (math.log10(eh / et + 1e-16) - s0)))
    for a, b, c in ((0.0, 0.0, 0.0), (0.5, 0.0, 0.0), (0.0, 0.3, 0.1), (0.2, 0.4, 0.3)):
        rho = 2.0 + a * Pp[:, None] + b * Pp[None, :] + c * P1[:, None] * P1[None, :]
        if p == 1:   # P_1 x P_1 is then itself a top mode in both directions
            continue
        st.u[...] = 0.0
        st.u[:, :, :, 0] = rho[:, :, None]
        st.u[:, :, :, 3] = 2.5
        s = orc.sense(st.u)
        eh = a * a + b * b
        et = 4.0 + a * a + b * b + c * c + 1e-16
        np.testing.assert_allclose(s, sig(eh, et), rtol=1e-9, atol=1e-30)


@pytest.mark.parametrize("form", ["br0", "dd"])
@pytest.mark.parametrize("bc", ["periodic", "transmissive"])
@pytest.mark.parametrize("axis", ["x", "y"])
def test_av_quadratic_exactness(form, bc, axis):
    """Constant eps on u = s^2 (s = x or y): both AV forms give du = 2 eps exactly for p >= 2
    (BR0: g = 2 eps s is continuous, the penalty vanishes; DD: eps/J^2 D(D u) = 2 eps)."""
    st, m, orc = make_case(3, 5, 4, bc=bc, domain=(0.0, 1.0, 0.0, 0.8))
    X, Y = st.metrics.x, st.metrics.y
    s = X if axis == "x" else Y
    for c in range(4):
        st.u[:, :, :, c] = s * s
    eps = np.full(st.mesh.n_elements, 0.037)
    du = orc.apply_br0(st.u, eps) if form == "br0" else orc.apply_dd(st.u, eps)
    if bc == "periodic" and form == "br0":
        # the periodic wrap joins s^2 at 0 with s^2 at L: only elements away from the wrap
        jx = np.arange(st.mesh.n_elements) % st.mesh.nx
        jy = np.arange(st.mesh.n_elements) // st.mesh.nx
        j = jx if axis == "x" else jy
        n = st.mesh.nx if axis == "x" else st.mesh.ny
        keep = (j > 0) & (j < n - 1)
        du = du[:, :, keep]
    np.testing.assert_allclose(du, 2 * 0.037, rtol=1e-10)


@pytest.mark.parametrize("capturing", ["null", "persson_br0"])
@pytest.mark.parametrize("flux", ["rusanov", "hllc"])
def test_periodic_conservation(capturing, flux):
    """Periodic telescoping: sum_e sum_ij w_i w_j J du = 0 for every conserved variable."""
    st, m, orc = make_case(3, 9, 7, bc="periodic", flux=flux, capturing=capturing, ic="riemann",
                           seed=4, noise=0.05)
    du = orc.residual(st.u)
    w = st.ops.w
    W = w[:, None, None] * w[None, :, None] * st.metrics.J
    for c in range(4):
        tot = np.sum(W * du[:, :, :, c])
        scale = np.sum(np.abs(W * du[:, :, :, c]))
        assert abs(tot) < 1e-13 * scale


def _mirror_x(st, u):
    """Reflect a state through x -> (x_l + x_r) - x on the element grid."""
    nx, ny, Np = st.mesh.nx, st.mesh.ny, st.ops.n_points
    v = u.reshape(Np, Np, nx, ny, 4, order="F")[::-1, :, ::-1, :, :].copy()   # (i, j, jx, jy, c)
    v[..., 1] *= -1
    return np.asfortranarray(v.reshape(Np, Np, nx * ny, 4, order="F"))


def _rotate(st, u):
    """Swap x and y (square mesh): u[i,j,(jy,jx)] -> u[j,i,(jx,jy)], rho u <-> rho v."""
    n, Np = st.mesh.nx, st.ops.n_points
    v = u.reshape(Np, Np, n, n, 4, order="F").transpose(1, 0, 3, 2, 4)[..., [0, 2, 1, 3]]
    return np.asfortranarray(v.reshape(Np, Np, n * n, 4, order="F"))


@pytest.mark.parametrize("capturing", ["null", "persson_br0", "persson_dd"])
@pytest.mark.parametrize("flux", ["rusanov", "hllc"])
def test_residual_symmetries(capturing, flux):
    """Mirror and 90-degree rotation equivariance of the residual on a transmissive square:
    catches any xi/eta or west/east/south/north bookkeeping slip."""
    st, m, orc = make_case(3, 6, 6, bc="transmissive", flux=flux, capturing=capturing,
                           ic="riemann", seed=9, noise=0.03)
    du = orc.residual(st.u)
    scale = np.max(np.abs(du))
    dm = orc.residual(_mirror_x(st, st.u))
    assert np.max(np.abs(dm - _mirror_x(st, du))) < 1e-11 * scale
    dr = orc.residual(_rotate(st, st.u))
    assert np.max(np.abs(dr - _rotate(st, du))) < 1e-11 * scale


def test_compute_dt_closed_form():
    st, m, orc = make_case(3, 6, 4, bc="periodic", domain=(0.0, 1.2, 0.0, 0.6))
    st.u[...] = primitives_to_conserved(EQ, 1.3, 0.4, -0.9, 0.8)
    c = math.sqrt(1.4 * 0.8 / 1.3)
    lam = max(abs(0.4) + c, abs(-0.9) + c)
    h = min(1.2 / 6, 0.6 / 4)
    assert math.isclose(orc.compute_dt(st.u, 0.25), 0.25 * h / (7 * lam), rel_tol=1e-14)


def test_divergence_and_step_bookkeeping():
    st, m, orc = make_case(2, 6, 6, bc="periodic", ic="vortex")
    u = st.u.copy(order="F")
    u[0, 0, 3, 2] = np.inf
    status, u1, t1 = orc.rk3_step(u, 0.25, 1e-3)
    assert status == "diverged" and t1 == 0.25
    np.testing.assert_array_equal(u1, u)
    status, n, t, _ = orc.run(st.u, 0.0, 1 / 64, dt=1.5 / 256)
    assert status == "ok" and n == 3 and t == 1 / 64     # two full steps, then the truncated one
    status, n, t, _ = orc.run(st.u, 0.0, 1.0, cfl=0.1, max_steps=2)
    assert status == "max_steps" and n == 2
