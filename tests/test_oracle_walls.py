"""Reflecting and Dirichlet boundaries of the oracle (exterior_state, BoundaryConditions.jl:43-110;
boundary_fhat!, Residual2D.jl:63-94), pinned by what the mathematics fixes:

* a closed box (four reflecting walls) conserves mass and energy: sum w_i w_j |J| du = 0 for the
  density and the energy (the wall flux of both vanishes when the ghost mirrors the velocity);
* mirror symmetry: on a state symmetric about x = 1/2 in a periodic domain, the interface at
  x = 1/2 sees the mirrored trace, so the half domain [1/2, 1] with reflecting walls at both x
  ends has the full domain's residual restricted to that half;
* a Dirichlet side changes the residual only in the elements on that side, by the correction
  -(F^_D - F^_T) sJ g'_L / J with F^_D the flux against the given state and F^_T the
  transmissive one.
"""
import numpy as np
import pytest

from frforge_amd import (Mesh2D, PeriodicBC, TransmissiveBC, allocate_state, build_operators,
                         set_initial_condition)
from frforge_amd.equations import primitives_to_conserved
from oracle.oracle import Oracle, flux_n

from helpers import EQ, oracle_ops

WALLS = {"left": "reflecting", "right": "reflecting", "bottom": "reflecting", "top": "reflecting"}


def _smooth(x, y):
    rho = 1.0 + 0.2 * np.sin(2 * np.pi * x) * np.cos(np.pi * y) + 0.1 * x
    u = 0.3 * np.cos(np.pi * y) + 0.2 * x
    v = -0.2 * np.sin(np.pi * x) + 0.1
    p = 1.0 + 0.1 * np.cos(2 * np.pi * (x + y))
    return primitives_to_conserved(EQ, rho, u, v, p)


def _case(p, nx, ny, x0=0.0, x1=1.0, periodic_x=False, periodic_y=False, f=_smooth):
    ops = build_operators(p)
    Bx = PeriodicBC if periodic_x else TransmissiveBC
    By = PeriodicBC if periodic_y else TransmissiveBC
    mesh = Mesh2D(x0, x1, 0.0, 1.0, nx, ny, left_bc=Bx(), right_bc=Bx(), bottom_bc=By(), top_bc=By())
    st = allocate_state(mesh, ops)
    set_initial_condition(st, f)
    return st


def _oracle(st, capturing="null", flux="rusanov", sides=None, solid=None):
    m = st.mesh
    bx = "periodic" if isinstance(m.left_bc, PeriodicBC) else "transmissive"
    by = "periodic" if isinstance(m.bottom_bc, PeriodicBC) else "transmissive"
    return Oracle(st.p, m.nx, m.ny, m.x_vertices, m.y_vertices, oracle_ops(st.ops), bc_x=bx, bc_y=by,
                  flux=flux, capturing=capturing, sides=sides, solid=solid)


@pytest.mark.parametrize("flux", ["rusanov", "hllc"])
@pytest.mark.parametrize("p", [1, 3])
def test_closed_box_conserves_mass_and_energy(p, flux):
    """Inviscid: BR0 keeps the reference's boundary viscous flux g^ = g_interior on every
    non-periodic side (PerssonAV2D.jl:310-348), which is not zero at a wall."""
    st = _case(p, 9, 7)
    orc = _oracle(st, "null", flux, WALLS)
    du = orc.residual(st.u)
    w = st.ops.w
    W = w[:, None, None] * w[None, :, None] * np.abs(st.metrics.J)
    for c in (0, 3):
        tot = np.sum(W * du[:, :, :, c])
        scale = np.sum(W * np.abs(du[:, :, :, c]))
        assert abs(tot) <= 1e-13 * scale, (c, tot, scale)
    # the x momentum is not conserved: the walls push
    assert abs(np.sum(W * du[:, :, :, 1])) > 1e-6 * np.sum(W * np.abs(du[:, :, :, 1]))


@pytest.mark.parametrize("flux", ["rusanov", "hllc"])
@pytest.mark.parametrize("p", [2, 3])
def test_mirror_symmetry_matches_reflecting_walls(p, flux):
    m = 6

    def sym(x, y):   # symmetric about x = 1/2 (and so about x = 0 = 1): rho, v, p even, u odd
        s = np.abs(x - 0.5)
        rho = 1.0 + 0.2 * np.cos(2 * np.pi * s) + 0.1 * np.sin(2 * np.pi * y)
        u = 0.3 * np.sin(2 * np.pi * (x - 0.5))
        v = 0.2 * np.cos(2 * np.pi * s) + 0.05 * np.cos(2 * np.pi * y)
        pr = 1.0 + 0.1 * np.cos(4 * np.pi * s)
        return primitives_to_conserved(EQ, rho, u, v, pr)

    full = _case(p, 2 * m, 5, periodic_x=True, periodic_y=True, f=sym)
    half = _case(p, m, 5, x0=0.5, x1=1.0, periodic_y=True, f=sym)
    du_full = _oracle(full, flux=flux).residual(full.u)
    du_half = _oracle(half, flux=flux, sides={"left": "reflecting", "right": "reflecting"}).residual(half.u)
    nx = 2 * m
    right = [jy * nx + jx for jy in range(5) for jx in range(m, nx)]
    ref = du_full[:, :, right, :]
    err = np.max(np.abs(du_half - ref)) / np.max(np.abs(ref))
    assert err <= 1e-11, err


@pytest.mark.parametrize("side", ["left", "top"])
def test_dirichlet_side_changes_only_its_elements(side):
    p, nx, ny = 3, 5, 4
    st = _case(p, nx, ny)
    UD = primitives_to_conserved(EQ, 1.3, 0.4, -0.1, 1.2)
    du_t = _oracle(st).residual(st.u)
    du_d = _oracle(st, sides={side: ("dirichlet", UD)}).residual(st.u)
    orc = _oracle(st)
    Np, ops = p + 1, st.ops
    # metric arrays: 7..9 W (nx, ny, sJ), 16..18 N; J = 6
    J = orc.metric(6).reshape(Np, Np, nx * ny, order="F")
    pred = np.zeros_like(du_t)
    u = st.u
    if side == "left":
        nxw, nyw, sJ = (orc.metric(k).reshape(Np, nx * ny, order="F") for k in (7, 8, 9))
        for jy in range(ny):
            e = jy * nx
            for j in range(Np):
                tr = np.array([np.dot(ops.ell_L, u[:, j, e, c]) for c in range(4)])
                fd = flux_n("rusanov", 1.4, UD, tr, -nxw[j, e], -nyw[j, e]) * sJ[j, e]
                ft = flux_n("rusanov", 1.4, tr, tr, -nxw[j, e], -nyw[j, e]) * sJ[j, e]
                for i in range(Np):
                    pred[i, j, e, :] = -((fd - ft) * ops.gL_xi[i]) / J[i, j, e]
    else:
        nxn, nyn, sJ = (orc.metric(k).reshape(Np, nx * ny, order="F") for k in (16, 17, 18))
        for jx in range(nx):
            e = (ny - 1) * nx + jx
            for i in range(Np):
                tr = np.array([np.dot(ops.ell_R, u[i, :, e, c]) for c in range(4)])
                fd = flux_n("rusanov", 1.4, tr, UD, nxn[i, e], nyn[i, e]) * sJ[i, e]
                ft = flux_n("rusanov", 1.4, tr, tr, nxn[i, e], nyn[i, e]) * sJ[i, e]
                for j in range(Np):
                    pred[i, j, e, :] = -((fd - ft) * ops.gR_xi[j]) / J[i, j, e]
    diff = du_d - du_t
    scale = np.max(np.abs(pred))
    assert scale > 1e-3
    assert np.max(np.abs(diff - pred)) <= 1e-11 * max(scale, np.max(np.abs(du_t)))


def test_reflecting_rest_state_is_steady():
    """A fluid at rest in a closed box: zero residual (the wall flux is the pressure alone)."""
    st = _case(3, 6, 5, f=lambda x, y: primitives_to_conserved(EQ, 1.0 + 0 * x, 0 * x, 0 * x, 1.0 + 0 * x))
    du = _oracle(st, sides=WALLS).residual(st.u)
    assert np.max(np.abs(du)) < 1e-12


def _mask(nx, ny, cells):
    m = np.zeros(nx * ny, dtype=bool)
    for jx, jy in cells:
        m[jy * nx + jx] = True
    return m


@pytest.mark.parametrize("flux", ["rusanov", "hllc"])
def test_solid_column_is_a_reflecting_wall(flux):
    """A column of solid elements at jx = 0 of [0,1] (8 columns) leaves the fluid of [1/8, 1] with
    a wall at x = 1/8: the residual equals the 7-column mesh's with a reflecting left side (the
    vertices are the same binary fractions), and solid elements get du = 0."""
    p, ny = 3, 5
    full = _case(p, 8, ny)
    cut = _case(p, 7, ny, x0=0.125, x1=1.0)
    mask = _mask(8, ny, [(0, jy) for jy in range(ny)])
    du_full = _oracle(full, flux=flux, sides={"left": "reflecting"}, solid=mask).residual(full.u)
    du_cut = _oracle(cut, flux=flux, sides={"left": "reflecting"}).residual(cut.u)
    fluid = [jy * 8 + jx for jy in range(ny) for jx in range(1, 8)]
    assert np.all(du_full[:, :, mask, :] == 0.0)
    err = np.max(np.abs(du_full[:, :, fluid, :] - du_cut)) / np.max(np.abs(du_cut))
    assert err <= 1e-13, err


def test_solid_block_in_a_closed_box_conserves_mass():
    p, nx, ny = 2, 10, 8
    st = _case(p, nx, ny)
    mask = _mask(nx, ny, [(jx, jy) for jx in (4, 5) for jy in (0, 1, 2)])   # a step on the floor
    du = _oracle(st, sides=WALLS, solid=mask).residual(st.u)
    assert np.all(du[:, :, mask, :] == 0.0)
    w = st.ops.w
    W = w[:, None, None] * w[None, :, None] * np.abs(st.metrics.J)
    for c in (0, 3):
        tot = np.sum(W * du[:, :, :, c])
        assert abs(tot) <= 1e-13 * np.sum(W * np.abs(du[:, :, :, c]))


def test_solid_mask_rejects_periodic_axes():
    st = _case(2, 4, 4, periodic_x=True)
    with pytest.raises(ValueError):
        _oracle(st, solid=_mask(4, 4, [(1, 1)]))


# --- split ghost sides (GhostStateBC with a moving switch point, Cases2D_Optional.jl:72-84) ---

U_LO = primitives_to_conserved(EQ, 1.7, 0.4, -0.2, 1.9)
U_HI = primitives_to_conserved(EQ, 0.8, -0.1, 0.05, 0.7)


def _split(coef, hi="reflecting"):
    return ("split", coef, U_LO, hi, U_HI if hi == "dirichlet" else None)


@pytest.mark.parametrize("side", ["left", "right", "bottom", "top"])
@pytest.mark.parametrize("hi", ["reflecting", "dirichlet"])
def test_split_side_far_switch_points_reduce_to_constant_sides(side, hi):
    """A switch point below every face point gives the hi kind everywhere, one above every face
    point the low Dirichlet state: the residuals equal those of the constant sides bit for bit."""
    st = _case(3, 8, 6)
    u = st.u
    below = _oracle(st, "persson_br0", "hllc", {side: _split((-1e9, 0.0, 0.0, 1.0), hi)}).residual(u)
    above = _oracle(st, "persson_br0", "hllc", {side: _split((1e9, 0.0, 0.0, 1.0), hi)}).residual(u)
    hi_spec = "reflecting" if hi == "reflecting" else ("dirichlet", U_HI)
    want_hi = _oracle(st, "persson_br0", "hllc", {side: hi_spec}).residual(u)
    want_lo = _oracle(st, "persson_br0", "hllc", {side: ("dirichlet", U_LO)}).residual(u)
    assert np.array_equal(below, want_hi)
    assert np.array_equal(above, want_lo)


def _split_at_edge(st, side, hi, coef, t, k):
    """Residual with the switch point at vertex k of the side's axis (time t), and the element
    by element composition of the two constant sides that it must equal: the Gauss-Legendre face
    points lie strictly inside their elements, so each element sees one of the two states."""
    m = st.mesh
    got = _oracle(st, "null", "rusanov", {side: _split(coef, hi)}).residual(st.u, t)
    hi_spec = "reflecting" if hi == "reflecting" else ("dirichlet", U_HI)
    r_hi = _oracle(st, "null", "rusanov", {side: hi_spec}).residual(st.u, t)
    r_lo = _oracle(st, "null", "rusanov", {side: ("dirichlet", U_LO)}).residual(st.u, t)
    e = np.arange(m.n_elements)
    along = (e % m.nx) if side in ("bottom", "top") else (e // m.nx)
    want = np.where((along < k)[None, None, :, None], r_lo, r_hi)
    return got, want, r_lo, r_hi


@pytest.mark.parametrize("side", ["bottom", "top", "left", "right"])
def test_split_side_switch_at_an_element_edge(side):
    """8 x 8 elements on the unit square (binary vertices): with the switch point at x (or y) =
    3/8 the elements left of it take the low state and the rest the hi kind."""
    st = _case(2, 8, 8)
    got, want, r_lo, r_hi = _split_at_edge(st, side, "reflecting", (0.375, 0.0, 0.0, 1.0), 0.0, 3)
    assert np.array_equal(got, want)
    assert not np.array_equal(got, r_lo) and not np.array_equal(got, r_hi)


def test_split_side_moves_with_time():
    """Switch point (1/4 + t) / 2 on the top: at t = 1/4 it sits on x = 1/4 (vertex 2), at
    t = 3/4 on x = 1/2 (vertex 4)."""
    st = _case(2, 8, 8)
    coef = (0.0, 0.25, 1.0, 2.0)
    for t, k in ((0.25, 2), (0.75, 4)):
        got, want, _, _ = _split_at_edge(st, "top", "dirichlet", coef, t, k)
        assert np.array_equal(got, want)


def test_split_side_stage_times_of_a_step():
    """One SSP-RK3 step from t0 = 0 with dt = 1/1024 and the switch point 1/4 + 256 t on the top
    equals the step composed from residuals at t0, t0 + dt and t0 + dt / 2 (SSP_RK3.jl:40-60):
    the three stage times put the switch point on three different element edges (x = 1/4,
    1/2 and 3/8), so a wrong stage time changes the result."""
    st = _case(1, 8, 8)
    sides = {"top": _split((0.25, 0.0, 256.0, 1.0), "dirichlet"), "bottom": "reflecting"}
    orc = _oracle(st, "null", "rusanov", sides)
    t0, dt = 0.0, 1.0 / 1024.0
    u0 = st.u.copy(order="F")
    status, n, t, u = orc.run(u0, t0, t0 + dt, dt=dt)
    assert (status, n, t) == ("ok", 1, t0 + dt)
    u1 = u0 + dt * orc.residual(u0, t0)
    u2 = 0.75 * u0 + 0.25 * u1 + 0.25 * dt * orc.residual(u1, t0 + dt)
    u3 = (1.0 / 3.0) * u0 + (2.0 / 3.0) * u2 + (2.0 / 3.0) * dt * orc.residual(u2, t0 + dt / 2.0)
    assert np.array_equal(u, u3)
    for t_bad in (t0, t0 + dt):
        wrong = (1.0 / 3.0) * u0 + (2.0 / 3.0) * u2 + (2.0 / 3.0) * dt * orc.residual(u2, t_bad)
        assert not np.array_equal(u, wrong)
