"""Reference-element operators: the host builder (the product's set-up) and the oracle's own
builder, against the reference's invariants (test/test_operators.jl:5-60) and textbook rules."""
import numpy as np
import pytest

from frforge_amd.operators import (build_operators, g_DG_endpoints, gauss_legendre_nodes_weights,
                                   gauss_lobatto_legendre_nodes_weights)
from oracle import oracle as O


@pytest.mark.parametrize("n", range(1, 7))
def test_gl_nodes_weights(n):
    xi, w = gauss_legendre_nodes_weights(n)
    assert len(xi) == n and np.all(np.diff(xi) > 0) and np.all(np.abs(xi) <= 1)
    assert abs(w.sum() - 2) < 1e-14
    assert abs(np.sum(w * xi)) < 1e-13
    xr, wr = np.polynomial.legendre.leggauss(n)   # independent rule (Newton on P_n)
    np.testing.assert_allclose(xi, xr, atol=2e-15)
    np.testing.assert_allclose(w, wr, atol=2e-15)
    # exact for degree 2n-1
    for k in range(2 * n):
        exact = 0.0 if k % 2 else 2.0 / (k + 1)
        assert abs(np.sum(w * xi**k) - exact) < 1e-13


@pytest.mark.parametrize("p", range(1, 6))
def test_differentiation_exact_on_polynomials(p):
    ops = build_operators(p)
    for k in range(p + 1):
        du = ops.D @ ops.xi**k
        ex = np.zeros_like(ops.xi) if k == 0 else k * ops.xi ** (k - 1)
        np.testing.assert_allclose(du, ex, atol=1e-12, rtol=1e-12)


@pytest.mark.parametrize("p", range(1, 6))
def test_correction_and_lagrange(p):
    ep = g_DG_endpoints(p)
    assert abs(ep["gL_m1"] - 1) < 1e-12 and abs(ep["gL_p1"]) < 1e-12
    assert abs(ep["gR_m1"]) < 1e-12 and abs(ep["gR_p1"] - 1) < 1e-12
    ops = build_operators(p)
    assert abs(ops.ell_L.sum() - 1) < 1e-13 and abs(ops.ell_R.sum() - 1) < 1e-13
    # l(+-1) interpolate polynomials of degree p exactly
    for k in range(p + 1):
        assert abs(ops.ell_L @ ops.xi**k - (-1.0) ** k) < 1e-12
        assert abs(ops.ell_R @ ops.xi**k - 1.0) < 1e-12
    # g_DG (right Radau based): g_L' + g_R' integrate to -1 and +1 over [-1, 1]
    assert abs(np.sum(ops.w * ops.gL_xi) + 1) < 1e-12
    assert abs(np.sum(ops.w * ops.gR_xi) - 1) < 1e-12
    # g_L(xi) = (-1)^p/2 (P_p - P_{p+1}) is the left Radau polynomial
    P = np.polynomial.legendre.Legendre
    gl = 0.5 * (-1) ** p * (P.basis(p) - P.basis(p + 1))
    np.testing.assert_allclose(ops.gL, gl(ops.xi), atol=1e-13)
    np.testing.assert_allclose(ops.gL_xi, gl.deriv()(ops.xi), atol=1e-12)
    V = np.polynomial.legendre.legvander(ops.xi, p)
    np.testing.assert_allclose(ops.V_legendre, V, atol=1e-14)


@pytest.mark.parametrize("p", range(1, 6))
def test_oracle_operators_match_host(p):
    ops = build_operators(p)
    oo = O.build_operators(p)
    for a, b in ((ops.xi, oo["xi"]), (ops.w, oo["w"]), (ops.D, oo["D"]), (ops.ell_L, oo["lL"]),
                 (ops.ell_R, oo["lR"]), (ops.gL_xi, oo["gLxi"]), (ops.gR_xi, oo["gRxi"]),
                 (ops.V_legendre, oo["V"])):
        np.testing.assert_allclose(a, b, atol=1e-13, rtol=1e-13)


@pytest.mark.parametrize("n", range(2, 7))
def test_gll(n):
    xi, w = gauss_lobatto_legendre_nodes_weights(n)
    assert xi[0] == -1 and xi[-1] == 1 and abs(w.sum() - 2) < 1e-14
    for k in range(2 * n - 2):
        exact = 0.0 if k % 2 else 2.0 / (k + 1)
        assert abs(np.sum(w * xi**k) - exact) < 1e-12
