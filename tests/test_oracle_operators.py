"""Pins the oracle's FR operators (reference basis.cpp:12-613).

Golden tables come from the reference's own FROperators compiled in place
(tests/golden/make_golden.py); the closed-form checks below catch a transposed
operand, a wrong edge orientation or a wrong correction polynomial
independently of those tables.
"""
import numpy as np
import pytest
from numpy.polynomial import legendre as npleg

from oracle import pyoracle as O

KEYS = ("sol_xi", "sol_eta", "qw", "diff_xi", "diff_eta", "interp", "corr")


@pytest.mark.parametrize("n", range(1, 9))
def test_gauss_matches_reference_and_integrates(golden_ops, n):
    p, w = O.gauss(n)
    assert np.array_equal(p, golden_ops[f"gauss{n}_pts"])
    assert np.array_equal(w, golden_ops[f"gauss{n}_wts"])
    assert abs(w.sum() - 2.0) < 1e-13
    # exact for polynomials of degree 2n-1 (the 6-point table has 15 digits)
    tol = 1e-13 if n != 6 else 1e-13 * 1e3
    for k in range(2 * n):
        exact = 0.0 if k % 2 else 2.0 / (k + 1)
        assert abs(np.dot(w, p ** k) - exact) < tol


@pytest.mark.parametrize("p", range(1, 6))
def test_operators_bit_exact_vs_reference(golden_ops, p):
    ops = O.quad_operators(p)
    for k in KEYS:
        assert np.array_equal(ops[k], golden_ops[f"P{p}_{k}"]), k
    assert int(golden_ops[f"P{p}_fluxtype_invariant"]) == 1


def _poly2(deg, rng):
    c = rng.standard_normal((deg + 1, deg + 1))
    f = lambda x, y: sum(c[a, b] * x ** a * y ** b for a in range(deg + 1) for b in range(deg + 1))
    fx = lambda x, y: sum(a * c[a, b] * x ** (a - 1) * y ** b
                          for a in range(1, deg + 1) for b in range(deg + 1))
    fy = lambda x, y: sum(b * c[a, b] * x ** a * y ** (b - 1)
                          for a in range(deg + 1) for b in range(1, deg + 1))
    return f, fx, fy


@pytest.mark.parametrize("p", range(1, 6))
def test_diff_and_interp_exact_on_polynomials(p):
    rng = np.random.default_rng(p)
    ops = O.quad_operators(p)
    n = p + 1; S = n * n
    xi, eta = ops["sol_xi"], ops["sol_eta"]
    f, fx, fy = _poly2(p, rng)
    u = f(xi, eta)
    Dx = ops["diff_xi"].reshape(S, S); De = ops["diff_eta"].reshape(S, S)
    assert np.allclose(Dx @ u, fx(xi, eta), atol=1e-11 * (1 + np.abs(u).max()))
    assert np.allclose(De @ u, fy(xi, eta), atol=1e-11 * (1 + np.abs(u).max()))
    # interpolation to the flux points: edge 0 (s,-1), 1 (1,s), 2 (-s,1), 3 (-1,-s)
    s, _ = O.gauss(n)
    I = ops["interp"].reshape(4, n, S)
    locs = [(s, -np.ones(n)), (np.ones(n), s), (-s, np.ones(n)), (-np.ones(n), -s)]
    for e, (xf, ef) in enumerate(locs):
        assert np.allclose(I[e] @ u, f(xf, ef), atol=1e-12 * (1 + np.abs(u).max())), e
    assert np.allclose(I.sum(axis=2), 1.0, atol=1e-14)


@pytest.mark.parametrize("p", range(1, 6))
def test_correction_is_radau_derivative(golden_ops, p):
    """g_R = (P_p + P_{p+1})/2, g_L = (-1)^p (P_p - P_{p+1})/2 (Huynh 2007, DG case);
    corr[e][sp] = g'(coordinate of sp) on every edge, no Lagrange factor != 1."""
    ops = O.quad_operators(p)
    n = p + 1; S = n * n
    xi, eta = ops["sol_xi"], ops["sol_eta"]
    cp = np.zeros(p + 2); cp[p] = 1
    cp1 = np.zeros(p + 2); cp1[p + 1] = 1
    dgR = lambda x: 0.5 * (npleg.legval(x, npleg.legder(cp)) + npleg.legval(x, npleg.legder(cp1)))
    dgL = lambda x: (-1) ** p * 0.5 * (npleg.legval(x, npleg.legder(cp)) - npleg.legval(x, npleg.legder(cp1)))
    C = ops["corr"].reshape(4, S)
    assert np.allclose(C[0], dgL(eta), atol=1e-12)
    assert np.allclose(C[2], dgR(eta), atol=1e-12)
    assert np.allclose(C[3], dgL(xi), atol=1e-12)
    assert np.allclose(C[1], dgR(xi), atol=1e-12)
    # the endpoint branch (basis.cpp:336-339) against the reference's values
    xs = golden_ops["corr_x"]
    gl = np.array([O.corr_deriv(p, x, False) for x in xs])
    gr = np.array([O.corr_deriv(p, x, True) for x in xs])
    assert np.array_equal(gl, golden_ops[f"corr_P{p}_left"])
    assert np.array_equal(gr, golden_ops[f"corr_P{p}_right"])
