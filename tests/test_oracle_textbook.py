"""Pins of the textbook FR oracle (oracle/zj_textbook.c), CPU only.

Each check is chosen so that a plausible mistake fails it: a wrong sign or
metric on one edge breaks freestream or conservation, a mirrored flux-point
index breaks the convergence order, a wrong correction function fails its
closed form or characterisation, and the Riemann solvers are compared with
the compiled reference's own outputs.
"""
import math

import numpy as np
import pytest

from oracle import pyoracle as O
from tests import fr_case
from zhijian_amd import cases


# ----------------------------------------------------------- Riemann solvers
@pytest.mark.parametrize("kind", ["rusanov", "roe", "hllc"])
def test_riemann_matches_reference_golden(golden_physics, kind):
    """RiemannSolvers::{rusanov,roe,hllc} of the compiled reference (euler.hpp)."""
    g = float(golden_physics["gamma"])
    for k in range(golden_physics["UL"].shape[0]):
        UL, UR = golden_physics["UL"][k], golden_physics["UR"][k]
        nx, ny = golden_physics["normals"][k]
        ref = golden_physics[kind][k]
        got = O.riemann(kind, UL, UR, nx, ny, g)
        assert np.allclose(got, ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max()), (kind, k)


@pytest.mark.parametrize("kind", ["rusanov", "roe", "hllc"])
def test_riemann_consistency_and_symmetry(kind):
    rng = np.random.default_rng(3)
    for _ in range(50):
        rho, u, v, p = rng.uniform(0.5, 2), rng.uniform(-1, 1), rng.uniform(-1, 1), rng.uniform(0.5, 3)
        U = cases.prim_to_cons(np.array(rho), np.array(u), np.array(v), np.array(p), 1.4)
        a = rng.uniform(0, 2 * math.pi)
        nx, ny = math.cos(a), math.sin(a)
        # F(U, U, n) = F(U) . n
        assert np.allclose(O.riemann(kind, U, U, nx, ny), O.flux_normal(U, nx, ny), rtol=1e-13, atol=1e-13)
        # conservation: F(UL, UR, n) = -F(UR, UL, -n)
        U2 = cases.prim_to_cons(np.array(rho * 1.2), np.array(u - 0.1), np.array(v + 0.2), np.array(p * 0.9), 1.4)
        assert np.allclose(O.riemann(kind, U, U2, nx, ny), -O.riemann(kind, U2, U, -nx, -ny),
                           rtol=1e-12, atol=1e-12)


# ------------------------------------------------------ correction functions
def _poly_from_deriv(p, x, d, at_minus1):
    """The degree-(p+1) polynomial with derivative sampled as d at x and value
    at_minus1 at -1 (numpy polynomial fit, exact for these degrees)."""
    c = np.polynomial.polynomial.polyfit(x, d, p)
    g = np.polynomial.polynomial.polyint(c)
    g[0] += at_minus1 - np.polynomial.polynomial.polyval(-1.0, g)
    return g


@pytest.mark.parametrize("p", [1, 2, 3, 4, 5])
def test_vcjh_dg_is_the_reference_radau(p):
    xs, _ = O.gauss(p + 1)
    gl, gr = O.vcjh_deriv(p, "dg", xs)
    for i, x in enumerate(xs):
        assert abs(gl[i] - O.corr_deriv(p, x, False)) <= 1e-12 * max(1, abs(gl[i]))
        assert abs(gr[i] - O.corr_deriv(p, x, True)) <= 1e-12 * max(1, abs(gr[i]))


@pytest.mark.parametrize("p", [1, 2, 3, 4, 5])
def test_vcjh_hu_is_huynh_g2(p):
    """Huynh (2007): g2_L = (-1)^p/2 [P_p - ((p+1) P_{p-1} + p P_{p+1})/(2p+1)]."""
    L = np.polynomial.legendre
    x = np.linspace(-0.97, 0.97, 9)
    gl, gr = O.vcjh_deriv(p, "hu", x)
    e = lambda n: np.eye(p + 2)[n]
    g2 = (-1) ** p / 2 * (e(p) - ((p + 1) * e(p - 1) + p * e(p + 1)) / (2 * p + 1))
    assert np.allclose(gl, L.legval(x, L.legder(g2)), rtol=1e-12, atol=1e-12)
    assert np.allclose(gr, -L.legval(-x, L.legder(g2)), rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("kind", ["sd", "ga"])
@pytest.mark.parametrize("p", [1, 2, 3, 4, 5])
def test_vcjh_sd_ga_characterisation(kind, p):
    """The SD / Gauss correction: degree p+1, g_L(-1) = 1, zero at +1 and at
    the p Gauss points of P_p (Huynh's g_Ga)."""
    x = np.linspace(-0.9, 0.9, p + 1)
    gl, gr = O.vcjh_deriv(p, kind, x)
    g = _poly_from_deriv(p, x, gl, 1.0)
    zeros = np.polynomial.legendre.leggauss(p)[0]
    assert abs(np.polynomial.polynomial.polyval(1.0, g)) < 1e-12
    assert np.allclose(np.polynomial.polynomial.polyval(zeros, g), 0.0, atol=1e-12)
    assert np.allclose(gr, -O.vcjh_deriv(p, kind, -x)[0], rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("kind", ["dg", "sd", "hu"])
def test_vcjh_endpoint_values(kind):
    for p in range(1, 6):
        x = np.linspace(-0.9, 0.9, p + 1)
        gl, gr = O.vcjh_deriv(p, kind, x)
        GL = _poly_from_deriv(p, x, gl, 1.0)
        GR = _poly_from_deriv(p, x, gr, 0.0)
        assert abs(np.polynomial.polynomial.polyval(1.0, GL)) < 1e-12
        assert abs(np.polynomial.polynomial.polyval(1.0, GR) - 1.0) < 1e-12


# -------------------------------------------------------------- the scheme
def _state(case, kind, amp=0.0, seed=0, **kw):
    pos = O.sp_coords(case.elem_nodes, case.xy, case.ops)
    U = cases.vortex_ic(pos, **kw) if kind == "vortex" else cases.tgv_ic(pos, mach=0.3)
    return cases.perturb(U, amp, seed).reshape(-1) if amp else U.reshape(-1)


@pytest.mark.parametrize("riemann", ["rusanov", "roe", "hllc"])
@pytest.mark.parametrize("corr", ["dg", "sd", "hu", "ga"])
def test_textbook_freestream_on_jittered_mesh(riemann, corr):
    mesh = cases.uniform_quad_mesh(5, 4, 0.0, 1.0, 0.0, 1.0, jitter=0.15, seed=2)
    case = fr_case.from_quadmesh(mesh, 3)
    tb = case.textbook(riemann, corr)
    U = np.tile(cases.prim_to_cons(np.array(1.2), np.array(0.4), np.array(-0.3), np.array(1.1), 1.4),
                (case.E * case.S, 1)).reshape(-1)
    assert np.abs(tb.residual(U)).max() < 1e-11


@pytest.mark.parametrize("p", [1, 3, 5])
def test_textbook_conserves(p):
    """sum_e sum_sp w_sp J dU/dt = 0 on a periodic jittered mesh (DG with
    Gauss points integrates the correction exactly)."""
    mesh = cases.uniform_quad_mesh(5, 4, 0.0, 1.0, 0.0, 1.0, jitter=0.15, seed=4)
    case = fr_case.from_quadmesh(mesh, p)
    tb = case.textbook("roe", "dg")
    U = _state(case, "tgv", amp=0.05, seed=5)
    R = tb.residual(U).reshape(case.E, case.S, 4)
    w = case.ops["qw"][None, :, None] * case.J.reshape(case.E, case.S)[:, :, None]
    tot = (w * R).sum(axis=(0, 1))
    scale = (w * np.abs(R)).sum(axis=(0, 1))
    assert np.all(np.abs(tot) <= 1e-12 * scale)


def _jacobian(f, U0, eps=1e-7):
    n = U0.size
    Jm = np.zeros((n, n))
    for k in range(n):
        d = np.zeros(n)
        d[k] = eps
        Jm[:, k] = (f(U0 + d) - f(U0 - d)) / (2 * eps)
    return Jm


@pytest.mark.parametrize("corr", ["dg", "sd", "hu"])
def test_textbook_scheme_is_stable(corr):
    """Linearised about a uniform flow on a periodic mesh, the textbook scheme
    with an upwind flux has no growing modes (the reference's executed
    arithmetic has Re(lambda) of the order of |lambda|)."""
    mesh = cases.uniform_quad_mesh(3, 3, 0.0, 1.0, 0.0, 1.0)
    case = fr_case.from_quadmesh(mesh, 2)
    tb = case.textbook("roe", corr)
    U0 = np.tile(cases.prim_to_cons(np.array(1.0), np.array(0.5), np.array(0.2), np.array(1 / 1.4), 1.4),
                 (case.E * case.S, 1)).reshape(-1)
    lam = np.linalg.eigvals(_jacobian(tb.residual, U0))
    assert lam.real.max() < 1e-6 * np.abs(lam).max()


@pytest.mark.parametrize("p", [2, 3])
def test_textbook_order_of_accuracy_vortex(p):
    """Isentropic vortex advected for t = 0.5 on [-5,5]^2 periodic; the
    error must fall at close to the design order p+1 from 8x8 to 16x16."""
    errs = []
    for n in (8, 16):
        mesh = cases.vortex_mesh(n)
        case = fr_case.from_quadmesh(mesh, p)
        tb = case.textbook("roe", "dg")
        U = _state(case, "vortex")
        t_end, steps = 0.5, 25 * n // 8
        dt = t_end / steps
        for _ in range(steps):
            U = tb.advance(U, dt, False)  # the limiter's SI clamps (p >= 100) do not suit this state
        pos = O.sp_coords(case.elem_nodes, case.xy, case.ops)
        ex = cases.vortex_ic(pos - np.array([t_end, 0.0])).reshape(-1)
        w = (case.ops["qw"][None, :] * case.J.reshape(case.E, case.S)).reshape(-1)
        d = (U - ex).reshape(-1, 4)[:, 0]
        errs.append(math.sqrt(float(np.sum(w * d * d))))
    order = math.log2(errs[0] / errs[1])
    assert order > p + 0.5, (errs, order)
