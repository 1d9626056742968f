"""Pins of the viscous (BR2) terms of the textbook oracle, CPU only.

* the viscous flux against the reference's own ViscousFlux (compiled from
  include/physics/navierstokes.hpp; golden vectors in tests/golden/viscous.npz)
  given the same viscosity and conductivity, and against an independent
  derivation from analytic primitive fields;
* the corrected gradient: exact for a continuous polynomial field, and the
  discrete Gauss identity sum_e int Q = 0 on periodic meshes;
* the residual: freestream, conservation, linear stability, and the decay
  rate of the 2D Taylor-Green vortex, exp(-2 nu k^2 t) with k^2 = 2 for its
  kinetic energy, at low Mach number.
"""
import math
import os

import numpy as np
import pytest

from oracle import pyoracle as O
from tests import fr_case
from zhijian_amd import cases

GOLD = os.path.join(os.path.dirname(__file__), "golden", "viscous.npz")


def test_visc_flux_golden():
    g = np.load(GOLD)
    for i in range(g["U"].shape[0]):
        Fv, Gv = O.visc_flux(g["U"][i], g["dUdx"][i], g["dUdy"][i], g["mu_k"][i, 0], g["mu_k"][i, 1],
                             float(g["gamma"]))
        assert np.array_equal(Fv, g["fx"][i]) and np.array_equal(Gv, g["fy"][i])


def test_visc_flux_from_primitive_fields():
    """rho, u, v, p as analytic fields at a point; the conservative gradient
    by the product rule; the flux from the primitive derivatives directly."""
    rng = np.random.default_rng(3)
    gam, mu, kap = 1.4, 0.013, 0.021
    for _ in range(20):
        r, u, v, p = math.exp(rng.normal()), rng.normal(), rng.normal(), math.exp(rng.normal())
        dr, du, dv, dp = (rng.normal(size=2) for _ in range(4))   # (d/dx, d/dy) of each
        E = p / (gam - 1) + 0.5 * r * (u * u + v * v)
        dE = dp / (gam - 1) + 0.5 * dr * (u * u + v * v) + r * (u * du + v * dv)
        U = np.array([r, r * u, r * v, E])
        dU = np.array([dr, dr * u + r * du, dr * v + r * dv, dE])      # [4][2]
        Fv, Gv = O.visc_flux(U, dU[:, 0], dU[:, 1], mu, kap, gam)
        T, dT = p / r, (dp * r - p * dr) / r ** 2
        S = np.array([[du[0], du[1]], [dv[0], dv[1]]])                  # velocity gradient
        tau = mu * (S + S.T) - 2.0 / 3.0 * mu * np.trace(S) * np.eye(2)
        vel = np.array([u, v])
        ref_x = [0.0, tau[0, 0], tau[1, 0], vel @ tau[:, 0] + kap * dT[0]]
        ref_y = [0.0, tau[0, 1], tau[1, 1], vel @ tau[:, 1] + kap * dT[1]]
        assert np.allclose(Fv, ref_x, rtol=1e-12, atol=1e-14) and np.allclose(Gv, ref_y, rtol=1e-12, atol=1e-14)


def _tb(mesh, p, periodic=True, corr="dg", riemann="roe", Re=100.0, eta=2.0):
    per = (mesh.Lx, mesh.Ly) if periodic else None
    c = fr_case.build(mesh.xy, mesh.elem_nodes, p, periodic=per)
    c.pos = O.sp_coords(c.elem_nodes, c.xy, c.ops).reshape(c.E, -1, 2)
    return c, c.textbook(riemann, corr).set_viscous(Re, eta=eta)


@pytest.mark.parametrize("p", [1, 2, 3, 4])
def test_gradient_exact_on_continuous_polynomials(p):
    mesh = cases.uniform_quad_mesh(3, 3, 0.0, 1.5, -0.5, 1.0)
    c, tb = _tb(mesh, p, periodic=False)
    x, y = c.pos[..., 0].reshape(-1), c.pos[..., 1].reshape(-1)
    rng = np.random.default_rng(p)
    U = np.zeros((x.size, 4)); dx = np.zeros_like(U); dy = np.zeros_like(U)
    for v in range(4):
        a = rng.normal(size=(p + 1, p + 1))
        for i in range(p + 1):
            for j in range(p + 1 - i):
                U[:, v] += a[i, j] * x ** i * y ** j
                dx[:, v] += a[i, j] * i * x ** max(i - 1, 0) * y ** j
                dy[:, v] += a[i, j] * j * x ** i * y ** max(j - 1, 0)
    Q = tb.gradient(U.reshape(-1))
    assert np.allclose(Q[:, :4], dx, atol=1e-10) and np.allclose(Q[:, 4:], dy, atol=1e-10)


@pytest.mark.parametrize("corr", ["dg", "sd", "hu", "ga"])
def test_gradient_gauss_identity(corr):
    """sum over elements of int Q = boundary sum of U^C n = 0 (periodic)."""
    mesh = cases.uniform_quad_mesh(4, 3, 0.0, 1.0, 0.0, 1.0, jitter=0.0)
    c, tb = _tb(mesh, 3, corr=corr)
    U = cases.perturb(cases.tgv_ic(c.pos * 2 * math.pi, mach=0.3), 0.1, 5).reshape(-1)
    Q = tb.gradient(U)
    w = np.asarray(c.ops["qw"]).reshape(-1)
    tot = np.einsum("es,s,esk->k", c.J.reshape(c.E, -1), w, Q.reshape(c.E, -1, 8))
    scale = np.einsum("es,s,esk->k", c.J.reshape(c.E, -1), w, np.abs(Q.reshape(c.E, -1, 8)))
    assert np.all(np.abs(tot) <= 1e-12 * scale)


def test_freestream_and_conservation():
    mesh = cases.uniform_quad_mesh(4, 4, 0.0, 1.0, 0.0, 1.0)
    c, tb = _tb(mesh, 3, Re=50.0)
    shp = c.pos.shape[:2]
    Uf = cases.prim_to_cons(np.full(shp, 1.1), np.full(shp, 0.4), np.full(shp, -0.2), np.full(shp, 2.0), 1.4)
    assert np.abs(tb.residual(Uf.reshape(-1))).max() < 1e-12
    U = cases.perturb(cases.tgv_ic(c.pos * 2 * math.pi, mach=0.3), 0.05, 6).reshape(-1)
    R = tb.residual(U).reshape(c.E, -1, 4)
    w = np.asarray(c.ops["qw"]).reshape(-1)
    Jw = c.J.reshape(c.E, -1) * w
    tot = np.einsum("es,esv->v", Jw, R)
    scale = np.einsum("es,esv->v", Jw, np.abs(R))
    assert np.all(np.abs(tot) <= 1e-12 * scale)


def _max_real_eig(mesh, p, corr, eta):
    c, tb = _tb(mesh, p, Re=5.0, eta=eta, corr=corr)
    shp = c.pos.shape[:2]
    U0 = cases.prim_to_cons(np.full(shp, 1.0), np.full(shp, 0.3), np.full(shp, 0.1), np.full(shp, 1 / 1.4),
                            1.4).reshape(-1)
    n = U0.size
    Jm = np.zeros((n, n))
    for k in range(n):
        d = np.zeros(n)
        d[k] = 1e-7
        Jm[:, k] = (tb.residual(U0 + d) - tb.residual(U0 - d)) / 2e-7
    lam = np.linalg.eigvals(Jm)
    return lam.real.max() / np.abs(lam).max()


@pytest.mark.parametrize("p,corr", [(1, "hu"), (2, "dg"), (3, "sd"), (4, "hu")])
def test_linear_stability(p, corr):
    """The default penalty eta = 2 keeps the linearised scheme stable; with
    the unpenalised central flux (eta = 0) it is not."""
    mesh = cases.uniform_quad_mesh(2, 2, 0.0, 1.0, 0.0, 1.0)
    assert _max_real_eig(mesh, p, corr, 2.0) < 1e-6
    assert _max_real_eig(mesh, p, corr, 0.0) > 1e-2


def test_tgv_decay_rate():
    """Kinetic energy of the low-Mach 2D Taylor-Green vortex decays as
    exp(-4 nu t) (velocity ~ exp(-2 nu t), |k|^2 = 2)."""
    Re, steps, dt = 10.0, 500, 5e-4
    mesh = cases.tgv_mesh(6)
    c, tb = _tb(mesh, 4, Re=Re)
    U = cases.tgv_ic(c.pos, mach=0.05).reshape(-1)
    w = np.asarray(c.ops["qw"]).reshape(-1)
    Jw = (c.J.reshape(c.E, -1) * w).reshape(-1)

    def ke(X):
        X = X.reshape(-1, 4)
        return float(Jw @ (0.5 * (X[:, 1] ** 2 + X[:, 2] ** 2) / X[:, 0]))

    k0 = ke(U)
    for _ in range(steps):
        U = tb.advance(U, dt, False)
    rate = math.log(ke(U) / k0) / (steps * dt)
    assert abs(rate / (-4.0 / Re) - 1.0) < 0.02, rate
