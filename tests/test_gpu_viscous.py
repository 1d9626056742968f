"""Viscous textbook terms (BR2) on an MI355X against the textbook oracle with
its viscous terms (oracle/zj_textbook.c, pinned on the CPU in
test_oracle_viscous.py).

Bar: FP64 residual within 1e-12 relative L2 of the oracle (P1-P5, Roe/HLLC,
DG/HU corrections); 30 steps within 1e-10; FP32-bulk residual within 1e-5;
the Taylor-Green kinetic-energy decay rate within 2% of -4/Re; FP32-bulk vs
FP64 after 100 steps of TGV at Re = 1600 within 1e-6.
"""
import math

import numpy as np
import pytest

from oracle import pyoracle as O
from tests import fr_case
from tests.test_gpu_parity import rel, say
from zhijian_amd import cases, solver
from zhijian_amd._lib import ZJError

pytestmark = pytest.mark.gpu


def _pair(mesh, p, riemann="roe", corr="dg", Re=100.0, precision="fp64", eta=2.0):
    per = (mesh.Lx, mesh.Ly)
    t = solver.build_tables(mesh.xy, mesh.elem_nodes, p, periodic=per)
    s = solver.Solver(t, precision=precision, compat="textbook", riemann=riemann, corr=corr, positivity=False,
                      viscous=True, Re=Re, Pr=0.72, br2_eta=eta)
    oc = fr_case.build(mesh.xy, mesh.elem_nodes, p, periodic=per)
    return s, oc.textbook(riemann, corr).set_viscous(Re, eta=eta), t


def _state(t, seed, mach=0.3):
    return cases.perturb(cases.tgv_ic(t.pos, mach=mach), 0.05, seed).reshape(-1)


@pytest.mark.parametrize("p", [1, 2, 3, 4, 5])
@pytest.mark.parametrize("riemann,corr", [("roe", "dg"), ("hllc", "hu")])
def test_viscous_residual(p, riemann, corr):
    say(f"case: viscous residual, TGV 5x5 P{p} {riemann}/{corr} Re 100")
    s, tb, t = _pair(cases.tgv_mesh(5), p, riemann, corr)
    U = _state(t, 30 + p)
    s.set_state(U)
    r = rel(s.residual(), tb.residual(U))
    say(f"  rel L2 {r:.3e}")
    assert r < 1e-12


def _sheared_mesh(n, L=2 * math.pi):
    """Periodic n x n mesh of parallelograms whose shear and height change
    from row to row (affine elements with different metrics)."""
    h = L / n
    ys = np.cumsum(np.r_[0.0, h * (1.0 + 0.3 * np.cos(2 * math.pi * np.arange(n) / n))])
    ys *= L / ys[-1]
    shift = 0.25 * h * (np.arange(n + 1) % 2)
    X = np.arange(n + 1)[None, :] * h + shift[:, None]
    Y = np.repeat(ys[:, None], n + 1, axis=1)
    m = cases.uniform_quad_mesh(n, n, 0.0, L, 0.0, L)
    m.xy = np.stack([X.reshape(-1), Y.reshape(-1)], axis=1)
    return m


@pytest.mark.parametrize("p", [2, 3])
def test_viscous_residual_sheared(p):
    say(f"case: viscous residual, sheared periodic 6x6 P{p} hllc/sd Re 100")
    s, tb, t = _pair(_sheared_mesh(6), p, "hllc", "sd")
    assert s.info()["affine"]
    U = _state(t, 50 + p)
    s.set_state(U)
    r = rel(s.residual(), tb.residual(U))
    say(f"  rel L2 {r:.3e}")
    assert r < 1e-12


def test_viscous_steps():
    say("case: viscous 30 steps, TGV 8x8 P3 roe/dg Re 50")
    s, tb, t = _pair(cases.tgv_mesh(8), 3, Re=50.0)
    U = _state(t, 7)
    s.set_state(U)
    dt = 1e-3
    V = U.copy()
    for _ in range(30):
        s.rk3_step(dt)
        V = tb.advance(V, dt, False)
    r = rel(s.get_state(), V)
    say(f"  rel L2 after 30 steps {r:.3e}")
    assert r < 1e-10


@pytest.mark.parametrize("p", [1, 2, 3, 4, 5])
def test_viscous_fp32bulk_residual(p):
    say(f"case: viscous FP32-bulk residual vs oracle, TGV 6x6 P{p} roe/hu Re 100")
    s, tb, t = _pair(cases.tgv_mesh(6), p, "roe", "hu", precision="fp32bulk")
    U = _state(t, 40 + p)
    s.set_state(U)
    r = rel(s.residual(), tb.residual(U))
    say(f"  rel L2 {r:.3e}")
    assert r < 1e-5


def test_tgv_decay_rate_gpu():
    say("case: TGV kinetic-energy decay, 8x8 P4, Re 10, Mach 0.05")
    Re = 10.0
    mesh = cases.tgv_mesh(8)
    t = solver.build_tables(mesh.xy, mesh.elem_nodes, 4, periodic=(mesh.Lx, mesh.Ly))
    s = solver.Solver(t, compat="textbook", riemann="roe", corr="dg", positivity=False, viscous=True, Re=Re)
    U = cases.tgv_ic(t.pos, mach=0.05).reshape(-1)
    w = np.asarray(t.ops["qw"]).reshape(-1)
    Jw = (t.J.reshape(t.E, -1) * w).reshape(-1)

    def ke(X):
        X = X.reshape(-1, 4)
        return float(Jw @ (0.5 * (X[:, 1] ** 2 + X[:, 2] ** 2) / X[:, 0]))

    s.set_state(U)
    steps, dt = 500, 5e-4
    for _ in range(steps):
        s.rk3_step(dt)
    rate = math.log(ke(s.get_state()) / ke(U)) / (steps * dt)
    say(f"  d ln KE/dt = {rate:.6f}, -4/Re = {-4.0 / Re:.6f}")
    assert abs(rate / (-4.0 / Re) - 1.0) < 0.02


def test_viscous_fp32_after_100_steps():
    say("case: viscous FP32-bulk vs FP64 after 100 steps, TGV 32x32 P4 roe/hu Re 1600 Mach 0.1")
    mesh = cases.tgv_mesh(32)
    t = solver.build_tables(mesh.xy, mesh.elem_nodes, 4, periodic=(mesh.Lx, mesh.Ly))
    kw = dict(compat="textbook", riemann="roe", corr="hu", positivity=False, viscous=True, Re=1600.0)
    a = solver.Solver(t, **kw)
    b = solver.Solver(t, precision="fp32bulk", **kw)
    U = cases.tgv_ic(t.pos, mach=0.1).reshape(-1)
    a.set_state(U)
    b.set_state(U)
    dt = a.dt_cfl(0.5)
    for _ in range(100):
        a.rk3_step(dt)
        b.rk3_step(dt)
    r = rel(b.get_state(), a.get_state())
    say(f"  FP32-bulk vs FP64 after 100 steps: rel L2 {r:.3e}")
    assert r < 1e-6


@pytest.mark.parametrize("Re", [20.0, 1600.0])
def test_viscous_dt_rule(Re):
    """zj_dt_cfl of a viscous handle against the rule evaluated here: the
    minimum over the points of cfl / ((2p+1)(|u|+c)/h + (p+1)^4 nu_k /
    (rho h^2)), h = sqrt(area), nu_k = mu max(4/3, gamma/Pr).  The BR2
    operator's spectral radius on squares of side h is at most 2.0 (p+1)^4
    nu_k / h^2 for P1-P5 (measured on the oracle's Jacobian), so at cfl 0.5
    the viscous part alone keeps dt * rho(J) near 1, inside SSP-RK3's real
    interval (-2.51, 0).  Sheared parallelograms: h varies by element."""
    p, g, Pr = 3, 1.4, 0.72
    s, _, t = _pair(_sheared_mesh(8), p, "hllc", "hu", Re=Re)
    U = _state(t, 60)
    s.set_state(U)
    q = U.reshape(t.E, t.S, 4)
    u, w = q[..., 1] / q[..., 0], q[..., 2] / q[..., 0]
    pr = (g - 1) * q[..., 0] * (q[..., 3] / q[..., 0] - 0.5 * (u * u + w * w))
    ws = np.sqrt(u * u + w * w) + np.sqrt(g * np.maximum(pr, 1e-10) / np.maximum(q[..., 0], 1e-10))
    h = np.sqrt(t.area)[:, None]
    visc = (p + 1) ** 4 * (1.0 / Re) * max(4.0 / 3.0, g / Pr)
    dt = float(np.min(0.5 / ((2 * p + 1) * ws / h + visc / (q[..., 0] * h * h))))
    got = s.dt_cfl(0.5)
    conv = float(np.min(0.5 * h / ((2 * p + 1) * ws)))
    say(f"  Re {Re:g}: dt {got:.6e} vs {dt:.6e} (convective alone {conv:.6e})")
    assert abs(got - dt) <= 1e-13 * dt
    assert got < conv


def test_viscous_rejects_unsupported_meshes():
    say("case: viscous terms on jittered / open meshes are refused")
    jit = cases.uniform_quad_mesh(4, 4, 0.0, 1.0, 0.0, 1.0, jitter=0.1, seed=1)
    t = solver.build_tables(jit.xy, jit.elem_nodes, 2, periodic=(1.0, 1.0))
    with pytest.raises(ZJError) as e:
        solver.Solver(t, compat="textbook", positivity=False, viscous=True)
    assert e.value.code == 4
    open_ = cases.uniform_quad_mesh(4, 4, 0.0, 1.0, 0.0, 1.0)
    t = solver.build_tables(open_.xy, open_.elem_nodes, 2)
    with pytest.raises(ZJError) as e:
        solver.Solver(t, compat="textbook", positivity=False, viscous=True)
    assert e.value.code == 4
