"""Discrete adjoint of the Navier-Stokes terms (BR2, textbook mode) on an
MI355X: zj_residual_vjp, zj_rk3_step_vjp and zj_adjoint_get_re of a viscous
handle.

The pin is the identity <w, J v> = <J^T w, v>, with J v a central finite
difference (eps = 1e-7 relative) of the textbook ORACLE's forward map with its
viscous terms (oracle/zj_textbook.c, pinned in test_oracle_viscous.py; the GPU
forward equals it to 1e-12, test_gpu_viscous.py) and J^T w the HIP adjoint;
bar 1e-6 relative.  dJ/dRe is checked against Richardson-extrapolated central
differences of the oracle in Re.  The meshes hold several adjoint tiles, so
edge cotangents of U and grad U cross tiles through HBM.
"""
import numpy as np
import pytest

from tests import fr_case
from tests.test_gpu_adjoint import TOL, direction, dot_check, fd
from tests.test_gpu_parity import say
from tests.test_gpu_viscous import _sheared_mesh, _state
from zhijian_amd import cases, solver

pytestmark = pytest.mark.gpu

ETA = 2.0


def _pair(mesh, p, riemann="roe", corr="dg", Re=100.0, precision="fp64"):
    per = (mesh.Lx, mesh.Ly)
    t = solver.build_tables(mesh.xy, mesh.elem_nodes, p, periodic=per)
    s = solver.Solver(t, precision=precision, compat="textbook", riemann=riemann, corr=corr, positivity=False,
                      viscous=True, Re=Re, Pr=0.72, br2_eta=ETA)
    tb = fr_case.build(mesh.xy, mesh.elem_nodes, p, periodic=per).textbook(riemann, corr)
    return s, tb.set_viscous(Re, eta=ETA), t


def _richardson_re(f, Re):
    """d f / d Re from central differences at h = 2% and 1% of Re (error O(h^4))."""
    def central(h):
        return (f(Re + h) - f(Re - h)) / (2 * h)
    return (4.0 * central(0.01 * Re) - central(0.02 * Re)) / 3.0


@pytest.mark.parametrize("p", [1, 2, 3, 4, 5])
@pytest.mark.parametrize("riemann,corr", [("roe", "dg"), ("hllc", "hu")])
def test_viscous_residual_vjp(p, riemann, corr):
    say(f"case: viscous residual VJP, TGV 5x5 P{p} {riemann}/{corr} Re 100")
    s, tb, t = _pair(cases.tgv_mesh(5), p, riemann, corr)
    U = _state(t, 30 + p)
    s.set_state(U)
    w = np.random.default_rng(p).standard_normal(U.shape)
    v = direction(U, 40 + p)
    assert dot_check(float(w @ fd(tb.residual, U, v)), float(s.residual_vjp(w) @ v), "residual") < TOL


@pytest.mark.parametrize("p", [2, 3])
def test_viscous_residual_vjp_sheared(p):
    """Parallelograms whose metric changes from row to row."""
    say(f"case: viscous residual VJP, sheared periodic 6x6 P{p} hllc/sd Re 100")
    s, tb, t = _pair(_sheared_mesh(6), p, "hllc", "sd")
    U = _state(t, 50 + p)
    s.set_state(U)
    w = np.random.default_rng(51).standard_normal(U.shape)
    v = direction(U, 52)
    assert dot_check(float(w @ fd(tb.residual, U, v)), float(s.residual_vjp(w) @ v), "sheared") < TOL


@pytest.mark.parametrize("precision", ["fp64", "fp32bulk"])
@pytest.mark.parametrize("p", [2, 3])
def test_viscous_step_vjp_dt_re(precision, p):
    """One SSP-RK3 step: the cotangents of the state, of dt and of Re."""
    say(f"case: viscous step VJP, TGV 6x6 P{p} roe/hu Re 100 {precision}")
    Re = 100.0
    s, tb, t = _pair(cases.tgv_mesh(6), p, "roe", "hu", Re=Re, precision=precision)
    U = _state(t, 70 + p)
    dt = 1e-3
    ubar = np.random.default_rng(71).standard_normal(U.shape)
    v = direction(U, 72)
    lhs = float(ubar @ fd(lambda X: tb.advance(X, dt, False), U, v))
    s.set_state(U)
    s.adjoint_set(ubar)
    dtb = s.rk3_step_vjp(dt)
    g = s.adjoint_get()
    dre = s.adjoint_get_re()
    tol = TOL if precision == "fp64" else 2e-4
    assert dot_check(lhs, float(g @ v), "step") < tol
    h = 1e-7
    dfd = (float(ubar @ tb.advance(U, dt + h, False)) - float(ubar @ tb.advance(U, dt - h, False))) / (2 * h)
    rd = abs(dtb - dfd) / abs(dfd)
    say(f"  dJ/ddt: adjoint {dtb:.12e}  FD {dfd:.12e}  rel {rd:.2e}")
    assert rd < tol

    def J_of_re(r):
        tb.set_viscous(r, eta=ETA)
        return float(ubar @ tb.advance(U, dt, False))
    refd = _richardson_re(J_of_re, Re)
    tb.set_viscous(Re, eta=ETA)
    rr = abs(dre - refd) / abs(refd)
    say(f"  dJ/dRe: adjoint {dre:.12e}  Richardson FD {refd:.12e}  rel {rr:.2e}")
    assert rr < tol


def test_viscous_multi_step_vjp():
    """Three steps backwards: the pending 12-word edge cotangents pass from
    one step's stage 1 to the previous step's stage 3; dJ/dRe accumulates."""
    say("case: viscous 3-step VJP, TGV 8x8 P3 hllc/dg Re 50")
    Re = 50.0
    s, tb, t = _pair(cases.tgv_mesh(8), 3, "hllc", "dg", Re=Re)
    U = _state(t, 80)
    dt, steps = 1e-3, 3
    traj = [U]
    for _ in range(steps):
        traj.append(tb.advance(traj[-1], dt, False))
    ubar = np.random.default_rng(81).standard_normal(U.shape)
    v = direction(U, 82)

    def fwd(X):
        for _ in range(steps):
            X = tb.advance(X, dt, False)
        return X
    lhs = float(ubar @ fd(fwd, U, v))
    s.adjoint_set(ubar)
    for n in range(steps - 1, -1, -1):
        s.set_state(traj[n])
        s.rk3_step_vjp(dt, want_dt=False)
    assert dot_check(lhs, float(s.adjoint_get() @ v), "3 steps") < TOL

    def J_of_re(r):
        tb.set_viscous(r, eta=ETA)
        return float(ubar @ fwd(U))
    refd = _richardson_re(J_of_re, Re)
    tb.set_viscous(Re, eta=ETA)
    dre = s.adjoint_get_re()
    rr = abs(dre - refd) / abs(refd)
    say(f"  dJ/dRe: adjoint {dre:.12e}  Richardson FD {refd:.12e}  rel {rr:.2e}")
    assert rr < TOL
