"""Discrete adjoint (zj_residual_vjp, zj_rk3_step_vjp) on an MI355X.

The adjoint has no reference counterpart; what pins it is the mathematics:
for any directions w and v,  <w, J v> = <J^T w, v>, where J v is taken by a
central finite difference of the ORACLE's forward map (FP64, eps = 1e-7
relative) and J^T w is the HIP adjoint.  The two sides agree to the FD
truncation and round-off, about 1e-8 relative; the bar is 1e-6.  A
transposed operand, a dropped term, a wrong sign or a mirrored index in any
VJP moves the HIP side by O(1) relative.
"""
import sys

import numpy as np
import pytest

from tests.test_gpu_parity import make, rel, say, tgv_state, vortex_state
from zhijian_amd import cases, solver

pytestmark = pytest.mark.gpu

EPS = 1e-7
TOL = 1e-6


def direction(U, seed):
    rng = np.random.default_rng(seed)
    return U * rng.standard_normal(U.shape)


def fd(f, U, v):
    return (f(U + EPS * v) - f(U - EPS * v)) / (2 * EPS)


def dot_check(lhs, rhs, what):
    r = abs(lhs - rhs) / max(abs(lhs), abs(rhs), 1e-300)
    say(f"  {what}: <w, J v> = {lhs:.12e}  <J^T w, v> = {rhs:.12e}  rel {r:.2e}")
    return r


@pytest.mark.parametrize("p", [1, 2, 3, 4, 5])
def test_residual_vjp_all_orders(p):
    """Jittered (per-point metric) and uniform (per-element metric) meshes."""
    for jit in (0.12, 0.0):
        say(f"case: residual VJP, periodic 5x4 P{p} jitter {jit}")
        mesh = cases.uniform_quad_mesh(5, 4, 0.0, 1.0, 0.0, 1.0, jitter=jit, seed=p)
        s, oc, t = make(mesh, p)
        U = tgv_state(t, amp=0.05, seed=10 + p, mach=0.3)
        s.set_state(U)
        o = oc.solver()
        w = np.random.default_rng(p).standard_normal(U.shape)
        v = direction(U, 20 + p)
        lhs = float(w @ fd(o.residual, U, v))
        rhs = float(s.residual_vjp(w) @ v)
        assert dot_check(lhs, rhs, "residual") < TOL


@pytest.mark.parametrize("bct", [1, 2, 3, 4, 6, 0])
def test_residual_vjp_boundaries(bct):
    """Every boundary face of one BC type (wall, slip wall, symmetry, far
    field, outflow, and the copy state of the remaining types)."""
    say(f"case: residual VJP, open 5x4 P3, boundary type {bct}")
    mesh = cases.uniform_quad_mesh(5, 4, 0.0, 1.0, 0.0, 1.0, jitter=0.15, seed=3)
    faces = solver._lib.build_faces(mesh.elem_nodes)
    types = np.zeros(faces.shape[0], np.int32)
    types[faces[:, 2] < 0] = bct
    fs = (1.1, 0.3, 0.1, 1.5e3)
    s, oc, t = make(mesh, 3, periodic=False, bc_types=types, fs=fs)
    shp = t.pos.shape[:2]
    U = cases.perturb(cases.prim_to_cons(np.full(shp, 1.0), np.full(shp, 0.3), np.full(shp, -0.1),
                                         np.full(shp, 2e3), 1.4), 0.05, 4).reshape(-1)
    s.set_state(U)
    o = oc.solver()
    w = np.random.default_rng(5).standard_normal(U.shape)
    v = direction(U, 6)
    assert dot_check(float(w @ fd(o.residual, U, v)), float(s.residual_vjp(w) @ v), f"residual, BC {bct}") < TOL


def test_residual_vjp_self_neighbour():
    say("case: residual VJP, one periodic element P2 (its own neighbour on every edge)")
    mesh = cases.uniform_quad_mesh(1, 1, 0.0, 1.0, 0.0, 1.0)
    s, oc, t = make(mesh, 2)
    U = tgv_state(t, amp=0.05, seed=3, mach=0.3)
    s.set_state(U)
    w = np.random.default_rng(7).standard_normal(U.shape)
    v = direction(U, 8)
    assert dot_check(float(w @ fd(oc.solver().residual, U, v)), float(s.residual_vjp(w) @ v), "1x1") < TOL


@pytest.mark.parametrize("precision", ["fp64", "fp32bulk"])
def test_step_vjp_and_dt(precision):
    """One SSP-RK3 step with the positivity limiter: state and dt cotangents."""
    say(f"case: step VJP, vortex 6x6 P3 {precision}")
    mesh = cases.vortex_mesh(6)
    s, oc, t = make(mesh, 3, precision=precision)
    U = vortex_state(t, amp=0.05, seed=11)
    o = oc.solver()
    dt = 2e-3
    ubar = np.random.default_rng(12).standard_normal(U.shape)
    v = direction(U, 13)
    lhs = float(ubar @ fd(lambda X: o.advance(X, dt, True), U, v))
    s.set_state(U)
    s.adjoint_set(ubar)
    dtb = s.rk3_step_vjp(dt)
    g = s.adjoint_get()
    tol = TOL if precision == "fp64" else 2e-4
    assert dot_check(lhs, float(g @ v), "step") < tol
    h = 1e-7
    dfd = (float(ubar @ o.advance(U, dt + h, True)) - float(ubar @ o.advance(U, dt - h, True))) / (2 * h)
    rd = abs(dtb - dfd) / abs(dfd)
    say(f"  dJ/ddt: adjoint {dtb:.12e}  FD {dfd:.12e}  rel {rd:.2e}")
    assert rd < (1e-6 if precision == "fp64" else 2e-4)


def test_multi_step_adjoint_tgv():
    """Three steps of TGV 8x8 P4 in reverse, the cotangent held on the device
    between steps (cross-tile edge cotangents folded in by the next launch)."""
    say("case: 3-step adjoint, TGV 8x8 P4")
    mesh = cases.tgv_mesh(8)
    s, oc, t = make(mesh, 4)
    U0 = tgv_state(t, amp=0.02, seed=14)
    o = oc.solver()
    dt = 1e-2
    steps = 3

    def run(X):
        for _ in range(steps):
            X = o.advance(X, dt, True)
        return X

    traj = [U0]
    for _ in range(steps):
        traj.append(o.advance(traj[-1], dt, True))
    ubar = np.random.default_rng(15).standard_normal(U0.shape)
    v = direction(U0, 16)
    lhs = float(ubar @ fd(run, U0, v))
    s.set_state(traj[-1])
    s.adjoint_set(ubar)
    for n in range(steps - 1, -1, -1):
        s.set_state(traj[n])
        s.rk3_step_vjp(dt)
    assert dot_check(lhs, float(s.adjoint_get() @ v), "3 steps") < TOL


def test_adjoint_affine_equals_per_point():
    say("case: step VJP with per-element vs per-point metric, TGV 8x8 P3")
    mesh = cases.tgv_mesh(8)
    t = solver.build_tables(mesh.xy, mesh.elem_nodes, 3, periodic=(mesh.Lx, mesh.Ly))
    a, b = solver.Solver(t), solver.Solver(t)
    assert a.set_geometry_mode(-1) == 1 and b.set_geometry_mode(0) == 0
    U = tgv_state(t, amp=0.02, seed=17)
    ubar = np.random.default_rng(18).standard_normal(U.shape)
    out = []
    for s in (a, b):
        s.set_state(U)
        s.adjoint_set(ubar)
        s.rk3_step_vjp(5e-3)
        out.append(s.adjoint_get())
    assert rel(out[0], out[1]) < 1e-12


@pytest.mark.parametrize("precision", ["fp64", "fp32bulk"])
def test_adjoint_c5_full_size(precision):
    """The adjoint configuration at full size (TGV 128x128 P3, Mach 0.1,
    16,384 elements, limiter on): the residual VJP and one step's VJP
    (state and dt cotangents) against central differences of the oracle's
    forward map, through the dot-product identity."""
    say(f"case: adjoint at full size, TGV 128x128 P3 {precision}: residual VJP, step VJP, dJ/ddt")
    mesh = cases.tgv_mesh(128)
    s, oc, t = make(mesh, 3, precision=precision)
    U = tgv_state(t, amp=0.02, seed=41, mach=0.1)
    o = oc.solver()
    tol = TOL if precision == "fp64" else 2e-4
    w = np.random.default_rng(42).standard_normal(U.shape)
    v = direction(U, 43)
    s.set_state(U)
    if precision == "fp64":
        lhs = float(w @ fd(o.residual, U, v))
        assert dot_check(lhs, float(s.residual_vjp(w) @ v), "residual") < tol
    dt = 0.5 * s.dt_cfl(0.5)
    lhs = float(w @ fd(lambda X: o.advance(X, dt, True), U, v))
    s.set_state(U)
    s.adjoint_set(w)
    dtb = s.rk3_step_vjp(dt)
    assert dot_check(lhs, float(s.adjoint_get() @ v), "step") < tol
    # the step is polynomial in dt (no clamp changes branch), so a wide central
    # difference is accurate to (h/dt)^2; a narrow one drowns in the round-off
    # of w.U (1e5 in size here) against 2h dJ/ddt
    h = 1e-4 * dt
    dfd = (float(w @ o.advance(U, dt + h, True)) - float(w @ o.advance(U, dt - h, True))) / (2 * h)
    rd = abs(dtb - dfd) / abs(dfd)
    say(f"  dJ/ddt: adjoint {dtb:.12e}  FD {dfd:.12e}  rel {rd:.2e}")
    assert rd < tol


def test_dt_cfl_vjp():
    """The CFL step's VJP (zj_dt_cfl_vjp): one point's four variables carry
    the derivative of the minimum; against central differences of the
    oracle's CFL step (fr_solver.cpp:784-816 restated) in three directions."""
    say("case: CFL step VJP, jittered 6x5 P3")
    mesh = cases.uniform_quad_mesh(6, 5, 0.0, 1.0, 0.0, 1.0, jitter=0.12, seed=21)
    s, oc, t = make(mesh, 3)
    U = tgv_state(t, amp=0.05, seed=22, mach=0.3)
    o = oc.solver()
    s.set_state(U)
    s.adjoint_set(np.zeros_like(U))
    s.dt_cfl_vjp(0.5, 1.0)
    g = s.adjoint_get()
    assert np.count_nonzero(g) == 4
    for k in range(3):
        v = direction(U, 23 + k)
        lhs = (o.dt_cfl(U + EPS * v, 0.5) - o.dt_cfl(U - EPS * v, 0.5)) / (2 * EPS)
        assert dot_check(lhs, float(g @ v), "dt_cfl") < TOL


def test_dt_cfl_vjp_viscous():
    """With the viscous terms the step is the minimum of
    cfl / ((2p+1) ws / h + (p+1)^4 nu_k / (rho h^2)); against central
    differences of the handle's own zj_dt_cfl (pinned by test_viscous_dt_rule)."""
    say("case: CFL step VJP with the viscous rate, TGV 6x6 P2 Re 20")
    mesh = cases.tgv_mesh(6)
    t = solver.build_tables(mesh.xy, mesh.elem_nodes, 2, periodic=(mesh.Lx, mesh.Ly))
    s = solver.Solver(t, compat="textbook", riemann="roe", corr="hu", positivity=False, viscous=True, Re=20.0)
    U = tgv_state(t, amp=0.05, seed=27, mach=0.3)
    s.set_state(U)
    s.adjoint_set(np.zeros_like(U))
    s.dt_cfl_vjp(0.5, 2.5)
    g = s.adjoint_get()

    def dt_of(X):
        s.set_state(X)
        return s.dt_cfl(0.5)

    for k in range(2):
        v = direction(U, 28 + k)
        lhs = 2.5 * (dt_of(U + EPS * v) - dt_of(U - EPS * v)) / (2 * EPS)
        assert dot_check(lhs, float(g @ v), "dt_cfl viscous") < TOL


def test_cfl_trajectory_gradient():
    """Three steps with dt_n = dt_cfl(U^n): J = <w, U^3>; the reverse sweep
    chains each step's dJ/d(dt_n) (zj_rk3_step_vjp) into the CFL step's VJP
    (zj_dt_cfl_vjp); against a central difference of the oracle's map with
    the step recomputed from each state."""
    say("case: gradient through 3 CFL-stepped steps, TGV 6x6 P3")
    mesh = cases.tgv_mesh(6)
    s, oc, t = make(mesh, 3)
    U0 = tgv_state(t, amp=0.02, seed=31, mach=0.3)
    o = oc.solver()
    steps = 3

    def run(X):
        for _ in range(steps):
            X = o.advance(X, o.dt_cfl(X, 0.5), True)
        return X

    traj, dts = [U0], []
    for _ in range(steps):
        dts.append(o.dt_cfl(traj[-1], 0.5))
        traj.append(o.advance(traj[-1], dts[-1], True))
    w = np.random.default_rng(32).standard_normal(U0.shape)
    v = direction(U0, 33)
    lhs = float(w @ fd(run, U0, v))
    s.adjoint_set(w)
    for n in range(steps - 1, -1, -1):
        s.set_state(traj[n])
        dtb = s.rk3_step_vjp(dts[n])
        s.dt_cfl_vjp(0.5, dtb)
    assert dot_check(lhs, float(s.adjoint_get() @ v), "3 CFL steps") < TOL
