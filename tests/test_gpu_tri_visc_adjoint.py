"""Discrete adjoint of the triangle stage with the Navier-Stokes terms (BR2,
textbook scheme) on an MI355X: zj_tri_residual_vjp and zj_tri_rk3_step_vjp on
viscous triangle handles (tri_visc_vjp_kernel).

The pin is the identity <w, J v> = <J^T w, v>, with J v a central finite
difference (eps = 1e-7 relative) of the triangle ORACLE's forward map with its
viscous terms (oracle/tri_oracle.py, FP64; pinned on the CPU in
test_tri_oracle.py) and J^T w the HIP adjoint; bar 1e-6 relative for FP64,
2e-4 for FP32-bulk.  The meshes hold several tiles, so the 12-word edge
cotangents that cross tiles (traces of U and of its gradient, written to HBM
by one workgroup and collected by the next launch or the completion kernel)
are on the path; a Hilbert-ordered handle and one in the caller's order (other
tiles, other cross-tile edges) must agree to round-off.
"""
import math

import numpy as np
import pytest

from oracle import tri_oracle as T
from tests.test_gpu_adjoint import TOL, direction, dot_check, fd
from tests.test_gpu_parity import say
from zhijian_amd import cases
from zhijian_amd.tri import TriSolver, sp_coords

pytestmark = pytest.mark.gpu

TOL32 = 2e-4


def _pair(n, p, riemann, corr="dg", Re=100.0, precision="fp64", order="hilbert", seed=1):
    m = cases.split_tri_mesh(n, jitter=0.2, seed=seed)
    s = TriSolver(m.xy, m.tris, p, periodic=(1.0, 1.0), riemann=riemann, precision=precision, order=order, corr=corr,
                  viscous=True, Re=Re)
    o = T.TriTextbook(p, m.xy, m.tris, s.faces, s.normals, riemann=riemann, corr=corr).set_viscous(Re)
    return s, o, m


def _state(m, p, seed, mach=0.3):
    pos = sp_coords(m.xy, m.tris, p)
    return cases.perturb(cases.tgv_ic(pos * 2 * math.pi / m.L, mach=mach), 0.05, seed).reshape(-1)


@pytest.mark.parametrize("p", [1, 2, 3])
@pytest.mark.parametrize("riemann,corr", [("roe", "dg"), ("hllc", "ga")])
def test_tri_viscous_residual_vjp(p, riemann, corr):
    say(f"case: triangle viscous residual VJP, periodic 6x6 jitter 0.2 P{p} {riemann}/{corr} Re 100")
    s, o, m = _pair(6, p, riemann, corr, seed=p)
    assert s.info()["cross_tile_edges"] > 0
    U = _state(m, p, 80 + p)
    s.set_state(U)
    w = np.random.default_rng(p).standard_normal(U.shape)
    v = direction(U, 90 + p)
    assert dot_check(float(w @ fd(o.residual, U, v)), float(s.residual_vjp(w) @ v), "residual") < TOL


@pytest.mark.parametrize("p", [2, 3, 4, 5])
def test_tri_viscous_residual_vjp_fp32bulk(p):
    say(f"case: triangle viscous residual VJP, FP32-bulk, periodic 6x6 P{p} roe/hu Re 100")
    s, o, m = _pair(6, p, "roe", "hu", precision="fp32bulk", seed=p)
    U = _state(m, p, 100 + p)
    s.set_state(U)
    w = np.random.default_rng(110 + p).standard_normal(U.shape)
    v = direction(U, 120 + p)
    assert dot_check(float(w @ fd(o.residual, U, v)), float(s.residual_vjp(w) @ v), "residual") < TOL32


@pytest.mark.parametrize("precision,p", [("fp64", 2), ("fp64", 3), ("fp32bulk", 3), ("fp32bulk", 5)])
def test_tri_viscous_step_vjp_and_dt(precision, p):
    """One SSP-RK3 step: the state's and dt's cotangents."""
    say(f"case: triangle viscous step VJP, periodic 6x6 P{p} hllc/ga Re 100 {precision}")
    s, o, m = _pair(6, p, "hllc", "ga", precision=precision, seed=p)
    U = _state(m, p, 130 + p)
    dt = 1e-3
    ubar = np.random.default_rng(131).standard_normal(U.shape)
    v = direction(U, 132)
    lhs = float(ubar @ fd(lambda X: o.advance(X, dt), U, v))
    s.set_state(U)
    s.adjoint_set(ubar)
    dtb = s.rk3_step_vjp(dt)
    tol = TOL if precision == "fp64" else TOL32
    assert dot_check(lhs, float(s.adjoint_get() @ v), "step") < tol
    h = 1e-7
    dfd = (float(ubar @ o.advance(U, dt + h)) - float(ubar @ o.advance(U, dt - h))) / (2 * h)
    rd = abs(dtb - dfd) / abs(dfd)
    say(f"  dJ/ddt: adjoint {dtb:.12e}  FD {dfd:.12e}  rel {rd:.2e}")
    assert rd < tol


def test_tri_viscous_multi_step_vjp():
    """Four steps backwards on a mesh of many tiles, the trajectory's states
    set before each step's VJP: the pending 12-word cross-tile cotangents
    pass from one step's stage 1 to the previous step's stage 3."""
    say("case: triangle viscous 4-step VJP, periodic 8x8 P3 roe/dg Re 200")
    s, o, m = _pair(8, 3, "roe", "dg", Re=200.0, seed=8)
    U = _state(m, 3, 140)
    dt, steps = 1e-3, 4
    traj = [U]
    for _ in range(steps):
        traj.append(o.advance(traj[-1], dt))
    ubar = np.random.default_rng(141).standard_normal(U.shape)
    v = direction(U, 142)

    def run(X):
        for _ in range(steps):
            X = o.advance(X, dt)
        return X

    lhs = float(ubar @ fd(run, U, v))
    s.adjoint_set(ubar)
    for n in range(steps - 1, -1, -1):
        s.set_state(traj[n])
        s.rk3_step_vjp(dt, want_dt=False)
    assert dot_check(lhs, float(s.adjoint_get() @ v), "4 steps") < TOL


def test_tri_viscous_vjp_numbering_independent():
    """A Hilbert-ordered handle and one in the caller's order (different
    tiles, so different cross-tile edges) give the same cotangent to round-off."""
    say("case: triangle viscous residual VJP, Hilbert vs caller order, periodic 12x12 P2 roe/dg")
    a, _, m = _pair(12, 2, "roe", order="hilbert", seed=12)
    b, _, _ = _pair(12, 2, "roe", order="none", seed=12)
    U = _state(m, 2, 150)
    w = np.random.default_rng(151).standard_normal(U.shape)
    out = []
    for s in (a, b):
        s.set_state(U)
        out.append(s.residual_vjp(w))
    r = np.linalg.norm(out[0] - out[1]) / np.linalg.norm(out[0])
    say(f"  rel L2 {r:.3e}")
    assert r < 1e-12


def test_tri_viscous_vjp_fp64_p5():
    """FP64 P5: the adjoint's 16-element tile takes 91 KiB of LDS (one
    workgroup per CU); residual and one step against the oracle."""
    say("case: triangle viscous FP64 P5 residual and step VJP, periodic 6x6 roe/dg Re 100")
    s, o, m = _pair(6, 5, "roe", seed=5)
    U = _state(m, 5, 160)
    s.set_state(U)
    w = np.random.default_rng(161).standard_normal(U.shape)
    v = direction(U, 162)
    assert dot_check(float(w @ fd(o.residual, U, v)), float(s.residual_vjp(w) @ v), "residual") < TOL
    dt = 5e-4
    lhs = float(w @ fd(lambda X: o.advance(X, dt), U, v))
    s.adjoint_set(w)
    s.rk3_step_vjp(dt, want_dt=False)
    assert dot_check(lhs, float(s.adjoint_get() @ v), "step") < TOL
