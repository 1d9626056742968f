"""Discrete adjoint on triangles (zj_tri_residual_vjp, zj_tri_rk3_step_vjp)
on an MI355X.

As for quadrilaterals (test_gpu_adjoint.py), the pin is the identity
<w, J v> = <J^T w, v>, with J v a central finite difference (eps = 1e-7
relative) of the triangle ORACLE's forward map (oracle/tri_oracle.py, FP64)
and J^T w the HIP adjoint; bar 1e-6 relative.  The meshes hold more than one
64-element tile, so the edge cotangents that cross tiles (written to HBM by
one workgroup, collected by the next launch or the completion kernel) are on
the path; the multi-step case carries them from one step to the next.
"""
import math

import numpy as np
import pytest

from oracle import tri_oracle as T
from tests.test_gpu_adjoint import TOL, direction, dot_check, fd
from tests.test_gpu_parity import say
from zhijian_amd import cases
from zhijian_amd.tri import TriSolver, sp_coords, tri_faces

pytestmark = pytest.mark.gpu


def _periodic(n, p, riemann, corr="dg", precision="fp64", order="hilbert", jitter=0.2, seed=1):
    m = cases.split_tri_mesh(n, jitter=jitter, seed=seed)
    s = TriSolver(m.xy, m.tris, p, periodic=(1.0, 1.0), riemann=riemann, precision=precision, order=order,
                  corr=corr)
    o = T.TriTextbook(p, m.xy, m.tris, s.faces, s.normals, riemann=riemann, corr=corr)
    return s, o, m


def _state(m, p, seed, mach=0.3):
    pos = sp_coords(m.xy, m.tris, p)
    return cases.perturb(cases.tgv_ic(pos * 2 * math.pi / m.L, mach=mach), 0.05, seed).reshape(-1)


def _advance(o, U, dt, steps=1):
    for _ in range(steps):
        U = o.advance(U, dt)
    return U


@pytest.mark.parametrize("p", [1, 2, 3, 4, 5])
@pytest.mark.parametrize("riemann", ["rusanov", "roe", "hllc"])
def test_tri_residual_vjp(p, riemann):
    say(f"case: triangle residual VJP, periodic 6x6 jitter 0.2 P{p} {riemann}")
    s, o, m = _periodic(6, p, riemann, seed=p)
    assert s.info()["cross_tile_edges"] > 0
    U = _state(m, p, 20 + p)
    s.set_state(U)
    w = np.random.default_rng(p).standard_normal(U.shape)
    v = direction(U, 30 + p)
    assert dot_check(float(w @ fd(o.residual, U, v)), float(s.residual_vjp(w) @ v), "residual") < TOL


@pytest.mark.parametrize("corr", ["sd", "hu", "ga"])
def test_tri_residual_vjp_corrections(corr):
    say(f"case: triangle residual VJP, periodic 6x6 P3 roe/{corr}")
    s, o, m = _periodic(6, 3, "roe", corr=corr, seed=3)
    U = _state(m, 3, 41)
    s.set_state(U)
    w = np.random.default_rng(42).standard_normal(U.shape)
    v = direction(U, 43)
    assert dot_check(float(w @ fd(o.residual, U, v)), float(s.residual_vjp(w) @ v), corr) < TOL


@pytest.mark.parametrize("bct", [1, 2, 4, 6, 0])
def test_tri_residual_vjp_boundaries(bct):
    """Open mesh, every boundary face of one type (wall, slip wall, far field,
    outflow, and the copy state)."""
    say(f"case: triangle residual VJP, open 6x6 P3 hllc, boundary type {bct}")
    m = cases.split_tri_mesh(6, jitter=0.15, seed=4)
    faces, nrm = tri_faces(m.xy, m.tris)
    bt = np.where(faces[:, 2] < 0, bct, 0).astype(np.int32)
    bmap = np.where(bt == 6, 1, 0).astype(np.int32)
    bcd = np.array([[1.1, 0.3, 0.1, 1.5e3, 0, 0, 0, 0], [1.4e3, 0, 0, 0, 0, 0, 0, 0]], np.float64)
    s = TriSolver(m.xy, m.tris, 3, faces=faces, normals=nrm, bc_type=bt, bc_face_map=bmap, bc_data=bcd,
                  riemann="hllc")
    o = T.TriTextbook(3, m.xy, m.tris, faces, nrm, riemann="hllc", bc_type=bt, bc_face_map=bmap, bc_data=bcd)
    pos = sp_coords(m.xy, m.tris, 3)
    shp = pos.shape[:2]
    U = cases.perturb(cases.prim_to_cons(np.full(shp, 1.0), np.full(shp, 0.3), np.full(shp, -0.1),
                                         np.full(shp, 2e3), 1.4), 0.05, 4).reshape(-1)
    s.set_state(U)
    w = np.random.default_rng(5).standard_normal(U.shape)
    v = direction(U, 6)
    assert dot_check(float(w @ fd(o.residual, U, v)), float(s.residual_vjp(w) @ v), f"BC {bct}") < TOL


@pytest.mark.parametrize("precision", ["fp64", "fp32bulk"])
@pytest.mark.parametrize("p", [2, 4])
def test_tri_step_vjp_and_dt(precision, p):
    """One SSP-RK3 step: the state's and dt's cotangents."""
    say(f"case: triangle step VJP, periodic 6x6 P{p} roe {precision}")
    s, o, m = _periodic(6, p, "roe", precision=precision, seed=p)
    U = _state(m, p, 50 + p)
    dt = 2e-3
    ubar = np.random.default_rng(51).standard_normal(U.shape)
    v = direction(U, 52)
    lhs = float(ubar @ fd(lambda X: o.advance(X, dt), U, v))
    s.set_state(U)
    s.adjoint_set(ubar)
    dtb = s.rk3_step_vjp(dt)
    tol = TOL if precision == "fp64" else 2e-4
    assert dot_check(lhs, float(s.adjoint_get() @ v), "step") < tol
    h = 1e-7
    dfd = (float(ubar @ o.advance(U, dt + h)) - float(ubar @ o.advance(U, dt - h))) / (2 * h)
    rd = abs(dtb - dfd) / abs(dfd)
    say(f"  dJ/ddt: adjoint {dtb:.12e}  FD {dfd:.12e}  rel {rd:.2e}")
    assert rd < tol


def test_tri_multi_step_vjp():
    """Four steps backwards, the states of the trajectory set before each
    step's VJP: the pending cross-tile cotangents pass from one step's
    stage 1 to the previous step's stage 3."""
    say("case: triangle 4-step VJP, periodic 8x8 P3 hllc/hu")
    s, o, m = _periodic(8, 3, "hllc", corr="hu", seed=8)
    U = _state(m, 3, 60)
    dt, steps = 1e-3, 4
    traj = [U]
    for _ in range(steps):
        traj.append(o.advance(traj[-1], dt))
    ubar = np.random.default_rng(61).standard_normal(U.shape)
    v = direction(U, 62)
    lhs = float(ubar @ fd(lambda X: _advance(o, X, dt, steps), U, v))
    s.adjoint_set(ubar)
    for n in range(steps - 1, -1, -1):
        s.set_state(traj[n])
        s.rk3_step_vjp(dt, want_dt=False)
    assert dot_check(lhs, float(s.adjoint_get() @ v), "4 steps") < TOL


def test_tri_vjp_numbering_independent():
    """A Hilbert-ordered handle and one in the caller's order (different
    tiles, so different cross-tile edges) give the same cotangent to round-off."""
    say("case: triangle residual VJP, Hilbert vs caller order, periodic 12x12 P2 roe")
    a, _, m = _periodic(12, 2, "roe", order="hilbert", seed=12)
    b, _, _ = _periodic(12, 2, "roe", order="none", seed=12)
    U = _state(m, 2, 70)
    w = np.random.default_rng(71).standard_normal(U.shape)
    out = []
    for s in (a, b):
        s.set_state(U)
        out.append(s.residual_vjp(w))
    r = np.linalg.norm(out[0] - out[1]) / np.linalg.norm(out[0])
    say(f"  rel L2 {r:.3e}")
    assert r < 1e-12
