"""Triangles (zj_tri_*) on an MI355X against the triangle oracle
(oracle/tri_oracle.py, pinned on the CPU in test_tri_oracle.py).

Bar: FP64 residual within 1e-12 relative L2 of the oracle for every order and
Riemann solver, periodic and open meshes; 20 steps within 1e-10 (the textbook
scheme is stable, so round-off does not grow); FP32-bulk residual within 1e-5.
The element numbering is internal: a Hilbert-ordered handle and one in the
caller's order give the same bits.
"""
import math

import numpy as np
import pytest

from oracle import pyoracle as O
from oracle import tri_oracle as T
from tests.test_gpu_parity import rel, say
from zhijian_amd import cases
from zhijian_amd.tri import TriSolver, sp_coords

pytestmark = pytest.mark.gpu

TOL = 1e-12


def _periodic(n, p, jitter, seed, riemann, precision="fp64", order="hilbert", L=1.0, x0=0.0, corr="dg"):
    m = cases.split_tri_mesh(n, jitter=jitter, seed=seed, L=L, x0=x0)
    s = TriSolver(m.xy, m.tris, p, periodic=(L, L), riemann=riemann, precision=precision, order=order, corr=corr)
    o = T.TriTextbook(p, m.xy, m.tris, s.faces, s.normals, riemann=riemann, corr=corr)
    return s, o, m


def _state(m, p, seed, mach=0.3):
    pos = sp_coords(m.xy, m.tris, p)
    return cases.perturb(cases.tgv_ic(pos * 2 * math.pi / m.L, mach=mach), 0.05, seed).reshape(-1)


@pytest.mark.parametrize("p", [1, 2, 3, 4, 5])
@pytest.mark.parametrize("riemann", ["rusanov", "roe", "hllc"])
def test_tri_residual(p, riemann):
    say(f"case: triangle residual, periodic 6x6 jitter 0.2 P{p} {riemann}")
    s, o, m = _periodic(6, p, 0.2, p, riemann)
    U = _state(m, p, 10 + p)
    s.set_state(U)
    r = rel(s.residual(), o.residual(U))
    say(f"  rel L2 {r:.3e}")
    assert r < TOL


def test_tri_boundaries():
    say("case: triangle residual, open 5x5 P3 hllc with walls, far field, outflow, inflow")
    m = cases.split_tri_mesh(5, jitter=0.15, seed=4)
    from zhijian_amd.tri import tri_faces
    faces, nrm = tri_faces(m.xy, m.tris)
    bt = np.zeros(faces.shape[0], np.int32)
    cycle = [1, 2, 3, 4, 6, 5, 7, 0]
    for k, f in enumerate(np.nonzero(faces[:, 2] < 0)[0]):
        bt[f] = cycle[k % len(cycle)]
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
    assert rel(s.residual(), o.residual(U)) < TOL


def test_tri_order_independent():
    say("case: triangle Hilbert-ordered vs caller-ordered handle, periodic 12x12 P3 roe, 3 steps")
    a, _, m = _periodic(12, 3, 0.2, 1, "roe", order="hilbert")
    b, _, _ = _periodic(12, 3, 0.2, 1, "roe", order="none")
    U = _state(m, 3, 5)
    for s in (a, b):
        s.set_state(U)
        for _ in range(3):
            s.rk3_step(1e-4)
    assert np.array_equal(a.get_state(), b.get_state())
    m = cases.split_tri_mesh(40)
    ce = [TriSolver(m.xy, m.tris, 3, periodic=(1.0, 1.0), order=o).info()["cross_tile_edges"]
          for o in ("hilbert", "none")]
    say(f"  cross-tile edge slots on 40x40: Hilbert {ce[0]}, caller order {ce[1]}")
    assert ce[0] < ce[1]


def test_tri_20_steps_vortex():
    say("case: triangle 20 steps, vortex 16x16 on [-5,5]^2 P3 roe")
    s, o, m = _periodic(16, 3, 0.0, 0, "roe", L=10.0, x0=-5.0)
    pos = sp_coords(m.xy, m.tris, 3)
    U = cases.vortex_ic(pos).reshape(-1)
    s.set_state(U)
    dt = 0.5 * s.dt_cfl(1.0)
    V = U.copy()
    for _ in range(20):
        s.rk3_step(dt)
        V = o.advance(V, dt)
    r = rel(s.get_state(), V)
    say(f"  dt {dt:.3e}  rel L2 after 20 steps {r:.3e}")
    assert r < 1e-10
    assert s.residual_l2() > 0


@pytest.mark.parametrize("p", [2, 4])
def test_tri_fp32bulk_residual(p):
    say(f"case: triangle FP32-bulk residual vs oracle (FP64), periodic 6x6 P{p} hllc")
    s, o, m = _periodic(6, p, 0.2, 3, "hllc", precision="fp32bulk")
    U = _state(m, p, 21)
    s.set_state(U)
    r = rel(s.residual(), o.residual(U))
    say(f"  rel L2 {r:.3e}")
    assert r < 1e-5


def test_tri_fp32bulk_steps():
    say("case: triangle FP32-bulk vs FP64 after 50 steps, periodic 32x32 P3 hllc, Fourier free stream")
    m = cases.split_tri_mesh(32, jitter=0.2, seed=7)
    a = TriSolver(m.xy, m.tris, 3, periodic=(1.0, 1.0), riemann="hllc")
    b = TriSolver(m.xy, m.tris, 3, periodic=(1.0, 1.0), riemann="hllc", precision="fp32bulk")
    U = cases.fourier_freestream_ic(sp_coords(m.xy, m.tris, 3), seed=1).reshape(-1)
    a.set_state(U)
    b.set_state(U)
    dt = a.dt_cfl(0.5)
    for _ in range(50):
        a.rk3_step(dt)
        b.rk3_step(dt)
    r = rel(b.get_state(), a.get_state())
    say(f"  FP32-bulk vs FP64 after 50 steps: rel L2 {r:.3e}")
    assert r < 1e-5


@pytest.mark.parametrize("n", [1, 2])
def test_tri_tiny_periodic_meshes(n):
    """1x1 and 2x2 periodic meshes: two triangles share several faces (the
    in-tile face points are solved once, by each face's left element, for
    both sides)."""
    say(f"case: triangle residual, periodic {n}x{n} P2 hllc and roe")
    for riemann in ("hllc", "roe"):
        s, o, m = _periodic(n, 2, 0.0, 0, riemann)
        U = _state(m, 2, 3 + n)
        s.set_state(U)
        r = rel(s.residual(), o.residual(U))
        say(f"  {riemann}: rel L2 {r:.3e}")
        assert r < TOL


@pytest.mark.parametrize("p", [1, 2, 3, 4, 5])
@pytest.mark.parametrize("corr", ["sd", "hu", "ga"])
def test_tri_residual_correction_families(p, corr):
    """Energy-stable correction fields (the lifting the handle reads is chosen
    per handle): residual against the oracle's (M + K)^-1 M filtered lifting."""
    say(f"case: triangle residual, periodic 6x6 jitter 0.2 P{p} hllc corr {corr}")
    s, o, m = _periodic(6, p, 0.2, p, "hllc", corr=corr)
    U = _state(m, p, 30 + p)
    s.set_state(U)
    r = rel(s.residual(), o.residual(U))
    say(f"  rel L2 {r:.3e}")
    assert r < TOL


def test_tri_ga_steps_and_fp32bulk():
    """The C3 configuration's correction (GA): 20 steps of the vortex against
    the oracle, FP32-bulk against FP64 after 50 steps, and handles of two
    families side by side keep their own lifting."""
    say("case: triangle GA, vortex 16x16 P3 roe 20 steps; FP32-bulk vs FP64 50 steps; DG and GA handles")
    s, o, m = _periodic(16, 3, 0.0, 0, "roe", L=10.0, x0=-5.0, corr="ga")
    d, _, _ = _periodic(16, 3, 0.0, 0, "roe", L=10.0, x0=-5.0, corr="dg")
    pos = sp_coords(m.xy, m.tris, 3)
    U = cases.vortex_ic(pos).reshape(-1)
    s.set_state(U)
    d.set_state(U)
    dt = 0.5 * s.dt_cfl(1.0)
    V = U.copy()
    for _ in range(20):
        s.rk3_step(dt)
        d.rk3_step(dt)
        V = o.advance(V, dt)
    r = rel(s.get_state(), V)
    say(f"  GA rel L2 after 20 steps {r:.3e}")
    assert r < 1e-10
    assert rel(d.get_state(), V) > 1e-8   # the DG handle is a different scheme
    m2 = cases.split_tri_mesh(24, jitter=0.2, seed=7)
    a = TriSolver(m2.xy, m2.tris, 3, periodic=(1.0, 1.0), riemann="hllc", corr="ga")
    b = TriSolver(m2.xy, m2.tris, 3, periodic=(1.0, 1.0), riemann="hllc", corr="ga", precision="fp32bulk")
    W = cases.fourier_freestream_ic(sp_coords(m2.xy, m2.tris, 3), seed=1).reshape(-1)
    a.set_state(W)
    b.set_state(W)
    dt2 = a.dt_cfl(0.5)
    for _ in range(50):
        a.rk3_step(dt2)
        b.rk3_step(dt2)
    r2 = rel(b.get_state(), a.get_state())
    say(f"  GA FP32-bulk vs FP64 after 50 steps: rel L2 {r2:.3e}")
    assert r2 < 1e-5


def test_tri_c3_full_size_properties():
    """The triangle configuration at full size (500x500 split quads = 500,000
    elements, P3, HLLC, GA), through properties that do not depend on size:
    a uniform flow has zero residual; the residual of the Fourier state is
    discretely conservative (sum over elements of J w^T dU/dt, w the nodal
    quadrature weights, vanishes to round-off on the periodic mesh); one step
    keeps the totals, and the FP32-bulk step agrees with the FP64 one.  (The
    FP32 residual itself differs from FP64 by ~5e-4 here: at h = 0.002 it is
    a small difference of fluxes dominated by the mean pressure, which FP32
    rounds at 1e-7 of the flux; the states that it updates differ by ~1e-8.)"""
    say("case: triangles C3 full size 500x500 P3 hllc GA: freestream, conservation, FP32-bulk, one step")
    m = cases.split_tri_mesh(500, jitter=0.2, seed=2024)
    a = TriSolver(m.xy, m.tris, 3, periodic=(1.0, 1.0), riemann="hllc", corr="ga")
    b = TriSolver(m.xy, m.tris, 3, periodic=(1.0, 1.0), riemann="hllc", corr="ga", precision="fp32bulk")
    Np = 10
    ops = T.operators(3)
    w = np.ones(Np) @ np.linalg.inv(ops["V"] @ ops["V"].T)          # nodal quadrature weights
    P = m.xy[m.tris]
    J = 0.25 * ((P[:, 1, 0] - P[:, 0, 0]) * (P[:, 2, 1] - P[:, 0, 1])
                - (P[:, 2, 0] - P[:, 0, 0]) * (P[:, 1, 1] - P[:, 0, 1]))   # area / 2
    free = np.tile(cases.prim_to_cons(np.array(1.0), np.array(0.8), np.array(0.3), np.array(1 / 1.4 / 0.25), 1.4),
                   (m.n_elem * Np, 1)).reshape(-1)
    a.set_state(free)
    assert np.abs(a.residual()).max() < 1e-9   # round-off of O(1) fluxes over h = 0.002; a defect is O(1/h)
    U = cases.fourier_freestream_ic(sp_coords(m.xy, m.tris, 3), seed=2024).reshape(-1)
    a.set_state(U)
    b.set_state(U)
    R = a.residual().reshape(m.n_elem, Np, 4)
    tot = np.einsum("e,i,eiv->v", J, w, R)
    scale = np.einsum("e,i,eiv->v", np.abs(J), np.abs(w), np.abs(R))
    say(f"  conservation |sum J w R| / sum|.|: {np.abs(tot / scale).max():.2e}")
    assert np.all(np.abs(tot) <= 1e-11 * scale)
    before = np.einsum("e,i,eiv->v", J, w, U.reshape(m.n_elem, Np, 4))
    dt = a.dt_cfl(0.5)
    a.rk3_step(dt)
    b.rk3_step(dt)
    Ua = a.get_state()
    after = np.einsum("e,i,eiv->v", J, w, Ua.reshape(m.n_elem, Np, 4))
    assert np.all(np.abs(after - before) <= 1e-12 * np.abs(before).max())
    r = rel(b.get_state(), Ua)
    say(f"  FP32-bulk vs FP64 after one step: rel L2 {r:.3e}")
    assert r < 1e-7
    a.close()
    b.close()
