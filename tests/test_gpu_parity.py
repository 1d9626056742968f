"""HIP path vs the CPU oracle (parity proper).  Runs on an MI355X.

Bar: FP64 residual and RK state within 1e-12 relative L2 of the oracle (the
oracle restates the reference's executed arithmetic; the only differences are
FMA contraction and summation order), connectivity bit-exact (checked on the
CPU in test_host_tables.py).  FP32-bulk is reported against FP64.
"""
import math
import sys

import numpy as np
import pytest

from oracle import pyoracle as O
from tests import fr_case
from tests.conftest import mesh_from_golden
from zhijian_amd import cases, solver

pytestmark = pytest.mark.gpu

TOL = 1e-12


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def make(mesh, p, periodic=True, precision="fp64", bc_types=None, fs=None, gamma=1.4):
    per = (mesh.Lx, mesh.Ly) if periodic else None
    t = solver.build_tables(mesh.xy, mesh.elem_nodes, p, periodic=per, bc_type_per_face=bc_types, fs=fs)
    s = solver.Solver(t, gamma=gamma, precision=precision)
    oc = fr_case.build(mesh.xy, mesh.elem_nodes, p, periodic=per, bc_type_per_face=bc_types, fs=fs,
                       gamma=gamma)
    return s, oc, t


def vortex_state(t, amp=0.05, seed=1):
    U = cases.vortex_ic(t.pos)
    return cases.perturb(U, amp, seed).reshape(-1)


def tgv_state(t, amp=0.05, seed=2, mach=0.1):
    return cases.perturb(cases.tgv_ic(t.pos, mach=mach), amp, seed).reshape(-1)


def test_vortex_c1_residual():
    say("case: vortex 10x10 P3 residual")
    s, oc, t = make(cases.vortex_mesh(10), 3)
    U = vortex_state(t)
    s.set_state(U)
    Rg = s.residual()
    Ro = oc.solver().residual(U)
    assert np.linalg.norm(Ro) > 1.0
    assert rel(Rg, Ro) < TOL
    assert np.array_equal(s.get_state(), U)


@pytest.mark.parametrize("p", [1, 2, 3, 4, 5])
def test_residual_all_orders_jittered(p):
    say(f"case: jittered 6x5 P{p} residual")
    mesh = cases.uniform_quad_mesh(6, 5, 0.0, 2 * math.pi, 0.0, 2 * math.pi, jitter=0.2, seed=p)
    s, oc, t = make(mesh, p)
    U = tgv_state(t, amp=0.1, seed=p, mach=0.3)
    s.set_state(U)
    assert rel(s.residual(), oc.solver().residual(U)) < TOL


def test_one_element_periodic_self_neighbour():
    say("case: 1x1 periodic P2 (element is its own neighbour)")
    mesh = cases.uniform_quad_mesh(1, 1, 0.0, 1.0, 0.0, 1.0)
    s, oc, t = make(mesh, 2)
    U = tgv_state(t, amp=0.2, seed=9, mach=0.5)
    s.set_state(U)
    assert rel(s.residual(), oc.solver().residual(U)) < TOL


def test_boundary_conditions_residual():
    say("case: open 5x4 P3 with every BC type")
    mesh = cases.uniform_quad_mesh(5, 4, 0.0, 1.0, 0.0, 1.0, jitter=0.15, seed=3)
    faces = solver._lib.build_faces(mesh.elem_nodes)
    types = np.zeros(faces.shape[0], np.int32)
    bnd = np.nonzero(faces[:, 2] < 0)[0]
    cycle = [1, 2, 3, 4, 5, 6, 7, 0]
    for n, f in enumerate(bnd):
        types[f] = cycle[n % len(cycle)]
    fs = (1.1, 0.3, 0.1, 1.5e3)  # far field inside the reference's ghost clamps
    s, oc, t = make(mesh, 3, periodic=False, bc_types=types, fs=fs)
    U = cases.perturb(cases.prim_to_cons(np.full(t.pos.shape[:2], 1.0), np.full(t.pos.shape[:2], 0.3),
                                         np.full(t.pos.shape[:2], -0.1), np.full(t.pos.shape[:2], 2e3), 1.4),
                      0.05, 4).reshape(-1)
    s.set_state(U)
    assert rel(s.residual(), oc.solver().residual(U)) < TOL


def test_reference_mesh_farfield_freestream(golden_meshes):
    say("case: uniform_quad_10x10.msh far field, freestream and perturbed")
    m = mesh_from_golden(golden_meshes, "uniform_quad_10x10")
    t = solver.build_tables(m["xy"], m["elem_nodes"], 3, periodic=None)
    s = solver.Solver(t)
    oc = fr_case.build(m["xy"], m["elem_nodes"], 3)
    rho, u, v, p = solver.freestream()
    U = np.tile(cases.prim_to_cons(np.array(rho), np.array(u), np.array(v), np.array(p), 1.4),
                (t.E * t.S, 1)).reshape(-1)
    s.set_state(U)
    R = s.residual()
    assert np.abs(R).max() < 1e-8 * p / 0.1
    Up = cases.perturb(U.reshape(t.E, t.S, 4), 0.02, 5).reshape(-1)
    s.set_state(Up)
    assert rel(s.residual(), oc.solver().residual(Up)) < TOL


def test_vortex_c1_steps_parity():
    """C1 (vortex 10x10 P3, dt=1e-3, limiter on).  The reference arithmetic is
    linearly unstable (test_reference_scheme_is_unstable): the state reaches the
    limiter clamps near step 28 and two oracle runs 1e-15 apart differ at O(1)
    by step 50.  Step parity is therefore checked over the first 20 steps; the
    full 100 steps are checked for what the limiter guarantees."""
    say("case: vortex 10x10 P3, 20 + 80 SSP-RK3 steps, dt=1e-3, limiter on")
    s, oc, t = make(cases.vortex_mesh(10), 3)
    U0 = cases.vortex_ic(t.pos).reshape(-1)
    s.set_state(U0)
    o = oc.solver()
    U = U0.copy()
    for _ in range(20):
        s.rk3_step(1e-3)
        U = o.advance(U, 1e-3, True)
    assert rel(s.get_state(), U) < TOL
    l2g = s.residual_l2()
    assert abs(l2g - o.residual_l2()) <= 1e-12 * o.residual_l2()
    for _ in range(80):
        s.rk3_step(1e-3)
    q = s.get_state().reshape(-1, 4)
    assert np.all(np.isfinite(q))
    assert q[:, 0].min() >= 0.01 and q[:, 0].max() <= 100.0
    u = q[:, 1] / q[:, 0]; v = q[:, 2] / q[:, 0]
    assert np.sqrt(u * u + v * v).max() <= 2000.0 * (1 + 1e-12)


@pytest.mark.parametrize("positivity,steps", [(True, 6), (False, 4)])
def test_tgv_steps_parity(positivity, steps):
    say(f"case: TGV 32x32 P4, {steps} steps, limiter={positivity}")
    mesh = cases.tgv_mesh(32)
    per = (mesh.Lx, mesh.Ly)
    t = solver.build_tables(mesh.xy, mesh.elem_nodes, 4, periodic=per)
    s = solver.Solver(t, positivity=positivity)
    oc = fr_case.build(mesh.xy, mesh.elem_nodes, 4, periodic=per)
    U0 = tgv_state(t, amp=0.01, seed=3)
    s.set_state(U0)
    o = oc.solver()
    dt = o.dt_cfl(U0, 0.5)
    assert abs(s.dt_cfl(0.5) - dt) <= 1e-15 * dt
    U = U0.copy()
    for _ in range(steps):
        s.rk3_step(dt)
        U = o.advance(U, dt, positivity)
    assert rel(s.get_state(), U) < TOL


def test_tgv_full_size_residual_and_freestream():
    say("case: TGV 256x256 P4 full size: one residual vs oracle, freestream")
    mesh = cases.tgv_mesh(256)
    t = solver.build_tables(mesh.xy, mesh.elem_nodes, 4, periodic=(mesh.Lx, mesh.Ly))
    s = solver.Solver(t)
    oc = fr_case.build(mesh.xy, mesh.elem_nodes, 4, periodic=(mesh.Lx, mesh.Ly))
    U = tgv_state(t, amp=0.01, seed=7)
    s.set_state(U)
    O.set_threads(8)
    try:
        o = oc.solver()
        Ro = o.residual(U)
        assert rel(s.residual(), Ro) < TOL
        assert abs(s.residual_l2() - o.residual_l2()) <= 1e-12 * o.residual_l2()
    finally:
        O.set_threads(1)
    Uf = np.tile(cases.prim_to_cons(np.array(1.0), np.array(0.3), np.array(0.2), np.array(71.4), 1.4),
                 (t.E * t.S, 1)).reshape(-1)
    s.set_state(Uf)
    assert np.abs(s.residual()).max() < 1e-9


def test_graph_replay_equals_stage_launches():
    say("case: hipGraph step vs three stage launches, repeated with host syncs")
    mesh = cases.tgv_mesh(16)
    t = solver.build_tables(mesh.xy, mesh.elem_nodes, 3, periodic=(mesh.Lx, mesh.Ly))
    a = solver.Solver(t)
    b = solver.Solver(t)
    U0 = tgv_state(t, amp=0.02, seed=5)
    a.set_state(U0); b.set_state(U0)
    for k in range(4):
        a.rk3_step(2e-3)
        for st in (1, 2, 3):
            b.rk3_stage(st, 2e-3)
        ga, gb = a.get_state(), b.get_state()
        assert np.array_equal(ga, gb), k


@pytest.mark.parametrize("precision", ["fp64", "fp32bulk"])
def test_launch_shape_does_not_change_results(precision):
    """Tile size (max vs half, also changed mid-run), and workgroup-to-tile
    order (per-XCD runs vs launch order): traces of edges inside a tile are
    exchanged in LDS, the rest through HBM, with identical arithmetic."""
    say(f"case: tiles max vs half, XCD order vs launch order, {precision}")
    mesh = cases.uniform_quad_mesh(9, 7, 0.0, 1.0, 0.0, 1.0, jitter=0.1, seed=8)
    t = solver.build_tables(mesh.xy, mesh.elem_nodes, 4, periodic=(1.0, 1.0))
    a = solver.Solver(t, precision=precision)
    b = solver.Solver(t, precision=precision)
    c = solver.Solver(t, precision=precision)
    half = b.info()["elems_per_block"] // 2
    b.set_tile(half)
    c.set_xcd_groups(1)
    U0 = tgv_state(t, amp=0.02, seed=6)
    a.set_state(U0); b.set_state(U0); c.set_state(U0)
    Ra = a.residual()
    assert np.array_equal(Ra, b.residual()) and np.array_equal(Ra, c.residual())
    assert a.residual_l2() == c.residual_l2()
    for k in range(4):
        a.rk3_step(1e-3); b.rk3_step(1e-3); c.rk3_step(1e-3)
        if k == 1:  # retile with a state in place: the traces are rebuilt
            b.set_tile(2 * half)
            c.set_tile(half)
    assert np.array_equal(a.get_state(), b.get_state())
    assert np.array_equal(a.get_state(), c.get_state())


def test_nonfinite_detection_and_scrub():
    say("case: NaN in the state: counted; the Rusanov scrub zeroes non-finite fluxes")
    mesh = cases.vortex_mesh(4)
    s, oc, t = make(mesh, 2)
    U = vortex_state(t)
    s.set_state(U)
    assert s.count_nonfinite() == 0
    U[17] = np.nan
    s.set_state(U)
    assert s.count_nonfinite() == 1


@pytest.mark.parametrize("p", [3, 4])
def test_fp32bulk_residual_vs_fp64(p):
    """FP32 bulk in the reference arithmetic: residual error and a short run
    (the reference scheme amplifies any round-off exponentially, so long-run
    FP32 error is only meaningful for a stable correction)."""
    say(f"case: FP32-bulk vs FP64 (reference arithmetic), TGV 32x32 P{p}")
    mesh = cases.tgv_mesh(32)
    t = solver.build_tables(mesh.xy, mesh.elem_nodes, p, periodic=(mesh.Lx, mesh.Ly))
    a = solver.Solver(t, precision="fp64")
    b = solver.Solver(t, precision="fp32bulk")
    U0 = cases.tgv_ic(t.pos).reshape(-1)
    a.set_state(U0); b.set_state(U0)
    dt = a.dt_cfl(0.5)
    Ra, Rb = a.residual(), b.residual()
    r_res = rel(Rb, Ra)
    for _ in range(3):
        a.rk3_step(dt); b.rk3_step(dt)
    Ua, Ub = a.get_state(), b.get_state()
    r_l2 = rel(Ub, Ua)
    say(f"  fp32bulk P{p}: residual rel L2 {r_res:.3e}; state after 3 steps rel L2 {r_l2:.3e}")
    assert r_res < 1e-2
    assert r_l2 < 1e-5


@pytest.mark.parametrize("precision", ["fp64", "fp32bulk"])
def test_affine_metric_matches_per_point(precision):
    """Per-element metric with contravariant fluxes (affine meshes) against the
    per-point metric the reference stores, every order; a jittered mesh has no
    per-element form."""
    tol = 1e-13 if precision == "fp64" else 2e-6
    for p in range(1, 6):
        say(f"case: affine vs per-point metric, vortex 6x6 P{p} {precision}")
        mesh = cases.vortex_mesh(6)
        t = solver.build_tables(mesh.xy, mesh.elem_nodes, p, periodic=(mesh.Lx, mesh.Ly))
        a = solver.Solver(t, precision=precision)
        b = solver.Solver(t, precision=precision)
        assert a.set_geometry_mode(-1) == 1
        assert b.set_geometry_mode(0) == 0
        U0 = vortex_state(t, seed=p)
        a.set_state(U0); b.set_state(U0)
        assert rel(a.residual(), b.residual()) < tol
        for _ in range(2):
            a.rk3_step(1e-3); b.rk3_step(1e-3)
        assert rel(a.get_state(), b.get_state()) < tol
    mesh = cases.uniform_quad_mesh(5, 5, 0.0, 1.0, 0.0, 1.0, jitter=0.1, seed=3)
    t = solver.build_tables(mesh.xy, mesh.elem_nodes, 3, periodic=(1.0, 1.0))
    s = solver.Solver(t)
    assert s.set_geometry_mode(-1) == 0
    with pytest.raises(RuntimeError):
        s.set_geometry_mode(1)


def test_fp32bulk_stage_launches_equal_graph():
    """FP32-bulk keeps the intermediate stages as FP32 increments over U^n; the
    stage-by-stage entry point and the graph step must agree bit for bit."""
    say("case: FP32-bulk hipGraph step vs three stage launches")
    mesh = cases.tgv_mesh(16)
    t = solver.build_tables(mesh.xy, mesh.elem_nodes, 4, periodic=(mesh.Lx, mesh.Ly))
    a = solver.Solver(t, precision="fp32bulk")
    b = solver.Solver(t, precision="fp32bulk")
    U0 = tgv_state(t, amp=0.02, seed=9)
    a.set_state(U0); b.set_state(U0)
    for k in range(3):
        a.rk3_step(2e-3)
        for st in (1, 2, 3):
            b.rk3_stage(st, 2e-3)
        assert np.array_equal(a.get_state(), b.get_state()), k


@pytest.mark.parametrize("p", [2, 4])
def test_sheared_mesh_paired_by_node_links(p):
    """A parallelogram domain made periodic by node links (a Gmsh $Periodic
    section's kind of pairing, zj_pair_periodic_nodes): the reference
    arithmetic's residual against the oracle on the same links, and
    freestream preserved; the textbook scheme's residual and 10 RK3 steps
    against its oracle (the reference arithmetic, whose correction ignores
    the metric, blows up within a few steps on elements this skewed)."""
    say(f"case: sheared 7x6 periodic by node links, P{p}, residual; textbook 10 steps")
    mesh = cases.sheared_quad_mesh(7, 6, 2 * math.pi, 2 * math.pi, 1.9, jitter=0.15, seed=p)
    links = cases.periodic_links(7, 6)
    t = solver.build_tables(mesh.xy, mesh.elem_nodes, p, periodic_nodes=links)
    assert np.all(t.faces[:, 2] >= 0)
    oc = fr_case.build(mesh.xy, mesh.elem_nodes, p, periodic_nodes=links)
    assert np.array_equal(t.faces, oc.faces.as_array())
    U = tgv_state(t, amp=0.1, seed=p, mach=0.3)
    free = np.tile(np.array([1.0, 0.3, -0.2, 2.5]), t.E * t.S)
    s = solver.Solver(t)
    s.set_state(U)
    assert rel(s.residual(), oc.solver().residual(U)) < TOL
    s.set_state(free)
    assert np.abs(s.residual()).max() < 1e-11
    tb = oc.textbook("roe", "hu")
    s2 = solver.Solver(t, compat="textbook", riemann="roe", corr="hu", positivity=False)
    s2.set_state(U)
    assert rel(s2.residual(), tb.residual(U)) < TOL
    V = U.copy()
    for _ in range(10):
        s2.rk3_step(2e-3)
        V = tb.advance(V, 2e-3, False)
    assert rel(s2.get_state(), V) < 1e-10
    s2.set_state(free)
    assert np.abs(s2.residual()).max() < 1e-11
