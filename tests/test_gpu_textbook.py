"""Textbook mode (ZJ_COMPAT_TEXTBOOK) on an MI355X against the textbook oracle
(oracle/zj_textbook.c, pinned on the CPU in test_oracle_textbook.py).

Bar: FP64 residual within 1e-12 relative L2 of the oracle for every Riemann
solver and correction function, on affine (per-element metric) and jittered
(per-point metric) meshes and with boundaries; 100 steps within 1e-10 (the
textbook scheme is stable, so round-off does not grow); FP32-bulk after 100
steps within 1e-5 of FP64.
"""
import numpy as np
import pytest

from oracle import pyoracle as O
from tests import fr_case
from tests.test_gpu_parity import rel, say, tgv_state, vortex_state
from zhijian_amd import cases, solver

pytestmark = pytest.mark.gpu

TOL = 1e-12


def _pair(mesh, p, riemann, corr, periodic=True, precision="fp64", bc_types=None, fs=None):
    per = (mesh.Lx, mesh.Ly) if periodic else None
    t = solver.build_tables(mesh.xy, mesh.elem_nodes, p, periodic=per, bc_type_per_face=bc_types, fs=fs)
    s = solver.Solver(t, precision=precision, compat="textbook", riemann=riemann, corr=corr)
    oc = fr_case.build(mesh.xy, mesh.elem_nodes, p, periodic=per, bc_type_per_face=bc_types, fs=fs)
    return s, oc.textbook(riemann, corr), t


@pytest.mark.parametrize("riemann", ["rusanov", "roe", "hllc"])
@pytest.mark.parametrize("corr", ["dg", "sd", "hu", "ga"])
def test_textbook_residual_solvers_and_corrections(riemann, corr):
    say(f"case: textbook residual, vortex 6x6 P3, {riemann}/{corr}")
    s, tb, t = _pair(cases.vortex_mesh(6), 3, riemann, corr)
    U = vortex_state(t, amp=0.05, seed=3)
    s.set_state(U)
    assert rel(s.residual(), tb.residual(U)) < TOL


@pytest.mark.parametrize("p", [2, 4])
def test_textbook_fp32bulk_residual(p):
    say(f"case: textbook FP32-bulk residual vs oracle (FP64), vortex 6x6 P{p}, roe/hu")
    s, tb, t = _pair(cases.vortex_mesh(6), p, "roe", "hu", precision="fp32bulk")
    U = vortex_state(t, amp=0.05, seed=9)
    s.set_state(U)
    r = rel(s.residual(), tb.residual(U))
    say(f"  rel L2 {r:.3e}")
    assert r < 1e-5


@pytest.mark.parametrize("p", [1, 2, 3, 4, 5])
def test_textbook_residual_orders_jittered(p):
    for jit in (0.0, 0.12):
        say(f"case: textbook residual, periodic 5x4 P{p} jitter {jit}, roe/hu")
        mesh = cases.uniform_quad_mesh(5, 4, 0.0, 1.0, 0.0, 1.0, jitter=jit, seed=p)
        s, tb, t = _pair(mesh, p, "roe", "hu")
        assert s.info()["affine"] == (jit == 0.0)
        U = tgv_state(t, amp=0.05, seed=p, mach=0.3)
        s.set_state(U)
        assert rel(s.residual(), tb.residual(U)) < TOL


def test_textbook_boundaries():
    say("case: textbook residual, open 5x4 P3 with every BC type, hllc/dg")
    mesh = cases.uniform_quad_mesh(5, 4, 0.0, 1.0, 0.0, 1.0, jitter=0.15, seed=3)
    faces = solver._lib.build_faces(mesh.elem_nodes)
    types = np.zeros(faces.shape[0], np.int32)
    bnd = np.nonzero(faces[:, 2] < 0)[0]
    cycle = [1, 2, 3, 4, 5, 6, 7, 0]
    for n, f in enumerate(bnd):
        types[f] = cycle[n % len(cycle)]
    fs = (1.1, 0.3, 0.1, 1.5e3)
    s, tb, t = _pair(mesh, 3, "hllc", "dg", periodic=False, bc_types=types, fs=fs)
    shp = t.pos.shape[:2]
    U = cases.perturb(cases.prim_to_cons(np.full(shp, 1.0), np.full(shp, 0.3), np.full(shp, -0.1),
                                         np.full(shp, 2e3), 1.4), 0.05, 4).reshape(-1)
    s.set_state(U)
    assert rel(s.residual(), tb.residual(U)) < TOL


def test_textbook_100_steps_vortex():
    """The textbook scheme is stable: 100 steps stay within round-off of the
    oracle (the reference arithmetic diverges from any other evaluation
    order within ~30 steps)."""
    say("case: textbook 100 steps, vortex 10x10 P3 roe/dg, no limiter")
    s, tb, t = _pair(cases.vortex_mesh(10), 3, "roe", "dg")
    s2 = solver.Solver(t, compat="textbook", riemann="roe", corr="dg", positivity=False)
    U = vortex_state(t, amp=0.0, seed=0)
    s2.set_state(U)
    dt = 1e-2
    V = U.copy()
    for _ in range(100):
        s2.rk3_step(dt)
        V = tb.advance(V, dt, False)
    assert rel(s2.get_state(), V) < 1e-10


def test_textbook_fp32_after_100_steps():
    say("case: textbook FP32-bulk vs FP64 after 100 steps, TGV 32x32 P4 roe/hu, no limiter")
    mesh = cases.tgv_mesh(32)
    t = solver.build_tables(mesh.xy, mesh.elem_nodes, 4, periodic=(mesh.Lx, mesh.Ly))
    a = solver.Solver(t, compat="textbook", riemann="roe", corr="hu", positivity=False)
    b = solver.Solver(t, compat="textbook", riemann="roe", corr="hu", positivity=False, precision="fp32bulk")
    U = cases.tgv_ic(t.pos, mach=0.1).reshape(-1)
    a.set_state(U)
    b.set_state(U)
    dt = a.dt_cfl(0.5)
    for _ in range(100):
        a.rk3_step(dt)
        b.rk3_step(dt)
    r = rel(b.get_state(), a.get_state())
    say(f"  FP32-bulk vs FP64 after 100 steps: rel L2 {r:.3e}")
    assert r < 1e-5
