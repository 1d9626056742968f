"""Kernel execution timing (zj_time_stages, zj_tri_time_stages) on an MI355X.

The timed launches are ordinary stage launches with runtime-recorded events,
so the state they leave must equal that of the same number of graph-replayed
steps bit for bit, and the three stage kernels' times cannot exceed a step.
"""
import numpy as np
import pytest

from tests.test_gpu_parity import say, tgv_state
from zhijian_amd import cases, solver

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", ["fp64", "fp32bulk"])
def test_time_stages_steps_and_bounds(precision):
    say(f"case: zj_time_stages on TGV 64x64 P4, {precision}")
    mesh = cases.tgv_mesh(64)
    t = solver.build_tables(mesh.xy, mesh.elem_nodes, 4, periodic=(mesh.Lx, mesh.Ly))
    U = tgv_state(t, amp=0.02, seed=12)
    a, b = solver.Solver(t, precision=precision), solver.Solver(t, precision=precision)
    a.set_state(U)
    b.set_state(U)
    st = a.time_stages(1e-3, 4)
    for _ in range(4):
        b.rk3_step(1e-3)
    assert np.array_equal(a.get_state(), b.get_state())
    assert all(x > 0 for x in st)
    import time
    b.synchronize()
    t0 = time.perf_counter()
    for _ in range(50):
        b.rk3_step(1e-3)
    b.synchronize()
    step_ms = (time.perf_counter() - t0) * 1e3 / 50
    say(f"  stage kernels {['%.4f' % x for x in st]} ms, step {step_ms:.4f} ms (host clock)")
    assert sum(st) <= step_ms * 1.05


def test_tri_time_stages_steps():
    say("case: zj_tri_time_stages on a 24x24 split-quad mesh, P3")
    from zhijian_amd.tri import TriSolver, sp_coords
    m = cases.split_tri_mesh(24, jitter=0.2, seed=5)
    U = cases.fourier_freestream_ic(sp_coords(m.xy, m.tris, 3), seed=5).reshape(-1)
    a = TriSolver(m.xy, m.tris, 3, periodic=(1.0, 1.0), riemann="hllc", corr="ga")
    b = TriSolver(m.xy, m.tris, 3, periodic=(1.0, 1.0), riemann="hllc", corr="ga")
    a.set_state(U)
    b.set_state(U)
    st = a.time_stages(1e-5, 3)
    for _ in range(3):
        b.rk3_step(1e-5)
    assert np.array_equal(a.get_state(), b.get_state())
    assert all(x > 0 for x in st)
