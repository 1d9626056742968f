"""Restart files from and into device state on an MI355X (zj_restart_save /
zj_restart_load, zj_tri_restart_save / zj_tri_restart_load).

Bar: N steps, save, a new handle loaded from the file, M more steps equals
N + M steps straight through, bit for bit (the file holds the FP64 state
exactly and every other buffer is rebuilt from it); mismatched meshes, orders
and gamma are refused.
"""
import math

import numpy as np
import pytest

from tests.test_gpu_parity import say
from zhijian_amd import cases, io, solver
from zhijian_amd._lib import ZJError
from zhijian_amd.tri import TriSolver, sp_coords

pytestmark = pytest.mark.gpu


def _vortex(n, p, precision, compat="reference"):
    mesh = cases.vortex_mesh(n)
    t = solver.build_tables(mesh.xy, mesh.elem_nodes, p, periodic=(mesh.Lx, mesh.Ly))
    kw = dict(precision=precision)
    if compat == "textbook":
        kw.update(compat="textbook", riemann="roe", corr="hu", positivity=False)
    return t, (lambda: solver.Solver(t, **kw))


@pytest.mark.parametrize("precision,compat", [("fp64", "reference"), ("fp32bulk", "reference"),
                                              ("fp32bulk", "textbook")])
def test_quad_resume_bit_exact(tmp_path, precision, compat):
    say(f"case: quad restart resume, vortex 12x12 P3 {precision} {compat}, 5 + 5 steps")
    t, make = _vortex(12, 3, precision, compat)
    U0 = cases.vortex_ic(t.pos).reshape(-1)
    a = make()
    a.set_state(U0)
    for _ in range(10):
        a.rk3_step(1e-3)
    b = make()
    b.set_state(U0)
    for _ in range(5):
        b.rk3_step(1e-3)
    path = tmp_path / "q.h5"
    b.save_restart(path, time=5e-3, iteration=5)
    r = io.read_restart(path)
    assert r["solution"].tobytes() == b.get_state().tobytes() and r["poly_order"] == 3
    c = make()
    assert c.load_restart(path) == (5e-3, 5)
    for _ in range(5):
        c.rk3_step(1e-3)
    assert c.get_state().tobytes() == a.get_state().tobytes()


def test_tri_resume_bit_exact(tmp_path):
    say("case: triangle restart resume, periodic 16x16 P3 hllc fp64 and fp32bulk, 4 + 4 steps")
    m = cases.split_tri_mesh(16, jitter=0.2, seed=3)
    U0 = cases.fourier_freestream_ic(sp_coords(m.xy, m.tris, 3), seed=2).reshape(-1)
    for prec in ("fp64", "fp32bulk"):
        def make():
            return TriSolver(m.xy, m.tris, 3, periodic=(1.0, 1.0), riemann="hllc", precision=prec)
        a, b = make(), make()
        for s in (a, b):
            s.set_state(U0)
        dt = 0.5 * a.dt_cfl(0.5)
        for _ in range(8):
            a.rk3_step(dt)
        for _ in range(4):
            b.rk3_step(dt)
        path = tmp_path / f"t_{prec}.h5"
        b.save_restart(path, time=4 * dt, iteration=4)
        c = make()
        t, it = c.load_restart(path)
        assert it == 4 and t == 4 * dt
        for _ in range(4):
            c.rk3_step(dt)
        assert c.get_state().tobytes() == a.get_state().tobytes(), prec


def test_load_refuses_mismatches(tmp_path):
    say("case: restart load refuses other meshes, orders and gamma")
    t3, make3 = _vortex(6, 3, "fp64")
    a = make3()
    a.set_state(cases.vortex_ic(t3.pos).reshape(-1))
    path = tmp_path / "a.h5"
    a.save_restart(path, 0.0, 0)
    _, make2 = _vortex(6, 2, "fp64")
    with pytest.raises(ZJError, match="do not match|poly_order"):
        make2().load_restart(path)
    _, make7 = _vortex(7, 3, "fp64")
    with pytest.raises(ZJError, match="do not match"):
        make7().load_restart(path)
    r = io.read_restart(path)
    io.write_restart(tmp_path / "g.h5", r["elem_n_sp"], r["solution"], 0.0, 0, 3, 1.3)
    with pytest.raises(ZJError, match="gamma"):
        make3().load_restart(tmp_path / "g.h5")
    m = cases.split_tri_mesh(4)
    s = TriSolver(m.xy, m.tris, 2, periodic=(1.0, 1.0))
    with pytest.raises(ZJError):
        s.load_restart(path)
    assert math.isfinite(float(np.sum(a.get_state())))
