"""Batched stepping (zj_batch_*) and the launch autotuner on an MI355X.

The batched launches run the same stage arithmetic as a single mesh's
kernels, so a batched step must equal each member's own step bit for bit, for
every order, both metric forms, every scheme (reference arithmetic, textbook
Euler, textbook Navier-Stokes) and every workgroup-grid policy.
"""
import numpy as np
import pytest

from tests.test_gpu_parity import say, tgv_state, vortex_state
from zhijian_amd import cases, solver

pytestmark = pytest.mark.gpu

MEMBERS = [  # (order, elements per side, case, node jitter)
    (1, 12, "tgv", 0.0), (2, 9, "vortex", 0.0), (3, 10, "tgv", 0.1), (4, 8, "vortex", 0.0),
    (5, 6, "tgv", 0.0), (4, 11, "tgv", 0.08), (3, 16, "vortex", 0.0), (2, 7, "tgv", 0.12),
]


def _member(p, n, kind, jit, precision, seed):
    if kind == "tgv":
        mesh = cases.uniform_quad_mesh(n, n, 0.0, 2 * np.pi, 0.0, 2 * np.pi, jitter=jit, seed=seed)
    else:
        mesh = cases.uniform_quad_mesh(n, n, -5.0, 5.0, -5.0, 5.0, jitter=jit, seed=seed)
    t = solver.build_tables(mesh.xy, mesh.elem_nodes, p, periodic=(mesh.Lx, mesh.Ly))
    U = tgv_state(t, amp=0.02, seed=seed, mach=0.3) if kind == "tgv" else vortex_state(t, amp=0.02, seed=seed)
    return t, U


@pytest.mark.parametrize("precision", ["fp64", "fp32bulk"])
def test_batch_equals_members(precision):
    say(f"case: batch of {len(MEMBERS)} meshes (P1-P5, affine and not) vs own steps, {precision}")
    built = [_member(*m, precision, seed=30 + k) for k, m in enumerate(MEMBERS)]
    a = [solver.Solver(t, precision=precision) for t, _ in built]
    b = [solver.Solver(t, precision=precision) for t, _ in built]
    for s1, s2, (t, U) in zip(a, b, built):
        s1.set_state(U)
        s2.set_state(U)
    batch = solver.Batch(a)
    dts = np.array([1e-3 * (1 + 0.1 * k) for k in range(len(a))])
    for step in range(3):
        if step == 2:
            batch.set_grid(2)  # persistent: two workgroups per CU looping over the items
        batch.rk3_step(dts)
        for s2, dt in zip(b, dts):
            s2.rk3_step(dt)
    batch.synchronize()
    for k, (s1, s2) in enumerate(zip(a, b)):
        assert np.array_equal(s1.get_state(), s2.get_state()), MEMBERS[k]
        assert s1.residual_l2() == s2.residual_l2(), MEMBERS[k]
    batch.close()


@pytest.mark.parametrize("precision", ["fp64", "fp32bulk"])
def test_single_launch_equals_members(precision):
    """One launch per stage over every member (orders P1-P5, both metric forms
    in the same kernel) gives each member's own step bit for bit, under both
    grid policies; the batch autotuner leaves the states as they were."""
    say(f"case: single-launch batch of {len(MEMBERS)} meshes vs own steps, {precision}")
    built = [_member(*m, precision, seed=50 + k) for k, m in enumerate(MEMBERS)]
    a = [solver.Solver(t, precision=precision) for t, _ in built]
    b = [solver.Solver(t, precision=precision) for t, _ in built]
    for s1, s2, (t, U) in zip(a, b, built):
        s1.set_state(U)
        s2.set_state(U)
    batch = solver.Batch(a)
    batch.set_launch(True)
    dts = np.array([1e-3 * (1 + 0.05 * k) for k in range(len(a))])
    for step in range(3):
        batch.set_grid(0 if step < 2 else 4)
        batch.rk3_step(dts)
        for s2, dt in zip(b, dts):
            s2.rk3_step(dt)
    batch.synchronize()
    for k, (s1, s2) in enumerate(zip(a, b)):
        assert np.array_equal(s1.get_state(), s2.get_state()), MEMBERS[k]
    before = [s1.get_state() for s1 in a]
    single, policy, ms = batch.autotune(dts, reps=2)
    say(f"  autotune: single={single} policy={policy} {ms:.3f} ms per step")
    assert ms > 0 and policy in (0, 2, 4)
    for k, s1 in enumerate(a):
        assert np.array_equal(s1.get_state(), before[k]), MEMBERS[k]
    batch.rk3_step(dts)
    for s2, dt in zip(b, dts):
        s2.rk3_step(dt)
    batch.synchronize()
    for k, (s1, s2) in enumerate(zip(a, b)):
        assert np.array_equal(s1.get_state(), s2.get_state()), MEMBERS[k]
    batch.close()


# (order, elements per side, scheme, Reynolds number, Riemann solver, correction, jitter)
SCHEME_MEMBERS = [
    (1, 10, "ns", 100.0, "roe", "hu", 0.0), (2, 9, "ns", 4000.0, "hllc", "dg", 0.0),
    (3, 8, "ns", 800.0, "roe", "hu", 0.0), (4, 7, "ns", 250.0, "roe", "sd", 0.0),
    (5, 6, "ns", 1600.0, "hllc", "hu", 0.0), (3, 11, "ns", 5000.0, "roe", "hu", 0.0),
    (2, 8, "euler", None, "roe", "hu", 0.0), (4, 9, "euler", None, "hllc", "dg", 0.1),
    (3, 10, "reference", None, None, None, 0.0), (5, 7, "reference", None, None, None, 0.08),
]


def _scheme_member(p, n, scheme, Re, riemann, corr, jit, precision, seed):
    mesh = cases.uniform_quad_mesh(n, n, 0.0, 2 * np.pi, 0.0, 2 * np.pi, jitter=jit, seed=seed)
    t = solver.build_tables(mesh.xy, mesh.elem_nodes, p, periodic=(mesh.Lx, mesh.Ly))
    U = tgv_state(t, amp=0.02, seed=seed, mach=0.3)
    if scheme == "reference":
        kw = {}
    else:
        kw = dict(compat="textbook", riemann=riemann, corr=corr, positivity=False)
        if scheme == "ns":
            kw.update(viscous=True, Re=Re, Pr=0.72)
    return t, U, kw


@pytest.mark.parametrize("precision", ["fp64", "fp32bulk"])
def test_batch_textbook_and_navier_stokes_members(precision):
    """Textbook Euler and Navier-Stokes (BR2, each its own Reynolds number)
    members batched with reference-arithmetic ones: one launch per (order,
    metric form, scheme) group; every member's step equals its own bit for
    bit under both grid policies, and the single launch (reference arithmetic
    or Navier-Stokes members only, not a mix) is refused."""
    say(f"case: batch of {len(SCHEME_MEMBERS)} meshes: NS Re 100-5000, textbook Euler, reference; {precision}")
    built = [_scheme_member(*m, precision, seed=70 + k) for k, m in enumerate(SCHEME_MEMBERS)]
    a = [solver.Solver(t, precision=precision, **kw) for t, _, kw in built]
    b = [solver.Solver(t, precision=precision, **kw) for t, _, kw in built]
    for s1, s2, (t, U, _) in zip(a, b, built):
        s1.set_state(U)
        s2.set_state(U)
    dts = np.array([0.5 * s.dt_cfl(0.5) for s in a])
    batch = solver.Batch(a)
    with pytest.raises(solver.ZJError) as ei:
        batch.set_launch(True)
    assert ei.value.code == 4
    for step in range(4):
        batch.set_grid(0 if step < 2 else 4)
        batch.rk3_step(dts)
        for s2, dt in zip(b, dts):
            s2.rk3_step(dt)
    batch.synchronize()
    for k, (s1, s2) in enumerate(zip(a, b)):
        assert np.array_equal(s1.get_state(), s2.get_state()), SCHEME_MEMBERS[k]
        assert s1.residual_l2() == s2.residual_l2(), SCHEME_MEMBERS[k]
    single, policy, ms = batch.autotune(dts, reps=2)
    say(f"  autotune: single={single} policy={policy} {ms:.3f} ms per step")
    assert not single and policy in (0, 2, 4)
    batch.rk3_step(dts)
    for s2, dt in zip(b, dts):
        s2.rk3_step(dt)
    batch.synchronize()
    for k, (s1, s2) in enumerate(zip(a, b)):
        assert np.array_equal(s1.get_state(), s2.get_state()), SCHEME_MEMBERS[k]
        assert np.all(np.isfinite(s1.get_state()))
    batch.close()


@pytest.mark.parametrize("precision", ["fp64", "fp32bulk"])
def test_single_launch_navier_stokes_members(precision):
    """A batch of Navier-Stokes members only (P1-P5, each its own Re, Roe or
    HLLC, HU or DG): one BR2 launch per stage across the orders gives every
    member's own step bit for bit, under both grid policies; the autotuner
    may pick it."""
    say(f"case: single-launch Navier-Stokes batch P1-P5 vs own steps, {precision}")
    ns = [m for m in SCHEME_MEMBERS if m[2] == "ns"]
    built = [_scheme_member(*m, precision, seed=130 + k) for k, m in enumerate(ns)]
    a = [solver.Solver(t, precision=precision, **kw) for t, _, kw in built]
    b = [solver.Solver(t, precision=precision, **kw) for t, _, kw in built]
    for s1, s2, (t, U, _) in zip(a, b, built):
        s1.set_state(U)
        s2.set_state(U)
    dts = np.array([0.5 * s.dt_cfl(0.5) for s in a])
    batch = solver.Batch(a)
    batch.set_launch(True)
    for step in range(3):
        batch.set_grid(0 if step < 2 else 4)
        batch.rk3_step(dts)
        for s2, dt in zip(b, dts):
            s2.rk3_step(dt)
    batch.synchronize()
    for k, (s1, s2) in enumerate(zip(a, b)):
        assert np.array_equal(s1.get_state(), s2.get_state()), ns[k]
    single, policy, ms = batch.autotune(dts, reps=2)
    say(f"  autotune: single={single} policy={policy} {ms:.3f} ms per step")
    batch.rk3_step(dts)
    for s2, dt in zip(b, dts):
        s2.rk3_step(dt)
    for k, (s1, s2) in enumerate(zip(a, b)):
        assert np.array_equal(s1.get_state(), s2.get_state()), ns[k]
    batch.close()


def test_member_calls_wait_for_the_batch_step():
    """A member's own calls right after zj_batch_rk3_step (no
    zj_batch_synchronize) see the stepped state: its state read, its next
    own step and its CFL step wait for the event the batch step records."""
    say("case: member calls after an unsynchronized batch step, 3 x TGV 96x96 P4 Navier-Stokes fp32bulk")
    built = [_scheme_member(4, 96, "ns", Re, "roe", "hu", 0.0, "fp32bulk", seed=110 + k)
             for k, Re in enumerate((200.0, 1000.0, 3000.0))]
    a = [solver.Solver(t, precision="fp32bulk", **kw) for t, _, kw in built]
    b = [solver.Solver(t, precision="fp32bulk", **kw) for t, _, kw in built]
    for s1, s2, (t, U, _) in zip(a, b, built):
        s1.set_state(U)
        s2.set_state(U)
    dts = np.array([0.5 * s.dt_cfl(0.5) for s in a])
    batch = solver.Batch(a)
    for s2, dt in zip(b, dts):
        s2.rk3_step(dt)
    batch.rk3_step(dts)
    for k, (s1, s2) in enumerate(zip(a, b)):
        assert np.array_equal(s1.get_state(), s2.get_state()), k
    batch.rk3_step(dts)
    for k, (s1, s2) in enumerate(zip(a, b)):  # own step right after an unsynchronized batch step
        s1.rk3_step(dts[k])
        s2.rk3_step(dts[k])
        s2.rk3_step(dts[k])
        assert s1.dt_cfl(0.5) == s2.dt_cfl(0.5), k
        assert np.array_equal(s1.get_state(), s2.get_state()), k
    batch.close()


def test_batch_member_changes():
    """While a batch holds a handle, changes that would leave the batch's
    launch arguments stale are refused (tile, metric form, autotune, destroy),
    and a boundary-table change that moves the member's bc_data buffer is
    taken up by the batch's next step."""
    say("case: batch member changes: refusals and a moved bc_data buffer")
    from tests import fr_case
    mesh = cases.uniform_quad_mesh(6, 5, 0.0, 1.0, 0.0, 1.0)
    t = solver.build_tables(mesh.xy, mesh.elem_nodes, 3, periodic=None)
    rho, u, v, p = fr_case.freestream()
    U = cases.perturb(np.tile(cases.prim_to_cons(np.array(rho), np.array(u), np.array(v), np.array(p), 1.4),
                              (t.E * t.S, 1)), 0.01, 9).reshape(-1)
    a, b = solver.Solver(t), solver.Solver(t)
    for s in (a, b):
        s.set_state(U)
    batch = solver.Batch([a])
    half = a.info()["elems_per_block"] // 2
    for call in (lambda: a.set_tile(half), lambda: a.autotune(1), lambda: a.close(),
                 lambda: a.set_geometry_mode(0 if a.set_geometry_mode(-1) else 1)):
        with pytest.raises(solver.ZJError) as ei:
            call()
        assert ei.value.code == 6
    batch.rk3_step([1e-6])
    b.rk3_step(1e-6)
    # a far-field state per boundary face instead of one shared row: the bc_data buffer grows
    bnd = t.faces[:, 2] < 0
    assert t.bc_data.size == 8 * int(bnd.sum()) and bnd.sum() == 22
    fs = np.array([rho * 1.01, u, v, p * 0.99, 0, 0, 0, 0])
    for s in (a, b):
        s.set_boundary(t.bc_type, np.where(bnd, np.cumsum(bnd) - 1, -1), np.tile(fs, 2 * int(bnd.sum())))
    batch.rk3_step([1e-6])
    b.rk3_step(1e-6)
    batch.synchronize()
    assert np.array_equal(a.get_state(), b.get_state())
    batch.close()
    a.close()


def test_autotune_keeps_results():
    say("case: autotune tile and XCD order, TGV 32x32 P4")
    mesh = cases.tgv_mesh(32)
    t = solver.build_tables(mesh.xy, mesh.elem_nodes, 4, periodic=(mesh.Lx, mesh.Ly))
    s = solver.Solver(t)
    U = tgv_state(t, amp=0.02, seed=40)
    s.set_state(U)
    R0 = s.residual()
    epb, xcd, ms = s.autotune(5)
    say(f"  chose elems_per_block={epb} xcd_groups={xcd} ({ms * 1e3:.1f} us per residual)")
    assert epb in (10, 5) and xcd >= 1 and ms > 0
    assert np.array_equal(s.residual(), R0)
    ref = solver.Solver(t)
    ref.set_state(U)
    for _ in range(2):
        s.rk3_step(1e-3)
        ref.rk3_step(1e-3)
    assert np.array_equal(s.get_state(), ref.get_state())


@pytest.mark.parametrize("scheme", ["reference", "navier_stokes"])
def test_batch_c4_full_size(scheme):
    """The batched configuration at full size (64 seeded meshes, N ~
    U{16..128} elements per side, P ~ U{1..5}, Mach ~ U[0.1, 0.8], alternating
    Taylor-Green / vortex, FP32-bulk, PCG64 seed 4; 35.8 M DOF) with the
    reference arithmetic, and as the benchmark builds it: textbook FR (VCJH
    HU, Roe) with the Navier-Stokes terms, Re ~ U[100, 5000] (PCG64 [4, 1]).
    Two batched steps (one workgroup per tile, then four per CU looping) equal
    every member's own steps bit for bit."""
    import math
    say(f"case: batch C4 full size, 64 seeded meshes P1-P5, {scheme}, fp32bulk, 2 steps, grid policies 0 and 4")
    rng = np.random.default_rng(4)
    re_rng = np.random.default_rng([4, 1])
    a, b, dts = [], [], []
    for k in range(64):
        n, p, mach = int(rng.integers(16, 129)), int(rng.integers(1, 6)), float(rng.uniform(0.1, 0.8))
        Re = float(re_rng.uniform(100.0, 5000.0))
        mesh = cases.tgv_mesh(n) if k % 2 == 0 else cases.vortex_mesh(n)
        t = solver.build_tables(mesh.xy, mesh.elem_nodes, p, periodic=(mesh.Lx, mesh.Ly))
        U = (cases.tgv_ic(t.pos, mach=mach) if k % 2 == 0
             else cases.vortex_ic(t.pos, u_inf=mach * math.sqrt(1.4))).reshape(-1)
        kw = {} if scheme == "reference" else dict(compat="textbook", riemann="roe", corr="hu", positivity=False,
                                                   viscous=True, Re=Re, Pr=0.72)
        for lst in (a, b):
            s = solver.Solver(t, precision="fp32bulk", cfl=0.5, **kw)
            s.set_state(U)
            lst.append(s)
        dts.append(a[-1].dt_cfl(0.5))
    assert sum(s.n for s in a) == 35785688
    dts = np.array(dts)
    batch = solver.Batch(a)
    for policy in (0, 4):
        batch.set_grid(policy)
        batch.rk3_step(dts)
        for s, dt in zip(b, dts):
            s.rk3_step(dt)
    batch.synchronize()
    for k, (s1, s2) in enumerate(zip(a, b)):
        assert np.array_equal(s1.get_state(), s2.get_state()), k
    batch.close()
    for s in a + b:
        s.close()


@pytest.mark.parametrize("precision", ["fp64", "fp32bulk"])
def test_batch_step_vjp_equals_members(precision):
    """Two SSP-RK3 steps backwards for a batch of every scheme (Navier-Stokes
    Re 100-5000, textbook Euler, reference arithmetic; P1-P5, both metric
    forms), one stage-VJP launch per group: each member's cotangent, dJ/d(dt)
    and dJ/dRe equal its own zj_rk3_step_vjp to round-off (the batched kernel
    is its own compilation of the same stage-VJP body, so the compiler may
    fuse multiplies and adds differently: bar 1e-13 relative in FP64, 1e-5 in
    FP32-bulk); the pending cross-tile edge cotangents pass from one batched
    step to the next."""
    say(f"case: batched step VJP of {len(SCHEME_MEMBERS)} meshes vs own step VJPs, {precision}")
    built = [_scheme_member(*m, precision, seed=110 + k) for k, m in enumerate(SCHEME_MEMBERS)]
    a = [solver.Solver(t, precision=precision, **kw) for t, _, kw in built]
    b = [solver.Solver(t, precision=precision, **kw) for t, _, kw in built]
    states = []
    for k, (s1, s2, (t, U, _)) in enumerate(zip(a, b, built)):
        s1.set_state(U)
        s2.set_state(U)
        rng = np.random.default_rng(120 + k)
        ubar = rng.standard_normal(U.shape)
        s1.adjoint_set(ubar)
        s2.adjoint_set(ubar)
        states.append((U, U * (1.0 + 1e-3 * rng.standard_normal(U.shape))))
    dts = np.array([0.5 * s.dt_cfl(0.5) for s in a])
    tol = 1e-13 if precision == "fp64" else 1e-5
    batch = solver.Batch(a)
    for step in range(2):
        for s1, s2, st in zip(a, b, states):
            s1.set_state(st[1 - step])
            s2.set_state(st[1 - step])
        dtb = batch.rk3_step_vjp(dts, want_dt=True)
        for k, (s2, dt) in enumerate(zip(b, dts)):
            d2 = s2.rk3_step_vjp(dt)
            assert abs(dtb[k] - d2) <= tol * abs(d2), (SCHEME_MEMBERS[k], step, dtb[k], d2)
    worst = 0.0
    for k, (s1, s2) in enumerate(zip(a, b)):
        x, y = s1.adjoint_get(), s2.adjoint_get()
        r = float(np.linalg.norm(x - y) / np.linalg.norm(y))
        worst = max(worst, r)
        say(f"  member {k} {SCHEME_MEMBERS[k][:3]}: rel L2 {r:.2e}")
        assert r <= tol, SCHEME_MEMBERS[k]
        if SCHEME_MEMBERS[k][2] == "ns":
            ra, rb = s1.adjoint_get_re(), s2.adjoint_get_re()
            assert abs(ra - rb) <= tol * abs(rb), (SCHEME_MEMBERS[k], ra, rb)
    say(f"  worst {worst:.2e}")
    batch.close()
