"""Discrete adjoint of the textbook scheme (ZJ_COMPAT_TEXTBOOK: VCJH correction
with metrics, Rusanov / Roe / HLLC of euler.hpp) on an MI355X.

As for the reference arithmetic (test_gpu_adjoint.py), the pin is the
identity <w, J v> = <J^T w, v> with J v a central finite difference of the
textbook ORACLE's forward map (oracle/zj_textbook.c, FP64) and J^T w the HIP
adjoint; bar 1e-6 relative.  The Riemann solvers' VJPs, the transposed VCJH
correction, the transposed interpolation of the transformed flux and the 1/J
of non-affine elements each move the HIP side by O(1) when wrong.
"""
import numpy as np
import pytest

from tests import fr_case
from tests.test_gpu_adjoint import TOL, direction, dot_check, fd
from tests.test_gpu_parity import say, tgv_state, vortex_state
from zhijian_amd import cases, solver

pytestmark = pytest.mark.gpu


def _pair(mesh, p, riemann, corr, periodic=True, precision="fp64", bc_types=None, fs=None, positivity=False):
    per = (mesh.Lx, mesh.Ly) if periodic else None
    t = solver.build_tables(mesh.xy, mesh.elem_nodes, p, periodic=per, bc_type_per_face=bc_types, fs=fs)
    s = solver.Solver(t, precision=precision, compat="textbook", riemann=riemann, corr=corr,
                      positivity=positivity)
    oc = fr_case.build(mesh.xy, mesh.elem_nodes, p, periodic=per, bc_type_per_face=bc_types, fs=fs)
    return s, oc.textbook(riemann, corr), t


@pytest.mark.parametrize("p", [1, 2, 3, 4, 5])
@pytest.mark.parametrize("riemann", ["rusanov", "roe", "hllc"])
def test_textbook_residual_vjp(p, riemann):
    """Every order and Riemann solver, affine and jittered (per-point metric) meshes."""
    for jit in (0.0, 0.12):
        say(f"case: textbook residual VJP, periodic 5x4 P{p} {riemann}/hu jitter {jit}")
        mesh = cases.uniform_quad_mesh(5, 4, 0.0, 1.0, 0.0, 1.0, jitter=jit, seed=p)
        s, tb, t = _pair(mesh, p, riemann, "hu")
        U = tgv_state(t, amp=0.05, seed=40 + p, mach=0.3)
        s.set_state(U)
        w = np.random.default_rng(p).standard_normal(U.shape)
        v = direction(U, 50 + p)
        assert dot_check(float(w @ fd(tb.residual, U, v)), float(s.residual_vjp(w) @ v), "residual") < TOL


@pytest.mark.parametrize("corr", ["dg", "sd", "ga"])
def test_textbook_residual_vjp_corrections(corr):
    say(f"case: textbook residual VJP, vortex 6x6 P3 roe/{corr}")
    s, tb, t = _pair(cases.vortex_mesh(6), 3, "roe", corr)
    U = vortex_state(t, amp=0.05, seed=61)
    s.set_state(U)
    w = np.random.default_rng(62).standard_normal(U.shape)
    v = direction(U, 63)
    assert dot_check(float(w @ fd(tb.residual, U, v)), float(s.residual_vjp(w) @ v), corr) < TOL


@pytest.mark.parametrize("bct", [1, 2, 4, 6])
@pytest.mark.parametrize("riemann", ["roe", "hllc"])
def test_textbook_residual_vjp_boundaries(bct, riemann):
    """Open jittered mesh, every boundary face of one type (wall, slip wall, far field, outflow)."""
    say(f"case: textbook residual VJP, open 5x4 P3 {riemann}, boundary type {bct}")
    mesh = cases.uniform_quad_mesh(5, 4, 0.0, 1.0, 0.0, 1.0, jitter=0.15, seed=3)
    faces = solver._lib.build_faces(mesh.elem_nodes)
    types = np.zeros(faces.shape[0], np.int32)
    types[faces[:, 2] < 0] = bct
    fs = (1.1, 0.3, 0.1, 1.5e3)
    s, tb, t = _pair(mesh, 3, riemann, "hu", periodic=False, bc_types=types, fs=fs)
    shp = t.pos.shape[:2]
    U = cases.perturb(cases.prim_to_cons(np.full(shp, 1.0), np.full(shp, 0.3), np.full(shp, -0.1),
                                         np.full(shp, 2e3), 1.4), 0.05, 4).reshape(-1)
    s.set_state(U)
    w = np.random.default_rng(5).standard_normal(U.shape)
    v = direction(U, 6)
    assert dot_check(float(w @ fd(tb.residual, U, v)), float(s.residual_vjp(w) @ v), f"BC {bct}") < TOL


@pytest.mark.parametrize("precision", ["fp64", "fp32bulk"])
@pytest.mark.parametrize("p", [2, 4])
def test_textbook_step_vjp_and_dt(precision, p):
    """One SSP-RK3 step of the textbook scheme: state and dt cotangents
    (P2 runs its adjoint at the half tile, P4 at the full one)."""
    say(f"case: textbook step VJP, TGV 6x6 P{p} roe/hu {precision}")
    mesh = cases.uniform_quad_mesh(6, 6, 0.0, 2 * np.pi, 0.0, 2 * np.pi, jitter=0.1, seed=p)
    s, tb, t = _pair(mesh, p, "roe", "hu", precision=precision)
    U = tgv_state(t, amp=0.02, seed=70 + p, mach=0.3)
    dt = 2e-3
    ubar = np.random.default_rng(71).standard_normal(U.shape)
    v = direction(U, 72)
    lhs = float(ubar @ fd(lambda X: tb.advance(X, dt, False), U, v))
    s.set_state(U)
    s.adjoint_set(ubar)
    dtb = s.rk3_step_vjp(dt)
    g = s.adjoint_get()
    tol = TOL if precision == "fp64" else 2e-4
    assert dot_check(lhs, float(g @ v), "step") < tol
    h = 1e-7
    dfd = (float(ubar @ tb.advance(U, dt + h, False)) - float(ubar @ tb.advance(U, dt - h, False))) / (2 * h)
    rd = abs(dtb - dfd) / abs(dfd)
    say(f"  dJ/ddt: adjoint {dtb:.12e}  FD {dfd:.12e}  rel {rd:.2e}")
    assert rd < (1e-6 if precision == "fp64" else 2e-4)


def test_textbook_step_vjp_with_limiter():
    """The textbook step with the reference's positivity limiter switched on."""
    say("case: textbook step VJP with the limiter, vortex 6x6 P3 hllc/ga")
    s, tb, t = _pair(cases.vortex_mesh(6), 3, "hllc", "ga", positivity=True)
    U = vortex_state(t, amp=0.05, seed=81)
    dt = 1e-3
    ubar = np.random.default_rng(82).standard_normal(U.shape)
    v = direction(U, 83)
    lhs = float(ubar @ fd(lambda X: tb.advance(X, dt, True), U, v))
    s.set_state(U)
    s.adjoint_set(ubar)
    s.rk3_step_vjp(dt)
    assert dot_check(lhs, float(s.adjoint_get() @ v), "step+limiter") < TOL



@pytest.mark.parametrize("p,viscous", [(2, False), (3, False), (3, True)])
def test_native_sweep_equals_stepwise(p, viscous):
    """zj_forward_checkpointed + zj_adjoint_sweep (segments recorded on the
    device tape, the last one short) against the step-by-step VJP of the same
    trajectory: state and dt cotangents to round-off."""
    say(f"case: native checkpointed sweep vs step-wise VJP, TGV 6x6 P{p} roe/hu viscous {viscous}")
    mesh = cases.uniform_quad_mesh(6, 6, 0.0, 2 * np.pi, 0.0, 2 * np.pi)
    t = solver.build_tables(mesh.xy, mesh.elem_nodes, p, periodic=(mesh.Lx, mesh.Ly))
    extra = dict(viscous=True, Re=200.0) if viscous else {}
    s = solver.Solver(t, compat="textbook", riemann="roe", corr="hu", positivity=False, **extra)
    U0 = tgv_state(t, amp=0.02, seed=90 + p, mach=0.3)
    dts = [1e-3 * (1.0 + 0.1 * k) for k in range(7)]
    ubar = np.random.default_rng(91).standard_normal(U0.shape)
    s.set_state(U0)
    s.forward_checkpointed(dts, 3)
    s.adjoint_set(ubar)
    dtb1 = s.adjoint_sweep(dts, 3, want_dt=True)
    g1 = s.adjoint_get()
    # without dt cotangents the VJP kernels skip the forward residual: same bits
    s.set_state(U0)
    s.forward_checkpointed(dts, 3)
    s.adjoint_set(ubar)
    assert s.adjoint_sweep(dts, 3, want_dt=False) is None
    assert np.array_equal(s.adjoint_get(), g1)
    traj = [U0]
    s.set_state(U0)
    for dt in dts:
        s.rk3_step(dt)
        traj.append(s.get_state())
    s.adjoint_set(ubar)
    dtb2 = [0.0] * len(dts)
    for n in range(len(dts) - 1, -1, -1):
        s.set_state(traj[n])
        dtb2[n] = s.rk3_step_vjp(dts[n])
    g2 = s.adjoint_get()
    r = np.linalg.norm(g1 - g2) / np.linalg.norm(g2)
    rd = max(abs(a - b) / abs(b) for a, b in zip(dtb1, dtb2))
    say(f"  cotangent rel L2 {r:.3e}, dt cotangents max rel {rd:.3e}")
    assert r < 1e-12 and rd < 1e-10


@pytest.mark.parametrize("viscous", [False, True])
def test_kinetic_energy_seed_on_device(viscous):
    """zj_adjoint_seed_kinetic_energy against the host objective and its
    gradient (adjoint.kinetic_energy / kinetic_energy_grad) on a jittered
    mesh, whose elements the library renumbers along a Hilbert curve: J to
    1e-13, the seeded cotangent (read back before any step) to 1e-14."""
    from zhijian_amd import adjoint
    from zhijian_amd.hiprt import DeviceArray
    say(f"case: kinetic-energy objective and seed on the GPU, TGV 9x9 jittered P3, viscous {viscous}")
    mesh = cases.uniform_quad_mesh(9, 9, 0.0, 2 * np.pi, 0.0, 2 * np.pi, jitter=0.0 if viscous else 0.15, seed=5)
    t = solver.build_tables(mesh.xy, mesh.elem_nodes, 3, periodic=(mesh.Lx, mesh.Ly))
    extra = dict(viscous=True, Re=300.0) if viscous else {}
    s = solver.Solver(t, compat="textbook", riemann="roe", corr="hu", positivity=False, **extra)
    U = tgv_state(t, amp=0.05, seed=97, mach=0.3)
    s.set_state(U)
    w = DeviceArray((t.ops["qw"][None, :] * t.J.reshape(t.E, t.S)).reshape(-1))
    J = s.adjoint_seed_kinetic_energy(w.ptr)
    Jh = adjoint.kinetic_energy(U, t)
    assert abs(J - Jh) <= 1e-13 * abs(Jh)
    g, gh = s.adjoint_get(), adjoint.kinetic_energy_grad(U, t)
    assert np.linalg.norm(g - gh) <= 1e-14 * np.linalg.norm(gh)


@pytest.mark.parametrize("precision,viscous", [("fp64", False), ("fp64", True), ("fp32bulk", True)])
def test_tape_recorded_by_forward_equals_recomputed(precision, viscous):
    """Segments recorded on the tape by the forward pass itself (none, the
    last one, all three) give the same final state, cotangents, dt and Re
    cotangents, bit for bit; a second sweep of the same trajectory after a
    sweep that recomputed reuses nothing stale."""
    say(f"case: tape recorded by the forward pass vs recomputed, TGV 8x8 P3 {precision} viscous {viscous}")
    mesh = cases.uniform_quad_mesh(8, 8, 0.0, 2 * np.pi, 0.0, 2 * np.pi)
    t = solver.build_tables(mesh.xy, mesh.elem_nodes, 3, periodic=(mesh.Lx, mesh.Ly))
    extra = dict(viscous=True, Re=300.0) if viscous else {}
    s = solver.Solver(t, precision=precision, compat="textbook", riemann="hllc", corr="hu", positivity=False,
                      **extra)
    U0 = tgv_state(t, amp=0.02, seed=95, mach=0.3)
    dts = [1e-3 * (1.0 + 0.05 * k) for k in range(8)]
    ubar = np.random.default_rng(96).standard_normal(U0.shape)
    out = []
    for segs in (0, 1, 3, 0):
        s.adjoint_tape_segments(segs)
        s.set_state(U0)
        s.forward_checkpointed(dts, 3)
        UN = s.get_state()
        s.adjoint_set(ubar)
        dtb = s.adjoint_sweep(dts, 3, want_dt=True)
        out.append((UN, s.adjoint_get(), dtb, s.adjoint_get_re()))
    for k in (1, 2, 3):
        assert np.array_equal(out[k][0], out[0][0]), k
        assert np.array_equal(out[k][1], out[0][1]), k
        assert out[k][2] == out[0][2], k
        assert out[k][3] == out[0][3], k
    # the handle steps on as before: the recorded segments left its state and traces consistent
    s.adjoint_tape_segments(3)
    s.set_state(U0)
    s.forward_checkpointed(dts, 3)
    s.rk3_step(dts[0])
    a = s.get_state()
    s.set_state(out[0][0])
    s.rk3_step(dts[0])
    assert np.array_equal(a, s.get_state())
