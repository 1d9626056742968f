"""Navier-Stokes terms on triangles (BR2, textbook scheme) on an MI355X
against the triangle oracle with its viscous terms (oracle/tri_oracle.py,
pinned on the CPU in test_tri_oracle.py: the viscous flux against the C
oracle's, the corrected gradient exact for polynomials, discrete Gauss
identity, freestream, conservation, linear stability with and without the
penalty, the Taylor-Green decay rate).

Bar: FP64 residual within 1e-12 relative L2 (P1-P3, Roe/HLLC, every
correction family); 20 steps within 1e-10; FP32-bulk residual within 1e-5;
the triangle configuration (C3: 500x500 split quads, P3, HLLC, GA, Re 5000,
Mach 0.5) at full size through freestream, conservation and FP32-bulk vs
FP64.
"""
import math

import numpy as np
import pytest

from oracle import tri_oracle as T
from tests.test_gpu_parity import rel, say
from zhijian_amd import cases
from zhijian_amd._lib import ZJError
from zhijian_amd.tri import TriSolver, sp_coords, tri_faces

pytestmark = pytest.mark.gpu

TOL = 1e-12


def _pair(n, p, riemann, corr="dg", Re=100.0, precision="fp64", jitter=0.2, seed=1, order="hilbert", L=1.0):
    m = cases.split_tri_mesh(n, jitter=jitter, seed=seed, L=L)
    s = TriSolver(m.xy, m.tris, p, periodic=(L, L), riemann=riemann, precision=precision, order=order, corr=corr,
                  viscous=True, Re=Re)
    o = T.TriTextbook(p, m.xy, m.tris, s.faces, s.normals, riemann=riemann, corr=corr).set_viscous(Re)
    return s, o, m


def _state(m, p, seed, mach=0.3):
    pos = sp_coords(m.xy, m.tris, p)
    return cases.perturb(cases.tgv_ic(pos * 2 * math.pi / m.L, mach=mach), 0.05, seed).reshape(-1)


@pytest.mark.parametrize("p", [1, 2, 3])
@pytest.mark.parametrize("riemann,corr", [("roe", "dg"), ("hllc", "ga")])
def test_tri_viscous_residual(p, riemann, corr):
    say(f"case: triangle viscous residual, periodic 6x6 jitter 0.2 P{p} {riemann}/{corr} Re 100")
    s, o, m = _pair(6, p, riemann, corr, seed=p)
    assert s.info()["cross_tile_edges"] > 0
    U = _state(m, p, 10 + p)
    s.set_state(U)
    r = rel(s.residual(), o.residual(U))
    say(f"  rel L2 {r:.3e}")
    assert r < TOL


@pytest.mark.parametrize("corr", ["sd", "hu"])
def test_tri_viscous_residual_families(corr):
    say(f"case: triangle viscous residual, periodic 6x6 P3 roe/{corr} Re 50")
    s, o, m = _pair(6, 3, "roe", corr, Re=50.0, seed=4)
    U = _state(m, 3, 20)
    s.set_state(U)
    r = rel(s.residual(), o.residual(U))
    say(f"  rel L2 {r:.3e}")
    assert r < TOL


def test_tri_viscous_20_steps():
    say("case: triangle viscous 20 steps, periodic 8x8 P3 hllc/ga Re 100")
    s, o, m = _pair(8, 3, "hllc", "ga", seed=8)
    U = _state(m, 3, 30)
    s.set_state(U)
    dt = 0.5 * s.dt_cfl(0.5)
    V = U.copy()
    for _ in range(20):
        s.rk3_step(dt)
        V = o.advance(V, dt)
    r = rel(s.get_state(), V)
    say(f"  dt {dt:.3e}  rel L2 after 20 steps {r:.3e}")
    assert r < 1e-10


def test_tri_viscous_fp64_p4_refused():
    """FP64 P4 with the viscous terms is switched off in the library (an
    earlier build of the kernel faulted on the GPU with no cause found)."""
    m = cases.split_tri_mesh(4)
    with pytest.raises(ZJError) as e:
        TriSolver(m.xy, m.tris, 4, periodic=(1.0, 1.0), viscous=True, Re=100.0)
    assert e.value.code == 4


@pytest.mark.parametrize("p", [2, 3, 4, 5])
def test_tri_viscous_fp32bulk_residual(p):
    say(f"case: triangle viscous FP32-bulk residual vs oracle (FP64), periodic 6x6 P{p} hllc")
    s, o, m = _pair(6, p, "hllc", "ga", precision="fp32bulk", seed=3)
    U = _state(m, p, 21)
    s.set_state(U)
    r = rel(s.residual(), o.residual(U))
    say(f"  rel L2 {r:.3e}")
    assert r < 1e-5


@pytest.mark.parametrize("viscous", [False, True])
def test_tri_dt_rule(viscous):
    """zj_tri_dt_cfl against the rule evaluated here: the minimum over the
    points of cfl / ((2p+1)(|u|+c)/d + (p+1)^4 nu_k / (rho d^2)), d the
    inscribed-circle diameter, nu_k = mu max(4/3, gamma/Pr) with the viscous
    terms (the BR2 operator's spectral radius is about 1.3 (p+1)^4 nu / d^2
    on triangles, measured on the oracle's Jacobian)."""
    p, Re, g, Pr = 3, 500.0, 1.4, 0.72
    m = cases.split_tri_mesh(8, jitter=0.2, seed=9)
    kw = dict(viscous=True, Re=Re) if viscous else {}
    s = TriSolver(m.xy, m.tris, p, periodic=(1.0, 1.0), riemann="hllc", corr="ga", **kw)
    U = _state(m, p, 40)
    s.set_state(U)
    P = m.xy[m.tris]
    a = np.linalg.norm(P[:, 1] - P[:, 0], axis=1)
    b = np.linalg.norm(P[:, 2] - P[:, 1], axis=1)
    c = np.linalg.norm(P[:, 0] - P[:, 2], axis=1)
    area = 0.5 * np.abs((P[:, 1, 0] - P[:, 0, 0]) * (P[:, 2, 1] - P[:, 0, 1])
                        - (P[:, 2, 0] - P[:, 0, 0]) * (P[:, 1, 1] - P[:, 0, 1]))
    d = (4.0 * area / (a + b + c))[:, None]
    q = U.reshape(m.n_elem, -1, 4)
    u, w = q[..., 1] / q[..., 0], q[..., 2] / q[..., 0]
    pr = (g - 1) * q[..., 0] * (q[..., 3] / q[..., 0] - 0.5 * (u * u + w * w))
    ws = np.sqrt(u * u + w * w) + np.sqrt(g * np.maximum(pr, 1e-10) / np.maximum(q[..., 0], 1e-10))
    visc = (p + 1) ** 4 * (1.0 / Re) * max(4.0 / 3.0, g / Pr) if viscous else 0.0
    dt = float(np.min(0.5 / ((2 * p + 1) * ws / d + visc / (q[..., 0] * d * d))))
    got = s.dt_cfl(0.5)
    say(f"  dt {got:.6e} vs {dt:.6e} (viscous {viscous})")
    assert abs(got - dt) <= 1e-13 * dt


def test_tri_viscous_numbering_independent():
    say("case: triangle viscous, Hilbert vs caller order, periodic 12x12 P2 roe, 3 steps")
    a, _, m = _pair(12, 2, "roe", order="hilbert", seed=12)
    b, _, _ = _pair(12, 2, "roe", order="none", seed=12)
    U = _state(m, 2, 5)
    out = []
    for s in (a, b):
        s.set_state(U)
        for _ in range(3):
            s.rk3_step(1e-4)
        out.append(s.get_state())
    r = rel(out[0], out[1])
    say(f"  rel L2 {r:.3e}")
    assert r < 1e-13


def test_tri_viscous_refusals():
    """Boundary faces with the viscous terms are refused with
    ZJ_ERR_UNSUPPORTED (the adjoint of viscous handles is covered in
    test_gpu_tri_visc_adjoint.py)."""
    m = cases.split_tri_mesh(3)
    faces, nrm = tri_faces(m.xy, m.tris)
    bt = np.where(faces[:, 2] < 0, 2, 0).astype(np.int32)
    with pytest.raises(ZJError) as e:
        TriSolver(m.xy, m.tris, 2, faces=faces, normals=nrm, bc_type=bt, viscous=True, Re=100.0)
    assert e.value.code == 4


def test_tri_viscous_c3_full_size_properties():
    """The triangle configuration with its Navier-Stokes terms at full size
    (500x500 split quads, interior nodes jittered 0.2 h, P3, HLLC, GA,
    Re 5000, Mach 0.5 at 2 degrees plus Fourier modes): zero residual of a
    uniform flow; discrete conservation of the residual of the Fourier state;
    one step keeps the totals; FP32-bulk against FP64 after 100 steps."""
    say("case: triangles C3 full size 500x500 P3 hllc GA Re 5000: freestream, conservation, FP32-bulk 100 steps")
    m = cases.split_tri_mesh(500, jitter=0.2, seed=2024)
    kw = dict(periodic=(1.0, 1.0), riemann="hllc", corr="ga", viscous=True, Re=5000.0)
    a = TriSolver(m.xy, m.tris, 3, **kw)
    b = TriSolver(m.xy, m.tris, 3, precision="fp32bulk", **kw)
    Np = 10
    ops = T.operators(3)
    w = np.ones(Np) @ np.linalg.inv(ops["V"] @ ops["V"].T)
    Pt = m.xy[m.tris]
    J = 0.25 * ((Pt[:, 1, 0] - Pt[:, 0, 0]) * (Pt[:, 2, 1] - Pt[:, 0, 1])
                - (Pt[:, 2, 0] - Pt[:, 0, 0]) * (Pt[:, 1, 1] - Pt[:, 0, 1]))
    free = np.tile(cases.prim_to_cons(np.array(1.0), np.array(0.8), np.array(0.3), np.array(1 / 1.4 / 0.25), 1.4),
                   (m.n_elem * Np, 1)).reshape(-1)
    a.set_state(free)
    # round-off of O(10) fluxes and state differences over h = 0.002 through the
    # lifting (measured 2e-9); a defect would be O(1/h)
    assert np.abs(a.residual()).max() < 1e-8
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
    after = np.einsum("e,i,eiv->v", J, w, a.get_state().reshape(m.n_elem, Np, 4))
    assert np.all(np.abs(after - before) <= 1e-12 * np.abs(before).max())
    a.set_state(U)
    for _ in range(100):
        a.rk3_step(dt)
        b.rk3_step(dt)
    Ua, Ub = a.get_state(), b.get_state()
    r = rel(Ub, Ua)
    rinf = np.abs(Ub - Ua).max() / np.abs(Ua).max()
    say(f"  FP32-bulk vs FP64 after 100 steps: rel L2 {r:.3e}, rel Linf {rinf:.3e}")
    assert r < 1e-6
    a.close()
    b.close()
