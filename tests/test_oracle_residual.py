"""Pins the oracle's residual kernels (reference kernels.cu:30-67, 157-194,
346-450, 528-676, 813-1034; fr_solver.cpp:351-608, 784-833).

What fixes each kernel independently of the oracle's own formula:
* divergence: exact -grad p for a rest state with polynomial pressure;
* interpolation: polynomial values at the physical flux points, with edges 2
  and 3 stored in reversed order;
* F_diff: slot placement written out from kernels.cu:402-425 (the right element
  uses its own slot fp, the common flux pairs U_L[fp] with U_R[Fp-1-fp]);
* correction: impulse responses written out from kernels.cu:971-1020;
* whole residual: freestream preservation, periodic and far field.
"""
import math

import numpy as np
import pytest

from oracle import pyoracle as O
from tests import fr_case
from zhijian_amd import cases


def _rest_state_poly(case, deg, seed, gamma=1.4):
    rng = np.random.default_rng(seed)
    c = rng.standard_normal((deg + 1, deg + 1))
    pos = O.sp_coords(case.elem_nodes, case.xy, case.ops)
    x, y = pos[..., 0], pos[..., 1]
    P = 5.0 + sum(c[a, b] * x ** a * y ** b for a in range(deg + 1) for b in range(deg + 1)
                  if a + b <= deg)
    dPx = sum(a * c[a, b] * x ** (a - 1) * y ** b for a in range(1, deg + 1) for b in range(deg + 1)
              if a + b <= deg)
    dPy = sum(b * c[a, b] * x ** a * y ** (b - 1) for a in range(deg + 1) for b in range(1, deg + 1)
              if a + b <= deg)
    U = np.stack([np.ones_like(P), 0 * P, 0 * P, P / (gamma - 1)], axis=-1)
    return U, dPx, dPy


@pytest.mark.parametrize("p", [1, 2, 3, 4, 5])
def test_divergence_exact_on_affine_elements(p):
    # sheared (affine, non-rectangular) elements: parallelogram mesh
    mesh = cases.uniform_quad_mesh(3, 2, 0.0, 3.0, 0.0, 2.0)
    xy = mesh.xy.copy()
    xy[:, 0] += 0.3 * xy[:, 1]
    case = fr_case.build(xy, mesh.elem_nodes, p)
    s = case.solver()
    U, dPx, dPy = _rest_state_poly(case, p, seed=p)
    s.k_flux_sp(U)
    div = s.k_divergence().reshape(case.E, case.S, 4)
    scale = np.abs(U[..., 3]).max()
    assert np.abs(div[..., 0]).max() < 1e-12 * scale
    assert np.abs(div[..., 3]).max() < 1e-12 * scale
    assert np.allclose(div[..., 1], -dPx, atol=1e-11 * scale)
    assert np.allclose(div[..., 2], -dPy, atol=1e-11 * scale)


@pytest.mark.parametrize("p", [1, 3, 5])
def test_interpolation_exact_at_physical_flux_points(p):
    mesh = cases.uniform_quad_mesh(2, 2, 0.0, 1.0, 0.0, 1.0)
    xy = mesh.xy.copy(); xy[:, 1] += 0.2 * xy[:, 0]
    case = fr_case.build(xy, mesh.elem_nodes, p)
    s = case.solver()
    f = lambda x, y: 1.0 + x ** p - 2.0 * y ** p + 0.5 * x * y ** (p - 1)  # total degree p
    pos = O.sp_coords(case.elem_nodes, case.xy, case.ops)
    U = np.stack([f(pos[..., 0], pos[..., 1])] * 4, axis=-1)
    ufp = s.k_interp(U).reshape(case.E, 4, p + 1, 4)
    g, _ = O.gauss(p + 1)
    for e in range(case.E):
        P = case.xy[case.elem_nodes[e]]
        for edge, (a, b) in enumerate([(0, 1), (1, 2), (2, 3), (3, 0)]):
            # edge k runs from corner k to corner k+1; slot fp sits at Gauss point g[fp]
            # of that traversal, which for edges 2 and 3 is the reversed reference direction
            t = 0.5 * (1 + g)
            X = (1 - t)[:, None] * P[a] + t[:, None] * P[b]
            assert np.allclose(ufp[e, edge, :, 0], f(X[:, 0], X[:, 1]), atol=1e-12), (e, edge)


def test_flux_diff_slot_placement():
    """Two elements side by side; face between elem0 edge1 (left) and elem1 edge3."""
    mesh = cases.uniform_quad_mesh(2, 1, 0.0, 2.0, 0.0, 1.0)
    p = 2
    case = fr_case.build(mesh.xy, mesh.elem_nodes, p)
    s = case.solver()
    Fp = p + 1
    rng = np.random.default_rng(5)
    ufp = np.zeros((2, 4, Fp, 4))
    for e in range(2):
        for k in range(4):
            for f in range(Fp):
                ufp[e, k, f] = [1 + 0.1 * rng.random(), 0.2 * rng.random(), 0.1 * rng.random(),
                                3 + rng.random()]
    s.set_ufp(ufp)
    fc = s.k_riemann().reshape(case.faces.n, Fp, 4)
    fd = s.k_flux_diff().reshape(2, 4, Fp, 4)
    a = [i for i in range(case.faces.n) if case.faces.re[i] >= 0]
    assert len(a) == 1
    fi = a[0]
    assert (case.faces.le[fi], case.faces.ll[fi], case.faces.re[fi], case.faces.rl[fi]) == (0, 1, 1, 3)
    n = case.normals[fi]
    assert np.allclose(n, [1.0, 0.0])
    for f in range(Fp):
        UL = ufp[0, 1, f]
        UR = ufp[1, 3, Fp - 1 - f]            # kernels.cu:577
        Fc = O.rusanov(UL, UR, n[0], n[1])
        assert np.allclose(fc[fi, f], Fc, rtol=1e-15)
        fxL, fyL = O.flux_xy(UL)
        assert np.allclose(fd[0, 1, f], Fc - (fxL * n[0] + fyL * n[1]), rtol=1e-14)
        fxR, fyR = O.flux_xy(ufp[1, 3, f])    # kernels.cu:403-405: same fp
        assert np.allclose(fd[1, 3, f], -Fc + (fxR * n[0] + fyR * n[1]), rtol=1e-14)


@pytest.mark.parametrize("p", [1, 2, 4])
def test_correction_impulse_responses(p):
    mesh = cases.uniform_quad_mesh(1, 1, 0.0, 1.0, 0.0, 1.0)
    case = fr_case.build(mesh.xy, mesh.elem_nodes, p)
    s = case.solver()
    Fp = p + 1; S = Fp * Fp
    corr = case.ops["corr"].reshape(4, S)
    for edge in range(4):
        for f in range(Fp):
            for v in (0, 3):
                fd = np.zeros((1, 4, Fp, 4)); fd[0, edge, f, v] = 1.0
                s.set_fdiff(fd)
                out = s.k_correction(np.zeros(S * 4)).reshape(S, 4)
                expect = np.zeros((S, 4))
                for sp in range(S):
                    eta_i, xi_j = sp // Fp, sp % Fp    # kernels.cu:966-967
                    hit = (xi_j == f) if edge in (0, 2) else (eta_i == f)
                    if hit:
                        expect[sp, v] = -corr[edge, sp]
                assert np.array_equal(out, expect), (edge, f, v)


def test_freestream_periodic_is_preserved():
    for p in (1, 3, 4):
        mesh = cases.uniform_quad_mesh(5, 4, 0.0, 1.0, 0.0, 1.0, jitter=0.2, seed=1)
        case = fr_case.from_quadmesh(mesh, p)
        s = case.solver()
        U = np.tile(cases.prim_to_cons(np.array(1.2), np.array(0.3), np.array(-0.4), np.array(2.0), 1.4),
                    (case.E * case.S, 1))
        R = s.residual(U)
        assert np.abs(R).max() < 1e-11, p


def test_freestream_farfield_reference_mesh(golden_meshes):
    """The reference's freestream intent: uniform_quad_10x10.msh, default
    SimParams (p_inf=101325, rho_inf=1.225, M=0.5), FarField boundaries."""
    from tests.conftest import mesh_from_golden
    m = mesh_from_golden(golden_meshes, "uniform_quad_10x10")
    case = fr_case.build(m["xy"], m["elem_nodes"], 3)
    assert int(np.sum(case.faces.re < 0)) == 40
    s = case.solver()
    rho, u, v, pinf = fr_case.freestream()
    U = np.tile(cases.prim_to_cons(np.array(rho), np.array(u), np.array(v), np.array(pinf), 1.4),
                (case.E * case.S, 1))
    R = s.residual(U)
    # flux scale p_inf/h = 1e6: round-off level only
    assert np.abs(R).max() < 1e-8 * pinf / 0.1


def _jacobian(s, U0, eps=1e-7):
    n = U0.size
    Jm = np.zeros((n, n))
    for k in range(n):
        d = np.zeros(n); d[k] = eps
        Jm[:, k] = (s.residual(U0 + d) - s.residual(U0 - d)) / (2 * eps)
    return Jm


@pytest.mark.parametrize("p", [1, 2])
def test_reference_scheme_is_unstable(p):
    """A property of the reference's executed arithmetic, recorded because it
    bounds every multi-step comparison: linearised about a uniform flow on a
    periodic mesh, the residual has eigenvalues with Re(lambda) of the order of
    |lambda| (the g_L terms carry the wrong sign, kernels.cu:1009-1020).  Any
    round-off difference therefore grows like exp(Re(lambda) t)."""
    mesh = cases.uniform_quad_mesh(3, 3, 0.0, 1.0, 0.0, 1.0)
    case = fr_case.from_quadmesh(mesh, p)
    s = case.solver()
    U0 = np.tile(cases.prim_to_cons(np.array(1.0), np.array(0.5), np.array(0.2), np.array(1 / 1.4), 1.4),
                 (case.E * case.S, 1)).reshape(-1)
    lam = np.linalg.eigvals(_jacobian(s, U0))
    assert lam.real.max() > 0.5 * np.abs(lam).max()


def test_residual_l2_is_sequential_sum():
    mesh = cases.vortex_mesh(4)
    case = fr_case.from_quadmesh(mesh, 2)
    s = case.solver()
    pos = O.sp_coords(case.elem_nodes, case.xy, case.ops)
    U = cases.perturb(cases.vortex_ic(pos), 0.05, 3)
    R = s.residual(U)
    l2 = s.residual_l2()
    assert abs(l2 - math.sqrt(math.fsum(R * R))) <= 1e-14 * l2
