"""The gpu:: free functions (zj_gpu_*, the reference's include/gpu/kernels.hpp)
against the oracle's restatement of each reference kernel (oz_k_*), through
the C ABI on device buffers.  Runs on an MI355X.

Bar: within 1e-12 relative L2 of the oracle (FP64; the differences are FMA
contraction and summation order: the matrix cores accumulate the operator
products in their own order), exact where the operation is a copy or a
single rounding.  The dense operator products run on both engines (FP64 MFMA
and VALU), checked the same way.
"""
import math
import sys

import numpy as np
import pytest

from oracle import pyoracle as O
from tests import fr_case
from zhijian_amd import cases, gpu_ops as G
from zhijian_amd.hiprt import DeviceArray as DA

pytestmark = pytest.mark.gpu

TOL = 1e-12


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


_live = []


@pytest.fixture(autouse=True)
def _device_buffers():
    """Buffers made by dptr() live until the test's kernels have finished."""
    yield
    from zhijian_amd.hiprt import hip
    hip().hipDeviceSynchronize()
    _live.clear()


def dptr(host):
    """Device copy of a host array, kept alive until the end of the test; returns its address."""
    a = DA(host)
    _live.append(a)
    return a.ptr


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300))


def setup(p, n=5, jitter=0.15, bc=False, seed=3):
    """A jittered mesh (periodic, or open with far-field/wall/outflow faces), its
    oracle case and a perturbed state."""
    mesh = cases.uniform_quad_mesh(n, n + 1, 0.0, 1.0, 0.0, 1.2, jitter=jitter, seed=seed)
    if bc:
        f0 = O.build_faces(np.ascontiguousarray(mesh.elem_nodes, np.int64), nv=4)
        types = np.zeros(f0.n, np.int32)
        kinds = [O.BC_FARFIELD, O.BC_WALL, O.BC_OUTFLOW, O.BC_SLIPWALL, O.BC_SYMMETRY]
        bidx = np.nonzero(f0.re < 0)[0]
        types[bidx] = [kinds[i % len(kinds)] for i in range(bidx.size)]
        oc = fr_case.build(mesh.xy, mesh.elem_nodes, p, periodic=None, bc_type_per_face=types)
    else:
        oc = fr_case.build(mesh.xy, mesh.elem_nodes, p, periodic=(mesh.Lx, mesh.Ly))
    pos = O.sp_coords(oc.elem_nodes, oc.xy, oc.ops).reshape(-1, 2)
    rho, u, v, pr = fr_case.freestream()
    U0 = np.tile(cases.prim_to_cons(np.array(rho), np.array(u), np.array(v), np.array(pr), 1.4), (pos.shape[0], 1))
    U0[:, 0] *= 1.0 + 0.1 * np.sin(2 * math.pi * pos[:, 0]) * np.cos(2 * math.pi * pos[:, 1])
    U = cases.perturb(U0, 0.05, seed).reshape(-1)
    return oc, U


@pytest.fixture(params=[ENG for ENG in (G.ENGINE_MFMA, G.ENGINE_VALU)], ids=["mfma", "valu"])
def engine(request):
    G.set_operator_engine(request.param)
    yield request.param
    G.set_operator_engine(G.ENGINE_MFMA)


@pytest.mark.parametrize("p", [1, 2, 3, 4, 5])
def test_operator_products(engine, p):
    """interpolateToFluxPoints, computeReferenceGradients, computeFluxDivergence
    (kernels.cu:30-108, 813-872): the oracle's K1/K11 and dense products."""
    say(f"case: operator products P{p} engine {engine}")
    oc, U = setup(p, n=7)
    o = oc.solver()
    E, S, Fp = oc.E, oc.S, p + 1
    ops = oc.ops
    dU = DA(U)
    # K1
    ufp = DA(E * 4 * Fp * 4)
    G.interpolate_to_flux_points(dU.ptr, ufp.ptr, dptr(ops["interp"]), E, S, 4, Fp)
    assert rel(ufp.get(), o.k_interp(U)) < TOL
    # K2: dU/dxi = diff_xi U per element (a plain product of the golden-pinned tables)
    gx, ge = DA(E * S * 4), DA(E * S * 4)
    G.compute_reference_gradients(dU.ptr, gx.ptr, ge.ptr, dptr(ops["diff_xi"]), dptr(ops["diff_eta"]), E, S)
    Ue = U.reshape(E, S, 4)
    want_x = np.einsum("sk,ekv->esv", ops["diff_xi"].reshape(S, S), Ue).reshape(-1)
    want_e = np.einsum("sk,ekv->esv", ops["diff_eta"].reshape(S, S), Ue).reshape(-1)
    assert rel(gx.get(), want_x) < TOL and rel(ge.get(), want_e) < TOL
    # K4 then K11
    fx, fy = DA(E * S * 4), DA(E * S * 4)
    G.compute_inviscid_flux_sp(dU.ptr, fx.ptr, fy.ptr, 1.4, E, S)
    ofx, ofy = o.k_flux_sp(U)
    assert rel(fx.get(), ofx) < 1e-15 and rel(fy.get(), ofy) < 1e-15
    div = DA(E * S * 4)
    G.compute_flux_divergence(fx.ptr, fy.ptr, div.ptr, dptr(ops["diff_xi"]), dptr(ops["diff_eta"]),
                              dptr(oc.Jinv), E, S)
    assert rel(div.get(), o.k_divergence()) < TOL


def test_operator_products_odd_sizes():
    """Element counts that leave the last group of four partly empty, and one element."""
    for E_want in (1, 2, 3, 5, 6, 7):
        p = 3
        mesh = cases.uniform_quad_mesh(E_want, 1, 0.0, 1.0, 0.0, 0.3, jitter=0.0)
        oc = fr_case.build(mesh.xy, mesh.elem_nodes, p, periodic=None)
        o = oc.solver()
        pos = O.sp_coords(oc.elem_nodes, oc.xy, oc.ops).reshape(-1, 2)
        U = cases.perturb(np.tile(cases.prim_to_cons(np.array(1.0), np.array(0.3), np.array(0.1), np.array(1.0), 1.4),
                                  (pos.shape[0], 1)), 0.05, E_want).reshape(-1)
        E, S, Fp = oc.E, oc.S, p + 1
        for eng in (G.ENGINE_MFMA, G.ENGINE_VALU):
            G.set_operator_engine(eng)
            ufp = DA(E * 4 * Fp * 4 + 8)
            ufp.upload(np.full(ufp.n, 7.0))
            G.interpolate_to_flux_points(dptr(U), ufp.ptr, dptr(oc.ops["interp"]), E, S, 4, Fp)
            got = ufp.get()
            assert rel(got[:-8], o.k_interp(U)) < TOL
            assert np.all(got[-8:] == 7.0)  # nothing written past the output
        G.set_operator_engine(G.ENGINE_MFMA)


@pytest.mark.parametrize("p", [1, 3, 4])
@pytest.mark.parametrize("bc", [False, True], ids=["periodic", "boundary"])
def test_face_kernels(p, bc):
    """computeRiemannFluxWithBC, computeFluxDiffForFR, applyFRCorrection
    (kernels.cu:346-701, 948-1034) against oz_k_riemann/flux_diff/correction,
    chained as FRSolver chains them; the full residual equals the oracle's."""
    say(f"case: face kernels P{p} bc={bc}")
    oc, U = setup(p, n=6, bc=bc)
    o = oc.solver()
    E, S, Fp, F = oc.E, oc.S, p + 1, oc.faces.n
    f = oc.faces
    le, ll, re, rl = (DA(np.ascontiguousarray(a, np.int32)) for a in (f.le, f.ll, f.re, f.rl))
    nrm = DA(oc.normals)
    ufp_o = o.k_interp(U)
    ufp = DA(ufp_o)
    fc = DA(F * Fp * 4)
    G.compute_riemann_flux_with_bc(ufp.ptr, fc.ptr, le.ptr, ll.ptr, re.ptr, rl.ptr, dptr(oc.bc_type),
                                   dptr(oc.bc_data), dptr(oc.bc_face_map), nrm.ptr, 1.4, 2, F, Fp, 4)
    fc_o = o.k_riemann()
    assert rel(fc.get(), fc_o) < TOL
    fd = DA(E * 4 * Fp * 4)  # zeroed, as fr_solver.cpp:501 zeroes it
    G.compute_flux_diff_for_fr(ufp.ptr, fc.ptr, fd.ptr, le.ptr, ll.ptr, re.ptr, rl.ptr, nrm.ptr, 1.4, F, Fp, 4)
    fd_o = o.k_flux_diff()
    assert rel(fd.get(), fd_o) < TOL
    o.k_flux_sp(U)
    div_o = o.k_divergence()
    dudt = DA(div_o)
    G.apply_fr_correction(dudt.ptr, fd.ptr, dptr(oc.ops["corr"]), E, S, 4, Fp)
    assert rel(dudt.get(), o.k_correction(div_o)) < TOL
    assert rel(dudt.get(), o.residual(U)) < TOL


def test_face_helpers():
    """computeInteriorFluxAtFace, computeInviscidFluxAtFP, computeFluxDifference,
    scatterFluxToElements (kernels.cu:198-340, 875-934) against the golden-pinned
    point flux."""
    oc, U = setup(2, n=4, bc=True)
    o = oc.solver()
    E, Fp, F = oc.E, 3, oc.faces.n
    f = oc.faces
    ufp_o = o.k_interp(U).reshape(E, 4, Fp, 4)
    ufp = DA(ufp_o.reshape(-1))
    le, ll, re, rl = (DA(np.ascontiguousarray(a, np.int32)) for a in (f.le, f.ll, f.re, f.rl))
    fint = DA(F * Fp * 4)
    G.compute_interior_flux_at_face(ufp.ptr, fint.ptr, le.ptr, ll.ptr, dptr(oc.normals), 1.4, F, Fp, 4)
    want = np.zeros((F, Fp, 4))
    for k in range(F):
        for q in range(Fp):
            fx, fy = O.flux_xy(ufp_o[f.le[k], f.ll[k], q])
            want[k, q] = fx * oc.normals[k, 0] + fy * oc.normals[k, 1]
    assert rel(fint.get(), want.reshape(-1)) < 1e-15
    ffp = DA(E * 4 * Fp * 4)
    G.compute_inviscid_flux_at_fp(ufp.ptr, ffp.ptr, 1.4, E, 4, Fp)
    ref_n = [(0.0, -1.0), (1.0, 0.0), (0.0, 1.0), (-1.0, 0.0)]
    want = np.zeros((E, 4, Fp, 4))
    for e in range(E):
        for k in range(4):
            for q in range(Fp):
                fx, fy = O.flux_xy(ufp_o[e, k, q])
                want[e, k, q] = fx * ref_n[k][0] + fy * ref_n[k][1]
    assert rel(ffp.get(), want.reshape(-1)) < 1e-15
    fc_o = o.k_riemann()
    fd = DA(F * Fp * 4)
    G.compute_flux_difference(dptr(fc_o), fint.ptr, fd.ptr, F * Fp * 4)
    assert np.array_equal(fd.get(), fc_o - fint.get())
    sc = DA(E * 4 * Fp * 4)
    G.scatter_flux_to_elements(dptr(fc_o), sc.ptr, le.ptr, ll.ptr, re.ptr, rl.ptr, F, Fp, 4)
    want = np.zeros((E, 4, Fp, 4))
    fco = fc_o.reshape(F, Fp, 4)
    for k in range(F):
        want[f.le[k], f.ll[k]] = fco[k]
        if f.re[k] >= 0:
            want[f.re[k], f.rl[k]] = fco[k]
    assert np.array_equal(sc.get(), want.reshape(-1))


@pytest.mark.parametrize("kind", ["rusanov", "roe", "hllc"])
def test_riemann_face_indexed(kind):
    """computeRiemannFlux on face-indexed pairs: the oracle's RiemannSolvers,
    which the compiled reference's golden vectors pin."""
    rng = np.random.default_rng(8)
    n = 300
    UL = cases.perturb(np.tile([1.0, 0.4, -0.2, 2.6], (n, 1)), 0.3, 1).reshape(-1, 4)
    UR = cases.perturb(np.tile([0.9, -0.1, 0.3, 2.2], (n, 1)), 0.3, 2).reshape(-1, 4)
    th = rng.uniform(0, 2 * math.pi, n)
    nrm = np.stack([np.cos(th), np.sin(th)], 1)
    F = DA(n * 4)
    G.compute_riemann_flux(dptr(UL.reshape(-1)), dptr(UR.reshape(-1)), F.ptr, dptr(nrm.reshape(-1)), 1.4,
                           O.RIEMANN[kind], n, 1)
    assert rel(F.get(), O.riemann_batch(kind, UL, UR, nrm).reshape(-1)) < TOL


def test_gradients_viscous_flux():
    """transformGradients and computeViscousFluxSP (kernels.cu:110-151,
    720-807): mu = T^0.7/Re per point, through the oracle's viscous flux."""
    oc, U = setup(3, n=4)
    E, S = oc.E, oc.S
    rng = np.random.default_rng(5)
    gx = rng.standard_normal(E * S * 4)
    ge = rng.standard_normal(E * S * 4)
    dx, dy = DA(E * S * 4), DA(E * S * 4)
    G.transform_gradients(dptr(gx), dptr(ge), dx.ptr, dy.ptr, dptr(oc.Jinv), E, S)
    Jv = oc.Jinv.reshape(E * S, 4)
    wx = gx.reshape(-1, 4) * Jv[:, 0:1] + ge.reshape(-1, 4) * Jv[:, 2:3]
    wy = gx.reshape(-1, 4) * Jv[:, 1:2] + ge.reshape(-1, 4) * Jv[:, 3:4]
    assert rel(dx.get(), wx.reshape(-1)) < 1e-15 and rel(dy.get(), wy.reshape(-1)) < 1e-15
    Re, Pr, g = 1600.0, 0.72, 1.4
    fvx, fvy = DA(E * S * 4), DA(E * S * 4)
    G.compute_viscous_flux_sp(dptr(U), dx.ptr, dy.ptr, fvx.ptr, fvy.ptr, g, Re, Pr, E, S)
    Ux, Uy = dx.get().reshape(-1, 4), dy.get().reshape(-1, 4)
    want_x, want_y = np.zeros((E * S, 4)), np.zeros((E * S, 4))
    for i, q in enumerate(U.reshape(-1, 4)):
        T = O._pressure(q, g) / q[0]
        mu = math.pow(T, 0.7) / Re
        want_x[i], want_y[i] = O.visc_flux(q, Ux[i], Uy[i], mu, mu * g / ((g - 1.0) * Pr), g)
    assert rel(fvx.get(), want_x.reshape(-1)) < TOL and rel(fvy.get(), want_y.reshape(-1)) < TOL


def test_boundary_conditions_characteristic():
    """applyBoundaryConditions (kernels.cu:1040-1211) against the oracle's
    restatement, every BC type, normals and data per boundary face."""
    oc, U = setup(3, n=4, bc=True)
    o = oc.solver()
    Fp = 4
    ufp_o = o.k_interp(U)
    f = oc.faces
    bidx = np.nonzero(f.re < 0)[0]
    nb = bidx.size
    types = np.array([O.BC_WALL, O.BC_SLIPWALL, O.BC_SYMMETRY, O.BC_FARFIELD, O.BC_OUTFLOW, O.BC_INFLOW,
                      O.BC_INTERIOR, O.BC_PERIODIC] * nb, np.int32)[:nb]
    rng = np.random.default_rng(3)
    data = np.zeros((nb, 8))
    rho, u, v, p = fr_case.freestream()
    for b in range(nb):
        if types[b] == O.BC_FARFIELD:  # alternate inflow and outflow branches
            s = 1.0 if b % 2 else -1.0
            data[b, :4] = [rho, s * u, s * v, p]
        elif types[b] == O.BC_OUTFLOW:
            data[b, 0] = 0.97 * p
        elif types[b] == O.BC_INFLOW:
            data[b, :4] = [1.02 * p, 300.0, rng.uniform(-1, 1), rng.uniform(-1, 1)]
    nrm = np.repeat(oc.normals[bidx], Fp, axis=0).reshape(-1)
    d = DA(ufp_o)
    ei = DA(np.ascontiguousarray(f.le[bidx], np.int32))
    lf = DA(np.ascontiguousarray(f.ll[bidx], np.int32))
    G.apply_boundary_conditions(d.ptr, ei.ptr, lf.ptr, dptr(types), dptr(data.reshape(-1)), dptr(nrm), 1.4,
                                nb, Fp, 4)
    want = O.apply_boundary_conditions(ufp_o, f.le[bidx], f.ll[bidx], types, data, nrm, 1.4, Fp, 4)
    assert not np.array_equal(want, ufp_o)
    assert rel(d.get(), want) < TOL


def test_positivity_rk_and_elementwise():
    """enforcePositivity, rkStage1-3 (zj_rk3_combine), setInitialCondition,
    copySolution, scaleArray, axpy (kernels.cu:1236-1349, 1526-1582)."""
    rng = np.random.default_rng(11)
    n = 1000
    U = np.stack([rng.uniform(-0.5, 200, n), rng.uniform(-3000, 3000, n), rng.uniform(-3000, 3000, n),
                  rng.uniform(-1e5, 1e9, n)], 1).reshape(-1)
    d = DA(U)
    G.enforce_positivity(d.ptr, n)
    assert rel(d.get(), O.positivity(U)) < 1e-15
    ic = DA(n * 4)
    G.set_initial_condition(ic.ptr, 1.2, 0.3, -0.4, 2.5, 1.4, 10, n // 10)
    q = np.array([1.2, 1.2 * 0.3, -1.2 * 0.4, 2.5 / 0.4 + 0.5 * 1.2 * (0.09 + 0.16)])
    assert rel(ic.get(), np.tile(q, n)) < 1e-15
    x = rng.standard_normal(4 * n)
    y = rng.standard_normal(4 * n)
    dx, dy = DA(x), DA(y)
    cp = DA(4 * n)
    G.copy_solution(cp.ptr, dx.ptr, 4 * n)
    assert np.array_equal(cp.get(), x)
    G.scale_array(cp.ptr, dx.ptr, -2.5, 4 * n)
    assert np.array_equal(cp.get(), -2.5 * x)
    G.axpy(dy.ptr, dx.ptr, 0.75, 4 * n)
    assert rel(dy.get(), y + 0.75 * x) < 1e-15


@pytest.mark.parametrize("p", [1, 2, 3, 4, 5])
def test_time_step_and_norm(p):
    """computeTimeStep (its per-element tree keeps only some points when n_sp is
    not a power of two) and computeResidualNorm (kernels.cu:1355-1524)."""
    oc, U = setup(p, n=5)
    E, S = oc.E, oc.S
    h_min = np.sqrt(oc.area) * (1.0 + 0.1 * np.arange(E) / E)
    dt = G.compute_time_step(dptr(U), dptr(oc.J), dptr(h_min), 0.5, 1.4, E, S)
    want = O.dt_gpu(U, oc.J, h_min, 0.5, 1.4, S)
    assert abs(dt - want) <= 1e-15 * want
    o = oc.solver()
    R = o.residual(U)
    nrm = G.compute_residual_norm(dptr(R), E, S)
    assert abs(nrm - math.sqrt(math.fsum(R * R))) <= 1e-14 * nrm
    assert abs(nrm - o.residual_l2()) <= 1e-14 * nrm


def test_argument_errors():
    from zhijian_amd._lib import ZJError
    with pytest.raises(ZJError) as e:
        G.interpolate_to_flux_points(None, None, None, 4, 100, 4, 3)
    assert e.value.code == 1
    with pytest.raises(ZJError):
        G.apply_fr_correction(None, None, None, 3, 16, 3, 4)
    with pytest.raises(ZJError):
        G.set_operator_engine(7)
    G.interpolate_to_flux_points(None, None, None, 0, 16, 4, 4)  # nothing to do
