"""Pins of the oracle's restatements of the reference's free functions that
FRSolver's residual does not call (oracle/pyoracle.py: dt_tree_points,
dt_gpu, apply_boundary_conditions).  CPU only.

- computeTimeStepKernel's per-element tree (kernels.cu:1389-1397): the points
  it keeps, derived by hand for each order below.
- applyBoundaryConditionsKernel (kernels.cu:1040-1211): each ghost state
  satisfies the relations its branch is built on (reflection, Riemann
  invariants and entropy, prescribed pressure, isentropic total conditions),
  checked on the ghost state alone.
"""
import math

import numpy as np

from oracle import pyoracle as O

G = 1.4


def test_dt_tree_points_by_hand():
    # powers of two: every point
    assert O.dt_tree_points(4) == list(range(4))
    assert O.dt_tree_points(16) == list(range(16))
    # P2, 9 points: s=4 folds 4..7 into 0..3 and leaves 8; then 2, 1
    assert O.dt_tree_points(9) == list(range(8))
    # P4, 25 points: s=12 leaves 24; s=6, s=3; then s=1 folds only 1 into 0, so
    # slot 2 (holding 2, 5, 8, 11, 14, 17, 20, 23) never reaches slot 0
    lost = {2, 5, 8, 11, 14, 17, 20, 23, 24}
    assert O.dt_tree_points(25) == [i for i in range(25) if i not in lost]
    # P5, 36 points: s=18, s=9, then s=4 leaves slot 8 (8, 17, 26, 35)
    assert O.dt_tree_points(36) == [i for i in range(36) if i not in {8, 17, 26, 35}]


def test_dt_gpu_is_min_over_kept_points():
    rng = np.random.default_rng(4)
    S, E = 25, 6
    U = np.tile([1.0, 0.3, -0.2, 2.5], (E * S, 1))
    U[:, 0] *= rng.uniform(0.8, 1.2, E * S)
    J = rng.uniform(0.5, 2.0, E * S)
    h = rng.uniform(0.1, 0.2, E)
    # make a dropped point the smallest of all: the result must not see it
    U[3 * S + 24] = [1.0, 50.0, 0.0, 2000.0]
    got = O.dt_gpu(U, J, h, 0.5, G, S)
    vals = []
    for e in range(E):
        for sp in O.dt_tree_points(S):
            i = e * S + sp
            vals.append(0.5 * h[e] / (O.max_wave_speed(U[i], G) * math.sqrt(J[i])))
    assert got == min(vals)
    ws_drop = O.max_wave_speed(U[3 * S + 24], G)
    assert 0.5 * h[3] / (ws_drop * math.sqrt(J[3 * S + 24])) < got


def _prim(U):
    rho = U[0]
    u, v = U[1] / rho, U[2] / rho
    p = (G - 1.0) * (U[3] - 0.5 * rho * (u * u + v * v))
    return rho, u, v, p


def _ghost(U, n, bct, data):
    ufp = np.zeros((1, 4, 1, 4))
    ufp[0, 2, 0] = U
    out = O.apply_boundary_conditions(ufp, [0], [2], [bct], np.asarray(data, float).reshape(1, 8),
                                      np.asarray(n, float).reshape(1, 1, 2), G, 1, 4)
    return out.reshape(4, 4)[2]


def test_bc_reflection():
    n = np.array([0.6, -0.8])
    U = np.array([1.1, 0.5, 0.7, 3.0])
    for bct in (O.BC_WALL, O.BC_SLIPWALL, O.BC_SYMMETRY):
        r, u, v, p = _prim(U)
        rg, ug, vg, pg = _prim(_ghost(U, n, bct, np.zeros(8)))
        t = np.array([-n[1], n[0]])
        assert abs(rg - r) < 1e-15 and abs(pg - p) < 1e-12
        assert abs((ug * n[0] + vg * n[1]) + (u * n[0] + v * n[1])) < 1e-14
        assert abs((ug * t[0] + vg * t[1]) - (u * t[0] + v * t[1])) < 1e-14


def test_bc_farfield_riemann_invariants():
    n = np.array([0.8, 0.6])
    t = np.array([-n[1], n[0]])
    U = np.array([1.05, 0.3, 0.25, 2.6])
    for sgn in (1.0, -1.0):
        inf = np.array([1.0, sgn * 0.4, sgn * 0.1, 1.0])  # outflow and inflow directions
        r, u, v, p = _prim(U)
        c = math.sqrt(G * p / r)
        cinf = math.sqrt(G * inf[3] / inf[0])
        Rp = (u * n[0] + v * n[1]) + 2 * c / (G - 1)
        Rm = (inf[1] * n[0] + inf[2] * n[1]) - 2 * cinf / (G - 1)
        rg, ug, vg, pg = _prim(_ghost(U, n, O.BC_FARFIELD, list(inf) + [0] * 4))
        vn, cg = ug * n[0] + vg * n[1], math.sqrt(G * pg / rg)
        # the ghost carries the outgoing invariant of the interior and the incoming one of the free stream
        assert abs(vn + 2 * cg / (G - 1) - Rp) < 1e-12
        assert abs(vn - 2 * cg / (G - 1) - Rm) < 1e-12
        src_s, src_t = ((p / r ** G, u * t[0] + v * t[1]) if vn >= 0 else
                        (inf[3] / inf[0] ** G, inf[1] * t[0] + inf[2] * t[1]))
        assert abs(pg / rg ** G - src_s) < 1e-12 * src_s
        assert abs((ug * t[0] + vg * t[1]) - src_t) < 1e-12


def test_bc_outflow_and_inflow():
    n = np.array([1.0, 0.0])
    U = np.array([0.9, 0.2, -0.1, 2.2])
    r, u, v, p = _prim(U)
    rg, ug, vg, pg = _prim(_ghost(U, n, O.BC_OUTFLOW, [0.7] + [0] * 7))
    assert abs(rg - r) < 1e-15 and abs(ug - u) < 1e-15 and abs(vg - v) < 1e-15 and abs(pg - 0.7) < 1e-14
    pt, Tt, d = 1.3 * p, 3.1, np.array([3.0, -4.0])
    rg, ug, vg, pg = _prim(_ghost(U, n, O.BC_INFLOW, [pt, Tt, d[0], d[1], 0, 0, 0, 0]))
    assert abs(pg - p) < 1e-12  # static pressure from the interior
    vel = math.hypot(ug, vg)
    assert abs(ug / vel - 0.6) < 1e-12 and abs(vg / vel + 0.8) < 1e-12  # along the normalised direction
    M2 = vel * vel / (G * pg / rg)
    assert abs(pt / pg - (1 + 0.5 * (G - 1) * M2) ** (G / (G - 1))) < 1e-12
    assert abs(Tt - (pg / rg) * (1 + 0.5 * (G - 1) * M2)) < 1e-12
    # no direction given: against the outward normal
    rg, ug, vg, pg = _prim(_ghost(U, n, O.BC_INFLOW, [pt, Tt, 0, 0, 0, 0, 0, 0]))
    assert ug < 0 and abs(vg) < 1e-15


def test_bc_other_types_copy():
    U = np.array([0.9, 0.2, -0.1, 2.2])
    for bct in (O.BC_INTERIOR, O.BC_PERIODIC):
        assert np.array_equal(_ghost(U, [0.0, 1.0], bct, np.zeros(8)), U)
