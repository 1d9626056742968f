"""Pins the oracle's point physics (reference euler.hpp, kernels.cu:453-668, 1236-1349).

Euler fluxes and wave speeds are compared bit for bit with the reference's
IdealGas; the inline Rusanov of kernels.cu against the reference's
RiemannSolvers::rusanov (same mathematics, different rounding) and against
consistency and conservation; ghost states, the limiter and the RK stages
against hand-derived values and closed forms.
"""
import math

import numpy as np
import pytest

from oracle import pyoracle as O


def test_euler_flux_bit_exact(golden_physics):
    g = float(golden_physics["gamma"])
    for i, U in enumerate(golden_physics["UL"]):
        fx, fy = O.flux_xy(U, g)
        assert np.array_equal(fx, golden_physics["fx"][i])
        assert np.array_equal(fy, golden_physics["fy"][i])


def test_inline_rusanov_vs_reference_rusanov(golden_physics):
    g = float(golden_physics["gamma"])
    for i in range(golden_physics["UL"].shape[0]):
        UL, UR = golden_physics["UL"][i], golden_physics["UR"][i]
        n = golden_physics["normals"][i]
        F = O.rusanov(UL, UR, n[0], n[1], g)
        ref = golden_physics["rusanov"][i]
        scale = np.abs(ref).max() + 1e-300
        assert np.abs(F - ref).max() <= 1e-13 * scale, i


def test_rusanov_consistency_and_conservation(golden_physics):
    g = 1.4
    for i in range(golden_physics["UL"].shape[0]):
        U = golden_physics["UL"][i]
        V = golden_physics["UR"][i]
        nx, ny = golden_physics["normals"][i]
        fx, fy = O.flux_xy(U, g)
        F = O.rusanov(U, U, nx, ny, g)
        assert np.allclose(F, fx * nx + fy * ny, rtol=1e-14, atol=1e-14 * np.abs(fx).max())
        assert np.array_equal(O.rusanov(U, V, nx, ny, g), -O.rusanov(V, U, -nx, -ny, g))


def test_rusanov_scrubs_nan():
    U = np.array([1.0, 0.1, 0.2, 2.5])
    V = np.array([1.0, np.nan, 0.2, 2.5])
    F = O.rusanov(U, V, 1.0, 0.0)
    assert np.all(np.isfinite(F))
    assert np.all(F == 0.0)


def _cons(rho, u, v, p, g=1.4):
    return np.array([rho, rho * u, rho * v, p / (g - 1) + 0.5 * rho * (u * u + v * v)])


def test_ghost_states_hand_derived():
    g = 1.4
    Ui = _cons(1.3, 0.4, -0.2, 2.0)
    # slip wall with n=(1,0): normal velocity flips, everything else kept
    G = O.ghost_state(Ui, 1.0, 0.0, O.BC_SLIPWALL, None, 0)
    assert np.allclose(G, _cons(1.3, -0.4, -0.2, 2.0), rtol=1e-15)
    # wall and symmetry behave the same (kernels.cu:459-472)
    for t in (O.BC_WALL, O.BC_SYMMETRY):
        assert np.array_equal(O.ghost_state(Ui, 1.0, 0.0, t, None, 0), G)
    # outflow: interior rho, u, v with the prescribed pressure
    bd = np.array([[7.5, 0, 0, 0, 0, 0, 0, 0]], float)
    assert np.allclose(O.ghost_state(Ui, 0.0, 1.0, O.BC_OUTFLOW, bd, 0), _cons(1.3, 0.4, -0.2, 7.5))
    # far field, inflow side (vn_inf > 0): free-stream rho/u/v, averaged pressure
    # clamped into [1e3, 1e7]: (p_inf + p_i)/2 = (1.0 + 2.0)/2 -> 1000
    bd = np.array([[1.1, 0.5, 0.0, 1.0, 0, 0, 0, 0]], float)
    assert np.allclose(O.ghost_state(Ui, 1.0, 0.0, O.BC_FARFIELD, bd, 0), _cons(1.1, 0.5, 0.0, 1000.0))
    # outflow side: interior rho/u/v; density clamped to [0.01, 100]
    Ub = _cons(150.0, 0.4, -0.2, 5e3)
    G = O.ghost_state(Ub, -1.0, 0.0, O.BC_FARFIELD, np.array([[1.1, 0.5, 0.0, 3e3, 0, 0, 0, 0]]), 0)
    assert np.allclose(G, _cons(100.0, 0.4, -0.2, 4e3), rtol=1e-13)
    # other types copy the interior state
    assert np.array_equal(O.ghost_state(Ui, 1.0, 0.0, O.BC_INFLOW, None, 0), Ui)


def test_positivity_limiter_hand_derived():
    # inactive clamps still rewrite E through gamma_pp = 1.4 and P_MIN = 100
    U = _cons(1.0, 2.0, 0.0, 500.0)
    out = O.positivity(U)
    assert np.allclose(out, U, rtol=1e-15)
    # pressure below P_MIN is raised to 100
    out = O.positivity(_cons(1.0, 0.0, 0.0, 1.0))
    assert np.allclose(out, _cons(1.0, 0.0, 0.0, 100.0))
    # density clamp and velocity cap at 2000 (u computed from the clamped rho)
    q = np.array([0.001, 0.001 * 30.0, 0.0, 1e6])
    out = O.positivity(q)
    rho = 0.01
    u = 0.001 * 30.0 / 0.01
    ke = 0.5 * rho * u * u
    p = min(max(0.4 * (1e6 - ke), 100.0), 1e8)
    assert np.allclose(out, [rho, rho * u, 0.0, p / 0.4 + ke])
    q = np.array([1.0, 3000.0, 4000.0, 1e9])
    out = O.positivity(q)
    ke = 0.5 * (1200.0 ** 2 + 1600.0 ** 2)
    assert np.allclose(out, [1.0, 1200.0, 1600.0, min(0.4 * (1e9 - ke), 1e8) / 0.4 + ke])


def test_ssprk3_amplification_polynomial():
    """On R = lambda U the three stages give 1 + z + z^2/2 + z^3/6."""
    lam = np.array([-1.0, -0.3 + 2.0j, 0.5j, -2.5 + 0.1j])
    dt = 0.37
    for l in lam:
        U0 = np.array([1.0, 0.0])  # (Re, Im)
        def R(U):
            c = (U[0] + 1j * U[1]) * l
            return np.array([c.real, c.imag])
        U1 = O.rk_stage(1, U0, U0, R(U0), dt)
        U2 = O.rk_stage(2, U0, U1, R(U1), dt)
        U3 = O.rk_stage(3, U0, U2, R(U2), dt)
        z = l * dt
        amp = 1 + z + z * z / 2 + z ** 3 / 6
        assert abs(U3[0] + 1j * U3[1] - amp) < 1e-14


def test_wave_speed_matches_reference(golden_physics):
    """maxWaveSpeed used by the CFL dt (euler.hpp:166-171), through the oracle's
    dt on one-point elements: dt = CFL*sqrt(area)/(ws*(2p+1))."""
    from tests import fr_case
    from zhijian_amd import cases
    mesh = cases.uniform_quad_mesh(1, 1, 0.0, 2.0, 0.0, 2.0)
    case = fr_case.from_quadmesh(mesh, 1, periodic=True)
    s = case.solver()
    for i in range(8):
        U = np.tile(golden_physics["UL"][i], (case.E * case.S, 1))
        dt = s.dt_cfl(U, 0.5)
        expect = 0.5 * math.sqrt(4.0) / (golden_physics["ws"][i] * 3.0)
        assert dt == expect
