"""Checkpointed reverse time stepping over the HIP discrete adjoint.

The objective is the kinetic energy of the final state,

    J(U) = sum_e sum_sp w_sp |J|_{e,sp} rho |u|^2 / 2,

with the reference's quadrature weights (FROperators quad_weights,
basis.cpp:484) and Jacobian determinants (FRSolver::initGeometry,
fr_solver.cpp:85-129).  The forward trajectory keeps one state every
`every` steps; the reverse sweep recomputes each segment from its checkpoint
and runs zj_rk3_step_vjp backwards through it.  Checkpoints and the
segment's states stay on the GPU (zj_state_slots / zj_state_save /
zj_state_load: no host copy per step), and the library holds the cotangent
on the device between steps; the host sees the final state (for J and its
gradient) and the result.
"""
from __future__ import annotations

from typing import List, Sequence, Tuple

import numpy as np

from .solver import Solver, Tables


def kinetic_energy(U: np.ndarray, t: Tables) -> float:
    """J of a state [E][S][4] (flattened)."""
    q = U.reshape(t.E, t.S, 4)
    w = t.ops["qw"][None, :] * t.J.reshape(t.E, t.S)
    ke = 0.5 * (q[..., 1] ** 2 + q[..., 2] ** 2) / q[..., 0]
    return float(np.sum(w * ke))


def kinetic_energy_grad(U: np.ndarray, t: Tables) -> np.ndarray:
    q = U.reshape(t.E, t.S, 4)
    w = t.ops["qw"][None, :] * t.J.reshape(t.E, t.S)
    u, v = q[..., 1] / q[..., 0], q[..., 2] / q[..., 0]
    g = np.zeros_like(q)
    g[..., 0] = -0.5 * w * (u * u + v * v)
    g[..., 1] = w * u
    g[..., 2] = w * v
    return g.reshape(-1)


def gradient(sol: Solver, t: Tables, U0: np.ndarray, dts: Sequence[float], every: int = 16, want_dt: bool = True,
             stream=None) -> Tuple[float, np.ndarray, List[float]]:
    """J(U^N) after len(dts) steps from U0, dJ/dU0 and dJ/d(dt_n) (None
    each with want_dt=False, which keeps the reverse sweep free of host
    synchronisation)."""
    steps = len(dts)
    n_ckpt = (steps + every - 1) // every
    # slots 0 .. n_ckpt-1: checkpoints; n_ckpt .. n_ckpt+every-2: the segment's later states
    sol.state_slots(n_ckpt + every - 1)
    sol.set_state(np.array(U0, np.float64).reshape(-1))
    for n in range(steps):
        if n % every == 0:
            sol.state_save(n // every, stream)
        sol.rk3_step(dts[n], stream)
    UN = sol.get_state()
    J = kinetic_energy(UN, t)
    sol.adjoint_set(kinetic_energy_grad(UN, t))
    dtb = [None] * steps
    for c in range(n_ckpt - 1, -1, -1):
        start = c * every
        end = min(start + every, steps)
        sol.state_load(c, stream)
        for n in range(start, end - 1):  # states start+1 .. end-1 of the segment
            sol.rk3_step(dts[n], stream)
            sol.state_save(n_ckpt + n - start, stream)
        for n in range(end - 1, start - 1, -1):
            sol.state_load(c if n == start else n_ckpt + n - start - 1, stream)
            dtb[n] = sol.rk3_step_vjp(dts[n], stream, want_dt=want_dt)
    g = sol.adjoint_get()
    sol.state_slots(0)
    return J, g, dtb
