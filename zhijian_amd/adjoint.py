"""Checkpointed reverse time stepping over the HIP discrete adjoint.

The objective is the kinetic energy of the final state,

    J(U) = sum_e sum_sp w_sp |J|_{e,sp} rho |u|^2 / 2,

with the reference's quadrature weights (FROperators quad_weights,
basis.cpp:484) and Jacobian determinants (FRSolver::initGeometry,
fr_solver.cpp:85-129).  The forward trajectory keeps one state every
`every` steps; the reverse sweep recomputes each segment from its checkpoint
and runs zj_rk3_step_vjp backwards through it.  The library holds the
cotangent on the device between steps.
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


def gradient(sol: Solver, t: Tables, U0: np.ndarray, dts: Sequence[float], every: int = 16
             ) -> Tuple[float, np.ndarray, List[float]]:
    """J(U^N) after len(dts) steps from U0, dJ/dU0 and dJ/d(dt_n)."""
    steps = len(dts)
    checkpoints = {0: np.array(U0, np.float64).reshape(-1)}
    sol.set_state(checkpoints[0])
    for n in range(steps):
        sol.rk3_step(dts[n])
        if (n + 1) % every == 0 and n + 1 < steps:
            checkpoints[n + 1] = sol.get_state()
    UN = sol.get_state()
    J = kinetic_energy(UN, t)
    sol.adjoint_set(kinetic_energy_grad(UN, t))
    dtb = [0.0] * steps
    seg_starts = sorted(checkpoints, reverse=True)
    for start in seg_starts:
        end = min(start + every, steps)
        # states start .. end-1 of this segment
        traj = [checkpoints[start]]
        sol.set_state(traj[0])
        for n in range(start, end - 1):
            sol.rk3_step(dts[n])
            traj.append(sol.get_state())
        for n in range(end - 1, start - 1, -1):
            sol.set_state(traj[n - start])
            dtb[n] = sol.rk3_step_vjp(dts[n])
    return J, sol.adjoint_get(), dtb
