"""Checkpointed reverse time stepping over the HIP discrete adjoint.

The objective is the kinetic energy of the final state,

    J(U) = sum_e sum_sp w_sp |J|_{e,sp} rho |u|^2 / 2,

with the reference's quadrature weights (FROperators quad_weights,
basis.cpp:484) and Jacobian determinants (FRSolver::initGeometry,
fr_solver.cpp:85-129).  The forward trajectory keeps one state every
`every` steps (zj_forward_checkpointed) and records its last segments on a
device tape as it steps; the reverse sweep (zj_adjoint_sweep) recomputes
every other segment from its checkpoint, recording the stage states and
traces of its steps on the tape, and runs the stage VJPs backwards from the
tape.  Both passes and the objective run in the library, on the GPU:
the host sees J and the resulting gradient.
"""
from __future__ import annotations

from typing import Sequence

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


def _device_weights(sol: Solver, t: Tables):
    """The objective's weights (quadrature weight x |J| per point, [E][S]) on
    the GPU, uploaded once per solver and tables."""
    from .hiprt import DeviceArray
    cached = getattr(sol, "_ke_weights", None)
    if cached is None or cached[0] is not t:
        w = np.ascontiguousarray((t.ops["qw"][None, :] * t.J.reshape(t.E, t.S)).reshape(-1), np.float64)
        cached = (t, DeviceArray(w))
        sol._ke_weights = cached
    return cached[1]


def gradient(sol: Solver, t: Tables, U0: np.ndarray, dts: Sequence[float], every: int = 16, want_dt: bool = True,
             stream=None, want_re: bool = False, tape_segments: int | None = None):
    """J(U^N) after len(dts) steps from U0, dJ/dU0 and dJ/d(dt_n) (None
    each with want_dt=False, which keeps the reverse sweep free of host
    synchronisation); with want_re, also dJ/dRe (Navier-Stokes terms) as a
    fourth value.  tape_segments: trailing segments the forward pass records
    on the device tape (None: the library's setting, 1 by default); with all
    of them recorded, no step is computed twice."""
    steps = len(dts)
    if tape_segments is not None:
        sol.adjoint_tape_segments(tape_segments)
    sol.set_state(np.array(U0, np.float64).reshape(-1))
    sol.forward_checkpointed(dts, every, stream)
    # J and its gradient on the GPU: the final state never leaves it
    J = sol.adjoint_seed_kinetic_energy(_device_weights(sol, t).ptr, stream)
    dtb = sol.adjoint_sweep(dts, every, want_dt, stream)
    if dtb is None:
        dtb = [None] * steps
    g = sol.adjoint_get()
    # the checkpoint slots stay allocated for the next gradient, as the tape does
    # (sol.state_slots(0) releases them)
    if want_re:
        return J, g, dtb, sol.adjoint_get_re()
    return J, g, dtb
