"""The adjoint configuration at its stated size: Taylor-Green vortex on 128x128
quadrilaterals, P3, Re 1600, Mach 0.1, objective = domain-integrated kinetic
energy after 256 SSP-RK3 steps, checkpoints every 16 steps kept on the GPU
(zhijian_amd.adjoint.gradient), the textbook scheme with its Navier-Stokes
terms (BR2, VCJH HU, Roe), FP64.

dJ/dU0 is checked on 32 seeded directions against central differences
(eps = 1e-6 relative) of the FP64 forward map; dJ/dMach through the
closed-form initial state and dJ/dRe against Richardson-extrapolated central
differences.  The forward map of these differences is the GPU viscous path,
which equals the textbook oracle with its viscous terms to 1e-10 over 30
steps (test_gpu_viscous.py); one of the directions is also differenced with
the oracle itself (oracle/zj_textbook.c) over the full 256 steps.  Bar: 1e-6
relative.  Runs on an MI355X.
"""
import os

import numpy as np
import pytest

from oracle import pyoracle as O
from tests import fr_case
from tests.test_gpu_parity import say
from zhijian_amd import adjoint, cases, solver

pytestmark = pytest.mark.gpu

N, P, RE, MACH, STEPS, EVERY, EPS = 128, 3, 1600.0, 0.1, 256, 16, 1e-6
ETA = 2.0  # BR2 penalty (the library's default)


def _solver(t, Re):
    return solver.Solver(t, compat="textbook", riemann="roe", corr="hu", positivity=False, viscous=True, Re=Re,
                         Pr=0.72, br2_eta=ETA)


@pytest.fixture(scope="module")
def c5():
    mesh = cases.tgv_mesh(N)
    t = solver.build_tables(mesh.xy, mesh.elem_nodes, P, periodic=(mesh.Lx, mesh.Ly))
    s, f = _solver(t, RE), _solver(t, RE)
    U0 = cases.tgv_ic(t.pos, mach=MACH).reshape(-1)
    s.set_state(U0)
    dt = s.dt_cfl(0.5)  # the CFL step of the initial state, then held fixed
    say(f"case: C5 gradient, TGV {N}x{N} P{P} Re {RE:g}, {STEPS} steps of dt {dt:.4e}, checkpoints every {EVERY}")
    J, g, _, dre = adjoint.gradient(s, t, U0, [dt] * STEPS, every=EVERY, want_dt=False, want_re=True)

    def run(sol, U):
        sol.set_state(U)
        for _ in range(STEPS):
            sol.rk3_step(dt)
        return adjoint.kinetic_energy(sol.get_state(), t)

    return dict(mesh=mesh, t=t, U0=U0, dt=dt, J=J, g=g, dre=dre, forward=lambda U: run(f, U), run=run)


def test_c5_objective_matches_forward(c5):
    """The gradient's J (summed on the GPU after the checkpointed forward
    pass) against the same 256 steps run one by one and J summed on the host:
    the trajectories are the same bits, the two sums differ in order only."""
    J, Jf = c5["J"], c5["forward"](c5["U0"])
    assert abs(J - Jf) <= 1e-13 * abs(Jf)


def test_c5_gradient_32_directions(c5):
    U0, g, fwd = c5["U0"], c5["g"], c5["forward"]
    rng = np.random.default_rng(1600)
    worst = 0.0
    for k in range(32):
        v = U0 * rng.standard_normal(U0.shape)
        fd = (fwd(U0 + EPS * v) - fwd(U0 - EPS * v)) / (2 * EPS)
        ad = float(g @ v)
        r = abs(ad - fd) / abs(fd)
        worst = max(worst, r)
        say(f"  direction {k:2d}: adjoint {ad:+.12e}  FD {fd:+.12e}  rel {r:.2e}")
    assert worst < 1e-6


def test_c5_gradient_mach(c5):
    """dJ/dMach = <dJ/dU0, dU0/dMach>: Mach enters through p0 = 1/(gamma M^2) of
    the initial state, so dU0/dMach = (0, 0, 0, -2 / (gamma (gamma-1) M^3)).
    The kinetic energy barely depends on the pressure over these 256 steps
    (dJ/dMach ~ 5e-4 against J ~ 10), so a difference quotient with
    eps = 1e-6 relative would be swamped by the round-off of J; the check
    uses Richardson's extrapolation of central differences at h = 2% and 1%
    of Mach (truncation O(h^4))."""
    t, g, fwd = c5["t"], c5["g"], c5["forward"]
    gam = 1.4
    dU = np.zeros((t.E * t.S, 4))
    dU[:, 3] = -2.0 / (gam * (gam - 1.0) * MACH ** 3)
    ad = float(g @ dU.reshape(-1))

    def central(h):
        return (fwd(cases.tgv_ic(t.pos, mach=MACH + h).reshape(-1)) -
                fwd(cases.tgv_ic(t.pos, mach=MACH - h).reshape(-1))) / (2 * h)
    d1, d2 = central(0.02 * MACH), central(0.01 * MACH)
    fd = (4.0 * d2 - d1) / 3.0
    r = abs(ad - fd) / abs(fd)
    say(f"  dJ/dMach: adjoint {ad:+.12e}  Richardson FD {fd:+.12e} (central {d2:+.12e})  rel {r:.2e}")
    assert r < 1e-6


def test_c5_gradient_re(c5):
    """dJ/dRe from the adjoint's accumulated <cotangent of F_v, F_v> against
    Richardson's extrapolation of central differences at h = 2% and 1% of Re."""
    t, U0, run = c5["t"], c5["U0"], c5["run"]

    def central(h):
        return (run(_solver(t, RE + h), U0) - run(_solver(t, RE - h), U0)) / (2 * h)
    d1, d2 = central(0.02 * RE), central(0.01 * RE)
    fd = (4.0 * d2 - d1) / 3.0
    ad = c5["dre"]
    r = abs(ad - fd) / abs(fd)
    say(f"  dJ/dRe: adjoint {ad:+.12e}  Richardson FD {fd:+.12e} (central {d2:+.12e})  rel {r:.2e}")
    assert r < 1e-6


def test_c5_gradient_against_oracle_forward(c5):
    """One direction differenced with the textbook oracle's own forward map."""
    mesh, U0, g, dt = c5["mesh"], c5["U0"], c5["g"], c5["dt"]
    O.set_threads(min(16, len(os.sched_getaffinity(0))))
    try:
        tb = fr_case.build(mesh.xy, mesh.elem_nodes, P, periodic=(mesh.Lx, mesh.Ly)).textbook("roe", "hu")
        tb.set_viscous(RE, eta=ETA)
        tt = c5["t"]

        def oracle_J(U):
            for _ in range(STEPS):
                U = tb.advance(U, dt, False)
            return adjoint.kinetic_energy(U, tt)
        v = U0 * np.random.default_rng(7).standard_normal(U0.shape)
        fd = (oracle_J(U0 + EPS * v) - oracle_J(U0 - EPS * v)) / (2 * EPS)
    finally:
        O.set_threads(1)
    ad = float(g @ v)
    r = abs(ad - fd) / abs(fd)
    say(f"  oracle direction: adjoint {ad:+.12e}  oracle FD {fd:+.12e}  rel {r:.2e}")
    assert r < 1e-6
