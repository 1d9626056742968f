"""Checkpointed reverse sweeps and the torch.autograd bindings (MI355X).

Pinned by the oracle's forward map: central finite differences (FP64,
eps = 1e-7 relative) of the kinetic-energy objective along seeded
directions, and of the Mach number of the initial state through a torch IC.
"""
import numpy as np
import pytest

from tests.test_gpu_parity import make, say, tgv_state
from zhijian_amd import adjoint, cases, solver

pytestmark = pytest.mark.gpu


def test_checkpointed_gradient_matches_fd():
    # Mach 0.05 keeps p above the limiter's 100 Pa floor (the limiter stays
    # inactive, so J is smooth), and the horizon is short against the growth
    # rate of the reference arithmetic.
    say("case: dJ/dU0 over 10 steps, checkpoints every 4, TGV 8x8 P3")
    mesh = cases.tgv_mesh(8)
    s, oc, t = make(mesh, 3)
    o = oc.solver()
    U0 = tgv_state(t, amp=0.02, seed=21, mach=0.05)
    dts = [1e-3] * 10
    J, g, dtb = adjoint.gradient(s, t, U0, dts, every=4)

    def Jof(X, dd=dts):
        for dt in dd:
            X = o.advance(X, dt, True)
        return adjoint.kinetic_energy(X, t)

    assert abs(J - Jof(U0)) <= 1e-12 * abs(J)
    v = U0 * np.random.default_rng(22).standard_normal(U0.shape)
    eps = 1e-7
    fd = (Jof(U0 + eps * v) - Jof(U0 - eps * v)) / (2 * eps)
    r = abs(fd - g @ v) / abs(fd)
    say(f"  <dJ/dU0, v>: adjoint {g @ v:.12e}  FD {fd:.12e}  rel {r:.2e}")
    assert r < 1e-6
    # dJ/d(dt_7)
    h = 1e-8
    dp, dm = list(dts), list(dts)
    dp[7] += h
    dm[7] -= h
    fdt = (Jof(U0, dp) - Jof(U0, dm)) / (2 * h)
    r = abs(fdt - dtb[7]) / abs(fdt)
    say(f"  dJ/d(dt_7): adjoint {dtb[7]:.12e}  FD {fdt:.12e}  rel {r:.2e}")
    assert r < 1e-5


def test_torch_autograd_mach_derivative():
    """torch IC(Mach) -> 3 RK3Step -> Residual -> loss; dloss/dMach against FD
    of the oracle.  Runs in a fresh interpreter that imports torch (and
    initialises its GPU) before the library loads, so both share one HIP
    runtime."""
    import os
    import subprocess
    import sys
    from tests.conftest import ROOT
    say("case: torch autograd through RK3Step/Residual, dloss/dMach vs FD (subprocess)")
    r = subprocess.run([sys.executable, "-u", os.path.join(ROOT, "tests", "torch_autograd_case.py")],
                       capture_output=True, text=True, timeout=240, cwd=ROOT)
    say(r.stdout[-2000:])
    if r.returncode == 5:
        pytest.skip("torch sees no GPU in this process: " + r.stdout.strip().splitlines()[-1])
    assert r.returncode == 0, r.stderr[-3000:]
