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
    say("case: dJ/dU0 over 20 steps, checkpoints every 8, TGV 8x8 P3")
    mesh = cases.tgv_mesh(8)
    s, oc, t = make(mesh, 3)
    o = oc.solver()
    U0 = tgv_state(t, amp=0.02, seed=21, mach=0.3)
    dts = [4e-3] * 20
    J, g, dtb = adjoint.gradient(s, t, U0, dts, every=8)

    def Jof(X, dd=dts):
        for dt in dd:
            X = o.advance(X, dt, True)
        return adjoint.kinetic_energy(X, t)

    assert abs(J - Jof(U0)) <= 1e-9 * abs(J)
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
    torch = pytest.importorskip("torch")
    from zhijian_amd import autograd as ZA

    say("case: torch IC(Mach) -> 3 RK3Step -> J; dJ/dMach vs FD, TGV 8x8 P3")
    mesh = cases.tgv_mesh(8)
    s, oc, t = make(mesh, 3)
    o = oc.solver()
    dev = torch.device("cuda:0")
    pos = torch.tensor(t.pos, dtype=torch.float64, device=dev)
    qw = torch.tensor(t.ops["qw"], dtype=torch.float64, device=dev)
    Jd = torch.tensor(t.J.reshape(t.E, t.S), dtype=torch.float64, device=dev)
    g = 1.4
    dt = 4e-3

    def ic(M):
        x, y = pos[..., 0], pos[..., 1]
        u = torch.sin(x) * torch.cos(y)
        v = -torch.cos(x) * torch.sin(y)
        rho = torch.ones_like(x)
        p = 1.0 / (g * M * M) + 0.25 * (torch.cos(2 * x) + torch.cos(2 * y))
        return torch.stack([rho, rho * u, rho * v, p / (g - 1) + 0.5 * rho * (u * u + v * v)], -1)

    def objective(U):
        q = U.reshape(t.E, t.S, 4)
        return torch.sum(qw[None, :] * Jd * 0.5 * (q[..., 1] ** 2 + q[..., 2] ** 2) / q[..., 0])

    M = torch.tensor(0.3, dtype=torch.float64, device=dev, requires_grad=True)
    U = ic(M)
    for _ in range(3):
        U = ZA.rk3_step(U, dt, s)
    R = ZA.residual(U, s)
    loss = objective(U) + 1e-3 * torch.sum(R * R)
    loss.backward()
    torch.cuda.synchronize()

    def loss_np(Mv):
        X = cases.tgv_ic(t.pos, mach=Mv).reshape(-1)
        for _ in range(3):
            X = o.advance(X, dt, True)
        Rr = o.residual(X)
        return adjoint.kinetic_energy(X, t) + 1e-3 * float(Rr @ Rr)

    assert abs(loss.item() - loss_np(0.3)) <= 1e-11 * abs(loss.item())
    h = 1e-6
    fd = (loss_np(0.3 + h) - loss_np(0.3 - h)) / (2 * h)
    r = abs(M.grad.item() - fd) / abs(fd)
    say(f"  dloss/dMach: torch+adjoint {M.grad.item():.12e}  FD {fd:.12e}  rel {r:.2e}")
    assert r < 1e-5
