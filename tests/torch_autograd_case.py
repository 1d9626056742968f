"""The torch.autograd case of tests/test_gpu_autograd.py, run as its own
process: torch is imported and its GPU initialised before zhijian_amd loads
its library.  Exit 0 = pass, 5 = torch sees no GPU, anything else = failure."""
import os
import sys

import torch

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))

if not torch.cuda.is_available():
    print("torch.cuda.is_available() is False")
    sys.exit(5)
torch.cuda.init()

from oracle import pyoracle as O  # noqa: E402
from tests import fr_case  # noqa: E402
from zhijian_amd import adjoint, autograd as ZA, cases, solver  # noqa: E402


def main():
    mesh = cases.tgv_mesh(8)
    t = solver.build_tables(mesh.xy, mesh.elem_nodes, 3, periodic=(mesh.Lx, mesh.Ly))
    s = solver.Solver(t)
    o = fr_case.build(mesh.xy, mesh.elem_nodes, 3, periodic=(mesh.Lx, mesh.Ly)).solver()
    dev = torch.device("cuda:0")
    pos = torch.tensor(t.pos, dtype=torch.float64, device=dev)
    qw = torch.tensor(t.ops["qw"], dtype=torch.float64, device=dev)
    Jd = torch.tensor(t.J.reshape(t.E, t.S), dtype=torch.float64, device=dev)
    g, dt, M0 = 1.4, 1e-3, 0.05

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

    M = torch.tensor(M0, dtype=torch.float64, device=dev, requires_grad=True)
    U = ic(M)
    for _ in range(3):
        U = ZA.rk3_step(U, dt, s)
    R = ZA.residual(U, s)
    loss = objective(U) + 1e-6 * torch.sum(R * R)
    loss.backward()
    torch.cuda.synchronize()

    def loss_np(Mv):
        X = cases.tgv_ic(t.pos, mach=Mv).reshape(-1)
        for _ in range(3):
            X = o.advance(X, dt, True)
        Rr = o.residual(X)
        return adjoint.kinetic_energy(X, t) + 1e-6 * float(Rr @ Rr)

    l0 = loss_np(M0)
    print(f"loss torch {loss.item():.15e} oracle {l0:.15e}")
    assert abs(loss.item() - l0) <= 1e-11 * abs(l0)
    h = 1e-7
    fd = (loss_np(M0 + h) - loss_np(M0 - h)) / (2 * h)
    r = abs(M.grad.item() - fd) / abs(fd)
    print(f"dloss/dMach: torch+adjoint {M.grad.item():.12e}  FD {fd:.12e}  rel {r:.2e}")
    assert r < 1e-5


if __name__ == "__main__":
    main()
