"""Synthetic workloads: meshes and initial conditions of the benchmark cases.

The reference ships no initial conditions beyond a uniform far field
(src/main.cpp:173-184) and no mesh generator; these closed forms define the
workloads the solver is measured and checked on:

* isentropic vortex, periodic [-5,5]^2, Shu's vortex (strength 5, freestream
  rho=1, u=1, v=0, p=1);
* 2D Taylor-Green vortex, periodic [0,2pi]^2, rho=1, p = 1/(gamma M^2) +
  (cos 2x + cos 2y)/4;
* triangles: a periodic square of n x n cells, each split along its
  (i,j)->(i+1,j+1) diagonal, interior nodes jittered; a uniform flow at Mach M
  and angle of attack alpha carrying seeded Fourier velocity modes.

Meshes follow the node/element numbering of the reference's uniform test mesh
(test_meshes/uniform_quad_10x10.msh:258-357): nodes row-major with x fastest,
elements row-major, corners counter-clockwise (i,j), (i+1,j), (i+1,j+1), (i,j+1).
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np


@dataclass
class QuadMesh:
    xy: np.ndarray          # [n_nodes, 2]
    elem_nodes: np.ndarray  # [n_elem, 4] int64, CCW corners
    nx: int
    ny: int
    Lx: float               # periods (domain extents)
    Ly: float

    @property
    def n_elem(self) -> int:
        return int(self.elem_nodes.shape[0])


def uniform_quad_mesh(nx: int, ny: int, x0: float, x1: float, y0: float, y1: float,
                      jitter: float = 0.0, seed: int = 0) -> QuadMesh:
    """nx*ny quads on [x0,x1]x[y0,y1]; interior nodes optionally jittered by
    U(-jitter*h, jitter*h) per coordinate (seeded), keeping boundary nodes fixed
    so that periodic partners still match."""
    xs = np.linspace(x0, x1, nx + 1)
    ys = np.linspace(y0, y1, ny + 1)
    X, Y = np.meshgrid(xs, ys, indexing="xy")
    xy = np.stack([X.reshape(-1), Y.reshape(-1)], axis=1).astype(np.float64)
    if jitter > 0:
        rng = np.random.default_rng(seed)
        hx, hy = (x1 - x0) / nx, (y1 - y0) / ny
        d = rng.uniform(-jitter, jitter, size=xy.shape) * np.array([hx, hy])
        ii = np.tile(np.arange(nx + 1), ny + 1)
        jj = np.repeat(np.arange(ny + 1), nx + 1)
        interior = (ii > 0) & (ii < nx) & (jj > 0) & (jj < ny)
        xy[interior] += d[interior]
    j, i = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    n0 = (j * (nx + 1) + i).reshape(-1)
    en = np.stack([n0, n0 + 1, n0 + nx + 2, n0 + nx + 1], axis=1).astype(np.int64)
    return QuadMesh(xy, en, nx, ny, x1 - x0, y1 - y0)


def sheared_quad_mesh(nx: int, ny: int, Lx: float, Ly: float, shear: float, jitter: float = 0.0,
                      seed: int = 0) -> QuadMesh:
    """nx*ny parallelogram cells of the domain spanned by (Lx, 0) and
    (shear, Ly), node (i, j) at (i Lx/nx + j shear/ny, j Ly/ny), numbered as
    uniform_quad_mesh; interior nodes optionally jittered.  Periodic in both
    spanning directions: the top row is the bottom row shifted by (shear, Ly),
    so pairing by an axis period misses it and pairing by node links
    (periodic_links) is needed."""
    m = uniform_quad_mesh(nx, ny, 0.0, Lx, 0.0, Ly, jitter=jitter, seed=seed)
    xy = m.xy.copy()
    xy[:, 0] += shear * xy[:, 1] / Ly
    return QuadMesh(xy, m.elem_nodes, nx, ny, Lx, Ly)


def periodic_links(nx: int, ny: int) -> np.ndarray:
    """(slave, master) node links of an (nx+1) x (ny+1) lattice numbered
    i + j (nx+1), periodic in i and j: the last column onto the first and the
    last row onto the first, corners included, as a Gmsh $Periodic section
    lists them."""
    col = [((j * (nx + 1) + nx), j * (nx + 1)) for j in range(ny + 1)]
    row = [((ny * (nx + 1) + i), i) for i in range(nx + 1)]
    return np.array(col + row, np.int64)


def gauss_points(n: int) -> np.ndarray:
    """1D Gauss-Legendre points (numpy's table; used only to place ICs)."""
    return np.polynomial.legendre.leggauss(n)[0]


def sp_coords(mesh: QuadMesh, p: int, sol_xi=None, sol_eta=None) -> np.ndarray:
    """Physical coordinates [E, S, 2] of the solution points through the bilinear
    map N = (1-+xi)(1-+eta)/4 (reference basis.cpp:741-747, 882-893)."""
    if sol_xi is None:
        g = gauss_points(p + 1)
        sol_xi = np.tile(g, p + 1)
        sol_eta = np.repeat(g, p + 1)
    xi = np.asarray(sol_xi)[None, :]
    eta = np.asarray(sol_eta)[None, :]
    N = [0.25 * (1 - xi) * (1 - eta), 0.25 * (1 + xi) * (1 - eta),
         0.25 * (1 + xi) * (1 + eta), 0.25 * (1 - xi) * (1 + eta)]
    P = mesh.xy[mesh.elem_nodes]  # [E,4,2]
    x = sum(N[k] * P[:, k, 0:1] for k in range(4))
    y = sum(N[k] * P[:, k, 1:2] for k in range(4))
    return np.stack([x, y], axis=-1)


def prim_to_cons(rho, u, v, p, gamma):
    return np.stack([rho, rho * u, rho * v, p / (gamma - 1.0) + 0.5 * rho * (u * u + v * v)],
                    axis=-1)


def vortex_ic(pos: np.ndarray, gamma: float = 1.4, beta: float = 5.0,
              u_inf: float = 1.0, v_inf: float = 0.0) -> np.ndarray:
    """Shu's isentropic vortex at the origin on (rho, u, v, p) = (1, 1, 0, 1)."""
    x, y = pos[..., 0], pos[..., 1]
    r2 = x * x + y * y
    f = np.exp(0.5 * (1.0 - r2))
    du = -beta / (2 * math.pi) * y * f
    dv = beta / (2 * math.pi) * x * f
    T = 1.0 - (gamma - 1.0) * beta * beta / (8.0 * gamma * math.pi ** 2) * np.exp(1.0 - r2)
    rho = T ** (1.0 / (gamma - 1.0))
    p = rho ** gamma
    return prim_to_cons(rho, u_inf + du, v_inf + dv, p, gamma)


def tgv_ic(pos: np.ndarray, gamma: float = 1.4, mach: float = 0.1) -> np.ndarray:
    """2D Taylor-Green vortex, rho = 1, p0 = 1/(gamma M^2)."""
    x, y = pos[..., 0], pos[..., 1]
    u = np.sin(x) * np.cos(y)
    v = -np.cos(x) * np.sin(y)
    rho = np.ones_like(x)
    p = 1.0 / (gamma * mach * mach) + 0.25 * (np.cos(2 * x) + np.cos(2 * y))
    return prim_to_cons(rho, u, v, p, gamma)


def perturb(U: np.ndarray, amp: float, seed: int, gamma: float = 1.4) -> np.ndarray:
    """Seeded perturbation of the primitive state (relative N(0,1)*amp on rho
    and p, absolute on u, v) keeping rho and p positive."""
    rng = np.random.default_rng(seed)
    rho = U[..., 0]
    u = U[..., 1] / rho
    v = U[..., 2] / rho
    p = (gamma - 1.0) * (U[..., 3] - 0.5 * rho * (u * u + v * v))
    rho2 = rho * np.exp(amp * rng.standard_normal(rho.shape))
    p2 = p * np.exp(amp * rng.standard_normal(p.shape))
    u2 = u + amp * rng.standard_normal(u.shape)
    v2 = v + amp * rng.standard_normal(v.shape)
    return prim_to_cons(rho2, u2, v2, p2, gamma)


def tgv_mesh(n: int) -> QuadMesh:
    return uniform_quad_mesh(n, n, 0.0, 2 * math.pi, 0.0, 2 * math.pi)


def vortex_mesh(n: int = 10) -> QuadMesh:
    return uniform_quad_mesh(n, n, -5.0, 5.0, -5.0, 5.0)


@dataclass
class TriMesh:
    xy: np.ndarray          # [n_nodes, 2]
    tris: np.ndarray        # [n_elem, 3] int64, counter-clockwise corners
    n: int
    L: float                # period in x and y

    @property
    def n_elem(self) -> int:
        return int(self.tris.shape[0])


def split_tri_mesh(n: int, jitter: float = 0.0, seed: int = 0, L: float = 1.0, x0: float = 0.0) -> TriMesh:
    """n x n periodic square [x0, x0+L]^2; cell (i, j) splits into (n00, n10,
    n11) and (n00, n11, n01), cells row-major.  Interior nodes move by
    U(-jitter h, jitter h) per coordinate (PCG64 with `seed`, one draw pair per
    interior node in node order), boundary nodes stay so that periodic
    partners match."""
    h = L / n
    xs = np.linspace(0.0, L, n + 1)
    X, Y = np.meshgrid(xs, xs)
    xy = np.stack([X.reshape(-1), Y.reshape(-1)], axis=1)
    if jitter > 0:
        rng = np.random.default_rng(seed)
        x, y = X.reshape(-1), Y.reshape(-1)
        inner = (x > 0) & (x < L) & (y > 0) & (y < L)
        xy[inner] += rng.uniform(-jitter * h, jitter * h, size=(int(inner.sum()), 2))
    j, i = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    n00 = (j * (n + 1) + i).reshape(-1)
    n10, n01 = n00 + 1, n00 + n + 1
    n11 = n01 + 1
    tris = np.stack([np.stack([n00, n10, n11], 1), np.stack([n00, n11, n01], 1)], 1).reshape(-1, 3)
    return TriMesh(xy + x0, tris.astype(np.int64), n, L)


def fourier_freestream_ic(pos: np.ndarray, mach: float = 0.5, aoa_deg: float = 2.0, modes: int = 8,
                          amp: float = 0.05, seed: int = 0, L: float = 1.0, gamma: float = 1.4) -> np.ndarray:
    """rho = 1, |u_inf| = 1 at angle aoa, p = 1/(gamma M^2), plus `modes`
    velocity modes amp*(a, b) sin(2 pi (kx x + ky y)/L + phi) with integer
    wavenumbers kx, ky in [-4, 4] (not both 0), unit directions (a, b) and
    phases phi drawn from PCG64(seed) in that order per mode."""
    rng = np.random.default_rng(seed)
    x, y = pos[..., 0], pos[..., 1]
    a = math.radians(aoa_deg)
    u = np.full_like(x, math.cos(a))
    v = np.full_like(x, math.sin(a))
    for _ in range(modes):
        kx, ky = 0, 0
        while kx == 0 and ky == 0:
            kx, ky = (int(k) for k in rng.integers(-4, 5, size=2))
        th = rng.uniform(0.0, 2 * math.pi)
        phi = rng.uniform(0.0, 2 * math.pi)
        w = np.sin(2 * math.pi * (kx * x + ky * y) / L + phi)
        u = u + amp * math.cos(th) * w
        v = v + amp * math.sin(th) * w
    rho = np.ones_like(x)
    p = np.full_like(x, 1.0 / (gamma * mach * mach))
    return prim_to_cons(rho, u, v, p, gamma)
