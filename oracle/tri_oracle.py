"""CPU oracle of the triangle path: operators and the textbook FR residual.

TEST INFRASTRUCTURE ONLY (imported by tests/ and the cpu_baseline leg of
bench.py, never by the product package).

The reference builds triangle operators as zero placeholders
(src/basis/basis.cpp:615-669) and its solver assumes quadrilaterals
(src/solver/fr_solver.cpp:138-140), so there is no reference semantics to
follow.  This restates the textbook scheme for straight-sided triangles:

* reference triangle (-1,-1), (1,-1), (-1,1); edges e0 = v0->v1, e1 = v1->v2,
  e2 = v2->v0 (counter-clockwise);
* solution points: the warp-and-blend nodes of Hesthaven & Warburton
  ("Nodal Discontinuous Galerkin Methods", 2008, section 6.1) with their
  optimised alpha, numbered row by row from v0;
* flux points: p+1 Gauss-Legendre points per edge, ordered along the edge's
  counter-clockwise direction (neighbours meet at slot p-f);
* orthonormal Dubiner basis (ibid., Simplex2DP / GradSimplex2DP);
* residual (DG-equivalent FR, c = 0; other correction fields below):
      dU/dt = -div F^D - sum_j LIFT[:, j] (|e_j| / (2 J)) (F^I - F^D) . n_out
  with F^D the polynomial interpolant of the flux, F^I the Riemann flux
  (euler.hpp:184-375 via the C oracle) once per face point, and
  LIFT[sp, j] = w_j sum_m psi_m(x_sp) psi_m(x_j), the L2 lifting of a unit
  value at flux point j.

Correction fields other than DG (esfr_lift): the energy-stable family on
triangles of Castonguay, Vincent & Jameson (J. Sci. Comput. 51, 2012), norm
int_T u^2 + (c/2) sum_v C(p,v) (d^p u / dr^v ds^(p-v))^2, whose correction
field is the DG lifting filtered by (M + K)^-1 M.  Restated nodally: M the
nodal mass matrix (V V^T)^-1, K = c sum_v C(p,v) d_v d_v^T with d_v the
p-th derivative of each nodal basis function, taken as a row of
Dr^v Ds^(p-v) (constant over the element; the reference triangle's area 2
times the c/2).  c is the one-dimensional VCJH value of the family (Vincent,
Castonguay & Jameson, J. Sci. Comput. 47, 2011), SD = GA.  Pinned by c = 0
giving the DG lifting, conservation (the volume weights of the lifting do
not depend on c), K annihilating degree < p, linear stability and the
library's independent modal construction.

Boundary faces take the ghost state of the quadrilateral path
(kernels.cu:453-525 via the C oracle) at each flux point.

Pins (tests/test_tri_oracle.py): polynomial exactness of interpolation and
differentiation; the lifting identity against an exact triangle quadrature;
freestream on jittered meshes (periodic, and open with every BC type whose
ghost state preserves a uniform flow); conservation; linear stability; order
of accuracy on the isentropic vortex.
"""
from __future__ import annotations

import math

import numpy as np

from . import pyoracle as O


# ------------------------------------------------------------ 1D helpers
def jacobi_p(x, alpha, beta, n):
    """Orthonormal Jacobi polynomial P_n^(alpha,beta) at x (Hesthaven JacobiP)."""
    x = np.asarray(x, np.float64)
    PL = np.zeros((n + 1,) + x.shape)
    g0 = 2 ** (alpha + beta + 1) / (alpha + beta + 1) * math.gamma(alpha + 1) * math.gamma(beta + 1) / \
        math.gamma(alpha + beta + 1)
    PL[0] = 1.0 / math.sqrt(g0)
    if n == 0:
        return PL[0]
    g1 = (alpha + 1) * (beta + 1) / (alpha + beta + 3) * g0
    PL[1] = ((alpha + beta + 2) * x / 2 + (alpha - beta) / 2) / math.sqrt(g1)
    if n == 1:
        return PL[1]
    aold = 2 / (2 + alpha + beta) * math.sqrt((alpha + 1) * (beta + 1) / (alpha + beta + 3))
    for i in range(1, n):
        h1 = 2 * i + alpha + beta
        anew = 2 / (h1 + 2) * math.sqrt((i + 1) * (i + 1 + alpha + beta) * (i + 1 + alpha) * (i + 1 + beta) /
                                        (h1 + 1) / (h1 + 3))
        bnew = -(alpha ** 2 - beta ** 2) / h1 / (h1 + 2)
        PL[i + 1] = 1 / anew * (-aold * PL[i - 1] + (x - bnew) * PL[i])
        aold = anew
    return PL[n]


def grad_jacobi_p(x, alpha, beta, n):
    if n == 0:
        return np.zeros_like(np.asarray(x, np.float64))
    return math.sqrt(n * (n + alpha + beta + 1)) * jacobi_p(x, alpha + 1, beta + 1, n - 1)


def jacobi_gl(n):
    """Gauss-Lobatto nodes of order n on [-1, 1] (alpha = beta = 0)."""
    if n == 1:
        return np.array([-1.0, 1.0])
    inner = np.polynomial.legendre.Legendre.basis(n).deriv().roots()
    return np.concatenate([[-1.0], np.sort(inner.real), [1.0]])


# ---------------------------------------------------------- triangle nodes
_ALPOPT = [0.0, 0.0, 1.4152, 0.1001, 0.2751, 0.9800, 1.0999, 1.2832, 1.3648, 1.4773]


def _warpfactor(n, rout):
    lgl = jacobi_gl(n)
    req = np.linspace(-1, 1, n + 1)
    veq = np.stack([jacobi_p(req, 0, 0, i) for i in range(n + 1)], axis=1)
    pmat = np.stack([jacobi_p(rout, 0, 0, i) for i in range(n + 1)], axis=0)
    lmat = np.linalg.solve(veq.T, pmat)
    warp = lmat.T @ (lgl - req)
    zerof = np.abs(rout) < 1.0 - 1e-10
    sf = 1.0 - (zerof * rout) ** 2
    return warp / sf + warp * (zerof - 1)


def solution_points(p):
    """Warp-and-blend nodes in (r, s), numbered row by row from v0."""
    alpha = _ALPOPT[p] if p < len(_ALPOPT) else 5.0 / 3.0
    L1, L3 = [], []
    for n in range(p + 1):
        for m in range(p + 1 - n):
            L1.append(n / p)
            L3.append(m / p)
    L1, L3 = np.array(L1), np.array(L3)
    L2 = 1.0 - L1 - L3
    x = -L2 + L3
    y = (-L2 - L3 + 2 * L1) / math.sqrt(3.0)
    b1, b2, b3 = 4 * L2 * L3, 4 * L1 * L3, 4 * L1 * L2
    w1 = b1 * _warpfactor(p, L3 - L2) * (1 + (alpha * L1) ** 2)
    w2 = b2 * _warpfactor(p, L1 - L3) * (1 + (alpha * L2) ** 2)
    w3 = b3 * _warpfactor(p, L2 - L1) * (1 + (alpha * L3) ** 2)
    x = x + w1 + math.cos(2 * math.pi / 3) * w2 + math.cos(4 * math.pi / 3) * w3
    y = y + math.sin(2 * math.pi / 3) * w2 + math.sin(4 * math.pi / 3) * w3
    # equilateral -> reference triangle
    l1 = (math.sqrt(3.0) * y + 1.0) / 3.0
    l2 = (-3.0 * x - math.sqrt(3.0) * y + 2.0) / 6.0
    l3 = (3.0 * x - math.sqrt(3.0) * y + 2.0) / 6.0
    r = -l2 + l3 - l1
    s = -l2 - l3 + l1
    return np.stack([r, s], axis=1)


EDGE_VERTS = np.array([[-1.0, -1.0], [1.0, -1.0], [-1.0, 1.0]])


def flux_points(p):
    """[3][p+1][2]: Gauss points on edge e from vertex e to vertex e+1, and
    the 1D Gauss weights."""
    xg, wg = O.gauss(p + 1)
    pts = np.zeros((3, p + 1, 2))
    for e in range(3):
        a, b = EDGE_VERTS[e], EDGE_VERTS[(e + 1) % 3]
        for f in range(p + 1):
            t = 0.5 * (1 + xg[f])
            pts[e, f] = a + t * (b - a)
    return pts, np.asarray(wg)


# ---------------------------------------------------------- Dubiner basis
def _rs_to_ab(r, s):
    a = np.where(np.abs(1 - s) > 1e-14, 2 * (1 + r) / np.where(np.abs(1 - s) > 1e-14, 1 - s, 1.0) - 1, -1.0)
    return a, s


def dubiner(p, rs):
    """[n_pts][Np] orthonormal basis values and d/dr, d/ds."""
    r, s = rs[:, 0], rs[:, 1]
    a, b = _rs_to_ab(r, s)
    V, Vr, Vs = [], [], []
    for i in range(p + 1):
        for j in range(p + 1 - i):
            h1, h2 = jacobi_p(a, 0, 0, i), jacobi_p(b, 2 * i + 1, 0, j)
            V.append(math.sqrt(2.0) * h1 * h2 * (1 - b) ** i)
            dh1, dh2 = grad_jacobi_p(a, 0, 0, i), grad_jacobi_p(b, 2 * i + 1, 0, j)
            dmodr = dh1 * h2
            if i > 0:
                dmodr = dmodr * (0.5 * (1 - b)) ** (i - 1)
            dmods = dh1 * (h2 * (0.5 * (1 + a)))
            if i > 0:
                dmods = dmods * (0.5 * (1 - b)) ** (i - 1)
            tmp = dh2 * (0.5 * (1 - b)) ** i
            if i > 0:
                tmp = tmp - 0.5 * i * h2 * (0.5 * (1 - b)) ** (i - 1)
            dmods = dmods + h1 * tmp
            Vr.append(dmodr * 2 ** (i + 0.5))
            Vs.append(dmods * 2 ** (i + 0.5))
    return np.stack(V, 1), np.stack(Vr, 1), np.stack(Vs, 1)


def operators(p):
    """Solution points, flux points, Dr, Ds [Np][Np], interpolation to flux
    points [3(p+1)][Np], lifting [Np][3(p+1)] (without the edge metric)."""
    sp = solution_points(p)
    fp, wg = flux_points(p)
    V, Vr, Vs = dubiner(p, sp)
    Vi = np.linalg.inv(V)
    Dr, Ds = Vr @ Vi, Vs @ Vi
    Vf, _, _ = dubiner(p, fp.reshape(-1, 2))
    I = Vf @ Vi
    w = np.tile(wg, 3)
    LIFT = (V @ Vf.T) * w[None, :]
    return dict(sp=sp, fp=fp, wg=np.asarray(wg), Dr=Dr, Ds=Ds, interp=I, lift=LIFT, V=V)


def vcjh_c(p, family):
    """One-dimensional VCJH c: DG 0, SD = GA 2p/((2p+1)(p+1)(a_p p!)^2),
    HU 2(p+1)/((2p+1) p (a_p p!)^2), a_p = C(2p, p) / 2^p."""
    ap_pf = math.comb(2 * p, p) / 2 ** p * math.factorial(p)
    if family == "dg":
        return 0.0
    if family in ("sd", "ga"):
        return 2.0 * p / ((2 * p + 1) * (p + 1) * ap_pf ** 2)
    if family == "hu":
        return 2.0 * (p + 1) / ((2 * p + 1) * p * ap_pf ** 2)
    raise ValueError(family)


def esfr_lift(p, family, ops=None):
    """Lifting [Np][3(p+1)] of the family's energy-stable correction field."""
    ops = operators(p) if ops is None else ops
    c = vcjh_c(p, family)
    if c == 0.0:
        return ops["lift"].copy()
    V, Dr, Ds = ops["V"], ops["Dr"], ops["Ds"]
    M = np.linalg.inv(V @ V.T)
    K = np.zeros_like(M)
    for v in range(p + 1):
        D = np.linalg.matrix_power(Dr, v) @ np.linalg.matrix_power(Ds, p - v)
        d = D.mean(axis=0)             # every row holds the same constant derivatives
        K += c * math.comb(p, v) * np.outer(d, d)
    return np.linalg.solve(M + K, M @ ops["lift"])


# ------------------------------------------------------------ meshes
def split_quad_mesh(n, jitter=0.0, seed=0, L=1.0):
    """n x n periodic square of side L, each cell split along (i,j)->(i+1,j+1)
    into (n00, n10, n11) and (n00, n11, n01); interior nodes jittered by
    U(-jitter h, jitter h), boundary nodes fixed."""
    h = L / n
    xs = np.linspace(0.0, L, n + 1)
    X, Y = np.meshgrid(xs, xs)
    xy = np.stack([X.reshape(-1), Y.reshape(-1)], axis=1)
    if jitter > 0:
        rng = np.random.default_rng(seed)
        inner = (X.reshape(-1) > 0) & (X.reshape(-1) < L) & (Y.reshape(-1) > 0) & (Y.reshape(-1) < L)
        xy[inner] += rng.uniform(-jitter * h, jitter * h, size=(int(inner.sum()), 2))
    idx = lambda i, j: j * (n + 1) + i
    tris = []
    for j in range(n):
        for i in range(n):
            tris.append([idx(i, j), idx(i + 1, j), idx(i + 1, j + 1)])
            tris.append([idx(i, j), idx(i + 1, j + 1), idx(i, j + 1)])
    return xy, np.array(tris, np.int64)


def geometry(xy, tris):
    """Per element: (r_x, r_y, s_x, s_y, J), edge lengths [E][3], centroids."""
    a, b, c = xy[tris[:, 0]], xy[tris[:, 1]], xy[tris[:, 2]]
    # x = a + (b-a)(r+1)/2 + (c-a)(s+1)/2
    xr, yr = 0.5 * (b - a)[:, 0], 0.5 * (b - a)[:, 1]
    xs, ys = 0.5 * (c - a)[:, 0], 0.5 * (c - a)[:, 1]
    J = xr * ys - xs * yr
    met = np.stack([ys / J, -xs / J, -yr / J, xr / J, J], axis=1)
    el = np.stack([np.linalg.norm(b - a, axis=1), np.linalg.norm(c - b, axis=1), np.linalg.norm(a - c, axis=1)], 1)
    return met, el


def sp_coords(xy, tris, ops):
    a, b, c = xy[tris[:, 0]], xy[tris[:, 1]], xy[tris[:, 2]]
    r, s = ops["sp"][:, 0], ops["sp"][:, 1]
    return (a[:, None, :] + 0.5 * (b - a)[:, None, :] * (r[None, :, None] + 1)
            + 0.5 * (c - a)[:, None, :] * (s[None, :, None] + 1))


class TriTextbook:
    """Textbook FR residual / SSP-RK3 on straight-sided triangles."""

    def __init__(self, p, xy, tris, faces, normals, riemann="roe", gamma=1.4, bc_type=None, bc_face_map=None,
                 bc_data=None, corr="dg"):
        self.p, self.g = p, gamma
        self.ops = operators(p)
        self.ops["lift"] = esfr_lift(p, corr, self.ops)
        self.Np = self.ops["sp"].shape[0]
        self.Fp = p + 1
        self.E = tris.shape[0]
        self.met, self.elen = geometry(xy, tris)
        self.faces = np.asarray(faces)
        self.normals = np.asarray(normals)
        self.kind = O.RIEMANN[riemann]
        self.bc_type = None if bc_type is None else np.asarray(bc_type)
        self.bc_face_map = None if bc_face_map is None else np.asarray(bc_face_map)
        self.bc_data = None if bc_data is None else np.asarray(bc_data, np.float64).reshape(-1)
        self.viscous = False

    def set_viscous(self, Re, Pr=0.72, eta=2.0):
        """Navier-Stokes terms (BR2) with mu = 1/Re, kappa = mu gamma / ((gamma-1) Pr)
        and BR2 penalty eta; periodic meshes only."""
        if (self.faces[:, 2] < 0).any():
            raise ValueError("the triangle viscous terms are built for meshes without boundary faces")
        self.mu = 1.0 / Re
        self.kappa = self.mu * self.g / ((self.g - 1.0) * Pr)
        self.eta = eta
        self.viscous = True
        return self

    def _visc_flux(self, U, Qx, Qy):
        """ViscousFlux::fluxX/fluxY (navierstokes.hpp:53-99) with mu, kappa as
        parameters, in the steps of oz_visc_flux, over arrays [..., 4]."""
        g, mu, kappa = self.g, self.mu, self.kappa
        rho = U[..., 0]
        u, v = U[..., 1] / rho, U[..., 2] / rho
        dudx, dudy = (Qx[..., 1] - u * Qx[..., 0]) / rho, (Qy[..., 1] - u * Qy[..., 0]) / rho
        dvdx, dvdy = (Qx[..., 2] - v * Qx[..., 0]) / rho, (Qy[..., 2] - v * Qy[..., 0]) / rho
        p = (g - 1.0) * rho * (U[..., 3] / rho - 0.5 * (u * u + v * v))
        dpdx = (g - 1.0) * (Qx[..., 3] - 0.5 * Qx[..., 0] * (u * u + v * v) - rho * (u * dudx + v * dvdx))
        dpdy = (g - 1.0) * (Qy[..., 3] - 0.5 * Qy[..., 0] * (u * u + v * v) - rho * (u * dudy + v * dvdy))
        T = p / rho
        dTdx, dTdy = (dpdx - T * Qx[..., 0]) / rho, (dpdy - T * Qy[..., 0]) / rho
        div, lam = dudx + dvdy, -2.0 / 3.0 * mu
        txx, tyy, txy = 2.0 * mu * dudx + lam * div, 2.0 * mu * dvdy + lam * div, mu * (dudy + dvdx)
        z = np.zeros_like(rho)
        Fv = np.stack([z, txx, txy, u * txx + v * txy + kappa * dTdx], -1)
        Gv = np.stack([z, txy, tyy, u * txy + v * tyy + kappa * dTdy], -1)
        return Fv, Gv

    def gradient(self, U):
        """Corrected gradient Q at the solution points ([E][Np][4] each of x, y):
        grad U plus the lifting of the jumps U^C - U_own, U^C = (U_L + U_R) / 2,
        along the outward normals (DG-equivalent FR applied to grad u)."""
        return self._viscous_parts(U.reshape(self.E, self.Np, 4))[:2]

    def _viscous_parts(self, U):
        E, Fp, p = self.E, self.Fp, self.p
        rx, ry, sx, sy, J = (self.met[:, k] for k in range(5))
        Dr, Ds, I, LIFT = self.ops["Dr"], self.ops["Ds"], self.ops["interp"], self.ops["lift"]
        Ur, Us = np.einsum("ij,ejv->eiv", Dr, U), np.einsum("ij,ejv->eiv", Ds, U)
        gx = Ur * rx[:, None, None] + Us * sx[:, None, None]
        gy = Ur * ry[:, None, None] + Us * sy[:, None, None]
        Uf = np.einsum("ij,ejv->eiv", I, U).reshape(E, 3, Fp, 4)
        f = self.faces
        le, ll, re, rl = f[:, 0], f[:, 1], f[:, 2], f[:, 3]
        fr = np.arange(Fp)
        UL = Uf[le[:, None], ll[:, None], fr[None, :]]
        UR = Uf[re[:, None], rl[:, None], (p - fr)[None, :]]
        UC = 0.5 * (UL + UR)
        jump = np.zeros((E, 3, Fp, 4))
        jump[le[:, None], ll[:, None], fr[None, :]] = UC - UL
        jump[re[:, None], rl[:, None], (p - fr)[None, :]] = UC - UR
        nrm_out = np.zeros((E, 3, 2))
        nrm_out[le, ll] = self.normals
        nrm_out[re, rl] = -self.normals
        sc = self.elen / (2 * J[:, None])                      # |e| / (2J) per element edge
        wx = (jump * (sc * nrm_out[:, :, 0])[:, :, None, None]).reshape(E, 3 * Fp, 4)
        wy = (jump * (sc * nrm_out[:, :, 1])[:, :, None, None]).reshape(E, 3 * Fp, 4)
        Qx = gx + np.einsum("ij,ejv->eiv", LIFT, wx)
        Qy = gy + np.einsum("ij,ejv->eiv", LIFT, wy)
        # BR2: each side's gradient at a face point is its own trace of grad U
        # plus eta times the lifting of this face's jumps alone, evaluated at
        # the face's points (LF = interp . lift restricted to the face)
        gxf = np.einsum("ij,ejv->eiv", I, gx).reshape(E, 3, Fp, 4)
        gyf = np.einsum("ij,ejv->eiv", I, gy).reshape(E, 3, Fp, 4)
        rxf, ryf = np.zeros_like(gxf), np.zeros_like(gyf)
        for k in range(3):
            LF = I[k * Fp:(k + 1) * Fp] @ LIFT[:, k * Fp:(k + 1) * Fp]
            lj = np.einsum("ab,ebv->eav", LF, jump[:, k])
            rxf[:, k] = self.eta * (sc[:, k] * nrm_out[:, k, 0])[:, None, None] * lj
            ryf[:, k] = self.eta * (sc[:, k] * nrm_out[:, k, 1])[:, None, None] * lj
        qx, qy = gxf + rxf, gyf + ryf
        Qbx = 0.5 * (qx[le[:, None], ll[:, None], fr[None, :]] + qx[re[:, None], rl[:, None], (p - fr)[None, :]])
        Qby = 0.5 * (qy[le[:, None], ll[:, None], fr[None, :]] + qy[re[:, None], rl[:, None], (p - fr)[None, :]])
        Fvf, Gvf = self._visc_flux(UC, Qbx, Qby)
        FvN = Fvf * self.normals[:, None, 0:1] + Gvf * self.normals[:, None, 1:2]   # along the left normal
        return Qx, Qy, FvN

    def _flux(self, U):
        rho = U[..., 0]
        u, v = U[..., 1] / rho, U[..., 2] / rho
        p = (self.g - 1) * (U[..., 3] - 0.5 * rho * (u * u + v * v))
        Fx = np.stack([U[..., 1], U[..., 1] * u + p, U[..., 1] * v, (U[..., 3] + p) * u], -1)
        Fy = np.stack([U[..., 2], U[..., 2] * u, U[..., 2] * v + p, (U[..., 3] + p) * v], -1)
        return Fx, Fy

    def residual(self, U):
        E, Np, Fp, p = self.E, self.Np, self.Fp, self.p
        U = U.reshape(E, Np, 4)
        Fx, Fy = self._flux(U)
        FvN = None
        if self.viscous:  # total flux F_inv - F_v(U, Q) at the points
            Qx, Qy, FvN = self._viscous_parts(U)
            Fv, Gv = self._visc_flux(U, Qx, Qy)
            Fx, Fy = Fx - Fv, Fy - Gv
        rx, ry, sx, sy, J = (self.met[:, k] for k in range(5))
        Dr, Ds, I, LIFT = self.ops["Dr"], self.ops["Ds"], self.ops["interp"], self.ops["lift"]
        div = (np.einsum("ij,ejv->eiv", Dr, Fx) * rx[:, None, None] + np.einsum("ij,ejv->eiv", Ds, Fx) * sx[:, None, None]
               + np.einsum("ij,ejv->eiv", Dr, Fy) * ry[:, None, None] + np.einsum("ij,ejv->eiv", Ds, Fy) * sy[:, None, None])
        Uf = np.einsum("ij,ejv->eiv", I, U).reshape(E, 3, Fp, 4)
        Fxf = np.einsum("ij,ejv->eiv", I, Fx).reshape(E, 3, Fp, 4)
        Fyf = np.einsum("ij,ejv->eiv", I, Fy).reshape(E, 3, Fp, 4)
        f = self.faces
        inner = f[:, 2] >= 0
        bnd = np.nonzero(~inner)[0]
        le, ll, re, rl = f[:, 0], f[:, 1], f[:, 2], f[:, 3]
        fr = np.arange(Fp)
        UL = Uf[le[:, None], ll[:, None], fr[None, :]]                 # [F][Fp][4]
        UR = Uf[np.maximum(re, 0)[:, None], rl[:, None], (p - fr)[None, :]]
        for fi in bnd:  # boundary faces: the ghost state of each point (kernels.cu:453-525)
            bci = int(self.bc_face_map[fi]) if self.bc_face_map is not None else 0
            for q in range(Fp):
                UR[fi, q] = O.ghost_state(UL[fi, q], self.normals[fi, 0], self.normals[fi, 1],
                                          int(self.bc_type[fi]), self.bc_data, max(bci, 0), self.g)
        nrm = np.repeat(self.normals[:, None, :], Fp, axis=1)
        Fc = O.riemann_batch(list(O.RIEMANN)[self.kind], UL, UR, nrm, self.g).reshape(-1, Fp, 4)
        if FvN is not None:  # common viscous normal flux F_v(U^C, Qbar) . n (BR2)
            Fc = Fc - FvN
        Fout = np.zeros((E, 3, Fp, 4))
        Fout[le[:, None], ll[:, None], fr[None, :]] = Fc
        Fout[re[inner][:, None], rl[inner][:, None], (p - fr)[None, :]] = -Fc[inner]
        nrm_out = np.zeros((E, 3, 2))
        nrm_out[le, ll] = self.normals
        nrm_out[re[inner], rl[inner]] = -self.normals[inner]
        Fdn = Fxf * nrm_out[:, :, None, 0:1] + Fyf * nrm_out[:, :, None, 1:2]
        delta = (Fout - Fdn) * (self.elen[:, :, None, None] / (2 * J[:, None, None, None]))
        res = -div - np.einsum("ij,ejv->eiv", LIFT, delta.reshape(E, 3 * Fp, 4))
        return res.reshape(-1)

    def advance(self, U, dt):
        U0 = U.copy()
        U1 = U0 + dt * self.residual(U0)
        U2 = 0.75 * U0 + 0.25 * U1 + 0.25 * dt * self.residual(U1)
        return U0 / 3.0 + 2.0 / 3.0 * U2 + 2.0 / 3.0 * dt * self.residual(U2)
