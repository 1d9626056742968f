"""ctypes bindings for the CPU oracle (oracle/_build/libzj_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, ``__graft_entry__.smoke()`` and
the ``cpu_baseline`` leg of bench.py, never by the product package.  See
oracle/zj_oracle.h for what each function restates and how it is pinned.
"""
from __future__ import annotations

import ctypes as C
import math
import os
from dataclasses import dataclass
from typing import Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libzj_oracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libzjref.so")

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_lp = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")

BC_INTERIOR, BC_WALL, BC_SLIPWALL, BC_SYMMETRY, BC_FARFIELD, BC_INFLOW, BC_OUTFLOW, BC_PERIODIC = range(8)

_lib = None


def build(quiet: bool = True) -> None:
    """Compile the oracle (and oracle/_ref when the reference sources exist)."""
    import subprocess
    subprocess.run(["make", "-C", HERE, "-s", "all"], check=True,
                   stdout=subprocess.DEVNULL if quiet else None)


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing: {LIB_PATH} (run make -C oracle)")
        L = C.CDLL(LIB_PATH)
        L.oz_gauss.argtypes = [C.c_int, _dp, _dp]
        L.oz_lagrange.argtypes = [_dp, C.c_int, C.c_int, C.c_double]
        L.oz_lagrange.restype = C.c_double
        L.oz_lagrange_deriv.argtypes = [_dp, C.c_int, C.c_int, C.c_double]
        L.oz_lagrange_deriv.restype = C.c_double
        L.oz_corr_left_deriv.argtypes = [C.c_int, C.c_double]
        L.oz_corr_left_deriv.restype = C.c_double
        L.oz_corr_right_deriv.argtypes = [C.c_int, C.c_double]
        L.oz_corr_right_deriv.restype = C.c_double
        L.oz_quad_operators.argtypes = [C.c_int] + [_dp] * 7
        L.oz_build_faces.argtypes = [C.c_int, C.c_int, C.c_int, _lp, _ip, _ip, _ip, _ip]
        L.oz_pair_periodic.argtypes = [C.c_int, _ip, _ip, _ip, _ip, C.c_int, C.c_int, _lp, _dp,
                                       C.c_double, C.c_double, C.c_double]
        L.oz_face_normals.argtypes = [C.c_int, _ip, _ip, C.c_int, C.c_int, _lp, _dp, _dp]
        L.oz_elem_areas.argtypes = [C.c_int, C.c_int, C.c_int, _lp, _dp, _dp]
        L.oz_geometry.argtypes = [C.c_int, C.c_int, C.c_int, _lp, _dp, C.c_int, _dp, _dp, _dp, _dp]
        L.oz_sp_coords.argtypes = [C.c_int, C.c_int, C.c_int, _lp, _dp, C.c_int, _dp, _dp, _dp]
        L.oz_create.argtypes = [C.c_void_p]
        L.oz_create.restype = C.c_void_p
        L.oz_destroy.argtypes = [C.c_void_p]
        L.oz_set_threads.argtypes = [C.c_int]
        for name in ("oz_k_flux_sp", "oz_k_interp"):
            getattr(L, name).argtypes = [C.c_void_p, _dp]
        for name in ("oz_k_riemann", "oz_k_flux_diff"):
            getattr(L, name).argtypes = [C.c_void_p]
        for name in ("oz_k_divergence", "oz_k_correction"):
            getattr(L, name).argtypes = [C.c_void_p, _dp]
        L.oz_residual.argtypes = [C.c_void_p, _dp, C.c_void_p]
        L.oz_rk_stage1.argtypes = [_dp, _dp, _dp, C.c_int, C.c_double]
        L.oz_rk_stage2.argtypes = [_dp, _dp, _dp, _dp, C.c_int, C.c_double]
        L.oz_rk_stage3.argtypes = [_dp, _dp, _dp, _dp, C.c_int, C.c_double]
        L.oz_positivity.argtypes = [_dp, C.c_int]
        L.oz_advance.argtypes = [C.c_void_p, _dp, C.c_double, C.c_int]
        L.oz_dt_cfl.argtypes = [C.c_void_p, _dp, C.c_double]
        L.oz_dt_cfl.restype = C.c_double
        L.oz_residual_l2.argtypes = [C.c_void_p]
        L.oz_residual_l2.restype = C.c_double
        L.oz_flux_normal.argtypes = [_dp, C.c_double, C.c_double, C.c_double, _dp]
        L.oz_max_wave_speed.argtypes = [_dp, C.c_double]
        L.oz_max_wave_speed.restype = C.c_double
        L.oz_riemann.argtypes = [C.c_int, _dp, _dp, C.c_double, C.c_double, C.c_double, _dp]
        L.oz_vcjh_deriv.argtypes = [C.c_int, C.c_int, _dp, C.c_int, _dp, _dp]
        L.oz_riemann_batch.argtypes = [C.c_int, C.c_int, _dp, _dp, _dp, C.c_double, _dp]
        L.oz_tb_create.argtypes = [C.c_void_p, _dp, C.c_int, C.c_int]
        L.oz_tb_create.restype = C.c_void_p
        L.oz_tb_destroy.argtypes = [C.c_void_p]
        L.oz_tb_residual.argtypes = [C.c_void_p, _dp, _dp]
        L.oz_tb_advance.argtypes = [C.c_void_p, _dp, C.c_double, C.c_int]
        L.oz_tb_set_viscous.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_double]
        L.oz_tb_gradient.argtypes = [C.c_void_p, _dp, _dp]
        L.oz_visc_flux.argtypes = [_dp, _dp, _dp, C.c_double, C.c_double, C.c_double, _dp, _dp]
        for name in ("oz_buf_fx", "oz_buf_fy", "oz_buf_ufp", "oz_buf_fface", "oz_buf_fdiff",
                     "oz_buf_dudt", "oz_buf_ufp_mut", "oz_buf_fdiff_mut"):
            getattr(L, name).argtypes = [C.c_void_p]
            getattr(L, name).restype = C.POINTER(C.c_double)
        L.oz_flux_xy.argtypes = [_dp, C.c_double, _dp, _dp]
        L.oz_rusanov_inline.argtypes = [_dp, _dp, C.c_double, C.c_double, C.c_double, _dp]
        L.oz_ghost_state.argtypes = [_dp, C.c_double, C.c_double, C.c_int, _dp, C.c_int,
                                     C.c_double, _dp]
        _lib = L
    return _lib


def set_threads(n: int) -> None:
    lib().oz_set_threads(int(n))


# --------------------------------------------------------------- operators
def gauss(n: int):
    p = np.zeros(n); w = np.zeros(n)
    lib().oz_gauss(n, p, w)
    return p, w


def quad_operators(p: int) -> dict:
    n = p + 1; S = n * n; Fp = n
    out = dict(sol_xi=np.zeros(S), sol_eta=np.zeros(S), qw=np.zeros(S),
               diff_xi=np.zeros(S * S), diff_eta=np.zeros(S * S),
               interp=np.zeros(4 * Fp * S), corr=np.zeros(4 * S))
    lib().oz_quad_operators(p, out["sol_xi"], out["sol_eta"], out["qw"], out["diff_xi"],
                            out["diff_eta"], out["interp"], out["corr"])
    return out


def corr_deriv(p: int, xi: float, right: bool) -> float:
    L = lib()
    return L.oz_corr_right_deriv(p, xi) if right else L.oz_corr_left_deriv(p, xi)


# -------------------------------------------------------------------- mesh
@dataclass
class Faces:
    le: np.ndarray
    ll: np.ndarray
    re: np.ndarray
    rl: np.ndarray

    @property
    def n(self) -> int:
        return int(self.le.shape[0])

    def as_array(self) -> np.ndarray:
        return np.stack([self.le, self.ll, self.re, self.rl], axis=1)


def build_faces(elem_nodes: np.ndarray, nv: Optional[int] = None) -> Faces:
    en = np.ascontiguousarray(elem_nodes, dtype=np.int64)
    E, stride = en.shape
    nv = stride if nv is None else nv
    cap = max(E * nv, 1)
    arrs = [np.zeros(cap, np.int32) for _ in range(4)]
    nf = lib().oz_build_faces(E, nv, stride, en.reshape(-1), *arrs)
    if nf < 0:
        raise MemoryError("oz_build_faces")
    return Faces(*[a[:nf].copy() for a in arrs])


def pair_periodic(f: Faces, elem_nodes: np.ndarray, xy: np.ndarray, Lx: float, Ly: float,
                  tol: float = 1e-9, nv: Optional[int] = None) -> Faces:
    en = np.ascontiguousarray(elem_nodes, dtype=np.int64)
    stride = en.shape[1]
    nv = stride if nv is None else nv
    arrs = [a.copy() for a in (f.le, f.ll, f.re, f.rl)]
    nf = lib().oz_pair_periodic(f.n, *arrs, nv, stride, en.reshape(-1),
                                np.ascontiguousarray(xy, np.float64).reshape(-1), Lx, Ly, tol)
    if nf < 0:
        raise ValueError("ambiguous periodic partners")
    return Faces(*[a[:nf].copy() for a in arrs])


def pair_periodic_nodes(f: Faces, elem_nodes: np.ndarray, xy, links, tol: Optional[float] = None,
                        nv: Optional[int] = None) -> Faces:
    """Periodic pairing by node correspondence (no reference counterpart: the
    reference's Gmsh reader skips $Periodic).  Restated with plain loops:
    node labels by propagation (each link sets both ends to the smaller label
    until nothing changes); the link offsets xy[master] - xy[slave]; then, in
    ascending face order, each unpaired boundary face f takes as partner the
    later unpaired boundary face g whose end-node labels equal f's and whose
    end nodes are f's moved by one common vector equal (within tol, either
    sign) to a link offset.  Two candidates: ValueError.  The lower-indexed
    face keeps its place with the partner's left side as its right side; the
    partner is removed.  Pinned in the tests by equality with pair_periodic on
    translation-periodic meshes, by the edge-cover and order invariants and
    partner midpoints one spanning vector apart on sheared meshes that
    pair_periodic cannot pair, and by wall edges left unpaired."""
    en = np.asarray(elem_nodes, np.int64)
    nv = en.shape[1] if nv is None else nv
    P = np.asarray(xy, np.float64).reshape(-1, 2)
    if tol is None:
        tol = 1e-9 * float(np.ptp(P, axis=0).max())
    lab = list(range(P.shape[0]))
    pairs = [(int(a), int(b)) for a, b in np.asarray(links, np.int64).reshape(-1, 2)]
    changed = True
    while changed:
        changed = False
        for a, b in pairs:
            m = min(lab[a], lab[b])
            if lab[a] != m or lab[b] != m:
                lab[a] = lab[b] = m
                changed = True
    offsets = [P[b] - P[a] for a, b in pairs]

    def is_offset(d):
        return any(max(abs(d[0] - s * o[0]), abs(d[1] - s * o[1])) <= tol for o in offsets for s in (1.0, -1.0))

    def ends(k):
        e, l = int(f.le[k]), int(f.ll[k])
        return int(en[e, l]), int(en[e, (l + 1) % nv])

    def moved(k, g):
        a0, a1 = ends(k)
        for c0, c1 in (ends(g), ends(g)[::-1]):
            d0, d1 = P[c0] - P[a0], P[c1] - P[a1]
            if max(abs(d0 - d1)) <= tol and is_offset(d0):
                return True
        return False

    bnd = [k for k in range(f.n) if f.re[k] < 0]
    key = {k: tuple(sorted(lab[n] for n in ends(k))) for k in bnd}
    partner, taken = {}, set()
    for x, k in enumerate(bnd):
        if k in taken:
            continue
        cands = [g for g in bnd[x + 1:] if g not in taken and key[g] == key[k] and moved(k, g)]
        if len(cands) > 1:
            raise ValueError("ambiguous periodic partners")
        if cands:
            partner[k] = cands[0]
            taken.update((k, cands[0]))
    dead = set(partner.values())
    rows = []
    for k in range(f.n):
        if k in dead:
            continue
        g = partner.get(k)
        rows.append((f.le[k], f.ll[k], f.le[g] if g is not None else f.re[k], f.ll[g] if g is not None else f.rl[k]))
    a = np.array(rows, np.int32).reshape(-1, 4)
    return Faces(*[a[:, c].copy() for c in range(4)])


def face_normals(f: Faces, elem_nodes, xy, nv=None) -> np.ndarray:
    en = np.ascontiguousarray(elem_nodes, dtype=np.int64)
    stride = en.shape[1]
    nv = stride if nv is None else nv
    out = np.zeros(2 * max(f.n, 1))
    lib().oz_face_normals(f.n, f.le, f.ll, nv, stride, en.reshape(-1),
                          np.ascontiguousarray(xy, np.float64).reshape(-1), out)
    return out[: 2 * f.n].reshape(f.n, 2)


def elem_areas(elem_nodes, xy, nv=None) -> np.ndarray:
    en = np.ascontiguousarray(elem_nodes, dtype=np.int64)
    E, stride = en.shape
    nv = stride if nv is None else nv
    out = np.zeros(max(E, 1))
    lib().oz_elem_areas(E, nv, stride, en.reshape(-1),
                        np.ascontiguousarray(xy, np.float64).reshape(-1), out)
    return out[:E]


def geometry(elem_nodes, xy, ops: dict, node_count: Optional[int] = None):
    en = np.ascontiguousarray(elem_nodes, dtype=np.int64)
    E, stride = en.shape
    nc = stride if node_count is None else node_count
    S = ops["sol_xi"].shape[0]
    J = np.zeros(max(E * S, 1)); Jinv = np.zeros(max(E * S * 4, 1))
    rc = lib().oz_geometry(E, nc, stride, en.reshape(-1),
                           np.ascontiguousarray(xy, np.float64).reshape(-1), S,
                           ops["sol_xi"], ops["sol_eta"], J, Jinv)
    if rc:
        raise ValueError("unsupported element node count %d" % nc)
    return J[: E * S], Jinv[: E * S * 4]


def sp_coords(elem_nodes, xy, ops: dict, node_count: Optional[int] = None) -> np.ndarray:
    en = np.ascontiguousarray(elem_nodes, dtype=np.int64)
    E, stride = en.shape
    nc = stride if node_count is None else node_count
    S = ops["sol_xi"].shape[0]
    pos = np.zeros(max(E * S * 2, 1))
    rc = lib().oz_sp_coords(E, nc, stride, en.reshape(-1),
                            np.ascontiguousarray(xy, np.float64).reshape(-1), S,
                            ops["sol_xi"], ops["sol_eta"], pos)
    if rc:
        raise ValueError("unsupported element node count %d" % nc)
    return pos[: E * S * 2].reshape(E, S, 2)


# ----------------------------------------------------------------- physics
def flux_xy(U, gamma=1.4):
    Fx = np.zeros(4); Fy = np.zeros(4)
    lib().oz_flux_xy(np.ascontiguousarray(U, np.float64), gamma, Fx, Fy)
    return Fx, Fy


def rusanov(UL, UR, nx, ny, gamma=1.4):
    F = np.zeros(4)
    lib().oz_rusanov_inline(np.ascontiguousarray(UL, np.float64),
                            np.ascontiguousarray(UR, np.float64), nx, ny, gamma, F)
    return F


def ghost_state(Ui, nx, ny, bc_type, bc_data, bc_idx, gamma=1.4):
    G = np.zeros(4)
    bd = np.ascontiguousarray(bc_data if bc_data is not None else np.zeros(8), np.float64).reshape(-1)
    lib().oz_ghost_state(np.ascontiguousarray(Ui, np.float64), nx, ny, bc_type, bd, bc_idx, gamma, G)
    return G


# ------------------------------------------------------------------ solver
class _Desc(C.Structure):
    _fields_ = [("p", C.c_int), ("n_elem", C.c_int), ("n_faces", C.c_int), ("gamma", C.c_double),
                ("diff_xi", C.c_void_p), ("diff_eta", C.c_void_p), ("interp", C.c_void_p),
                ("corr", C.c_void_p), ("J", C.c_void_p), ("Jinv", C.c_void_p),
                ("normals", C.c_void_p), ("face_le", C.c_void_p), ("face_ll", C.c_void_p),
                ("face_re", C.c_void_p), ("face_rl", C.c_void_p), ("bc_type", C.c_void_p),
                ("bc_face_map", C.c_void_p), ("bc_data", C.c_void_p), ("n_bc_data", C.c_int),
                ("elem_area", C.c_void_p)]


def _ptr(a):
    return None if a is None else a.ctypes.data


class OracleSolver:
    """The reference residual / SSP-RK3 path on the CPU (one instance per mesh)."""

    def __init__(self, p, ops, J, Jinv, faces: Faces, normals, gamma=1.4, bc_type=None,
                 bc_face_map=None, bc_data=None, elem_area=None):
        self.p = p
        self.S = (p + 1) ** 2
        self.Fp = p + 1
        self.E = J.shape[0] // self.S
        self.F = faces.n
        keep = []

        def c(a, dt):
            if a is None:
                return None
            a = np.ascontiguousarray(a, dtype=dt)
            keep.append(a)
            return a
        d = _Desc()
        d.p = p; d.n_elem = self.E; d.n_faces = self.F; d.gamma = gamma
        d.diff_xi = _ptr(c(ops["diff_xi"], np.float64)); d.diff_eta = _ptr(c(ops["diff_eta"], np.float64))
        d.interp = _ptr(c(ops["interp"], np.float64)); d.corr = _ptr(c(ops["corr"], np.float64))
        d.J = _ptr(c(J, np.float64)); d.Jinv = _ptr(c(Jinv, np.float64))
        d.normals = _ptr(c(normals, np.float64))
        d.face_le = _ptr(c(faces.le, np.int32)); d.face_ll = _ptr(c(faces.ll, np.int32))
        d.face_re = _ptr(c(faces.re, np.int32)); d.face_rl = _ptr(c(faces.rl, np.int32))
        d.bc_type = _ptr(c(bc_type, np.int32)); d.bc_face_map = _ptr(c(bc_face_map, np.int32))
        bd = c(bc_data, np.float64)
        d.bc_data = _ptr(bd); d.n_bc_data = 0 if bd is None else bd.size // 8
        d.elem_area = _ptr(c(elem_area, np.float64))
        h = lib().oz_create(C.byref(d))
        if not h:
            raise MemoryError("oz_create failed")
        self._h = h
        self.n = self.E * self.S * 4
        self.desc = (d, keep)

    def textbook(self, edge_half_len, riemann="roe", corr="dg") -> "TextbookSolver":
        return TextbookSolver(self, self.desc, edge_half_len, riemann, corr)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().oz_destroy(h)
            self._h = None

    def _view(self, fn, count):
        ptr = getattr(lib(), fn)(self._h)
        return np.ctypeslib.as_array(ptr, shape=(count,)).copy()

    def residual(self, U: np.ndarray) -> np.ndarray:
        U = np.ascontiguousarray(U, np.float64).reshape(-1)
        out = np.zeros(self.n)
        lib().oz_residual(self._h, U, out.ctypes.data)
        return out

    def advance(self, U: np.ndarray, dt: float, positivity: bool = True) -> np.ndarray:
        U = np.ascontiguousarray(U, np.float64).reshape(-1).copy()
        lib().oz_advance(self._h, U, dt, int(positivity))
        return U

    def dt_cfl(self, U, cfl):
        return lib().oz_dt_cfl(self._h, np.ascontiguousarray(U, np.float64).reshape(-1), cfl)

    def residual_l2(self):
        return lib().oz_residual_l2(self._h)

    # staged access for kernel-level tests
    def k_flux_sp(self, U):
        lib().oz_k_flux_sp(self._h, np.ascontiguousarray(U, np.float64).reshape(-1))
        return self._view("oz_buf_fx", self.n), self._view("oz_buf_fy", self.n)

    def k_interp(self, U):
        lib().oz_k_interp(self._h, np.ascontiguousarray(U, np.float64).reshape(-1))
        return self._view("oz_buf_ufp", self.E * 4 * self.Fp * 4)

    def set_ufp(self, ufp):
        ptr = lib().oz_buf_ufp_mut(self._h)
        np.ctypeslib.as_array(ptr, shape=(self.E * 4 * self.Fp * 4,))[:] = np.asarray(ufp).reshape(-1)

    def set_fdiff(self, fd):
        ptr = lib().oz_buf_fdiff_mut(self._h)
        np.ctypeslib.as_array(ptr, shape=(self.E * 4 * self.Fp * 4,))[:] = np.asarray(fd).reshape(-1)

    def k_riemann(self):
        lib().oz_k_riemann(self._h)
        return self._view("oz_buf_fface", self.F * self.Fp * 4)

    def k_flux_diff(self):
        lib().oz_k_flux_diff(self._h)
        return self._view("oz_buf_fdiff", self.E * 4 * self.Fp * 4)

    def k_divergence(self):
        out = np.zeros(self.n)
        lib().oz_k_divergence(self._h, out)
        return out

    def k_correction(self, dudt):
        out = np.ascontiguousarray(dudt, np.float64).reshape(-1).copy()
        lib().oz_k_correction(self._h, out)
        return out


def rk_stage(k, U0, Uk, R, dt):
    n = U0.size
    out = np.zeros(n)
    L = lib()
    if k == 1:
        L.oz_rk_stage1(out, U0, R, n, dt)
    elif k == 2:
        L.oz_rk_stage2(out, U0, Uk, R, n, dt)
    else:
        L.oz_rk_stage3(out, U0, Uk, R, n, dt)
    return out


def positivity(U):
    U = np.ascontiguousarray(U, np.float64).reshape(-1).copy()
    lib().oz_positivity(U, U.size // 4)
    return U


# ------------------------------------------------------------ textbook scheme
RIEMANN = {"rusanov": 0, "roe": 1, "hllc": 2}
CORR = {"dg": 0, "sd": 1, "hu": 2, "ga": 3}


def flux_normal(U, nx, ny, gamma=1.4):
    out = np.zeros(4)
    lib().oz_flux_normal(np.ascontiguousarray(U, np.float64), nx, ny, gamma, out)
    return out


def riemann(kind, UL, UR, nx, ny, gamma=1.4):
    out = np.zeros(4)
    lib().oz_riemann(RIEMANN[kind], np.ascontiguousarray(UL, np.float64),
                     np.ascontiguousarray(UR, np.float64), nx, ny, gamma, out)
    return out


def vcjh_deriv(p, kind, x):
    x = np.ascontiguousarray(x, np.float64).reshape(-1)
    gl, gr = np.zeros(x.size), np.zeros(x.size)
    if lib().oz_vcjh_deriv(p, CORR[kind], x, x.size, gl, gr) != 0:
        raise ValueError("bad order or kind")
    return gl, gr


def edge_half_lengths(elem_nodes, xy):
    """[E][4]: half the length of edge k = (node k, node k+1) of each element."""
    en = np.asarray(elem_nodes)[:, :4]
    a = xy[en]
    b = xy[np.roll(en, -1, axis=1)]
    return 0.5 * np.linalg.norm(b - a, axis=2)


class TextbookSolver:
    """Textbook FR residual / SSP-RK3 (zj_textbook.c) on the CPU."""

    def __init__(self, base: OracleSolver, keep_desc, edge_half_len, riemann="roe", corr="dg"):
        self.n = base.n
        self._ehl = np.ascontiguousarray(edge_half_len, np.float64).reshape(-1)
        self._desc, self._keep = keep_desc
        h = lib().oz_tb_create(C.byref(self._desc), self._ehl, RIEMANN[riemann], CORR[corr])
        if not h:
            raise ValueError("oz_tb_create failed")
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().oz_tb_destroy(h)
            self._h = None

    def residual(self, U):
        U = np.ascontiguousarray(U, np.float64).reshape(-1)
        out = np.zeros(self.n)
        lib().oz_tb_residual(self._h, U, out)
        return out

    def advance(self, U, dt, positivity=True):
        U = np.ascontiguousarray(U, np.float64).reshape(-1).copy()
        lib().oz_tb_advance(self._h, U, dt, int(positivity))
        return U

    def set_viscous(self, Re, Pr=0.72, eta=1.0, gamma=1.4):
        """BR2 viscous terms with mu = 1/Re, kappa = mu gamma / ((gamma-1) Pr)."""
        mu = 1.0 / Re
        lib().oz_tb_set_viscous(self._h, mu, mu * gamma / ((gamma - 1.0) * Pr), eta)
        return self

    def gradient(self, U):
        """Corrected gradient [E*S][8] (d/dx of the 4 variables, then d/dy)."""
        U = np.ascontiguousarray(U, np.float64).reshape(-1)
        out = np.zeros(2 * self.n)
        lib().oz_tb_gradient(self._h, U, out)
        return out.reshape(-1, 8)


def visc_flux(U, Qx, Qy, mu, kappa, gamma=1.4):
    Fv, Gv = np.zeros(4), np.zeros(4)
    lib().oz_visc_flux(*(np.ascontiguousarray(a, np.float64) for a in (U, Qx, Qy)), gamma, mu, kappa, Fv, Gv)
    return Fv, Gv


def riemann_batch(kind, UL, UR, nrm, gamma=1.4):
    UL = np.ascontiguousarray(UL, np.float64).reshape(-1, 4)
    UR = np.ascontiguousarray(UR, np.float64).reshape(-1, 4)
    nrm = np.ascontiguousarray(nrm, np.float64).reshape(-1, 2)
    out = np.zeros_like(UL)
    lib().oz_riemann_batch(RIEMANN[kind], UL.shape[0], UL, UR, nrm, gamma, out)
    return out


# --------------------------------------------- the reference's gpu:: helpers
# Restatements of the free functions of kernels.cu that FRSolver's residual
# does not call, for the tests of the library's gpu:: layer (small inputs,
# plain Python loops).

def _pressure(U, gamma):
    """IdealGas::pressure (euler.hpp:16-23)."""
    rho = U[0]
    u, v = U[1] / rho, U[2] / rho
    return (gamma - 1.0) * rho * (U[3] / rho - 0.5 * (u * u + v * v))


def _sound_speed(U, gamma):
    """IdealGas::soundSpeed (euler.hpp:32-41)."""
    return math.sqrt(gamma * max(_pressure(U, gamma), 1e-10) / max(U[0], 1e-10))


def _prim2cons(rho, u, v, p, gamma):
    return np.array([rho, rho * u, rho * v, p / (gamma - 1.0) + 0.5 * rho * (u * u + v * v)])


def apply_boundary_conditions(U_fp, elem_idx, local_face, bc_type, bc_data, normals, gamma, nfp, nfe):
    """applyBoundaryConditionsKernel (kernels.cu:1040-1211): returns a copy of
    U_fp [E][nfe][nfp][4] with the ghost state written over each boundary
    face's interior flux-point states; normals [n_bc][nfp][2], bc_data
    [n_bc][8], both indexed by boundary face.  Pinned by the relations each
    branch must satisfy (tests/test_gpu_api_oracle.py)."""
    out = np.array(U_fp, np.float64).reshape(-1, nfe, nfp, 4).copy()
    nrm = np.asarray(normals, np.float64).reshape(-1, nfp, 2)
    bd = np.asarray(bc_data, np.float64).reshape(-1, 8)
    g = gamma
    for b in range(len(elem_idx)):
        for f in range(nfp):
            U = out[elem_idx[b], local_face[b], f].copy()
            nx, ny = nrm[b, f]
            t = int(bc_type[b])
            if t in (BC_WALL, BC_SLIPWALL, BC_SYMMETRY):
                rho = U[0]; u = U[1] / rho; v = U[2] / rho; p = _pressure(U, g)
                vn = u * nx + v * ny
                G = _prim2cons(rho, u - 2.0 * vn * nx, v - 2.0 * vn * ny, p, g)
            elif t == BC_FARFIELD:
                rinf, uinf, vinf, pinf = bd[b, :4]
                cinf = _sound_speed(_prim2cons(rinf, uinf, vinf, pinf, g), g)
                ri = U[0]; ui = U[1] / ri; vi = U[2] / ri; ci = _sound_speed(U, g)
                vni = ui * nx + vi * ny
                vninf = uinf * nx + vinf * ny
                Rp = vni + 2.0 * ci / (g - 1.0)
                Rm = vninf - 2.0 * cinf / (g - 1.0)
                vnb = 0.5 * (Rp + Rm)
                cb = 0.25 * (g - 1.0) * (Rp - Rm)
                if vnb >= 0:
                    vt = ui * (-ny) + vi * nx
                    s = _pressure(U, g) / ri ** g
                else:
                    vt = uinf * (-ny) + vinf * nx
                    s = pinf / rinf ** g
                rb = (cb * cb / (g * s)) ** (1.0 / (g - 1.0))
                pb = rb * cb * cb / g
                G = _prim2cons(rb, vnb * nx - vt * ny, vnb * ny + vt * nx, pb, g)
            elif t == BC_OUTFLOW:
                rho = U[0]
                G = _prim2cons(rho, U[1] / rho, U[2] / rho, bd[b, 0], g)
            elif t == BC_INFLOW:
                pt, Tt, dx, dy = bd[b, :4]
                dm = math.sqrt(dx * dx + dy * dy)
                if dm > 1e-10:
                    dx, dy = dx / dm, dy / dm
                else:
                    dx, dy = -nx, -ny
                ps = _pressure(U, g)
                M2 = max(2.0 / (g - 1.0) * ((pt / ps) ** ((g - 1.0) / g) - 1.0), 0.0)
                M = math.sqrt(M2)
                Ts = Tt / (1.0 + 0.5 * (g - 1.0) * M2)
                rb = ps / Ts
                vel = M * math.sqrt(g * ps / rb)
                G = _prim2cons(rb, vel * dx, vel * dy, ps, g)
            else:
                G = U
            out[elem_idx[b], local_face[b], f] = G
    return out.reshape(-1)


def dt_tree_points(n_sp):
    """Solution points whose value reaches computeTimeStepKernel's per-element
    minimum (kernels.cu:1389-1397, blockDim = n_sp): the tree
    `for s = n_sp/2; s > 0; s >>= 1: if sp < s and sp + s < n_sp` run on index
    sets."""
    sets = [{i} for i in range(n_sp)]
    s = n_sp // 2
    while s > 0:
        for sp in range(s):
            if sp + s < n_sp:
                sets[sp] |= sets[sp + s]
        s >>= 1
    return sorted(sets[0])


def dt_gpu(U, J, h_min, cfl, gamma, n_sp):
    """gpu::computeTimeStep (kernels.cu:1355-1474): min of
    CFL h_min[e] / (maxWaveSpeed(U) sqrt|J|) over the points dt_tree_points keeps."""
    U = np.asarray(U, np.float64).reshape(-1, n_sp, 4)
    J = np.asarray(J, np.float64).reshape(-1, n_sp)
    keep = dt_tree_points(n_sp)
    best = 1e20
    for e in range(U.shape[0]):
        for sp in keep:
            ws = max_wave_speed(U[e, sp], gamma)
            best = min(best, cfl * h_min[e] / (ws * math.sqrt(abs(J[e, sp]))))
    return best


def max_wave_speed(U, gamma=1.4):
    """IdealGas::maxWaveSpeed (euler.hpp:166-171) through the oracle library."""
    return lib().oz_max_wave_speed(np.ascontiguousarray(U, np.float64), gamma)
