"""Triangle solver handle (Python face of the zj_tri_* entry points).

The reference declares triangles but leaves their operators zero
(src/basis/basis.cpp:615-669); ``TriSolver`` runs the textbook scheme its
settings name on straight-sided triangles (nodal FR, the energy-stable
correction field of the configured family, configured Riemann solver,
SSP-RK3).  Faces, periodic pairing and normals come
from the library's native mesh routines (Mesh::buildFaces /
Mesh::faceNormal counterparts with three corners per element).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import ZJError, lib
from .solver import CORRS, PRECISIONS, RIEMANNS

ZJ_CORR_DG = 0


def _check(rc, h=None):
    if rc != _lib.ZJ_OK:
        msg = lib().zj_tri_last_error(h)
        raise ZJError(rc, msg.decode() if msg else "")
    return rc


def tri_faces(xy, tris, periodic=None, periodic_nodes=None):
    """Faces [F][4] (left elem, left edge, right elem, right edge) and unit
    normals [F][2]; `periodic` = (Lx, Ly) pairs the boundary faces one period
    apart, `periodic_nodes` = [n][2] (slave, master) node links those whose
    end nodes correspond."""
    faces = _lib.build_faces(tris, nv=3)
    if periodic is not None:
        faces = _lib.pair_periodic(faces, tris, xy, periodic[0], periodic[1], nv=3)
    if periodic_nodes is not None:
        faces = _lib.pair_periodic_nodes(faces, tris, xy, periodic_nodes, nv=3)
    return faces, _lib.face_normals(faces, tris, xy, nv=3)


class TriSolver:
    """One triangle mesh on one GPU.  State arrays are [E][Np][4] (AoS),
    solution points in the order of ``_lib.tri_operators(p)["sp"]``."""

    def __init__(self, xy, tris, p, faces=None, normals=None, periodic=None, bc_type=None, bc_face_map=None,
                 bc_data=None, gamma=1.4, precision="fp64", riemann="roe", device=0, order="hilbert",
                 corr="dg", viscous=False, Re=1600.0, Pr=0.72, br2_eta=0.0):
        xy = np.ascontiguousarray(xy, np.float64)
        tris = np.ascontiguousarray(tris, np.int64)
        if faces is None:
            faces, normals = tri_faces(xy, tris, periodic)
        faces = np.ascontiguousarray(faces, np.int32)
        normals = np.ascontiguousarray(normals, np.float64)
        keep = [xy, tris, faces, normals]
        cols = [np.ascontiguousarray(faces[:, k]) for k in range(4)]
        keep += cols
        d = _lib.TriDesc()
        d.poly_order = int(p)
        d.n_nodes = xy.shape[0]
        d.xy = xy.ctypes.data
        d.n_elem = tris.shape[0]
        d.tri = tris.ctypes.data
        d.n_faces = faces.shape[0]
        d.face_left_elem, d.face_left_local, d.face_right_elem, d.face_right_local = (c.ctypes.data for c in cols)
        d.normals = normals.ctypes.data
        if bc_type is not None:
            bt = np.ascontiguousarray(bc_type, np.int32); keep.append(bt); d.bc_type = bt.ctypes.data
        if bc_face_map is not None:
            bm = np.ascontiguousarray(bc_face_map, np.int32); keep.append(bm); d.bc_face_map = bm.ctypes.data
        if bc_data is not None:
            bd = np.ascontiguousarray(bc_data, np.float64).reshape(-1)
            keep.append(bd); d.bc_data = bd.ctypes.data; d.n_bc_data = bd.size // 8
        d.reorder = 1 if order == "hilbert" else 0
        prm = _lib.Params(gamma, 0.5, PRECISIONS[precision], _lib.ZJ_COMPAT_TEXTBOOK, 0, device,
                          RIEMANNS[riemann], CORRS[corr], int(bool(viscous)), float(Re), float(Pr), float(br2_eta))
        h = C.c_void_p()
        _check(lib().zj_tri_create(C.byref(h), C.byref(d), C.byref(prm)))
        self._h = h
        self.p, self.E = int(p), int(tris.shape[0])
        self.Np = (self.p + 1) * (self.p + 2) // 2
        self.n = self.E * self.Np * 4
        self.precision = precision
        self.faces, self.normals = faces, normals

    def close(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            lib().zj_tri_destroy(h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc):
        return _check(rc, self._h)

    def set_state(self, U):
        U = np.ascontiguousarray(U, np.float64).reshape(-1)
        assert U.size == self.n
        self._chk(lib().zj_tri_set_state(self._h, U))

    def get_state(self):
        out = np.zeros(self.n)
        self._chk(lib().zj_tri_get_state(self._h, out))
        return out

    def save_restart(self, path, time=0.0, iteration=0):
        """Device state to a restart file (zj_tri_restart_save)."""
        from . import io
        io.save_handle("tri", self._h, path, time, iteration)

    def load_restart(self, path):
        """Restart file to the device state (zj_tri_restart_load); returns (time, iteration)."""
        from . import io
        return io.load_handle("tri", self._h, path)

    def set_state_device(self, ptr, stream=None):
        self._chk(lib().zj_tri_set_state_device(self._h, ptr, stream))

    def get_state_device(self, ptr, stream=None):
        self._chk(lib().zj_tri_get_state_device(self._h, ptr, stream))

    def residual(self, fetch=True, stream=None):
        out = np.zeros(self.n) if fetch else None
        self._chk(lib().zj_tri_residual(self._h, None if out is None else out.ctypes.data, stream))
        return out

    def residual_device(self, out_ptr, stream=None):
        self._chk(lib().zj_tri_residual_device(self._h, out_ptr, stream))

    def rk3_step(self, dt, stream=None):
        self._chk(lib().zj_tri_rk3_step(self._h, float(dt), stream))

    def adjoint_set(self, ubar):
        """Seed the state cotangent (zj_tri_adjoint_set)."""
        ubar = np.ascontiguousarray(ubar, np.float64).reshape(-1)
        assert ubar.size == self.n
        self._chk(lib().zj_tri_adjoint_set(self._h, ubar))

    def adjoint_get(self):
        out = np.zeros(self.n)
        self._chk(lib().zj_tri_adjoint_get(self._h, out))
        return out

    def rk3_step_vjp(self, dt, stream=None, want_dt=True):
        """One SSP-RK3 step backwards from the current state (zj_tri_rk3_step_vjp);
        returns dJ/d(dt) (None with want_dt=False)."""
        v = C.c_double()
        self._chk(lib().zj_tri_rk3_step_vjp(self._h, float(dt), C.byref(v) if want_dt else None, stream))
        return v.value if want_dt else None

    def residual_vjp(self, w):
        """(dR/dU)^T w at the current state (zj_tri_residual_vjp)."""
        w = np.ascontiguousarray(w, np.float64).reshape(-1)
        assert w.size == self.n
        out = np.zeros(self.n)
        self._chk(lib().zj_tri_residual_vjp(self._h, w, out))
        return out

    def time_stages(self, dt, steps, stream=None):
        """zj_tri_time_stages: each stage kernel's mean execution time (ms)."""
        out = np.zeros(3)
        self._chk(lib().zj_tri_time_stages(self._h, float(dt), int(steps), out, stream))
        return [float(x) for x in out]

    def dt_cfl(self, cfl):
        v = C.c_double()
        self._chk(lib().zj_tri_dt_cfl(self._h, float(cfl), C.byref(v)))
        return v.value

    def residual_l2(self):
        v = C.c_double()
        self._chk(lib().zj_tri_residual_l2(self._h, C.byref(v)))
        return v.value

    def synchronize(self):
        self._chk(lib().zj_tri_synchronize(self._h))

    def set_xcd_groups(self, groups):
        self._chk(lib().zj_tri_set_xcd_groups(self._h, int(groups)))

    def info(self):
        n, e, s, ce = C.c_int64(), C.c_int(), C.c_int(), C.c_int64()
        self._chk(lib().zj_tri_get_info(self._h, C.byref(n), C.byref(e), C.byref(s), C.byref(ce)))
        return dict(n_dof=n.value, elems_per_block=e.value, n_sp=s.value, cross_tile_edges=ce.value)


def sp_coords(xy, tris, p):
    """Physical coordinates [E][Np][2] of the solution points."""
    sp = _lib.tri_operators(p)["sp"]
    a, b, c = xy[tris[:, 0]], xy[tris[:, 1]], xy[tris[:, 2]]
    r, s = sp[:, 0], sp[:, 1]
    return (a[:, None, :] + 0.5 * (b - a)[:, None, :] * (r[None, :, None] + 1)
            + 0.5 * (c - a)[:, None, :] * (s[None, :, None] + 1))
