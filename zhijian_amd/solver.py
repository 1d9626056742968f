"""Host-side handle of one device solver (Python face of include/zhijian_amd.h).

``Solver`` owns a zj_handle.  ``Solver.from_mesh`` builds every table the
reference's FRSolver::initialize builds (operators, faces with periodic
pairing, normals, Jacobians, areas, BC tables) with the library's native host
code, then creates the device solver.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass
from typing import Optional

import numpy as np

from . import _lib
from ._lib import ZJError, check, lib

PRECISIONS = {"fp64": _lib.ZJ_PREC_FP64, "fp32bulk": _lib.ZJ_PREC_FP32BULK}
BC_FARFIELD = 4
BC_OUTFLOW = 6


def _ptr(a):
    return None if a is None else a.ctypes.data


@dataclass
class Tables:
    """Everything uploaded at zj_create, in the reference's layouts."""
    p: int
    ops: dict
    faces: np.ndarray        # [F,4] (left elem, left edge, right elem, right edge)
    normals: np.ndarray      # [F,2]
    J: np.ndarray
    Jinv: np.ndarray
    area: np.ndarray
    bc_type: np.ndarray
    bc_face_map: np.ndarray
    bc_data: np.ndarray
    pos: Optional[np.ndarray] = None
    edge_half_len: Optional[np.ndarray] = None   # [E][4], textbook mode
    centroid: Optional[np.ndarray] = None        # [E][2], device element ordering

    @property
    def E(self):
        return self.area.shape[0]

    @property
    def S(self):
        return (self.p + 1) ** 2


def freestream(gamma=1.4, p_inf=101325.0, rho_inf=1.225, mach=0.5, aoa_deg=0.0):
    """Far-field state as FRSolver builds it (fr_solver.cpp:438-443)."""
    c = math.sqrt(gamma * p_inf / rho_inf)
    vel = mach * c
    a = aoa_deg * math.pi / 180.0
    return rho_inf, vel * math.cos(a), vel * math.sin(a), p_inf


def bc_tables(faces: np.ndarray, bc_type_per_face=None, fs=None, outflow_p=None):
    """fr_solver.cpp:432-482: one 8-double entry per boundary face, face order."""
    F = faces.shape[0]
    bct = np.zeros(F, np.int32)
    bmap = -np.ones(F, np.int32)
    bnd = np.nonzero(faces[:, 2] < 0)[0]
    fs = freestream() if fs is None else fs
    data = np.zeros((max(len(bnd), 1), 8))
    for n, f in enumerate(bnd):
        t = BC_FARFIELD if bc_type_per_face is None else int(bc_type_per_face[f])
        bct[f] = t
        bmap[f] = n
        if t == BC_FARFIELD:
            data[n, :4] = fs
        elif t == BC_OUTFLOW:
            data[n, 0] = fs[3] if outflow_p is None else outflow_p
    return bct, bmap, data.reshape(-1)


def build_tables(xy, elem_nodes, p, periodic=None, bc_type_per_face=None, fs=None,
                 node_count=None, want_pos=True, periodic_nodes=None) -> Tables:
    """Host tables of a quadrilateral mesh.  periodic=(Lx, Ly) joins boundary
    faces one period apart; periodic_nodes=[n][2] (slave, master) node links,
    e.g. a Gmsh file's $Periodic section, joins the faces whose end nodes
    correspond."""
    en = np.ascontiguousarray(elem_nodes, np.int64)
    ops = _lib.quad_operators(p)
    faces = _lib.build_faces(en[:, :4], nv=4)
    if periodic is not None:
        faces = _lib.pair_periodic(faces, en[:, :4], xy, periodic[0], periodic[1])
    if periodic_nodes is not None:
        faces = _lib.pair_periodic_nodes(faces, en[:, :4], xy, periodic_nodes)
    normals = _lib.face_normals(faces, en[:, :4], xy, nv=4)
    geo = _lib.geometry(en, xy, ops, node_count=node_count, want_pos=want_pos)
    area = _lib.elem_areas(en[:, :4], xy, nv=4)
    bct, bmap, bd = bc_tables(faces, bc_type_per_face, fs)
    corners = np.asarray(xy, np.float64)[en[:, :4]]
    ehl = 0.5 * np.linalg.norm(np.roll(corners, -1, axis=1) - corners, axis=2)
    return Tables(p, ops, faces, normals, geo[0], geo[1], area, bct, bmap, bd,
                  geo[2] if want_pos else None, ehl, corners.mean(axis=1))


COMPATS = {"reference": 0, "textbook": 1}
RIEMANNS = {"rusanov": 0, "roe": 1, "hllc": 2}
CORRS = {"dg": 0, "sd": 1, "hu": 2, "ga": 3}


class Solver:
    def __init__(self, t: Tables, gamma=1.4, cfl=0.5, precision="fp64", positivity=True, device=0,
                 compat="reference", riemann="rusanov", corr="dg", order="hilbert", viscous=False, Re=1000.0,
                 Pr=0.72, br2_eta=0.0):
        self.tables = t
        self.p, self.E, self.S = t.p, t.E, t.S
        self.n = self.E * self.S * 4
        keep = []

        def c(a, dt):
            a = np.ascontiguousarray(a, dtype=dt)
            keep.append(a)
            return a
        fc = c(t.faces, np.int32)
        le, ll, re, rl = (c(fc[:, k], np.int32) for k in range(4))
        d = _lib.MeshDesc()
        d.poly_order = t.p; d.n_elem = t.E; d.n_faces = fc.shape[0]
        d.diff_xi = _ptr(c(t.ops["diff_xi"], np.float64)); d.diff_eta = _ptr(c(t.ops["diff_eta"], np.float64))
        d.interp = _ptr(c(t.ops["interp"], np.float64)); d.corr = _ptr(c(t.ops["corr"], np.float64))
        d.J = _ptr(c(t.J, np.float64)); d.Jinv = _ptr(c(t.Jinv, np.float64))
        d.normals = _ptr(c(t.normals, np.float64))
        d.face_left_elem = _ptr(le); d.face_left_local = _ptr(ll)
        d.face_right_elem = _ptr(re); d.face_right_local = _ptr(rl)
        d.bc_type = _ptr(c(t.bc_type, np.int32)); d.bc_face_map = _ptr(c(t.bc_face_map, np.int32))
        bd = c(t.bc_data, np.float64)
        d.bc_data = _ptr(bd); d.n_bc_data = bd.size // 8
        d.elem_area = _ptr(c(t.area, np.float64))
        if t.edge_half_len is not None:
            d.edge_half_len = _ptr(c(t.edge_half_len, np.float64))
        if order == "hilbert" and t.centroid is not None:
            d.elem_centroid = _ptr(c(t.centroid, np.float64))
        prm = _lib.Params(gamma, cfl, PRECISIONS[precision], COMPATS[compat], int(positivity), device,
                          RIEMANNS[riemann], CORRS[corr], int(viscous), float(Re), float(Pr), float(br2_eta))
        h = C.c_void_p()
        check(lib().zj_create(C.byref(h), C.byref(d), C.byref(prm)))
        self._h = h
        self.precision = precision

    @classmethod
    def from_mesh(cls, mesh, p, periodic=True, **kw):
        t = build_tables(mesh.xy, mesh.elem_nodes, p, periodic=(mesh.Lx, mesh.Ly) if periodic else None)
        return cls(t, **kw)

    def close(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            self._chk(lib().zj_destroy(h))
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    def _chk(self, rc):
        return check(rc, self._h)

    def set_state(self, U):
        U = np.ascontiguousarray(U, np.float64).reshape(-1)
        assert U.size == self.n
        self._chk(lib().zj_set_state(self._h, U))

    def get_state(self):
        out = np.zeros(self.n)
        self._chk(lib().zj_get_state(self._h, out))
        return out

    def save_restart(self, path, time=0.0, iteration=0):
        """Device state to a restart file (zj_restart_save)."""
        from . import io
        io.save_handle("quad", self._h, path, time, iteration)

    def load_restart(self, path):
        """Restart file to the device state (zj_restart_load); returns (time, iteration)."""
        from . import io
        return io.load_handle("quad", self._h, path)

    # device-pointer forms (ints from e.g. torch's data_ptr(); stream = hipStream_t or None)
    def set_state_device(self, ptr, stream=None):
        self._chk(lib().zj_set_state_device(self._h, ptr, stream))

    def get_state_device(self, ptr, stream=None):
        self._chk(lib().zj_get_state_device(self._h, ptr, stream))

    def state_slots(self, n):
        """Allocates n device-resident state slots (zj_state_slots)."""
        self._chk(lib().zj_state_slots(self._h, int(n)))

    def state_save(self, slot, stream=None):
        self._chk(lib().zj_state_save(self._h, int(slot), stream))

    def state_load(self, slot, stream=None):
        self._chk(lib().zj_state_load(self._h, int(slot), stream))

    def adjoint_set_device(self, ptr, stream=None):
        self._chk(lib().zj_adjoint_set_device(self._h, ptr, stream))

    def adjoint_seed_kinetic_energy(self, weights_ptr, stream=None):
        """J = sum of w (m_x^2 + m_y^2) / (2 rho) of the current state, computed
        on the GPU, and the adjoint seeded with dJ/dU there
        (zj_adjoint_seed_kinetic_energy); weights_ptr: device [E][S]."""
        J = C.c_double()
        self._chk(lib().zj_adjoint_seed_kinetic_energy(self._h, weights_ptr, C.byref(J), stream))
        return J.value

    def adjoint_get_device(self, ptr, stream=None):
        self._chk(lib().zj_adjoint_get_device(self._h, ptr, stream))

    def residual_vjp_device(self, w_ptr, out_ptr, stream=None):
        self._chk(lib().zj_residual_vjp_device(self._h, w_ptr, out_ptr, stream))

    def residual_device(self, out_ptr, stream=None):
        """dU/dt into a device AoS buffer [E][S][4]."""
        self._chk(lib().zj_residual_device(self._h, out_ptr, stream))

    def residual(self, fetch=True, stream=None):
        out = np.zeros(self.n) if fetch else None
        self._chk(lib().zj_residual(self._h, None if out is None else out.ctypes.data, stream))
        return out

    def rk3_step(self, dt, stream=None):
        self._chk(lib().zj_rk3_step(self._h, float(dt), stream))

    def rk3_stage(self, stage, dt, stream=None):
        self._chk(lib().zj_rk3_stage(self._h, int(stage), float(dt), stream))

    def synchronize(self):
        self._chk(lib().zj_synchronize(self._h))

    def time_stages(self, dt, steps, stream=None):
        """zj_time_stages: advances `steps` steps, each stage launched on its
        own; returns each stage kernel's mean execution time (ms, list of 3)."""
        out = np.zeros(3)
        self._chk(lib().zj_time_stages(self._h, float(dt), int(steps), out, stream))
        return [float(x) for x in out]

    def dt_cfl(self, cfl):
        v = C.c_double()
        self._chk(lib().zj_dt_cfl(self._h, float(cfl), C.byref(v)))
        return v.value

    def residual_l2(self):
        v = C.c_double()
        self._chk(lib().zj_residual_l2(self._h, C.byref(v)))
        return v.value

    def count_nonfinite(self):
        v = C.c_int64()
        self._chk(lib().zj_check_finite(self._h, C.byref(v)))
        return v.value

    def set_boundary(self, bc_type, bc_face_map, bc_data):
        """zj_set_boundary: new BC tables in the layout of zj_create (per face
        of the creation face list; 8 doubles per bc_data entry)."""
        bt = np.ascontiguousarray(bc_type, np.int32)
        bm = np.ascontiguousarray(bc_face_map, np.int32)
        bd = np.ascontiguousarray(bc_data, np.float64).reshape(-1)
        self._chk(lib().zj_set_boundary(self._h, bt.ctypes.data, bm.ctypes.data, bd.ctypes.data, bd.size // 8))

    def set_tile(self, epb):
        self._chk(lib().zj_set_tile(self._h, int(epb)))

    # ---- discrete adjoint
    def residual_vjp(self, w):
        """(dR/dU)^T w at the current state."""
        w = np.ascontiguousarray(w, np.float64).reshape(-1)
        assert w.size == self.n
        out = np.empty(self.n)
        self._chk(lib().zj_residual_vjp(self._h, w, out))
        return out

    def adjoint_set(self, ubar):
        ubar = np.ascontiguousarray(ubar, np.float64).reshape(-1)
        assert ubar.size == self.n
        self._chk(lib().zj_adjoint_set(self._h, ubar))

    def adjoint_get(self):
        out = np.empty(self.n)
        self._chk(lib().zj_adjoint_get(self._h, out))
        return out

    def forward_checkpointed(self, dts, every, stream=None):
        """len(dts) steps, U^{c*every} kept in state slot c (zj_forward_checkpointed)."""
        d = np.ascontiguousarray(dts, np.float64)
        self._chk(lib().zj_forward_checkpointed(self._h, d, d.size, int(every), stream))

    def adjoint_tape_segments(self, n):
        """Trailing segments that forward_checkpointed records on the device
        tape as it steps, so that adjoint_sweep need not recompute them
        (zj_adjoint_tape_segments; default 1, 0 = none)."""
        self._chk(lib().zj_adjoint_tape_segments(self._h, int(n)))

    def adjoint_sweep(self, dts, every, want_dt=True, stream=None):
        """Every step backwards from the seeded cotangent (zj_adjoint_sweep);
        returns dJ/d(dt_n) per step, or None with want_dt=False (asynchronous)."""
        d = np.ascontiguousarray(dts, np.float64)
        out = np.zeros(d.size) if want_dt else None
        self._chk(lib().zj_adjoint_sweep(self._h, d, d.size, int(every), None if out is None else out.ctypes.data,
                                         stream))
        return None if out is None else [float(x) for x in out]

    def adjoint_get_re(self):
        """dJ/dRe accumulated since adjoint_set (Navier-Stokes terms; 0 without)."""
        v = C.c_double()
        self._chk(lib().zj_adjoint_get_re(self._h, C.byref(v)))
        return v.value

    def dt_cfl_vjp(self, cfl, dt_bar, stream=None):
        """Adds dt_bar * d(dt_cfl)/dU at the held state to the held cotangent
        (zj_dt_cfl_vjp)."""
        self._chk(lib().zj_dt_cfl_vjp(self._h, float(cfl), float(dt_bar), stream))

    def rk3_step_vjp(self, dt, stream=None, want_dt=True):
        """Maps the held cotangent of U^{n+1} to that of U^n (the state set on
        the handle); returns dJ/d(dt) (synchronising), or None with
        want_dt=False (asynchronous)."""
        if not want_dt:
            self._chk(lib().zj_rk3_step_vjp(self._h, float(dt), None, stream))
            return None
        dtb = C.c_double()
        self._chk(lib().zj_rk3_step_vjp(self._h, float(dt), C.byref(dtb), stream))
        return dtb.value

    def autotune(self, reps=20):
        """Picks the fastest tile and XCD order; returns (epb, xcd_groups, ms)."""
        e, x, ms = C.c_int(), C.c_int(), C.c_double()
        self._chk(lib().zj_autotune(self._h, int(reps), C.byref(e), C.byref(x), C.byref(ms)))
        return e.value, x.value, ms.value

    def set_xcd_groups(self, groups):
        self._chk(lib().zj_set_xcd_groups(self._h, int(groups)))

    def set_geometry_mode(self, mode):
        """-1 query, 0 per-point metric, 1 per-element metric (affine meshes);
        returns the mode in use."""
        a = C.c_int()
        self._chk(lib().zj_set_geometry_mode(self._h, int(mode), C.byref(a)))
        return a.value

    def info(self):
        n = C.c_int64(); e = C.c_int(); s = C.c_int()
        self._chk(lib().zj_get_info(self._h, C.byref(n), C.byref(e), C.byref(s)))
        ce = C.c_int64(); ro = C.c_int()
        self._chk(lib().zj_get_tiling(self._h, C.byref(ce), C.byref(ro)))
        return dict(n_dof=n.value, elems_per_block=e.value, n_sp=s.value,
                    affine=bool(self.set_geometry_mode(-1)), cross_tile_edges=ce.value,
                    reordered=bool(ro.value))


class Batch:
    """Many Solvers stepped together (zj_batch_*).  Keeps its members alive."""

    def __init__(self, members):
        self.members = list(members)
        arr = (C.c_void_p * len(self.members))(*[m._h.value for m in self.members])
        h = C.c_void_p()
        rc = lib().zj_batch_create(C.byref(h), arr, len(self.members))
        if rc != 0:
            msg = lib().zj_batch_last_error(None)
            raise ZJError(rc, msg.decode() if msg else "")
        self._h = h

    def _chk(self, rc):
        if rc != 0:
            msg = lib().zj_batch_last_error(self._h)
            raise ZJError(rc, msg.decode() if msg else "")

    def rk3_step(self, dts, stream=None):
        """One SSP-RK3 step of every member, dts[k] for member k.  The batch
        launches read the members' device buffers, so every member must
        still be open."""
        d = np.ascontiguousarray(dts, np.float64).reshape(-1)
        if d.size != len(self.members):
            raise ValueError(f"{d.size} time steps for {len(self.members)} members")
        if not all(getattr(m, "_h", None) and m._h.value for m in self.members):
            raise RuntimeError("a member of the batch has been closed")
        self._chk(lib().zj_batch_rk3_step(self._h, d, stream))

    def rk3_step_vjp(self, dts, stream=None, want_dt=False):
        """One SSP-RK3 step backwards for every member (each holding its
        state U^n and an adjoint seed); returns dJ/d(dts) per member when
        want_dt, else None."""
        d = np.ascontiguousarray(dts, np.float64).reshape(-1)
        if d.size != len(self.members):
            raise ValueError(f"{d.size} time steps for {len(self.members)} members")
        if not all(getattr(m, "_h", None) and m._h.value for m in self.members):
            raise RuntimeError("a member of the batch has been closed")
        out = np.zeros(d.size) if want_dt else None
        self._chk(lib().zj_batch_rk3_step_vjp(self._h, d, out.ctypes.data if want_dt else None, stream))
        return out

    def synchronize(self):
        self._chk(lib().zj_batch_synchronize(self._h))

    def set_grid(self, policy):
        self._chk(lib().zj_batch_set_grid(self._h, int(policy)))

    def set_launch(self, single):
        """True: one launch per stage over every member; False: one per (order, metric) group."""
        self._chk(lib().zj_batch_set_launch(self._h, int(bool(single))))

    def autotune(self, dts, reps=5):
        """Times each launch form and grid policy, keeps the fastest and leaves
        every member's state unchanged; returns (single, policy, ms per step)."""
        d = np.ascontiguousarray(dts, np.float64).reshape(-1)
        if d.size != len(self.members):
            raise ValueError(f"{d.size} time steps for {len(self.members)} members")
        sg, pol, ms = C.c_int(), C.c_int(), C.c_double()
        self._chk(lib().zj_batch_autotune(self._h, d, int(reps), C.byref(sg), C.byref(pol), C.byref(ms)))
        return bool(sg.value), pol.value, ms.value

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            lib().zj_batch_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
