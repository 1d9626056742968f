"""ctypes binding of libzhijian_amd.so (include/zhijian_amd.h).

The library is built in-tree by ``python -m zhijian_amd.build``.  There is no
fallback: every entry point raises if the library is missing or a call fails.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libzhijian_amd.so")

ZJ_OK = 0
ZJ_PREC_FP64 = 0
ZJ_PREC_FP32BULK = 1
ZJ_COMPAT_REFERENCE = 0
ZJ_COMPAT_TEXTBOOK = 1

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_lp = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")


class ZJError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"zhijian_amd error {code}: {msg}")
        self.code = code


class MeshDesc(C.Structure):
    _fields_ = [("poly_order", C.c_int), ("n_elem", C.c_int), ("n_faces", C.c_int),
                ("diff_xi", C.c_void_p), ("diff_eta", C.c_void_p), ("interp", C.c_void_p),
                ("corr", C.c_void_p), ("J", C.c_void_p), ("Jinv", C.c_void_p),
                ("normals", C.c_void_p), ("face_left_elem", C.c_void_p),
                ("face_left_local", C.c_void_p), ("face_right_elem", C.c_void_p),
                ("face_right_local", C.c_void_p), ("bc_type", C.c_void_p),
                ("bc_face_map", C.c_void_p), ("bc_data", C.c_void_p), ("n_bc_data", C.c_int),
                ("elem_area", C.c_void_p), ("edge_half_len", C.c_void_p), ("elem_centroid", C.c_void_p)]


class TriDesc(C.Structure):
    _fields_ = [("poly_order", C.c_int), ("n_nodes", C.c_int64), ("xy", C.c_void_p), ("n_elem", C.c_int),
                ("tri", C.c_void_p), ("n_faces", C.c_int), ("face_left_elem", C.c_void_p),
                ("face_left_local", C.c_void_p), ("face_right_elem", C.c_void_p),
                ("face_right_local", C.c_void_p), ("normals", C.c_void_p), ("bc_type", C.c_void_p),
                ("bc_face_map", C.c_void_p), ("bc_data", C.c_void_p), ("n_bc_data", C.c_int),
                ("reorder", C.c_int)]


class Params(C.Structure):
    _fields_ = [("gamma", C.c_double), ("cfl", C.c_double), ("precision", C.c_int),
                ("compat", C.c_int), ("positivity", C.c_int), ("device", C.c_int),
                ("riemann", C.c_int), ("corr", C.c_int), ("viscous", C.c_int), ("Re", C.c_double),
                ("Pr", C.c_double), ("br2_eta", C.c_double)]


class RestartInfo(C.Structure):
    _fields_ = [("n_elem", C.c_int64), ("n_values", C.c_int64), ("time", C.c_double), ("iteration", C.c_int),
                ("poly_order", C.c_int), ("gamma", C.c_double)]


class VisMesh(C.Structure):
    _fields_ = [("n_nodes", C.c_int64), ("xy", C.c_void_p), ("n_elem", C.c_int64), ("elem_type", C.c_void_p),
                ("elem_order", C.c_void_p), ("elem_node_start", C.c_void_p), ("elem_nodes", C.c_void_p)]


_lib = None

# every exported symbol of include/zhijian_amd.h
EXPORTS = ("zj_create", "zj_destroy", "zj_last_error", "zj_version", "zj_set_state", "zj_get_state",
           "zj_set_state_device", "zj_get_state_device", "zj_residual", "zj_rk3_step",
           "zj_rk3_stage", "zj_dt_cfl", "zj_residual_l2", "zj_check_finite", "zj_set_tile",
           "zj_get_info", "zj_quad_operators", "zj_build_faces", "zj_pair_periodic",
           "zj_face_normals", "zj_elem_areas", "zj_geometry", "zj_copy_probe", "zj_set_boundary",
           "zj_rk3_combine", "zj_euler_update", "zj_synchronize", "zj_set_geometry_mode",
           "zj_set_xcd_groups", "zj_adjoint_set", "zj_adjoint_set_device", "zj_adjoint_get",
           "zj_adjoint_get_device", "zj_rk3_step_vjp", "zj_dt_cfl_vjp", "zj_residual_vjp", "zj_residual_vjp_device",
           "zj_residual_device", "zj_batch_create", "zj_batch_destroy", "zj_batch_last_error",
           "zj_batch_rk3_step", "zj_batch_rk3_step_vjp", "zj_batch_synchronize", "zj_batch_set_grid", "zj_autotune",
           "zj_batch_set_launch", "zj_batch_autotune", "zj_state_slots", "zj_state_save", "zj_state_load",
           "zj_gmsh_read", "zj_gmsh_error", "zj_gmsh_free", "zj_gmsh_counts", "zj_gmsh_nodes",
           "zj_gmsh_elements", "zj_gmsh_edge_tag", "zj_gmsh_group", "zj_get_tiling",
           "zj_tri_create", "zj_tri_destroy", "zj_tri_last_error", "zj_tri_set_state", "zj_tri_get_state",
           "zj_tri_set_state_device", "zj_tri_get_state_device", "zj_tri_residual", "zj_tri_residual_device",
           "zj_tri_rk3_step", "zj_tri_dt_cfl", "zj_tri_residual_l2", "zj_tri_synchronize",
           "zj_tri_set_xcd_groups", "zj_tri_get_info", "zj_tri_operators", "zj_get_scheme",
           "zj_tri_get_scheme", "zj_io_last_error", "zj_restart_write", "zj_restart_query", "zj_restart_read",
           "zj_restart_save", "zj_restart_load", "zj_tri_restart_save", "zj_tri_restart_load", "zj_vtk_write",
           "zj_pvtu_write", "zj_gmsh_periodic", "zj_pair_periodic_nodes",
           "zj_tri_lift", "zj_time_stages", "zj_tri_time_stages", "zj_adjoint_get_re",
           "zj_forward_checkpointed", "zj_adjoint_sweep", "zj_adjoint_tape_segments",
           "zj_adjoint_seed_kinetic_energy",
           "zj_tri_adjoint_set", "zj_tri_adjoint_set_device", "zj_tri_adjoint_get", "zj_tri_adjoint_get_device",
           "zj_tri_rk3_step_vjp", "zj_tri_residual_vjp", "zj_tri_residual_vjp_device",
           # the gpu:: free functions (zhijian_amd.gpu_ops binds them)
           "zj_gpu_set_operator_engine", "zj_gpu_interpolate_to_flux_points", "zj_gpu_reference_gradients",
           "zj_gpu_transform_gradients", "zj_gpu_inviscid_flux_sp", "zj_gpu_riemann_flux",
           "zj_gpu_riemann_flux_with_bc", "zj_gpu_viscous_flux_sp", "zj_gpu_interior_flux_at_face",
           "zj_gpu_inviscid_flux_at_fp", "zj_gpu_flux_difference", "zj_gpu_flux_diff_for_fr",
           "zj_gpu_flux_divergence", "zj_gpu_scatter_flux_to_elements", "zj_gpu_apply_fr_correction",
           "zj_gpu_apply_boundary_conditions", "zj_gpu_enforce_positivity", "zj_gpu_compute_time_step",
           "zj_gpu_residual_norm", "zj_gpu_residual_norm_noalloc", "zj_gpu_set_initial_condition",
           "zj_gpu_copy_solution", "zj_gpu_scale_array", "zj_gpu_axpy")


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m zhijian_amd.build`")
        L = C.CDLL(LIB_PATH)
        H = C.c_void_p
        L.zj_create.argtypes = [C.POINTER(H), C.POINTER(MeshDesc), C.POINTER(Params)]
        L.zj_destroy.argtypes = [H]
        L.zj_last_error.argtypes = [H]
        L.zj_last_error.restype = C.c_char_p
        L.zj_set_state.argtypes = [H, _dp]
        L.zj_get_state.argtypes = [H, _dp]
        L.zj_set_state_device.argtypes = [H, C.c_void_p, C.c_void_p]
        L.zj_get_state_device.argtypes = [H, C.c_void_p, C.c_void_p]
        L.zj_residual.argtypes = [H, C.c_void_p, C.c_void_p]
        L.zj_rk3_step.argtypes = [H, C.c_double, C.c_void_p]
        L.zj_rk3_stage.argtypes = [H, C.c_int, C.c_double, C.c_void_p]
        L.zj_dt_cfl.argtypes = [H, C.c_double, C.POINTER(C.c_double)]
        L.zj_residual_l2.argtypes = [H, C.POINTER(C.c_double)]
        L.zj_check_finite.argtypes = [H, C.POINTER(C.c_int64)]
        L.zj_set_tile.argtypes = [H, C.c_int]
        L.zj_set_geometry_mode.argtypes = [H, C.c_int, C.POINTER(C.c_int)]
        L.zj_set_xcd_groups.argtypes = [H, C.c_int]
        L.zj_adjoint_set.argtypes = [H, _dp]
        L.zj_adjoint_set_device.argtypes = [H, C.c_void_p, C.c_void_p]
        L.zj_adjoint_get.argtypes = [H, _dp]
        L.zj_adjoint_get_device.argtypes = [H, C.c_void_p, C.c_void_p]
        L.zj_rk3_step_vjp.argtypes = [H, C.c_double, C.POINTER(C.c_double), C.c_void_p]
        L.zj_dt_cfl_vjp.argtypes = [H, C.c_double, C.c_double, C.c_void_p]
        L.zj_residual_vjp.argtypes = [H, _dp, _dp]
        L.zj_residual_vjp_device.argtypes = [H, C.c_void_p, C.c_void_p, C.c_void_p]
        L.zj_residual_device.argtypes = [H, C.c_void_p, C.c_void_p]
        L.zj_batch_create.argtypes = [C.POINTER(H), C.POINTER(H), C.c_int]
        L.zj_batch_destroy.argtypes = [H]
        L.zj_batch_last_error.argtypes = [H]
        L.zj_batch_last_error.restype = C.c_char_p
        L.zj_batch_rk3_step.argtypes = [H, _dp, C.c_void_p]
        L.zj_batch_rk3_step_vjp.argtypes = [H, _dp, C.c_void_p, C.c_void_p]
        L.zj_batch_synchronize.argtypes = [H]
        L.zj_batch_set_grid.argtypes = [H, C.c_int]
        L.zj_batch_set_launch.argtypes = [H, C.c_int]
        L.zj_state_slots.argtypes = [H, C.c_int]
        L.zj_state_save.argtypes = [H, C.c_int, C.c_void_p]
        L.zj_state_load.argtypes = [H, C.c_int, C.c_void_p]
        L.zj_batch_autotune.argtypes = [H, _dp, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                        C.POINTER(C.c_double)]
        L.zj_autotune.argtypes = [H, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_double)]
        L.zj_gmsh_read.argtypes = [C.c_char_p, C.POINTER(H)]
        L.zj_gmsh_error.restype = C.c_char_p
        L.zj_gmsh_free.argtypes = [H]
        L.zj_gmsh_counts.argtypes = [H, C.POINTER(C.c_int64), C.POINTER(C.c_int), C.POINTER(C.c_int),
                                     C.POINTER(C.c_int)]
        L.zj_gmsh_nodes.argtypes = [H, C.c_void_p]
        L.zj_gmsh_elements.argtypes = [H, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int]
        L.zj_gmsh_edge_tag.argtypes = [H, C.c_int64, C.c_int64, C.POINTER(C.c_int)]
        L.zj_gmsh_group.argtypes = [H, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int), C.c_char_p, C.c_int]
        L.zj_gmsh_periodic.argtypes = [H, C.POINTER(C.c_int64), C.c_void_p, C.c_void_p]
        L.zj_pair_periodic_nodes.argtypes = [C.POINTER(C.c_int), _ip, _ip, _ip, _ip, C.c_int, C.c_int, C.c_int, _lp, _dp,
                                             C.c_int64, C.c_int64, C.c_void_p, C.c_void_p, C.c_double]
        L.zj_get_info.argtypes = [H, C.POINTER(C.c_int64), C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.zj_get_tiling.argtypes = [H, C.POINTER(C.c_int64), C.POINTER(C.c_int)]
        L.zj_quad_operators.argtypes = [C.c_int] + [_dp] * 7
        L.zj_build_faces.argtypes = [C.c_int, C.c_int, C.c_int, _lp, C.POINTER(C.c_int), _ip, _ip, _ip, _ip]
        L.zj_pair_periodic.argtypes = [C.POINTER(C.c_int), _ip, _ip, _ip, _ip, C.c_int, C.c_int, _lp, _dp,
                                       C.c_double, C.c_double, C.c_double]
        L.zj_face_normals.argtypes = [C.c_int, _ip, _ip, C.c_int, C.c_int, _lp, _dp, _dp]
        L.zj_elem_areas.argtypes = [C.c_int, C.c_int, C.c_int, _lp, _dp, _dp]
        L.zj_set_boundary.argtypes = [H, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int]
        L.zj_rk3_combine.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64,
                                     C.c_double, C.c_void_p]
        L.zj_euler_update.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.c_double, C.c_void_p]
        L.zj_synchronize.argtypes = [H]
        L.zj_copy_probe.argtypes = [C.c_int, C.c_int64, C.c_int, C.POINTER(C.c_double)]
        L.zj_geometry.argtypes = [C.c_int, C.c_int, C.c_int, _lp, _dp, C.c_int, _dp, _dp, _dp, _dp, C.c_void_p]
        L.zj_tri_create.argtypes = [C.POINTER(H), C.POINTER(TriDesc), C.POINTER(Params)]
        L.zj_tri_destroy.argtypes = [H]
        L.zj_tri_last_error.argtypes = [H]
        L.zj_tri_last_error.restype = C.c_char_p
        L.zj_tri_set_state.argtypes = [H, _dp]
        L.zj_tri_get_state.argtypes = [H, _dp]
        L.zj_tri_set_state_device.argtypes = [H, C.c_void_p, C.c_void_p]
        L.zj_tri_get_state_device.argtypes = [H, C.c_void_p, C.c_void_p]
        L.zj_tri_residual.argtypes = [H, C.c_void_p, C.c_void_p]
        L.zj_tri_residual_device.argtypes = [H, C.c_void_p, C.c_void_p]
        L.zj_tri_rk3_step.argtypes = [H, C.c_double, C.c_void_p]
        L.zj_tri_time_stages.argtypes = [H, C.c_double, C.c_int, _dp, C.c_void_p]
        L.zj_time_stages.argtypes = [H, C.c_double, C.c_int, _dp, C.c_void_p]
        L.zj_tri_dt_cfl.argtypes = [H, C.c_double, C.POINTER(C.c_double)]
        L.zj_tri_residual_l2.argtypes = [H, C.POINTER(C.c_double)]
        L.zj_tri_synchronize.argtypes = [H]
        L.zj_tri_set_xcd_groups.argtypes = [H, C.c_int]
        L.zj_tri_get_info.argtypes = [H, C.POINTER(C.c_int64), C.POINTER(C.c_int), C.POINTER(C.c_int),
                                      C.POINTER(C.c_int64)]
        L.zj_adjoint_get_re.argtypes = [H, C.POINTER(C.c_double)]
        L.zj_forward_checkpointed.argtypes = [H, _dp, C.c_int, C.c_int, C.c_void_p]
        L.zj_adjoint_sweep.argtypes = [H, _dp, C.c_int, C.c_int, C.c_void_p, C.c_void_p]
        L.zj_adjoint_tape_segments.argtypes = [H, C.c_int]
        L.zj_adjoint_seed_kinetic_energy.argtypes = [H, C.c_void_p, C.POINTER(C.c_double), C.c_void_p]
        L.zj_tri_adjoint_set.argtypes = [H, _dp]
        L.zj_tri_adjoint_get.argtypes = [H, _dp]
        L.zj_tri_adjoint_set_device.argtypes = [H, C.c_void_p, C.c_void_p]
        L.zj_tri_adjoint_get_device.argtypes = [H, C.c_void_p, C.c_void_p]
        L.zj_tri_rk3_step_vjp.argtypes = [H, C.c_double, C.POINTER(C.c_double), C.c_void_p]
        L.zj_tri_residual_vjp.argtypes = [H, _dp, _dp]
        L.zj_tri_residual_vjp_device.argtypes = [H, C.c_void_p, C.c_void_p, C.c_void_p]
        L.zj_tri_operators.argtypes = [C.c_int] + [C.c_void_p] * 7
        L.zj_tri_lift.argtypes = [C.c_int, C.c_int, C.c_void_p]
        L.zj_get_scheme.argtypes = [H, C.POINTER(C.c_int), C.POINTER(C.c_double)]
        L.zj_tri_get_scheme.argtypes = [H, C.POINTER(C.c_int), C.POINTER(C.c_double)]
        L.zj_io_last_error.argtypes = []
        L.zj_io_last_error.restype = C.c_char_p
        L.zj_restart_write.argtypes = [C.c_char_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_double, C.c_int,
                                       C.c_int, C.c_double]
        L.zj_restart_query.argtypes = [C.c_char_p, C.POINTER(RestartInfo)]
        L.zj_restart_read.argtypes = [C.c_char_p, C.c_int64, C.c_void_p, C.c_int64, C.c_void_p]
        for f in ("zj_restart_save", "zj_tri_restart_save"):
            getattr(L, f).argtypes = [H, C.c_char_p, C.c_double, C.c_int]
        for f in ("zj_restart_load", "zj_tri_restart_load"):
            getattr(L, f).argtypes = [H, C.c_char_p, C.POINTER(C.c_double), C.POINTER(C.c_int)]
        L.zj_vtk_write.argtypes = [C.c_char_p, C.POINTER(VisMesh), C.c_void_p, C.c_void_p, C.c_int, C.c_double,
                                   C.c_int, C.c_int]
        L.zj_pvtu_write.argtypes = [C.c_char_p, C.c_int, C.POINTER(C.c_char_p)]
        _lib = L
    return _lib


def check(rc, h=None):
    if rc != ZJ_OK:
        msg = lib().zj_last_error(h)
        raise ZJError(rc, msg.decode() if msg else "")
    return rc


# ----------------------------------------------------------- host tables
def quad_operators(p: int) -> dict:
    n = p + 1; S = n * n
    out = dict(sol_xi=np.zeros(S), sol_eta=np.zeros(S), qw=np.zeros(S),
               diff_xi=np.zeros(S * S), diff_eta=np.zeros(S * S), interp=np.zeros(4 * n * S),
               corr=np.zeros(4 * S))
    check(lib().zj_quad_operators(p, out["sol_xi"], out["sol_eta"], out["qw"], out["diff_xi"],
                                  out["diff_eta"], out["interp"], out["corr"]))
    return out


def build_faces(elem_nodes: np.ndarray, nv: int | None = None):
    en = np.ascontiguousarray(elem_nodes, np.int64)
    E, stride = en.shape
    nv = stride if nv is None else nv
    cap = max(E * nv, 1)
    arrs = [np.zeros(cap, np.int32) for _ in range(4)]
    nf = C.c_int(0)
    check(lib().zj_build_faces(E, nv, stride, en.reshape(-1), C.byref(nf), *arrs))
    return np.stack([a[: nf.value] for a in arrs], axis=1)


def pair_periodic(faces: np.ndarray, elem_nodes, xy, Lx, Ly, tol=None, nv=None):
    en = np.ascontiguousarray(elem_nodes, np.int64)
    stride = en.shape[1]
    nv = stride if nv is None else nv
    arrs = [np.ascontiguousarray(faces[:, k], np.int32).copy() for k in range(4)]
    nf = C.c_int(faces.shape[0])
    tol = 1e-9 * max(Lx, Ly) if tol is None else tol
    check(lib().zj_pair_periodic(C.byref(nf), *arrs, nv, stride, en.reshape(-1),
                                 np.ascontiguousarray(xy, np.float64).reshape(-1), Lx, Ly, tol))
    return np.stack([a[: nf.value] for a in arrs], axis=1)


def pair_periodic_nodes(faces: np.ndarray, elem_nodes, xy, links, tol=None, nv=None):
    """zj_pair_periodic_nodes: links [n][2] (slave node, master node); tol
    defaults to 1e-9 of the mesh's extent."""
    en = np.ascontiguousarray(elem_nodes, np.int64)
    stride = en.shape[1]
    nv = stride if nv is None else nv
    xy = np.ascontiguousarray(xy, np.float64).reshape(-1, 2)
    if tol is None:
        tol = 1e-9 * float(np.ptp(xy, axis=0).max()) if xy.size else 0.0
    lk = np.ascontiguousarray(np.asarray(links, np.int64).reshape(-1, 2))
    sl, ms = np.ascontiguousarray(lk[:, 0]), np.ascontiguousarray(lk[:, 1])
    arrs = [np.ascontiguousarray(faces[:, k], np.int32).copy() for k in range(4)]
    nf = C.c_int(faces.shape[0])
    check(lib().zj_pair_periodic_nodes(C.byref(nf), *arrs, nv, stride, en.shape[0], en.reshape(-1), xy.reshape(-1),
                                       xy.shape[0],
                                       lk.shape[0], sl.ctypes.data, ms.ctypes.data, float(tol)))
    return np.stack([a[: nf.value] for a in arrs], axis=1)


def face_normals(faces, elem_nodes, xy, nv=None):
    en = np.ascontiguousarray(elem_nodes, np.int64)
    stride = en.shape[1]
    nv = stride if nv is None else nv
    F = faces.shape[0]
    out = np.zeros(max(2 * F, 1))
    check(lib().zj_face_normals(F, np.ascontiguousarray(faces[:, 0], np.int32),
                                np.ascontiguousarray(faces[:, 1], np.int32), nv, stride,
                                en.reshape(-1), np.ascontiguousarray(xy, np.float64).reshape(-1), out))
    return out[: 2 * F].reshape(F, 2)


def elem_areas(elem_nodes, xy, nv=None):
    en = np.ascontiguousarray(elem_nodes, np.int64)
    E, stride = en.shape
    nv = stride if nv is None else nv
    out = np.zeros(max(E, 1))
    check(lib().zj_elem_areas(E, nv, stride, en.reshape(-1), np.ascontiguousarray(xy, np.float64).reshape(-1), out))
    return out[:E]


def geometry(elem_nodes, xy, ops, node_count=None, want_pos=False):
    en = np.ascontiguousarray(elem_nodes, np.int64)
    E, stride = en.shape
    nc = stride if node_count is None else node_count
    S = ops["sol_xi"].shape[0]
    J = np.zeros(max(E * S, 1)); Jinv = np.zeros(max(E * S * 4, 1))
    pos = np.zeros(max(E * S * 2, 1)) if want_pos else None
    check(lib().zj_geometry(E, nc, stride, en.reshape(-1), np.ascontiguousarray(xy, np.float64).reshape(-1),
                            S, ops["sol_xi"], ops["sol_eta"], J, Jinv,
                            None if pos is None else pos.ctypes.data))
    out = (J[: E * S], Jinv[: E * S * 4])
    if want_pos:
        out = out + (pos[: E * S * 2].reshape(E, S, 2),)
    return out


def tri_operators(p: int) -> dict:
    """Triangle operators of order p as the library builds them."""
    np_, nf = (p + 1) * (p + 2) // 2, 3 * (p + 1)
    out = dict(sp=np.zeros((np_, 2)), fp=np.zeros((nf, 2)), wg=np.zeros(p + 1), Dr=np.zeros((np_, np_)),
               Ds=np.zeros((np_, np_)), interp=np.zeros((nf, np_)), lift=np.zeros((np_, nf)))
    check(lib().zj_tri_operators(p, *(out[k].ctypes.data for k in ("sp", "fp", "wg", "Dr", "Ds", "interp", "lift"))))
    return out


def tri_lift(p: int, corr: int) -> np.ndarray:
    """zj_tri_lift: the lifting matrix [Np][3(p+1)] of correction family corr."""
    out = np.zeros(((p + 1) * (p + 2) // 2, 3 * (p + 1)))
    check(lib().zj_tri_lift(p, corr, out.ctypes.data))
    return out


def copy_probe(nbytes: int, reps: int = 20, device: int = 0) -> float:
    """Achievable HBM bandwidth (GB/s) of a plain copy kernel."""
    v = C.c_double()
    check(lib().zj_copy_probe(device, int(nbytes), int(reps), C.byref(v)))
    return v.value
