"""ctypes binding of the gpu:: free functions (include/zhijian_amd.h, zj_gpu_*),
the reference's include/gpu/kernels.hpp:17-266 over raw device pointers.

Arguments are device addresses (ints, e.g. ``hiprt.DeviceArray.ptr``) and
sizes, in the reference's order; ``stream`` is a hipStream_t address or None.
Every call raises ``ZJError`` with the library's message when it fails.
"""
from __future__ import annotations

import ctypes as C

from zhijian_amd._lib import ZJError, lib

P, I, D = C.c_void_p, C.c_int, C.c_double

# name -> argument types, in declaration order (the trailing void* stream included)
_SIG = {
    "zj_gpu_set_operator_engine": [I],
    "zj_gpu_interpolate_to_flux_points": [P, P, P, I, I, I, I, P],
    "zj_gpu_reference_gradients": [P, P, P, P, P, I, I, P],
    "zj_gpu_transform_gradients": [P, P, P, P, P, I, I, P],
    "zj_gpu_inviscid_flux_sp": [P, P, P, D, I, I, P],
    "zj_gpu_riemann_flux": [P, P, P, P, D, I, I, I, P],
    "zj_gpu_riemann_flux_with_bc": [P, P, P, P, P, P, P, P, P, P, D, I, I, I, I, P],
    "zj_gpu_viscous_flux_sp": [P, P, P, P, P, D, D, D, I, I, P],
    "zj_gpu_interior_flux_at_face": [P, P, P, P, P, D, I, I, I, P],
    "zj_gpu_inviscid_flux_at_fp": [P, P, P, P, P, D, I, I, I, P],
    "zj_gpu_flux_difference": [P, P, P, I, P],
    "zj_gpu_flux_diff_for_fr": [P, P, P, P, P, P, P, P, D, I, I, I, P],
    "zj_gpu_flux_divergence": [P, P, P, P, P, P, P, I, I, P],
    "zj_gpu_scatter_flux_to_elements": [P, P, P, P, P, P, I, I, I, P],
    "zj_gpu_apply_fr_correction": [P, P, P, P, I, I, I, I, P],
    "zj_gpu_apply_boundary_conditions": [P, P, P, P, P, P, D, I, I, I, P],
    "zj_gpu_enforce_positivity": [P, I, P],
    "zj_gpu_compute_time_step": [P, P, P, D, D, I, I, C.POINTER(D), P],
    "zj_gpu_residual_norm": [P, I, I, C.POINTER(D), P],
    "zj_gpu_residual_norm_noalloc": [P, I, I, C.POINTER(D), P],
    "zj_gpu_set_initial_condition": [P, D, D, D, D, D, I, I, P],
    "zj_gpu_copy_solution": [P, P, I, P],
    "zj_gpu_scale_array": [P, P, D, I, P],
    "zj_gpu_axpy": [P, P, D, I, P],
}
_bound = False

ENGINE_MFMA = 0
ENGINE_VALU = 1


def _L():
    global _bound
    L = lib()
    if not _bound:
        for name, args in _SIG.items():
            getattr(L, name).argtypes = args
        _bound = True
    return L


def _call(name, *args):
    rc = getattr(_L(), name)(*args)
    if rc != 0:
        raise ZJError(rc, lib().zj_last_error(None).decode())


def set_operator_engine(engine):
    _call("zj_gpu_set_operator_engine", int(engine))


def interpolate_to_flux_points(U_sp, U_fp, interp, n_elem, n_sp, n_edges, n_fp, stream=None):
    _call("zj_gpu_interpolate_to_flux_points", U_sp, U_fp, interp, n_elem, n_sp, n_edges, n_fp, stream)


def compute_reference_gradients(U, dUdxi, dUdeta, diff_xi, diff_eta, n_elem, n_sp, stream=None):
    _call("zj_gpu_reference_gradients", U, dUdxi, dUdeta, diff_xi, diff_eta, n_elem, n_sp, stream)


def transform_gradients(dUdxi, dUdeta, dUdx, dUdy, Jinv, n_elem, n_sp, stream=None):
    _call("zj_gpu_transform_gradients", dUdxi, dUdeta, dUdx, dUdy, Jinv, n_elem, n_sp, stream)


def compute_inviscid_flux_sp(U, Fx, Fy, gamma, n_elem, n_sp, stream=None):
    _call("zj_gpu_inviscid_flux_sp", U, Fx, Fy, gamma, n_elem, n_sp, stream)


def compute_riemann_flux(U_L, U_R, F, normals, gamma, solver, n_faces, n_fp, stream=None):
    _call("zj_gpu_riemann_flux", U_L, U_R, F, normals, gamma, solver, n_faces, n_fp, stream)


def compute_riemann_flux_with_bc(U_fp, F, le, ll, re, rl, bc_type, bc_data, bc_face_map, normals, gamma, solver,
                                 n_faces, n_fp, n_faces_per_elem, stream=None):
    _call("zj_gpu_riemann_flux_with_bc", U_fp, F, le, ll, re, rl, bc_type, bc_data, bc_face_map, normals, gamma,
          solver, n_faces, n_fp, n_faces_per_elem, stream)


def compute_viscous_flux_sp(U, dUdx, dUdy, Fvx, Fvy, gamma, Re, Pr, n_elem, n_sp, stream=None):
    _call("zj_gpu_viscous_flux_sp", U, dUdx, dUdy, Fvx, Fvy, gamma, Re, Pr, n_elem, n_sp, stream)


def compute_interior_flux_at_face(U_fp, F_int, le, ll, normals, gamma, n_faces, n_fp, n_faces_per_elem,
                                  stream=None):
    _call("zj_gpu_interior_flux_at_face", U_fp, F_int, le, ll, normals, gamma, n_faces, n_fp, n_faces_per_elem,
          stream)


def compute_inviscid_flux_at_fp(U_fp, F_fp, gamma, n_elem, n_edges, n_fp, stream=None):
    _call("zj_gpu_inviscid_flux_at_fp", U_fp, F_fp, None, None, None, gamma, n_elem, n_edges, n_fp, stream)


def compute_flux_difference(F_common, F_int, F_diff, n, stream=None):
    _call("zj_gpu_flux_difference", F_common, F_int, F_diff, n, stream)


def compute_flux_diff_for_fr(U_fp, F_common, F_diff, le, ll, re, rl, normals, gamma, n_faces, n_fp,
                             n_faces_per_elem, stream=None):
    _call("zj_gpu_flux_diff_for_fr", U_fp, F_common, F_diff, le, ll, re, rl, normals, gamma, n_faces, n_fp,
          n_faces_per_elem, stream)


def compute_flux_divergence(Fx, Fy, div_F, diff_xi, diff_eta, Jinv, n_elem, n_sp, stream=None):
    _call("zj_gpu_flux_divergence", Fx, Fy, div_F, diff_xi, diff_eta, Jinv, None, n_elem, n_sp, stream)


def scatter_flux_to_elements(F_face, F_elem, le, ll, re, rl, n_faces, n_fp, n_faces_per_elem, stream=None):
    _call("zj_gpu_scatter_flux_to_elements", F_face, F_elem, le, ll, re, rl, n_faces, n_fp, n_faces_per_elem,
          stream)


def apply_fr_correction(div_F, F_diff, corr, n_elem, n_sp, n_edges, n_fp, stream=None):
    _call("zj_gpu_apply_fr_correction", div_F, F_diff, corr, None, n_elem, n_sp, n_edges, n_fp, stream)


def apply_boundary_conditions(U_fp, elem_idx, local_face, bc_type, bc_data, normals, gamma, n_bc, n_fp,
                              n_faces_per_elem, stream=None):
    _call("zj_gpu_apply_boundary_conditions", U_fp, elem_idx, local_face, bc_type, bc_data, normals, gamma, n_bc,
          n_fp, n_faces_per_elem, stream)


def enforce_positivity(U, n_states, stream=None):
    _call("zj_gpu_enforce_positivity", U, n_states, stream)


def compute_time_step(U, J, h_min, cfl, gamma, n_elem, n_sp, stream=None):
    out = C.c_double()
    _call("zj_gpu_compute_time_step", U, J, h_min, cfl, gamma, n_elem, n_sp, C.byref(out), stream)
    return out.value


def compute_residual_norm(R, n_elem, n_sp, stream=None):
    out = C.c_double()
    _call("zj_gpu_residual_norm", R, n_elem, n_sp, C.byref(out), stream)
    return out.value


def compute_residual_norm_noalloc(R, n_elem, n_sp, stream=None):
    """zj_gpu_residual_norm_noalloc: the same norm with no device allocation."""
    out = C.c_double()
    _call("zj_gpu_residual_norm_noalloc", R, n_elem, n_sp, C.byref(out), stream)
    return out.value


def set_initial_condition(U, rho, u, v, p, gamma, n_elem, n_sp, stream=None):
    _call("zj_gpu_set_initial_condition", U, rho, u, v, p, gamma, n_elem, n_sp, stream)


def copy_solution(dst, src, n, stream=None):
    _call("zj_gpu_copy_solution", dst, src, n, stream)


def scale_array(dst, src, alpha, n, stream=None):
    _call("zj_gpu_scale_array", dst, src, alpha, n, stream)


def axpy(y, x, alpha, n, stream=None):
    _call("zj_gpu_axpy", y, x, alpha, n, stream)
