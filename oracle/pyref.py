"""ctypes bindings for oracle/_ref/libzjref.so, the reference's own host code.

TEST INFRASTRUCTURE ONLY.  The library is built by oracle/Makefile from the
reference sources where they lie (src/basis/basis.cpp, src/mesh/mesh.cpp,
src/mesh/gmsh_reader.cpp, include/physics/euler.hpp) plus oracle/ref_shim.cpp.
It exists only where the reference sources were present at build time.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import pyoracle

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_lp = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")

_lib = None


def available() -> bool:
    return os.path.exists(pyoracle.REF_PATH)


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        L = C.CDLL(pyoracle.REF_PATH)
        L.ref_gauss.argtypes = [C.c_int, _dp, _dp]
        L.ref_corr_deriv.argtypes = [C.c_int, C.c_int, C.c_double, C.c_int]
        L.ref_corr_deriv.restype = C.c_double
        L.ref_quad_operators.argtypes = [C.c_int, C.c_int] + [_dp] * 7
        L.ref_mesh_from_arrays.argtypes = [C.c_int, _dp, C.c_int, C.c_int, _lp]
        L.ref_mesh_from_arrays.restype = C.c_void_p
        L.ref_mesh_read_gmsh.argtypes = [C.c_char_p]
        L.ref_mesh_read_gmsh.restype = C.c_void_p
        L.ref_mesh_free.argtypes = [C.c_void_p]
        L.ref_mesh_counts.argtypes = [C.c_void_p] + [C.POINTER(C.c_int)] * 4
        L.ref_mesh_dump.argtypes = [C.c_void_p, _dp, C.c_int, _lp, _ip, _ip, _ip, _ip, _ip, _ip,
                                    _ip, _dp, _dp, _dp]
        L.ref_mesh_geometry.argtypes = [C.c_void_p, C.c_int, _dp, _dp]
        L.ref_flux_xy.argtypes = [_dp, C.c_double, _dp, _dp]
        L.ref_max_wave_speed.argtypes = [_dp, C.c_double]
        L.ref_max_wave_speed.restype = C.c_double
        L.ref_riemann.argtypes = [C.c_int, _dp, _dp, C.c_double, C.c_double, C.c_double, _dp]
        L.ref_visc_flux.argtypes = [_dp, _dp, _dp, C.c_double, C.c_double, C.c_double, _dp, _dp, _dp]
        _lib = L
    return _lib


def gauss(n):
    p = np.zeros(n); w = np.zeros(n)
    lib().ref_gauss(n, p, w)
    return p, w


def quad_operators(p, flux_type=0):
    n = p + 1; S = n * n
    out = dict(sol_xi=np.zeros(S), sol_eta=np.zeros(S), qw=np.zeros(S),
               diff_xi=np.zeros(S * S), diff_eta=np.zeros(S * S),
               interp=np.zeros(4 * n * S), corr=np.zeros(4 * S))
    lib().ref_quad_operators(p, flux_type, out["sol_xi"], out["sol_eta"], out["qw"],
                             out["diff_xi"], out["diff_eta"], out["interp"], out["corr"])
    return out


def corr_deriv(flux_type, p, xi, right):
    return lib().ref_corr_deriv(flux_type, p, xi, int(right))


def _dump(h, p_geom=None):
    L = lib()
    nn, ne, nf, mx = C.c_int(), C.c_int(), C.c_int(), C.c_int()
    L.ref_mesh_counts(h, C.byref(nn), C.byref(ne), C.byref(nf), C.byref(mx))
    nn, ne, nf, mx = nn.value, ne.value, nf.value, mx.value
    xy = np.zeros(max(2 * nn, 1)); en = np.zeros(max(ne * mx, 1), np.int64)
    info = np.zeros(max(3 * ne, 1), np.int32)
    fa = [np.zeros(max(nf, 1), np.int32) for _ in range(6)]
    normals = np.zeros(max(2 * nf, 1)); area = np.zeros(max(ne, 1)); flen = np.zeros(max(nf, 1))
    L.ref_mesh_dump(h, xy, mx, en, info, *fa, normals, area, flen)
    out = dict(xy=xy[: 2 * nn].reshape(nn, 2), elem_nodes=en[: ne * mx].reshape(ne, mx),
               elem_info=info[: 3 * ne].reshape(ne, 3),
               faces=np.stack([a[:nf] for a in fa[:4]], axis=1),
               face_bc_type=fa[4][:nf].copy(), face_bc_tag=fa[5][:nf].copy(),
               normals=normals[: 2 * nf].reshape(nf, 2), area=area[:ne], face_len=flen[:nf])
    if p_geom is not None:
        S = (p_geom + 1) ** 2
        J = np.zeros(max(ne * S, 1)); Jinv = np.zeros(max(ne * S * 4, 1))
        L.ref_mesh_geometry(h, p_geom, J, Jinv)
        out["J"] = J[: ne * S]; out["Jinv"] = Jinv[: ne * S * 4]
    return out


def mesh_from_arrays(xy, elem_nodes, p_geom=None):
    xy = np.ascontiguousarray(xy, np.float64)
    en = np.ascontiguousarray(elem_nodes, np.int64)
    h = lib().ref_mesh_from_arrays(xy.shape[0], xy.reshape(-1), en.shape[0], en.shape[1],
                                   en.reshape(-1))
    try:
        return _dump(h, p_geom)
    finally:
        lib().ref_mesh_free(h)


def mesh_read_gmsh(path, p_geom=None):
    h = lib().ref_mesh_read_gmsh(path.encode())
    if not h:
        raise IOError("reference GmshReader failed on " + path)
    try:
        return _dump(h, p_geom)
    finally:
        lib().ref_mesh_free(h)


def flux_xy(U, gamma=1.4):
    Fx = np.zeros(4); Fy = np.zeros(4)
    lib().ref_flux_xy(np.ascontiguousarray(U, np.float64), gamma, Fx, Fy)
    return Fx, Fy


def max_wave_speed(U, gamma=1.4):
    return lib().ref_max_wave_speed(np.ascontiguousarray(U, np.float64), gamma)


def riemann(which, UL, UR, nx, ny, gamma=1.4):
    F = np.zeros(4)
    lib().ref_riemann(which, np.ascontiguousarray(UL, np.float64),
                      np.ascontiguousarray(UR, np.float64), nx, ny, gamma, F)
    return F


def visc_flux(U, dUdx, dUdy, gamma=1.4, Re=1000.0, Pr=0.72):
    """ViscousFlux::fluxX/fluxY and the (mu, k) they used."""
    Fx, Fy, mk = np.zeros(4), np.zeros(4), np.zeros(2)
    lib().ref_visc_flux(*(np.ascontiguousarray(a, np.float64) for a in (U, dUdx, dUdy)), gamma, Re, Pr,
                        Fx, Fy, mk)
    return Fx, Fy, mk
