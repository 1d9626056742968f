"""Restart files and visualisation output (zj_restart_*, zj_vtk_write,
zj_pvtu_write of include/zhijian_amd.h).

Restart files follow the reference's RestartIO (src/io/restart.cpp:8-165):
HDF5 with root attributes time, iteration, poly_order, gamma and datasets
"solution" (float64, element by element, point by point, variable fastest)
and "elem_n_sp" (int32).  The library's own HDF5 codec writes and reads them.

VTK output follows the reference's VTKWriter (src/io/visualization.cpp):
each element subdivided into a lattice of `subdivisions` cells, split into
triangles, with Density, Pressure, Mach and Velocity at the lattice points.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib

VTU, LEGACY, BINARY, INTERPOLATE = 0, 1, 2, 4


def _check(rc):
    if rc != _lib.ZJ_OK:
        msg = _lib.lib().zj_io_last_error()
        raise _lib.ZJError(rc, msg.decode() if msg else "")


def _path(p) -> bytes:
    return str(p).encode()


def write_restart(path, elem_n_sp, solution, time=0.0, iteration=0, poly_order=1, gamma=1.4):
    """RestartIO::write.  elem_n_sp [E] int, solution [sum(elem_n_sp) * 4]."""
    nsp = np.ascontiguousarray(elem_n_sp, np.int32).reshape(-1)
    sol = np.ascontiguousarray(solution, np.float64).reshape(-1)
    if sol.size != int(nsp.sum()) * 4:
        raise ValueError("solution size does not match elem_n_sp")
    _check(_lib.lib().zj_restart_write(_path(path), nsp.size, nsp.ctypes.data, sol.ctypes.data, float(time),
                                       int(iteration), int(poly_order), float(gamma)))


def read_restart(path) -> dict:
    """RestartIO::read: dict(time, iteration, poly_order, gamma, elem_n_sp, solution)."""
    info = _lib.RestartInfo()
    _check(_lib.lib().zj_restart_query(_path(path), C.byref(info)))
    nsp = np.zeros(info.n_elem, np.int32)
    sol = np.zeros(info.n_values, np.float64)
    _check(_lib.lib().zj_restart_read(_path(path), info.n_elem, nsp.ctypes.data, info.n_values, sol.ctypes.data))
    return dict(time=info.time, iteration=info.iteration, poly_order=info.poly_order, gamma=info.gamma,
                elem_n_sp=nsp, solution=sol)


def save_handle(kind: str, handle, path, time=0.0, iteration=0):
    f = _lib.lib().zj_restart_save if kind == "quad" else _lib.lib().zj_tri_restart_save
    _check(f(handle, _path(path), float(time), int(iteration)))


def load_handle(kind: str, handle, path):
    f = _lib.lib().zj_restart_load if kind == "quad" else _lib.lib().zj_tri_restart_load
    t, it = C.c_double(0.0), C.c_int(0)
    _check(f(handle, _path(path), C.byref(t), C.byref(it)))
    return t.value, it.value


class VisMesh:
    """Element connectivity for zj_vtk_write.  elements: list of node-id
    sequences or an [E][k] array; elem_type 0 triangle / 1 quadrilateral
    (per element or one value); elem_order geometric order (1, 2, 4)."""

    def __init__(self, xy, elements, elem_type=1, elem_order=1):
        self.xy = np.ascontiguousarray(xy, np.float64).reshape(-1, 2)
        if isinstance(elements, np.ndarray) and elements.ndim == 2:
            counts = np.full(elements.shape[0], elements.shape[1], np.int64)
            nodes = np.ascontiguousarray(elements, np.int64).reshape(-1)
        else:
            counts = np.array([len(e) for e in elements], np.int64)
            nodes = np.array([n for e in elements for n in e], np.int64)
        E = counts.size
        self.start = np.zeros(E + 1, np.int64)
        np.cumsum(counts, out=self.start[1:])
        self.nodes = nodes if nodes.size else np.zeros(1, np.int64)
        self.type = np.ascontiguousarray(np.broadcast_to(np.asarray(elem_type, np.int32), (E,)))
        self.order = np.ascontiguousarray(np.broadcast_to(np.asarray(elem_order, np.int32), (E,)))
        self.desc = _lib.VisMesh(self.xy.shape[0], self.xy.ctypes.data, E, self.type.ctypes.data,
                                 self.order.ctypes.data, self.start.ctypes.data, self.nodes.ctypes.data)


def write_vtk(path, mesh: VisMesh, elem_n_sp, solution, poly_order, gamma=1.4, subdivisions=4, flags=VTU):
    nsp = np.ascontiguousarray(np.broadcast_to(np.asarray(elem_n_sp, np.int32), (mesh.type.size,)))
    sol = np.ascontiguousarray(solution, np.float64).reshape(-1)
    if sol.size != int(nsp.sum()) * 4:
        raise ValueError("solution size does not match elem_n_sp")
    _check(_lib.lib().zj_vtk_write(_path(path), C.byref(mesh.desc), nsp.ctypes.data,
                                   sol.ctypes.data if sol.size else None, int(poly_order), float(gamma),
                                   int(subdivisions), int(flags)))


def write_pvtu(path, pieces):
    arr = (C.c_char_p * max(len(pieces), 1))(*[_path(p) for p in pieces])
    _check(_lib.lib().zj_pvtu_write(_path(path), len(pieces), arr))


def read_vtu_binary(path) -> dict:
    """Parses a .vtu written with BINARY (appended raw, UInt64 headers):
    name -> array (points as [N][3])."""
    import xml.etree.ElementTree as ET
    raw = open(path, "rb").read()
    mark = raw.index(b"<AppendedData encoding=\"raw\">")
    us = raw.index(b"_", mark) + 1
    root = ET.fromstring(raw[:mark].decode() + "</VTKFile>")
    types = {"Float64": np.float64, "Int32": np.int32, "Int64": np.int64, "UInt8": np.uint8}
    out = {}
    for da in root.iter("DataArray"):
        off = us + int(da.get("offset"))
        n = int(np.frombuffer(raw, np.uint64, 1, off)[0])
        a = np.frombuffer(raw, types[da.get("type")], n // np.dtype(types[da.get("type")]).itemsize, off + 8)
        nc = int(da.get("NumberOfComponents", "1"))
        out[da.get("Name", "Points")] = a.reshape(-1, nc) if nc > 1 else a
    return out
