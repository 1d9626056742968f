"""Gmsh MSH 2.2 / 4.1 ASCII ingestion through the native reader
(zj_gmsh_* in include/zhijian_amd.h), with the reference's conventions
(src/mesh/gmsh_reader.cpp:161-760) and its face tagging
(GmshReader::applyBoundaryTags, :736-758).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import Dict, Tuple

import numpy as np

from . import _lib
from ._lib import ZJError, lib

BC_FARFIELD = 4


@dataclass
class GmshMesh:
    xy: np.ndarray          # [n][2]
    elem_nodes: np.ndarray  # [E][max nodes], -1 padded
    elem_kind: np.ndarray   # ElementType: 0 triangle, 1 quadrilateral
    elem_order: np.ndarray
    elem_n_nodes: np.ndarray
    groups: Dict[int, Tuple[str, int]]  # physical tag -> (name, BC type)
    _edge_tags: Dict[Tuple[int, int], int]
    periodic_nodes: np.ndarray = None   # [n][2] $Periodic (slave, master) node indices; the reference skips them

    def edge_tag(self, a: int, b: int) -> int:
        return self._edge_tags.get((min(a, b), max(a, b)), 0)

    def faces(self, periodic: bool = False):
        """Faces (zj_build_faces) with each boundary face's physical tag and BC
        type stamped as GmshReader::applyBoundaryTags does; untagged boundary
        faces keep buildFaces' FarField / tag 0, interior faces Interior / tag -1
        (mesh.cpp:40-57).  periodic=True first joins the faces the file's
        $Periodic links make correspond (zj_pair_periodic_nodes)."""
        kinds = set(self.elem_kind.tolist())
        if len(kinds) != 1:
            raise ValueError("mixed triangle/quadrilateral meshes are not supported here")
        nv = 3 if 0 in kinds else 4
        en = np.ascontiguousarray(self.elem_nodes[:, :nv])
        f = _lib.build_faces(en, nv=nv)
        if periodic:
            f = _lib.pair_periodic_nodes(f, en, self.xy, self.periodic_nodes, nv=nv)
        tag = np.full(f.shape[0], -1, np.int32)   # interior faces: tag -1, type Interior
        typ = np.zeros(f.shape[0], np.int32)
        for i in np.nonzero(f[:, 2] < 0)[0]:
            le, ll = int(f[i, 0]), int(f[i, 1])
            v0 = int(self.elem_nodes[le, ll % nv])
            v1 = int(self.elem_nodes[le, (ll + 1) % nv])
            t = self.edge_tag(v0, v1)
            typ[i] = BC_FARFIELD
            tag[i] = 0
            if t:
                tag[i] = t
                if t in self.groups:
                    typ[i] = self.groups[t][1]
        return f, tag, typ


def read(path: str) -> GmshMesh:
    L = lib()
    h = C.c_void_p()
    rc = L.zj_gmsh_read(path.encode(), C.byref(h))
    if rc != 0:
        raise ZJError(rc, L.zj_gmsh_error().decode())
    try:
        nn, ne, mx, ng = C.c_int64(), C.c_int(), C.c_int(), C.c_int()
        L.zj_gmsh_counts(h, C.byref(nn), C.byref(ne), C.byref(mx), C.byref(ng))
        xy = np.zeros((nn.value, 2))
        L.zj_gmsh_nodes(h, xy.ctypes.data)
        E, M = ne.value, mx.value
        kind, order, cnt = np.zeros(E, np.int32), np.zeros(E, np.int32), np.zeros(E, np.int32)
        nodes = np.zeros((E, M), np.int64)
        L.zj_gmsh_elements(h, kind.ctypes.data, order.ctypes.data, cnt.ctypes.data, nodes.ctypes.data, M)
        groups = {}
        buf = C.create_string_buffer(256)
        for i in range(ng.value):
            tag, typ = C.c_int(), C.c_int()
            L.zj_gmsh_group(h, i, C.byref(tag), C.byref(typ), buf, 256)
            groups[tag.value] = (buf.value.decode(), typ.value)
        # boundary-edge tags for every element edge (vertex pairs)
        edge_tags = {}
        t = C.c_int()
        for e in range(E):
            nv = 3 if kind[e] == 0 else 4
            for k in range(nv):
                a, b = int(nodes[e, k]), int(nodes[e, (k + 1) % nv])
                L.zj_gmsh_edge_tag(h, a, b, C.byref(t))
                if t.value:
                    edge_tags[(min(a, b), max(a, b))] = t.value
        npl = C.c_int64()
        L.zj_gmsh_periodic(h, C.byref(npl), None, None)
        links = np.zeros((npl.value, 2), np.int64)
        sl, ms = np.zeros(npl.value, np.int64), np.zeros(npl.value, np.int64)
        L.zj_gmsh_periodic(h, C.byref(npl), sl.ctypes.data, ms.ctypes.data)
        links[:, 0], links[:, 1] = sl, ms
        return GmshMesh(xy, nodes, kind, order, cnt, groups, edge_tags, links)
    finally:
        L.zj_gmsh_free(h)
