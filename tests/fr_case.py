"""Builds one solver case (mesh, operators, geometry, BC tables) for the tests.

The mesh-derived tables are built twice and independently: by the oracle here
(``OracleCase``) and by the product library in the GPU tests; the tests compare
the two.  BC tables follow fr_solver.cpp:432-482 (boundary faces in face order,
8 doubles each; FarField gets the free stream).
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import Optional

import numpy as np

from oracle import pyoracle as O


@dataclass
class OracleCase:
    p: int
    ops: dict
    faces: O.Faces
    normals: np.ndarray
    J: np.ndarray
    Jinv: np.ndarray
    area: np.ndarray
    bc_type: np.ndarray
    bc_face_map: np.ndarray
    bc_data: np.ndarray
    gamma: float
    E: int
    xy: np.ndarray
    elem_nodes: np.ndarray

    @property
    def S(self):
        return (self.p + 1) ** 2

    def solver(self) -> O.OracleSolver:
        return O.OracleSolver(self.p, self.ops, self.J, self.Jinv, self.faces, self.normals,
                              gamma=self.gamma, bc_type=self.bc_type,
                              bc_face_map=self.bc_face_map, bc_data=self.bc_data,
                              elem_area=self.area)

    def textbook(self, riemann="roe", corr="dg") -> O.TextbookSolver:
        return self.solver().textbook(O.edge_half_lengths(self.elem_nodes, self.xy), riemann, corr)


def freestream(gamma=1.4, p_inf=101325.0, rho_inf=1.225, mach=0.5, aoa_deg=0.0):
    c = math.sqrt(gamma * p_inf / rho_inf)
    vel = mach * c
    return rho_inf, vel * math.cos(aoa_deg * math.pi / 180.0), vel * math.sin(aoa_deg * math.pi / 180.0), p_inf


def bc_tables(faces: O.Faces, bc_type_per_face: Optional[np.ndarray], fs, outflow_p=None):
    F = faces.n
    bct = np.zeros(F, np.int32)
    bmap = -np.ones(F, np.int32)
    data = []
    for f in range(F):
        if faces.re[f] >= 0:
            continue
        t = int(bc_type_per_face[f]) if bc_type_per_face is not None else O.BC_FARFIELD
        bct[f] = t
        bmap[f] = len(data)
        if t == O.BC_FARFIELD:
            data.append([fs[0], fs[1], fs[2], fs[3], 0, 0, 0, 0])
        elif t == O.BC_OUTFLOW:
            data.append([fs[3] if outflow_p is None else outflow_p, 0, 0, 0, 0, 0, 0, 0])
        else:
            data.append([0.0] * 8)
    bd = np.array(data, np.float64).reshape(-1) if data else np.zeros(8)
    return bct, bmap, bd


def build(xy, elem_nodes, p, periodic=None, gamma=1.4, bc_type_per_face=None, fs=None,
          node_count=None, periodic_nodes=None) -> OracleCase:
    en = np.ascontiguousarray(elem_nodes, np.int64)
    nv = 4
    ops = O.quad_operators(p)
    faces = O.build_faces(en, nv=nv)
    if periodic is not None:
        faces = O.pair_periodic(faces, en, xy, periodic[0], periodic[1], 1e-9 * max(periodic), nv=nv)
    if periodic_nodes is not None:
        faces = O.pair_periodic_nodes(faces, en[:, :nv], xy, periodic_nodes, nv=nv)
    normals = O.face_normals(faces, en, xy, nv=nv)
    J, Jinv = O.geometry(en, xy, ops, node_count=node_count)
    area = O.elem_areas(en, xy, nv=nv)
    fs = freestream(gamma) if fs is None else fs
    bct, bmap, bd = bc_tables(faces, bc_type_per_face, fs)
    return OracleCase(p, ops, faces, normals, J, Jinv, area, bct, bmap, bd, gamma,
                      en.shape[0], np.asarray(xy, np.float64), en)


def from_quadmesh(mesh, p, periodic=True, **kw) -> OracleCase:
    return build(mesh.xy, mesh.elem_nodes, p, periodic=(mesh.Lx, mesh.Ly) if periodic else None, **kw)
