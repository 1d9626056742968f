"""Pins the oracle's mesh path (reference mesh.cpp:18-194, fr_solver.cpp:85-129).

Face tables, normals and areas are compared bit for bit with the reference's
own Mesh (compiled in place) on its nine test meshes and five synthetic ones;
the periodic pairing (absent from the reference) is checked against what the
convention fixes: face counts, a perfect edge cover, partner midpoints one
period apart and normals pointing from left to right across the seam.
"""
import numpy as np
import pytest

from oracle import pyoracle as O
from tests.conftest import mesh_from_golden
from zhijian_amd import cases


def _names(g):
    return [str(n) for n in g["mesh_names"]]


def _corners(m):
    nv = 3 if int(m["elem_info"][0, 0]) == 0 else 4
    return nv


def test_golden_has_all_reference_meshes(golden_meshes):
    names = _names(golden_meshes)
    for n in ("quad_v41", "tri_v41", "tri_v22", "uniform_quad_10x10", "serendipity_quad_v41",
              "tri_order2_v41", "no_phys_v41", "no_phys_nosaveall_v41"):
        assert n in names


def test_faces_normals_areas_bit_exact(golden_meshes):
    for name in _names(golden_meshes):
        m = mesh_from_golden(golden_meshes, name)
        nv = _corners(m)
        f = O.build_faces(m["elem_nodes"], nv=nv)
        assert np.array_equal(f.as_array(), m["faces"]), name
        nrm = O.face_normals(f, m["elem_nodes"], m["xy"], nv=nv)
        assert np.array_equal(nrm, m["normals"]), name
        area = O.elem_areas(m["elem_nodes"], m["xy"], nv=nv)
        assert np.array_equal(area, m["area"]), name
        # boundary faces default to FarField (bc_type 4, tag 0) in the reference
        bnd = m["faces"][:, 2] < 0
        assert np.all(m["face_bc_type"][~bnd] == 0)


def test_geometry_vs_reference(golden_meshes):
    ops = O.quad_operators(3)
    checked = 0
    for name in _names(golden_meshes):
        m = mesh_from_golden(golden_meshes, name)
        if "J" not in m:
            continue
        nc = int(m["elem_info"][0, 2])
        J, Jinv = O.geometry(m["elem_nodes"], m["xy"], ops, node_count=nc)
        if nc == 4:
            assert np.array_equal(J, m["J"]) and np.array_equal(Jinv, m["Jinv"]), name
        else:  # biquadratic: same map, different rounding of the shape functions
            assert np.allclose(J, m["J"], rtol=1e-13, atol=0), name
            assert np.allclose(Jinv, m["Jinv"], rtol=1e-12, atol=1e-12), name
        checked += 1
    assert checked >= 5


def test_uniform_jacobian_closed_form():
    mesh = cases.uniform_quad_mesh(4, 3, 0.0, 2.0, 0.0, 1.5)
    ops = O.quad_operators(2)
    J, Jinv = O.geometry(mesh.elem_nodes, mesh.xy, ops)
    h = 0.5
    assert np.allclose(J, h * h / 4, rtol=1e-15)
    Ji = Jinv.reshape(-1, 4)
    assert np.allclose(Ji[:, 0], 2 / h) and np.allclose(Ji[:, 3], 2 / h)
    assert np.allclose(Ji[:, 1], 0) and np.allclose(Ji[:, 2], 0)


@pytest.mark.parametrize("nx,ny,jit", [(10, 10, 0.0), (3, 5, 0.0), (6, 4, 0.2), (1, 1, 0.0), (2, 1, 0.0)])
def test_periodic_pairing_convention(nx, ny, jit):
    mesh = cases.uniform_quad_mesh(nx, ny, -1.0, 2.0, 0.5, 2.5, jitter=jit, seed=11)
    f0 = O.build_faces(mesh.elem_nodes)
    f = O.pair_periodic(f0, mesh.elem_nodes, mesh.xy, mesh.Lx, mesh.Ly, 1e-9)
    E = mesh.n_elem
    assert f0.n == 2 * nx * ny + nx + ny
    assert f.n == 2 * E
    # every element edge appears exactly once on some face
    seen = np.zeros((E, 4), int)
    for a in range(f.n):
        seen[f.le[a], f.ll[a]] += 1
        seen[f.re[a], f.rl[a]] += 1
    assert np.all(seen == 1)
    # kept faces stay in their original relative order
    orig = {tuple(r[:2]): i for i, r in enumerate(f0.as_array())}
    idx = [orig[(f.le[a], f.ll[a])] for a in range(f.n)]
    assert idx == sorted(idx)
    # partners: the kept face has the lower original index; midpoints one period apart
    xy, en = mesh.xy, mesh.elem_nodes
    nrm = O.face_normals(f, en, xy)
    mid = lambda e, k: 0.5 * (xy[en[e, k]] + xy[en[e, (k + 1) % 4]])
    for a in range(f.n):
        ml, mr = mid(f.le[a], f.ll[a]), mid(f.re[a], f.rl[a])
        d = mr - ml
        if np.allclose(d, 0):
            continue  # interior face
        assert orig[(f.le[a], f.ll[a])] < [i for i, r in enumerate(f0.as_array())
                                             if r[0] == f.re[a] and r[1] == f.rl[a]][0]
        assert (np.isclose(abs(d[0]), mesh.Lx) and np.isclose(d[1], 0)) or \
               (np.isclose(abs(d[1]), mesh.Ly) and np.isclose(d[0], 0))
        # the normal points out of the left element, i.e. against the jump to the partner
        assert np.dot(nrm[a], d) < 0


def test_periodic_pairing_vortex_counts():
    mesh = cases.vortex_mesh(10)
    f0 = O.build_faces(mesh.elem_nodes)
    assert f0.n == 220 and int(np.sum(f0.re < 0)) == 40
    f = O.pair_periodic(f0, mesh.elem_nodes, mesh.xy, 10.0, 10.0)
    assert f.n == 200 and np.all(f.re >= 0)
