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


@pytest.mark.parametrize("nx,ny,jit", [(10, 10, 0.0), (3, 5, 0.0), (6, 4, 0.2), (2, 3, 0.1), (2, 2, 0.0),
                                       (1, 1, 0.0), (2, 1, 0.0)])
def test_node_pairing_equals_period_pairing(nx, ny, jit):
    """On a translation-periodic mesh, pairing by the lattice's node links
    gives the same face table as pairing by the periods, including meshes one
    or two elements across, where the end-node classes alone do not decide."""
    mesh = cases.uniform_quad_mesh(nx, ny, -1.0, 2.0, 0.5, 2.5, jitter=jit, seed=5)
    f0 = O.build_faces(mesh.elem_nodes)
    geo = O.pair_periodic(f0, mesh.elem_nodes, mesh.xy, mesh.Lx, mesh.Ly, 1e-9)
    nod = O.pair_periodic_nodes(f0, mesh.elem_nodes, mesh.xy, cases.periodic_links(nx, ny))
    assert np.array_equal(geo.as_array(), nod.as_array())
    links = cases.periodic_links(nx, ny)[::-1, ::-1]   # other order, slave and master swapped
    assert np.array_equal(O.pair_periodic_nodes(f0, mesh.elem_nodes, mesh.xy, links).as_array(),
                          geo.as_array())


@pytest.mark.parametrize("nx,ny,shear,jit", [(4, 3, 0.37, 0.0), (5, 6, -1.3, 0.2), (2, 2, 0.5, 0.0)])
def test_node_pairing_on_sheared_mesh(nx, ny, shear, jit):
    """A parallelogram domain: every edge is covered once, kept faces keep
    their order and the lower index, and partner midpoints differ by one of
    the spanning vectors (Lx, 0) or (shear, Ly).  Pairing by the axis
    periods pairs only the x sides here."""
    mesh = cases.sheared_quad_mesh(nx, ny, 2.0, 1.5, shear, jitter=jit, seed=2)
    en, xy = mesh.elem_nodes, mesh.xy
    f0 = O.build_faces(en)
    f = O.pair_periodic_nodes(f0, en, xy, cases.periodic_links(nx, ny))
    E = en.shape[0]
    assert f.n == 2 * E
    seen = np.zeros((E, 4), int)
    for a in range(f.n):
        seen[f.le[a], f.ll[a]] += 1
        seen[f.re[a], f.rl[a]] += 1
    assert np.all(seen == 1)
    rows = [tuple(r[:2]) for r in f0.as_array()]
    orig = {r: i for i, r in enumerate(rows)}
    idx = [orig[(f.le[a], f.ll[a])] for a in range(f.n)]
    assert idx == sorted(idx)
    mid = lambda e, k: 0.5 * (xy[en[e, k]] + xy[en[e, (k + 1) % 4]])
    spans = [np.array([2.0, 0.0]), np.array([shear, 1.5])]
    n_periodic = 0
    for a in range(f.n):
        d = mid(f.re[a], f.rl[a]) - mid(f.le[a], f.ll[a])
        if f0.re[orig[(f.le[a], f.ll[a])]] >= 0:
            continue  # interior before pairing
        n_periodic += 1
        assert orig[(f.le[a], f.ll[a])] < orig[(f.re[a], f.rl[a])]
        assert any(np.allclose(d, v) or np.allclose(d, -v) for v in spans)
    assert n_periodic == nx + ny
    by_axis = O.pair_periodic(f0, en, xy, 2.0, 1.5, 1e-9)
    assert by_axis.n == f.n + nx   # the sheared top and bottom rows stay unpaired


@pytest.mark.parametrize("nx", [2, 3])
def test_node_pairing_leaves_walls_alone(nx):
    """Periodic in x only, walls top and bottom: with two elements across,
    both bottom wall edges join a corner class to the middle class, yet they
    are not translates by a link offset, so every wall edge stays a boundary
    face."""
    mesh = cases.uniform_quad_mesh(nx, 3, 0.0, 1.0, 0.0, 1.0)
    links = cases.periodic_links(nx, 3)[:4]   # the column links only
    f0 = O.build_faces(mesh.elem_nodes)
    f = O.pair_periodic_nodes(f0, mesh.elem_nodes, mesh.xy, links)
    geo = O.pair_periodic(f0, mesh.elem_nodes, mesh.xy, 1.0, 0.0, 1e-9)
    assert np.array_equal(f.as_array(), geo.as_array())
    assert int(np.sum(f.re < 0)) == 2 * nx


def test_node_pairing_ambiguous_and_unlinked():
    """Links that also identify nodes half a period apart give a face two
    candidates (refused); without links nothing changes."""
    mesh = cases.uniform_quad_mesh(2, 2, 0.0, 2.0, 0.0, 2.0)
    links = np.concatenate([cases.periodic_links(2, 2), [[1, 0], [4, 3], [7, 6]]])
    f0 = O.build_faces(mesh.elem_nodes)
    with pytest.raises(ValueError):
        O.pair_periodic_nodes(f0, mesh.elem_nodes, mesh.xy, links)
    same = O.pair_periodic_nodes(f0, mesh.elem_nodes, mesh.xy, np.zeros((0, 2), np.int64))
    assert np.array_equal(same.as_array(), f0.as_array())
