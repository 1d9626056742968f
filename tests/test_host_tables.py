"""The product's native host tables against the reference (golden) and the oracle,
and the C ABI's exports and argument checks (no GPU needed)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from oracle import pyoracle as O
from tests.conftest import ROOT, mesh_from_golden
from zhijian_amd import _lib, cases, solver


def test_library_exports_every_declared_symbol():
    hdr = open(os.path.join(ROOT, "include", "zhijian_amd.h")).read()
    declared = set(re.findall(r"\b(zj_[a-z0-9_]+)\s*\(", hdr))
    assert declared == set(_lib.EXPORTS)
    L = _lib.lib()
    for name in declared:
        assert hasattr(L, name), name
    assert L.zj_version() >= 1


@pytest.mark.parametrize("p", range(1, 6))
def test_operators_bit_exact(golden_ops, p):
    ops = _lib.quad_operators(p)
    for k in ("sol_xi", "sol_eta", "qw", "diff_xi", "diff_eta", "interp", "corr"):
        assert np.array_equal(ops[k], golden_ops[f"P{p}_{k}"]), k


def test_faces_normals_areas_geometry_bit_exact(golden_meshes):
    for name in [str(n) for n in golden_meshes["mesh_names"]]:
        m = mesh_from_golden(golden_meshes, name)
        nv = 3 if int(m["elem_info"][0, 0]) == 0 else 4
        f = _lib.build_faces(m["elem_nodes"], nv=nv)
        assert np.array_equal(f, m["faces"]), name
        assert np.array_equal(_lib.face_normals(f, m["elem_nodes"], m["xy"], nv=nv), m["normals"]), name
        assert np.array_equal(_lib.elem_areas(m["elem_nodes"], m["xy"], nv=nv), m["area"]), name
        if "J" in m:
            nc = int(m["elem_info"][0, 2])
            J, Jinv = _lib.geometry(m["elem_nodes"], m["xy"], _lib.quad_operators(3), node_count=nc)
            if nc == 4:
                assert np.array_equal(J, m["J"]) and np.array_equal(Jinv, m["Jinv"]), name
            else:
                assert np.allclose(J, m["J"], rtol=1e-13, atol=0), name


@pytest.mark.parametrize("nx,ny,jit", [(10, 10, 0.0), (7, 3, 0.2), (1, 1, 0.0), (1, 4, 0.0), (16, 16, 0.1)])
def test_periodic_pairing_matches_oracle(nx, ny, jit):
    mesh = cases.uniform_quad_mesh(nx, ny, -1.0, 2.0, 0.0, 1.5, jitter=jit, seed=nx)
    f = _lib.build_faces(mesh.elem_nodes)
    fp = _lib.pair_periodic(f, mesh.elem_nodes, mesh.xy, mesh.Lx, mesh.Ly)
    of = O.pair_periodic(O.build_faces(mesh.elem_nodes), mesh.elem_nodes, mesh.xy, mesh.Lx, mesh.Ly,
                         1e-9 * max(mesh.Lx, mesh.Ly))
    assert np.array_equal(fp, of.as_array())
    assert fp.shape[0] == 2 * nx * ny


@pytest.mark.parametrize("kind,n,shear,jit", [("quad", 10, 0.0, 0.0), ("quad", 7, 0.0, 0.2), ("quad", 2, 0.0, 0.0),
                                              ("quad", 1, 0.0, 0.0), ("sheared", 6, 0.45, 0.1),
                                              ("sheared", 2, -0.7, 0.0), ("tri", 8, 0.0, 0.2), ("tri", 2, 0.0, 0.0)])
def test_node_pairing_matches_oracle(kind, n, shear, jit):
    """zj_pair_periodic_nodes against the oracle's restatement, bit for bit,
    on square, sheared and triangle meshes with lattice links; on the
    translation-periodic ones also against zj_pair_periodic."""
    if kind == "tri":
        m = cases.split_tri_mesh(n, jitter=jit, seed=3)
        en, xy, nv = m.tris, m.xy, 3
    elif kind == "quad":
        m = cases.uniform_quad_mesh(n, n + 1, 0.0, 1.0, 0.0, 1.2, jitter=jit, seed=n)
        en, xy, nv = m.elem_nodes, m.xy, 4
    else:
        m = cases.sheared_quad_mesh(n, n + 1, 1.0, 1.2, shear, jitter=jit, seed=n)
        en, xy, nv = m.elem_nodes, m.xy, 4
    ny = n if kind == "tri" else n + 1
    links = cases.periodic_links(n, ny)
    f = _lib.build_faces(en, nv=nv)
    got = _lib.pair_periodic_nodes(f, en, xy, links, nv=nv)
    want = O.pair_periodic_nodes(O.build_faces(en, nv=nv), en, xy, links, nv=nv)
    assert np.array_equal(got, want.as_array())
    assert np.all(got[:, 2] >= 0)
    if kind != "sheared":
        L = (1.0, 1.0) if kind == "tri" else (1.0, 1.2)
        assert np.array_equal(got, _lib.pair_periodic(f, en, xy, L[0], L[1], nv=nv))


def test_node_pairing_argument_checks():
    mesh = cases.uniform_quad_mesh(2, 2, 0.0, 2.0, 0.0, 2.0)
    f = _lib.build_faces(mesh.elem_nodes)
    with pytest.raises(_lib.ZJError):   # a link to a node that does not exist
        _lib.pair_periodic_nodes(f, mesh.elem_nodes, mesh.xy, [[2, 99]])
    with pytest.raises(_lib.ZJError):   # half-period links: two candidates
        _lib.pair_periodic_nodes(f, mesh.elem_nodes, mesh.xy,
                                 np.concatenate([cases.periodic_links(2, 2), [[1, 0], [4, 3], [7, 6]]]))
    same = _lib.pair_periodic_nodes(f, mesh.elem_nodes, mesh.xy, np.zeros((0, 2), np.int64))
    assert np.array_equal(same, f)
    links = cases.periodic_links(2, 2)
    for col, bad in ((0, 4), (0, -1), (1, 4), (2, 7)):   # a face naming an element or edge that does not exist
        g = f.copy()
        g[np.nonzero(g[:, 2] < 0)[0][0] if col != 2 else 0, col] = bad
        with pytest.raises(_lib.ZJError):
            _lib.pair_periodic_nodes(g, mesh.elem_nodes, mesh.xy, links)


def test_tables_match_oracle_case():
    from tests import fr_case
    mesh = cases.uniform_quad_mesh(5, 4, 0.0, 1.0, 0.0, 1.0, jitter=0.2, seed=4)
    t = solver.build_tables(mesh.xy, mesh.elem_nodes, 3, periodic=(mesh.Lx, mesh.Ly))
    oc = fr_case.from_quadmesh(mesh, 3)
    assert np.array_equal(t.faces, oc.faces.as_array())
    assert np.array_equal(t.normals, oc.normals)
    assert np.array_equal(t.J, oc.J) and np.array_equal(t.Jinv, oc.Jinv)
    assert np.array_equal(t.area, oc.area)
    assert np.array_equal(t.bc_type, oc.bc_type) and np.array_equal(t.bc_face_map, oc.bc_face_map)


def _desc_for(t, corrupt=None):
    keep = []

    def c(a, dt):
        a = np.ascontiguousarray(a, dt).copy()
        keep.append(a)
        return a
    ops = {k: c(v, np.float64) for k, v in t.ops.items()}
    if corrupt:
        corrupt(ops)
    fc = c(t.faces, np.int32)
    cols = [c(fc[:, k], np.int32) for k in range(4)]
    d = _lib.MeshDesc()
    d.poly_order = t.p; d.n_elem = t.E; d.n_faces = fc.shape[0]
    d.diff_xi = ops["diff_xi"].ctypes.data; d.diff_eta = ops["diff_eta"].ctypes.data
    d.interp = ops["interp"].ctypes.data; d.corr = ops["corr"].ctypes.data
    d.Jinv = c(t.Jinv, np.float64).ctypes.data
    d.normals = c(t.normals, np.float64).ctypes.data
    d.face_left_elem, d.face_left_local, d.face_right_elem, d.face_right_local = (a.ctypes.data for a in cols)
    return d, keep, cols


def _create(d, prec=0, compat=0, gamma=1.4):
    h = C.c_void_p()
    prm = _lib.Params(gamma, 0.5, prec, compat, 1, 0)
    rc = _lib.lib().zj_create(C.byref(h), C.byref(d), C.byref(prm))
    return rc, _lib.lib().zj_last_error(None).decode()


def test_create_rejects_bad_input_before_touching_the_gpu():
    mesh = cases.uniform_quad_mesh(3, 3, 0.0, 1.0, 0.0, 1.0)
    t = solver.build_tables(mesh.xy, mesh.elem_nodes, 2, periodic=(1.0, 1.0))
    # dense operators that are not Gauss tensor products
    def bad_diff(ops):
        ops["diff_xi"][1] += 1e-3
    d, keep, _ = _desc_for(t, bad_diff)
    rc, msg = _create(d)
    assert rc == 4 and "tensor" in msg
    def bad_interp(ops):
        S = 9
        I = ops["interp"].reshape(4, 3, S)
        I[2] = I[2][::-1].copy()  # edge 2 in forward order: not the reference's layout
        ops["interp"][:] = I.reshape(-1)
    d, keep, _ = _desc_for(t, bad_interp)
    rc, msg = _create(d)
    assert rc == 4 and "reversed" in msg
    # an element edge used twice
    d, keep, cols = _desc_for(t)
    cols[1][1] = cols[1][0]; cols[0][1] = cols[0][0]
    rc, msg = _create(d)
    assert rc == 1 and "two faces" in msg
    # unsupported arithmetic mode, bad order, bad gamma
    d, keep, _ = _desc_for(t)
    assert _create(d, compat=7)[0] == 1
    rc, msg = _create(d, compat=1)  # textbook mode without the edge lengths it needs
    assert rc == 1 and "edge_half_len" in msg
    assert _create(d, gamma=1.0)[0] == 1
    d.poly_order = 7
    assert _create(d)[0] == 1
    # far-field faces without bc_data
    open_t = solver.build_tables(mesh.xy, mesh.elem_nodes, 2, periodic=None)
    d, keep, _ = _desc_for(open_t)
    bct = np.full(open_t.faces.shape[0], 4, np.int32); keep.append(bct)
    bmap = np.arange(open_t.faces.shape[0], dtype=np.int32); keep.append(bmap)
    d.bc_type = bct.ctypes.data; d.bc_face_map = bmap.ctypes.data
    rc, msg = _create(d)
    assert rc == 1 and "bc_data" in msg


def test_batch_step_argument_checks_before_the_library():
    """Batch.rk3_step refuses a dts array of the wrong length and a closed
    member before any call into the library (no GPU needed for the checks)."""
    class _Member:
        def __init__(self, v):
            self._h = C.c_void_p(v)
    b = solver.Batch.__new__(solver.Batch)
    b.members = [_Member(1), _Member(2)]
    b._h = C.c_void_p()
    with pytest.raises(ValueError, match="2 members"):
        b.rk3_step(np.zeros(3))
    b.members[1]._h = C.c_void_p()
    with pytest.raises(RuntimeError, match="closed"):
        b.rk3_step(np.zeros(2))


def test_adjoint_entry_points_reject_null_handles():
    """The step-backwards entry points added for batches, the CFL step and
    triangles return ZJ_ERR_ARG on a null handle or null arrays, before any
    device work (no GPU needed); Batch.rk3_step_vjp checks its dts like
    rk3_step."""
    L = _lib.lib()
    dts = np.zeros(2)
    assert L.zj_batch_rk3_step_vjp(None, dts, None, None) == 1
    assert L.zj_dt_cfl_vjp(None, 0.5, 1.0, None) == 1
    assert L.zj_tri_rk3_step_vjp(None, 1e-3, None, None) == 1
    assert L.zj_tri_residual_vjp(None, dts, dts) == 1

    class _Member:
        def __init__(self, v):
            self._h = C.c_void_p(v)
    b = solver.Batch.__new__(solver.Batch)
    b.members = [_Member(1), _Member(2)]
    b._h = C.c_void_p()
    with pytest.raises(ValueError, match="2 members"):
        b.rk3_step_vjp(np.zeros(3))
    b.members[0]._h = C.c_void_p()
    with pytest.raises(RuntimeError, match="closed"):
        b.rk3_step_vjp(np.zeros(2))
