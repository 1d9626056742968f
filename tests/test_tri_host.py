"""Host side of the triangle path, CPU only: the library's operator tables
against the oracle's, argument checking of zj_tri_create (every check runs
before the first HIP call), and the synthetic triangle workload."""
import math

import numpy as np
import pytest

from oracle import tri_oracle as T
from zhijian_amd import _lib, cases
from zhijian_amd._lib import ZJError
from zhijian_amd.tri import TriSolver, sp_coords, tri_faces


@pytest.mark.parametrize("p", [1, 2, 3, 4, 5])
def test_operators_match_oracle(p):
    a, b = _lib.tri_operators(p), T.operators(p)
    for k in ("sp", "wg", "Dr", "Ds", "interp", "lift"):
        ref = np.asarray(b[k]).reshape(a[k].shape)
        assert np.abs(a[k] - ref).max() <= 1e-12 * max(1.0, np.abs(ref).max()), k
    assert np.abs(a["fp"] - b["fp"].reshape(-1, 2)).max() <= 1e-15


@pytest.mark.parametrize("p", [1, 2, 3, 4, 5])
def test_esfr_lift_matches_oracle(p):
    """zj_tri_lift (modal construction: monomial fit of the Dubiner modes'
    p-th derivatives) against the oracle's nodal one (powers of Dr and Ds)."""
    ops = T.operators(p)
    assert np.array_equal(_lib.tri_lift(p, 0), _lib.tri_operators(p)["lift"])
    for code, fam in ((1, "sd"), (2, "hu"), (3, "ga")):
        ref = T.esfr_lift(p, fam, ops)
        assert np.abs(_lib.tri_lift(p, code) - ref).max() <= 1e-11 * np.abs(ref).max(), fam
    with pytest.raises(ZJError):
        _lib.tri_lift(p, 4)


def test_mesh_matches_oracle_mesh():
    for n, j, seed in ((4, 0.2, 3), (7, 0.0, 0), (9, 0.15, 11)):
        m = cases.split_tri_mesh(n, jitter=j, seed=seed)
        xy, tris = T.split_quad_mesh(n, jitter=j, seed=seed)
        assert np.array_equal(m.xy, xy) and np.array_equal(m.tris, tris)


def test_periodic_faces_pair_every_edge():
    m = cases.split_tri_mesh(6, jitter=0.2, seed=1)
    faces, nrm = tri_faces(m.xy, m.tris, periodic=(1.0, 1.0))
    assert faces.shape[0] == 3 * m.n_elem // 2 and np.all(faces[:, 2] >= 0)
    assert np.allclose(np.linalg.norm(nrm, axis=1), 1.0)


def _expect(code, **kw):
    m = cases.split_tri_mesh(3)
    with pytest.raises(ZJError) as ei:
        TriSolver(m.xy, kw.pop("tris", m.tris), 2, periodic=kw.pop("periodic", (1.0, 1.0)), **kw)
    assert ei.value.code == code


def test_create_rejects():
    m = cases.split_tri_mesh(3)
    _expect(1, tris=m.tris[:, ::-1].copy())            # clockwise corners
    _expect(1, periodic=None)                          # boundary faces without bc_type
    with pytest.raises(KeyError):
        TriSolver(m.xy, m.tris, 2, periodic=(1.0, 1.0), riemann="exact")


def test_create_rejects_quad_only_modes():
    import ctypes as C
    from zhijian_amd._lib import lib
    m = cases.split_tri_mesh(2)
    faces, nrm = tri_faces(m.xy, m.tris, periodic=(1.0, 1.0))
    cols = [np.ascontiguousarray(faces[:, k], np.int32) for k in range(4)]
    xy = np.ascontiguousarray(m.xy)
    tris = np.ascontiguousarray(m.tris)
    d = _lib.TriDesc()
    d.poly_order, d.n_nodes, d.xy, d.n_elem, d.tri = 2, xy.shape[0], xy.ctypes.data, tris.shape[0], tris.ctypes.data
    d.n_faces = faces.shape[0]
    d.face_left_elem, d.face_left_local, d.face_right_elem, d.face_right_local = (c.ctypes.data for c in cols)
    d.normals = nrm.ctypes.data
    for compat, positivity, corr, code in ((0, 0, 0, 4), (1, 1, 0, 4), (1, 0, 9, 1), (1, 0, -1, 1), (7, 0, 0, 4)):
        prm = _lib.Params(1.4, 0.5, 0, compat, positivity, 0, 1, corr)
        h = C.c_void_p()
        assert lib().zj_tri_create(C.byref(h), C.byref(d), C.byref(prm)) == code
        assert not h.value
    d.poly_order = 6
    prm = _lib.Params(1.4, 0.5, 0, 1, 0, 0, 1, 0)
    assert lib().zj_tri_create(C.byref(C.c_void_p()), C.byref(d), C.byref(prm)) == 1


def test_fourier_freestream():
    m = cases.split_tri_mesh(16)
    pos = sp_coords(m.xy, m.tris, 3)
    U = cases.fourier_freestream_ic(pos, mach=0.5, aoa_deg=2.0, seed=3)
    assert np.allclose(U[..., 0], 1.0)
    p = 0.4 * (U[..., 3] - 0.5 * (U[..., 1] ** 2 + U[..., 2] ** 2) / U[..., 0])
    assert np.allclose(p, 1.0 / (1.4 * 0.25))
    # each mode is periodic on the unit square: the mean velocity is the free stream
    a = math.radians(2.0)
    assert abs(U[..., 1].mean() - math.cos(a)) < 0.02 and abs(U[..., 2].mean() - math.sin(a)) < 0.02
    assert np.array_equal(U, cases.fourier_freestream_ic(pos, mach=0.5, aoa_deg=2.0, seed=3))
    assert not np.array_equal(U, cases.fourier_freestream_ic(pos, mach=0.5, aoa_deg=2.0, seed=4))
