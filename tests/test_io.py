"""Restart files and VTK output, CPU only (zj_restart_*, zj_vtk_write,
zj_pvtu_write).

Restart files are pinned against libhdf5 itself in both directions: the
committed fixtures were written by libhdf5 with the reference's call sequence
(tests/golden/make_restart_h5.c) and must read back bit for bit; files this
library writes must read back bit for bit through libhdf5 (h5py, when an
installed libhdf5 is present: the conda HDF5 1.10.6 of this image).

VTK text is compared byte for byte with the restatement of the reference's
writers (oracle/vtk_oracle.py), which is itself pinned here by closed forms:
lattice coordinates on affine elements, counts, XML well-formedness, and the
gas relations evaluated independently.
"""
import math
import os
import shutil
import subprocess
import xml.etree.ElementTree as ET

import numpy as np
import pytest

from oracle import vtk_oracle as V
from zhijian_amd import io
from zhijian_amd._lib import ZJError

GOLD = os.path.join(os.path.dirname(__file__), "golden")
CONDA_PY = "/opt/conda/bin/python3.9"
HAVE_H5PY = os.path.exists(CONDA_PY) and subprocess.run([CONDA_PY, "-c", "import h5py"],
                                                         capture_output=True).returncode == 0


def _sample(n_elem, seed):
    rng = np.random.default_rng(seed)
    nsp = rng.choice([4, 9, 16, 6], size=n_elem).astype(np.int32)
    sol = rng.normal(size=int(nsp.sum()) * 4)
    return nsp, sol


# ----------------------------------------------------------------- restart
def test_restart_round_trip_bit_exact(tmp_path):
    nsp, sol = _sample(37, 1)
    sol[:5] = [0.0, -0.0, 5e-324, -1.7976931348623157e308, np.inf]
    p = tmp_path / "r.h5"
    io.write_restart(p, nsp, sol, time=0.125 + 1e-17, iteration=2 ** 31 - 1, poly_order=3, gamma=1.4)
    r = io.read_restart(p)
    assert r["time"] == 0.125 + 1e-17 and r["iteration"] == 2 ** 31 - 1 and r["poly_order"] == 3
    assert r["gamma"] == 1.4
    assert np.array_equal(r["elem_n_sp"], nsp)
    assert r["solution"].tobytes() == sol.tobytes()


def test_restart_deterministic_bytes(tmp_path):
    nsp, sol = _sample(11, 2)
    for name in ("a.h5", "b.h5"):
        io.write_restart(tmp_path / name, nsp, sol, 1.0, 5, 2, 1.4)
    assert (tmp_path / "a.h5").read_bytes() == (tmp_path / "b.h5").read_bytes()


def test_restart_empty(tmp_path):
    p = tmp_path / "e.h5"
    io.write_restart(p, np.zeros(0, np.int32), np.zeros(0), 0.0, 0, 1, 1.4)
    r = io.read_restart(p)
    assert r["elem_n_sp"].size == 0 and r["solution"].size == 0


def test_restart_reads_libhdf5_fixture():
    """restart_libhdf5.h5: written by libhdf5 1.10.6 with the reference's
    RestartIO::write call sequence (make_restart_h5.c)."""
    exp = np.load(os.path.join(GOLD, "restart_expected.npz"))
    r = io.read_restart(os.path.join(GOLD, "restart_libhdf5.h5"))
    assert r["time"] == float(exp["time"]) and r["iteration"] == int(exp["iteration"])
    assert r["poly_order"] == int(exp["poly_order"]) and r["gamma"] == float(exp["gamma"])
    assert np.array_equal(r["elem_n_sp"], exp["elem_n_sp"])
    assert r["solution"].tobytes() == exp["solution"].astype(np.float64).tobytes()


def test_restart_parallel_fixture_is_not_a_serial_restart():
    """The writeParallel fallback file has no elem_n_sp / poly_order: the
    serial reader refuses it with a message instead of reading garbage."""
    with pytest.raises(ZJError) as e:
        io.read_restart(os.path.join(GOLD, "restart_libhdf5_par.h5.rank0"))
    assert "poly_order" in str(e.value) or "elem_n_sp" in str(e.value)


@pytest.mark.skipif(not HAVE_H5PY, reason="no installed libhdf5/h5py to check against")
def test_libhdf5_reads_our_restart(tmp_path):
    nsp, sol = _sample(200, 3)
    p = tmp_path / "ours.h5"
    io.write_restart(p, nsp, sol, time=3.25, iteration=77, poly_order=4, gamma=1.3)
    script = (
        "import h5py, numpy as np, sys\n"
        "f = h5py.File(sys.argv[1], 'r')\n"
        "a = f.attrs\n"
        "assert sorted(f.keys()) == ['elem_n_sp', 'solution'], list(f.keys())\n"
        "assert f['solution'].dtype == np.float64 and f['elem_n_sp'].dtype == np.int32\n"
        "assert a['iteration'].dtype == np.int32 and a['poly_order'].dtype == np.int32\n"
        "np.save(sys.argv[2], f['solution'][:]); np.save(sys.argv[3], f['elem_n_sp'][:])\n"
        "print(repr(float(a['time'])), int(a['iteration']), int(a['poly_order']), repr(float(a['gamma'])))\n")
    out = subprocess.run([CONDA_PY, "-c", script, str(p), str(tmp_path / "s.npy"), str(tmp_path / "n.npy")],
                         capture_output=True, text=True, cwd=str(tmp_path), timeout=120)
    assert out.returncode == 0, out.stderr
    t, it, po, g = out.stdout.split()
    assert float(t) == 3.25 and int(it) == 77 and int(po) == 4 and float(g) == 1.3
    assert np.load(tmp_path / "s.npy").tobytes() == sol.tobytes()
    assert np.array_equal(np.load(tmp_path / "n.npy"), nsp)


@pytest.mark.skipif(not (os.path.exists("/opt/conda/include/hdf5.h") and shutil.which("gcc")),
                    reason="no installed libhdf5 headers to build the fixture writer")
def test_reads_large_libhdf5_file(tmp_path):
    """50 000 elements written by libhdf5 (not committed: made here)."""
    big = tmp_path / "big.h5"
    subprocess.run(["python3", os.path.join(GOLD, "make_restart_golden.py"), "--big", str(big)], check=True,
                   timeout=300, capture_output=True)
    exp = np.load(str(big) + ".npz")
    r = io.read_restart(big)
    assert np.array_equal(r["elem_n_sp"], exp["elem_n_sp"])
    assert r["solution"].tobytes() == exp["solution"].tobytes()
    assert r["iteration"] == int(exp["iteration"]) and r["time"] == float(exp["time"])


def test_restart_errors(tmp_path):
    with pytest.raises(ZJError, match="cannot open"):
        io.read_restart(tmp_path / "missing.h5")
    (tmp_path / "junk.h5").write_bytes(b"not an hdf5 file" * 10)
    with pytest.raises(ZJError, match="not an HDF5 file"):
        io.read_restart(tmp_path / "junk.h5")
    nsp, sol = _sample(20, 4)
    good = tmp_path / "g.h5"
    io.write_restart(good, nsp, sol, 0.0, 1, 2, 1.4)
    raw = good.read_bytes()
    for cut in (40, 300, len(raw) // 2, len(raw) - 8):
        bad = tmp_path / f"cut{cut}.h5"
        bad.write_bytes(raw[:cut])
        with pytest.raises(ZJError):
            io.read_restart(bad)
    with pytest.raises(ZJError, match="cannot create"):
        io.write_restart(tmp_path / "no_dir" / "x.h5", nsp, sol)


def test_restart_claimed_size_bounded_by_file(tmp_path):
    """A dataspace and layout that claim more data than the file holds are
    refused before any caller sizes an array from them."""
    nsp, sol = _sample(4, 8)
    good = tmp_path / "g.h5"
    io.write_restart(good, nsp, sol, 0.0, 1, 2, 1.4)
    raw = bytearray(good.read_bytes())
    n = sol.size
    # the solution dataset's dims and max dims, and its layout size, all say n doubles
    old_dims = np.array([n, n], np.uint64).tobytes()
    old_size = np.array([8 * n], np.uint64).tobytes()
    assert raw.count(old_dims) == 1 and raw.count(old_size) >= 1
    huge = 1 << 40
    raw = raw.replace(old_dims, np.array([huge, huge], np.uint64).tobytes())
    raw = raw.replace(old_size, np.array([8 * huge], np.uint64).tobytes(), 1)
    (tmp_path / "b.h5").write_bytes(bytes(raw))
    with pytest.raises(ZJError, match="past the end|smaller than"):
        io.read_restart(tmp_path / "b.h5")


def test_restart_corrupt_bytes_never_crash(tmp_path):
    """Flipped bytes anywhere in the metadata: an error or a successful read,
    never a crash (the reader bounds-checks every access)."""
    nsp, sol = _sample(9, 5)
    good = tmp_path / "g.h5"
    io.write_restart(good, nsp, sol, 0.0, 1, 2, 1.4)
    raw = bytearray(good.read_bytes())
    rng = np.random.default_rng(0)
    meta = raw.index(sol[:1].tobytes())
    for k in range(300):
        b = bytearray(raw)
        for _ in range(3):
            b[int(rng.integers(0, meta))] ^= int(rng.integers(1, 256))
        p = tmp_path / "c.h5"
        p.write_bytes(bytes(b))
        try:
            io.read_restart(p)
        except ZJError:
            pass


# ----------------------------------------------------------------- VTK
def _mixed_mesh(seed=0):
    """2 x 2 jittered Q1 quads, 2 triangles, 1 curved Q2 quad, 1 P2 triangle."""
    rng = np.random.default_rng(seed)
    xy = [(i + 0.1 * rng.normal(), j + 0.1 * rng.normal()) for j in range(3) for i in range(3)]
    els = [[0, 1, 4, 3], [1, 2, 5, 4], [3, 4, 7, 6], [4, 5, 8, 7]]
    kinds, orders = [1] * 4, [1] * 4
    base = len(xy)
    xy += [(3.0, 0.0), (4.0, 0.2), (3.1, 1.0)]
    els += [[base, base + 1, base + 2], [base + 1, 5, base + 2]]
    kinds += [0, 0]
    orders += [1, 1]
    q = len(xy)   # Q2: corners, edge midpoints (bulged), centre
    xy += [(5, 0), (6, 0), (6, 1), (5, 1), (5.5, -0.1), (6.1, 0.5), (5.5, 1.1), (4.9, 0.5), (5.5, 0.5)]
    els.append(list(range(q, q + 9)))
    kinds.append(1)
    orders.append(2)
    t = len(xy)   # P2 triangle
    xy += [(7, 0), (8, 0), (7, 1), (7.5, -0.05), (7.55, 0.55), (6.95, 0.5)]
    els.append(list(range(t, t + 6)))
    kinds.append(0)
    orders.append(2)
    return np.array(xy, np.float64), els, kinds, orders


def _states(n_elem, nsp, seed):
    rng = np.random.default_rng(seed)
    sol = []
    for e in range(n_elem):
        for _ in range(nsp[e]):
            rho, u, v, p = 1 + 0.2 * rng.random(), rng.normal(), rng.normal(), 1 + rng.random()
            sol += [rho, rho * u, rho * v, p / 0.4 + 0.5 * rho * (u * u + v * v)]
    return np.array(sol)


def _oracle(xy, els, kinds, orders, nsp, sol, n, gamma=1.4):
    offs = np.concatenate([[0], np.cumsum(np.asarray(nsp) * 4)])
    states = [tuple(sol[offs[e]:offs[e] + 4]) if nsp[e] > 0 else None for e in range(len(els))]
    return V.lattice([tuple(p) for p in xy], els, kinds, orders, states, n, gamma)


@pytest.mark.parametrize("n", [1, 3, 4])
def test_vtu_ascii_matches_reference_text(tmp_path, n):
    xy, els, kinds, orders = _mixed_mesh(1)
    nsp = [9, 4, 16, 9, 6, 3, 9, 6]
    sol = _states(len(els), nsp, 2)
    m = io.VisMesh(xy, els, kinds, orders)
    io.write_vtk(tmp_path / "a.vtu", m, nsp, sol, poly_order=2, subdivisions=n)
    want = V.vtu_text(*_oracle(xy, els, kinds, orders, nsp, sol, n))
    assert (tmp_path / "a.vtu").read_text() == want
    ET.parse(tmp_path / "a.vtu")


def test_legacy_matches_reference_text(tmp_path):
    xy, els, kinds, orders = _mixed_mesh(3)
    nsp = [4] * 4 + [3, 3, 4, 3]
    sol = _states(len(els), nsp, 4)
    m = io.VisMesh(xy, els, kinds, orders)
    io.write_vtk(tmp_path / "a.vtk", m, nsp, sol, poly_order=1, subdivisions=4, flags=io.LEGACY)
    assert (tmp_path / "a.vtk").read_text() == V.legacy_text(*_oracle(xy, els, kinds, orders, nsp, sol, 4))


def test_vtu_element_without_points(tmp_path):
    """An element with no solution points carries a zero State (NaN velocity
    and pressure), printed as glibc prints NaN."""
    xy, els, kinds, orders = _mixed_mesh(5)
    nsp = [4, 0, 4, 4, 3, 3, 0, 3]
    sol = _states(len(els), nsp, 6)
    m = io.VisMesh(xy, els, kinds, orders)
    for flags, fn in ((io.VTU, V.vtu_text), (io.LEGACY, V.legacy_text)):
        io.write_vtk(tmp_path / "z", m, nsp, sol, poly_order=1, subdivisions=2, flags=flags)
        assert (tmp_path / "z").read_text() == fn(*_oracle(xy, els, kinds, orders, nsp, sol, 2))


def test_pvtu_text(tmp_path):
    pieces = ["out/case_0.vtu", "out/case_1.vtu", "out/case_2.vtu"]
    io.write_pvtu(tmp_path / "c.pvtu", pieces)
    assert (tmp_path / "c.pvtu").read_text() == V.pvtu_text(pieces)
    ET.parse(tmp_path / "c.pvtu")


def test_vtu_binary_equals_oracle_values(tmp_path):
    xy, els, kinds, orders = _mixed_mesh(7)
    nsp = [9, 4, 16, 9, 6, 3, 9, 6]
    sol = _states(len(els), nsp, 8)
    m = io.VisMesh(xy, els, kinds, orders)
    io.write_vtk(tmp_path / "b.vtu", m, nsp, sol, poly_order=2, subdivisions=4, flags=io.BINARY)
    got = io.read_vtu_binary(tmp_path / "b.vtu")
    P, T, rho, pres, mac, vel = _oracle(xy, els, kinds, orders, nsp, sol, 4)
    assert np.array_equal(got["Points"][:, :2], np.array(P)) and not got["Points"][:, 2].any()
    assert np.array_equal(got["connectivity"].reshape(-1, 3), np.array(T))
    assert np.array_equal(got["offsets"], 3 * np.arange(1, len(T) + 1))
    assert np.all(got["types"] == 5)
    for k, ref in (("Density", rho), ("Pressure", pres), ("Mach", mac)):
        assert np.array_equal(got[k], np.array(ref)), k
    assert np.array_equal(got["Velocity"][:, :2], np.array(vel))


def test_oracle_lattice_closed_forms():
    """Pins of the restatement: on an affine quad [a,b]x[c,d] the lattice is
    the uniform grid; on an affine triangle the points are corner
    combinations; counts; every cell has positive orientation; values are the
    gas relations of the first point."""
    n = 4
    xy = [(1.0, 2.0), (3.0, 2.0), (3.0, 5.0), (1.0, 5.0), (4.0, 0.0), (6.0, 1.0), (4.5, 3.0)]
    st = (1.3, 0.26, -0.39, 2.8)
    P, T, rho, pres, mac, vel = V.lattice(xy, [[0, 1, 2, 3], [4, 5, 6]], [1, 0], [1, 1], [st, st], n, 1.4)
    assert len(P) == (n + 1) ** 2 + (n + 1) * (n + 2) // 2 and len(T) == 2 * n * n + n * n
    gx, gy = np.meshgrid(np.linspace(1, 3, n + 1), np.linspace(2, 5, n + 1))
    assert np.allclose(np.array(P[:(n + 1) ** 2]), np.stack([gx.ravel(), gy.ravel()], 1), atol=1e-15)
    k = (n + 1) ** 2
    a, b, c = (np.array(xy[i]) for i in (4, 5, 6))
    want = [a + (b - a) * i / n + (c - a) * j / n for j in range(n + 1) for i in range(n + 1 - j)]
    assert np.allclose(np.array(P[k:]), np.array(want), atol=1e-14)
    Pa = np.array(P)
    for t in T:
        u, v = Pa[t[1]] - Pa[t[0]], Pa[t[2]] - Pa[t[0]]
        assert u[0] * v[1] - u[1] * v[0] > 0
    r, ru, rv, rE = st
    p = 0.4 * (rE - 0.5 * (ru * ru + rv * rv) / r)
    assert np.allclose(pres, p, rtol=1e-14) and np.allclose(rho, r)
    assert np.allclose(mac, math.hypot(ru / r, rv / r) / math.sqrt(1.4 * p / r), rtol=1e-14)
    assert np.allclose(np.array(vel), [ru / r, rv / r])


def _poly(x, y, c):
    return c[0] + c[1] * x + c[2] * y + c[3] * x * y + c[4] * x * x + c[5] * y * y


@pytest.mark.parametrize("p", [2, 3, 5])
def test_vtu_interpolated_exact_for_polynomials(tmp_path, p):
    """ZJ_VTK_INTERPOLATE: a state that is a polynomial of degree <= 2 in
    (x, y) on affine elements is reproduced exactly at every lattice point
    (quadrilateral Gauss points, triangle warp-and-blend points)."""
    from zhijian_amd import _lib
    from zhijian_amd.tri import sp_coords
    rng = np.random.default_rng(p)
    C = rng.normal(size=(4, 6)) * 0.1
    C[0, 0] += 2.0
    C[3, 0] += 5.0
    xy = np.array([(0, 0), (1, 0.2), (1.3, 1.1), (0.2, 1.0), (2.0, 0.0), (2.6, 0.9)], np.float64)
    quad, tri = [0, 1, 2, 3], [1, 4, 5]
    ops = _lib.quad_operators(p)
    xi, eta = ops["sol_xi"], ops["sol_eta"]
    N = np.stack([0.25 * (1 - xi) * (1 - eta), 0.25 * (1 + xi) * (1 - eta), 0.25 * (1 + xi) * (1 + eta),
                  0.25 * (1 - xi) * (1 + eta)], 1)
    qpos = N @ xy[quad]
    tpos = sp_coords(xy, np.array([tri]), p)[0]
    sol = np.concatenate([np.stack([_poly(pp[:, 0], pp[:, 1], C[v]) for v in range(4)], 1).ravel()
                          for pp in (qpos, tpos)])
    m = io.VisMesh(xy, [quad, tri], [1, 0], [1, 1])
    nsp = [(p + 1) ** 2, (p + 1) * (p + 2) // 2]
    io.write_vtk(tmp_path / "h.vtu", m, nsp, sol, poly_order=p, subdivisions=5, flags=io.BINARY | io.INTERPOLATE)
    got = io.read_vtu_binary(tmp_path / "h.vtu")
    X, Y = got["Points"][:, 0], got["Points"][:, 1]
    rho = _poly(X, Y, C[0])
    assert np.allclose(got["Density"], rho, rtol=0, atol=1e-12)
    assert np.allclose(got["Velocity"][:, 0], _poly(X, Y, C[1]) / rho, rtol=0, atol=1e-11)
    E = _poly(X, Y, C[3])
    pres = 0.4 * rho * (E / rho - 0.5 * ((_poly(X, Y, C[1]) / rho) ** 2 + (_poly(X, Y, C[2]) / rho) ** 2))
    assert np.allclose(got["Pressure"], pres, rtol=0, atol=1e-11)


def test_vtk_errors(tmp_path):
    xy, els, kinds, orders = _mixed_mesh(0)
    m = io.VisMesh(xy, els, kinds, orders)
    nsp = [4] * 8
    sol = _states(8, nsp, 0)
    with pytest.raises(ZJError, match="do not match"):
        io.write_vtk(tmp_path / "x.vtu", m, nsp, sol, poly_order=2, flags=io.INTERPOLATE)
    with pytest.raises(ZJError, match="unknown flags"):
        io.write_vtk(tmp_path / "x.vtu", m, nsp, sol, poly_order=1, flags=64)
    with pytest.raises(ZJError, match="subdivisions"):
        io.write_vtk(tmp_path / "x.vtu", m, nsp, sol, poly_order=1, subdivisions=0)
    bad = io.VisMesh(xy, [[0, 1, 4, 99]], 1, 1)
    with pytest.raises(ZJError, match="out of range"):
        io.write_vtk(tmp_path / "x.vtu", bad, [4], sol[:16], poly_order=1)
    with pytest.raises(ZJError, match="Failed to open"):
        io.write_vtk(tmp_path / "nodir" / "x.vtu", m, nsp, sol, poly_order=1)


def test_restart_attribute_and_name_limits(tmp_path):
    """Argument checks of zj_restart_write: a solution that does not match
    elem_n_sp, a negative element count, negative points per element."""
    import ctypes as C
    from zhijian_amd import _lib
    nsp, sol = _sample(3, 9)
    with pytest.raises(ValueError):
        io.write_restart(tmp_path / "x.h5", nsp, sol[:-1])
    rc = _lib.lib().zj_restart_write(str(tmp_path / "y.h5").encode(), -1, None, None, 0.0, 0, 1, 1.4)
    assert rc != 0 and b"elem_n_sp" in _lib.lib().zj_io_last_error()
    bad = np.array([4, -1, 4], np.int32)
    rc = _lib.lib().zj_restart_write(str(tmp_path / "z.h5").encode(), 3, bad.ctypes.data, sol.ctypes.data,
                                     0.0, 0, 1, 1.4)
    assert rc != 0 and b"negative" in _lib.lib().zj_io_last_error()


@pytest.mark.parametrize("p", [2, 4])
def test_vtu_interpolated_exact_on_curved_quads(tmp_path, p):
    """A curved Q2 quadrilateral: a state polynomial of degree <= p in the
    reference coordinates (xi, eta) is reproduced at every lattice point
    (xi_i, eta_j) = (-1 + 2i/n, -1 + 2j/n), whatever the element's geometry."""
    from zhijian_amd import _lib
    ops = _lib.quad_operators(p)
    xi, eta = ops["sol_xi"], ops["sol_eta"]
    rng = np.random.default_rng(10 + p)
    c = rng.normal(size=(4, p + 1, p + 1)) * 0.1
    c[0, 0, 0] += 2.0
    c[3, 0, 0] += 6.0

    def field(v, X, Y):
        return sum(c[v, a, b] * X ** a * Y ** b for a in range(p + 1) for b in range(p + 1 - a))

    sol = np.stack([field(v, xi, eta) for v in range(4)], 1).ravel()
    xy = [(0, 0), (2, 0.1), (2.2, 1.9), (-0.1, 2), (1, -0.2), (2.3, 1), (1.1, 2.2), (0.1, 1.05), (1.05, 0.95)]
    m = io.VisMesh(np.array(xy, float), [list(range(9))], 1, 2)
    n = 5
    io.write_vtk(tmp_path / "c.vtu", m, (p + 1) ** 2, sol, poly_order=p, subdivisions=n,
                 flags=io.BINARY | io.INTERPOLATE)
    got = io.read_vtu_binary(tmp_path / "c.vtu")
    X, Y = np.meshgrid(-1 + 2 * np.arange(n + 1) / n, -1 + 2 * np.arange(n + 1) / n)
    X, Y = X.ravel(), Y.ravel()
    assert np.allclose(got["Density"], field(0, X, Y), rtol=0, atol=1e-12)
    assert np.allclose(got["Velocity"][:, 0], field(1, X, Y) / field(0, X, Y), rtol=0, atol=1e-11)
    # the lattice points themselves follow the Q2 map (compare with the oracle's shape functions)
    want = [V.subdivide(1, 2, xy, None, n)[0][k] for k in range(len(X))]
    assert np.allclose(got["Points"][:, :2], np.array(want), rtol=0, atol=1e-14)


def test_restart_read_checks_counts_before_allocating(tmp_path):
    """zj_restart_read with an element count other than the file's is refused
    (ZJ_ERR_ARG) from the dataset's header, before anything is sized from the
    caller's count; the right counts read the file."""
    from zhijian_amd import _lib
    nsp, sol = _sample(6, 4)
    path = tmp_path / "r.h5"
    io.write_restart(path, nsp, sol, 0.5, 3, 2, 1.4)
    L = _lib.lib()
    for n_elem in (5, 7, 1 << 40):
        rc = L.zj_restart_read(str(path).encode(), n_elem, None, sol.size, None)
        assert rc == 1 and b"elem_n_sp holds 6 entries" in L.zj_io_last_error()
    got_nsp = np.zeros(6, np.int32)
    got = np.zeros(sol.size)
    assert L.zj_restart_read(str(path).encode(), 6, got_nsp.ctypes.data, sol.size, got.ctypes.data) == 0
    assert got.tobytes() == sol.tobytes() and got_nsp.tolist() == list(nsp)


def test_vtk_refuses_bad_node_offsets(tmp_path):
    """Element node offsets that are negative or decreasing are refused
    before any node is read through them."""
    xy, els, kinds, orders = _mixed_mesh(0)
    m = io.VisMesh(xy, els, kinds, orders)
    nsp = [4] * 8
    sol = _states(8, nsp, 0)
    m.start[0] = -4
    with pytest.raises(ZJError, match="too few nodes"):
        io.write_vtk(tmp_path / "x.vtu", m, nsp, sol, poly_order=1)
    m = io.VisMesh(xy, els, kinds, orders)
    m.start[3], m.start[4] = m.start[4], m.start[3]
    with pytest.raises(ZJError, match="too few nodes"):
        io.write_vtk(tmp_path / "x.vtu", m, nsp, sol, poly_order=1)
