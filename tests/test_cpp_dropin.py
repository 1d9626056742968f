"""The C++ drop-in API (include/zhijian/fr_solver.hpp) through a driver program
that uses it the way the reference's main() does (tests/cpp/frsolver_driver.cpp).
"""
import json
import os
import subprocess

import numpy as np
import pytest

from oracle import pyoracle as O
from tests import fr_case
from tests.conftest import ROOT
from zhijian_amd import cases

DRIVER = os.path.join(ROOT, "tests", "cpp", "_build", "frsolver_driver")


def _need_driver():
    if not os.path.exists(DRIVER):
        pytest.fail(f"{DRIVER} missing: run `python -m zhijian_amd.build`")


def test_mesh_api_matches_oracle(tmp_path):
    """Mesh::buildFaces + pairPeriodic + faceNormal (C++ API, no GPU) against the oracle."""
    _need_driver()
    out = tmp_path / "m.bin"
    subprocess.run([DRIVER, "mesh", "7", str(out)], check=True, capture_output=True)
    got = np.fromfile(out, dtype=np.float64).reshape(-1, 6)
    mesh = cases.uniform_quad_mesh(7, 7, 0.0, 1.0, 0.0, 1.0)
    f = O.pair_periodic(O.build_faces(mesh.elem_nodes), mesh.elem_nodes, mesh.xy, 1.0, 1.0, 1e-9)
    assert np.array_equal(got[:, :4].astype(np.int64), f.as_array())
    assert np.allclose(got[:, 4:], O.face_normals(f, mesh.elem_nodes, mesh.xy), rtol=0, atol=1e-15)


@pytest.mark.parametrize("name", ["quad_v41", "tri_v22", "serendipity_quad_v41", "uniform_quad_10x10"])
def test_gmsh_reader_api_matches_reference(tmp_path, golden_meshes, name):
    """GmshReader::read + Mesh::buildFaces + boundary tags (C++ API, no GPU)
    against the compiled reference on its own test meshes."""
    _need_driver()
    from tests.conftest import mesh_from_golden
    out = tmp_path / "f.bin"
    subprocess.run([DRIVER, "gmsh", os.path.join(ROOT, "tests", "golden", "msh", name + ".msh"), str(out)],
                   check=True, capture_output=True)
    got = np.fromfile(out, dtype=np.float64).reshape(-1, 6).astype(np.int64)
    g = mesh_from_golden(golden_meshes, name)
    assert np.array_equal(got[:, :4], g["faces"])
    assert np.array_equal(got[:, 4], g["face_bc_type"])
    assert np.array_equal(got[:, 5], g["face_bc_tag"])


@pytest.mark.parametrize("v4", [False, True])
def test_gmsh_periodic_links_cpp_api(tmp_path, v4):
    """GmshReader keeps a file's $Periodic links (Mesh::periodicNodes) and
    Mesh::pairPeriodicNodes joins the faces they link, as the C ABI and the
    oracle do, on a sheared doubly periodic mesh (no GPU)."""
    _need_driver()
    from tests.test_gmsh import _msh_with_periodic
    n = 5
    m = cases.sheared_quad_mesh(n, n, 1.0, 1.0, -0.4, jitter=0.15, seed=4)
    links = cases.periodic_links(n, n)
    msh = tmp_path / "s.msh"
    msh.write_text(_msh_with_periodic(m.xy, m.elem_nodes, 3, links[: n + 1], links[n + 1:], v4))
    out = tmp_path / "f.bin"
    r = subprocess.run([DRIVER, "gmsh", str(msh), str(out), "periodic"], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    info = json.loads(r.stdout.strip().splitlines()[-1])
    assert info["periodic_links"] == len(links) and info["faces"] == 2 * n * n
    got = np.fromfile(out, dtype=np.float64).reshape(-1, 6).astype(np.int64)
    want = O.pair_periodic_nodes(O.build_faces(m.elem_nodes), m.elem_nodes, m.xy, links)
    assert np.array_equal(got[:, :4], want.as_array())
    assert np.all(got[:, 4] == 0) and np.all(got[:, 5] == -1)   # every face interior


def _run(tmp_path, kind, n, p, steps, dt_max, prec):
    prefix = str(tmp_path / f"{kind}{n}p{p}")
    print(f"driver: {kind} n={n} P{p} steps={steps} {prec}", flush=True)
    r = subprocess.run([DRIVER, "run", kind, str(n), str(p), str(steps), repr(dt_max), prec, prefix],
                       capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stderr
    info = json.loads(r.stdout.strip().splitlines()[-1])
    U0 = np.fromfile(prefix + "_U0.bin", dtype=np.float64)
    U = np.fromfile(prefix + "_U.bin", dtype=np.float64)
    return info, U0, U


@pytest.mark.gpu
def test_frsolver_vortex_matches_oracle(tmp_path):
    """C1 geometry (10x10 on [-5,5]^2, P3), dt = min(CFL dt, 1e-3), 10 steps
    (inside the reference scheme's pre-blow-up window)."""
    _need_driver()
    info, U0, U = _run(tmp_path, "vortex", 10, 3, 10, 1e-3, "fp64")
    assert info["iteration"] == 10 and len(info["dt"]) == 10
    mesh = cases.vortex_mesh(10)
    o = fr_case.from_quadmesh(mesh, 3).solver()
    V = U0.copy()
    for dt in info["dt"]:
        assert dt <= 1e-3
        V = o.advance(V, dt, True)
    assert np.linalg.norm(U - V) / np.linalg.norm(V) < 1e-12
    assert abs(info["residual"] - o.residual_l2()) <= 1e-12 * o.residual_l2()
    assert abs(info["time"] - sum(info["dt"])) < 1e-15


@pytest.mark.gpu
def test_frsolver_tgv_fp32bulk_runs(tmp_path):
    _need_driver()
    info, U0, U = _run(tmp_path, "tgv", 32, 4, 3, 0.0, "fp32bulk")
    _, _, U64 = _run(tmp_path, "tgv", 32, 4, 3, 0.0, "fp64")
    assert np.all(np.isfinite(U))
    assert np.linalg.norm(U - U64) / np.linalg.norm(U64) < 1e-5


def _driver_mesh(n, x0, x1):
    """The driver's uniform_mesh: nodes x0 + i h (i fastest), CCW quads."""
    h = (x1 - x0) / n
    xy = [(x0 + i * h, x0 + j * h) for j in range(n + 1) for i in range(n + 1)]
    els = []
    for j in range(n):
        for i in range(n):
            n0 = j * (n + 1) + i
            els.append([n0, n0 + 1, n0 + n + 2, n0 + n + 1])
    return xy, els


def test_io_api(tmp_path):
    """RestartIO, VTKWriter, ResidualWriter (C++ API, no GPU): restart round
    trips, the reference's VTU / legacy / PVTU text byte for byte, binary
    and interpolated VTU equal to the C ABI's output, residual CSV text."""
    _need_driver()
    from oracle import vtk_oracle as V
    from zhijian_amd import io
    r = subprocess.run([DRIVER, "io", str(tmp_path)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    info = json.loads(r.stdout.strip().splitlines()[-1])
    assert all(info.values()), info
    U = np.fromfile(tmp_path / "U.bin", dtype=np.float64)
    xy, els = _driver_mesh(3, 0.0, 1.0)
    E, S = len(els), 9
    first = [tuple(U[e * S * 4:e * S * 4 + 4]) for e in range(E)]

    rs = io.read_restart(tmp_path / "r.h5")
    assert rs["solution"].tobytes() == U.tobytes() and np.all(rs["elem_n_sp"] == S)
    assert rs["time"] == 0.75 and rs["iteration"] == 120 and rs["poly_order"] == 2 and rs["gamma"] == 1.4

    for name, n, fn in (("a.vtu", 4, V.vtu_text), ("a.vtk", 4, V.legacy_text), ("par_0.vtu", 2, V.vtu_text)):
        want = fn(*V.lattice(xy, els, [1] * E, [1] * E, first, n, 1.4))
        assert (tmp_path / name).read_text() == want, name
    pieces = [str(tmp_path / "par_0.vtu"), str(tmp_path / "par_1.vtu")]
    assert (tmp_path / "par.pvtu").read_text() == V.pvtu_text(pieces)

    m = io.VisMesh(np.array(xy), np.array(els), 1, 1)
    for name, flags in (("b.vtu", io.BINARY), ("h.vtu", io.BINARY | io.INTERPOLATE)):
        io.write_vtk(tmp_path / ("c_" + name), m, S, U, poly_order=2, subdivisions=4, flags=flags)
        assert (tmp_path / name).read_bytes() == (tmp_path / ("c_" + name)).read_bytes(), name
    assert (tmp_path / "b.vtu").read_bytes() != (tmp_path / "h.vtu").read_bytes()

    want = ("# Iteration, Time, Residual, Rho_res, RhoU_res, RhoV_res, RhoE_res\n"
            "10, %.8e, %.8e, %.8e, %.8e, %.8e, %.8e\n" % (0.5, 1.25e-3, 1.0, 2.5e-7, -3.0, 4.125) +
            "20, %.8e, %.8e, %.8e, %.8e, %.8e, %.8e\n" % (1.0, 6.5e-4, 0, 0, 0, 0))
    assert (tmp_path / "res.csv").read_text() == want


@pytest.mark.gpu
def test_frsolver_resume_from_restart_bit_exact(tmp_path):
    """main()'s restart path completed: 6 steps, RestartIO::write, a new
    solver from RestartIO::read + setSolution, 4 more steps == 10 steps
    straight through, bit for bit (vortex 10 x 10, P3)."""
    _need_driver()
    print("driver: resume vortex n=10 P3 6+4 steps", flush=True)
    r = subprocess.run([DRIVER, "resume", str(tmp_path), "10", "3", "6", "4"], capture_output=True, text=True,
                       timeout=240)
    assert r.returncode == 0, r.stderr
    info = json.loads(r.stdout.strip().splitlines()[-1])
    a = np.fromfile(tmp_path / "resumed.bin", dtype=np.float64)
    b = np.fromfile(tmp_path / "straight.bin", dtype=np.float64)
    assert a.tobytes() == b.tobytes()
    assert info["time"] == info["time_straight"] and info["iteration"] == info["iteration_straight"] == 10


@pytest.mark.gpu
def test_frsolver_on_periodic_gmsh_file(tmp_path):
    """GmshReader + Mesh::pairPeriodicNodes + FRSolver (textbook, Roe, HU) on
    a sheared doubly periodic mesh given as an MSH 4.1 file with $Periodic
    links: 5 steps against the textbook oracle on the same face table."""
    _need_driver()
    import math
    from tests.test_gmsh import _msh_with_periodic
    n, p, steps, dt = 6, 3, 5, 2e-3
    m = cases.sheared_quad_mesh(n, n, 2 * math.pi, 2 * math.pi, 1.1, jitter=0.1, seed=6)
    links = cases.periodic_links(n, n)
    msh = tmp_path / "tgv.msh"
    msh.write_text(_msh_with_periodic(m.xy, m.elem_nodes, 3, links[: n + 1], links[n + 1:], True))
    prefix = str(tmp_path / "g")
    print(f"driver: gmshrun sheared {n}x{n} P{p} {steps} steps", flush=True)
    r = subprocess.run([DRIVER, "gmshrun", str(msh), str(p), str(steps), repr(dt), prefix],
                       capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stderr
    info = json.loads(r.stdout.strip().splitlines()[-1])
    assert info["boundary_faces"] == 0 and info["faces"] == 2 * n * n
    U0 = np.fromfile(prefix + "_U0.bin", dtype=np.float64)
    U = np.fromfile(prefix + "_U.bin", dtype=np.float64)
    oc = fr_case.build(m.xy, m.elem_nodes, p, periodic_nodes=links)
    tb = oc.textbook("roe", "hu")
    V = U0.copy()
    for _ in range(steps):
        V = tb.advance(V, dt, False)
    assert np.linalg.norm(U - V) / np.linalg.norm(V) < 1e-10
    assert np.linalg.norm(U - U0) > 1e-6 * np.linalg.norm(U0)
