"""AddressSanitizer / UBSan run of the library's host-side parsers and mesh
routines (zj_h5.cpp, zj_gmsh.cpp, zj_host.cpp), the code that reads files a
user hands in.  tests/cpp/sanitize_driver.cpp is built with g++ (no GPU code)
and fed the committed fixtures, then seeded mutations of them: byte flips,
tail cuts, dropped spans and edited digits.  Every call must return, with
success or an error code; the sanitizers abort the driver on any invalid
access, leak or undefined behaviour.  CPU only."""
import json
import os
import shutil
import subprocess

import pytest

from tests.conftest import ROOT

CSRC = os.path.join(ROOT, "zhijian_amd", "csrc")
SRC = os.path.join(ROOT, "tests", "cpp", "sanitize_driver.cpp")
OUT = os.path.join(ROOT, "tests", "cpp", "_build_san", "sanitize_driver")
GOLD = os.path.join(ROOT, "tests", "golden")


@pytest.fixture(scope="module")
def driver():
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    os.makedirs(os.path.dirname(OUT), exist_ok=True)
    out = "%s.%d" % (OUT, os.getpid())
    tmp = out + ".tmp"
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
           "-fno-omit-frame-pointer", "-I", os.path.join(ROOT, "include"), "-I", CSRC, SRC,
           os.path.join(CSRC, "zj_h5.cpp"), os.path.join(CSRC, "zj_gmsh.cpp"), os.path.join(CSRC, "zj_host.cpp"),
           "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0 and "asan" in (r.stderr + r.stdout).lower():
        pytest.skip("the sanitizer runtime is not installed: " + r.stderr[-200:])
    assert r.returncode == 0, r.stderr
    # each test process builds its own copy and swaps it in whole, so parallel
    # workers never execute a file another one is still writing
    os.replace(tmp, out)
    return out


def _run(args, tmp_path):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1", UBSAN_OPTIONS="halt_on_error=1")
    r = subprocess.run(args, capture_output=True, text=True, timeout=300, env=env)
    assert r.returncode == 0, (r.stdout[-2000:], r.stderr[-4000:])
    return json.loads(r.stdout.strip().splitlines()[-1])


@pytest.mark.parametrize("name", ["restart_libhdf5.h5", "restart_libhdf5_par.h5.rank0"])
def test_hdf5_reader_under_sanitizers(driver, tmp_path, name):
    out = _run([driver, "h5", os.path.join(GOLD, name), "600", "11", str(tmp_path)], tmp_path)
    assert out["base_ok"] >= 1 and out["runs"] == 601


def test_hdf5_writer_output_under_sanitizers(driver, tmp_path):
    import numpy as np
    from zhijian_amd import io
    p = tmp_path / "ours.h5"
    io.write_restart(p, np.full(7, 9, np.int32), np.arange(7 * 9 * 4, dtype=np.float64), 0.25, 3, 2, 1.4)
    out = _run([driver, "h5", str(p), "600", "12", str(tmp_path)], tmp_path)
    assert out["base_ok"] >= 1


@pytest.mark.parametrize("name", ["quad_v41", "tri_v22", "serendipity_quad_v41", "tri_order2_v41"])
def test_gmsh_reader_under_sanitizers(driver, tmp_path, name):
    out = _run([driver, "gmsh", os.path.join(GOLD, "msh", name + ".msh"), "400", "13", str(tmp_path)], tmp_path)
    assert out["base_ok"] == 1


def test_gmsh_periodic_section_under_sanitizers(driver, tmp_path):
    from tests.test_gmsh import _msh_with_periodic
    from zhijian_amd import cases
    m = cases.sheared_quad_mesh(4, 4, 1.0, 1.0, 0.2)
    links = cases.periodic_links(4, 4)
    for v4 in (False, True):
        p = tmp_path / f"per{int(v4)}.msh"
        p.write_text(_msh_with_periodic(m.xy, m.elem_nodes, 3, links[:5], links[5:], v4))
        out = _run([driver, "gmsh", str(p), "400", "14", str(tmp_path)], tmp_path)
        assert out["base_ok"] == 1


def test_pairing_under_sanitizers(driver, tmp_path):
    out = _run([driver, "pair", "200", "15"], tmp_path)
    assert out["ok"] > 0


def test_oracle_under_sanitizers(tmp_path):
    """The CPU oracle itself (test infrastructure) under gcc AddressSanitizer
    + UBSan: every order's reference-arithmetic residual, limited steps, CFL dt
    and norm, and the textbook residual for every Riemann solver and
    correction family with the BR2 terms (tests/cpp/sanitize_oracle.c)."""
    if shutil.which("gcc") is None:
        pytest.skip("gcc not available")
    exe = os.path.join(ROOT, "tests", "cpp", "_build_san", "sanitize_oracle")
    ora = os.path.join(ROOT, "oracle")
    cmd = ["gcc", "-std=gnu11", "-O1", "-g", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
           "-fno-omit-frame-pointer", "-ffp-contract=off", "-I", ora,
           os.path.join(ROOT, "tests", "cpp", "sanitize_oracle.c"), os.path.join(ora, "zj_oracle.c"),
           os.path.join(ora, "zj_textbook.c"), "-lm", "-o", exe]
    os.makedirs(os.path.dirname(exe), exist_ok=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0 and "asan" in (r.stderr + r.stdout).lower():
        pytest.skip("the sanitizer runtime is not installed: " + r.stderr[-200:])
    assert r.returncode == 0, r.stderr
    out = _run([exe, "4"], tmp_path)
    assert out["finite"] == 1
