"""Native Gmsh reader (zj_gmsh_*) against the compiled reference's
GmshReader + Mesh::buildFaces on the reference's own test meshes (copied as
fixtures to tests/golden/msh/; expected outputs in tests/golden/meshes.npz,
written by tests/golden/make_golden.py from oracle/_ref).  Bit-exact."""
import os

import numpy as np
import pytest

from tests.conftest import ROOT, mesh_from_golden
from zhijian_amd import gmsh

MSH = os.path.join(ROOT, "tests", "golden", "msh")
NAMES = sorted(f[:-4] for f in os.listdir(MSH) if f.endswith(".msh"))


@pytest.mark.parametrize("name", NAMES)
def test_reader_matches_reference(golden_meshes, name):
    g = mesh_from_golden(golden_meshes, name)
    m = gmsh.read(os.path.join(MSH, name + ".msh"))
    assert np.array_equal(m.xy, g["xy"])
    info = np.stack([m.elem_kind, m.elem_order, m.elem_n_nodes], axis=1)
    assert np.array_equal(info, g["elem_info"])
    w = g["elem_nodes"].shape[1]
    en = np.full((m.elem_nodes.shape[0], max(w, m.elem_nodes.shape[1])), -1, np.int64)
    en[:, : m.elem_nodes.shape[1]] = m.elem_nodes
    assert np.array_equal(en[:, :w], g["elem_nodes"])
    faces, tag, typ = m.faces()
    assert np.array_equal(faces, g["faces"])
    assert np.array_equal(tag, g["face_bc_tag"])
    assert np.array_equal(typ, g["face_bc_type"])


def test_reader_errors(tmp_path):
    p = tmp_path / "bad.msh"
    p.write_text("$MeshFormat\n4.1 1 8\n$EndMeshFormat\n")
    with pytest.raises(RuntimeError, match="ASCII"):
        gmsh.read(str(p))
    p.write_text("$MeshFormat\n2.2 0 8\n$EndMeshFormat\n$Nodes\n1\n1 0 0 0\n$EndNodes\n"
                 "$Elements\n1\n1 3 2 0 1 1 2 3 4\n$EndElements\n")
    with pytest.raises(RuntimeError, match="unknown node tag"):
        gmsh.read(str(p))
    with pytest.raises(RuntimeError, match="cannot open"):
        gmsh.read(str(tmp_path / "missing.msh"))
