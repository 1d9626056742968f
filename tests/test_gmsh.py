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


V2_HEAD = "$MeshFormat\n2.2 0 8\n$EndMeshFormat\n"
V4_HEAD = "$MeshFormat\n4.1 0 8\n$EndMeshFormat\n"
QUAD_V2 = ("$Nodes\n4\n1 0 0 0\n2 1 0 0\n3 1 1 0\n4 0 1 0\n$EndNodes\n"
           "$Elements\n1\n1 3 2 0 1 1 2 3 4\n$EndElements\n")


@pytest.mark.parametrize("text,match", [
    # counts that promise more lines than the file holds
    (V2_HEAD + "$Nodes\n4\n1 0 0 0\n2 1 0 0\n", "end of file in \\$Nodes"),
    (V2_HEAD + "$Nodes\n4\n1 0 0 0\n2 1 0 0\n3 1 1 0\n4 0 1 0\n$EndNodes\n$Elements\n3\n1 3 2 0 1 1 2 3 4\n",
     "end of file in \\$Elements"),
    # entity blocks that hold more nodes than the section header declares
    (V4_HEAD + "$Nodes\n1 2 1 2\n2 1 0 3\n1\n2\n3\n0 0 0\n1 0 0\n1 1 0\n$EndNodes\n", "entity block"),
    # counts and tags that are not numbers, or negative
    (V2_HEAD + "$Nodes\n-3\n$EndNodes\n", "\\$Nodes header"),
    (V2_HEAD + "$Nodes\nfour\n$EndNodes\n", "\\$Nodes header"),
    (V4_HEAD + "$Nodes\n1 1 1 1\n2 1 0 1\nx\n0 0 0\n$EndNodes\n", "node tag"),
    (V2_HEAD + "$Nodes\n1\n1 0 zero 0\n$EndNodes\n", "node line"),
    (V2_HEAD + "$PhysicalNames\nmany\n$EndPhysicalNames\n" + QUAD_V2, "\\$PhysicalNames header"),
    (V4_HEAD + "$Entities\n0 x 0 0\n$EndEntities\n", "\\$Entities header"),
    (V2_HEAD + "$Elements\n0\n$EndElements\n", "\\$Elements before \\$Nodes"),
    # an element line cut short, and repeated sections
    (V2_HEAD + "$Nodes\n4\n1 0 0 0\n2 1 0 0\n3 1 1 0\n4 0 1 0\n$EndNodes\n$Elements\n1\n1 3 2 0 1 1 2 3\n"
     "$EndElements\n", "fewer than 4 node tags"),
    (V2_HEAD + "$Nodes\n1\n1 0 0 0\n$EndNodes\n" + QUAD_V2, "more than one \\$Nodes"),
    (V2_HEAD + QUAD_V2 + "$Elements\n0\n$EndElements\n", "more than one \\$Elements"),
    ("$MeshFormat\n$EndMeshFormat\n" + QUAD_V2, "\\$MeshFormat line"),
])
def test_reader_refuses_malformed_files(tmp_path, text, match):
    """Truncated sections, counts that disagree with the lines that follow,
    non-numeric fields and repeated sections end in an error naming the fault;
    no count read from the file sizes an allocation or an index."""
    p = tmp_path / "bad.msh"
    p.write_text(text)
    with pytest.raises(RuntimeError, match=match):
        gmsh.read(str(p))


def test_reader_small_quad_files(tmp_path):
    """The well-formed one-element files the malformed cases are cut from read
    as one bilinear quadrilateral, in both format versions."""
    p = tmp_path / "ok.msh"
    p.write_text(V2_HEAD + QUAD_V2)
    m = gmsh.read(str(p))
    assert m.xy.tolist() == [[0, 0], [1, 0], [1, 1], [0, 1]]
    assert m.elem_kind.tolist() == [1] and m.elem_order.tolist() == [1]
    p.write_text(V4_HEAD + "$Nodes\n1 4 1 4\n2 1 0 4\n1\n2\n3\n4\n0 0 0\n1 0 0\n1 1 0\n0 1 0\n$EndNodes\n"
                 "$Elements\n1 1 1 1\n2 1 3 1\n1 1 2 3 4\n$EndElements\n")
    m4 = gmsh.read(str(p))
    assert m4.xy.tolist() == m.xy.tolist()
    assert m4.elem_nodes[:, :4].tolist() == [[0, 1, 2, 3]]
