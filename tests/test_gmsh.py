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


def _msh_with_periodic(xy, elems, etype, links_x, links_y, v4):
    """MSH text of a 2D mesh (one node block, one element block) with a
    $Periodic section of two curve links, x then y, in the file format's own
    layout (tags 1-based); MSH 2.2 with an Affine line on the first link only."""
    o = ["$MeshFormat\n", "4.1 0 8\n" if v4 else "2.2 0 8\n", "$EndMeshFormat\n", "$Nodes\n"]
    N, E = len(xy), len(elems)
    if v4:
        o += [f"1 {N} 1 {N}\n", f"2 1 0 {N}\n"] + [f"{i + 1}\n" for i in range(N)]
        o += [f"{float(x)!r} {float(y)!r} 0\n" for x, y in xy]
    else:
        o += [f"{N}\n"] + [f"{i + 1} {float(x)!r} {float(y)!r} 0\n" for i, (x, y) in enumerate(xy)]
    o += ["$EndNodes\n", "$Elements\n"]
    if v4:
        o += [f"1 {E} 1 {E}\n", f"2 1 {etype} {E}\n"]
        o += [f"{e + 1} " + " ".join(str(int(n) + 1) for n in el) + "\n" for e, el in enumerate(elems)]
    else:
        o += [f"{E}\n"] + [f"{e + 1} {etype} 2 0 1 " + " ".join(str(int(n) + 1) for n in el) + "\n"
                           for e, el in enumerate(elems)]
    o += ["$EndElements\n", "$Periodic\n", "2\n"]
    for k, links in enumerate((links_x, links_y)):
        o.append(f"1 {2 + k} {4 - 3 * k}\n")
        if v4:
            o.append("16 1 0 0 -1 0 1 0 0 0 0 1 0 0 0 0 1\n" if k == 0 else "0\n")
        elif k == 0:
            o.append("Affine 1 0 0 -1 0 1 0 0 0 0 1 0 0 0 0 1\n")
        o.append(f"{len(links)}\n")
        o += [f"{int(a) + 1} {int(b) + 1}\n" for a, b in links]
    o.append("$EndPeriodic\n")
    return "".join(o)


@pytest.mark.parametrize("v4", [False, True])
@pytest.mark.parametrize("kind", ["quad", "sheared", "tri"])
def test_periodic_section_pairs_faces(tmp_path, v4, kind):
    """$Periodic node links read in both format versions, in file order, as
    0-based node indices; faces(periodic=True) joins the linked faces.  On
    the square meshes this equals pairing by the periods; on the sheared
    mesh every boundary face finds its partner, which period pairing cannot
    do.  Without periodic=True the faces are the reference reader's."""
    from zhijian_amd import _lib, cases
    n = 4
    if kind == "tri":
        m = cases.split_tri_mesh(n, jitter=0.2, seed=1)
        xy, els, etype = m.xy, m.tris, 2
    elif kind == "quad":
        m = cases.uniform_quad_mesh(n, n, 0.0, 1.0, 0.0, 1.0, jitter=0.2, seed=1)
        xy, els, etype = m.xy, m.elem_nodes, 3
    else:
        m = cases.sheared_quad_mesh(n, n, 1.0, 1.0, 0.3, jitter=0.2, seed=1)
        xy, els, etype = m.xy, m.elem_nodes, 3
    links = cases.periodic_links(n, n)
    p = tmp_path / "per.msh"
    p.write_text(_msh_with_periodic(xy, els, etype, links[: n + 1], links[n + 1:], v4))
    g = gmsh.read(str(p))
    assert np.array_equal(g.periodic_nodes, links)
    assert np.array_equal(g.xy, xy)
    plain, _, _ = g.faces()
    paired, tag, typ = g.faces(periodic=True)
    nv = 3 if kind == "tri" else 4
    assert np.array_equal(plain, _lib.build_faces(np.asarray(els), nv=nv))
    assert np.all(paired[:, 2] >= 0) and paired.shape[0] == len(els) * nv // 2
    assert np.all(tag == -1) and np.all(typ == 0)
    if kind != "sheared":
        assert np.array_equal(paired, _lib.pair_periodic(plain, np.asarray(els), xy, 1.0, 1.0, nv=nv))


@pytest.mark.parametrize("text,match", [
    (V2_HEAD + QUAD_V2 + "$Periodic\n1\n1 2 4\n1\n9 1\n$EndPeriodic\n", "unknown node tag 9"),
    (V2_HEAD + QUAD_V2 + "$Periodic\n1\n1 2 4\n3\n2 1\n", "end of file in \\$Periodic"),
    (V2_HEAD + QUAD_V2 + "$Periodic\n1\n1 2 4\nx\n$EndPeriodic\n", "periodic node count"),
    (V4_HEAD + "$Nodes\n1 4 1 4\n2 1 0 4\n1\n2\n3\n4\n0 0 0\n1 0 0\n1 1 0\n0 1 0\n$EndNodes\n"
     "$Elements\n1 1 1 1\n2 1 3 1\n1 1 2 3 4\n$EndElements\n$Periodic\n1\n1 2 4\n-16\n$EndPeriodic\n",
     "affine line"),
])
def test_periodic_section_errors(tmp_path, text, match):
    p = tmp_path / "bad.msh"
    p.write_text(text)
    with pytest.raises(RuntimeError, match=match):
        gmsh.read(str(p))
