"""gmsh 2.2 input (Mesh(fname) -> Mesh::parse_gmsh, src/mesh/parse_gmsh.cpp:
80-151) through the C ABI.  Host code: these run without a GPU."""
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def H():
    from yafel_amd import hip
    return hip


def test_square_msh_matches_the_reference_test(H):
    """test/test_GmshMesh.cpp:12-44 on the reference's test/square.msh: five
    nodes and four triangles with 0-based connectivity {3,0,4}, {1,4,0},
    {1,2,4}, {2,3,4}."""
    g = H.GmshMesh(os.path.join(HERE, "golden", "square.msh"))
    x = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0.5, 0.5, 0]], float)
    assert np.array_equal(g.nodes, x)
    assert [list(g.cell(c)) for c in range(4)] == [[3, 0, 4], [1, 4, 0], [1, 2, 4], [2, 3, 4]]
    assert list(g.cell_types) == [H.CELL_TRI3] * 4
    assert all(list(g.cell_tags(c)) == [1, 10] for c in range(4))
    assert list(g.cell_offsets) == [0, 3, 6, 9, 12]


def test_minsquare_msh_counts_and_lines(H):
    """The reference's test/minsquare.msh: the declared counts, and nodes /
    cells exactly as written on their lines."""
    path = os.path.join(HERE, "golden", "minsquare.msh")
    g = H.GmshMesh(path)
    lines = open(path).read().split("\n")
    i = lines.index("$Nodes")
    nn = int(lines[i + 1])
    assert g.nodes.shape == (nn, 3) == (568, 3)
    for ln in lines[i + 2:i + 2 + nn]:
        w = ln.split()
        assert np.array_equal(g.nodes[int(w[0]) - 1], [float(v) for v in w[1:4]])
    j = lines.index("$Elements")
    ne = int(lines[j + 1])
    assert len(g.cell_types) == ne
    for ln in lines[j + 2:j + 2 + ne]:
        w = [int(v) for v in ln.split()]
        c = w[0] - 1
        assert list(g.cell_tags(c)) == w[3:3 + w[2]]
        assert list(g.cell(c)) == [v - 1 for v in w[3 + w[2]:]]
    assert g.cell_offsets[-1] == g.cell_nodes.size


def _write_msh(path, nodes, cells, celltype, extra=(), crlf=False, spaces=" "):
    nl = "\r\n" if crlf else "\n"
    out = ["$MeshFormat", "2.2 0 8", "$EndMeshFormat", "$PhysicalNames", "1", '2 9 "domain"', "$EndPhysicalNames",
           "$Nodes", str(len(nodes))]
    out += [spaces.join([str(i + 1)] + [repr(float(v)) for v in x]) for i, x in enumerate(nodes)]
    out += ["$EndNodes", "$Elements", str(len(extra) + len(cells))]
    k = 1
    for typ, tags, nd in extra:
        out.append(" ".join(str(v) for v in [k, typ, len(tags), *tags, *[n + 1 for n in nd]]))
        k += 1
    for c in cells:
        out.append(" ".join(str(v) for v in [k, celltype, 2, 9, 1, *[n + 1 for n in c]]))
        k += 1
    out += ["$EndElements"]
    with open(path, "w", newline="") as f:
        f.write(nl.join(out) + nl)


@pytest.mark.parametrize("crlf,spaces", [(False, " "), (True, "  ")])
def test_round_trip_of_a_quad_mesh_with_boundary_cells(H, tmp_path, crlf, spaces):
    nodes, cells = H.structured_mesh(2, 5, 4, 0, 2)
    extra = [(1, (7, 1), (0, 1)), (15, (8, 1), (0,))]
    p = tmp_path / "q.msh"
    _write_msh(p, nodes, cells, 3, extra, crlf, spaces)
    g = H.GmshMesh(p)
    assert np.array_equal(g.nodes, nodes)
    assert list(g.cell_types) == [H.CELL_LINE2, H.CELL_POINT1] + [H.CELL_QUAD4] * len(cells)
    assert np.array_equal(g.volume_cells(2), cells)
    assert list(g.cell_tags(0)) == [7, 1] and list(g.cell(1)) == [0]


def test_hex_mesh_volume_cells_skip_faces(H, tmp_path):
    nodes, cells = H.structured_mesh(3, 2, 2, 2, 1)
    p = tmp_path / "h.msh"
    _write_msh(p, nodes, cells, 5, [(3, (5, 2), (0, 1, 4, 3))])
    g = H.GmshMesh(p)
    assert np.array_equal(g.volume_cells(3), cells)
    with pytest.raises(H.YafelHipError):
        H.GmshMesh(os.path.join(HERE, "golden", "square.msh")).volume_cells(2)  # triangles


def test_unknown_type_is_none_and_errors_are_reported(H, tmp_path):
    p = tmp_path / "u.msh"
    p.write_text("$Nodes\n2\n1 0 0 0\n2 1 0 0\n$EndNodes\n$Elements\n1\n1 8 0 1 2\n$EndElements\n")
    assert list(H.GmshMesh(p).cell_types) == [H.CELL_NONE]
    with pytest.raises(H.YafelHipError):
        H.GmshMesh(tmp_path / "missing.msh")
    bad = tmp_path / "b.msh"
    bad.write_text("$Nodes\n2\n1 0 0 0\n3 1 0 0\n$EndNodes\n")  # node id beyond the count
    with pytest.raises(H.YafelHipError):
        H.GmshMesh(bad)
    bad.write_text("$Nodes\n1\n1 0 x 0\n$EndNodes\n")
    with pytest.raises(H.YafelHipError):
        H.GmshMesh(bad)
