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


# ------------------------------------------------------------ VTU output
def _parse_vtu(path):
    import xml.etree.ElementTree as ET
    root = ET.parse(path).getroot()
    piece = root.find("UnstructuredGrid/Piece")
    arr = lambda node: np.array(node.text.split(), float)
    pts = arr(piece.find("Points/DataArray")).reshape(-1, 3)
    cells = {d.get("Name"): arr(d).astype(int) for d in piece.find("Cells")}
    pdata = {d.get("Name"): (int(d.get("NumberOfComponents")), arr(d)) for d in piece.findall("PointData/DataArray")}
    cdata = {d.get("Name"): (int(d.get("NumberOfComponents")), arr(d)) for d in piece.findall("CellData/DataArray")}
    return piece, pts, cells, pdata, cdata


def _signed_area(P):
    x, y = P[:, 0], P[:, 1]
    return 0.5 * np.sum(x * np.roll(y, -1) - np.roll(x, -1) * y)


def _hex_volume(P):
    # VTK hexahedron: bottom 0-1-2-3 counter-clockwise seen from above, top 4-7
    tets = [(0, 1, 3, 4), (1, 2, 3, 6), (1, 3, 4, 6), (1, 4, 5, 6), (3, 4, 6, 7)]
    return sum(np.linalg.det(np.stack([P[b] - P[a], P[c] - P[a], P[d] - P[a]])) / 6 for a, b, c, d in tets)


@pytest.mark.parametrize("dim,p", [(2, 1), (2, 2), (2, 3), (3, 1), (3, 2)])
def test_vtu_sub_cells_tile_each_element_with_positive_orientation(H, tmp_path, dim, p, orc):
    """OutputMesh draws each order-p element as p^dim sub-cells of its local
    node grid: every sub-cell has positive area / volume, the sub-cells of an
    element use each of its nodes and together cover its area / volume."""
    el, dn = orc.mesh(dim, p, 2)
    u = np.sin(dn[:, 0]) + dn[:, 1]
    base = tmp_path / "m"
    H.write_vtu(base, dim, p, dn, el, [("Solution_0", u, [1], H.SCALAR, H.POINT_DATA)])
    piece, pts, cells, pdata, _ = _parse_vtu(str(base) + ".vtu")
    nv = 4 if dim == 2 else 8
    nsub = p ** dim
    assert int(piece.get("NumberOfPoints")) == len(dn) and int(piece.get("NumberOfCells")) == len(el) * nsub
    assert np.allclose(pts, dn, rtol=1e-5, atol=1e-6)
    conn = cells["connectivity"].reshape(-1, nv)
    assert np.array_equal(cells["offsets"], nv * np.arange(1, len(conn) + 1))
    assert set(cells["types"]) == {9 if dim == 2 else 12}
    h = 1.0 / 2
    for e in range(len(el)):
        sub = conn[e * nsub:(e + 1) * nsub]
        assert set(sub.ravel()) == set(el[e])
        meas = [(_signed_area(pts[s, :2]) if dim == 2 else _hex_volume(pts[s])) for s in sub]
        assert min(meas) > 0
        assert abs(sum(meas) - h ** dim) < 1e-5
    nc, vals = pdata["Solution_0"]
    assert nc == 1 and np.allclose(vals, u, rtol=1e-5, atol=1e-6)


def test_vtu_data_arrays_and_number_format(H, tmp_path):
    """Point vectors with two of three components (padded with 0), a tensor
    array (written like a vector, as the reference does), cell data repeated
    on each sub-cell, and the ostream-default (%g, 6 digits) number text."""
    nodes, cells = H.structured_mesh(2, 2, 1, 0, 1)
    n = len(nodes)
    V = np.arange(3 * n, dtype=float).reshape(n, 3) / 7.0
    base = tmp_path / "d"
    H.write_vtu(base, 2, 1, nodes, cells,
                [("v", V.ravel(), [1, 0, 1], H.VECTOR, H.POINT_DATA), ("t", V.ravel(), [0, 1, 0], H.TENSOR, H.POINT_DATA),
                 ("c", np.array([1e-7, 123456789.0]), [1], H.SCALAR, H.CELL_DATA)])
    text = open(str(base) + ".vtu").read()
    _, _, _, pdata, cdata = _parse_vtu(str(base) + ".vtu")
    nc, v = pdata["v"]
    assert nc == 3 and np.allclose(v.reshape(n, 3), np.stack([V[:, 0], V[:, 2], 0 * V[:, 0]], 1), rtol=1e-5)
    nc, t = pdata["t"]
    assert nc == 9 and np.allclose(t.reshape(n, 3), np.stack([V[:, 1], 0 * V[:, 0], 0 * V[:, 0]], 1), rtol=1e-5)
    nc, c = cdata["c"]
    assert nc == 1 and list(c) == [1e-7, 1.23457e8]  # six significant digits
    assert "\n1e-07\n1.23457e+08\n" in text and "0.142857 " in text
    assert text.startswith('<?xml version="1.0"?>\n<VTKFile type="UnstructuredGrid" version="0.1">\n<UnstructuredGrid>\n')
    assert text.endswith("</Piece>\n</UnstructuredGrid>\n</VTKFile>\n")


def test_vtu_rejects_what_the_reference_rejects(H, tmp_path):
    nodes, cells = H.structured_mesh(2, 1, 1, 0, 1)
    u = np.zeros(2 * len(nodes))
    with pytest.raises(H.YafelHipError):  # two components in a scalar array
        H.write_vtu(tmp_path / "x", 2, 1, nodes, cells, [("s", u, [1, 1], H.SCALAR, H.POINT_DATA)])
    with pytest.raises(H.YafelHipError):  # node index outside the points
        H.write_vtu(tmp_path / "x", 2, 1, nodes, cells + 10, [])
