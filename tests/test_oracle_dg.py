"""CPU checks that pin the DG oracle (DGAssembly / DGRK4 restated for Quad4
meshes): face lists by brute force, dof nodes by direct interpolation, and
closed forms of the residual -- a constant source (exercises the element mass
inverse), a linear field under constant advection (the volume and face terms
together), the outflow balance of a constant field (the face terms alone) and
RK4 on a constant right-hand side."""
import numpy as np
import pytest

from oracle.oracle import physics

DG = 3  # YAFEL_HIP_PHYS_DG_ADVECTION


def rect_mesh(orc, nx, ny, Lx=1.0, Ly=1.0, jitter=0.0, seed=0):
    """Counter-clockwise (gmsh) Quad4 mesh of [0,Lx]x[0,Ly]; interior nodes
    optionally jittered (non-affine cells)."""
    nodes, cells = orc.structured_mesh(2, nx, ny, 0, 1)
    nodes[:, 0] *= Lx
    nodes[:, 1] *= Ly
    if jitter:
        rng = np.random.default_rng(seed)
        ix = np.rint(nodes[:, 0] / Lx * nx).astype(int)
        iy = np.rint(nodes[:, 1] / Ly * ny).astype(int)
        inner = (ix > 0) & (ix < nx) & (iy > 0) & (iy < ny)
        nodes[inner, 0] += jitter * Lx / nx * rng.uniform(-1, 1, inner.sum())
        nodes[inner, 1] += jitter * Ly / ny * rng.uniform(-1, 1, inner.sum())
    return nodes, cells


def brute_faces(cells):
    """buildInternalFaces restated directly: one record per corner pair,
    left = the cell that walks the pair upwards."""
    rec = {}
    for c, cell in enumerate(cells):
        for f in range(4):
            a, b = int(cell[f]), int(cell[(f + 1) % 4])
            key = (min(a, b), max(a, b))
            r = rec.setdefault(key, [-1, -1, -1, -1])
            if a < b:
                r[0], r[2] = max(r[0], c), max(r[2], f)
            else:
                r[1], r[3] = max(r[1], c), max(r[3], f)
    keys = sorted(rec)
    return [np.array([rec[k][j] for k in keys], np.int32) for j in range(4)]


@pytest.mark.parametrize("nx,ny", [(1, 1), (4, 3), (7, 5)])
def test_faces_brute_force(orc, nx, ny):
    nodes, cells = rect_mesh(orc, nx, ny)
    got = orc.dg_faces(cells)
    exp = brute_faces(cells)
    for g, e in zip(got, exp):
        assert np.array_equal(g, e)
    left, right = got[0], got[1]
    assert len(left) == nx * (ny + 1) + ny * (nx + 1)
    assert np.sum((left < 0) | (right < 0)) == 2 * (nx + ny)


@pytest.mark.parametrize("p", [1, 2, 3])
def test_dof_nodes_are_bilinear_images_of_gll_nodes(orc, p):
    nodes, cells = rect_mesh(orc, 3, 2, 1.5, 0.7, jitter=0.2, seed=1)
    X = orc.dg_nodes(p, nodes, cells)
    xl = orc.local_nodes(2, p)
    for e, cell in enumerate(cells):
        c = nodes[cell]
        xi, eta = xl[:, 0], xl[:, 1]
        s = np.stack([(1 - xi) * (1 - eta), (1 + xi) * (1 - eta), (1 + xi) * (1 + eta), (1 - xi) * (1 + eta)], 1) / 4
        np.testing.assert_allclose(X[e * (p + 1) ** 2:(e + 1) * (p + 1) ** 2], s @ c, rtol=0, atol=1e-15)


@pytest.mark.parametrize("p", [1, 2, 3, 4])
def test_constant_source_gives_the_source(orc, p):
    # c = 0: no face flux; M^-1 (s * int N) = s since the basis sums to 1
    nodes, cells = rect_mesh(orc, 4, 3, jitter=0.25, seed=2)
    ph = physics(DG, 2, f=2.5, U=(0, 0, 0))
    U = np.random.default_rng(3).standard_normal(cells.shape[0] * (p + 1) ** 2)
    R, Rf = orc.dg_residual(p, nodes, cells, ph, U, face=True)
    assert np.abs(Rf).max() == 0
    np.testing.assert_allclose(R, 2.5, rtol=0, atol=1e-13)


@pytest.mark.parametrize("p", [1, 2, 3])
def test_linear_field_element_part_is_minus_c_dot_grad(orc, p):
    # U = a.x + b is continuous, so the upwind flux is (c.n) U; integration by
    # parts makes R_face + R_vol = -(c.a) int N, and M^-1 of that is -(c.a),
    # on every element without an inflow face (affine cells: exact quadrature)
    nx, ny = 5, 4
    nodes, cells = rect_mesh(orc, nx, ny, 1.3, 0.9)
    c = (0.7, -0.3)
    a = np.array([1.3, -0.4])
    ph = physics(DG, 2, U=(c[0], c[1], 0))
    X = orc.dg_nodes(p, nodes, cells)
    U = X[:, 0] * a[0] + X[:, 1] * a[1] + 0.25
    R, Rf = orc.dg_residual(p, nodes, cells, ph, U, face=True)
    nen = (p + 1) ** 2
    d = (R - Rf).reshape(-1, nen)
    ex = -(c[0] * a[0] + c[1] * a[1])
    # inflow faces: x = 0 (c.n = -0.7) and y = top (c.n = -0.3)
    for e in range(cells.shape[0]):
        i, j = e % nx, e // nx
        if i == 0 or j == ny - 1:
            continue
        np.testing.assert_allclose(d[e], ex, rtol=0, atol=1e-12)


@pytest.mark.parametrize("p", [1, 2, 3])
def test_linear_field_on_parallelograms(orc, p):
    # as above on a sheared mesh (x -> x + 0.35 y): affine cells whose J is not
    # diagonal, so the volume term goes through J^-T and the face terms through
    # slanted normals; the inflow faces are still the left column and top row
    nx, ny = 5, 4
    nodes, cells = rect_mesh(orc, nx, ny, 1.3, 0.9)
    nodes[:, 0] += 0.35 * nodes[:, 1]
    c = (0.7, -0.3)
    a = np.array([1.3, -0.4])
    ph = physics(DG, 2, U=(c[0], c[1], 0))
    X = orc.dg_nodes(p, nodes, cells)
    U = X[:, 0] * a[0] + X[:, 1] * a[1] + 0.25
    R, Rf = orc.dg_residual(p, nodes, cells, ph, U, face=True)
    d = (R - Rf).reshape(-1, (p + 1) ** 2)
    ex = -(c[0] * a[0] + c[1] * a[1])
    checked = 0
    for e in range(cells.shape[0]):
        i, j = e % nx, e // nx
        if i == 0 or j == ny - 1:
            continue
        np.testing.assert_allclose(d[e], ex, rtol=0, atol=5e-12)   # M^-1 of sheared cells: a few 1e-12
        checked += 1
    assert checked == (nx - 1) * (ny - 1)


@pytest.mark.parametrize("c", [(1.0, 0.0), (-1.0, 0.5), (0.3, 0.8)])
def test_constant_field_face_terms_are_the_outflow(orc, c):
    # U = 1: interior fluxes cancel between the two cells; the sum of the face
    # terms is minus the outflow through the boundary, sum over out faces of
    # (c.n) * length
    Lx, Ly = 1.5, 0.7
    nodes, cells = rect_mesh(orc, 6, 4, Lx, Ly, jitter=0.2, seed=5)
    p = 2
    U = np.ones(cells.shape[0] * (p + 1) ** 2)
    _, Rf = orc.dg_residual(p, nodes, cells, physics(DG, 2, U=(c[0], c[1], 0)), U, face=True)
    out = max(-c[0], 0) * Ly + max(c[0], 0) * Ly + max(-c[1], 0) * Lx + max(c[1], 0) * Lx
    assert abs(Rf.sum() + out) < 1e-12


def test_rk4_on_a_constant_rate_is_exact(orc):
    nodes, cells = rect_mesh(orc, 3, 3, jitter=0.2, seed=7)
    p = 2
    U0 = np.random.default_rng(8).standard_normal(cells.shape[0] * (p + 1) ** 2)
    U = orc.dg_rk4(p, nodes, cells, physics(DG, 2, f=-1.25, U=(0, 0, 0)), U0, 0.01, 7)
    np.testing.assert_allclose(U, U0 - 1.25 * 0.07, rtol=0, atol=1e-13)
