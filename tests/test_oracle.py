"""CPU checks that pin the oracle: golden vectors from the reference's own
sources, closed forms, invariants and brute force on tiny inputs."""
import math

import numpy as np
import pytest

from oracle.oracle import physics

LAM, MU = 0.3 / (1.3 * 0.4), 1 / 2.6


# ---------------------------------------------------------------- quadrature
@pytest.mark.parametrize("npts", [2, 3, 4, 5, 6])
def test_quadrature_1d_matches_reference(orc, golden, npts):
    q = golden("quadrature.npz")
    x, w = orc.gauss_legendre(npts)
    assert np.array_equal(x, q[f"gl{npts}_x"]) and np.array_equal(w, q[f"gl{npts}_w"])
    x, w = orc.gauss_lobatto(npts)
    assert np.array_equal(x, q[f"gll{npts}_x"]) and np.array_equal(w, q[f"gll{npts}_w"])


@pytest.mark.parametrize("dim", [1, 2, 3])
@pytest.mark.parametrize("order", [2, 4, 6, 8])
def test_tensor_rule_matches_reference(orc, golden, dim, order):
    q = golden("quadrature.npz")
    xi, w = orc.tensor_rule(0, dim, order)
    assert np.array_equal(xi, q[f"tp{dim}_{order}_xi"]) and np.array_equal(w, q[f"tp{dim}_{order}_w"])


@pytest.mark.parametrize("npts", [2, 3, 4, 5])
def test_quadrature_invariants(orc, npts):
    # weights sum to 2 (test/test_GaussLegQuadrature.cpp:10-17) and exactness
    x, w = orc.gauss_legendre(npts)
    assert abs(w.sum() - 2) < 1e-14
    xr, wr = np.polynomial.legendre.leggauss(npts)
    assert np.allclose(x, xr, atol=1e-14) and np.allclose(w, wr, atol=1e-14)
    for k in range(2 * npts):
        exact = (1 - (-1) ** (k + 1)) / (k + 1)
        assert abs((w * x ** k).sum() - exact) < 1e-13
    xl, wl = orc.gauss_lobatto(npts)
    assert xl[0] == -1 and xl[-1] == 1 and abs(wl.sum() - 2) < 1e-14
    for k in range(2 * npts - 2):
        exact = (1 - (-1) ** (k + 1)) / (k + 1)
        assert abs((wl * xl ** k).sum() - exact) < 1e-13


# ------------------------------------------------------------ shape functions
def lagrange_1d(nodes, t):
    """textbook Lagrange basis values/derivatives at points t (independent of the oracle)"""
    n = len(nodes)
    V = np.ones((len(t), n))
    D = np.zeros((len(t), n))
    for i in range(n):
        others = [m for m in range(n) if m != i]
        den = np.prod([nodes[i] - nodes[m] for m in others])
        V[:, i] = np.prod([t - nodes[m] for m in others], axis=0) / den
        for k in others:
            rest = [m for m in others if m != k]
            D[:, i] += np.prod([t - nodes[m] for m in rest], axis=0) / den if rest else 1 / den
    return V, D


@pytest.mark.parametrize("dim,p", [(2, 1), (2, 2), (2, 3), (2, 4), (3, 1), (3, 2)])
def test_shape_tables_closed_form(orc, dim, p):
    N, dN, w, xi = orc.shape_tables(dim, p)
    gll, _ = orc.gauss_lobatto(p + 1)
    n1 = p + 1
    nen = n1 ** dim
    idx = np.array([[(a // n1 ** d) % n1 for d in range(dim)] for a in range(nen)])
    Nc = np.ones_like(N)
    dNc = np.ones_like(dN)
    for d in range(dim):
        V, D = lagrange_1d(gll, xi[:, d])
        Nc *= V[:, idx[:, d]]
        for g in range(dim):
            dNc[:, :, g] *= (D if g == d else V)[:, idx[:, d]]
    assert np.abs(N - Nc).max() < 1e-13
    assert np.abs(dN - dNc).max() < 1e-12
    assert np.abs(N.sum(axis=1) - 1).max() < 1e-13          # partition of unity
    assert np.abs(dN.sum(axis=1)).max() < 1e-12            # gradients sum to zero
    assert abs(w.sum() - 2 ** dim) < 1e-13


# ------------------------------------------------------------ DoF numbering
@pytest.mark.parametrize("dim,p,n", [(2, 2, 3), (2, 4, 2), (3, 2, 2), (2, 3, 4)])
def test_dof_numbering_high_order_is_row_major_gll_grid(orc, dim, p, n):
    # p >= 2: the (z, y, x) coordinate sort numbers the (p n + 1)^d grid x fastest
    el, dn = orc.mesh(dim, p, n)
    m = p * n + 1
    assert dn.shape[0] == m ** dim
    gll, _ = orc.gauss_lobatto(p + 1)
    pts1 = np.unique(np.round(np.concatenate([(c + (gll + 1) / 2) / n for c in range(n)]), 12))
    assert len(pts1) == m
    grid = np.stack(np.meshgrid(*([pts1] * dim), indexing="ij")[::-1], axis=-1).reshape(-1, dim)
    assert np.abs(dn[:, :dim] - grid).max() < 1e-12   # grid was rounded to 12 digits
    # brute force: every element's local node (x fastest) sits at its GLL position
    xl = orc.local_nodes(dim, p)
    for e in range(el.shape[0]):
        cell = np.unravel_index(e, [n] * dim, order="F")
        for a in range(el.shape[1]):
            pos = [(cell[d] + (xl[a, d] + 1) / 2) / n for d in range(dim)]
            assert np.abs(dn[el[e, a], :dim] - pos).max() < 1e-14


def test_dof_numbering_p1_passthrough(orc):
    nodes, cells = orc.structured_mesh(2, 4, 3, 0, 0)
    el, dn = orc.dofmanager(2, 1, nodes, cells)
    assert np.array_equal(el, cells) and np.array_equal(dn, nodes)


# ------------------------------------------------------------ element matrices
def gll_1d_mats(orc, p, h):
    """1-D matrices on [0,h] of the GLL Lagrange basis, by high-order numpy Gauss."""
    gll, _ = orc.gauss_lobatto(p + 1)
    t, wt = np.polynomial.legendre.leggauss(20)
    V, D = lagrange_1d(gll, t)
    M = (V.T * wt) @ V * (h / 2)
    A = (D.T * wt) @ D * (2 / h)
    C = (D.T * wt) @ V          # int l_a' l_b  (scale (2/h)(h/2) = 1)
    return M, A, C


def kron_all(mats):
    out = np.ones((1, 1))
    for m in mats[::-1]:        # x fastest: last factor is x
        out = np.kron(out, m)
    return out


def grad_integrals(orc, dim, p, h):
    """I[j][l][A,B] = int d_j N_A d_l N_B over an axis-aligned cube of side h"""
    M, A, C = gll_1d_mats(orc, p, h)
    I = [[None] * dim for _ in range(dim)]
    for j in range(dim):
        for l in range(dim):
            fac = []
            for d in range(dim):
                if d == j and d == l:
                    fac.append(A)
                elif d == j:
                    fac.append(C)
                elif d == l:
                    fac.append(C.T)
                else:
                    fac.append(M)
            I[j][l] = kron_all(fac)
    return I


@pytest.mark.parametrize("p", [1, 2, 4])
def test_poisson_element_kronecker(orc, p):
    el, dn = orc.mesh(2, p, 3)
    coef = np.linspace(0.5, 2.0, el.shape[0])
    ph = physics(0, 2, coef=coef, f=3.0)
    I = grad_integrals(orc, 2, p, 1 / 3)
    M, _, _ = gll_1d_mats(orc, p, 1 / 3)
    for e in [0, 4, 8]:
        K, R = orc.element_matrix(2, p, 1, dn, el, e, ph)
        Kc = -coef[e] * (I[0][0] + I[1][1])
        assert np.abs(K - Kc).max() < 1e-12 * np.abs(Kc).max()
        Rc = 3.0 * kron_all([M, M]).sum(axis=1)
        assert np.abs(R - Rc).max() < 1e-13


def test_q1_poisson_stencil(orc):
    # interior row of the Q1 Laplacian under the -K sign: -8/3 centre, +1/3 neighbours
    el, dn = orc.mesh(2, 1, 4)
    rows, cols, vals, R = orc.assemble(2, 1, 1, dn, el, physics(0, 2, f=1.0))
    A = orc.csr(rows, cols, vals)
    ptr, col, val = A.arrays()
    r = 12  # node (2,2)
    row = dict(zip(col[ptr[r]:ptr[r + 1]], val[ptr[r]:ptr[r + 1]]))
    assert abs(row[r] + 8 / 3) < 1e-14
    assert all(abs(v - 1 / 3) < 1e-14 for c, v in row.items() if c != r)
    assert abs(R[r] - (1 / 4) ** 2) < 1e-15      # load f h^2 per interior node


@pytest.mark.parametrize("p", [1, 2])
def test_elasticity_element_closed_form(orc, p):
    el, dn = orc.mesh(3, p, 2)
    coef = np.array([1.0, 2.0, 0.5, 1.5, 1.0, 1.0, 3.0, 1.0])
    ph = physics(1, 3, coef=coef, lam=LAM, mu=MU, body=(0.1, -0.2, -1.0))
    I = grad_integrals(orc, 3, p, 0.5)
    M, _, _ = gll_1d_mats(orc, p, 0.5)
    nen = el.shape[1]
    for e in [0, 6]:
        K, R = orc.element_matrix(3, p, 3, dn, el, e, ph)
        lam, mu = coef[e] * LAM, coef[e] * MU
        tr = sum(I[j][j] for j in range(3))
        Kc = np.zeros_like(K)
        for i in range(3):
            for k in range(3):
                blk = lam * I[i][k] + mu * I[k][i] + (mu * tr if i == k else 0)
                Kc[i::3, k::3] = -blk
        assert np.abs(K - Kc).max() < 1e-12 * np.abs(Kc).max()
        assert np.abs(K - K.T).max() < 1e-14
        mass = kron_all([M, M, M]).sum(axis=1)
        for i, bi in enumerate((0.1, -0.2, -1.0)):
            assert np.abs(R[i::3] - bi * mass).max() < 1e-14
        assert nen == (p + 1) ** 3


def test_supg_element_closed_form(orc):
    p, n = 2, 4
    el, dn = orc.mesh(2, p, n)
    U = (math.cos(math.pi / 6), math.sin(math.pi / 6), 0)
    h, kd, al = 1 / n, 0.01, 0.7
    ph = physics(2, 2, U=U, kdiff=kd, h=h, alpha=al)
    I = grad_integrals(orc, 2, p, h)
    M, A1, C = gll_1d_mats(orc, p, h)
    # int N_A d_c N_B : derivative on N_B only -> C^T factor in direction c
    NdN = [kron_all([C.T, M]), kron_all([M, C.T])]
    nU = math.hypot(U[0], U[1])
    Kt = kd * np.eye(2) + (al * h / (2 * nU)) * np.outer(U[:2], U[:2])
    K, R = orc.element_matrix(2, p, 1, dn, el, 5, ph)
    Kc = 2 * (U[0] * NdN[0] + U[1] * NdN[1])      # advection added once per A component
    for a in range(2):
        for b in range(2):
            Kc = Kc + Kt[a, b] * I[a][b]
    assert np.abs(K - Kc).max() < 1e-12 * np.abs(Kc).max()


# ------------------------------------------- cells where J is not diagonal
# A wrong transformation of the shape gradients (J^-1 for J^-T), a wrong
# Jacobian index order or a wrong det J cannot be seen on axis-aligned cells,
# where J is diagonal.  These pins use jittered and sheared meshes.
from meshes import SHEAR, corner_mesh, interior  # noqa: E402


def linear_field(dn, dim, dpn, seed=0):
    """u = a + B x at the dof nodes (dof = node * dpn + comp); linear fields
    lie in the isoparametric space of every cell."""
    rng = np.random.default_rng(seed)
    a, B = rng.standard_normal(dpn), rng.standard_normal((dpn, dim))
    u = a[None, :] + dn[:, :dim] @ B.T
    return u.ravel(), B


@pytest.mark.parametrize("dim,p", [(2, 1), (2, 2), (2, 4), (3, 1), (3, 2)])
@pytest.mark.parametrize("kind", ["jitter", "shear"])
@pytest.mark.parametrize("phys", ["poisson", "elasticity"])
def test_patch_test_on_distorted_cells(orc, dim, p, kind, phys):
    """K (a + B x) vanishes at every interior row: the element gradients of a
    linear field are exact on any cell, and the fluxes of neighbouring cells
    cancel.  With J^-1 in place of J^-T the computed gradient is J^-1 J^T B,
    which is not B where J is not symmetric."""
    n = 4 if dim == 2 else 3
    nodes, cells = corner_mesh(orc, dim, n, p, kind)
    el, dn = orc.dofmanager(dim, p, nodes, cells)
    dpn = 1 if phys == "poisson" else dim
    ph = physics(0, dim, f=0.0) if phys == "poisson" else physics(1, dim, lam=LAM, mu=MU)
    rows, cols, vals, _ = orc.assemble(dim, p, dpn, dn, el, ph)
    A = orc.csr(rows, cols, vals)
    u, _ = linear_field(dn, dim, dpn)
    r = A.spmv(u)
    inner = np.repeat(interior(dn, dim, kind), dpn)
    scale = np.abs(vals).max() * np.abs(u).max()
    assert inner.sum() > 0 and np.abs(r[~inner]).max() > 1e-3 * scale   # boundary rows carry the flux
    assert np.abs(r[inner]).max() < 1e-12 * scale


@pytest.mark.parametrize("dim,p", [(2, 1), (2, 3), (3, 1), (3, 2)])
def test_load_integrates_the_domain_on_jittered_cells(orc, dim, p):
    """R_A = f int N_A: on the jittered unit square / cube the loads add up to
    f |Omega| = f (det J, its sign and the weights), and a body force b gives
    b_i |Omega| in each component."""
    n = 4 if dim == 2 else 3
    nodes, cells = corner_mesh(orc, dim, n, p, "jitter", seed=3)
    el, dn = orc.dofmanager(dim, p, nodes, cells)
    _, _, _, R = orc.assemble(dim, p, 1, dn, el, physics(0, dim, f=2.5), want=("R",))
    assert abs(R.sum() - 2.5) < 1e-13
    body = (0.3, -0.7, 1.1)[:dim] + (0.0,) * (3 - dim)
    _, _, _, R = orc.assemble(dim, p, dim, dn, el, physics(1, dim, lam=LAM, mu=MU, body=body), want=("R",))
    for i in range(dim):
        assert abs(R[i::dim].sum() - body[i]) < 1e-13


def parallelogram_grad_integrals(orc, dim, p, J):
    """X[j][l][A, B] = int d_xj N_A d_xl N_B over a cell with constant J
    (J_ij = d x_i / d xi_j): det J sum_mn Jinv_mj Jinv_nl Iref[m][n], Iref the
    reference-cell (side 2) integrals of the 1-D GLL matrices."""
    Iref = grad_integrals(orc, dim, p, 2.0)
    Ji, det = np.linalg.inv(J), np.linalg.det(J)
    X = [[sum(Ji[m, j] * Ji[k, l] * Iref[m][k] for m in range(dim) for k in range(dim)) * det
          for l in range(dim)] for j in range(dim)]
    return X


@pytest.mark.parametrize("dim,p", [(2, 1), (2, 2), (2, 3), (3, 1), (3, 2)])
def test_parallelogram_element_closed_form(orc, dim, p):
    """On a sheared mesh every cell has J = (h/2) S: the Poisson and elasticity
    element matrices equal the closed forms built from the 1-D GLL matrices
    through J^-1 J^-T (a transposed J, J^-1 for J^-T or a lost det J fails)."""
    n = 3 if dim == 2 else 2
    nodes, cells = corner_mesh(orc, dim, n, p, "shear")
    el, dn = orc.dofmanager(dim, p, nodes, cells)
    J = SHEAR[dim] * (0.5 / n)
    X = parallelogram_grad_integrals(orc, dim, p, J)
    coef = np.linspace(0.5, 2.0, el.shape[0])
    e = el.shape[0] // 2
    K, _ = orc.element_matrix(dim, p, 1, dn, el, e, physics(0, dim, coef=coef, f=0.0))
    Kc = -coef[e] * sum(X[j][j] for j in range(dim))
    assert np.abs(K - Kc).max() < 1e-12 * np.abs(Kc).max()
    K, _ = orc.element_matrix(dim, p, dim, dn, el, e, physics(1, dim, coef=coef, lam=LAM, mu=MU))
    lam, mu = coef[e] * LAM, coef[e] * MU
    tr = sum(X[j][j] for j in range(dim))
    Kc = np.zeros_like(K)
    for i in range(dim):
        for k in range(dim):
            Kc[i::dim, k::dim] = -(lam * X[i][k] + mu * X[k][i] + (mu * tr if i == k else 0))
    assert np.abs(K - Kc).max() < 1e-12 * np.abs(Kc).max()


@pytest.mark.parametrize("p", [1, 2])
def test_supg_linear_field_on_jittered_cells(orc, p):
    """SUPG K (a + b.x) at an interior node is NSD (U.b) int N_A (the
    advection term, added once per component; the diffusion fluxes cancel),
    and int N_A is the Poisson load with f = 1 on the same mesh."""
    n = 5
    nodes, cells = corner_mesh(orc, 2, n, p, "jitter", seed=7)
    el, dn = orc.dofmanager(2, p, nodes, cells)
    U = (math.cos(0.4), math.sin(0.4), 0.0)
    rows, cols, vals, _ = orc.assemble(2, p, 1, dn, el, physics(2, 2, U=U, kdiff=0.01, h=1 / n, alpha=0.8))
    A = orc.csr(rows, cols, vals)
    u, B = linear_field(dn, 2, 1, seed=5)
    _, _, _, mass = orc.assemble(2, p, 1, dn, el, physics(0, 2, f=1.0), want=("R",))
    inner = interior(dn, 2, "jitter")
    expect = 2 * (U[0] * B[0, 0] + U[1] * B[0, 1]) * mass
    r = A.spmv(u)
    assert np.abs(r[inner] - expect[inner]).max() < 1e-12 * np.abs(expect).max()


# ------------------------------------------------------------ sparse build
@pytest.mark.parametrize("nthreads", [1, 3])
def test_threaded_assembly_equals_the_element_order_loop(orc, nthreads):
    """The CPU baseline's assembly (CGAssembly's OpenMP structure) emits the
    serial loop's triplets, in another order, and the same residual."""
    el, dn = orc.mesh(3, 1, 3)
    ph = physics(1, 3, lam=LAM, mu=MU, body=(0, 0, -1))
    r1, c1, v1, R1 = orc.assemble(3, 1, 3, dn, el, ph)
    r2, c2, v2, R2 = orc.assemble_threads(nthreads, 3, 1, 3, dn, el, ph)
    k1, k2 = np.lexsort((v1, c1, r1)), np.lexsort((v2, c2, r2))
    assert np.array_equal(r1[k1], r2[k2]) and np.array_equal(c1[k1], c2[k2]) and np.array_equal(v1[k1], v2[k2])
    assert np.allclose(R1, R2, rtol=0, atol=1e-15)


@pytest.mark.parametrize("tag", ["q1", "q2", "q4", "hex"])
def test_csr_from_triplets_matches_reference(orc, golden, tag):
    s = golden("sparse_build.npz")
    A = orc.csr(s[f"{tag}_rows"], s[f"{tag}_cols"], s[f"{tag}_vals"])
    ptr, col, val = A.arrays()
    assert np.array_equal(ptr, s[f"{tag}_ptr"]) and np.array_equal(col, s[f"{tag}_col"])
    assert np.array_equal(val, s[f"{tag}_val"])        # same summation order: bit-exact


def test_bcsr_from_triplets_matches_reference(orc, golden):
    s = golden("sparse_build.npz")
    A = orc.csr(s["hex_rows"], s["hex_cols"], s["hex_vals"], block=3)
    ptr, col, val = A.arrays()
    assert np.array_equal(ptr, s["hex_bptr"]) and np.array_equal(col, s["hex_bcol"])
    assert np.abs(val - s["hex_bval"]).max() <= 1e-15 * np.abs(val).max()


@pytest.mark.parametrize("blk,tag", [(1, "dup_csr"), (3, "dup_bcsr")])
def test_duplicate_accumulation(orc, golden, blk, tag):
    s = golden("sparse_build.npz")
    N = 60
    A = orc.csr(np.array([N - 1, 0] * N), np.array([N - 1, 0] * N), np.ones(2 * N), block=blk)
    ptr, col, val = A.arrays()
    assert np.array_equal(ptr, s[f"{tag}_ptr"]) and np.array_equal(val, s[f"{tag}_val"])


def test_random_integer_sum(orc, golden):
    s = golden("sparse_build.npz")
    A = orc.csr(s["rnd_rows"], s["rnd_cols"], s["rnd_vals"])
    ptr, col, val = A.arrays()
    assert np.array_equal(val, s["rnd_val"]) and val.sum() == s["rnd_vals"].sum()


def test_pattern_brute_force(orc):
    el, dn = orc.mesh(2, 2, 3)
    rows, cols, vals, _ = orc.assemble(2, 2, 1, dn, el, physics(0, 2, f=1.0))
    A = orc.csr(rows, cols, vals)
    ptr, col, _ = A.arrays()
    pairs = sorted({(int(a), int(b)) for e in el for a in e for b in e})
    got = [(r, int(c)) for r in range(len(ptr) - 1) for c in col[ptr[r]:ptr[r + 1]]]
    assert got == pairs
    # closed-form row count: (cells*(p+1)^2 - (cells-1))^d pairs
    assert len(col) == (3 * 9 - 2) ** 2


# ------------------------------------------------------------ SpMV / dot
def test_spmv_dense_brute_force(orc):
    rng = np.random.default_rng(0)
    D = rng.standard_normal((30, 30)) * (rng.random((30, 30)) < 0.2)
    r, c = np.nonzero(D)
    A = orc.csr(r, c, D[r, c])
    x = rng.standard_normal(30)
    assert np.abs(A.spmv(x)[:A.rows] - (D @ x)[:A.rows]).max() < 1e-13
    B = orc.csr(r, c, D[r, c], block=3)
    assert np.abs(B.spmv(x)[:B.rows] - (D @ x)[:B.rows]).max() < 1e-13


def test_dot_matches_reference_order(orc):
    from oracle.oracle import Reference
    if not Reference.available():
        pytest.skip("reference harness not built here")
    rng = np.random.default_rng(1)
    a, b = rng.standard_normal(1001), rng.standard_normal(1001)
    assert orc.dot(a, b) == Reference().dot(a, b)


# ------------------------------------------------------------ Dirichlet
def test_dirichlet_closed_form(orc):
    el, dn = orc.mesh(2, 1, 4)
    rows, cols, vals, R = orc.assemble(2, 1, 1, dn, el, physics(0, 2, f=1.0))
    A = orc.csr(rows, cols, vals)
    ptr0, col0, val0 = A.arrays()
    mask = (np.abs(dn[:, 0]) < 1e-6).astype(np.uint8)
    g = np.where(mask, 2.0, 0.0)
    Ad = np.zeros((25, 25))
    for r in range(25):
        Ad[r, col0[ptr0[r]:ptr0[r + 1]]] = val0[ptr0[r]:ptr0[r + 1]]
    b = A.dirichlet(R, mask, g)
    ptr, col, val = A.arrays()
    assert np.array_equal(ptr, ptr0) and np.array_equal(col, col0)   # pattern kept
    exp_b = R - Ad @ g
    exp_b[mask == 1] = 2.0
    assert np.abs(b - exp_b).max() < 1e-14
    for r in range(25):
        for k in range(ptr[r], ptr[r + 1]):
            c = col[k]
            if mask[r] or mask[c]:
                assert val[k] == (1.0 if r == c else 0.0)
            else:
                assert val[k] == val0[k]


# ------------------------------------------------------------ solvers
def test_solver_kats(orc, golden):
    k = golden("kat_solvers.npz")
    I = orc.csr(np.arange(10), np.arange(10), np.ones(10))
    x, _ = I.cg(k["id_rhs"])
    assert np.array_equal(x, k["id_rhs"]) and np.array_equal(x, k["id_cg_x"])
    T = orc.csr(k["tri_rows"], k["tri_cols"], k["tri_vals"])
    x, _ = T.cg(k["tri_b"])
    assert np.array_equal(x, k["tri_cg_x"])
    assert ((x - np.arange(10)) ** 2).sum() < 1e-10
    x, _ = T.bicgstab(k["tri_b"])
    assert np.array_equal(x, k["tri_bicgstab_x"])


def test_solves_match_reference(orc, golden):
    v = golden("solves.npz")
    A = orc.wrap(1, len(v["p_ptr"]) - 1, len(v["p_ptr"]) - 1, v["p_ptr"], v["p_col"], v["p_val"])
    x, _ = A.cg(v["p_b"], tol=1e-20)
    assert np.array_equal(x, v["p_cg_x"])
    x, _ = A.pcg_jacobi(v["p_b"], tol=1e-20)
    assert np.array_equal(x, v["p_pcg_x"])


def test_cg_true_residual_and_monotone_energy(orc):
    el, dn = orc.mesh(2, 1, 16)
    rows, cols, vals, R = orc.assemble(2, 1, 1, dn, el, physics(0, 2, f=1.0))
    A = orc.csr(rows, cols, vals)
    mask = ((dn[:, 0] < 1e-6) | (dn[:, 0] > 1 - 1e-6) | (dn[:, 1] < 1e-6) | (dn[:, 1] > 1 - 1e-6)).astype(np.uint8)
    b = A.dirichlet(R, mask)
    x, it = A.cg(b, tol=1e-20)
    assert np.linalg.norm(b - A.spmv(x)) / np.linalg.norm(b) < 1e-9
    assert 0 < it < len(b)


def test_bicgstab_jacobi_nonsymmetric(orc):
    el, dn = orc.mesh(2, 2, 8)
    ph = physics(2, 2, U=(0.866, 0.5, 0), kdiff=1e-3, h=1 / 8, alpha=0.9)
    rows, cols, vals, _ = orc.assemble(2, 2, 1, dn, el, ph)
    A = orc.csr(rows, cols, vals)
    mask = ((dn[:, 0] < 1e-6) | (dn[:, 0] > 1 - 1e-6) | (dn[:, 1] < 1e-6) | (dn[:, 1] > 1 - 1e-6)).astype(np.uint8)
    b = A.dirichlet(orc.normal(7, len(dn)), mask)
    x, it = A.bicgstab_jacobi(b, 1e-10, 10000)
    assert np.linalg.norm(b - A.spmv(x)) / np.linalg.norm(b) < 1e-10


def test_rng_box_muller_moments(orc):
    z = orc.normal(42, 200001)
    assert abs(z.mean()) < 0.01 and abs(z.std() - 1) < 0.01


# ------------------------------------------------------------ gradient recovery
def _graded(dn, dim):
    """x -> x + 0.15 x (1 - x) sin(3x) per axis: rectangles stay rectangles
    (p = 1), with unequal neighbouring cells."""
    X = dn.copy()
    for d in range(dim):
        X[:, d] = dn[:, d] + 0.15 * dn[:, d] * (1 - dn[:, d]) * np.sin(3 * dn[:, d])
    return X


@pytest.mark.parametrize("dim,p,dpn", [(2, 1, 1), (2, 2, 1), (2, 3, 1), (3, 1, 1), (3, 1, 3), (3, 2, 3), (2, 1, 2)])
def test_local_smoothing_recovers_linear_fields_exactly(orc, dim, p, dpn):
    """u_c = a_c + B_c . x lies in every element space, so every gradient
    sample is B_c and the weighted average is B_c at every dof."""
    el, dn = orc.mesh(dim, p, 3)
    if p == 1:
        dn = _graded(dn, dim)
    rng = np.random.default_rng(dim * 10 + p + dpn)
    a, B = rng.standard_normal(dpn), rng.standard_normal((dpn, dim))
    u = (a[None, :] + dn[:, :dim] @ B.T).ravel()
    G = orc.grad_local_smoothing(dim, p, dpn, dn, el, u)
    expect = np.tile(B, (dn.shape[0], 1))
    assert np.abs(G - expect).max() < 1e-12


@pytest.mark.parametrize("dim,p,dpn", [(2, 1, 1), (2, 2, 2), (2, 3, 1), (3, 1, 3), (3, 2, 1)])
@pytest.mark.parametrize("kind", ["jitter", "shear"])
def test_local_smoothing_linear_fields_on_distorted_cells(orc, dim, p, dpn, kind):
    """As above on cells whose J is not diagonal: the element gradients of a
    linear field are exact only through J^-T."""
    nodes, cells = corner_mesh(orc, dim, 3, p, kind, seed=dim + p)
    el, dn = orc.dofmanager(dim, p, nodes, cells)
    rng = np.random.default_rng(dim * 10 + p + dpn)
    a, B = rng.standard_normal(dpn), rng.standard_normal((dpn, dim))
    u = (a[None, :] + dn[:, :dim] @ B.T).ravel()
    G = orc.grad_local_smoothing(dim, p, dpn, dn, el, u)
    assert np.abs(G - np.tile(B, (dn.shape[0], 1))).max() < 1e-11


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("kind", ["jitter", "shear"])
def test_zz_linear_fields_on_distorted_cells(orc, dim, kind):
    nodes, cells = corner_mesh(orc, dim, 4 if dim == 2 else 3, 1, kind, seed=2 * dim)
    el, dn = orc.dofmanager(dim, 1, nodes, cells)
    rng = np.random.default_rng(17 + dim)
    a, b = rng.standard_normal(), rng.standard_normal(dim)
    G = orc.grad_zz(dim, dn, el, a + dn[:, :dim] @ b)
    assert np.abs(G - b[None, :]).max() < 1e-10


@pytest.mark.parametrize("dim", [2, 3])
def test_local_smoothing_q1_quadratic_closed_form(orc, dim):
    """u = x^2 on a uniform Q1 grid (h = 1/n): the interpolant's x-derivative
    is 2 x_mid in each cell column.  The two cells around an interior node
    carry equal weights (sum N jxw), so the average is 2 x_n; at x = 0 only the
    right cells count (h), at x = 1 only the left ones (2 - h)."""
    n = 4
    h = 1.0 / n
    el, dn = orc.mesh(dim, 1, n)
    G = orc.grad_local_smoothing(dim, 1, 1, dn, el, dn[:, 0] ** 2)
    x = dn[:, 0]
    expect = np.where(x < 1e-12, h, np.where(x > 1 - 1e-12, 2 - h, 2 * x))
    assert np.abs(G[:, 0] - expect).max() < 1e-13
    assert np.abs(G[:, 1:]).max() < 1e-13


def test_local_smoothing_q2_exact_for_quadratics(orc):
    """x^2 + x y - y^2 lies in Q2: every sample is the exact gradient."""
    el, dn = orc.mesh(2, 2, 3)
    x, y = dn[:, 0], dn[:, 1]
    G = orc.grad_local_smoothing(2, 2, 1, dn, el, x * x + x * y - y * y)
    assert np.abs(G[:, 0] - (2 * x + y)).max() < 1e-12
    assert np.abs(G[:, 1] - (x - 2 * y)).max() < 1e-12


@pytest.mark.parametrize("dim", [2, 3])
def test_zz_recovers_linear_fields_exactly(orc, dim):
    el, dn = orc.mesh(dim, 1, 4)
    dn = _graded(dn, dim)
    rng = np.random.default_rng(7 + dim)
    a, b = rng.standard_normal(), rng.standard_normal(dim)
    G = orc.grad_zz(dim, dn, el, a + dn[:, :dim] @ b)
    assert np.abs(G - b[None, :]).max() < 1e-11


def test_zz_graded_patch_least_squares_closed_form(orc):
    """u = x^2 on a 2-D Q1 grid graded in x only: at a node with cells of
    widths h1 (left) and h2 (right), the samples are the per-column constant
    slopes 2 x_n - h1 and 2 x_n + h2 at the Gauss abscissae x_n - h1 (1 -+ 1/sqrt 3)/2
    and x_n + h2 (1 +- 1/sqrt 3)/2 (two y points per cell, two cells in y in the
    interior).  The fit of that step with a line in x (y carries no signal)
    is a weighted least-squares problem solved here by numpy's SVD-based
    lstsq from those closed-form samples; ZZ evaluates it at x_n."""
    n = 5
    el, dn = orc.mesh(2, 1, n)
    dn = _graded(dn, 1)
    xs = np.unique(np.round(dn[:, 0], 14))
    G = orc.grad_zz(2, dn, el, dn[:, 0] ** 2)
    g = 1 / math.sqrt(3)
    for k, xn in enumerate(xs):
        samples = []
        if k > 0:
            h1 = xn - xs[k - 1]
            samples += [(xn - h1 * (1 - s) / 2, 2 * xn - h1) for s in (g, -g)]
        if k < len(xs) - 1:
            h2 = xs[k + 1] - xn
            samples += [(xn + h2 * (1 + s) / 2, 2 * xn + h2) for s in (-g, g)]
        P = np.array([[1.0, sx - xn] for sx, _ in samples])
        b = np.array([v for _, v in samples])
        coef = np.linalg.lstsq(P, b, rcond=None)[0]
        # every node row with this x: interior rows in y double every sample,
        # which leaves the fit unchanged
        rows = np.abs(dn[:, 0] - xn) < 1e-12
        assert np.abs(G[rows, 0] - coef[0]).max() < 1e-12, (k, xn)
        assert np.abs(G[rows, 1]).max() < 1e-12


def test_zz_node_of_no_element(orc):
    """An unreferenced node gets the empty fit's zero; the others are unchanged."""
    el, dn = orc.mesh(2, 1, 4)
    dn = _graded(dn, 2)
    u = np.random.default_rng(1).standard_normal(dn.shape[0] + 1)
    G0 = orc.grad_zz(2, dn, el, u[:-1])
    G1 = orc.grad_zz(2, np.vstack([dn, [[3.0, 3.0, 0.0]]]), el, u)
    assert np.array_equal(G1[-1], [0.0, 0.0]) and np.array_equal(G1[:-1], G0)
