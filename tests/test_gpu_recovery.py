"""Gradient recovery on the device against the oracle: LocalSmoothingGradient
(every element order and dof count the library assembles) and ZZ patch
recovery (p = 1, scalar), on meshes with unequal cells and on a solved field.
Both sum in element order; the device's sums group the quadrature points of
one element first and the ZZ fit solves centred normal equations instead of a
pivoted QR, so agreement is to round-off, not bitwise."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def H():
    from yafel_amd import hip
    return hip


def _graded(nodes, dim):
    X = nodes.copy()
    for d in range(dim):
        X[:, d] = nodes[:, d] + 0.15 * nodes[:, d] * (1 - nodes[:, d]) * np.sin(3 * nodes[:, d])
    return X


def _mesh(H, ctx, dim, p, dpn, n, graded):
    nodes, cells = H.structured_mesh(dim, n, n, n if dim == 3 else 0, p)
    if graded:
        nodes = _graded(nodes, dim)
    return H.Mesh(ctx, dim, p, dpn, nodes, cells)


@pytest.mark.parametrize("dim,p,dpn,n", [(2, 1, 1, 24), (2, 2, 1, 12), (2, 3, 1, 8), (2, 4, 1, 6),
                                         (3, 1, 1, 8), (3, 1, 3, 8), (3, 2, 3, 5), (2, 1, 2, 10)])
def test_local_smoothing_matches_oracle(H, ctx, orc, dim, p, dpn, n):
    m = _mesh(H, ctx, dim, p, dpn, n, graded=(p == 1))
    el, dn = m.download()
    u = np.random.default_rng(100 * dim + 10 * p + dpn).standard_normal(m.ndofs)
    G = m.gradient_local_smoothing(ctx.upload(u)).download().reshape(m.ndofs, dim)
    Go = orc.grad_local_smoothing(dim, p, dpn, dn, el, u)
    scale = np.abs(Go).max()
    print(f"smoothing dim {dim} p {p} dpn {dpn}: max |diff| / max|G| = {np.abs(G - Go).max() / scale:.2e}", flush=True)
    assert np.abs(G - Go).max() <= 1e-12 * scale


@pytest.mark.parametrize("dim,n", [(2, 20), (3, 7)])
def test_zz_matches_oracle(H, ctx, orc, dim, n):
    m = _mesh(H, ctx, dim, 1, 1, n, graded=True)
    el, dn = m.download()
    u = np.random.default_rng(5 + dim).standard_normal(m.ndofs)
    G = m.gradient_zz(ctx.upload(u)).download().reshape(m.n_nodes, dim)
    Go = orc.grad_zz(dim, dn, el, u)
    scale = np.abs(Go).max()
    print(f"zz dim {dim}: max |diff| / max|G| = {np.abs(G - Go).max() / scale:.2e}", flush=True)
    assert np.abs(G - Go).max() <= 1e-10 * scale


@pytest.mark.parametrize("kind", ["jitter", "shear"])
@pytest.mark.parametrize("dim,p,dpn,n", [(2, 1, 1, 12), (2, 3, 1, 5), (3, 1, 3, 5), (3, 2, 1, 3)])
def test_local_smoothing_on_distorted_cells(H, ctx, orc, kind, dim, p, dpn, n):
    """On cells whose J is not diagonal (the oracle is pinned there by the
    exact recovery of linear fields)."""
    from meshes import corner_mesh
    nodes, cells = corner_mesh(orc, dim, n, p, kind, seed=21)
    m = H.Mesh(ctx, dim, p, dpn, nodes, cells)
    el, dn = m.download()
    u = np.random.default_rng(300 + dim + p).standard_normal(m.ndofs)
    G = m.gradient_local_smoothing(ctx.upload(u)).download().reshape(m.ndofs, dim)
    Go = orc.grad_local_smoothing(dim, p, dpn, dn, el, u)
    assert np.abs(G - Go).max() <= 1e-12 * np.abs(Go).max()
    if p == 1 and dpn == 1:
        G = m.gradient_zz(ctx.upload(u)).download().reshape(m.n_nodes, dim)
        Go = orc.grad_zz(dim, dn, el, u)
        assert np.abs(G - Go).max() <= 1e-10 * np.abs(Go).max()


def test_recovery_of_a_solved_poisson_field(H, ctx, orc):
    """Solve -lap u = 1 (Q1 64^2, homogeneous Dirichlet), recover the gradient
    both ways: each agrees with the oracle's recovery of the same solution, and
    by the problem's symmetry about x = 1/2 the x-gradient is odd."""
    m = H.Mesh.structured(ctx, 2, 1, 1, 64)
    A = m.pattern(H.CSR, H.FP64)
    R = ctx.upload(np.zeros(m.ndofs))
    m.assemble(H.physics(H.POISSON, 2, f=1.0), A, R)
    el, dn = m.download()
    mask = ((dn[:, 0] < 1e-9) | (dn[:, 0] > 1 - 1e-9) | (dn[:, 1] < 1e-9) | (dn[:, 1] > 1 - 1e-9)).astype(np.uint8)
    A.dirichlet(ctx.upload(mask), R)
    x = ctx.alloc(m.ndofs)
    x.upload(np.zeros(m.ndofs))
    info = A.solve(R, x, method=H.PCG_JACOBI, precision=H.FP64, stop=H.STOP_RELRES, tol=1e-12)
    assert info.status == 0
    u = x.download()
    for name, dev, ref in (("smoothing", m.gradient_local_smoothing(x), orc.grad_local_smoothing(2, 1, 1, dn, el, u)),
                           ("zz", m.gradient_zz(x), orc.grad_zz(2, dn, el, u))):
        G = dev.download().reshape(-1, 2)
        assert np.abs(G - ref).max() <= 1e-10 * np.abs(ref).max(), name
        # mirror x -> 1 - x maps node i to j with the same y: G_x(j) = -G_x(i)
        ix, iy = np.round(dn[:, 0] * 64).astype(int), np.round(dn[:, 1] * 64).astype(int)
        at = {(a, b): i for i, (a, b) in enumerate(zip(ix, iy))}
        mirror = np.array([at[(64 - a, b)] for a, b in zip(ix, iy)])
        assert np.abs(G[:, 0] + G[mirror, 0]).max() <= 1e-8 * np.abs(G[:, 0]).max(), name


def test_recovery_rejects_what_the_reference_does_not_define(H, ctx):
    m = H.Mesh.structured(ctx, 2, 2, 1, 4)
    with pytest.raises(H.YafelHipError):
        m.gradient_zz(ctx.upload(np.zeros(m.ndofs)))  # ZZ is for p = 1
    m3 = H.Mesh.structured(ctx, 3, 1, 3, 3)
    with pytest.raises(H.YafelHipError):
        m3.gradient_zz(ctx.upload(np.zeros(m3.ndofs)))  # scalar fields only


def test_zz_node_of_no_element_is_zero(H, ctx, orc):
    """A node no cell references has an empty patch: its ZZ gradient is zero
    (the least-squares solve of an empty system) and the other nodes' values
    are those of the mesh without it."""
    nodes, cells = H.structured_mesh(2, 9, 9, 0, 1)
    nodes = _graded(nodes, 2)
    extra = np.vstack([nodes, [[2.0, 2.0, 0.0]]])
    m0, m1 = H.Mesh(ctx, 2, 1, 1, nodes, cells), H.Mesh(ctx, 2, 1, 1, extra, cells)
    u = np.random.default_rng(3).standard_normal(m1.ndofs)
    G0 = m0.gradient_zz(ctx.upload(u[:-1])).download().reshape(-1, 2)
    G1 = m1.gradient_zz(ctx.upload(u)).download().reshape(-1, 2)
    assert np.array_equal(G1[-1], [0.0, 0.0])
    assert np.array_equal(G1[:-1], G0)
    el, dn = m1.download()
    assert np.abs(G1 - orc.grad_zz(2, dn, el, u)).max() <= 1e-10 * np.abs(G0).max()
