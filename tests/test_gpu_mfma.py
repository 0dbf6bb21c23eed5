"""FP32 matrix-core element matrices (2-D Poisson, p = 3, 4) against the FP64
scalar element path: same pattern, values within FP32 rounding, same FP64
residual, and a solve whose error against the FP64-assembled system is at
the FP32 level."""
import numpy as np
import pytest

from tests.test_oracle_dg import rect_mesh

pytestmark = pytest.mark.gpu


def system(ctx, orc, p, n, flag):
    from yafel_amd import hip as H
    nodes, cells = rect_mesh(orc, n, n, 1.0, 1.0, jitter=0.2, seed=21)
    coef = np.exp(0.5 * np.random.default_rng(22).standard_normal(n * n))
    ph = H.physics(H.POISSON, 2, coef=coef, f=1.5)
    mesh = H.Mesh(ctx, 2, p, 1, nodes, cells)
    A = mesh.pattern(H.CSR, H.FP64 | H.FP32)
    R = ctx.upload(np.zeros(mesh.ndofs))
    mesh.assemble(ph, A, R, H.TANGENT | H.RESIDUAL | flag)
    return mesh, A, R


@pytest.mark.parametrize("p", [3, 4])
def test_fp32_mfma_element_matrices_match_fp64(ctx, orc, p):
    from yafel_amd import hip as H
    _, A64, R64 = system(ctx, orc, p, 12, 0)
    _, A32, R32 = system(ctx, orc, p, 12, H.FP32_ELEMENTS)
    p64, c64, v64 = A64.arrays()
    p32, c32, v32 = A32.arrays()
    assert np.array_equal(p64, p32) and np.array_equal(c64, c32)
    scale = np.abs(v64).max()
    err = np.abs(v32 - v64).max() / scale
    assert err < 2e-6, err
    assert err > 0  # the FP32 path is really FP32
    np.testing.assert_allclose(R32.download(), R64.download(), rtol=1e-13, atol=1e-15)


def test_fp32_elements_solve_within_fp32_rounding_of_fp64(ctx, orc):
    from yafel_amd import hip as H
    p = 4
    m64, A64, R64 = system(ctx, orc, p, 10, 0)
    m32, A32, R32 = system(ctx, orc, p, 10, H.FP32_ELEMENTS)
    mask = ctx.upload(np.zeros(m64.ndofs, np.uint8))
    for d in range(2):
        m64.bc_mask_plane(mask, d, 0.0)
        m64.bc_mask_plane(mask, d, 1.0)
    A64.dirichlet(mask, R64)
    A32.dirichlet(mask, R32)
    b = R64.download()
    x64, i64 = A64.solve_host(b, method=H.PCG_JACOBI, precision=H.FP64, stop=H.STOP_RELRES, tol=1e-10)
    x32, i32 = A32.solve_host(b, method=H.PCG_JACOBI, precision=H.MIXED, stop=H.STOP_RELRES, tol=1e-10)
    assert i64.status == 0 and i32.status == 0
    assert np.abs(x32 - x64).max() <= 1e-3 * np.abs(x64).max()


def test_fp32_elements_reject_unsupported_cases(ctx, orc):
    from yafel_amd import hip as H
    nodes, cells = rect_mesh(orc, 4, 4)
    mesh = H.Mesh(ctx, 2, 2, 1, nodes, cells)
    A = mesh.pattern(H.CSR, H.FP64)
    with pytest.raises(H.YafelHipError) as e:
        mesh.assemble(H.physics(H.POISSON, 2), A, None, H.TANGENT | H.FP32_ELEMENTS)
    assert e.value.code == 1


@pytest.mark.parametrize("p", [3, 4])
def test_fp32_element_matrices_against_fp64_elementwise(ctx, orc, p):
    from yafel_amd import hip as H
    nodes, cells = rect_mesh(orc, 6, 5, 1.3, 0.8, jitter=0.25, seed=31)
    coef = np.exp(0.5 * np.random.default_rng(32).standard_normal(cells.shape[0]))
    mesh = H.Mesh(ctx, 2, p, 1, nodes, cells)
    K64 = mesh.element_matrices(H.physics(H.POISSON, 2, coef=coef))
    cd = ctx.upload(coef)
    K32 = mesh.element_matrices_fp32(H.physics(H.POISSON, 2, coef_dev=cd)).download().reshape(K64.shape)
    for e in range(K64.shape[0]):
        assert np.abs(K32[e] - K64[e]).max() <= 2e-6 * np.abs(K64[e]).max()


@pytest.mark.parametrize("p", [3, 4])
def test_fp32_assembly_writes_the_element_order_fp64_sums(ctx, orc, p):
    """The matrix-core assembly writes entries with one contributing element
    straight into the matrix and sums shared entries from a compact staging:
    every stored value is bit for bit the FP64 sum, in element order, of the
    FP32 element matrices."""
    from yafel_amd import hip as H
    n = 7
    nodes, cells = rect_mesh(orc, n, n + 1, 1.1, 0.9, jitter=0.2, seed=41)
    coef = np.exp(0.5 * np.random.default_rng(42).standard_normal(cells.shape[0]))
    mesh = H.Mesh(ctx, 2, p, 1, nodes, cells)
    cd = ctx.upload(coef)
    ph = H.physics(H.POISSON, 2, coef_dev=cd)
    A = mesh.pattern(H.CSR, H.FP64)
    mesh.assemble(ph, A, None, H.TANGENT | H.FP32_ELEMENTS)
    mesh.assemble(ph, A, None, H.TANGENT | H.FP32_ELEMENTS)   # the map is reused
    ptr, col, val = A.arrays()
    el, _ = mesh.download()
    nen = (p + 1) ** 2
    K32 = mesh.element_matrices_fp32(ph).download().reshape(-1, nen, nen)
    rows = np.repeat(el, nen, axis=1).ravel()
    cols = np.tile(el, (1, nen)).ravel()
    nd = mesh.ndofs
    acc = np.zeros(nd * nd)
    np.add.at(acc, rows.astype(np.int64) * nd + cols, K32.astype(np.float64).ravel())
    r = np.repeat(np.arange(nd), np.diff(ptr))
    assert np.array_equal(val, acc[r * nd + col])
    shared = np.bincount(rows.astype(np.int64) * nd + cols, minlength=nd * nd)[r * nd + col]
    assert (shared > 1).any() and (shared == 1).any()
