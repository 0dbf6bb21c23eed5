"""The allocation-free solve (yafel_hip_solve_noalloc) against yafel_hip_solve:
the same kernels on the same inputs, so the same iterates bit for bit; two
calls in a row, the second on a smaller system; no device memory allocated
by the call; the size limit and the missing-buffer errors."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def system(H, ctx, dim, dpn, n):
    """An assembled, Dirichlet-constrained elasticity (dpn 3) or Poisson system."""
    m = H.Mesh.structured(ctx, dim, 1, dpn, n)
    A = m.pattern(H.BCSR3 if dpn == 3 else H.CSR, H.FP64)
    if dpn == 3:
        ph = H.physics(H.ELASTICITY, 3, lam=0.5769230769, mu=0.3846153846, body=(0.0, 0.0, -1.0))
    else:
        ph = H.physics(H.POISSON, 2, coef=np.linspace(0.5, 2.0, n * n), f=1.0)
    b = ctx.upload(np.zeros(m.ndofs))
    m.assemble(ph, A, b)
    mask = ctx.upload(np.zeros(m.ndofs, np.uint8))
    m.bc_mask_plane(mask, 0, 0.0)
    if dpn == 1:
        m.bc_mask_plane(mask, 0, 1.0)
    A.dirichlet(mask, b)
    return A, b, m.ndofs


def build_diagonal(H, ctx, A, nd):
    """yafel_hip_matrix_diagonal: builds the diagonal (allocating its buffers)."""
    d = ctx.alloc(nd)
    H.check(H.lib.yafel_hip_matrix_diagonal(A.h, d.ptr), ctx.h)


def c_solve(H, fn, A, b, x, **kw):
    """One solve through the C ABI entry point `fn`, from x = 0."""
    x.upload(np.zeros(x.n))
    o = A.opts(**kw)
    info = H.SolveInfo()
    H.check(fn(A.h, C.byref(o), b.ptr, x.ptr, C.byref(info)), A.ctx.h)
    return x.download(), info


def free_bytes():
    hip = C.CDLL("libamdhip64.so")
    free, total = C.c_size_t(), C.c_size_t()
    assert hip.hipMemGetInfo(C.byref(free), C.byref(total)) == 0
    return free.value


@pytest.mark.parametrize("method,prec", [("PCG_JACOBI", "MIXED"), ("CG", "FP64"), ("BICGSTAB_JACOBI", "MIXED")])
def test_noalloc_solve_matches_the_allocating_solve(method, prec):
    from yafel_amd import hip as H
    ctx = H.Context(0)
    kw = dict(method=getattr(H, method), precision=getattr(H, prec), stop=H.STOP_RELRES, tol=1e-8)
    # two calls in a row, the second on a smaller system (the scratch keeps
    # the first call's contents)
    for dim, dpn, n in ((3, 3, 12), (2, 1, 20)):
        A, b, nd = system(H, ctx, dim, dpn, n)
        build_diagonal(H, ctx, A, nd)
        A.sync_fp32()
        x = ctx.alloc(nd)
        x_ref, i_ref = c_solve(H, H.lib.yafel_hip_solve, A, b, x, **kw)
        x_na, i_na = c_solve(H, H.lib.yafel_hip_solve_noalloc, A, b, x, **kw)
        assert i_ref.status == 0 and i_ref.rel_residual < 1e-8, i_ref.as_dict()
        assert i_na.status == i_ref.status and i_na.iterations == i_ref.iterations
        assert i_na.refinements == i_ref.refinements and i_na.rel_residual == i_ref.rel_residual
        assert np.array_equal(x_na, x_ref)
    ctx.close()


def test_noalloc_solve_allocates_nothing_and_reports_its_limits():
    from yafel_amd import hip as H
    ctx = H.Context(0)
    kw = dict(method=H.PCG_JACOBI, precision=H.MIXED, stop=H.STOP_RELRES, tol=1e-8)
    A, b, nd = system(H, ctx, 2, 1, 16)
    x = ctx.alloc(nd)
    # no diagonal and no FP32 copy yet: refused, nothing allocated for them
    with pytest.raises(H.YafelHipError) as e:
        c_solve(H, H.lib.yafel_hip_solve_noalloc, A, b, x, **kw)
    assert e.value.code == H.EINVAL and "diagonal" in str(e.value)
    build_diagonal(H, ctx, A, nd)
    with pytest.raises(H.YafelHipError) as e:
        c_solve(H, H.lib.yafel_hip_solve_noalloc, A, b, x, **kw)
    assert e.value.code == H.EINVAL and "FP32" in str(e.value)
    A.sync_fp32()
    c_solve(H, H.lib.yafel_hip_solve_noalloc, A, b, x, **kw)
    ctx.sync()
    before = free_bytes()
    _, info = c_solve(H, H.lib.yafel_hip_solve_noalloc, A, b, x, **kw)
    ctx.sync()
    assert info.status == 0 and free_bytes() == before
    # more rows than the scratch holds: refused before any work
    big = H.Mesh.structured(ctx, 2, 1, 1, 1060)          # 1,125,721 rows
    Ab = big.pattern(H.CSR, H.FP64)
    xb = ctx.alloc(big.ndofs)
    with pytest.raises(H.YafelHipError) as e:
        c_solve(H, H.lib.yafel_hip_solve_noalloc, Ab, xb, xb, **kw)
    assert e.value.code == H.EINVAL and "YAFEL_HIP_NOALLOC_MAX_ROWS" in str(e.value)
    ctx.close()
