"""Solves large enough that every Krylov kernel runs on many workgroups (up to
the grid cap), where the per-workgroup folds of the partial sums and the
versioned solver state have to agree across all workgroups and XCDs.

- Poisson Q1 400x400 (160,801 dofs, hundreds of workgroups per pass): the FP64
  Jacobi-PCG iteration count under the reference criterion equals the
  oracle's within +-2, and the solution agrees.
- Elasticity Q1 48^3 (352,947 dofs): the same solve repeated gives the same
  iteration count and bit-identical solutions, FP64 and mixed."""
import numpy as np
import pytest

from oracle.oracle import physics as ophysics

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def H():
    from yafel_amd import hip
    return hip


def test_poisson_400_pcg_matches_the_oracle_count(H, ctx, orc):
    n = 400
    ph, oph = H.physics(H.POISSON, 2, f=1.0), ophysics(0, 2, f=1.0)
    m = H.Mesh.structured(ctx, 2, 1, 1, n)
    A = m.pattern(H.CSR, H.FP64 | H.FP32)
    R = ctx.upload(np.zeros(m.ndofs))
    m.assemble(ph, A, R)
    el, dn = orc.mesh(2, 1, n)
    mask = np.zeros(len(dn), bool)
    for d in range(2):
        mask |= (np.abs(dn[:, d]) < 1e-6) | (np.abs(dn[:, d] - 1) < 1e-6)
    mask = mask.astype(np.uint8)
    A.dirichlet(ctx.upload(mask), R)
    rows, cols, vals, Ro = orc.assemble(2, 1, 1, dn, el, oph)
    O = orc.csr(rows, cols, vals)
    bo = O.dirichlet(Ro, mask)
    b = R.download()
    x, info = A.solve_host(b, method=H.PCG_JACOBI, tol=1e-16)
    xo, ito = O.pcg_jacobi(bo, tol=1e-16)
    print(f"poisson {n}^2: gpu {info.iterations} oracle {ito}", flush=True)
    assert ito > 300
    assert abs(info.iterations - ito) <= 2
    assert np.abs(x - xo).max() <= 1e-6 * np.abs(xo).max()


@pytest.mark.parametrize("prec", ["FP64", "MIXED"])
def test_elasticity_48_repeats_bit_identically(H, ctx, prec):
    from yafel_amd import workloads
    w = workloads.c2(ctx, n=48)
    kw = dict(method=H.PCG_JACOBI, precision=getattr(H, prec), stop=H.STOP_RELRES, tol=1e-8)
    runs = []
    for _ in range(3):
        info = w.step(**kw)
        runs.append((info.iterations, info.refinements, w.x.download()))
        assert info.status == 0 and info.rel_residual < 1e-8
    print(f"elasticity 48^3 {prec}: iterations {[r[0] for r in runs]}", flush=True)
    for it, rf, x in runs[1:]:
        assert it == runs[0][0] and rf == runs[0][1]
        assert np.array_equal(x, runs[0][2])



@pytest.mark.parametrize("fmt", ["BCSR3", "CSR"])
@pytest.mark.parametrize("prec", ["FP64", "MIXED", "MIXED_MATRIX"])
@pytest.mark.parametrize("method", ["CG", "PCG_JACOBI"])
@pytest.mark.parametrize("stop", ["STOP_REFERENCE", "STOP_RELRES"])
def test_cg_two_and_three_passes_are_bit_identical(H, ctx, fmt, prec, method, stop):
    """The two-pass CG iteration (beta, x += alpha p and p = z + beta p formed
    inside the next SpMV) runs the same arithmetic as the three-pass one, so
    its iteration counts and iterates are bit-identical: elasticity Q1 20^3
    (27,783 dofs) in both storage formats, every precision, both stopping
    rules."""
    from yafel_amd import workloads
    w = workloads.c2(ctx, n=20, fmt=getattr(H, fmt))
    w.assemble()
    tol = 1e-8 if stop == "STOP_RELRES" else 1e-14
    kw = dict(method=getattr(H, method), precision=getattr(H, prec), stop=getattr(H, stop), tol=tol)
    out = {}
    for passes in (3, 2):
        w.A.set_cg_passes(passes)
        info = w.solve(**kw)
        out[passes] = (info.status, info.iterations, info.refinements, w.x.download())
    w.A.set_cg_passes(0)
    print(f"{fmt} {prec} {method} {stop}: iterations {out[3][1]} / {out[2][1]}", flush=True)
    assert out[3][0] == 0 and out[3][1] > 20
    assert out[2][:3] == out[3][:3]
    assert np.array_equal(out[2][3], out[3][3])
