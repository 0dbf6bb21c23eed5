"""DG on the device against the oracle: face list, dof nodes, the residual
(R and its face part) on jittered meshes for p = 1..4, the closed forms, and
DGRK4 through the captured graph."""
import numpy as np
import pytest

from oracle.oracle import physics as ophysics
from tests.test_oracle_dg import rect_mesh

pytestmark = pytest.mark.gpu

DG = 3


def make(ctx, orc, p, nx, ny, c=(-1.0, 0.35), f=0.0, jitter=0.2, seed=11):
    from yafel_amd import hip as H
    nodes, cells = rect_mesh(orc, nx, ny, 1.25, 0.8, jitter=jitter, seed=seed)
    dg = H.DG(ctx, p, nodes, cells)
    dg.ph = H.physics(H.DG_ADVECTION, 2, f=f, U=(c[0], c[1], 0.0))
    return dg, nodes, cells, ophysics(DG, 2, f=f, U=(c[0], c[1], 0.0))


def test_faces_and_nodes_match_the_oracle(ctx, orc):
    for p in (1, 3):
        dg, nodes, cells, _ = make(ctx, orc, p, 7, 5)
        got = dg.faces()
        exp = orc.dg_faces(cells)
        for g, e in zip(got, exp):
            assert np.array_equal(g, e)
        assert dg.nboundary == 2 * (7 + 5)
        np.testing.assert_allclose(dg.nodes(), orc.dg_nodes(p, nodes, cells), rtol=0, atol=1e-15)


@pytest.mark.parametrize("p", [1, 2, 3, 4])
def test_residual_matches_the_oracle(ctx, orc, p):
    dg, nodes, cells, oph = make(ctx, orc, p, 9, 6, f=0.3)
    U = np.random.default_rng(p).standard_normal(dg.ndofs)
    Ud, Rd, Fd = ctx.upload(U), ctx.alloc(dg.ndofs), ctx.alloc(dg.ndofs)
    dg.residual(dg.ph, Ud, Rd, Fd)
    R, Rf = orc.dg_residual(p, nodes, cells, oph, U, face=True)
    r, f = Rd.download(), Fd.download()
    scale = np.abs(R).max()
    assert np.abs(f - Rf).max() <= 1e-12 * np.abs(Rf).max()
    assert np.abs(r - R).max() <= 1e-12 * scale


@pytest.mark.parametrize("p", [1, 3, 4])
@pytest.mark.parametrize("shape", ["rectangles", "parallelograms", "mixed"])
def test_affine_cells_match_the_oracle(ctx, orc, p, shape):
    """Parallelogram cells take the shared reference mass inverse scaled by
    1/det J; the residual still matches the oracle's per-cell inverse on
    rectangles, on a sheared (parallelogram) mesh and on a mesh where only
    some cells are parallelograms."""
    from yafel_amd import hip as H
    nx, ny = 9, 7
    jit = 0.0 if shape != "mixed" else 0.2
    nodes, cells = rect_mesh(orc, nx, ny, 1.25, 0.8, jitter=jit, seed=5)
    if shape == "parallelograms":
        nodes = nodes.copy()
        nodes[:, 0] += 0.4 * nodes[:, 1]
    if shape == "mixed":
        # straighten the nodes of the left half: those cells are rectangles again
        flat, _ = rect_mesh(orc, nx, ny, 1.25, 0.8, jitter=0.0, seed=5)
        left = flat[:, 0] < 0.5 * 1.25 + 1e-9
        nodes = nodes.copy()
        nodes[left] = flat[left]
    dg = H.DG(ctx, p, nodes, cells)
    if shape == "mixed":
        assert 0 < dg.naffine < len(cells)
    else:
        assert dg.naffine == len(cells)
    ph = H.physics(H.DG_ADVECTION, 2, f=0.3, U=(-1.0, 0.35, 0.0))
    oph = ophysics(DG, 2, f=0.3, U=(-1.0, 0.35, 0.0))
    U = np.random.default_rng(p + 17).standard_normal(dg.ndofs)
    Rd, Fd = ctx.alloc(dg.ndofs), ctx.alloc(dg.ndofs)
    dg.residual(ph, ctx.upload(U), Rd, Fd)
    R, Rf = orc.dg_residual(p, nodes, cells, oph, U, face=True)
    assert np.abs(Fd.download() - Rf).max() <= 1e-12 * np.abs(Rf).max()
    assert np.abs(Rd.download() - R).max() <= 1e-12 * np.abs(R).max()


def test_closed_forms_on_the_device(ctx, orc):
    # constant source, no advection: R == s
    dg, *_ = make(ctx, orc, 3, 5, 4, c=(0.0, 0.0), f=-0.75)
    U = np.random.default_rng(1).standard_normal(dg.ndofs)
    Rd = ctx.alloc(dg.ndofs)
    dg.residual(dg.ph, ctx.upload(U), Rd)
    np.testing.assert_allclose(Rd.download(), -0.75, rtol=0, atol=1e-13)
    # constant field: face terms sum to minus the outflow, here (1.0 * 0.8 + 0.35 * 1.25)
    dg, *_ = make(ctx, orc, 2, 6, 5, c=(-1.0, 0.35))
    Fd, Rd = ctx.alloc(dg.ndofs), ctx.alloc(dg.ndofs)
    dg.residual(dg.ph, ctx.upload(np.ones(dg.ndofs)), Rd, Fd)
    assert abs(Fd.download().sum() + (1.0 * 0.8 + 0.35 * 1.25)) < 1e-12


def test_rk4_graph_matches_the_oracle_and_replays_consistently(ctx, orc):
    p, dt = 2, 2e-3
    dg, nodes, cells, oph = make(ctx, orc, p, 8, 6)
    X = dg.nodes()
    U0 = np.sin(3 * X[:, 0]) * np.cos(2 * X[:, 1])
    exp = orc.dg_rk4(p, nodes, cells, oph, U0, dt, 6)
    Ud = ctx.upload(U0)
    t = dg.rk4(dg.ph, Ud, dt, 6)
    assert abs(t - 6 * dt) < 1e-15
    got = Ud.download()
    assert np.abs(got - exp).max() <= 1e-12 * np.abs(exp).max()
    # the same six steps one call at a time, with other device work in between
    Ud2 = ctx.upload(U0)
    scratch = ctx.alloc(dg.ndofs)
    t = 0.0
    for k in range(6):
        t = dg.rk4(dg.ph, Ud2, dt, 1, t)
        dg.residual(dg.ph, Ud, scratch)
    assert np.array_equal(Ud2.download(), got)
    assert abs(t - 6 * dt) < 1e-15


def test_errors_are_reported(ctx, orc):
    from yafel_amd import hip as H
    nodes, cells = rect_mesh(orc, 2, 2)
    with pytest.raises(H.YafelHipError) as e:
        H.DG(ctx, 5, nodes, cells)
    assert e.value.code == 1
    dg = H.DG(ctx, 2, nodes, cells)
    U, R = ctx.alloc(dg.ndofs), ctx.alloc(dg.ndofs)
    with pytest.raises(H.YafelHipError) as e:
        dg.residual(H.physics(H.POISSON, 2), U, R)
    assert e.value.code == 1
