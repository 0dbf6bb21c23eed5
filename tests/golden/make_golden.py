"""Writes the golden fixtures under tests/golden/ from the reference's own
Eigen-free sources (oracle/_ref/libyafel_ref.so, built by oracle/build_ref.sh
from the reference tree in this container).

    python tests/golden/make_golden.py

Fixtures (numpy .npz, no pickles):
  kat_solvers.npz   the known-answer cases of test/test_cg_solve.cpp:12-62,
                    test/test_bicgstab_solve.cpp:12-62, test/test_csr_spmv.cpp:13-30,
                    test/test_bcsr_spmv.cpp:13-30 run through the reference templates
  quadrature.npz    QuadratureRule 1-D Gauss-Legendre/Lobatto and tensor rules
  sparse_build.npz  sparse_csr / sparse_bcsr<3> built from assembly triplets of
                    small meshes (the triplets come from the oracle's assembly;
                    the CSR/BCSR arrays from the reference constructors)
  solves.npz        cg_solve / pcg_solve(Jacobi) / bicgstab_solve on small
                    assembled systems
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle.oracle import Oracle, Reference, physics  # noqa: E402


def tridiag(n):
    r, c, v = [], [], []
    for i in range(n):
        r.append(i); c.append(i); v.append(-2.0)
    for i in range(1, n):
        r += [i, i - 1]; c += [i - 1, i]; v += [1.0, 1.0]
    return np.array(r), np.array(c), np.array(v)


def main():
    ref = Reference()
    orc = Oracle()
    out = {}
    # test_cg_solve.cpp test_1: identity, rhs(i) = rand()%100 after srand(0)
    n = 10
    rhs = ref.rand_mod(0, n, 100)
    I = ref.matrix(np.arange(n), np.arange(n), np.ones(n))
    out["id_rhs"] = rhs
    out["id_cg_x"] = I.cg(rhs)
    out["id_bicgstab_x"] = I.bicgstab(rhs)
    # test_cg_solve.cpp test_2: tridiag(1,-2,1), xref = i
    r, c, v = tridiag(n)
    T = ref.matrix(r, c, v)
    xref = np.arange(n, dtype=float)
    b = T.spmv(xref)
    out["tri_rows"], out["tri_cols"], out["tri_vals"] = r, c, v
    out["tri_b"] = b
    out["tri_cg_x"] = T.cg(b)
    out["tri_bicgstab_x"] = T.bicgstab(b)
    # test_csr_spmv.cpp / test_bcsr_spmv.cpp: identity * u, u(i) = i+1
    u = np.arange(1, 31, dtype=float)
    B3 = ref.matrix(np.arange(30), np.arange(30), np.ones(30), block=3)
    out["bcsr_id_y"] = B3.spmv(u)
    out["csr_id_y"] = ref.matrix(np.arange(10), np.arange(10), np.ones(10)).spmv(u[:10])
    np.savez(os.path.join(HERE, "kat_solvers.npz"), **out)

    q = {}
    for t, name in [(0, "gl"), (1, "gll")]:
        for npts in range(2, 7):
            x, w = ref.quadrature_1d(t, npts)
            q[f"{name}{npts}_x"], q[f"{name}{npts}_w"] = x, w
    for dim in (1, 2, 3):
        for order in (2, 4, 6, 8):
            xi, w = ref.tensor_rule(0, dim, order)
            q[f"tp{dim}_{order}_xi"], q[f"tp{dim}_{order}_w"] = xi, w
    np.savez(os.path.join(HERE, "quadrature.npz"), **q)

    # sparse construction from assembly triplets of small meshes
    s = {}
    cases = [("q1", 2, 1, 1, 6), ("q2", 2, 2, 1, 3), ("q4", 2, 4, 1, 2), ("hex", 3, 1, 3, 2)]
    lam, mu = 0.3 / (1.3 * 0.4), 1 / 2.6
    for tag, dim, p, dpn, nc in cases:
        el, dn = orc.mesh(dim, p, nc)
        if dpn == 1:
            ph = physics(0, dim, f=1.0)
        else:
            ph = physics(1, dim, lam=lam, mu=mu, body=(0, 0, -1))
        rows, cols, vals, R = orc.assemble(dim, p, dpn, dn, el, ph)
        A = ref.matrix(rows, cols, vals, block=1)
        ptr, col, val = A.arrays()
        s[f"{tag}_rows"], s[f"{tag}_cols"], s[f"{tag}_vals"] = rows, cols, vals
        s[f"{tag}_ptr"], s[f"{tag}_col"], s[f"{tag}_val"] = ptr, col, val
        s[f"{tag}_R"] = R
        if dpn == 3:
            Bm = ref.matrix(rows, cols, vals, block=3)
            bp, bc, bv = Bm.arrays()
            s[f"{tag}_bptr"], s[f"{tag}_bcol"], s[f"{tag}_bval"] = bp, bc, bv
    # duplicate accumulation (test_sparse_csr.cpp:50-61, test_sparse_bcsr.cpp:80-104)
    N = 60
    rr = np.array([N - 1, 0] * N)
    cc = np.array([N - 1, 0] * N)
    vv = np.ones(2 * N)
    for blk, tag in [(1, "dup_csr"), (3, "dup_bcsr")]:
        M = ref.matrix(rr, cc, vv, block=blk)
        p_, c_, v_ = M.arrays()
        s[f"{tag}_ptr"], s[f"{tag}_col"], s[f"{tag}_val"] = p_, c_, v_
        s[f"{tag}_dims"] = np.array([M.rows, M.cols])
    # random duplicates with integer values (test_sparse_csr.cpp test_4, seeded here)
    rng = np.random.default_rng(4)
    ri, ci = rng.integers(0, 20, 500), rng.integers(0, 20, 500)
    vi = rng.integers(-10, 11, 500).astype(float)
    M = ref.matrix(ri, ci, vi)
    s["rnd_rows"], s["rnd_cols"], s["rnd_vals"] = ri, ci, vi
    s["rnd_ptr"], s["rnd_col"], s["rnd_val"] = M.arrays()
    np.savez(os.path.join(HERE, "sparse_build.npz"), **s)

    # solves on an assembled Poisson system (Q1 8x8, all edges Dirichlet)
    v = {}
    el, dn = orc.mesh(2, 1, 8)
    rows, cols, vals, R = orc.assemble(2, 1, 1, dn, el, physics(0, 2, f=1.0))
    A = orc.csr(rows, cols, vals)
    mask = ((np.abs(dn[:, 0]) < 1e-6) | (np.abs(dn[:, 0] - 1) < 1e-6) |
            (np.abs(dn[:, 1]) < 1e-6) | (np.abs(dn[:, 1] - 1) < 1e-6)).astype(np.uint8)
    b = A.dirichlet(R, mask)
    ptr, col, val = A.arrays()
    Rm = ref.matrix(np.repeat(np.arange(len(ptr) - 1), np.diff(ptr)), col, val)
    v["p_ptr"], v["p_col"], v["p_val"], v["p_b"] = ptr, col, val, b
    v["p_cg_x"] = Rm.cg(b, tol=1e-20)
    v["p_pcg_x"] = Rm.pcg_jacobi(b, tol=1e-20)
    np.savez(os.path.join(HERE, "solves.npz"), **v)
    print("wrote fixtures to", HERE)


if __name__ == "__main__":
    main()
