"""ctypes front end of the CPU oracle (oracle/oracle.cpp) and of the reference
harness (oracle/_ref/libyafel_ref*.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke entry point as the
checker and by bench.py's CPU-baseline leg.  The product never imports it.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_P = C.c_void_p
_I64 = C.c_int64
_D = C.c_double


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


class Physics(C.Structure):
    _fields_ = [("kind", C.c_int), ("nsd", C.c_int), ("coef", C.c_void_p), ("f", C.c_double),
                ("lam", C.c_double), ("mu", C.c_double), ("body", C.c_double * 3),
                ("U", C.c_double * 3), ("kdiff", C.c_double), ("h", C.c_double),
                ("alpha", C.c_double), ("Q", C.c_double)]


def physics(kind, nsd, coef=None, f=0.0, lam=0.0, mu=0.0, body=(0, 0, 0), U=(0, 0, 0),
            kdiff=0.0, h=0.0, alpha=0.0, Q=0.0):
    ph = Physics()
    ph.kind, ph.nsd, ph.f, ph.lam, ph.mu = kind, nsd, f, lam, mu
    ph._coef = None if coef is None else np.ascontiguousarray(coef, dtype=np.float64)
    ph.coef = None if coef is None else ph._coef.ctypes.data
    ph.body[:] = list(body)
    ph.U[:] = list(U)
    ph.kdiff, ph.h, ph.alpha, ph.Q = kdiff, h, alpha, Q
    return ph


def _load(name):
    lib = C.CDLL(os.path.join(HERE, name))
    return lib


class Oracle:
    """The C++ restatement of the reference CPU path."""

    def __init__(self, fast=False):
        self.lib = _load("liboracle_fast.so" if fast else "liboracle.so")
        L = self.lib
        for fn, res in [("orc_csr_from_triplets", _P), ("orc_bcsr3_from_triplets", _P),
                        ("orc_csr_wrap", _P), ("orc_csr_rows", _I64), ("orc_csr_cols", _I64),
                        ("orc_csr_nnz", _I64), ("orc_csr_nptr", _I64), ("orc_dot", _D),
                        ("orc_cg_solve", _I64), ("orc_pcg_jacobi_solve", _I64),
                        ("orc_bicgstab_solve", _I64), ("orc_bicgstab_jacobi_solve", _I64),
                        ("orc_dofmanager_cg", _I64), ("orc_tensor_rule", C.c_int)]:
            getattr(L, fn).restype = res
        L.orc_csr_from_triplets.argtypes = [_I64, _P, _P, _P]
        L.orc_bcsr3_from_triplets.argtypes = [_I64, _P, _P, _P]
        L.orc_csr_wrap.argtypes = [C.c_int, _I64, _I64, _P, _P, _P]
        for fn in ["orc_csr_rows", "orc_csr_cols", "orc_csr_nnz", "orc_csr_nptr", "orc_csr_free"]:
            getattr(L, fn).argtypes = [_P]
        L.orc_csr_copy.argtypes = [_P, _P, _P, _P]
        L.orc_spmv.argtypes = [_P, _P, _P]
        L.orc_dot.argtypes = [_I64, _P, _P]
        L.orc_dirichlet_apply.argtypes = [_P, _P, _P, _P]
        L.orc_cg_solve.argtypes = [_P, _P, _P, _D, _P]
        L.orc_pcg_jacobi_solve.argtypes = [_P, _P, _P, _D, _P]
        L.orc_bicgstab_solve.argtypes = [_P, _P, _P, _D, _I64, _P]
        L.orc_bicgstab_jacobi_solve.argtypes = [_P, _P, _D, _I64, _P]
        L.orc_dofmanager_cg.argtypes = [C.c_int, C.c_int, _I64, _P, _I64, _P, _P, _P]
        L.orc_cg_assembly.argtypes = [C.c_int, C.c_int, C.c_int, _I64, _P, _I64, _P, _P, _P, _P,
                                      _P, _P]
        L.orc_cg_assembly_threads.restype = _I64
        L.orc_cg_assembly_threads.argtypes = [C.c_int] * 4 + [_I64, _P, _I64, _P, _P, _P, _P, _P, _P]
        L.orc_element_matrix.argtypes = [C.c_int, C.c_int, C.c_int, _P, _P, _I64, _P, _P, _P]
        L.orc_structured_mesh.argtypes = [C.c_int] * 5 + [_P, _P]
        L.orc_shape_tables.argtypes = [C.c_int, C.c_int, _P, _P, _P, _P]
        L.orc_local_nodes.argtypes = [C.c_int, C.c_int, _P]
        L.orc_tensor_rule.argtypes = [C.c_int, C.c_int, C.c_int, _P, _P]
        L.orc_gauss_legendre_1d.argtypes = [C.c_int, _P, _P]
        L.orc_gauss_lobatto_1d.argtypes = [C.c_int, _P, _P]
        L.orc_normal_fill.argtypes = [C.c_uint64, _I64, _D, _D, _P]
        L.orc_dg_faces.restype = _I64
        L.orc_dg_faces.argtypes = [_I64, _P, _P, _P, _P, _P]
        L.orc_dg_nodes.argtypes = [C.c_int, _P, _I64, _P, _P]
        L.orc_dg_residual.argtypes = [C.c_int, _P, _I64, _P, _P, _D, _P, _P, _P]
        L.orc_dg_rk4.argtypes = [C.c_int, _P, _I64, _P, _P, _D, _I64, _D, _P]
        L.orc_grad_local_smoothing.argtypes = [C.c_int, C.c_int, C.c_int, _I64, _P, _I64, _P, _P, _P]
        L.orc_grad_zz.argtypes = [C.c_int, _I64, _P, _I64, _P, _P, _P]

    # -- quadrature / shape functions -------------------------------------
    def gauss_legendre(self, n):
        x, w = np.zeros(n), np.zeros(n)
        self.lib.orc_gauss_legendre_1d(n, _ptr(x), _ptr(w))
        return x, w

    def gauss_lobatto(self, n):
        x, w = np.zeros(n), np.zeros(n)
        self.lib.orc_gauss_lobatto_1d(n, _ptr(x), _ptr(w))
        return x, w

    def tensor_rule(self, qtype, dim, order):
        xi, w = np.zeros((1024, 3)), np.zeros(1024)
        n = self.lib.orc_tensor_rule(qtype, dim, order, _ptr(xi), _ptr(w))
        return xi[:n].copy(), w[:n].copy()

    def shape_tables(self, dim, p):
        nen = (p + 1) ** dim
        N = np.zeros((nen, nen))
        dN = np.zeros((nen, nen, dim))
        w = np.zeros(nen)
        xi = np.zeros((nen, 3))
        self.lib.orc_shape_tables(dim, p, _ptr(N), _ptr(dN), _ptr(w), _ptr(xi))
        return N, dN, w, xi

    def local_nodes(self, dim, p):
        xl = np.zeros(((p + 1) ** dim, 3))
        self.lib.orc_local_nodes(dim, p, _ptr(xl))
        return xl

    # -- meshes / dofs -------------------------------------------------------
    def structured_mesh(self, dim, nx, ny, nz=0, cell_order=0):
        nn = (nx + 1) * (ny + 1) * ((nz + 1) if dim == 3 else 1)
        nc = nx * ny * (nz if dim == 3 else 1)
        nodes = np.zeros((nn, 3))
        cells = np.zeros((nc, 2 ** dim), dtype=np.int32)
        self.lib.orc_structured_mesh(dim, nx, ny, nz, cell_order, _ptr(nodes), _ptr(cells))
        return nodes, cells

    def dofmanager(self, dim, p, nodes, cells):
        nen = (p + 1) ** dim
        nc = cells.shape[0]
        el = np.zeros((nc, nen), dtype=np.int32)
        dn = np.zeros((max(nc * nen, nodes.shape[0]), 3))
        nu = self.lib.orc_dofmanager_cg(dim, p, nodes.shape[0], _ptr(np.ascontiguousarray(nodes)),
                                        nc, _ptr(np.ascontiguousarray(cells)), _ptr(el), _ptr(dn))
        return el, dn[:nu].copy()

    def mesh(self, dim, p, n, nz=None):
        """Structured unit-domain mesh, numbered like DoFManager: (elements, dof_nodes)."""
        nodes, cells = self.structured_mesh(dim, n, n, (nz or n) if dim == 3 else 0,
                                            0 if p == 1 else 1)
        return self.dofmanager(dim, p, nodes, cells)

    # -- assembly ----------------------------------------------------------
    def assemble(self, dim, p, dpn, dof_nodes, elements, ph, want=("K", "R")):
        nel = elements.shape[0]
        nld = elements.shape[1] * dpn
        T = nel * nld * nld
        rows = np.zeros(T, np.int64) if "K" in want else None
        cols = np.zeros(T, np.int64) if "K" in want else None
        vals = np.zeros(T) if "K" in want else None
        R = np.zeros(dof_nodes.shape[0] * dpn) if "R" in want else None
        self.lib.orc_cg_assembly(dim, p, dpn, dof_nodes.shape[0], _ptr(dof_nodes), nel,
                                 _ptr(elements), C.byref(ph), _ptr(rows), _ptr(cols), _ptr(vals),
                                 _ptr(R))
        return rows, cols, vals, R

    def assemble_threads(self, nthreads, dim, p, dpn, dof_nodes, elements, ph):
        """CGAssembly with its OpenMP structure on nthreads threads (triplets in
        append order); the timed CPU baseline's element loop."""
        nel = elements.shape[0]
        nld = elements.shape[1] * dpn
        T = nel * nld * nld
        rows, cols, vals = np.zeros(T, np.int64), np.zeros(T, np.int64), np.zeros(T)
        R = np.zeros(dof_nodes.shape[0] * dpn)
        nt = self.lib.orc_cg_assembly_threads(nthreads, dim, p, dpn, dof_nodes.shape[0], _ptr(dof_nodes), nel,
                                              _ptr(elements), C.byref(ph), _ptr(rows), _ptr(cols), _ptr(vals),
                                              _ptr(R))
        assert nt == T
        return rows, cols, vals, R

    # -- gradient recovery -------------------------------------------------
    def grad_local_smoothing(self, dim, p, dpn, dof_nodes, elements, u):
        """LocalSmoothingGradient: (n_nodes*dpn, dim)."""
        dof_nodes = np.ascontiguousarray(dof_nodes, np.float64)
        elements = np.ascontiguousarray(elements, np.int32)
        u = np.ascontiguousarray(u, np.float64)
        G = np.zeros((dof_nodes.shape[0] * dpn, dim))
        self.lib.orc_grad_local_smoothing(dim, p, dpn, dof_nodes.shape[0], _ptr(dof_nodes), elements.shape[0],
                                          _ptr(elements), _ptr(u), _ptr(G))
        return G

    def grad_zz(self, dim, dof_nodes, elements, u):
        """ZZGradientRecovery (p = 1, scalar): (n_nodes, dim)."""
        dof_nodes = np.ascontiguousarray(dof_nodes, np.float64)
        elements = np.ascontiguousarray(elements, np.int32)
        u = np.ascontiguousarray(u, np.float64)
        G = np.zeros((dof_nodes.shape[0], dim))
        self.lib.orc_grad_zz(dim, dof_nodes.shape[0], _ptr(dof_nodes), elements.shape[0], _ptr(elements),
                             _ptr(u), _ptr(G))
        return G

    def element_matrix(self, dim, p, dpn, dof_nodes, elements, e, ph):
        nld = elements.shape[1] * dpn
        K = np.zeros((nld, nld))
        R = np.zeros(nld)
        self.lib.orc_element_matrix(dim, p, dpn, _ptr(dof_nodes), _ptr(elements), e, C.byref(ph),
                                    _ptr(K), _ptr(R))
        return K, R

    # -- sparse ------------------------------------------------------------
    def csr(self, rows, cols, vals, block=1):
        fn = self.lib.orc_csr_from_triplets if block == 1 else self.lib.orc_bcsr3_from_triplets
        h = fn(len(rows), _ptr(np.ascontiguousarray(rows, np.int64)),
               _ptr(np.ascontiguousarray(cols, np.int64)), _ptr(np.ascontiguousarray(vals, np.float64)))
        return OMatrix(self, h, block)

    def wrap(self, block, nrows, ncols, ptr, col, val):
        h = self.lib.orc_csr_wrap(block, nrows, ncols, _ptr(np.ascontiguousarray(ptr, np.int64)),
                                  _ptr(np.ascontiguousarray(col, np.int64)),
                                  _ptr(np.ascontiguousarray(val, np.float64)))
        return OMatrix(self, h, block)

    def dot(self, a, b):
        return self.lib.orc_dot(len(a), _ptr(np.ascontiguousarray(a, np.float64)),
                                _ptr(np.ascontiguousarray(b, np.float64)))

    def normal(self, seed, n, mean=0.0, sd=1.0):
        out = np.zeros(n)
        self.lib.orc_normal_fill(seed, n, mean, sd, _ptr(out))
        return out

    # -- DG (Quad4 cells, dpn = 1) -------------------------------------------
    def dg_faces(self, cells):
        cells = np.ascontiguousarray(cells, np.int32)
        n = 4 * cells.shape[0]
        l, r, lf, rf = (np.zeros(n, np.int32) for _ in range(4))
        nf = self.lib.orc_dg_faces(cells.shape[0], _ptr(cells), _ptr(l), _ptr(r), _ptr(lf), _ptr(rf))
        return l[:nf].copy(), r[:nf].copy(), lf[:nf].copy(), rf[:nf].copy()

    def dg_nodes(self, p, nodes, cells):
        cells = np.ascontiguousarray(cells, np.int32)
        X = np.zeros((cells.shape[0] * (p + 1) ** 2, 3))
        self.lib.orc_dg_nodes(p, _ptr(np.ascontiguousarray(nodes, np.float64)), cells.shape[0], _ptr(cells),
                              _ptr(X))
        return X

    def dg_residual(self, p, nodes, cells, ph, U, t=0.0, face=False):
        cells = np.ascontiguousarray(cells, np.int32)
        U = np.ascontiguousarray(U, np.float64)
        R = np.zeros_like(U)
        Rf = np.zeros_like(U) if face else None
        self.lib.orc_dg_residual(p, _ptr(np.ascontiguousarray(nodes, np.float64)), cells.shape[0], _ptr(cells),
                                 C.byref(ph), t, _ptr(U), _ptr(R), _ptr(Rf))
        return (R, Rf) if face else R

    def dg_rk4(self, p, nodes, cells, ph, U, dt, nsteps, t0=0.0):
        cells = np.ascontiguousarray(cells, np.int32)
        U = np.array(U, dtype=np.float64)
        self.lib.orc_dg_rk4(p, _ptr(np.ascontiguousarray(nodes, np.float64)), cells.shape[0], _ptr(cells),
                            C.byref(ph), dt, nsteps, t0, _ptr(U))
        return U


class OMatrix:
    def __init__(self, orc, h, block):
        self.o, self.h, self.block = orc, h, block

    def __del__(self):
        try:
            self.o.lib.orc_csr_free(self.h)
        except Exception:
            pass

    @property
    def rows(self):
        return self.o.lib.orc_csr_rows(self.h)

    @property
    def nnz(self):
        return self.o.lib.orc_csr_nnz(self.h)

    def arrays(self):
        L = self.o.lib
        nptr, nnz = L.orc_csr_nptr(self.h), L.orc_csr_nnz(self.h)
        ptr = np.zeros(nptr, np.int64)
        col = np.zeros(nnz, np.int64)
        val = np.zeros(nnz * self.block * self.block)
        L.orc_csr_copy(self.h, _ptr(ptr), _ptr(col), _ptr(val))
        return ptr, col, val

    def spmv(self, x):
        x = np.ascontiguousarray(x, np.float64)
        y = np.zeros(self.rows)
        self.o.lib.orc_spmv(self.h, _ptr(x), _ptr(y))
        return y

    def dirichlet(self, rhs, mask, g=None):
        rhs = np.array(rhs, dtype=np.float64)
        g = np.zeros(len(rhs)) if g is None else np.ascontiguousarray(g, np.float64)
        self.o.lib.orc_dirichlet_apply(self.h, _ptr(rhs), _ptr(np.ascontiguousarray(mask, np.uint8)),
                                       _ptr(g))
        return rhs

    def cg(self, b, x0=None, tol=1e-14):
        b = np.ascontiguousarray(b, np.float64)
        x0 = np.zeros_like(b) if x0 is None else np.ascontiguousarray(x0, np.float64)
        x = np.zeros_like(b)
        it = self.o.lib.orc_cg_solve(self.h, _ptr(b), _ptr(x0), tol, _ptr(x))
        return x, it

    def pcg_jacobi(self, b, x0=None, tol=1e-14):
        b = np.ascontiguousarray(b, np.float64)
        x0 = np.zeros_like(b) if x0 is None else np.ascontiguousarray(x0, np.float64)
        x = np.zeros_like(b)
        it = self.o.lib.orc_pcg_jacobi_solve(self.h, _ptr(b), _ptr(x0), tol, _ptr(x))
        return x, it

    def bicgstab(self, b, x0=None, tol=1e-14, max_iter=0):
        b = np.ascontiguousarray(b, np.float64)
        x0 = np.zeros_like(b) if x0 is None else np.ascontiguousarray(x0, np.float64)
        x = np.zeros_like(b)
        it = self.o.lib.orc_bicgstab_solve(self.h, _ptr(b), _ptr(x0), tol, max_iter, _ptr(x))
        return x, it

    def bicgstab_jacobi(self, b, rtol=1e-8, max_iter=100000):
        b = np.ascontiguousarray(b, np.float64)
        x = np.zeros_like(b)
        it = self.o.lib.orc_bicgstab_jacobi_solve(self.h, _ptr(b), rtol, max_iter, _ptr(x))
        return x, it


class Reference:
    """The reference's own handmade sparse/Krylov templates and quadrature,
    compiled from the reference tree by oracle/build_ref.sh."""

    def __init__(self, fast=False):
        path = os.path.join(HERE, "_ref", "libyafel_ref_fast.so" if fast else "libyafel_ref.so")
        self.lib = C.CDLL(path)
        L = self.lib
        L.ref_csr_from_triplets.restype = _P
        L.ref_bcsr3_from_triplets.restype = _P
        L.ref_csr_from_triplets.argtypes = [_I64, _P, _P, _P]
        L.ref_bcsr3_from_triplets.argtypes = [_I64, _P, _P, _P]
        L.ref_matrix_info.argtypes = [_P, _P]
        L.ref_matrix_copy.argtypes = [_P, _P, _P, _P]
        L.ref_matrix_free.argtypes = [_P]
        L.ref_spmv.argtypes = [_P, _P, _P]
        L.ref_dot.restype = _D
        L.ref_dot.argtypes = [_I64, _P, _P]
        L.ref_cg_solve.argtypes = [_P, _P, _P, _D, _P]
        L.ref_pcg_jacobi_solve.argtypes = [_P, _P, _P, _D, _P]
        L.ref_bicgstab_solve.argtypes = [_P, _P, _P, _P]
        L.ref_quadrature_1d.argtypes = [C.c_int, C.c_int, _P, _P]
        L.ref_tensor_rule.restype = C.c_int
        L.ref_tensor_rule.argtypes = [C.c_int, C.c_int, C.c_int, _P, _P]
        L.ref_rand_mod.argtypes = [C.c_uint, C.c_int, C.c_int, _P]
        L.ref_matrix_from_arrays.restype = _P
        L.ref_matrix_from_arrays.argtypes = [C.c_int, _I64, _P, _P, _P]
        L.ref_cg_iterations.restype = _D
        L.ref_cg_iterations.argtypes = [_P, _P, C.c_int]

    @staticmethod
    def available(fast=False):
        return os.path.exists(os.path.join(HERE, "_ref", "libyafel_ref_fast.so" if fast else "libyafel_ref.so"))

    def matrix(self, rows, cols, vals, block=1):
        fn = self.lib.ref_csr_from_triplets if block == 1 else self.lib.ref_bcsr3_from_triplets
        h = fn(len(rows), _ptr(np.ascontiguousarray(rows, np.int64)),
               _ptr(np.ascontiguousarray(cols, np.int64)), _ptr(np.ascontiguousarray(vals, np.float64)))
        return RMatrix(self, h, block)

    def matrix_from_arrays(self, block, nrows, ptr, col, val):
        h = self.lib.ref_matrix_from_arrays(block, nrows, _ptr(np.ascontiguousarray(ptr, np.int64)),
                                            _ptr(np.ascontiguousarray(col, np.int64)),
                                            _ptr(np.ascontiguousarray(val, np.float64)))
        return RMatrix(self, h, block)

    def quadrature_1d(self, qtype, n):
        x, w = np.zeros(n), np.zeros(n)
        self.lib.ref_quadrature_1d(qtype, n, _ptr(x), _ptr(w))
        return x, w

    def tensor_rule(self, qtype, dim, order):
        xi, w = np.zeros((1024, 3)), np.zeros(1024)
        n = self.lib.ref_tensor_rule(qtype, dim, order, _ptr(xi), _ptr(w))
        return xi[:n].copy(), w[:n].copy()

    def dot(self, a, b):
        return self.lib.ref_dot(len(a), _ptr(np.ascontiguousarray(a, np.float64)),
                                _ptr(np.ascontiguousarray(b, np.float64)))

    def rand_mod(self, seed, n, mod):
        out = np.zeros(n)
        self.lib.ref_rand_mod(seed, n, mod, _ptr(out))
        return out


class RMatrix:
    def __init__(self, ref, h, block):
        self.r, self.h, self.block = ref, h, block
        info = np.zeros(4, np.int64)
        ref.lib.ref_matrix_info(h, _ptr(info))
        self.rows, self.cols, self.nnz, self.nptr = (int(v) for v in info)

    def __del__(self):
        try:
            self.r.lib.ref_matrix_free(self.h)
        except Exception:
            pass

    def arrays(self):
        ptr = np.zeros(self.nptr, np.int64)
        col = np.zeros(self.nnz, np.int64)
        val = np.zeros(self.nnz * self.block * self.block)
        self.r.lib.ref_matrix_copy(self.h, _ptr(ptr), _ptr(col), _ptr(val))
        return ptr, col, val

    def spmv(self, x):
        y = np.zeros(self.rows)
        self.r.lib.ref_spmv(self.h, _ptr(np.ascontiguousarray(x, np.float64)), _ptr(y))
        return y

    def cg(self, b, x0=None, tol=1e-14):
        b = np.ascontiguousarray(b, np.float64)
        x0 = np.zeros_like(b) if x0 is None else np.ascontiguousarray(x0, np.float64)
        x = np.zeros_like(b)
        self.r.lib.ref_cg_solve(self.h, _ptr(b), _ptr(x0), tol, _ptr(x))
        return x

    def pcg_jacobi(self, b, x0=None, tol=1e-14):
        b = np.ascontiguousarray(b, np.float64)
        x0 = np.zeros_like(b) if x0 is None else np.ascontiguousarray(x0, np.float64)
        x = np.zeros_like(b)
        self.r.lib.ref_pcg_jacobi_solve(self.h, _ptr(b), _ptr(x0), tol, _ptr(x))
        return x

    def bicgstab(self, b, x0=None):
        b = np.ascontiguousarray(b, np.float64)
        x0 = np.zeros_like(b) if x0 is None else np.ascontiguousarray(x0, np.float64)
        x = np.zeros_like(b)
        self.r.lib.ref_bicgstab_solve(self.h, _ptr(b), _ptr(x0), _ptr(x))
        return x

    def cg_iterations(self, b, k):
        return self.r.lib.ref_cg_iterations(self.h, _ptr(np.ascontiguousarray(b, np.float64)), k)
