"""ctypes binding of libyafel_hip.so (the C ABI declared in include/yafel_hip.h).

Raises at import time when the HIP library has not been built: there is no
CPU fallback.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libyafel_hip.so")

if not os.path.exists(LIB_PATH):
    raise ImportError(f"{LIB_PATH} is missing: run `python -m yafel_amd.build` "
                      "(there is no CPU fallback)")

lib = C.CDLL(LIB_PATH)

OK, EINVAL, EDEVICE, ENOMEM, ESTATE = 0, 1, 2, 3, 4
FP64, FP32, MIXED, MIXED_MATRIX = 1, 2, 3, 4
CSR, BCSR3 = 1, 3
CG, PCG_JACOBI, BICGSTAB, BICGSTAB_JACOBI = 0, 1, 2, 3
STOP_REFERENCE, STOP_RELRES = 0, 1
STATUS = {0: "converged", 1: "maxiter", 2: "breakdown", 3: "nan"}
POISSON, ELASTICITY, SUPG, DG_ADVECTION = 0, 1, 2, 3
TANGENT, RESIDUAL, FP32_ELEMENTS = 1, 2, 4

_P, _I64, _D, _I = C.c_void_p, C.c_int64, C.c_double, C.c_int


class MatrixInfo(C.Structure):
    _fields_ = [("format", _I), ("precisions", _I), ("rows", _I64), ("cols", _I64),
                ("nnz", _I64), ("nblocks", _I64), ("ptr_len", _I64)]


class SolveOpts(C.Structure):
    _fields_ = [("method", _I), ("precision", _I), ("stop", _I), ("tol", _D),
                ("max_iter", _I64), ("check_every", _I), ("refine_delta", _D)]


class SolveInfo(C.Structure):
    _fields_ = [("status", _I), ("iterations", _I64), ("refinements", _I64),
                ("rel_residual", _D), ("criterion", _D), ("ms", _D)]

    def as_dict(self):
        return {"status": STATUS.get(self.status, self.status), "iterations": self.iterations,
                "refinements": self.refinements, "rel_residual": self.rel_residual,
                "criterion": self.criterion, "ms": self.ms}


class Physics(C.Structure):
    _fields_ = [("kind", _I), ("nsd", _I), ("coef", _P), ("coef_on_device", _I), ("f", _D),
                ("lam", _D), ("mu", _D), ("body", _D * 3), ("U", _D * 3), ("kdiff", _D),
                ("h", _D), ("alpha", _D), ("Q", _D)]


class Request(C.Structure):
    _fields_ = [("dim", _I), ("p", _I), ("dof_per_node", _I), ("n_corner", _I64), ("corner_nodes", _P),
                ("ncell", _I64), ("cells", _P), ("phys", Physics), ("load", _P), ("load_len", _I64),
                ("bc", _I), ("tol", _D), ("max_iter", _I64)]


class ServerStats(C.Structure):
    _fields_ = [("submitted", _I64), ("admitted", _I64), ("completed", _I64), ("failed", _I64),
                ("cancelled", _I64), ("queued", _I64), ("active", _I64), ("chunks", _I64),
                ("batched_launches", _I64), ("hbm_in_use", C.c_size_t), ("hbm_peak", C.c_size_t),
                ("hbm_budget", C.c_size_t), ("setup_ms", C.c_double), ("graph_captures", _I64),
                ("graph_replays", _I64), ("two_pass_chunks", _I64), ("three_pass_chunks", _I64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


def _sig(name, res, args):
    fn = getattr(lib, name)
    fn.restype = res
    fn.argtypes = args


for _n, _a in [
    ("yafel_hip_ctx_create", [_I, C.POINTER(_P)]), ("yafel_hip_ctx_destroy", [_P]),
    ("yafel_hip_ctx_sync", [_P]), ("yafel_hip_malloc", [_P, C.c_size_t, C.POINTER(_P)]),
    ("yafel_hip_free", [_P, _P]), ("yafel_hip_memcpy_h2d", [_P, _P, _P, C.c_size_t]),
    ("yafel_hip_memcpy_d2h", [_P, _P, _P, C.c_size_t]),
    ("yafel_hip_memcpy_d2d", [_P, _P, _P, C.c_size_t]), ("yafel_hip_memzero", [_P, _P, C.c_size_t]),
    ("yafel_hip_timer_begin", [_P]), ("yafel_hip_timer_end", [_P, C.POINTER(_D)]),
    ("yafel_hip_csr_create", [_P, _I64, _I64, _P, _P, _P, _I, C.POINTER(_P)]),
    ("yafel_hip_bcsr3_create", [_P, _I64, _I64, _P, _P, _P, _I, C.POINTER(_P)]),
    ("yafel_hip_csr_from_triplets", [_P, _I64, _P, _P, _P, _I, C.POINTER(_P)]),
    ("yafel_hip_bcsr3_from_triplets", [_P, _I64, _P, _P, _P, _I, C.POINTER(_P)]),
    ("yafel_hip_matrix_destroy", [_P]), ("yafel_hip_matrix_info_get", [_P, C.POINTER(MatrixInfo)]),
    ("yafel_hip_matrix_download", [_P, _P, _P, _P]), ("yafel_hip_matrix_sync_fp32", [_P]),
    ("yafel_hip_matrix_diagonal", [_P, _P]), ("yafel_hip_matrix_set_unroll", [_P, _I]),
    ("yafel_hip_matrix_set_cg_passes", [_P, _I]), ("yafel_hip_matrix_index_bytes", [_P, _P]),
    ("yafel_hip_matrix_set_cache_bytes", [_P, _I64]),
    ("yafel_hip_spmv", [_P, _I, _P, _P]), ("yafel_hip_spmv_host", [_P, _P, _P]),
    ("yafel_hip_spmm", [_P, _I, _I, _P, _P]), ("yafel_hip_dot", [_P, _I64, _P, _P, C.POINTER(_D)]),
    ("yafel_hip_copy_bandwidth", [_P, C.c_size_t, _I, C.POINTER(_D)]),
    ("yafel_hip_read_bandwidth", [_P, C.c_size_t, _I, C.POINTER(_D)]),
    ("yafel_hip_kernel_timing", [_P, _I]), ("yafel_hip_kernel_time", [_P, C.POINTER(_D), C.POINTER(_I64)]),
    ("yafel_hip_solve_opts_default", [C.POINTER(SolveOpts)]),
    ("yafel_hip_solve", [_P, C.POINTER(SolveOpts), _P, _P, C.POINTER(SolveInfo)]),
    ("yafel_hip_solve_noalloc", [_P, C.POINTER(SolveOpts), _P, _P, C.POINTER(SolveInfo)]),
    ("yafel_hip_solve_host", [_P, C.POINTER(SolveOpts), _P, _P, _P, C.POINTER(SolveInfo)]),
    ("yafel_hip_solve_multi", [_P, C.POINTER(SolveOpts), _I, _P, _P, C.POINTER(SolveInfo)]),
    ("yafel_hip_mesh_create_cg", [_P, _I, _I, _I, _I64, _P, _I64, _P, C.POINTER(_P)]),
    ("yafel_hip_mesh_destroy", [_P]), ("yafel_hip_mesh_info", [_P, _P]),
    ("yafel_hip_mesh_download", [_P, _P, _P]), ("yafel_hip_mesh_nodes_dev", [_P, C.POINTER(_P)]),
    ("yafel_hip_mesh_pattern", [_P, _I, _I, C.POINTER(_P)]),
    ("yafel_hip_assemble", [_P, C.POINTER(Physics), _I, _P, _P]),
    ("yafel_hip_element_matrices", [_P, C.POINTER(Physics), _P]),
    ("yafel_hip_element_matrices_fp32", [_P, C.POINTER(Physics), _P]),
    ("yafel_hip_dirichlet_apply", [_P, _P, _P, _P]),
    ("yafel_hip_bc_mask_plane", [_P, _I, _D, _D, _I, _P]),
    ("yafel_hip_gradient_local_smoothing", [_P, _P, _P]), ("yafel_hip_gradient_zz", [_P, _P, _P]),
    ("yafel_hip_gmsh_read", [C.c_char_p, _P]), ("yafel_hip_gmsh_sizes", [_P, _P]),
    ("yafel_hip_gmsh_arrays", [_P] * 7), ("yafel_hip_gmsh_destroy", [_P]),
    ("yafel_hip_vtu_write", [C.c_char_p, _I, _I, _I64, _P, _I64, _P, _I, _P]),
    ("yafel_hip_structured_mesh", [_I, _I, _I, _I, _I, _P, _P]),
    ("yafel_hip_normal_fill", [C.c_uint64, _I64, _D, _D, _P]),
    ("yafel_hip_dg_create", [_P, _I, _I64, _P, _I64, _P, C.POINTER(_P)]),
    ("yafel_hip_dg_destroy", [_P]), ("yafel_hip_dg_info", [_P, _P]), ("yafel_hip_dg_affine_cells", [_P, _P]), ("yafel_hip_dg_nodes", [_P, _P]),
    ("yafel_hip_dg_faces", [_P, _P, _P, _P, _P]),
    ("yafel_hip_dg_residual", [_P, C.POINTER(Physics), _D, _P, _P, _P]),
    ("yafel_hip_dg_rk4", [_P, C.POINTER(Physics), _D, _I64, C.POINTER(_D), _P]),
    ("yafel_hip_server_create", [_P, C.c_size_t, _I, C.POINTER(_P)]),
    ("yafel_hip_server_create_ex", [_P, C.c_size_t, _I, _I, C.POINTER(_P)]),
    ("yafel_hip_server_destroy", [_P]),
    ("yafel_hip_server_submit", [_P, C.POINTER(Request), C.POINTER(_I64)]),
    ("yafel_hip_server_progress", [_P, _P, _I, C.POINTER(_I)]),
    ("yafel_hip_server_query", [_P, _I64, C.POINTER(_I), C.POINTER(_I64)]),
    ("yafel_hip_server_result", [_P, _I64, _P, _I64, C.POINTER(SolveInfo)]),
    ("yafel_hip_server_cancel", [_P, _I64]),
    ("yafel_hip_server_times", [_P, _I64, _P]),
    ("yafel_hip_server_stats_get", [_P, C.POINTER(ServerStats)]),
    ("yafel_hip_server_set_graphs", [_P, _I]),
    ("yafel_hip_server_set_cg_passes", [_P, _I, _I64]),
]:
    _sig(_n, _I, _a)
_sig("yafel_hip_last_error", C.c_char_p, [_P])
_sig("yafel_hip_ctx_stream", _P, [_P])


class YafelHipError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"yafel_hip error {code}: {msg}")
        self.code = code


def check(rc, ctx=None):
    if rc != OK:
        msg = lib.yafel_hip_last_error(ctx)
        raise YafelHipError(rc, msg.decode() if msg else "")


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class Context:
    def __init__(self, device=0):
        h = _P()
        check(lib.yafel_hip_ctx_create(device, C.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            lib.yafel_hip_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def sync(self):
        check(lib.yafel_hip_ctx_sync(self.h), self.h)

    def timer_begin(self):
        check(lib.yafel_hip_timer_begin(self.h), self.h)

    def timer_end(self):
        ms = _D()
        check(lib.yafel_hip_timer_end(self.h, C.byref(ms)), self.h)
        return ms.value

    def alloc(self, n, dtype=np.float64):
        return DeviceArray(self, n, dtype)

    def upload(self, a):
        a = np.ascontiguousarray(a)
        d = DeviceArray(self, a.size, a.dtype)
        d.upload(a)
        return d

    def kernel_timing(self, enable=True):
        check(lib.yafel_hip_kernel_timing(self.h, 1 if enable else 0), self.h)

    def kernel_time(self):
        """(summed ms, launches) of the CG kernel-A launches that ran the product
        since timing was enabled (launches that returned at once are left out)."""
        ms, n = C.c_double(), _I64()
        check(lib.yafel_hip_kernel_time(self.h, C.byref(ms), C.byref(n)), self.h)
        return ms.value, n.value

    def copy_bandwidth(self, nbytes=1 << 30, reps=20):
        g = _D()
        check(lib.yafel_hip_copy_bandwidth(self.h, nbytes, reps, C.byref(g)), self.h)
        return g.value

    def read_bandwidth(self, nbytes=1 << 30, reps=20):
        g = _D()
        check(lib.yafel_hip_read_bandwidth(self.h, nbytes, reps, C.byref(g)), self.h)
        return g.value

    def dot(self, a, b):
        r = _D()
        check(lib.yafel_hip_dot(self.h, a.n, a.ptr, b.ptr, C.byref(r)), self.h)
        return r.value


class DeviceArray:
    def __init__(self, ctx, n, dtype=np.float64):
        self.ctx, self.n, self.dtype = ctx, int(n), np.dtype(dtype)
        p = _P()
        check(lib.yafel_hip_malloc(ctx.h, max(self.n * self.dtype.itemsize, 8), C.byref(p)), ctx.h)
        self.ptr = p

    @property
    def nbytes(self):
        return self.n * self.dtype.itemsize

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.size == self.n
        check(lib.yafel_hip_memcpy_h2d(self.ctx.h, self.ptr, _ptr(a), self.nbytes), self.ctx.h)
        return self

    def download(self):
        out = np.empty(self.n, dtype=self.dtype)
        check(lib.yafel_hip_memcpy_d2h(self.ctx.h, _ptr(out), self.ptr, self.nbytes), self.ctx.h)
        return out

    def copy_from(self, other):
        check(lib.yafel_hip_memcpy_d2d(self.ctx.h, self.ptr, other.ptr, self.nbytes), self.ctx.h)

    def zero(self):
        check(lib.yafel_hip_memzero(self.ctx.h, self.ptr, self.nbytes), self.ctx.h)
        return self

    def free(self):
        if self.ptr:
            lib.yafel_hip_free(self.ctx.h, self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Matrix:
    def __init__(self, ctx, h):
        self.ctx, self.h = ctx, h

    @classmethod
    def csr(cls, ctx, rows, cols, row_ptr, col_index, data, precisions=FP64):
        h = _P()
        check(lib.yafel_hip_csr_create(ctx.h, rows, cols, _ptr(np.ascontiguousarray(row_ptr, np.int64)),
                                       _ptr(np.ascontiguousarray(col_index, np.int64)),
                                       _ptr(np.ascontiguousarray(data, np.float64)), precisions,
                                       C.byref(h)), ctx.h)
        return cls(ctx, h)

    @classmethod
    def bcsr3(cls, ctx, brows, bcols, brow_ptr, bcol_index, data, precisions=FP64):
        h = _P()
        check(lib.yafel_hip_bcsr3_create(ctx.h, brows, bcols, _ptr(np.ascontiguousarray(brow_ptr, np.int64)),
                                         _ptr(np.ascontiguousarray(bcol_index, np.int64)),
                                         _ptr(np.ascontiguousarray(data, np.float64)), precisions,
                                         C.byref(h)), ctx.h)
        return cls(ctx, h)

    @classmethod
    def from_triplets(cls, ctx, rows, cols, vals, block=1, precisions=FP64):
        h = _P()
        fn = lib.yafel_hip_csr_from_triplets if block == 1 else lib.yafel_hip_bcsr3_from_triplets
        check(fn(ctx.h, len(rows), _ptr(np.ascontiguousarray(rows, np.int64)),
                 _ptr(np.ascontiguousarray(cols, np.int64)), _ptr(np.ascontiguousarray(vals, np.float64)),
                 precisions, C.byref(h)), ctx.h)
        return cls(ctx, h)

    def close(self):
        if self.h:
            lib.yafel_hip_matrix_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self):
        i = MatrixInfo()
        check(lib.yafel_hip_matrix_info_get(self.h, C.byref(i)), self.ctx.h)
        return i

    @property
    def rows(self):
        return self.info().rows

    def arrays(self):
        i = self.info()
        ptr = np.zeros(i.ptr_len, np.int64)
        col = np.zeros(i.nblocks, np.int64)
        val = np.zeros(i.nnz)
        check(lib.yafel_hip_matrix_download(self.h, _ptr(ptr), _ptr(col), _ptr(val)), self.ctx.h)
        return ptr, col, val

    def set_unroll(self, unroll):
        check(lib.yafel_hip_matrix_set_unroll(self.h, unroll), self.ctx.h)

    def set_cache_bytes(self, nbytes):
        """Bytes of each matrix stream read with the caching policy (-1: default)."""
        check(lib.yafel_hip_matrix_set_cache_bytes(self.h, int(nbytes)), self.ctx.h)

    def index_bytes(self):
        """Bytes per column index in the SpMV streams (2: 16-bit offsets)."""
        b = C.c_int(0)
        check(lib.yafel_hip_matrix_index_bytes(self.h, C.byref(b)), self.ctx.h)
        return b.value

    def set_cg_passes(self, passes):
        """Kernels per single-solve CG iteration: 2, 3, or 0 (by format)."""
        check(lib.yafel_hip_matrix_set_cg_passes(self.h, passes), self.ctx.h)

    def sync_fp32(self):
        check(lib.yafel_hip_matrix_sync_fp32(self.h), self.ctx.h)

    def spmv_host(self, x):
        x = np.ascontiguousarray(x, np.float64)
        y = np.zeros(self.info().rows)
        check(lib.yafel_hip_spmv_host(self.h, _ptr(x), _ptr(y)), self.ctx.h)
        return y

    def spmv(self, x_dev, y_dev, precision=FP64):
        check(lib.yafel_hip_spmv(self.h, precision, x_dev.ptr, y_dev.ptr), self.ctx.h)

    def spmm(self, X_dev, Y_dev, k, precision=FP64):
        check(lib.yafel_hip_spmm(self.h, precision, k, X_dev.ptr, Y_dev.ptr), self.ctx.h)

    @staticmethod
    def opts(method=CG, precision=FP64, stop=STOP_REFERENCE, tol=1e-14, max_iter=0, check_every=16,
             refine_delta=0.1):
        o = SolveOpts()
        lib.yafel_hip_solve_opts_default(C.byref(o))
        o.method, o.precision, o.stop, o.tol = method, precision, stop, tol
        o.max_iter, o.check_every, o.refine_delta = max_iter, check_every, refine_delta
        return o

    def solve_host(self, b, x0=None, **kw):
        o = self.opts(**kw)
        b = np.ascontiguousarray(b, np.float64)
        x = np.zeros_like(b)
        info = SolveInfo()
        x0p = None if x0 is None else _ptr(np.ascontiguousarray(x0, np.float64))
        check(lib.yafel_hip_solve_host(self.h, C.byref(o), _ptr(b), x0p, _ptr(x), C.byref(info)),
              self.ctx.h)
        return x, info

    def solve(self, b_dev, x_dev, **kw):
        o = self.opts(**kw)
        info = SolveInfo()
        check(lib.yafel_hip_solve(self.h, C.byref(o), b_dev.ptr, x_dev.ptr, C.byref(info)), self.ctx.h)
        return info

    def solve_noalloc(self, b_dev, x_dev, **kw):
        """yafel_hip_solve with the library's own device scratch: allocates nothing."""
        o = self.opts(**kw)
        info = SolveInfo()
        check(lib.yafel_hip_solve_noalloc(self.h, C.byref(o), b_dev.ptr, x_dev.ptr, C.byref(info)), self.ctx.h)
        return info

    def solve_multi(self, B_dev, X_dev, k, **kw):
        o = self.opts(**kw)
        infos = (SolveInfo * k)()
        check(lib.yafel_hip_solve_multi(self.h, C.byref(o), k, B_dev.ptr, X_dev.ptr, infos), self.ctx.h)
        return list(infos)

    def dirichlet(self, mask_dev, rhs_dev, g_dev=None):
        check(lib.yafel_hip_dirichlet_apply(self.h, mask_dev.ptr, None if g_dev is None else g_dev.ptr,
                                            rhs_dev.ptr), self.ctx.h)


def structured_mesh(dim, nx, ny, nz=0, p=1):
    nn = (nx + 1) * (ny + 1) * ((nz + 1) if dim == 3 else 1)
    nc = nx * ny * (nz if dim == 3 else 1)
    nodes = np.zeros((nn, 3))
    cells = np.zeros((nc, 2 ** dim), np.int32)
    check(lib.yafel_hip_structured_mesh(dim, nx, ny, nz, p, _ptr(nodes), _ptr(cells)))
    return nodes, cells


CELL_LINE2, CELL_TRI3, CELL_TET4, CELL_QUAD4, CELL_HEX8, CELL_POINT1, CELL_NONE = range(7)
SCALAR, VECTOR, TENSOR = 0, 1, 2
POINT_DATA, CELL_DATA = 0, 1


class OutputData(C.Structure):
    _fields_ = [("name", C.c_char_p), ("data", C.c_void_p), ("on_device", C.c_int), ("dof_per_loc", C.c_int),
                ("mask", C.c_uint32), ("data_type", C.c_int), ("location", C.c_int)]


def write_vtu(fname_base, dim, p, points, cells, data=()):
    """SimulationOutput(fname_base, VTU) + captureFrame: writes fname_base.vtu.
    points (n, 3); cells (n_cells, (p+1)^dim) in the DoFManager's local node
    order; data: tuples (name, values, dof_mask, data_type, location) with values
    a numpy array or a DeviceArray, dof_mask a list of 0/1 (OutputData)."""
    points = np.ascontiguousarray(points, np.float64)
    cells = np.ascontiguousarray(cells, np.int32)
    recs, keep = (OutputData * max(len(data), 1))(), []
    for k, (name, vals, mask, dtype, loc) in enumerate(data):
        dev = not isinstance(vals, np.ndarray)
        if not dev:
            vals = np.ascontiguousarray(vals, np.float64)
            keep.append(vals)
        bits = sum(1 << c for c, m in enumerate(mask) if m)
        nb = name.encode()
        keep.append(nb)
        recs[k] = OutputData(nb, vals.ptr if dev else vals.ctypes.data, 1 if dev else 0, len(mask), bits, dtype, loc)
    check(lib.yafel_hip_vtu_write(os.fsencode(str(fname_base)), dim, p, points.shape[0], _ptr(points), cells.shape[0],
                                  _ptr(cells), len(data), recs))


class GmshMesh:
    """Mesh(const std::string &fname) (src/mesh/parse_gmsh.cpp:80-151): the
    arrays of a gmsh 2.2 ASCII file -- nodes (n, 3); cell_nodes (0-based) with
    cell_offsets (cells + 1); cell_types (CELL_*); tags with tag_offsets."""

    def __init__(self, path):
        h = _P()
        check(lib.yafel_hip_gmsh_read(os.fsencode(path), C.byref(h)))
        try:
            sz = np.zeros(4, np.int64)
            check(lib.yafel_hip_gmsh_sizes(h, _ptr(sz)))
            nn, nc, ncn, nt = (int(v) for v in sz)
            self.nodes = np.zeros((nn, 3))
            self.cell_nodes = np.zeros(ncn, np.int32)
            self.cell_offsets = np.zeros(nc + 1, np.int32)
            self.cell_types = np.zeros(nc, np.int32)
            self.tag_offsets = np.zeros(nc + 1, np.int32)
            self.tags = np.zeros(nt, np.int32)
            check(lib.yafel_hip_gmsh_arrays(h, _ptr(self.nodes), _ptr(self.cell_nodes), _ptr(self.cell_offsets),
                                            _ptr(self.cell_types), _ptr(self.tag_offsets), _ptr(self.tags)))
        finally:
            lib.yafel_hip_gmsh_destroy(h)

    def cell(self, c):
        return self.cell_nodes[self.cell_offsets[c]:self.cell_offsets[c + 1]]

    def cell_tags(self, c):
        return self.tags[self.tag_offsets[c]:self.tag_offsets[c + 1]]

    def volume_cells(self, dim):
        """The Quad4 (dim 2) or Hex8 (dim 3) cells as an (n, 2^dim) array in file
        order -- the cells CGAssembly assembles (topoDim == nsd); lines and
        points are boundary tags only."""
        want = CELL_QUAD4 if dim == 2 else CELL_HEX8
        # lower-dimensional cells (boundary lines, faces, points) are skipped
        lower = {CELL_LINE2, CELL_POINT1} | ({CELL_QUAD4, CELL_TRI3} if dim == 3 else set())
        other = {int(t) for t in self.cell_types} - lower - {want}
        if other:
            raise YafelHipError(EINVAL, f"gmsh mesh has cells the device elements do not cover: {sorted(other)}")
        sel = [c for c in range(len(self.cell_types)) if self.cell_types[c] == want]
        nen = 2 ** dim
        return np.array([self.cell(c) for c in sel], np.int32).reshape(len(sel), nen)


def normal(seed, n, mean=0.0, sd=1.0):
    out = np.zeros(n)
    check(lib.yafel_hip_normal_fill(seed, n, mean, sd, _ptr(out)))
    return out


def physics(kind, nsd, coef=None, coef_dev=None, f=0.0, lam=0.0, mu=0.0, body=(0, 0, 0),
            U=(0, 0, 0), kdiff=0.0, h=0.0, alpha=0.0, Q=0.0):
    ph = Physics()
    ph.kind, ph.nsd, ph.f, ph.lam, ph.mu = kind, nsd, f, lam, mu
    if coef_dev is not None:
        ph.coef, ph.coef_on_device = coef_dev.ptr.value, 1
        ph._keep = coef_dev
    elif coef is not None:
        ph._keep = np.ascontiguousarray(coef, np.float64)
        ph.coef, ph.coef_on_device = ph._keep.ctypes.data, 0
    else:
        ph.coef, ph.coef_on_device = None, 0
    ph.body[:] = list(body)
    ph.U[:] = list(U)
    ph.kdiff, ph.h, ph.alpha, ph.Q = kdiff, h, alpha, Q
    return ph


class DG:
    """DoFManager(M, DG, p, 1) + DGAssembly / DGRK4 on the device (2-D Quad4)."""

    def __init__(self, ctx, p, corner_nodes, cells):
        self.ctx, self.p = ctx, p
        h = _P()
        corner_nodes = np.ascontiguousarray(corner_nodes, np.float64)
        cells = np.ascontiguousarray(cells, np.int32)
        check(lib.yafel_hip_dg_create(ctx.h, p, corner_nodes.shape[0], _ptr(corner_nodes), cells.shape[0],
                                      _ptr(cells), C.byref(h)), ctx.h)
        self.h = h
        info = np.zeros(4, np.int64)
        check(lib.yafel_hip_dg_info(h, _ptr(info)), ctx.h)
        self.ndofs, self.nfaces, self.nboundary, self.nen = (int(v) for v in info)
        na = _I64(0)
        check(lib.yafel_hip_dg_affine_cells(h, C.byref(na)), ctx.h)
        self.naffine = int(na.value)

    def close(self):
        if self.h:
            lib.yafel_hip_dg_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def nodes(self):
        d = self.ctx.alloc(self.ndofs * 3)
        check(lib.yafel_hip_dg_nodes(self.h, d.ptr), self.ctx.h)
        return d.download().reshape(-1, 3)

    def faces(self):
        out = [np.zeros(self.nfaces, np.int32) for _ in range(4)]
        check(lib.yafel_hip_dg_faces(self.h, *[_ptr(a) for a in out]), self.ctx.h)
        return tuple(out)

    def residual(self, ph, U_dev, R_dev, Rface_dev=None, time=0.0):
        check(lib.yafel_hip_dg_residual(self.h, C.byref(ph), time, U_dev.ptr, R_dev.ptr,
                                        None if Rface_dev is None else Rface_dev.ptr), self.ctx.h)

    def rk4(self, ph, U_dev, dt, nsteps, time=0.0):
        t = C.c_double(time)
        check(lib.yafel_hip_dg_rk4(self.h, C.byref(ph), dt, nsteps, C.byref(t), U_dev.ptr), self.ctx.h)
        return t.value


class Mesh:
    """DoFManager(M, CG, p, dpn) on the device."""

    def __init__(self, ctx, dim, p, dpn, corner_nodes, cells):
        self.ctx = ctx
        h = _P()
        corner_nodes = np.ascontiguousarray(corner_nodes, np.float64)
        cells = np.ascontiguousarray(cells, np.int32)
        check(lib.yafel_hip_mesh_create_cg(ctx.h, dim, p, dpn, corner_nodes.shape[0], _ptr(corner_nodes),
                                           cells.shape[0], _ptr(cells), C.byref(h)), ctx.h)
        self.h = h
        info = np.zeros(8, np.int64)
        check(lib.yafel_hip_mesh_info(h, _ptr(info)), ctx.h)
        self.n_nodes, self.nel, self.nen, self.ndofs, self.dim, self.p, self.dpn = (int(v) for v in info[:7])

    @classmethod
    def structured(cls, ctx, dim, p, dpn, n, nz=None):
        nodes, cells = structured_mesh(dim, n, n, (nz or n) if dim == 3 else 0, p)
        return cls(ctx, dim, p, dpn, nodes, cells)

    def close(self):
        if self.h:
            lib.yafel_hip_mesh_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def download(self):
        el = np.zeros((self.nel, self.nen), np.int32)
        dn = np.zeros((self.n_nodes, 3))
        check(lib.yafel_hip_mesh_download(self.h, _ptr(el), _ptr(dn)), self.ctx.h)
        return el, dn

    def pattern(self, fmt=CSR, precisions=FP64):
        h = _P()
        check(lib.yafel_hip_mesh_pattern(self.h, fmt, precisions, C.byref(h)), self.ctx.h)
        return Matrix(self.ctx, h)

    def assemble(self, ph, A=None, R_dev=None, req=TANGENT | RESIDUAL):
        check(lib.yafel_hip_assemble(self.h, C.byref(ph), req, None if A is None else A.h,
                                     None if R_dev is None else R_dev.ptr), self.ctx.h)

    def element_matrices(self, ph):
        nld = self.nen * self.dpn
        K = self.ctx.alloc(self.nel * nld * nld)
        check(lib.yafel_hip_element_matrices(self.h, C.byref(ph), K.ptr), self.ctx.h)
        return K.download().reshape(self.nel, nld, nld)

    def element_matrices_fp32(self, ph, K_dev=None):
        """FP32 matrix-core element matrices into K_dev (allocated when None)."""
        nld = self.nen * self.dpn
        K = K_dev if K_dev is not None else self.ctx.alloc(self.nel * nld * nld, np.float32)
        check(lib.yafel_hip_element_matrices_fp32(self.h, C.byref(ph), K.ptr), self.ctx.h)
        return K

    def bc_mask_plane(self, mask_dev, axis, value, tol=1e-6, comp=-1):
        check(lib.yafel_hip_bc_mask_plane(self.h, axis, value, tol, comp, mask_dev.ptr), self.ctx.h)

    def gradient_local_smoothing(self, u_dev, grad_dev=None):
        """LocalSmoothingGradient: (ndofs, dim) gradient per dof into grad_dev
        (allocated when None); returns the device array."""
        G = grad_dev if grad_dev is not None else self.ctx.alloc(self.ndofs * self.dim)
        check(lib.yafel_hip_gradient_local_smoothing(self.h, u_dev.ptr, G.ptr), self.ctx.h)
        return G

    def gradient_zz(self, u_dev, grad_dev=None):
        """ZZGradientRecovery (p = 1, scalar field): (n_nodes, dim) recovered
        nodal gradient into grad_dev (allocated when None)."""
        G = grad_dev if grad_dev is not None else self.ctx.alloc(self.n_nodes * self.dim)
        check(lib.yafel_hip_gradient_zz(self.h, u_dev.ptr, G.ptr), self.ctx.h)
        return G


BC_BOX, BC_XMIN = 0, 1
REQ_QUEUED, REQ_RUNNING, REQ_DONE, REQ_FAILED, REQ_CANCELLED = 0, 1, 2, 3, 4


class Server:
    """Device-resident request loop with continuous batching (yafel_hip_server_*)."""

    def __init__(self, ctx, hbm_budget, max_slots=64, setup_threads=0):
        self.ctx = ctx
        h = _P()
        check(lib.yafel_hip_server_create_ex(ctx.h, int(hbm_budget), max_slots, setup_threads, C.byref(h)), ctx.h)
        self.h = h
        self._ids = (_I64 * 4096)()

    def close(self):
        if self.h:
            lib.yafel_hip_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, dim, p, dpn, nodes, cells, ph, load=None, bc=BC_BOX, tol=1e-8, max_iter=0):
        nodes = np.ascontiguousarray(nodes, np.float64)
        cells = np.ascontiguousarray(cells, np.int32)
        q = Request()
        q.dim, q.p, q.dof_per_node = dim, p, dpn
        q.n_corner, q.corner_nodes = nodes.shape[0], nodes.ctypes.data
        q.ncell, q.cells = cells.shape[0], cells.ctypes.data
        q.phys = ph
        if load is not None:
            load = np.ascontiguousarray(load, np.float64)
            q.load, q.load_len = load.ctypes.data, load.size
        else:
            q.load, q.load_len = None, 0
        q.bc, q.tol, q.max_iter = bc, tol, max_iter
        rid = _I64()
        check(lib.yafel_hip_server_submit(self.h, C.byref(q), C.byref(rid)), self.ctx.h)
        return rid.value

    def progress(self):
        n = _I()
        check(lib.yafel_hip_server_progress(self.h, self._ids, len(self._ids), C.byref(n)), self.ctx.h)
        return [self._ids[i] for i in range(min(n.value, len(self._ids)))]

    def query(self, rid):
        st, nd = _I(), _I64()
        check(lib.yafel_hip_server_query(self.h, rid, C.byref(st), C.byref(nd)), self.ctx.h)
        return st.value, nd.value

    def result(self, rid):
        st, nd = self.query(rid)
        x = np.zeros(max(nd, 1))
        info = SolveInfo()
        check(lib.yafel_hip_server_result(self.h, rid, _ptr(x), x.size, C.byref(info)), self.ctx.h)
        return x[:nd], info

    def times(self, rid):
        """(admission wait, setup, wait to join, solve) in ms of a finished request."""
        ms = (C.c_double * 4)()
        check(lib.yafel_hip_server_times(self.h, rid, ms), self.ctx.h)
        return tuple(ms)

    def cancel(self, rid):
        check(lib.yafel_hip_server_cancel(self.h, rid), self.ctx.h)

    def set_cg_passes(self, passes, two_pass_limit=0):
        """Kernels per batched CG iteration with direct submission: 2, 3, or 0 (the
        default: two while the active slots store at most two_pass_limit scalars)."""
        check(lib.yafel_hip_server_set_cg_passes(self.h, int(passes), int(two_pass_limit)), self.ctx.h)

    def set_graphs(self, on):
        """Replay each chunk from a HIP graph (the slot table read from device memory)."""
        check(lib.yafel_hip_server_set_graphs(self.h, 1 if on else 0), self.ctx.h)

    def stats(self):
        s = ServerStats()
        check(lib.yafel_hip_server_stats_get(self.h, C.byref(s)), self.ctx.h)
        return s.as_dict()
