/*
 * yafel_hip.h -- C ABI of the MI355X (gfx950) implementation of YAFEL's
 * assembly -> sparse -> Krylov hot path.
 *
 * Conventions
 *  - Every call returns an int status (YAFEL_HIP_OK = 0); yafel_hip_last_error()
 *    gives the message.  No C++ exception crosses this boundary.
 *  - Host arrays use int64 indices (the reference's size_t size_type,
 *    old_handmade_linalg/sparse_matrix.hpp:16); device storage uses int32.
 *  - Pointers named *_dev are device pointers; everything else is host memory.
 *  - A context owns one HIP stream; calls on one context are stream-ordered.
 *    Calls that take or return host data are blocking.  A context is meant for
 *    one host thread at a time.
 *
 * Each entry point names the reference interface it replaces (paths relative
 * to the YAFEL source tree).
 */
#ifndef YAFEL_HIP_H
#define YAFEL_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define YAFEL_HIP_OK 0
#define YAFEL_HIP_EINVAL 1     /* bad argument (sizes, null pointers, unsupported element) */
#define YAFEL_HIP_EDEVICE 2    /* HIP runtime error */
#define YAFEL_HIP_ENOMEM 3     /* device allocation failed / budget exceeded */
#define YAFEL_HIP_ESTATE 4     /* object used in the wrong state (e.g. matrix not from this mesh) */

/* precision flags for matrix copies and solver arithmetic */
#define YAFEL_HIP_FP64 1
#define YAFEL_HIP_FP32 2
#define YAFEL_HIP_MIXED 3      /* solver: FP32 inner iteration inside FP64 refinement */
#define YAFEL_HIP_MIXED_MATRIX 4 /* solver: FP32 matrix streams, FP64 vectors/scalars, FP64 refinement */

/* matrix formats */
#define YAFEL_HIP_CSR 1
#define YAFEL_HIP_BCSR3 3

typedef struct yafel_hip_ctx_s *yafel_hip_ctx;
typedef struct yafel_hip_matrix_s *yafel_hip_matrix;
typedef struct yafel_hip_mesh_s *yafel_hip_mesh;

/* ---------------------------------------------------------------- context */
/* No reference counterpart (the reference has no device state). */
int yafel_hip_ctx_create(int device, yafel_hip_ctx *out);
/* Drops the caller's reference.  Matrices, meshes and servers created on ctx
 * hold references of their own, so they stay usable and may be destroyed
 * after the context; its stream and memory go with the last reference. */
int yafel_hip_ctx_destroy(yafel_hip_ctx ctx);
/* last error message of ctx (or of the calling thread when ctx is NULL) */
const char *yafel_hip_last_error(yafel_hip_ctx ctx);
int yafel_hip_ctx_sync(yafel_hip_ctx ctx);
/* the context's hipStream_t */
void *yafel_hip_ctx_stream(yafel_hip_ctx ctx);
/* device memory helpers (stream-ordered on the context stream) */
int yafel_hip_malloc(yafel_hip_ctx ctx, size_t bytes, void **dptr);
int yafel_hip_free(yafel_hip_ctx ctx, void *dptr);
int yafel_hip_memcpy_h2d(yafel_hip_ctx ctx, void *dst_dev, const void *src, size_t bytes);
int yafel_hip_memcpy_d2h(yafel_hip_ctx ctx, void *dst, const void *src_dev, size_t bytes);
int yafel_hip_memcpy_d2d(yafel_hip_ctx ctx, void *dst_dev, const void *src_dev, size_t bytes);
/* zero `bytes` of device memory (the Vector(n) zero fill, old_handmade_linalg/
 * Vector.hpp:29-33, kept on the device) */
int yafel_hip_memzero(yafel_hip_ctx ctx, void *dst_dev, size_t bytes);
/* HIP-event timer on the context stream: begin, then end returns elapsed ms */
int yafel_hip_timer_begin(yafel_hip_ctx ctx);
int yafel_hip_timer_end(yafel_hip_ctx ctx, double *ms);

/* --------------------------------------------------------------- matrices */
typedef struct {
  int format;        /* YAFEL_HIP_CSR or YAFEL_HIP_BCSR3 */
  int precisions;    /* OR of YAFEL_HIP_FP64 / YAFEL_HIP_FP32 copies held on the device */
  int64_t rows;      /* scalar rows (BCSR: 3 * block rows) */
  int64_t cols;      /* scalar cols */
  int64_t nnz;       /* stored scalars (BCSR: 9 * blocks) */
  int64_t nblocks;   /* BCSR blocks (CSR: = nnz) */
  int64_t ptr_len;   /* length of the host row_ptr (rows + 1, or brows + 1) */
} yafel_hip_matrix_info;

/* sparse_csr<double> from its public arrays row_ptr/col_index/_data
 * (old_handmade_linalg/sparse_csr.hpp:63-65). */
int yafel_hip_csr_create(yafel_hip_ctx ctx, int64_t rows, int64_t cols, const int64_t *row_ptr,
                         const int64_t *col_index, const double *data, int precisions,
                         yafel_hip_matrix *out);
/* sparse_bcsr<3,double> from brow_ptr/bcol_index/data, 3x3 row-major blocks
 * (old_handmade_linalg/sparse_bcsr.hpp:31-33, 227-229). */
int yafel_hip_bcsr3_create(yafel_hip_ctx ctx, int64_t brows, int64_t bcols, const int64_t *brow_ptr,
                           const int64_t *bcol_index, const double *data, int precisions,
                           yafel_hip_matrix *out);
/* sparse_csr<double>(const std::vector<triplet>&) on the device: sort by
 * (row, col, value), sum duplicates in that order (sparse_csr.hpp:90-149). */
int yafel_hip_csr_from_triplets(yafel_hip_ctx ctx, int64_t nt, const int64_t *rows,
                                const int64_t *cols, const double *vals, int precisions,
                                yafel_hip_matrix *out);
/* sparse_bcsr<3,double>(triplets) on the device (sparse_bcsr.hpp:93-176). */
int yafel_hip_bcsr3_from_triplets(yafel_hip_ctx ctx, int64_t nt, const int64_t *rows,
                                  const int64_t *cols, const double *vals, int precisions,
                                  yafel_hip_matrix *out);
int yafel_hip_matrix_destroy(yafel_hip_matrix A);
int yafel_hip_matrix_info_get(yafel_hip_matrix A, yafel_hip_matrix_info *info);
/* copy back in the reference layout (row_ptr/col_index/_data, or BCSR with
 * 3x3 row-major blocks); any pointer may be NULL */
int yafel_hip_matrix_download(yafel_hip_matrix A, int64_t *row_ptr, int64_t *col_index,
                              double *data);
/* refresh the FP32 copy from the FP64 values (allocating it if needed) */
int yafel_hip_matrix_sync_fp32(yafel_hip_matrix A);
/* slots per step (1, 2, 4 or 8) of the sliced SpMV kernel behind
 * yafel_hip_spmv (the Krylov solvers' SpMV uses 8 when set, else 4); the
 * default is 8 for CSR matrices averaging 24 or more slots per slice and 4
 * otherwise.  A tuning knob: results are identical for every value. */
int yafel_hip_matrix_set_unroll(yafel_hip_matrix A, int unroll);
/* kernels per iteration of the single-solve CG / Jacobi-PCG on this matrix:
 * 3 (SpMV; alpha and r, z; beta, x and p) or 2 (p = z + beta p formed in the
 * SpMV's gather, with x += alpha p, in the same kernel); 0 (the default)
 * picks 2 for MIXED solves up to 16 Mi stored scalars and FP64-vector solves
 * on CSR up to 1.5 Mi, else 3.  A tuning knob: the iterates are
 * bit-identical. */
int yafel_hip_matrix_set_cg_passes(yafel_hip_matrix A, int passes);
/* bytes of each of the matrix's streams (the values of one precision with
 * their column indices) that the SpMV kernels read with the caching policy;
 * the rest is read non-temporally and does not allocate in the 256 MB
 * Infinity Cache.  -1 (the default) is 224 MiB: the leading part of a stream
 * larger than the cache stays resident across the products of a solve.  A
 * tuning knob: results are identical for every value. */
int yafel_hip_matrix_set_cache_bytes(yafel_hip_matrix A, int64_t bytes);
/* No reference counterpart: bytes per stored column index of the SpMV
 * streams, 2 when every entry's column is within +-32767 of its stored row
 * position (16-bit offsets, a banded mesh numbering) and 4 otherwise. */
int yafel_hip_matrix_index_bytes(yafel_hip_matrix A, int *bytes);
/* device pointer to the FP64 diagonal (computed on first use) */
int yafel_hip_matrix_diagonal(yafel_hip_matrix A, double *diag_dev);

/* ------------------------------------------------------------------- SpMV */
/* y = A x; precision selects the FP64 or FP32 copy; x_dev/y_dev are double*
 * or float* accordingly.  Replaces csr_spmv (old_handmade_linalg/csr_spmv.hpp:25-56)
 * and bcsr_spmv (old_handmade_linalg/bcsr_spmv.hpp:23-62). */
int yafel_hip_spmv(yafel_hip_matrix A, int precision, const void *x_dev, void *y_dev);
/* Vector<T> operator*(A, x) (old_handmade_linalg/operators.hpp:117-131): host in, host out. */
int yafel_hip_spmv_host(yafel_hip_matrix A, const double *x, double *y);
/* Y = A X for k right-hand sides stored row-major (n x k, the k values of a
 * row contiguous).  In FP64 each column of Y equals yafel_hip_spmv of that
 * column of X bit for bit (rows summed in column order, one FMA per entry).
 * No reference counterpart: multi-RHS form of csr_spmv. */
int yafel_hip_spmm(yafel_hip_matrix A, int precision, int k, const void *X_dev, void *Y_dev);

/* ------------------------------------------------------------ vector ops */
/* VectorExpression::dot (old_handmade_linalg/VectorExpression.hpp:31-55), FP64,
 * deterministic two-stage reduction. */
int yafel_hip_dot(yafel_hip_ctx ctx, int64_t n, const double *a_dev, const double *b_dev,
                  double *result);

/* ------------------------------------------------------------ measurement */
/* No reference counterpart: achievable HBM bandwidth of a plain copy kernel
 * (read + write bytes / time) over `bytes` per buffer, `reps` launches. */
int yafel_hip_copy_bandwidth(yafel_hip_ctx ctx, size_t bytes, int reps, double *gbs);
/* No reference counterpart: achievable HBM read bandwidth of a read-only
 * stream (16-byte non-temporal loads) over `bytes`, `reps` launches; the
 * ceiling for read-dominated kernels such as the SpMV. */
int yafel_hip_read_bandwidth(yafel_hip_ctx ctx, size_t bytes, int reps, double *gbs);

/* No reference counterpart: HIP-event timing of every launch of the CG
 * iteration's fused SpMV kernel (kernel A) on ctx, for roofline figures.
 * Each timed launch carries its start/stop event pair (hipExtLaunchKernel),
 * so an event pair spans that dispatch alone, and a work word that the launch
 * sets when it runs the product.  Enabling (or disabling) resets the record;
 * kernel_time returns the summed duration and the count of the launches that
 * ran the product -- launches that returned at once, after convergence or at
 * a refinement pause, are left out -- over the first 65,536 timed launches
 * (stream-synchronising). */
int yafel_hip_kernel_timing(yafel_hip_ctx ctx, int enable);
int yafel_hip_kernel_time(yafel_hip_ctx ctx, double *ms_total, int64_t *launches);

/* ---------------------------------------------------------------- solvers */
#define YAFEL_HIP_CG 0              /* cg_solve.hpp:33-81 */
#define YAFEL_HIP_PCG_JACOBI 1      /* pcg_solve.hpp:126-178 + JacobiPreconditioner.hpp */
#define YAFEL_HIP_BICGSTAB 2        /* bicgstab_solve.hpp:27-95 */
#define YAFEL_HIP_BICGSTAB_JACOBI 3 /* BiCGSTAB with right Jacobi preconditioning */

/* stopping rules */
#define YAFEL_HIP_STOP_REFERENCE 0  /* the reference's own test: squared ratio rr/rr0 (CG),
                                       rz/rz0 (PCG), ||Ax-b||^2/||b||^2 (BiCGSTAB) < tol */
#define YAFEL_HIP_STOP_RELRES 1     /* true relative residual ||b - A x|| / ||b|| < tol */

#define YAFEL_HIP_STATUS_CONVERGED 0
#define YAFEL_HIP_STATUS_MAXITER 1
#define YAFEL_HIP_STATUS_BREAKDOWN 2
#define YAFEL_HIP_STATUS_NAN 3

typedef struct {
  int method;          /* YAFEL_HIP_CG ... */
  int precision;       /* YAFEL_HIP_FP64, YAFEL_HIP_MIXED or YAFEL_HIP_MIXED_MATRIX */
  int stop;            /* YAFEL_HIP_STOP_REFERENCE or YAFEL_HIP_STOP_RELRES */
  double tol;
  int64_t max_iter;    /* 0: the reference's limit (n for CG/PCG, 10 n for BiCGSTAB) */
  int check_every;     /* iterations per host status poll (0: default 16) */
  double refine_delta; /* MIXED*: refine when the inner residual fell by this factor (0: 0.1) */
} yafel_hip_solve_opts;

typedef struct {
  int status;           /* YAFEL_HIP_STATUS_* */
  int64_t iterations;   /* Krylov iterations (inner iterations for MIXED*) */
  int64_t refinements;  /* FP64 residual replacements (MIXED*) */
  double rel_residual;  /* true ||b - A x|| / ||b|| in FP64 at exit */
  double criterion;     /* last value of the stopping quantity */
  double ms;            /* device time of the solve */
} yafel_hip_solve_info;

int yafel_hip_solve_opts_default(yafel_hip_solve_opts *opts);
/* Solve A x = b; x_dev holds x0 on entry.  b_dev/x_dev are FP64 device vectors. */
int yafel_hip_solve(yafel_hip_matrix A, const yafel_hip_solve_opts *opts, const double *b_dev,
                    double *x_dev, yafel_hip_solve_info *info);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (a tight serving loop, a harness that forbids hipMalloc): the same
 * solve as yafel_hip_solve, with the same kernels and the same results, whose
 * workspace (the Krylov vectors, the partial sums and the solver state) is a
 * file-scope device array of the library instead of memory it allocates.
 * The array holds YAFEL_HIP_NOALLOC_MAX_ROWS rows (about 85 MiB on each GPU
 * the library runs on, held until the process ends); a larger matrix is
 * refused with YAFEL_HIP_EINVAL.  The call allocates nothing: the matrix's
 * diagonal and, for the MIXED precisions, its FP32 copy must exist already
 * (yafel_hip_matrix_diagonal and yafel_hip_matrix_sync_fp32 once beforehand;
 * after an assembly they are refreshed in place).  Every call on one GPU
 * shares that GPU's array: calls that overlap on one GPU race. */
#define YAFEL_HIP_NOALLOC_MAX_ROWS 1114112
int yafel_hip_solve_noalloc(yafel_hip_matrix A, const yafel_hip_solve_opts *opts, const double *b_dev,
                            double *x_dev, yafel_hip_solve_info *info);
/* host-vector form of cg_solve / pcg_solve / bicgstab_solve: x0 may be NULL (zeros) */
int yafel_hip_solve_host(yafel_hip_matrix A, const yafel_hip_solve_opts *opts, const double *b,
                         const double *x0, double *x, yafel_hip_solve_info *info);
/* k right-hand sides solved together (B_dev, X_dev row-major n x k, FP64);
 * info is an array of k.  Method must be BICGSTAB(_JACOBI) or CG/PCG_JACOBI.
 * No reference counterpart: the reference solves one right-hand side per call. */
int yafel_hip_solve_multi(yafel_hip_matrix A, const yafel_hip_solve_opts *opts, int k,
                          const double *B_dev, double *X_dev, yafel_hip_solve_info *info);

/* --------------------------------------------------------- mesh + assembly */
/* DoFManager(M, CG, p, dof_per_node) (src/utils/DoFManager.cpp:11-36, 143-229,
 * 296-414) on the device.  corner_nodes is n_corner*3 (the reference's 3-D
 * coordinate<>), cells is ncell*2^dim corner ids: lexicographic (x fastest)
 * for p = 1, gmsh counter-clockwise for p >= 2 -- the orders the reference's
 * p = 1 pass-through and p >= 2 corner interpolation assume.  All cells are
 * Quad4 (dim 2) or Hex8 (dim 3). */
int yafel_hip_mesh_create_cg(yafel_hip_ctx ctx, int dim, int p, int dof_per_node,
                             int64_t n_corner, const double *corner_nodes, int64_t ncell,
                             const int32_t *cells, yafel_hip_mesh *out);
int yafel_hip_mesh_destroy(yafel_hip_mesh m);
/* info[0]=dof nodes, [1]=elements, [2]=nodes per element, [3]=dofs, [4]=dim, [5]=p, [6]=dpn */
int yafel_hip_mesh_info(yafel_hip_mesh m, int64_t *info);
/* DoFManager::elements and ::dof_nodes (n_nodes*3) back to the host */
int yafel_hip_mesh_download(yafel_hip_mesh m, int32_t *elements, double *dof_nodes);
/* device pointer to dof_nodes (n_nodes*3 doubles) */
int yafel_hip_mesh_nodes_dev(yafel_hip_mesh m, const double **nodes_dev);

/* The global tangent's sparsity pattern (the closure of element connectivity,
 * explicit zeros kept, as sparse_csr(triplets) of CGAssembly's triplets yields;
 * include/assembly/CGAssembly.hpp:149-156 + sparse_csr.hpp:90-149) with zero
 * values, plus the deterministic gather plan used by yafel_hip_assemble. */
int yafel_hip_mesh_pattern(yafel_hip_mesh m, int format, int precisions, yafel_hip_matrix *out);

#define YAFEL_HIP_PHYS_POISSON 0     /* apps/BasicPoisson/main.cpp:25-61 */
#define YAFEL_HIP_PHYS_ELASTICITY 1  /* apps/BasicLinearElasticity/main.cpp:41-113 */
#define YAFEL_HIP_PHYS_SUPG 2        /* apps/supg_learning/supg_advection.hpp:19-132 */
#define YAFEL_HIP_PHYS_DG_ADVECTION 3 /* apps/BasicAdvection_DG/main.cpp:14-98 (velocity U, source f) */

/* The Physics template parameter of CGAssembly as a descriptor (a GPU cannot
 * call a host functor). */
typedef struct {
  int kind;              /* YAFEL_HIP_PHYS_* */
  int nsd;               /* must equal the mesh dim */
  const double *coef;    /* per-element scale of the tangent (k_e, or E_e for elasticity); NULL = 1 */
  int coef_on_device;    /* coef is a device pointer */
  double f;              /* Poisson source */
  double lambda, mu;     /* Lame parameters */
  double body[3];        /* elasticity body force */
  double U[3];           /* SUPG velocity */
  double kdiff, h, alpha, Q; /* SUPG diffusivity, element size, upwinding, source */
} yafel_hip_physics;

#define YAFEL_HIP_ASSEMBLE_TANGENT 1  /* AssemblyRequirement::Tangent */
#define YAFEL_HIP_ASSEMBLE_RESIDUAL 2 /* AssemblyRequirement::Residual */
/* No reference counterpart: element tangents computed in FP32 on the matrix
 * cores (G^T G as 32x32x2 FP32 MFMA steps) and gathered in FP64 -- values
 * carry FP32 rounding (about 1e-7 relative).  2-D Poisson with p = 3 or 4. */
#define YAFEL_HIP_ASSEMBLE_FP32_ELEMENTS 4

/* CGAssembly<Physics>(FESystem&, reqs) (include/assembly/CGAssembly.hpp:27-186):
 * overwrites A's values (A from yafel_hip_mesh_pattern of this mesh, FP64
 * values; the FP32 copy is not touched) and ACCUMULATES into R_dev (FP64,
 * n_dofs), as the reference accumulates GlobalResidual. */
int yafel_hip_assemble(yafel_hip_mesh m, const yafel_hip_physics *phys, int requirements,
                       yafel_hip_matrix A, double *R_dev);
/* element matrices only: Ke_dev holds nel*(nen*dpn)^2 doubles, element-major,
 * row-major (the local_tangent buffer of CGAssembly.hpp:117-118). */
int yafel_hip_element_matrices(yafel_hip_mesh m, const yafel_hip_physics *phys, double *Ke_dev);
/* the FP32 matrix-core element matrices of YAFEL_HIP_ASSEMBLE_FP32_ELEMENTS
 * (same layout, float; coefficients, if any, on the device); stream-ordered */
int yafel_hip_element_matrices_fp32(yafel_hip_mesh m, const yafel_hip_physics *phys, float *Ke_dev);

/* DirichletBC::apply (src/boundary_conditions/DirichletBC.cpp:27-56):
 * rhs -= A g; entries in BC rows or columns zeroed, diagonal set to 1,
 * rhs(bc) = g(bc); the pattern is kept.  mask_dev is one byte per dof,
 * g_dev may be NULL for homogeneous values. */
int yafel_hip_dirichlet_apply(yafel_hip_matrix A, const uint8_t *mask_dev, const double *g_dev,
                              double *rhs_dev);
/* DirichletBC::selectByFunction for axis-aligned planes: marks component comp
 * (or all components when comp < 0) of every node with |x_axis - value| < tol. */
int yafel_hip_bc_mask_plane(yafel_hip_mesh m, int axis, double value, double tol, int comp,
                            uint8_t *mask_dev);

/* ------------------------------------------------------------ mesh input */
/* Mesh(const std::string &fname) -> Mesh::parse_gmsh (src/mesh/Mesh.cpp:32-36,
 * src/mesh/parse_gmsh.cpp:80-151): a gmsh 2.2 ASCII file.  Node id i is stored
 * at i-1 (3 coordinates); cell id c at c-1 with its gmsh tags and 0-based
 * nodes at cell_offsets[c-1] (running offset in file order, the last entry
 * the total); cell types are the reference's CellType
 * (include/mesh/mesh_typedefs.hpp:42-50), an unknown gmsh type giving NONE.
 * Host only; no context needed. */
#define YAFEL_HIP_CELL_LINE2 0
#define YAFEL_HIP_CELL_TRI3 1
#define YAFEL_HIP_CELL_TET4 2
#define YAFEL_HIP_CELL_QUAD4 3
#define YAFEL_HIP_CELL_HEX8 4
#define YAFEL_HIP_CELL_POINT1 5
#define YAFEL_HIP_CELL_NONE 6
typedef struct yafel_hip_gmsh_s *yafel_hip_gmsh;
int yafel_hip_gmsh_read(const char *path, yafel_hip_gmsh *out);
/* sizes[0..3] = nodes, cells, cell node entries, tags */
int yafel_hip_gmsh_sizes(yafel_hip_gmsh g, int64_t *sizes);
/* any pointer may be NULL; cell_offsets and tag_offsets hold cells+1 entries */
int yafel_hip_gmsh_arrays(yafel_hip_gmsh g, double *nodes, int32_t *cell_nodes, int32_t *cell_offsets,
                          int32_t *cell_types, int32_t *tag_offsets, int32_t *tags);
int yafel_hip_gmsh_destroy(yafel_hip_gmsh g);

/* ---------------------------------------------------------------- output */
/* OutputData (include/output/OutputData.hpp:19-57): a named array of
 * n_locations * dof_per_loc values, the bits of `mask` standing for the
 * reference's dof_mask entries (the components written). */
typedef struct {
  const char *name;
  const double *data;  /* host pointer, or a device pointer when on_device */
  int on_device;
  int dof_per_loc;     /* the dof_mask length, 1..32 */
  uint32_t mask;       /* bit c set: component c written */
  int data_type;       /* 0 scalar, 1 vector, 2 tensor (OutputData::DataType) */
  int location;        /* 0 point (n_points locations), 1 cell (n_cells) */
} yafel_hip_output_data;
/* SimulationOutput(fname_base, BackendType::VTU).captureFrame: VTUBackend
 * (src/output/VTUBackend.cpp:19-230) writes fname_base.vtu with the OutputMesh
 * of the cells (src/output/OutputMesh.cpp:11-48): each element of order p,
 * given by its (p+1)^dim nodes in the DoFManager's local order, is drawn as
 * p^dim quads / hexes.  The text is the reference's, number formatting
 * included.  Scalars take one component; vectors and tensors up to three,
 * padded with zeros to three (as the reference writes both); cell data has
 * one value per cell.  Host side; device data is copied back. */
int yafel_hip_vtu_write(const char *fname_base, int dim, int p, int64_t n_points, const double *points,
                        int64_t n_cells, const int32_t *cells, int n_data, const yafel_hip_output_data *data);

/* ------------------------------------------------------ gradient recovery */
/* LocalSmoothingGradient<Physics>(FESystem&) (include/assembly/
 * LocalSmoothingGradient.hpp:20-94): grad_dev (n_dofs x dim, row-major, the
 * reference's RowMajor SolutionGradient) gets, for each dof, the volume-
 * weighted average over its node's elements and quadrature points of the
 * gradient of that dof's component: sum N_a jxw grad u_c / sum N_a jxw.
 * u_dev is the solution (n_dofs).  Stream-ordered on the mesh's context
 * stream; the per-element staging is kept with the mesh. */
int yafel_hip_gradient_local_smoothing(yafel_hip_mesh m, const double *u_dev, double *grad_dev);
/* ZZGradientRecovery<Physics>(FESystem&) (include/assembly/ZZGradientRecovery.hpp:
 * 33-141), p = 1 and one dof per node: at every node, the least-squares
 * linear fit of the gradient samples at the quadrature points of the adjacent
 * elements, evaluated at the node; grad_dev is n_nodes x dim, row-major.
 * Stream-ordered, like the smoothing. */
int yafel_hip_gradient_zz(yafel_hip_mesh m, const double *u_dev, double *grad_dev);

/* -------------------------------------------------------------------- DG */
/* DoFManager(M, DG, p, 1) + Mesh::buildInternalFaces (src/utils/DoFManager.cpp:
 * 166-293, src/mesh/build_faces.cpp:25-96) for 2-D Quad4 meshes with gmsh
 * (counter-clockwise) corners and p = 1..4.  Dof d = e * (p+1)^2 + a belongs
 * to element e alone.  The element mass inverses (the reference's cached
 * PartialPivLU, DGAssembly.hpp:274-277) are built here, once. */
typedef struct yafel_hip_dg_s *yafel_hip_dg;
int yafel_hip_dg_create(yafel_hip_ctx ctx, int p, int64_t n_corner, const double *corner_nodes,
                        int64_t ncell, const int32_t *cells, yafel_hip_dg *out);
int yafel_hip_dg_destroy(yafel_hip_dg dg);
/* info[0] dofs, info[1] faces, info[2] boundary faces, info[3] dofs per element */
int yafel_hip_dg_info(yafel_hip_dg dg, int64_t *info);
/* No reference counterpart: the number of parallelogram cells.  Their mass
 * matrix is det J times the reference one, so the residual applies the shared
 * reference inverse divided by det J (the same operator as the reference's
 * per-cell PartialPivLU inverse, to round-off) and reads one 64-byte record
 * instead of the cell's inverse and node coordinates. */
int yafel_hip_dg_affine_cells(yafel_hip_dg dg, int64_t *count);
/* dof node coordinates (dofs * 3 doubles, device) */
int yafel_hip_dg_nodes(yafel_hip_dg dg, double *nodes_dev);
/* the face list in the reference's order (host arrays of `faces` entries; -1 = none) */
int yafel_hip_dg_faces(yafel_hip_dg dg, int32_t *left, int32_t *right, int32_t *left_flocal,
                       int32_t *right_flocal);
/* DGAssembly<Physics>(FESystem&, P, {Residual, DtMass}) (include/assembly/
 * DGAssembly.hpp:32-305) with the advection physics of apps/BasicAdvection_DG/
 * main.cpp:14-98 (phys->kind YAFEL_HIP_PHYS_DG_ADVECTION, nsd 2, constant
 * velocity phys->U, source phys->f): R_dev = R_face + M^-1 (R_face + R_vol),
 * the reference's result; Rface_dev (may be NULL) receives R_face alone. */
int yafel_hip_dg_residual(yafel_hip_dg dg, const yafel_hip_physics *phys, double time,
                          const double *U_dev, double *R_dev, double *Rface_dev);
/* nsteps of DGRK4::step (include/time_integration/DGRK4.hpp:33-69) on U_dev in
 * place; *time advances by nsteps * dt.  A step is captured once into a HIP
 * graph (per dt, physics and U_dev) and replayed. */
int yafel_hip_dg_rk4(yafel_hip_dg dg, const yafel_hip_physics *phys, double dt, int64_t nsteps,
                     double *time, double *U_dev);

/* ---------------------------------------------------------------- serving */
/* No reference counterpart: a device-resident request loop.  Requests are
 * admitted under an HBM budget and a slot limit; every admitted request gets
 * its own DoF numbering, pattern, assembled FP64/FP32 matrix and Krylov state
 * on the device, and all admitted requests advance together in batched
 * Jacobi-PCG launches (FP32 inner iterations, FP64 refinement), each on its
 * own block range with its own scalars, so a result does not depend on what
 * else is in the batch.  Requests that arrive while others run join at the
 * next chunk boundary; finished ones retire and free their memory. */
typedef struct yafel_hip_server_s *yafel_hip_server;

#define YAFEL_HIP_BC_BOX 0      /* homogeneous Dirichlet on every face of the mesh's bounding box */
#define YAFEL_HIP_BC_XMIN 1     /* clamp every component on the x = min face */

typedef struct {
  int dim, p, dof_per_node;
  int64_t n_corner;
  const double *corner_nodes;   /* n_corner*3, host */
  int64_t ncell;
  const int32_t *cells;         /* ncell*2^dim, host, order as yafel_hip_mesh_create_cg */
  yafel_hip_physics phys;       /* coef (per element) must be a host pointer */
  const double *load;           /* right-hand side per dof (host); NULL: the assembled residual */
  int64_t load_len;             /* length of load (must equal the request's dof count) */
  int bc;                       /* YAFEL_HIP_BC_* */
  double tol;                   /* ||b - A x|| / ||b|| target */
  int64_t max_iter;             /* 0: number of dofs */
} yafel_hip_request;

#define YAFEL_HIP_REQ_QUEUED 0
#define YAFEL_HIP_REQ_RUNNING 1
#define YAFEL_HIP_REQ_DONE 2
#define YAFEL_HIP_REQ_FAILED 3
#define YAFEL_HIP_REQ_CANCELLED 4

typedef struct {
  int64_t submitted, admitted, completed, failed, cancelled, queued, active;
  int64_t chunks, batched_launches;
  size_t hbm_in_use, hbm_peak, hbm_budget;
  double setup_ms;              /* host time the setup worker spent building requests */
  int64_t graph_captures;       /* chunk graphs captured (graph replay on) */
  int64_t graph_replays;        /* chunks replayed from a graph */
  int64_t two_pass_chunks;      /* directly submitted chunks of two-pass CG iterations */
  int64_t three_pass_chunks;    /* ... and of three-pass iterations */
} yafel_hip_server_stats;

int yafel_hip_server_create(yafel_hip_ctx ctx, size_t hbm_budget, int max_slots, yafel_hip_server *out);
/* as yafel_hip_server_create, with the number of setup worker threads (each
 * with its own setup stream; 0: the default of 3).  Admission charges each
 * request's estimated bytes to hbm_budget; on top of that the workers keep
 * released set-up buffers for reuse, together at most hbm_budget / 8 (and at
 * most 4 GiB per worker), which the hbm_* statistics do not count. */
int yafel_hip_server_create_ex(yafel_hip_ctx ctx, size_t hbm_budget, int max_slots, int setup_threads,
                               yafel_hip_server *out);
int yafel_hip_server_destroy(yafel_hip_server s);
/* copies the request's inputs; YAFEL_HIP_ENOMEM when it could never fit the budget */
int yafel_hip_server_submit(yafel_hip_server s, const yafel_hip_request *req, int64_t *id);
/* admits what fits, runs one chunk of batched iterations for every active
 * request, retires finished ones; ids of requests that finished are written
 * to done_ids (up to max_ids), their count to n_done */
int yafel_hip_server_progress(yafel_hip_server s, int64_t *done_ids, int max_ids, int *n_done);
/* state (YAFEL_HIP_REQ_*) and, for finished requests, the dof count */
int yafel_hip_server_query(yafel_hip_server s, int64_t id, int *state, int64_t *ndofs);
/* copies a finished request's solution (DoFManager numbering) and releases it;
 * for a failed or cancelled request returns YAFEL_HIP_ESTATE (the message says
 * which) and releases it too, so its id is unknown afterwards */
int yafel_hip_server_result(yafel_hip_server s, int64_t id, double *x, int64_t n, yafel_hip_solve_info *info);
/* where a finished request's time went, in ms (host clock, before
 * yafel_hip_server_result releases it): ms[0] queued until a setup worker
 * took it (admission wait), ms[1] its setup (numbering, pattern, assembly,
 * Dirichlet), ms[2] set up until its first batched chunk was launched,
 * ms[3] from then until a progress call found it finished */
int yafel_hip_server_times(yafel_hip_server s, int64_t id, double *ms);
/* cancels a queued request: it leaves the queue and its inputs are freed at
 * once; yafel_hip_server_query reports YAFEL_HIP_REQ_CANCELLED until
 * yafel_hip_server_result releases it */
int yafel_hip_server_cancel(yafel_hip_server s, int64_t id);
int yafel_hip_server_stats_get(yafel_hip_server s, yafel_hip_server_stats *st);
/* on != 0: replay each chunk from a HIP graph whose kernels read the slot
 * table and its totals from device memory (uploaded when a request joins or
 * retires; a graph is captured again only when the totals outgrow the grids
 * it was captured with), one graph per chunk form and in-flight status slot;
 * 0 (the default): submit every launch directly.  Results are bit-identical
 * either way.  No reference counterpart (the reference has no request
 * server). */
int yafel_hip_server_set_graphs(yafel_hip_server s, int on);
/* kernels per batched CG iteration with direct submission: 3 (SpMV; alpha
 * and r, z; beta, x and p), 2 (after a chunk's first iteration the SpMV also
 * folds the update partials, takes beta and the stopping test, applies
 * x += alpha p and forms p = z + beta p in its gather: a launch fewer, but
 * two vectors gathered) or 0 (the default: a chunk takes two passes while its
 * active slots store at most two_pass_limit scalars in all, three otherwise;
 * two_pass_limit <= 0 keeps the default of 4 Mi), also with graph replay.  The
 * iterates are bit-identical whatever the choice.  No reference counterpart
 * (the reference has no request server). */
int yafel_hip_server_set_cg_passes(yafel_hip_server s, int passes, int64_t two_pass_limit);

/* ----------------------------------------------------- synthetic inputs */
/* Structured unit-square/cube corner mesh (cells ordered as
 * yafel_hip_mesh_create_cg expects for p).  No reference counterpart: the
 * reference reads gmsh files; this builds the synthetic workloads. */
int yafel_hip_structured_mesh(int dim, int nx, int ny, int nz, int p, double *corner_nodes,
                              int32_t *cells);
/* splitmix64 + Box-Muller normal deviates (host) */
int yafel_hip_normal_fill(uint64_t seed, int64_t n, double mean, double sd, double *out);

#ifdef __cplusplus
}
#endif
#endif
