/*
 * CPU oracle for the YAFEL assembly -> sparse -> Krylov path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: tests/, the smoke
 * entry point and the CPU-baseline leg of bench.py may load it; the product
 * (libyafel_hip.so and the C++ headers under include/yafel_amd) never does.
 *
 * Every function restates the reference algorithm named in its comment
 * (paths are relative to the reference tree).  Plain C types only, so it can
 * be driven from ctypes.
 */
#ifndef YAFEL_ORACLE_H
#define YAFEL_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- quadrature (src/quadrature/QuadratureRule.cpp) ---------------------- */
void orc_gauss_legendre_1d(int npts, double *x, double *w);
void orc_gauss_lobatto_1d(int npts, double *x, double *w);
/* type 0 = Gauss-Legendre, 1 = Gauss-Lobatto; xi is npts*3; returns npts */
int orc_tensor_rule(int type, int topo_dim, int poly_order, double *xi, double *w);

/* ---- shape tables (src/element/ShapeFunctionUtils.cpp:64-143) ------------ */
/* nen = (p+1)^dim local nodes, nqp = (p+1)^dim Gauss points (multiplier 2).
 * N is nqp*nen, dN is nqp*nen*dim (row-major per qp), w is nqp, xi nqp*3. */
void orc_shape_tables(int dim, int p, double *N, double *dN, double *w, double *xi);
/* local GLL node positions (make_tensorproduct_element.cpp:158-193), nen*3 */
void orc_local_nodes(int dim, int p, double *xl);

/* ---- structured synthetic meshes --------------------------------------- */
/* Unit square/cube, nx*ny(*nz) cells, corner nodes numbered x-fastest.
 * cell_order 0 = lexicographic corners (what p=1 shape functions assume),
 * 1 = gmsh counter-clockwise corners (what p>=2 corner interpolation assumes). */
void orc_structured_mesh(int dim, int nx, int ny, int nz, int cell_order,
                         double *nodes /* (n_nodes)*3 */, int32_t *cells /* ncell*2^dim */);

/* ---- DoF numbering (src/utils/DoFManager.cpp:143-229, 296-414) ---------- */
/* Returns the number of unique dof nodes; elements is ncell*(p+1)^dim,
 * dof_nodes must hold ncell*(p+1)^dim*3 doubles (raw upper bound). */
int64_t orc_dofmanager_cg(int dim, int p, int64_t n_corner_nodes, const double *corner_nodes,
                          int64_t ncell, const int32_t *cells,
                          int32_t *elements, double *dof_nodes);

/* ---- physics descriptor ------------------------------------------------ */
typedef struct {
  int kind;             /* 0 Poisson, 1 isotropic elasticity, 2 SUPG advection-diffusion */
  int nsd;
  const double *coef;   /* per-element scale (k_e for Poisson / elasticity); NULL = 1 */
  double f;             /* Poisson source */
  double lambda, mu;    /* elasticity */
  double body[3];       /* elasticity body force */
  double U[3];          /* SUPG velocity */
  double kdiff, h, alpha, Q;  /* SUPG diffusivity, length, upwind factor, source */
} orc_physics;

/* ---- CG assembly (include/assembly/CGAssembly.hpp:27-186) --------------- */
/* Emits all nel*(nen*dpn)^2 triplets in element order (rows, cols, vals) and
 * accumulates R (length n_nodes*dpn, caller zeroes).  Any of rows/cols/vals/R
 * may be NULL to skip. */
void orc_cg_assembly(int dim, int p, int dpn, int64_t n_nodes, const double *dof_nodes,
                     int64_t nel, const int32_t *elements, const orc_physics *phys,
                     int64_t *rows, int64_t *cols, double *vals, double *R);
/* CGAssembly's OpenMP structure (per-thread triplets and residual, appended
 * under a critical section) on nthreads threads, for the timed CPU baseline;
 * rows/cols/vals and R as above (all required), triplets in append order.
 * Returns the triplet count. */
int64_t orc_cg_assembly_threads(int nthreads, int dim, int p, int dpn, int64_t n_nodes,
                                const double *dof_nodes, int64_t nel, const int32_t *elements,
                                const orc_physics *phys, int64_t *rows, int64_t *cols, double *vals,
                                double *R);
/* local element matrix only (nen*dpn)^2 row-major, for element e */
void orc_element_matrix(int dim, int p, int dpn, const double *dof_nodes,
                        const int32_t *elements, int64_t e, const orc_physics *phys,
                        double *Ke, double *Re);

/* ---- sparse_csr / sparse_bcsr from triplets ----------------------------- */
typedef struct orc_csr orc_csr;
/* old_handmade_linalg/sparse_csr.hpp:90-149 */
orc_csr *orc_csr_from_triplets(int64_t nt, const int64_t *r, const int64_t *c, const double *v);
/* old_handmade_linalg/sparse_bcsr.hpp:93-176 (BLOCK = 3) */
orc_csr *orc_bcsr3_from_triplets(int64_t nt, const int64_t *r, const int64_t *c, const double *v);
int64_t orc_csr_rows(const orc_csr *A);   /* scalar rows (bcsr: brows*3) */
int64_t orc_csr_cols(const orc_csr *A);
int64_t orc_csr_nnz(const orc_csr *A);    /* csr: stored scalars; bcsr: blocks */
int64_t orc_csr_nptr(const orc_csr *A);   /* length of row_ptr */
void orc_csr_copy(const orc_csr *A, int64_t *row_ptr, int64_t *col, double *val);
void orc_csr_free(orc_csr *A);
/* wraps caller arrays (copied) as a CSR (block=1) or BCSR3 (block=3) matrix */
orc_csr *orc_csr_wrap(int block, int64_t nrows, int64_t ncols, const int64_t *row_ptr,
                      const int64_t *col, const double *val);

/* ---- SpMV / vector ops -------------------------------------------------- */
void orc_spmv(const orc_csr *A, const double *x, double *y);  /* csr_spmv / bcsr_spmv<3> */
double orc_dot(int64_t n, const double *a, const double *b); /* VectorExpression::dot */

/* ---- Dirichlet (src/boundary_conditions/DirichletBC.cpp:27-56) ----------- */
void orc_dirichlet_apply(orc_csr *A, double *rhs, const uint8_t *mask, const double *g);

/* ---- Krylov solvers ---------------------------------------------------- */
/* cg_solve.hpp:33-81; returns iterations (k at exit) */
int64_t orc_cg_solve(const orc_csr *A, const double *b, const double *x0, double tol, double *x);
/* pcg_solve.hpp:126-178 with JacobiPreconditioner.hpp:124-153 */
int64_t orc_pcg_jacobi_solve(const orc_csr *A, const double *b, const double *x0, double tol,
                             double *x);
/* bicgstab_solve.hpp:27-95; max_iter guards the reference's unbounded loop (0 = none) */
int64_t orc_bicgstab_solve(const orc_csr *A, const double *b, const double *x0, double tol,
                           int64_t max_iter, double *x);
/* Jacobi-preconditioned BiCGSTAB (right preconditioning) with the true-residual
 * stopping rule ||b-Ax||/||b|| < rtol checked every iteration; used for the
 * multi-RHS advection-diffusion workload.  Returns iterations. */
int64_t orc_bicgstab_jacobi_solve(const orc_csr *A, const double *b, double rtol, int64_t max_iter,
                                  double *x);

/* ---- DG (include/assembly/DGAssembly.hpp:32-305): Quad4 cells, dpn = 1 ---- */
/* Mesh::buildInternalFaces (src/mesh/build_faces.cpp:25-96) for Quad4 cells;
 * returns the face count, the arrays hold up to 4*ncell entries */
int64_t orc_dg_faces(int64_t ncell, const int32_t *cells, int32_t *left, int32_t *right,
                     int32_t *left_flocal, int32_t *right_flocal);
/* DG numbering (DoFManager::make_raw_dofs, DoFManager.cpp:176-229): element e
 * owns dofs e*nen .. e*nen+nen-1, nen = (p+1)^2; dof_nodes is ncell*nen*3 */
void orc_dg_nodes(int p, const double *corner_nodes, int64_t ncell, const int32_t *cells,
                  double *dof_nodes);
/* DGAssembly<AdvectionPhysics> with {Residual, DtMass}
 * (apps/BasicAdvection_DG/main.cpp:14-98): velocity phys->U, source phys->f.
 * R (ncell*nen) is overwritten; R_face, when not NULL, receives the residual
 * after the face loop alone. */
void orc_dg_residual(int p, const double *corner_nodes, int64_t ncell, const int32_t *cells,
                     const orc_physics *phys, double t, const double *U, double *R, double *R_face);
/* nsteps of DGRK4::step (include/time_integration/DGRK4.hpp:33-69); U in place */
void orc_dg_rk4(int p, const double *corner_nodes, int64_t ncell, const int32_t *cells,
                const orc_physics *phys, double dt, int64_t nsteps, double t0, double *U);

/* ---- seeded synthetic inputs ------------------------------------------- */
uint64_t orc_splitmix64(uint64_t *state);
void orc_normal_fill(uint64_t seed, int64_t n, double mean, double sd, double *out);

/* ---- gradient recovery ------------------------------------------------- */
/* LocalSmoothingGradient.hpp:20-94: grad is (n_nodes*dpn) x dim, row-major. */
void orc_grad_local_smoothing(int dim, int p, int dpn, int64_t n_nodes, const double *dof_nodes, int64_t nel,
                              const int32_t *elements, const double *u, double *grad);
/* ZZGradientRecovery.hpp:33-141 (p = 1, dpn = 1): grad is n_nodes x dim. */
void orc_grad_zz(int dim, int64_t n_nodes, const double *dof_nodes, int64_t nel, const int32_t *elements,
                 const double *u, double *grad);

#ifdef __cplusplus
}
#endif
#endif
