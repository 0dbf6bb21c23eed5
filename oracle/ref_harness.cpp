/*
 * C-ABI harness around the reference's own Eigen-free sources (the handmade
 * sparse/Krylov headers under include/old_handmade_linalg and
 * src/quadrature/QuadratureRule.cpp).  TEST INFRASTRUCTURE ONLY: compiled by
 * oracle/build_ref.sh straight from the reference tree into oracle/_ref/, used
 * to pin the oracle, to write golden fixtures and as the CPU baseline.  No
 * reference source is copied; this file only calls the reference templates.
 */
#include "yafel_globals.hpp"
#include "old_handmade_linalg/Vector.hpp"
#include "old_handmade_linalg/sparse_coo.hpp"
#include "old_handmade_linalg/sparse_csr.hpp"
#include "old_handmade_linalg/sparse_bcsr.hpp"
#include "old_handmade_linalg/operators.hpp"
#include "old_handmade_linalg/solver/iterative/cg_solve.hpp"
#include "old_handmade_linalg/solver/iterative/pcg_solve.hpp"
#include "old_handmade_linalg/solver/iterative/bicgstab_solve.hpp"
#include "old_handmade_linalg/solver/iterative/JacobiPreconditioner.hpp"
#include "quadrature/QuadratureRule.hpp"

#include <cstdint>
#include <cstdlib>
#include <vector>

using namespace yafel;

namespace {
using CSR = sparse_csr<double>;
using BCSR = sparse_bcsr<3, double>;
struct Handle {
  int block;
  CSR *csr = nullptr;
  BCSR *bcsr = nullptr;
};
std::vector<std::tuple<std::size_t, std::size_t, double>> triplets(int64_t nt, const int64_t *r,
                                                                   const int64_t *c, const double *v) {
  std::vector<std::tuple<std::size_t, std::size_t, double>> ts(nt);
  for (int64_t i = 0; i < nt; ++i) ts[i] = std::make_tuple((std::size_t)r[i], (std::size_t)c[i], v[i]);
  return ts;
}
Vector<double> vec(int64_t n, const double *x) {
  Vector<double> v(n, 0.0);
  for (int64_t i = 0; i < n; ++i) v(i) = x[i];
  return v;
}
void out(const Vector<double> &v, double *y) {
  for (std::size_t i = 0; i < v.size(); ++i) y[i] = v(i);
}
}  // namespace

extern "C" {

void ref_quadrature_1d(int type, int npts, double *x, double *w) {
  QuadratureRule q(npts, type == 0 ? QuadratureRule::QuadratureType::GAUSS_LEGENDRE
                                   : QuadratureRule::QuadratureType::GAUSS_LOBATTO);
  for (int i = 0; i < (int)q.weights.size(); ++i) {
    x[i] = q.nodes[i](0);
    w[i] = q.weights[i];
  }
}

int ref_tensor_rule(int type, int dim, int order, double *xi, double *w) {
  auto q = QuadratureRule::make_tensor_product(
      type == 0 ? QuadratureRule::QuadratureType::GAUSS_LEGENDRE : QuadratureRule::QuadratureType::GAUSS_LOBATTO,
      dim, order);
  for (int i = 0; i < (int)q.weights.size(); ++i) {
    for (int d = 0; d < 3; ++d) xi[3 * i + d] = q.nodes[i](d);
    w[i] = q.weights[i];
  }
  return (int)q.weights.size();
}

void *ref_csr_from_triplets(int64_t nt, const int64_t *r, const int64_t *c, const double *v) {
  auto *h = new Handle{1};
  h->csr = new CSR(triplets(nt, r, c, v));
  return h;
}

void *ref_bcsr3_from_triplets(int64_t nt, const int64_t *r, const int64_t *c, const double *v) {
  auto *h = new Handle{3};
  h->bcsr = new BCSR(triplets(nt, r, c, v));
  return h;
}

/* info[0]=rows, info[1]=cols, info[2]=nnz (csr scalars / bcsr blocks), info[3]=len(row_ptr) */
void ref_matrix_info(void *hp, int64_t *info) {
  auto *h = (Handle *)hp;
  if (h->block == 1) {
    info[0] = h->csr->rows(); info[1] = h->csr->cols();
    info[2] = h->csr->col_index.size(); info[3] = h->csr->row_ptr.size();
  } else {
    info[0] = h->bcsr->rows(); info[1] = h->bcsr->cols();
    info[2] = h->bcsr->bcol_index.size(); info[3] = h->bcsr->brow_ptr.size();
  }
}

void ref_matrix_copy(void *hp, int64_t *ptr, int64_t *col, double *val) {
  auto *h = (Handle *)hp;
  if (h->block == 1) {
    for (std::size_t i = 0; i < h->csr->row_ptr.size(); ++i) ptr[i] = h->csr->row_ptr[i];
    for (std::size_t i = 0; i < h->csr->col_index.size(); ++i) col[i] = h->csr->col_index[i];
    for (std::size_t i = 0; i < h->csr->_data.size(); ++i) val[i] = h->csr->_data[i];
  } else {
    for (std::size_t i = 0; i < h->bcsr->brow_ptr.size(); ++i) ptr[i] = h->bcsr->brow_ptr[i];
    for (std::size_t i = 0; i < h->bcsr->bcol_index.size(); ++i) col[i] = h->bcsr->bcol_index[i];
    for (std::size_t i = 0; i < h->bcsr->data.size(); ++i) val[i] = h->bcsr->data[i];
  }
}

void ref_matrix_free(void *hp) {
  auto *h = (Handle *)hp;
  delete h->csr;
  delete h->bcsr;
  delete h;
}

static int64_t nrows(Handle *h) { return h->block == 1 ? h->csr->rows() : h->bcsr->rows(); }

void ref_spmv(void *hp, const double *x, double *y) {
  auto *h = (Handle *)hp;
  auto xv = vec(h->block == 1 ? h->csr->cols() : h->bcsr->cols(), x);
  if (h->block == 1) out((*h->csr) * xv, y);
  else out((*h->bcsr) * xv, y);
}

double ref_dot(int64_t n, const double *a, const double *b) { return vec(n, a).dot(vec(n, b)); }

void ref_cg_solve(void *hp, const double *b, const double *x0, double tol, double *x) {
  auto *h = (Handle *)hp;
  int64_t n = nrows(h);
  if (h->block == 1) out(cg_solve(*h->csr, vec(n, b), vec(n, x0), tol), x);
  else out(cg_solve(*h->bcsr, vec(n, b), vec(n, x0), tol), x);
}

void ref_pcg_jacobi_solve(void *hp, const double *b, const double *x0, double tol, double *x) {
  auto *h = (Handle *)hp;
  int64_t n = nrows(h);
  JacobiPreconditioner<double> M(*h->csr);
  out(pcg_solve(*h->csr, vec(n, b), M, vec(n, x0), tol), x);
}

void ref_bicgstab_solve(void *hp, const double *b, const double *x0, double *x) {
  auto *h = (Handle *)hp;
  int64_t n = nrows(h);
  if (h->block == 1) out(bicgstab_solve(*h->csr, vec(n, b), vec(n, x0)), x);
  else out(bicgstab_solve(*h->bcsr, vec(n, b), vec(n, x0)), x);
}

/* A sparse_bcsr<3> / sparse_csr whose public arrays are filled from caller
 * arrays (the dimension bookkeeping comes from two zero triplets at the
 * corners), for timing the reference kernels on a matrix assembled elsewhere. */
void *ref_matrix_from_arrays(int block, int64_t nrows, const int64_t *ptr, const int64_t *col,
                             const double *val) {
  std::vector<std::tuple<std::size_t, std::size_t, double>> ts = {
      std::make_tuple((std::size_t)0, (std::size_t)0, 0.0),
      std::make_tuple((std::size_t)(nrows - 1), (std::size_t)(nrows - 1), 0.0)};
  auto *h = new Handle{block};
  const int64_t nr = block == 1 ? nrows : nrows / 3;
  const int64_t nnz = ptr[nr];
  if (block == 1) {
    h->csr = new CSR(ts);
    h->csr->row_ptr.assign(ptr, ptr + nr + 1);
    h->csr->col_index.assign(col, col + nnz);
    h->csr->_data.assign(val, val + nnz);
  } else {
    h->bcsr = new BCSR(ts);
    h->bcsr->brow_ptr.assign(ptr, ptr + nr + 1);
    h->bcsr->bcol_index.assign(col, col + nnz);
    h->bcsr->data.assign(val, val + nnz * 9);
  }
  return h;
}

/* k iterations of cg_solve's loop body (cg_solve.hpp:56-77) with the
 * reference's own Vector / operator* / dot / expression templates, for a
 * bounded CPU timing sample (cg_solve itself only stops on its tolerance). */
double ref_cg_iterations(void *hp, const double *b, int k) {
  auto *h = (Handle *)hp;
  int64_t n = nrows(h);
  Vector<double> x(n, 0.0);
  Vector<double> r = vec(n, b);
  Vector<double> p(r);
  double rr_old = r.dot(r);
  for (int it = 0; it < k; ++it) {
    Vector<double> Ap = h->block == 1 ? (*h->csr) * p : (*h->bcsr) * p;
    double alpha = rr_old / p.dot(Ap);
    x += p * alpha;
    r += -alpha * Ap;
    double rr_new = r.dot(r);
    double beta = rr_new / rr_old;
    rr_old = rr_new;
    p *= beta;
    p += r;
  }
  return rr_old;
}

/* glibc rand() stream after srand(seed), as the reference's tests draw it
 * (test/test_cg_solve.cpp:16-27: rhs(i) = rand()%100 after srand(0)). */
void ref_rand_mod(unsigned seed, int n, int mod, double *out_v) {
  srand(seed);
  for (int i = 0; i < n; ++i) out_v[i] = double(rand() % mod);
}

}  // extern "C"
