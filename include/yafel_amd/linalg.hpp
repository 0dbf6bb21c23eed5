// YAFEL's sparse / Krylov interface (old_handmade_linalg/*) over the HIP C ABI.
//
// Same names, argument meaning and results as the reference headers:
//   sparse_csr<double>(triplets)        old_handmade_linalg/sparse_csr.hpp:31, 90-149
//   sparse_bcsr<3,double>(triplets)     old_handmade_linalg/sparse_bcsr.hpp:93-176
//   Vector<double> operator*(A, x)      old_handmade_linalg/operators.hpp:117-131
//   cg_solve / pcg_solve / bicgstab_solve  old_handmade_linalg/solver/iterative/*.hpp
//   JacobiPreconditioner<double>        old_handmade_linalg/solver/iterative/JacobiPreconditioner.hpp
//   LinearSolve::solve(A, b, Tag)       lin_alg/linear_solvers/LinearSolve.hpp:36-67
// The matrices live on the device; the public row_ptr / col_index / _data
// (brow_ptr / bcol_index / data) members are host copies kept for code that
// reads them, as the reference exposes them.  Every product and solve runs on
// the GPU; errors from the C ABI are thrown as std::runtime_error.
#pragma once

#include <cmath>
#include <cstddef>
#include <memory>
#include <stdexcept>
#include <string>
#include <tuple>
#include <type_traits>
#include <vector>

#include "yafel_hip.h"

namespace yafel {

namespace hipdetail {
inline void check(int rc, yafel_hip_ctx ctx = nullptr) {
  if (rc != YAFEL_HIP_OK) throw std::runtime_error(std::string("yafel_hip: ") + yafel_hip_last_error(ctx));
}
// one device context per process (device 0), created on first use
inline yafel_hip_ctx context() {
  struct Holder {
    yafel_hip_ctx c = nullptr;
    Holder() { check(yafel_hip_ctx_create(0, &c)); }
    ~Holder() { yafel_hip_ctx_destroy(c); }
  };
  static Holder h;
  return h.c;
}
struct MatrixDeleter {
  void operator()(yafel_hip_matrix_s *m) const { yafel_hip_matrix_destroy(m); }
};
using MatrixPtr = std::shared_ptr<yafel_hip_matrix_s>;
}  // namespace hipdetail

// ---------------------------------------------------------------- Vector
template <typename dataType = double>
class Vector {
 public:
  using container_type = std::vector<dataType>;
  using value_type = dataType;
  using size_type = std::size_t;

  Vector() = default;
  explicit Vector(size_type n, dataType v = dataType(0)) : d_(n, v) {}
  Vector(const container_type &v) : d_(v) {}

  size_type size() const { return d_.size(); }
  dataType &operator()(size_type i) { return d_[i]; }
  dataType operator()(size_type i) const { return d_[i]; }
  dataType *data() { return d_.data(); }
  const dataType *data() const { return d_.data(); }

  // VectorExpression::dot (old_handmade_linalg/VectorExpression.hpp:31-55), on the device
  dataType dot(const Vector &o) const {
    static_assert(std::is_same<dataType, double>::value, "device dot is FP64");
    if (size() != o.size()) throw std::invalid_argument("Vector::dot dimension mismatch");
    auto ctx = hipdetail::context();
    void *a = nullptr, *b = nullptr;
    hipdetail::check(yafel_hip_malloc(ctx, sizeof(double) * size(), &a), ctx);
    hipdetail::check(yafel_hip_malloc(ctx, sizeof(double) * size(), &b), ctx);
    hipdetail::check(yafel_hip_memcpy_h2d(ctx, a, data(), sizeof(double) * size()), ctx);
    hipdetail::check(yafel_hip_memcpy_h2d(ctx, b, o.data(), sizeof(double) * size()), ctx);
    double r = 0;
    int rc = yafel_hip_dot(ctx, (int64_t)size(), (const double *)a, (const double *)b, &r);
    yafel_hip_free(ctx, a);
    yafel_hip_free(ctx, b);
    hipdetail::check(rc, ctx);
    return r;
  }
  bool operator==(const Vector &o) const { return d_ == o.d_; }
  Vector &operator+=(const Vector &o) {
    for (size_type i = 0; i < size(); ++i) d_[i] += o.d_[i];
    return *this;
  }
  Vector &operator-=(const Vector &o) {
    for (size_type i = 0; i < size(); ++i) d_[i] -= o.d_[i];
    return *this;
  }
  Vector &operator*=(dataType a) {
    for (auto &v : d_) v *= a;
    return *this;
  }
  friend Vector operator+(Vector a, const Vector &b) { return a += b; }
  friend Vector operator-(Vector a, const Vector &b) { return a -= b; }
  friend Vector operator*(Vector a, dataType s) { return a *= s; }
  friend Vector operator*(dataType s, Vector a) { return a *= s; }

 private:
  container_type d_;
};

using triplet = std::tuple<std::size_t, std::size_t, double>;

namespace hipdetail {
inline void split(const std::vector<triplet> &ts, std::vector<int64_t> &r, std::vector<int64_t> &c,
                  std::vector<double> &v) {
  r.resize(ts.size());
  c.resize(ts.size());
  v.resize(ts.size());
  for (std::size_t i = 0; i < ts.size(); ++i) {
    r[i] = (int64_t)std::get<0>(ts[i]);
    c[i] = (int64_t)std::get<1>(ts[i]);
    v[i] = std::get<2>(ts[i]);
  }
}
}  // namespace hipdetail

// ------------------------------------------------------------ sparse_csr
template <typename dataType = double>
class sparse_csr {
  static_assert(std::is_same<dataType, double>::value, "device matrices hold FP64 values (plus an FP32 copy)");

 public:
  using size_type = std::size_t;
  using value_type = dataType;
  using triplet = std::tuple<size_type, size_type, dataType>;

  // sparse_csr.hpp:90-149 on the device
  explicit sparse_csr(const std::vector<triplet> &ts) {
    std::vector<int64_t> r, c;
    std::vector<double> v;
    hipdetail::split(ts, r, c, v);
    auto ctx = hipdetail::context();
    yafel_hip_matrix m = nullptr;
    hipdetail::check(yafel_hip_csr_from_triplets(ctx, (int64_t)ts.size(), r.data(), c.data(), v.data(),
                                                 YAFEL_HIP_FP64, &m),
                     ctx);
    adopt(m);
  }
  // wrap a device matrix (e.g. the tangent assembled by CGAssembly)
  explicit sparse_csr(yafel_hip_matrix m) { adopt(m, false); }

  size_type rows() const { return rows_; }
  size_type cols() const { return cols_; }
  size_type nnz() const { return _data.size(); }
  yafel_hip_matrix handle() const { return m_.get(); }

  // operator()(i,j): value if stored, else 0 (sparse_csr.hpp:151-163)
  value_type operator()(size_type i, size_type j) const {
    if (i >= rows_ || j >= cols_) throw std::out_of_range("sparse_csr::operator() out of bounds");
    for (size_type k = row_ptr[i]; k < row_ptr[i + 1]; ++k)
      if (col_index[k] == j) return _data[k];
    return value_type(0);
  }
  // re-read the host copies after device-side changes (assembly, Dirichlet)
  void refresh() {
    yafel_hip_matrix_info inf;
    hipdetail::check(yafel_hip_matrix_info_get(m_.get(), &inf));
    std::vector<int64_t> p(inf.ptr_len), c(inf.nblocks);
    _data.resize(inf.nnz);
    hipdetail::check(yafel_hip_matrix_download(m_.get(), p.data(), c.data(), _data.data()));
    row_ptr.assign(p.begin(), p.end());
    col_index.assign(c.begin(), c.end());
    rows_ = inf.rows;
    cols_ = inf.cols;
  }

  std::vector<size_type> row_ptr;
  std::vector<size_type> col_index;
  std::vector<dataType> _data;

 private:
  void adopt(yafel_hip_matrix m, bool own = true) {
    if (own) m_ = hipdetail::MatrixPtr(m, hipdetail::MatrixDeleter());
    else m_ = hipdetail::MatrixPtr(m, [](yafel_hip_matrix_s *) {});
    refresh();
  }
  hipdetail::MatrixPtr m_;
  size_type rows_ = 0, cols_ = 0;
};

// ----------------------------------------------------------- sparse_bcsr
template <unsigned BLOCK, typename dataType = double>
class sparse_bcsr {
  static_assert(BLOCK == 3, "the device BCSR format holds 3x3 blocks");
  static_assert(std::is_same<dataType, double>::value, "device matrices hold FP64 values");

 public:
  using size_type = std::size_t;
  using value_type = dataType;
  using triplet = std::tuple<size_type, size_type, dataType>;

  explicit sparse_bcsr(const std::vector<triplet> &ts) {
    std::vector<int64_t> r, c;
    std::vector<double> v;
    hipdetail::split(ts, r, c, v);
    auto ctx = hipdetail::context();
    yafel_hip_matrix m = nullptr;
    hipdetail::check(yafel_hip_bcsr3_from_triplets(ctx, (int64_t)ts.size(), r.data(), c.data(), v.data(),
                                                   YAFEL_HIP_FP64, &m),
                     ctx);
    m_ = hipdetail::MatrixPtr(m, hipdetail::MatrixDeleter());
    refresh();
  }
  explicit sparse_bcsr(yafel_hip_matrix m) {
    m_ = hipdetail::MatrixPtr(m, [](yafel_hip_matrix_s *) {});
    refresh();
  }

  size_type rows() const { return rows_; }
  size_type cols() const { return cols_; }
  size_type nnz() const { return data.size(); }
  yafel_hip_matrix handle() const { return m_.get(); }

  value_type operator()(size_type i, size_type j) const {
    if (i >= rows_ || j >= cols_) throw std::out_of_range("sparse_bcsr::operator() out of bounds");
    const size_type I = i / BLOCK, J = j / BLOCK;
    for (size_type k = brow_ptr[I]; k < brow_ptr[I + 1]; ++k)
      if (bcol_index[k] == J) return data[k * BLOCK * BLOCK + (i % BLOCK) * BLOCK + j % BLOCK];
    return value_type(0);
  }
  void refresh() {
    yafel_hip_matrix_info inf;
    hipdetail::check(yafel_hip_matrix_info_get(m_.get(), &inf));
    std::vector<int64_t> p(inf.ptr_len), c(inf.nblocks);
    data.resize(inf.nnz);
    hipdetail::check(yafel_hip_matrix_download(m_.get(), p.data(), c.data(), data.data()));
    brow_ptr.assign(p.begin(), p.end());
    bcol_index.assign(c.begin(), c.end());
    rows_ = inf.rows;
    cols_ = inf.cols;
  }

  std::vector<size_type> brow_ptr;
  std::vector<size_type> bcol_index;
  std::vector<dataType> data;

 private:
  hipdetail::MatrixPtr m_;
  size_type rows_ = 0, cols_ = 0;
};

// ------------------------------------------------------------- operator*
namespace hipdetail {
inline Vector<double> spmv(yafel_hip_matrix A, std::size_t rows, const Vector<double> &x) {
  Vector<double> y(rows, 0.0);
  check(yafel_hip_spmv_host(A, x.data(), y.data()));
  return y;
}
}  // namespace hipdetail

// csr_spmv (old_handmade_linalg/csr_spmv.hpp:25-56)
inline Vector<double> csr_spmv(const sparse_csr<double> &A, const Vector<double> &x) {
  if (A.cols() != x.size()) throw std::invalid_argument("csr_spmv dimension mismatch");
  return hipdetail::spmv(A.handle(), A.rows(), x);
}
// bcsr_spmv (old_handmade_linalg/bcsr_spmv.hpp:23-62)
inline Vector<double> bcsr_spmv(const sparse_bcsr<3, double> &A, const Vector<double> &x) {
  if (A.cols() != x.size()) throw std::invalid_argument("bcsr_spmv dimension mismatch");
  return hipdetail::spmv(A.handle(), A.rows(), x);
}
inline Vector<double> operator*(const sparse_csr<double> &A, const Vector<double> &x) { return csr_spmv(A, x); }
inline Vector<double> operator*(const sparse_bcsr<3, double> &A, const Vector<double> &x) {
  return bcsr_spmv(A, x);
}

// ------------------------------------------------------------- solvers
constexpr double CG_SOLVER_TOL = 1.0e-14;
constexpr double PCG_SOLVER_TOL = 1.0e-14;
constexpr double BICGSTAB_SOLVER_TOL = 1.0e-14;

namespace hipdetail {
inline Vector<double> solve(yafel_hip_matrix A, int method, int precision, int stop, double tol,
                            const Vector<double> &b, const Vector<double> *x0, yafel_hip_solve_info *out = nullptr) {
  yafel_hip_solve_opts o;
  yafel_hip_solve_opts_default(&o);
  o.method = method;
  o.precision = precision;
  o.stop = stop;
  o.tol = tol;
  Vector<double> x(b.size(), 0.0);
  yafel_hip_solve_info info;
  check(yafel_hip_solve_host(A, &o, b.data(), x0 ? x0->data() : nullptr, x.data(), &info));
  if (out) *out = info;
  return x;
}
}  // namespace hipdetail

// cg_solve (old_handmade_linalg/solver/iterative/cg_solve.hpp:23-81): stops on
// r.r / r0.r0 < TOL, at most n iterations, returns x.
template <class M>
Vector<double> cg_solve(const M &A, const Vector<double> &rhs, const Vector<double> &x0,
                        double TOL = CG_SOLVER_TOL) {
  return hipdetail::solve(A.handle(), YAFEL_HIP_CG, YAFEL_HIP_FP64, YAFEL_HIP_STOP_REFERENCE, TOL, rhs, &x0);
}
template <class M>
Vector<double> cg_solve(const M &A, const Vector<double> &rhs, double TOL = CG_SOLVER_TOL) {
  return hipdetail::solve(A.handle(), YAFEL_HIP_CG, YAFEL_HIP_FP64, YAFEL_HIP_STOP_REFERENCE, TOL, rhs, nullptr);
}

// JacobiPreconditioner (JacobiPreconditioner.hpp:11-54): d^-1 of the diagonal.
template <typename dataType = double>
class JacobiPreconditioner {
 public:
  template <class M>
  explicit JacobiPreconditioner(const M &A) : A_(A.handle()), n_(A.rows()) {}
  // z <- D^-1 z
  void solve(Vector<double> &z) const {
    auto ctx = hipdetail::context();
    void *d = nullptr;
    hipdetail::check(yafel_hip_malloc(ctx, sizeof(double) * n_, &d), ctx);
    int rc = yafel_hip_matrix_diagonal(A_, (double *)d);
    std::vector<double> diag(n_);
    if (rc == YAFEL_HIP_OK) rc = yafel_hip_memcpy_d2h(ctx, diag.data(), d, sizeof(double) * n_);
    yafel_hip_free(ctx, d);
    hipdetail::check(rc, ctx);
    for (std::size_t i = 0; i < n_; ++i) z(i) *= 1.0 / diag[i];
  }
  yafel_hip_matrix matrix() const { return A_; }

 private:
  yafel_hip_matrix A_;
  std::size_t n_;
};

// pcg_solve (pcg_solve.hpp:115-178): stops on r.z / r0.z0 < TOL.  The
// reference's 3-argument form drops M and does not compile (pcg_solve.hpp:30-31);
// here it runs with M.
template <class M>
Vector<double> pcg_solve(const M &A, const Vector<double> &rhs, const JacobiPreconditioner<double> &,
                         const Vector<double> &x0, double TOL = PCG_SOLVER_TOL) {
  return hipdetail::solve(A.handle(), YAFEL_HIP_PCG_JACOBI, YAFEL_HIP_FP64, YAFEL_HIP_STOP_REFERENCE, TOL, rhs, &x0);
}
template <class M>
Vector<double> pcg_solve(const M &A, const Vector<double> &rhs, const JacobiPreconditioner<double> &,
                         double TOL = PCG_SOLVER_TOL) {
  return hipdetail::solve(A.handle(), YAFEL_HIP_PCG_JACOBI, YAFEL_HIP_FP64, YAFEL_HIP_STOP_REFERENCE, TOL, rhs,
                          nullptr);
}

// bicgstab_solve (bicgstab_solve.hpp:19-95): ||Ax-b||^2/||b||^2 < 1e-14,
// returns rhs when ||b||^2 < 1e-14.  The reference loops without bound; here
// the limit is 10 n iterations.
template <class M>
Vector<double> bicgstab_solve(const M &A, const Vector<double> &rhs, const Vector<double> &x0) {
  return hipdetail::solve(A.handle(), YAFEL_HIP_BICGSTAB, YAFEL_HIP_FP64, YAFEL_HIP_STOP_REFERENCE,
                          BICGSTAB_SOLVER_TOL, rhs, &x0);
}
template <class M>
Vector<double> bicgstab_solve(const M &A, const Vector<double> &rhs) {
  return hipdetail::solve(A.handle(), YAFEL_HIP_BICGSTAB, YAFEL_HIP_FP64, YAFEL_HIP_STOP_REFERENCE,
                          BICGSTAB_SOLVER_TOL, rhs, nullptr);
}

// ------------------------------------------------------------ LinearSolve
// Tag dispatch (lin_alg/linear_solvers/LinearSolve.hpp:36-67): the HIP tags
// slot in beside the Eigen / ViennaCL ones.
struct HipConjugateGradientTag {
  double tol = 1.0e-8;        // ||b - A x|| / ||b||
  yafel_hip_solve_info info{};
};
struct HipPCGJacobiTag {
  double tol = 1.0e-8;
  yafel_hip_solve_info info{};
};
struct HipMixedPCGJacobiTag {  // FP32 inner iterations, FP64 refinement
  double tol = 1.0e-8;
  yafel_hip_solve_info info{};
};
struct HipBiCGSTABTag {        // right-Jacobi BiCGSTAB, FP64
  double tol = 1.0e-8;
  yafel_hip_solve_info info{};
};

namespace LinearSolve {
namespace detail {
template <class M>
void solve_impl(Vector<double> &x, const M &A, const Vector<double> &b, HipConjugateGradientTag &t) {
  x = hipdetail::solve(A.handle(), YAFEL_HIP_CG, YAFEL_HIP_FP64, YAFEL_HIP_STOP_RELRES, t.tol, b, &x, &t.info);
}
template <class M>
void solve_impl(Vector<double> &x, const M &A, const Vector<double> &b, HipPCGJacobiTag &t) {
  x = hipdetail::solve(A.handle(), YAFEL_HIP_PCG_JACOBI, YAFEL_HIP_FP64, YAFEL_HIP_STOP_RELRES, t.tol, b, &x, &t.info);
}
template <class M>
void solve_impl(Vector<double> &x, const M &A, const Vector<double> &b, HipMixedPCGJacobiTag &t) {
  x = hipdetail::solve(A.handle(), YAFEL_HIP_PCG_JACOBI, YAFEL_HIP_MIXED, YAFEL_HIP_STOP_RELRES, t.tol, b, &x, &t.info);
}
template <class M>
void solve_impl(Vector<double> &x, const M &A, const Vector<double> &b, HipBiCGSTABTag &t) {
  x = hipdetail::solve(A.handle(), YAFEL_HIP_BICGSTAB_JACOBI, YAFEL_HIP_FP64, YAFEL_HIP_STOP_RELRES, t.tol, b, &x,
                       &t.info);
}
}  // namespace detail

template <class M, class Tag>
void solve_inplace(Vector<double> &result, const M &A, const Vector<double> &b, Tag &tag) {
  if (A.rows() != b.size() || A.rows() != A.cols()) throw std::invalid_argument("Incorrect linear system dimensions");
  detail::solve_impl(result, A, b, tag);
}
template <class M, class Tag>
Vector<double> solve(const M &A, const Vector<double> &b, Tag &tag) {
  Vector<double> result(b.size(), 0.0);
  solve_inplace(result, A, b, tag);
  return result;
}
}  // namespace LinearSolve

}  // namespace yafel
