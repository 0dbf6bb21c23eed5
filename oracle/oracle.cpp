/*
 * CPU oracle: a restatement of YAFEL's CPU hot path (assembly -> sparse ->
 * Krylov).  TEST INFRASTRUCTURE ONLY -- the checker for the HIP library, never
 * linked into it.  Each function names the reference file:line it follows
 * (paths relative to the reference tree).
 *
 * Built with -O2 -ffp-contract=off (no fast-math) so that, where it follows
 * the reference's operation order, it agrees bit-for-bit with the reference
 * headers compiled the same way.
 */
#include "oracle.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <tuple>
#include <vector>

namespace {

using u64 = uint64_t;
using i64 = int64_t;

/* ------------------------------------------------------------------------ */
/* Quadrature: src/quadrature/QuadratureRule.cpp                              */
/* ------------------------------------------------------------------------ */

// Gauss-Lobatto, QuadratureRule.cpp:25-63 (Newton on Legendre recurrence).
void gauss_lobatto(int N, double *xs, double *ws) {
  const double PI = std::atan2(1., 1.) * 4;
  const double TOL = 1.0e-15;
  for (int xi = 0; xi < N; ++xi) {
    double x = -std::cos((PI * xi) / (N - 1));
    double xold = x;
    double Pold, Pn, Pnew;
    do {
      xold = x;
      Pold = 1;
      Pn = x;
      for (int n = 2; n < N; ++n) {
        Pnew = ((2 * n - 1) * x * Pn - (n - 1) * Pold) / n;
        Pold = Pn;
        Pn = Pnew;
      }
      x = x - (x - Pold / Pn) / N;
    } while (std::abs(x - xold) > TOL);
    xs[xi] = x;
    ws[xi] = 2 / (N * (N - 1) * Pn * Pn);
  }
}

// Gauss-Legendre, QuadratureRule.cpp:66-106.
void gauss_legendre(int N, double *xs, double *ws) {
  const double PI = std::atan2(1., 1.) * 4;
  const double TOL = 1.0e-15;
  for (int xi = 0; xi < N; ++xi) {
    double x = -std::cos(PI * (1 + xi - 0.25) / (N + 0.5));
    double Pold, Pn, Pnew, dP;
    do {
      Pold = 1;
      Pn = x;
      for (int n = 2; n <= N; ++n) {
        Pnew = ((2 * n - 1) * x * Pn - (n - 1) * Pold) / n;
        Pold = Pn;
        Pn = Pnew;
      }
      dP = (N / (x * x - 1)) * (x * Pn - Pold);
      x -= Pn / dP;
    } while (std::abs(Pn) > TOL);
    xs[xi] = x;
    ws[xi] = 2.0 / ((1 - x * x) * (dP * dP));
  }
}

// Tensor product rule, QuadratureRule.cpp:109-175 (x fastest, weights multiply).
int tensor_rule(int type, int dim, int order, double *xi, double *w) {
  int n1;
  std::vector<double> x1(32), w1(32);
  if (type == 0) {
    n1 = (order + 1) / 2 + (order + 1) % 2;
    gauss_legendre(n1, x1.data(), w1.data());
  } else {
    n1 = (order + 3) / 2 + (order + 3) % 2;
    gauss_lobatto(n1, x1.data(), w1.data());
  }
  int q = 0;
  if (dim == 1) {
    for (int i = 0; i < n1; ++i, ++q) {
      if (xi) { xi[3 * q] = x1[i]; xi[3 * q + 1] = 0; xi[3 * q + 2] = 0; }
      if (w) w[q] = w1[i];
    }
  } else if (dim == 2) {
    for (int i = 0; i < n1; ++i)
      for (int j = 0; j < n1; ++j, ++q) {
        if (xi) { xi[3 * q] = x1[j]; xi[3 * q + 1] = x1[i]; xi[3 * q + 2] = 0; }
        if (w) w[q] = w1[i] * w1[j];
      }
  } else {
    for (int i = 0; i < n1; ++i)
      for (int j = 0; j < n1; ++j)
        for (int k = 0; k < n1; ++k, ++q) {
          if (xi) { xi[3 * q] = x1[k]; xi[3 * q + 1] = x1[j]; xi[3 * q + 2] = x1[i]; }
          if (w) w[q] = w1[i] * w1[j] * w1[k];
        }
  }
  return q;
}

/* ------------------------------------------------------------------------ */
/* Shape functions: src/element/ShapeFunctionUtils.cpp                        */
/* ------------------------------------------------------------------------ */

// jacobi(n, 0, 0) coefficients, highest degree first (ShapeFunctionUtils.cpp:16-62).
std::vector<double> legendre_coeffs(int n) {
  const double alpha = 0, beta = 0;
  std::vector<double> c1{(alpha + beta + 2) / 2, (alpha - beta) / 2};
  std::vector<double> c2{1.0};
  if (n == 0) return c2;
  if (n == 1) return c1;
  std::vector<double> c(c1);
  for (int i = 1; i < n; ++i) {
    double tmp = i + alpha + beta + 1;
    double a = (2 * i + alpha + beta + 1) * (2 * i + alpha + beta + 2) / (2 * (i + 1) * tmp);
    double b = (beta * beta - alpha * alpha) * (2 * i + alpha + beta + 1) /
               (2 * (i + 1) * tmp * (2 * i + alpha + beta));
    double cc = (i + alpha) * (i + beta) * (2 * i + alpha + beta + 2) /
                ((i + 1) * tmp * (2 * i + alpha + beta));
    c.push_back(0);
    for (auto &v : c) v *= a;
    auto rit = c.rbegin();
    for (auto r1 = c1.rbegin(); r1 != c1.rend(); ++rit, ++r1) *rit -= b * (*r1);
    rit = c.rbegin();
    for (auto r2 = c2.rbegin(); r2 != c2.rend(); ++rit, ++r2) *rit -= cc * (*r2);
    c2.swap(c1);
    c1.assign(c.begin(), c.end());
  }
  return c;
}

// Horner evaluation with derivative (poly_eval over a dual number,
// include/element/ShapeFunctionUtils.hpp:42-50).
inline void poly_eval_d(const std::vector<double> &c, double x, double &v, double &dv) {
  v = 0;
  dv = 0;
  for (double ci : c) {
    dv = dv * x + v;
    v = v * x + ci;
  }
}

// Local node positions: GLL points, x fastest (make_tensorproduct_element.cpp:21-25, 158-193).
void local_nodes(int dim, int p, double *xl) {
  int n1 = p + 1;
  std::vector<double> g(n1), gw(n1);
  gauss_lobatto(n1, g.data(), gw.data());
  int idx = 0;
  if (dim == 1) {
    for (int i = 0; i < n1; ++i, ++idx) { xl[3 * idx] = g[i]; xl[3 * idx + 1] = 0; xl[3 * idx + 2] = 0; }
  } else if (dim == 2) {
    for (int i = 0; i < n1; ++i)
      for (int j = 0; j < n1; ++j, ++idx) {
        xl[3 * idx + 1] = g[i]; xl[3 * idx] = g[j]; xl[3 * idx + 2] = 0;
      }
  } else {
    for (int i = 0; i < n1; ++i)
      for (int j = 0; j < n1; ++j)
        for (int k = 0; k < n1; ++k, ++idx) {
          xl[3 * idx + 2] = g[i]; xl[3 * idx + 1] = g[j]; xl[3 * idx] = g[k];
        }
  }
}

// Solve M y = rhs (M dense n x n, row-major) with partial-pivot LU (Eigen's
// PartialPivLU in the reference).  Multiple right-hand sides (columns of R).
void lu_solve(int n, std::vector<double> M, int nrhs, std::vector<double> &R) {
  std::vector<int> piv(n);
  for (int k = 0; k < n; ++k) {
    int pr = k;
    for (int i = k + 1; i < n; ++i)
      if (std::abs(M[i * n + k]) > std::abs(M[pr * n + k])) pr = i;
    piv[k] = pr;
    if (pr != k) {
      for (int j = 0; j < n; ++j) std::swap(M[k * n + j], M[pr * n + j]);
      for (int j = 0; j < nrhs; ++j) std::swap(R[k * nrhs + j], R[pr * nrhs + j]);
    }
    for (int i = k + 1; i < n; ++i) {
      double l = M[i * n + k] / M[k * n + k];
      M[i * n + k] = l;
      for (int j = k + 1; j < n; ++j) M[i * n + j] -= l * M[k * n + j];
      for (int j = 0; j < nrhs; ++j) R[i * nrhs + j] -= l * R[k * nrhs + j];
    }
  }
  for (int i = n - 1; i >= 0; --i)
    for (int j = 0; j < nrhs; ++j) {
      double s = R[i * nrhs + j];
      for (int k = i + 1; k < n; ++k) s -= M[i * n + k] * R[k * nrhs + j];
      R[i * nrhs + j] = s / M[i * n + i];
    }
}

// tensor_product_shape_functions (ShapeFunctionUtils.cpp:64-143): Legendre
// tensor basis n -> (i=n/p', j=n%p') with x fastest, Vandermonde V(i,j) =
// phi_j(node_i), values/gradients = V^{-T} phi(xi_q).
void shape_tables(int dim, int p, double *N, double *dN, double *w, double *xi_out) {
  int n1 = p + 1;
  int nen = 1;
  for (int d = 0; d < dim; ++d) nen *= n1;
  std::vector<double> xl(nen * 3);
  local_nodes(dim, p, xl.data());
  std::vector<double> qx(1024 * 3), qw(1024);
  int nqp = tensor_rule(0, dim, 2 * p, qx.data(), qw.data());
  std::vector<std::vector<double>> jc;
  for (int n = 0; n < n1; ++n) jc.push_back(legendre_coeffs(n));

  auto phi = [&](int n, const double *x, double *val, double *grad) {
    int idx[3] = {0, 0, 0};
    if (dim == 1) { idx[0] = n; }
    else if (dim == 2) { idx[1] = n / n1; idx[0] = n % n1; }
    else { idx[2] = n / (n1 * n1); idx[1] = (n % (n1 * n1)) / n1; idx[0] = n % n1; }
    double v[3], dv[3];
    for (int d = 0; d < dim; ++d) poly_eval_d(jc[idx[d]], x[d], v[d], dv[d]);
    double pv = 1;
    for (int d = 0; d < dim; ++d) pv *= v[d];
    *val = pv;
    if (grad)
      for (int g = 0; g < dim; ++g) {
        double pg = 1;
        for (int d = 0; d < dim; ++d) pg *= (d == g) ? dv[d] : v[d];
        grad[g] = pg;
      }
  };

  // V^T(j, i) = V(i, j) = phi_j(node_i)
  std::vector<double> VT(nen * nen);
  for (int i = 0; i < nen; ++i)
    for (int j = 0; j < nen; ++j) {
      double v;
      phi(j, &xl[3 * i], &v, nullptr);
      VT[j * nen + i] = v;
    }
  for (int q = 0; q < nqp; ++q) {
    int nr = 1 + dim;
    std::vector<double> R(nen * nr);
    for (int j = 0; j < nen; ++j) {
      double v, g[3];
      phi(j, &qx[3 * q], &v, g);
      R[j * nr] = v;
      for (int d = 0; d < dim; ++d) R[j * nr + 1 + d] = g[d];
    }
    lu_solve(nen, VT, nr, R);
    for (int a = 0; a < nen; ++a) {
      N[q * nen + a] = R[a * nr];
      for (int d = 0; d < dim; ++d) dN[(q * nen + a) * dim + d] = R[a * nr + 1 + d];
    }
    w[q] = qw[q];
    if (xi_out)
      for (int d = 0; d < 3; ++d) xi_out[3 * q + d] = qx[3 * q + d];
  }
}

/* ------------------------------------------------------------------------ */
/* DoF numbering: src/utils/DoFManager.cpp                                    */
/* ------------------------------------------------------------------------ */

// interpolate_from_corners, DoFManager.cpp:296-338 (gmsh corner order).
void interp_corners(int dim, const double *xi, const double *c /* corners*3 */, double *out) {
  for (int d = 0; d < 3; ++d) {
    if (dim == 2) {
      double a = (1 - xi[0]) * (1 - xi[1]), b = (1 + xi[0]) * (1 - xi[1]);
      double e = (1 + xi[0]) * (1 + xi[1]), f = (1 - xi[0]) * (1 + xi[1]);
      out[d] = (1.0 / 4.0) * (a * c[0 * 3 + d] + b * c[1 * 3 + d] + e * c[2 * 3 + d] + f * c[3 * 3 + d]);
    } else {
      double s[8];
      s[0] = (1 - xi[0]) * (1 - xi[1]) * (1 - xi[2]);
      s[1] = (1 + xi[0]) * (1 - xi[1]) * (1 - xi[2]);
      s[2] = (1 + xi[0]) * (1 + xi[1]) * (1 - xi[2]);
      s[3] = (1 - xi[0]) * (1 + xi[1]) * (1 - xi[2]);
      s[4] = (1 - xi[0]) * (1 - xi[1]) * (1 + xi[2]);
      s[5] = (1 + xi[0]) * (1 - xi[1]) * (1 + xi[2]);
      s[6] = (1 + xi[0]) * (1 + xi[1]) * (1 + xi[2]);
      s[7] = (1 - xi[0]) * (1 + xi[1]) * (1 + xi[2]);
      double acc = s[0] * c[d];
      for (int k = 1; k < 8; ++k) acc += s[k] * c[k * 3 + d];
      out[d] = (1.0 / 8.0) * acc;
    }
  }
}

/* ------------------------------------------------------------------------ */
/* Element geometry: include/element/Element.hpp:109-149 and the NSD<=3      */
/* det/inverse of lin_alg/tensor/tensor_functions/rank2_specializations.hpp.   */
/* ------------------------------------------------------------------------ */
void jac_inv(int nsd, const double *J, double &det, double *Jinv) {
  if (nsd == 2) {
    det = J[0] * J[3] - J[1] * J[2];
    Jinv[0] = J[3] / det; Jinv[1] = -J[1] / det;
    Jinv[2] = -J[2] / det; Jinv[3] = J[0] / det;
  } else {
    const double a = J[0], b = J[1], c = J[2], d = J[3], e = J[4], f = J[5], g = J[6], h = J[7], i = J[8];
    det = a * (e * i - f * h) - b * (d * i - f * g) + c * (d * h - e * g);
    Jinv[0] = (e * i - f * h) / det; Jinv[1] = (c * h - b * i) / det; Jinv[2] = (b * f - c * e) / det;
    Jinv[3] = (f * g - d * i) / det; Jinv[4] = (a * i - c * g) / det; Jinv[5] = (c * d - a * f) / det;
    Jinv[6] = (d * h - e * g) / det; Jinv[7] = (b * g - a * h) / det; Jinv[8] = (a * e - b * d) / det;
  }
}

struct ElementTables {
  int dim, p, nen, nqp;
  std::vector<double> N, dN, w;
};

ElementTables make_tables(int dim, int p) {
  ElementTables t;
  t.dim = dim; t.p = p;
  t.nen = 1;
  for (int d = 0; d < dim; ++d) t.nen *= (p + 1);
  t.nqp = t.nen;  // (p+1)^dim Gauss points with multiplier 2
  t.N.resize(t.nqp * t.nen);
  t.dN.resize(t.nqp * t.nen * dim);
  t.w.resize(t.nqp);
  shape_tables(dim, p, t.N.data(), t.dN.data(), t.w.data(), nullptr);
  return t;
}

// One element: qp loop of CGAssembly.hpp:129-145 with Element::update and the
// physics LocalTangent/LocalResidual.  K is (nen*dpn)^2 row-major, zeroed here.
void element_kernel(const ElementTables &T, int dpn, const double *X, const int32_t *conn,
                    i64 e, const orc_physics *ph, double *K, double *R) {
  const int nsd = T.dim, nen = T.nen, nld = nen * dpn;
  std::fill(K, K + nld * nld, 0.0);
  std::fill(R, R + nld, 0.0);
  const double ke = ph->coef ? ph->coef[e] : 1.0;
  std::vector<double> G(nen * nsd);
  for (int q = 0; q < T.nqp; ++q) {
    double J[9] = {0};
    for (int A = 0; A < nen; ++A) {
      const double *x = &X[3 * (i64)conn[A]];
      for (int i = 0; i < nsd; ++i)
        for (int j = 0; j < nsd; ++j) J[i * nsd + j] += x[i] * T.dN[(q * nen + A) * nsd + j];
    }
    double det, Ji[9];
    jac_inv(nsd, J, det, Ji);
    const double jxw = det * T.w[q];
    // shapeGrad(A,d) = sum_k JinvT(d,k) dN(A,k) = sum_k Jinv(k,d) dN(A,k)
    for (int A = 0; A < nen; ++A)
      for (int d = 0; d < nsd; ++d) {
        double s = 0;
        for (int k = 0; k < nsd; ++k) s += Ji[k * nsd + d] * T.dN[(q * nen + A) * nsd + k];
        G[A * nsd + d] = s;
      }
    const double *Nq = &T.N[q * nen];
    if (ph->kind == 0) {
      // apps/BasicPoisson/main.cpp:45-58: K(A,B) -= grad_A . grad_B * jxw, B>=A, mirrored.
      for (int A = 0; A < nen; ++A)
        for (int B = A; B < nen; ++B) {
          double dt = 0;
          for (int d = 0; d < nsd; ++d) dt += G[A * nsd + d] * G[B * nsd + d];
          K[A * nld + B] -= ke * dt * jxw;
          K[B * nld + A] = K[A * nld + B];
        }
      // apps/BasicPoisson/main.cpp:31-38: R += f * jxw * N.
      for (int A = 0; A < nen; ++A) R[A] += (ph->f * jxw) * Nq[A];
    } else if (ph->kind == 1) {
      // apps/BasicLinearElasticity/main.cpp:64-112: C_ikjl = lam d_ij d_kl + mu (d_ik d_jl + d_il d_jk);
      // K_(A,i),(B,k) -= sum_jl C_ikjl dN_A,j dN_B,l jxw (diagonal block full, B>A mirrored).
      const double lam = ke * ph->lambda, mu = ke * ph->mu;
      for (int An = 0; An < nen; ++An)
        for (int i = 0; i < nsd; ++i) {
          const int a = An * nsd + i;
          for (int Bn = An; Bn < nen; ++Bn)
            for (int k = 0; k < nsd; ++k) {
              const int b = Bn * nsd + k;
              double s = 0;
              for (int j = 0; j < nsd; ++j)
                for (int l = 0; l < nsd; ++l) {
                  double C = lam * (i == j) * (k == l) + mu * ((i == k) * (j == l) + (i == l) * (j == k));
                  s += C * (G[An * nsd + j] * G[Bn * nsd + l]);
                }
              K[a * nld + b] -= s * jxw;
              if (Bn != An) K[b * nld + a] = K[a * nld + b];
            }
        }
      // R_(A,i) += b_i N_A jxw (body force; apps/BasicLinearElasticity/main.cpp:53-61 uses 10 in all comps)
      for (int A = 0; A < nen; ++A)
        for (int i = 0; i < nsd; ++i) R[A * nsd + i] += ph->body[i] * Nq[A] * jxw;
    } else {
      // apps/supg_learning/supg_advection.hpp:65-107: K(A,B) += sum_{Ac,Bc}
      //   (N_A U_Bc dN_B,Bc + dN_A,Ac dN_B,Bc Kt(Ac,Bc)) jxw, Kt = k I + alpha h/(2|U|) U(x)U.
      double Kt[9];
      double nU = 0;
      for (int d = 0; d < nsd; ++d) nU += ph->U[d] * ph->U[d];
      nU = std::sqrt(nU);
      for (int a = 0; a < nsd; ++a)
        for (int b = 0; b < nsd; ++b) {
          Kt[a * nsd + b] = (a == b) ? ph->kdiff : 0.0;
          if (nU > 0) Kt[a * nsd + b] += (ph->alpha * ph->h / (nU * 2)) * ph->U[a] * ph->U[b];
        }
      for (int A = 0; A < nen; ++A)
        for (int Ac = 0; Ac < nsd; ++Ac)
          for (int B = 0; B < nen; ++B)
            for (int Bc = 0; Bc < nsd; ++Bc)
              K[A * nld + B] += (Nq[A] * ph->U[Bc] * G[B * nsd + Bc] +
                                 G[A * nsd + Ac] * G[B * nsd + Bc] * Kt[Ac * nsd + Bc]) * jxw;
      // supg_advection.hpp:30-62 (steady): R(A) -= (N_A Q + Uhat_Ac dN_A,Ac) jxw for each Ac.
      double Uh[3] = {0, 0, 0};
      if (nU > 0)
        for (int d = 0; d < nsd; ++d) Uh[d] = (ph->alpha * ph->h / 2) * ph->U[d] / nU;
      for (int A = 0; A < nen; ++A)
        for (int Ac = 0; Ac < nsd; ++Ac) R[A] -= (Nq[A] * ph->Q + Uh[Ac] * G[A * nsd + Ac]) * jxw;
    }
  }
}

/* ------------------------------------------------------------------------ */
/* Sparse matrices                                                            */
/* ------------------------------------------------------------------------ */
}  // namespace

struct orc_csr {
  int block;        // 1 = CSR, 3 = BCSR3
  i64 rows, cols;   // csr: scalar rows/cols; bcsr: block rows/cols
  std::vector<i64> ptr, col;
  std::vector<double> val;
};

namespace {
using triplet = std::tuple<u64, u64, double>;

std::vector<triplet> make_triplets(i64 nt, const i64 *r, const i64 *c, const double *v) {
  std::vector<triplet> ts(nt);
  for (i64 i = 0; i < nt; ++i) ts[i] = triplet((u64)r[i], (u64)c[i], v[i]);
  return ts;
}
}  // namespace

extern "C" {

void orc_gauss_legendre_1d(int npts, double *x, double *w) { gauss_legendre(npts, x, w); }
void orc_gauss_lobatto_1d(int npts, double *x, double *w) { gauss_lobatto(npts, x, w); }
int orc_tensor_rule(int type, int topo_dim, int poly_order, double *xi, double *w) {
  return tensor_rule(type, topo_dim, poly_order, xi, w);
}
void orc_shape_tables(int dim, int p, double *N, double *dN, double *w, double *xi) {
  shape_tables(dim, p, N, dN, w, xi);
}
void orc_local_nodes(int dim, int p, double *xl) { local_nodes(dim, p, xl); }

void orc_structured_mesh(int dim, int nx, int ny, int nz, int cell_order, double *nodes,
                         int32_t *cells) {
  if (dim == 2) nz = 0;
  const i64 sx = nx + 1, sy = ny + 1;
  for (i64 k = 0; k <= nz; ++k)
    for (i64 j = 0; j <= ny; ++j)
      for (i64 i = 0; i <= nx; ++i) {
        i64 n = i + sx * (j + sy * k);
        nodes[3 * n] = (double)i / nx;
        nodes[3 * n + 1] = (double)j / ny;
        nodes[3 * n + 2] = dim == 3 ? (double)k / nz : 0.0;
      }
  auto id = [&](i64 i, i64 j, i64 k) { return (int32_t)(i + sx * (j + sy * k)); };
  i64 c = 0;
  if (dim == 2) {
    for (i64 j = 0; j < ny; ++j)
      for (i64 i = 0; i < nx; ++i, ++c) {
        int32_t *o = &cells[4 * c];
        o[0] = id(i, j, 0); o[1] = id(i + 1, j, 0);
        if (cell_order == 0) { o[2] = id(i, j + 1, 0); o[3] = id(i + 1, j + 1, 0); }
        else { o[2] = id(i + 1, j + 1, 0); o[3] = id(i, j + 1, 0); }
      }
  } else {
    for (i64 k = 0; k < nz; ++k)
      for (i64 j = 0; j < ny; ++j)
        for (i64 i = 0; i < nx; ++i, ++c) {
          int32_t *o = &cells[8 * c];
          for (int l = 0; l < 2; ++l) {
            o[4 * l + 0] = id(i, j, k + l); o[4 * l + 1] = id(i + 1, j, k + l);
            if (cell_order == 0) { o[4 * l + 2] = id(i, j + 1, k + l); o[4 * l + 3] = id(i + 1, j + 1, k + l); }
            else { o[4 * l + 2] = id(i + 1, j + 1, k + l); o[4 * l + 3] = id(i, j + 1, k + l); }
          }
        }
  }
}

// DoFManager::make_cg_dofs (DoFManager.cpp:143-164): p=1 passes the mesh through;
// p>=2 makes raw nodes per cell (make_raw_dofs :176-229) and recombines
// duplicates by a snapped (z,y,x) sort (recombine_all_duplicates :340-414).
int64_t orc_dofmanager_cg(int dim, int p, int64_t n_corner_nodes, const double *corner_nodes,
                          int64_t ncell, const int32_t *cells, int32_t *elements,
                          double *dof_nodes) {
  const int ncorner = dim == 2 ? 4 : 8;
  if (p == 1) {
    std::memcpy(elements, cells, sizeof(int32_t) * ncell * ncorner);
    std::memcpy(dof_nodes, corner_nodes, sizeof(double) * 3 * n_corner_nodes);
    return n_corner_nodes;
  }
  int nen = 1;
  for (int d = 0; d < dim; ++d) nen *= (p + 1);
  std::vector<double> xl(nen * 3);
  local_nodes(dim, p, xl.data());
  const i64 N = ncell * nen;
  std::vector<double> raw(3 * N);
  for (i64 c = 0; c < ncell; ++c) {
    double cc[8 * 3];
    for (int k = 0; k < ncorner; ++k)
      for (int d = 0; d < 3; ++d) cc[3 * k + d] = corner_nodes[3 * (i64)cells[c * ncorner + k] + d];
    for (int a = 0; a < nen; ++a) interp_corners(dim, &xl[3 * a], cc, &raw[3 * (c * nen + a)]);
    for (int a = 0; a < nen; ++a) elements[c * nen + a] = (int32_t)(c * nen + a);
  }
  double bmin[3], bmax[3];
  for (int d = 0; d < 3; ++d) bmin[d] = bmax[d] = raw[d];
  for (i64 i = 0; i < N; ++i)
    for (int d = 0; d < 3; ++d) {
      bmin[d] = std::min(bmin[d], raw[3 * i + d]);
      bmax[d] = std::max(bmax[d], raw[3 * i + d]);
    }
  double maxdim = std::max(bmax[0] - bmin[0], std::max(bmax[1] - bmin[1], bmax[2] - bmin[2]));
  const double snapv = maxdim * 1.0e-6;
  std::vector<double> snapped(3 * N);
  for (i64 i = 0; i < 3 * N; ++i) snapped[i] = std::round(raw[i] / snapv) * snapv;
  std::vector<i64> idx(N);
  std::iota(idx.begin(), idx.end(), 0);
  // lexicographic compare of reversed coordinate data, i.e. (z, y, x)
  std::stable_sort(idx.begin(), idx.end(), [&](i64 L, i64 R) {
    for (int d = 2; d >= 0; --d) {
      if (snapped[3 * L + d] < snapped[3 * R + d]) return true;
      if (snapped[3 * R + d] < snapped[3 * L + d]) return false;
    }
    return false;
  });
  std::vector<i64> rev(N), uniq(N), revu(N);
  for (i64 i = 0; i < N; ++i) rev[idx[i]] = i;
  std::vector<i64> uu(1, 0);
  uniq[0] = 0;
  auto close = [&](i64 a, i64 b) {
    double s = 0;
    for (int d = 0; d < 3; ++d) {
      double t = raw[3 * a + d] - raw[3 * b + d];
      s += t * t;
    }
    return std::sqrt(s) < 1.0e-6;
  };
  for (i64 i = 1; i < N; ++i) {
    if (close(idx[i], idx[i - 1])) uniq[i] = uniq[i - 1];
    else { uu.push_back(i); uniq[i] = i; }
  }
  i64 cur = 0;
  revu[0] = 0;
  for (i64 i = 1; i < N; ++i) revu[i] = (uniq[i] == uniq[i - 1]) ? cur : ++cur;
  for (i64 u = 0; u < (i64)uu.size(); ++u)
    for (int d = 0; d < 3; ++d) dof_nodes[3 * u + d] = raw[3 * idx[uu[u]] + d];
  for (i64 i = 0; i < N; ++i) elements[i] = (int32_t)revu[uniq[rev[elements[i]]]];
  return (int64_t)uu.size();
}

void orc_element_matrix(int dim, int p, int dpn, const double *dof_nodes, const int32_t *elements,
                        int64_t e, const orc_physics *phys, double *Ke, double *Re) {
  ElementTables T = make_tables(dim, p);
  std::vector<double> R(T.nen * dpn);
  element_kernel(T, dpn, dof_nodes, &elements[e * T.nen], e, phys, Ke, Re ? Re : R.data());
}

// CGAssembly.hpp:86-162: per element K_e/R_e, emit all local pairs (zeros included)
// with dof = node*dpn + comp (DoFManager.cpp:39-52), R accumulated.
void orc_cg_assembly(int dim, int p, int dpn, int64_t n_nodes, const double *dof_nodes,
                     int64_t nel, const int32_t *elements, const orc_physics *phys, int64_t *rows,
                     int64_t *cols, double *vals, double *R) {
  (void)n_nodes;
  ElementTables T = make_tables(dim, p);
  const int nld = T.nen * dpn;
  std::vector<double> K(nld * nld), Re(nld);
  std::vector<i64> gd(nld);
  for (i64 e = 0; e < nel; ++e) {
    const int32_t *conn = &elements[e * T.nen];
    element_kernel(T, dpn, dof_nodes, conn, e, phys, K.data(), Re.data());
    for (int A = 0; A < T.nen; ++A)
      for (int c = 0; c < dpn; ++c) gd[A * dpn + c] = (i64)conn[A] * dpn + c;
    const i64 base = e * (i64)nld * nld;
    for (int a = 0; a < nld; ++a) {
      for (int b = 0; b < nld; ++b) {
        if (rows) rows[base + a * nld + b] = gd[a];
        if (cols) cols[base + a * nld + b] = gd[b];
        if (vals) vals[base + a * nld + b] = K[a * nld + b];
      }
      if (R) R[gd[a]] += Re[a];
    }
  }
}

// CGAssembly.hpp:73-185 with its OpenMP structure, for the timed CPU baseline:
// every thread builds its own element tables (one ElementFactory per thread,
// :84), keeps its elements' triplets and a full-length residual private
// (:80, :149-159), and appends them to the shared list under a critical
// section (:175-181).  The append order follows thread timing, so the result
// is orc_cg_assembly's up to the order of the triplets and the rounding of R.
// Returns the number of triplets written (nel * nld^2).
int64_t orc_cg_assembly_threads(int nthreads, int dim, int p, int dpn, int64_t n_nodes, const double *dof_nodes,
                                int64_t nel, const int32_t *elements, const orc_physics *phys, int64_t *rows,
                                int64_t *cols, double *vals, double *R) {
  i64 filled = 0;
#ifdef _OPENMP
#pragma omp parallel num_threads(nthreads)
#endif
  {
    ElementTables T = make_tables(dim, p);
    const int nld = T.nen * dpn;
    std::vector<double> K(nld * nld), Re(nld);
    std::vector<i64> gd(nld);
    std::vector<i64> tr, tc;
    std::vector<double> tv, Rt(R ? (size_t)n_nodes * dpn : 0, 0.0);
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
    for (i64 e = 0; e < nel; ++e) {
      const int32_t *conn = &elements[e * T.nen];
      element_kernel(T, dpn, dof_nodes, conn, e, phys, K.data(), Re.data());
      for (int A = 0; A < T.nen; ++A)
        for (int c = 0; c < dpn; ++c) gd[A * dpn + c] = (i64)conn[A] * dpn + c;
      for (int a = 0; a < nld; ++a) {
        for (int b = 0; b < nld; ++b) {
          tr.push_back(gd[a]);
          tc.push_back(gd[b]);
          tv.push_back(K[a * nld + b]);
        }
        if (R) Rt[gd[a]] += Re[a];
      }
    }
#ifdef _OPENMP
#pragma omp critical
#endif
    {
      std::copy(tr.begin(), tr.end(), rows + filled);
      std::copy(tc.begin(), tc.end(), cols + filled);
      std::copy(tv.begin(), tv.end(), vals + filled);
      filled += (i64)tv.size();
      if (R)
        for (size_t i = 0; i < Rt.size(); ++i) R[i] += Rt[i];
    }
  }
  return filled;
}

// sparse_csr<T>::sparse_csr(triplets), sparse_csr.hpp:90-149: sort on full
// (row, col, val) tuples, sum duplicates in that order, row_ptr of max_row+2.
orc_csr *orc_csr_from_triplets(int64_t nt, const int64_t *r, const int64_t *c, const double *v) {
  orc_csr *A = new orc_csr();
  A->block = 1;
  A->rows = A->cols = 0;
  if (nt == 0) return A;
  auto ts = make_triplets(nt, r, c, v);
  std::sort(ts.begin(), ts.end());
  A->ptr.assign(std::get<0>(ts.back()) + 2, 0);
  u64 cr = std::get<0>(ts[0]), cc = std::get<1>(ts[0]);
  A->rows = cr + 1;
  A->cols = cc + 1;
  A->col.push_back(cc);
  A->val.push_back(std::get<2>(ts[0]));
  i64 idx = 0;
  for (i64 i = 1; i < nt; ++i) {
    u64 row = std::get<0>(ts[i]), col = std::get<1>(ts[i]);
    double val = std::get<2>(ts[i]);
    A->rows = std::max<i64>(A->rows, row + 1);
    A->cols = std::max<i64>(A->cols, col + 1);
    if (row == cr && col == cc) {
      A->val[idx] += val;
    } else {
      ++idx;
      if (row != cr) {
        for (u64 rr = cr + 1; rr <= row; ++rr) A->ptr[rr] = idx;
        cr = row;
      }
      A->col.push_back(col);
      cc = col;
      A->val.push_back(val);
    }
  }
  A->ptr[A->rows] = A->val.size();
  return A;
}

// sparse_bcsr<3>::sparse_bcsr(triplets), sparse_bcsr.hpp:93-176: sort on
// (I, J, i, j), sum in that (stable) order into zero-filled 3x3 row-major
// blocks; _brows/_bcols are only updated from the second triplet on.
orc_csr *orc_bcsr3_from_triplets(int64_t nt, const int64_t *r, const int64_t *c, const double *v) {
  const u64 B = 3;
  orc_csr *A = new orc_csr();
  A->block = 3;
  A->rows = A->cols = 0;
  if (nt == 0) return A;
  auto ts = make_triplets(nt, r, c, v);
  auto key = [&](const triplet &t) {
    return std::make_tuple(std::get<0>(t) / B, std::get<1>(t) / B, std::get<0>(t) % B, std::get<1>(t) % B);
  };
  std::stable_sort(ts.begin(), ts.end(), [&](const triplet &a, const triplet &b) { return key(a) < key(b); });
  const u64 brows = std::get<0>(ts.back()) / B + 1;
  A->ptr.assign(brows + 1, 0);
  u64 cb = std::get<0>(ts[0]) / B, cbc = std::get<1>(ts[0]) / B;
  i64 nblocks = 1;
  for (i64 i = 1; i < nt; ++i) {
    u64 I = std::get<0>(ts[i]) / B, J = std::get<1>(ts[i]) / B;
    if (!(cb == I && cbc == J)) ++nblocks;
    cb = I;
    cbc = J;
  }
  A->col.assign(nblocks, 0);
  A->val.assign(nblocks * 9, 0.0);
  cb = std::get<0>(ts[0]) / B;
  cbc = std::get<1>(ts[0]) / B;
  A->col[0] = cbc;
  A->val[(std::get<0>(ts[0]) % B) * B + std::get<1>(ts[0]) % B] = std::get<2>(ts[0]);
  i64 bidx = 0;
  u64 nbr = 0, nbc = 0;
  for (i64 t = 1; t < nt; ++t) {
    u64 I = std::get<0>(ts[t]) / B, J = std::get<1>(ts[t]) / B;
    u64 i = std::get<0>(ts[t]) % B, j = std::get<1>(ts[t]) % B;
    double val = std::get<2>(ts[t]);
    nbr = std::max(nbr, I + 1);
    nbc = std::max(nbc, J + 1);
    if (I == cb && J == cbc) {
      A->val[bidx * 9 + i * B + j] += val;
    } else {
      ++bidx;
      if (cb != I) {
        for (u64 rr = cb + 1; rr <= I; ++rr) A->ptr[rr] = bidx;
        cb = I;
      }
      cbc = J;
      A->col[bidx] = J;
      A->val[bidx * 9 + i * B + j] += val;
    }
  }
  A->ptr[nbr] = nblocks;
  A->rows = nbr;
  A->cols = nbc;
  return A;
}

int64_t orc_csr_rows(const orc_csr *A) { return A->rows * A->block; }
int64_t orc_csr_cols(const orc_csr *A) { return A->cols * A->block; }
int64_t orc_csr_nnz(const orc_csr *A) { return (int64_t)A->col.size(); }
int64_t orc_csr_nptr(const orc_csr *A) { return (int64_t)A->ptr.size(); }
void orc_csr_copy(const orc_csr *A, int64_t *row_ptr, int64_t *col, double *val) {
  if (row_ptr) std::copy(A->ptr.begin(), A->ptr.end(), row_ptr);
  if (col) std::copy(A->col.begin(), A->col.end(), col);
  if (val) std::copy(A->val.begin(), A->val.end(), val);
}
void orc_csr_free(orc_csr *A) { delete A; }

orc_csr *orc_csr_wrap(int block, int64_t nrows, int64_t ncols, const int64_t *row_ptr,
                      const int64_t *col, const double *val) {
  orc_csr *A = new orc_csr();
  A->block = block;
  A->rows = nrows;
  A->cols = ncols;
  A->ptr.assign(row_ptr, row_ptr + nrows + 1);
  i64 nnz = row_ptr[nrows];
  A->col.assign(col, col + nnz);
  A->val.assign(val, val + nnz * block * block);
  return A;
}

// csr_spmv.hpp:25-56 (sequential per row in column order) and
// bcsr_spmv.hpp:23-62 (per block row, accum[i] += B(i,j) x(3J+j), j fastest).
void orc_spmv(const orc_csr *A, const double *x, double *y) {
  if (A->block == 1) {
    for (i64 r = 0; r < A->rows; ++r) {
      double acc = 0;
      for (i64 k = A->ptr[r]; k < A->ptr[r + 1]; ++k) acc += A->val[k] * x[A->col[k]];
      y[r] = acc;
    }
  } else {
    for (i64 br = 0; br < A->rows; ++br) {
      double acc[3] = {0, 0, 0};
      for (i64 k = A->ptr[br]; k < A->ptr[br + 1]; ++k)
        for (int i = 0; i < 3; ++i)
          for (int j = 0; j < 3; ++j) acc[i] += A->val[k * 9 + i * 3 + j] * x[A->col[k] * 3 + j];
      for (int i = 0; i < 3; ++i) y[3 * br + i] = acc[i];
    }
  }
}

// VectorExpression::dot, VectorExpression.hpp:31-55: four interleaved partial
// sums, tail into s0, returns s0+s1+s2+s3.
double orc_dot(int64_t n, const double *a, const double *b) {
  double s0 = 0, s1 = 0, s2 = 0, s3 = 0;
  i64 NN = 4 * (n / 4);
  for (i64 i = 0; i < NN; i += 4) {
    s0 += a[i] * b[i];
    s1 += a[i + 1] * b[i + 1];
    s2 += a[i + 2] * b[i + 2];
    s3 += a[i + 3] * b[i + 3];
  }
  for (i64 i = NN; i < n; ++i) s0 += a[i] * b[i];
  return s0 + s1 + s2 + s3;
}

// DirichletBC::apply (RowMajor form), DirichletBC.cpp:27-56: rhs -= A g, then
// every stored entry in a BC row or column is zeroed, the diagonal set to 1
// and rhs(r) = g(r).  The pattern is kept.
void orc_dirichlet_apply(orc_csr *A, double *rhs, const uint8_t *mask, const double *g) {
  const i64 n = A->rows * A->block;
  std::vector<double> Ag(n);
  orc_spmv(A, g, Ag.data());
  for (i64 i = 0; i < n; ++i) rhs[i] -= Ag[i];
  const int B = A->block;
  for (i64 br = 0; br < A->rows; ++br)
    for (i64 k = A->ptr[br]; k < A->ptr[br + 1]; ++k)
      for (int i = 0; i < B; ++i)
        for (int j = 0; j < B; ++j) {
          i64 r = br * B + i, c = A->col[k] * B + j;
          if (mask[r] || mask[c]) {
            A->val[k * B * B + i * B + j] = 0;
            if (r == c) {
              rhs[r] = g[c];
              A->val[k * B * B + i * B + j] = 1;
            }
          }
        }
}

// cg_solve.hpp:33-81.
int64_t orc_cg_solve(const orc_csr *A, const double *b, const double *x0, double tol, double *x) {
  const i64 n = A->rows * A->block;
  std::vector<double> r(n), p(n), Ap(n);
  std::copy(x0, x0 + n, x);
  orc_spmv(A, x0, Ap.data());
  for (i64 i = 0; i < n; ++i) r[i] = b[i] - Ap[i];
  p = r;
  i64 k = 0;
  const i64 maxiter = n;
  double rr_old = orc_dot(n, r.data(), r.data());
  const double rr0 = rr_old;
  if (rr_old == 0.0) return 0;
  while (k++ < maxiter) {
    orc_spmv(A, p.data(), Ap.data());
    double alpha = rr_old / orc_dot(n, p.data(), Ap.data());
    for (i64 i = 0; i < n; ++i) x[i] += p[i] * alpha;
    for (i64 i = 0; i < n; ++i) r[i] += -alpha * Ap[i];
    double rr_new = orc_dot(n, r.data(), r.data());
    if (rr_new / rr0 < tol) break;
    double beta = rr_new / rr_old;
    rr_old = rr_new;
    for (i64 i = 0; i < n; ++i) p[i] *= beta;
    for (i64 i = 0; i < n; ++i) p[i] += r[i];
  }
  return k;
}

// pcg_solve.hpp:126-178 with JacobiPreconditioner (d^-1 from A(i,i), 1/0 kept
// as the reference does after its warning), JacobiPreconditioner.hpp:124-153.
int64_t orc_pcg_jacobi_solve(const orc_csr *A, const double *b, const double *x0, double tol,
                             double *x) {
  const i64 n = A->rows * A->block;
  const int B = A->block;
  std::vector<double> dinv(n, 0.0);
  for (i64 br = 0; br < A->rows; ++br) {
    for (int i = 0; i < B; ++i) {
      double aii = 0;
      for (i64 k = A->ptr[br]; k < A->ptr[br + 1]; ++k)
        if (A->col[k] == br) { aii = A->val[k * B * B + i * B + i]; break; }
      dinv[br * B + i] = 1.0 / aii;
    }
  }
  std::vector<double> r(n), z(n), p(n), Ap(n);
  std::copy(x0, x0 + n, x);
  orc_spmv(A, x0, Ap.data());
  for (i64 i = 0; i < n; ++i) r[i] = b[i] - Ap[i];
  for (i64 i = 0; i < n; ++i) z[i] = r[i] * dinv[i];
  p = z;
  i64 k = 0;
  const i64 maxiter = n;
  double rz_old = orc_dot(n, r.data(), z.data());
  const double rz0 = rz_old;
  if (rz_old == 0.0) return 0;
  while (k++ < maxiter) {
    orc_spmv(A, p.data(), Ap.data());
    double alpha = rz_old / orc_dot(n, p.data(), Ap.data());
    for (i64 i = 0; i < n; ++i) x[i] += p[i] * alpha;
    for (i64 i = 0; i < n; ++i) r[i] += -alpha * Ap[i];
    for (i64 i = 0; i < n; ++i) z[i] = r[i] * dinv[i];
    double rz_new = orc_dot(n, r.data(), z.data());
    if (rz_new / rz0 < tol) break;
    double beta = rz_new / rz_old;
    rz_old = rz_new;
    for (i64 i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
  }
  return k;
}

// bicgstab_solve.hpp:27-95 (van der Vorst, unpreconditioned; true residual
// checked every max(n/100,1) iterations against ||b||^2; rhs returned when
// ||b||^2 < TOL).
int64_t orc_bicgstab_solve(const orc_csr *A, const double *b, const double *x0, double tol,
                           int64_t max_iter, double *x) {
  const i64 n = A->rows * A->block;
  std::vector<double> r(n), rhat(n), v(n, 0.0), p(n, 0.0), s(n), t(n), tmp(n);
  orc_spmv(A, x0, tmp.data());
  for (i64 i = 0; i < n; ++i) r[i] = b[i] - tmp[i];
  rhat = r;
  std::copy(x0, x0 + n, x);
  const i64 freq = (n / 100 > 1) ? n / 100 : 1;
  const double res0 = orc_dot(n, b, b);
  if (res0 < tol) {
    std::copy(b, b + n, x);
    return 0;
  }
  double rho_old = 1, alpha = 1, w_old = 1, rho_new, w_new;
  i64 k = 0;
  while (true) {
    rho_new = orc_dot(n, rhat.data(), r.data());
    double beta = (rho_new / rho_old) * (alpha / w_old);
    for (i64 i = 0; i < n; ++i) p[i] = r[i] + p[i] * beta - v[i] * (w_old * beta);
    orc_spmv(A, p.data(), v.data());
    alpha = rho_new / orc_dot(n, rhat.data(), v.data());
    for (i64 i = 0; i < n; ++i) s[i] = r[i] - v[i] * alpha;
    orc_spmv(A, s.data(), t.data());
    double tt = orc_dot(n, t.data(), t.data());
    w_new = (tt == 0) ? 0 : orc_dot(n, t.data(), s.data()) / tt;
    for (i64 i = 0; i < n; ++i) x[i] += p[i] * alpha + s[i] * w_new;
    if (k % freq == 0) {
      orc_spmv(A, x, tmp.data());
      for (i64 i = 0; i < n; ++i) tmp[i] -= b[i];
      double res = orc_dot(n, tmp.data(), tmp.data());
      if (res / res0 < tol) break;
    }
    for (i64 i = 0; i < n; ++i) r[i] = s[i] - t[i] * w_new;
    rho_old = rho_new;
    w_old = w_new;
    ++k;
    if (max_iter > 0 && k >= max_iter) break;
  }
  return k + 1;
}

// Right-Jacobi-preconditioned BiCGSTAB on the same recurrence (p_hat = D^-1 p,
// s_hat = D^-1 s), stopping on the true residual ||b - A x|| / ||b|| < rtol.
int64_t orc_bicgstab_jacobi_solve(const orc_csr *A, const double *b, double rtol, int64_t max_iter,
                                  double *x) {
  const i64 n = A->rows * A->block;
  const int B = A->block;
  std::vector<double> dinv(n, 1.0);
  for (i64 br = 0; br < A->rows; ++br)
    for (int i = 0; i < B; ++i)
      for (i64 k = A->ptr[br]; k < A->ptr[br + 1]; ++k)
        if (A->col[k] == br) { dinv[br * B + i] = 1.0 / A->val[k * B * B + i * B + i]; break; }
  std::vector<double> r(b, b + n), rhat(b, b + n), v(n, 0.0), p(n, 0.0), ph(n), s(n), sh(n), t(n), tmp(n);
  std::fill(x, x + n, 0.0);
  const double bb = orc_dot(n, b, b);
  if (bb == 0) return 0;
  double rho_old = 1, alpha = 1, w_old = 1;
  i64 k = 0;
  while (k < max_iter) {
    double rho_new = orc_dot(n, rhat.data(), r.data());
    double beta = (rho_new / rho_old) * (alpha / w_old);
    for (i64 i = 0; i < n; ++i) p[i] = r[i] + beta * (p[i] - w_old * v[i]);
    for (i64 i = 0; i < n; ++i) ph[i] = dinv[i] * p[i];
    orc_spmv(A, ph.data(), v.data());
    alpha = rho_new / orc_dot(n, rhat.data(), v.data());
    for (i64 i = 0; i < n; ++i) s[i] = r[i] - alpha * v[i];
    for (i64 i = 0; i < n; ++i) sh[i] = dinv[i] * s[i];
    orc_spmv(A, sh.data(), t.data());
    double tt = orc_dot(n, t.data(), t.data());
    double w_new = (tt == 0) ? 0 : orc_dot(n, t.data(), s.data()) / tt;
    for (i64 i = 0; i < n; ++i) x[i] += alpha * ph[i] + w_new * sh[i];
    for (i64 i = 0; i < n; ++i) r[i] = s[i] - w_new * t[i];
    rho_old = rho_new;
    w_old = w_new;
    ++k;
    orc_spmv(A, x, tmp.data());
    for (i64 i = 0; i < n; ++i) tmp[i] = b[i] - tmp[i];
    if (std::sqrt(orc_dot(n, tmp.data(), tmp.data()) / bb) < rtol) break;
  }
  return k;
}

// splitmix64 (Steele, Lea, Flood 2014) and Box-Muller: the seeded generator
// shared by the synthetic workloads.
/* ------------------------------------------------------------------------ */
/* DG: include/assembly/DGAssembly.hpp:32-305 for Quad4 cells, dpn = 1,      */
/* with the advection physics of apps/BasicAdvection_DG/main.cpp:14-98.      */
/* ------------------------------------------------------------------------ */
}  // extern "C"

namespace {

struct DGFace {
  int32_t a, b;                 // global corner ids, a < b after orient()
  int32_t left, right, lf, rf;  // cells and local faces (-1: none)
};

// Mesh::buildInternalFaces (build_faces.cpp:25-96) with CellFace::orient
// (CellFace.cpp:74-83) for 2-node faces {f, (f+1) % 4} (CellFace.cpp:18-21).
std::vector<DGFace> dg_faces(i64 ncell, const int32_t *cells) {
  std::vector<DGFace> F;
  F.reserve(4 * ncell);
  for (i64 c = 0; c < ncell; ++c)
    for (int f = 0; f < 4; ++f) {
      DGFace x{cells[4 * c + f], cells[4 * c + (f + 1) % 4], (int32_t)c, -1, f, -1};
      if (x.b < x.a) {
        std::swap(x.a, x.b);
        std::swap(x.left, x.right);
        std::swap(x.lf, x.rf);
      }
      F.push_back(x);
    }
  std::sort(F.begin(), F.end(), [](const DGFace &l, const DGFace &r) {
    return l.a != r.a ? l.a < r.a : l.b < r.b;
  });
  for (size_t i = 1; i < F.size(); ++i)
    if (F[i].a == F[i - 1].a && F[i].b == F[i - 1].b) {
      DGFace &P = F[i - 1];
      P.left = std::max(P.left, F[i].left);
      P.right = std::max(P.right, F[i].right);
      P.lf = std::max(P.lf, F[i].lf);
      P.rf = std::max(P.rf, F[i].rf);
    }
  auto e = std::unique(F.begin(), F.end(), [](const DGFace &l, const DGFace &r) {
    return l.a == r.a && l.b == r.b;
  });
  F.erase(e, F.end());
  return F;
}

struct DGSetup {
  int p, n1, nen;
  i64 ncell;
  ElementTables vol, bnd;
  std::vector<double> X;             // dof node coordinates, ncell*nen*3
  std::vector<std::vector<int>> fn;  // face_nodes[f][i]
  std::vector<DGFace> faces;
  std::vector<std::vector<int>> fl, fr;  // per face: left / right local node lists
};

// Element::build_element_faces for TensorProduct topoDim 2
// (element_boundary.cpp:10-43, 82-93): the element node on face f whose face
// coordinate equals boundary node i.
std::vector<std::vector<int>> dg_face_nodes(int p) {
  const int n1 = p + 1, nen = n1 * n1;
  std::vector<double> xl(nen * 3), g(n1), gw(n1);
  local_nodes(2, p, xl.data());
  gauss_lobatto(n1, g.data(), gw.data());
  std::vector<std::vector<int>> out(4, std::vector<int>(n1, -1));
  for (int f = 0; f < 4; ++f)
    for (int a = 0; a < nen; ++a) {
      const double x = xl[3 * a], y = xl[3 * a + 1];
      bool on;
      double sc;
      if (f == 0) { on = std::abs(y + 1) < 1e-6; sc = x; }
      else if (f == 1) { on = std::abs(x - 1) < 1e-6; sc = y; }
      else if (f == 2) { on = std::abs(y - 1) < 1e-6; sc = -x; }
      else { on = std::abs(x + 1) < 1e-6; sc = -y; }
      if (!on) continue;
      for (int i = 0; i < n1; ++i)
        if (std::abs(sc - g[i]) < 1e-6) out[f][i] = a;
    }
  return out;
}

DGSetup dg_setup(int p, const double *corners, i64 ncell, const int32_t *cells) {
  DGSetup S;
  S.p = p;
  S.n1 = p + 1;
  S.nen = S.n1 * S.n1;
  S.ncell = ncell;
  S.vol = make_tables(2, p);
  S.bnd = make_tables(1, p);
  // make_raw_dofs (DoFManager.cpp:176-229): every cell gets its own GLL nodes
  std::vector<double> xl(S.nen * 3);
  local_nodes(2, p, xl.data());
  S.X.resize(ncell * S.nen * 3);
  for (i64 c = 0; c < ncell; ++c) {
    double cc[12];
    for (int k = 0; k < 4; ++k)
      for (int d = 0; d < 3; ++d) cc[3 * k + d] = corners[3 * cells[4 * c + k] + d];
    for (int a = 0; a < S.nen; ++a) interp_corners(2, &xl[3 * a], cc, &S.X[3 * (c * S.nen + a)]);
  }
  S.fn = dg_face_nodes(p);
  S.faces = dg_faces(ncell, cells);
  // match_face_nodes (DoFManager.cpp:231-293)
  for (const DGFace &F : S.faces) {
    std::vector<int> L, R;
    if (F.left < 0) {
      for (int nr : S.fn[F.rf]) { L.push_back(-1); R.push_back(nr); }
    } else if (F.right < 0) {
      for (int nl : S.fn[F.lf]) { L.push_back(nl); R.push_back(-1); }
    } else {
      for (int nl : S.fn[F.lf]) {
        L.push_back(nl);
        const double *xl_ = &S.X[3 * ((i64)F.left * S.nen + nl)];
        for (int nr : S.fn[F.rf]) {
          const double *xr = &S.X[3 * ((i64)F.right * S.nen + nr)];
          const double d0 = xl_[0] - xr[0], d1 = xl_[1] - xr[1], d2 = xl_[2] - xr[2];
          if (std::sqrt(d0 * d0 + d1 * d1 + d2 * d2) < 1.0e-6) {
            R.push_back(nr);
            break;
          }
        }
      }
    }
    S.fl.push_back(L);
    S.fr.push_back(R);
  }
  return S;
}

// Element::face_update<2> (Element.hpp:151-199): unit outward normal
// (J(1,0), -J(0,0)) / |.| and jxw = |.| * w.
void dg_face_update(const DGSetup &S, i64 el, int fq, const std::vector<int> &lnodes, double n[2],
                    double &jxw) {
  double J0 = 0, J1 = 0;
  for (size_t A = 0; A < lnodes.size(); ++A) {
    const double *x = &S.X[3 * (el * S.nen + lnodes[A])];
    const double g = S.bnd.dN[fq * S.n1 + A];
    J0 += x[0] * g;
    J1 += x[1] * g;
  }
  double nx = J1, ny = -J0;
  const double det = std::sqrt(nx * nx + ny * ny);
  n[0] = nx / det;
  n[1] = ny / det;
  jxw = det * S.bnd.w[fq];
}

void dg_residual(const DGSetup &S, const orc_physics *ph, double t, const double *U, double *R,
                 double *R_face) {
  (void)t;
  const int n1 = S.n1, nen = S.nen;
  const double c0 = ph->U[0], c1 = ph->U[1];
  const i64 nd = S.ncell * nen;
  for (i64 i = 0; i < nd; ++i) R[i] = 0;  // GlobalResidual *= 0 (DGAssembly.hpp:43)
  // face loop (DGAssembly.hpp:106-189)
  for (size_t fi = 0; fi < S.faces.size(); ++fi) {
    const DGFace &F = S.faces[fi];
    if (F.left < 0 || F.right < 0) {
      const i64 el = F.left >= 0 ? F.left : F.right;
      const std::vector<int> &fnodes = F.left >= 0 ? S.fl[fi] : S.fr[fi];
      for (int fq = 0; fq < n1; ++fq) {
        double n[2], jxw;
        dg_face_update(S, el, fq, fnodes, n, jxw);
        double u = 0;
        for (int i = 0; i < n1; ++i) u += U[el * nen + fnodes[i]] * S.bnd.N[fq * n1 + i];
        // AdvectionPhysics::BoundaryFlux: outflow only
        const double vn = n[0] * c0 + n[1] * c1;
        const double flux = vn > 0 ? vn * u : 0 * vn * u;
        for (int A = 0; A < n1; ++A) R[el * nen + fnodes[A]] -= S.bnd.N[fq * n1 + A] * flux * jxw;
      }
    } else {
      const std::vector<int> &L = S.fl[fi], &Rn = S.fr[fi];
      for (int fq = 0; fq < n1; ++fq) {
        double nl[2], nr[2], jl, jr;
        dg_face_update(S, F.left, fq, L, nl, jl);
        dg_face_update(S, F.right, fq, Rn, nr, jr);
        double ul = 0, ur = 0;
        for (int i = 0; i < n1; ++i) {
          const double sv = S.bnd.N[fq * n1 + i];
          ul += U[(i64)F.left * nen + L[i]] * sv;
          ur += U[(i64)F.right * nen + Rn[i]] * sv;
        }
        // AdvectionPhysics::Flux (upwind)
        const double vn = nl[0] * c0 + nl[1] * c1;
        const double flux = 0.5 * (vn * (ul + ur) + std::abs(vn) * (ul - ur));
        for (int i = 0; i < n1; ++i) {
          const double sv = S.bnd.N[fq * n1 + i];
          R[(i64)F.left * nen + L[i]] -= sv * flux * jl;
          R[(i64)F.right * nen + Rn[i]] += sv * flux * jr;
        }
      }
    }
  }
  if (R_face)
    for (i64 i = 0; i < nd; ++i) R_face[i] = R[i];
  // element loop (DGAssembly.hpp:194-299): local = M^-1 (R_face + R_vol), R += local
  std::vector<double> loc(nen), M(nen * nen), grad(nen * 2);
  for (i64 e = 0; e < S.ncell; ++e) {
    for (int a = 0; a < nen; ++a) loc[a] = R[e * nen + a];
    std::fill(M.begin(), M.end(), 0.0);
    for (int q = 0; q < S.vol.nqp; ++q) {
      double J[4] = {0, 0, 0, 0};
      for (int A = 0; A < nen; ++A) {
        const double *x = &S.X[3 * (e * nen + A)];
        for (int i = 0; i < 2; ++i)
          for (int j = 0; j < 2; ++j) J[2 * i + j] += x[i] * S.vol.dN[(q * nen + A) * 2 + j];
      }
      double det, Ji[4];
      jac_inv(2, J, det, Ji);
      const double jxw = det * S.vol.w[q];
      for (int A = 0; A < nen; ++A) {
        const double g0 = S.vol.dN[(q * nen + A) * 2], g1 = S.vol.dN[(q * nen + A) * 2 + 1];
        grad[2 * A] = Ji[0] * g0 + Ji[2] * g1;      // (J^-T g)_0
        grad[2 * A + 1] = Ji[1] * g0 + Ji[3] * g1;  // (J^-T g)_1
      }
      // LocalResidual (main.cpp:53-71)
      double u = 0;
      for (int A = 0; A < nen; ++A) u += U[e * nen + A] * S.vol.N[q * nen + A];
      for (int i = 0; i < nen; ++i) {
        const double val = u * (grad[2 * i] * c0 + grad[2 * i + 1] * c1) + ph->f * S.vol.N[q * nen + i];
        loc[i] += val * jxw;
      }
      // LocalMass (main.cpp:73-78)
      for (int i = 0; i < nen; ++i)
        for (int j = 0; j < nen; ++j) M[i * nen + j] += (jxw * S.vol.N[q * nen + i]) * S.vol.N[q * nen + j];
    }
    lu_solve(nen, M, 1, loc);
    for (int a = 0; a < nen; ++a) R[e * nen + a] += loc[a];
  }
}

}  // namespace

extern "C" {

int64_t orc_dg_faces(int64_t ncell, const int32_t *cells, int32_t *left, int32_t *right,
                     int32_t *left_flocal, int32_t *right_flocal) {
  auto F = dg_faces(ncell, cells);
  for (size_t i = 0; i < F.size(); ++i) {
    left[i] = F[i].left;
    right[i] = F[i].right;
    left_flocal[i] = F[i].lf;
    right_flocal[i] = F[i].rf;
  }
  return (int64_t)F.size();
}

void orc_dg_nodes(int p, const double *corner_nodes, int64_t ncell, const int32_t *cells,
                  double *dof_nodes) {
  DGSetup S = dg_setup(p, corner_nodes, ncell, cells);
  std::memcpy(dof_nodes, S.X.data(), S.X.size() * sizeof(double));
}

void orc_dg_residual(int p, const double *corner_nodes, int64_t ncell, const int32_t *cells,
                     const orc_physics *phys, double t, const double *U, double *R, double *R_face) {
  DGSetup S = dg_setup(p, corner_nodes, ncell, cells);
  dg_residual(S, phys, t, U, R, R_face);
}

void orc_dg_rk4(int p, const double *corner_nodes, int64_t ncell, const int32_t *cells,
                const orc_physics *phys, double dt, int64_t nsteps, double t0, double *U) {
  DGSetup S = dg_setup(p, corner_nodes, ncell, cells);
  const i64 nd = ncell * S.nen;
  std::vector<double> U0(nd), k1(nd), k2(nd), k3(nd), k4(nd), W(nd);
  double t = t0;
  for (i64 s = 0; s < nsteps; ++s) {
    for (i64 i = 0; i < nd; ++i) U0[i] = U[i];
    dg_residual(S, phys, t, U0.data(), k1.data(), nullptr);
    for (i64 i = 0; i < nd; ++i) W[i] = U0[i] + (dt / 2) * k1[i];
    dg_residual(S, phys, t + dt / 2, W.data(), k2.data(), nullptr);
    for (i64 i = 0; i < nd; ++i) W[i] = U0[i] + (dt / 2) * k2[i];
    dg_residual(S, phys, t + dt / 2, W.data(), k3.data(), nullptr);
    for (i64 i = 0; i < nd; ++i) W[i] = U0[i] + dt * k3[i];
    dg_residual(S, phys, t + dt, W.data(), k4.data(), nullptr);
    for (i64 i = 0; i < nd; ++i)
      U[i] = U0[i] + (dt / 6) * k1[i] + (dt / 3) * k2[i] + (dt / 3) * k3[i] + (dt / 6) * k4[i];
    t += dt;
  }
}

uint64_t orc_splitmix64(uint64_t *state) {
  uint64_t z = (*state += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

void orc_normal_fill(uint64_t seed, int64_t n, double mean, double sd, double *out) {
  uint64_t st = seed;
  const double two_pi = 6.283185307179586476925286766559;
  for (int64_t i = 0; i < n; i += 2) {
    double u1 = ((orc_splitmix64(&st) >> 11) + 1) * (1.0 / 9007199254740992.0);
    double u2 = (orc_splitmix64(&st) >> 11) * (1.0 / 9007199254740992.0);
    double r = std::sqrt(-2.0 * std::log(u1));
    out[i] = mean + sd * r * std::cos(two_pi * u2);
    if (i + 1 < n) out[i + 1] = mean + sd * r * std::sin(two_pi * u2);
  }
}

}  // extern "C"

/* ------------------------------------------------------------------------ */
/* Gradient recovery: include/assembly/LocalSmoothingGradient.hpp:20-94 and  */
/* include/assembly/ZZGradientRecovery.hpp:33-141.                            */
/* ------------------------------------------------------------------------ */
namespace {

// Element::update at one qp (Element.hpp:109-149): jxw and shapeGrad (nen x nsd)
void element_update(const ElementTables &T, const double *X, const int32_t *conn, int q, double &jxw,
                    std::vector<double> &G) {
  const int nsd = T.dim, nen = T.nen;
  double J[9] = {0};
  for (int A = 0; A < nen; ++A) {
    const double *x = &X[3 * (i64)conn[A]];
    for (int i = 0; i < nsd; ++i)
      for (int j = 0; j < nsd; ++j) J[i * nsd + j] += x[i] * T.dN[(q * nen + A) * nsd + j];
  }
  double det, Ji[9];
  jac_inv(nsd, J, det, Ji);
  jxw = det * T.w[q];
  for (int A = 0; A < nen; ++A)
    for (int d = 0; d < nsd; ++d) {
      double s = 0;
      for (int k = 0; k < nsd; ++k) s += Ji[k * nsd + d] * T.dN[(q * nen + A) * nsd + k];
      G[A * nsd + d] = s;
    }
}

// Least squares min ||P a - b|| (P m x n, full column rank, row-major; b m x
// nr) by Householder QR with column pivoting (Businger & Golub 1965), the
// algorithm of Eigen's ColPivHouseholderQR that ZZGradientRecovery.hpp:127
// calls: at step k the remaining column of largest norm is swapped in, a
// reflector zeroes it below the diagonal and is applied to the rest of P and
// to b; R a = Q^T b is then solved by back substitution and un-permuted.
void lsq_colpiv_qr(int m, int n, std::vector<double> P, int nr, std::vector<double> b, double *a) {
  std::vector<int> perm(n);
  std::iota(perm.begin(), perm.end(), 0);
  for (int k = 0; k < n && k < m; ++k) {
    int best = k;
    double bn = -1;
    for (int j = k; j < n; ++j) {
      double s = 0;
      for (int i = k; i < m; ++i) s += P[i * n + j] * P[i * n + j];
      if (s > bn) { bn = s; best = j; }
    }
    if (best != k) {
      for (int i = 0; i < m; ++i) std::swap(P[i * n + k], P[i * n + best]);
      std::swap(perm[k], perm[best]);
    }
    double nrm = std::sqrt(bn);
    if (nrm == 0) continue;
    const double alpha = P[k * n + k] > 0 ? -nrm : nrm;
    std::vector<double> v(m, 0.0);
    for (int i = k; i < m; ++i) v[i] = P[i * n + k];
    v[k] -= alpha;
    double vv = 0;
    for (int i = k; i < m; ++i) vv += v[i] * v[i];
    if (vv == 0) continue;
    for (int j = k; j < n; ++j) {
      double s = 0;
      for (int i = k; i < m; ++i) s += v[i] * P[i * n + j];
      s = 2 * s / vv;
      for (int i = k; i < m; ++i) P[i * n + j] -= s * v[i];
    }
    for (int c = 0; c < nr; ++c) {
      double s = 0;
      for (int i = k; i < m; ++i) s += v[i] * b[i * nr + c];
      s = 2 * s / vv;
      for (int i = k; i < m; ++i) b[i * nr + c] -= s * v[i];
    }
  }
  std::vector<double> y(n * nr);
  for (int c = 0; c < nr; ++c)
    for (int r = n - 1; r >= 0; --r) {
      double s = b[r * nr + c];
      for (int j = r + 1; j < n; ++j) s -= P[r * n + j] * y[j * nr + c];
      y[r * nr + c] = s / P[r * n + r];
    }
  for (int j = 0; j < n; ++j)
    for (int c = 0; c < nr; ++c) a[perm[j] * nr + c] = y[j * nr + c];
}

}  // namespace

extern "C" {

// LocalSmoothingGradient.hpp:20-94: element loop, qp loop; qp_grad(comp, i)
// = sum_A u(A) shapeGrad(node(A), i) over local dofs A (node-major); then
// Volume(g(A)) += N jxw and VolTimesGrad(g(A), i) += N qp_grad(comp, i) jxw;
// finally SolutionGradient(A, i) = VolTimesGrad(A, i) / Volume(A).
void orc_grad_local_smoothing(int dim, int p, int dpn, int64_t n_nodes, const double *dof_nodes, int64_t nel,
                              const int32_t *elements, const double *u, double *grad) {
  ElementTables T = make_tables(dim, p);
  const int nsd = dim, nen = T.nen, nld = nen * dpn;
  const i64 ndofs = n_nodes * dpn;
  std::vector<double> vtg(ndofs * nsd, 0.0), vol(ndofs, 0.0), G(nen * nsd), qg(dpn * nsd), ul(nld);
  for (i64 e = 0; e < nel; ++e) {
    const int32_t *conn = &elements[e * nen];
    for (int A = 0; A < nld; ++A) ul[A] = u[(i64)conn[A / dpn] * dpn + A % dpn];
    for (int q = 0; q < T.nqp; ++q) {
      std::fill(qg.begin(), qg.end(), 0.0);
      double jxw;
      element_update(T, dof_nodes, conn, q, jxw, G);
      for (int A = 0; A < nld; ++A) {
        const int comp = A % dpn, node = A / dpn;
        for (int i = 0; i < nsd; ++i) qg[comp * nsd + i] += ul[A] * G[node * nsd + i];
      }
      for (int A = 0; A < nld; ++A) {
        const int comp = A % dpn, node = A / dpn;
        const i64 g = (i64)conn[node] * dpn + comp;
        const double N = T.N[q * nen + node];
        vol[g] += N * jxw;
        for (int i = 0; i < nsd; ++i) vtg[g * nsd + i] += N * qg[comp * nsd + i] * jxw;
      }
    }
  }
  for (i64 A = 0; A < ndofs; ++A)
    for (int i = 0; i < nsd; ++i) grad[A * nsd + i] = vtg[A * nsd + i] / vol[A];
}

// ZZGradientRecovery.hpp:33-141 (p = 1, scalar field): node -> adjacent
// element sets (ascending), samples at every qp of the patch, P = [1 x (y (z))],
// b = grad u at the qp, a = colPivHouseholderQr(P).solve(b), grad(n) = [1 x_n] a.
void orc_grad_zz(int dim, int64_t n_nodes, const double *dof_nodes, int64_t nel, const int32_t *elements,
                 const double *u, double *grad) {
  ElementTables T = make_tables(dim, 1);
  const int nsd = dim, nen = T.nen, nb = nsd + 1;
  std::vector<std::vector<i64>> adj(n_nodes);
  for (i64 e = 0; e < nel; ++e)
    for (int A = 0; A < nen; ++A) {
      auto &v = adj[elements[e * nen + A]];
      if (v.empty() || v.back() != e) v.push_back(e);
    }
  std::vector<double> G(nen * nsd);
  for (i64 nd = 0; nd < n_nodes; ++nd) {
    const auto &ae = adj[nd];
    const int m = (int)ae.size() * T.nqp;
    if (m == 0) {
      // a node of no element: Eigen's solve of the empty system is zero
      for (int i = 0; i < nsd; ++i) grad[nd * nsd + i] = 0;
      continue;
    }
    std::vector<double> P(m * nb), b(m * nsd);
    int idx = 0;
    for (i64 e : ae) {
      const int32_t *conn = &elements[e * nen];
      for (int q = 0; q < T.nqp; ++q) {
        double jxw, xq[3] = {0, 0, 0}, fg[3] = {0, 0, 0};
        element_update(T, dof_nodes, conn, q, jxw, G);
        for (int A = 0; A < nen; ++A) {
          for (int d = 0; d < 3; ++d) xq[d] += T.N[q * nen + A] * dof_nodes[3 * (i64)conn[A] + d];
          for (int d = 0; d < nsd; ++d) fg[d] += u[conn[A]] * G[A * nsd + d];
        }
        P[idx * nb] = 1;
        for (int i = 0; i < nsd; ++i) {
          P[idx * nb + i + 1] = xq[i];
          b[idx * nsd + i] = fg[i];
        }
        ++idx;
      }
    }
    std::vector<double> a(nb * nsd);
    lsq_colpiv_qr(m, nb, P, nsd, b, a.data());
    for (int i = 0; i < nsd; ++i) {
      double g = a[i];
      for (int j = 0; j < nsd; ++j) g += dof_nodes[3 * nd + j] * a[(j + 1) * nsd + i];
      grad[nd * nsd + i] = g;
    }
  }
}

}  // extern "C"
