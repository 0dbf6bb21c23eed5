// Gradient recovery after a solve, on the device:
//  - LocalSmoothingGradient (include/assembly/LocalSmoothingGradient.hpp:20-94):
//    grad(dof, i) = sum_{e,q} N_a jxw (sum_B u_B dN_B/dx_i) / sum_{e,q} N_a jxw
//    over the elements e around the dof's node (a = its local node).
//  - ZZGradientRecovery (include/assembly/ZZGradientRecovery.hpp:33-141), p = 1:
//    at each node, a least-squares linear fit g(x) = a0 + a.x of the gradient
//    samples at the quadrature points of the adjacent elements, evaluated at
//    the node.
//
// Both run as an element pass and a node pass.  The element pass stages
// per-element values (per (element, local node) weighted sums, or per
// (element, qp) samples); the node pass walks the node's incidence list
// (element order, the residual assembly's plan) -- deterministic, no atomics.
// ZZ fits in coordinates centred on the node: the same least-squares
// solution, evaluated at the node as its constant term, from at most 4 x 4
// normal equations whose conditioning no longer depends on where the node is.
#include <algorithm>

#include "internal.hpp"
#include "mesh.hpp"

using namespace yh;

namespace {

constexpr int NT = 256;

template <int DIM, int P>
struct RShape {
  static constexpr int N1 = P + 1;
  static constexpr int NEN = DIM == 2 ? N1 * N1 : N1 * N1 * N1;
  static constexpr int NQP = NEN;
  static constexpr int EPB = NT / NEN > 0 ? NT / NEN : 1;
};

// J, det and J^-1 at one quadrature point from the element's coordinates;
// the same operation order as the assembly's element pass (Element::update,
// include/element/Element.hpp:116-147).
template <int DIM>
__device__ __forceinline__ void qp_geometry(const double (*sX)[DIM], const double *__restrict__ tabdN, int q,
                                            int nen, double &det, double *Ji) {
  double J[DIM * DIM];
#pragma unroll
  for (int k = 0; k < DIM * DIM; ++k) J[k] = 0;
  for (int A = 0; A < nen; ++A) {
    const double *dn = &tabdN[(q * nen + A) * DIM];
#pragma unroll
    for (int r = 0; r < DIM; ++r)
#pragma unroll
      for (int c = 0; c < DIM; ++c) J[r * DIM + c] += sX[A][r] * dn[c];
  }
  if (DIM == 2) {
    det = J[0] * J[3] - J[1] * J[2];
    Ji[0] = J[3] / det;
    Ji[1] = -J[1] / det;
    Ji[2] = -J[2] / det;
    Ji[3] = J[0] / det;
  } else {
    const double a = J[0], b = J[1], c = J[2], d = J[3], e = J[4], f = J[5], g = J[6], h = J[7], i = J[8];
    det = a * (e * i - f * h) - b * (d * i - f * g) + c * (d * h - e * g);
    Ji[0] = (e * i - f * h) / det;
    Ji[1] = (c * h - b * i) / det;
    Ji[2] = (b * f - c * e) / det;
    Ji[3] = (f * g - d * i) / det;
    Ji[4] = (a * i - c * g) / det;
    Ji[5] = (c * d - a * f) / det;
    Ji[6] = (d * h - e * g) / det;
    Ji[7] = (b * g - a * h) / det;
    Ji[8] = (a * e - b * d) / det;
  }
}

// grad_x N_A = J^-T grad_xi N_A (Element.hpp:131-147)
template <int DIM>
__device__ __forceinline__ void shape_grad(const double *Ji, const double *dn, double *g) {
#pragma unroll
  for (int d = 0; d < DIM; ++d) {
    double s = 0;
#pragma unroll
    for (int k = 0; k < DIM; ++k) s += Ji[k * DIM + d] * dn[k];
    g[d] = s;
  }
}

// --------------------------------------------------------- local smoothing
// Element pass: a thread per (element, qp) forms jxw and the field gradient
// qg[c][i] = sum_B u(B, c) dN_B/dx_i (LocalSmoothingGradient.hpp:56-67) into
// LDS; then a thread per (element, local node a) sums over the qps
//   S[(e*NEN + a)*W + 0]            = sum_q N_a jxw
//   S[(e*NEN + a)*W + 1 + c*DIM + i] = sum_q N_a qg[c][i] jxw,  W = 1 + DPN*DIM.
template <int DIM, int P, int DPN>
__global__ __launch_bounds__(NT) void k_smooth_elem(int64_t nel, const int32_t *__restrict__ el,
                                                    const double *__restrict__ X, const double *__restrict__ tabN,
                                                    const double *__restrict__ tabdN,
                                                    const double *__restrict__ tabw, const double *__restrict__ U,
                                                    double *__restrict__ S) {
  using RS = RShape<DIM, P>;
  constexpr int NEN = RS::NEN, NQP = RS::NQP, EPB = RS::EPB, W = 1 + DPN * DIM;
  __shared__ double sX[EPB][NEN][DIM];
  __shared__ double sU[EPB][NEN][DPN];
  __shared__ double sQ[EPB][NQP][DPN * DIM + 1];  // qg, then jxw
  const int64_t e0 = (int64_t)blockIdx.x * EPB;
  for (int i = threadIdx.x; i < EPB * NEN; i += NT) {
    const int le = i / NEN, a = i - le * NEN;
    const int64_t e = e0 + le;
    if (e < nel) {
      const int64_t node = el[e * NEN + a];
#pragma unroll
      for (int d = 0; d < DIM; ++d) sX[le][a][d] = X[3 * node + d];
#pragma unroll
      for (int c = 0; c < DPN; ++c) sU[le][a][c] = U[node * DPN + c];
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < EPB * NQP; i += NT) {
    const int le = i / NQP, q = i - le * NQP;
    if (e0 + le >= nel) continue;
    double det, Ji[DIM * DIM];
    qp_geometry<DIM>(sX[le], tabdN, q, NEN, det, Ji);
    double qg[DPN][DIM];
#pragma unroll
    for (int c = 0; c < DPN; ++c)
#pragma unroll
      for (int d = 0; d < DIM; ++d) qg[c][d] = 0;
    for (int B = 0; B < NEN; ++B) {
      double g[DIM];
      shape_grad<DIM>(Ji, &tabdN[(q * NEN + B) * DIM], g);
#pragma unroll
      for (int c = 0; c < DPN; ++c)
#pragma unroll
        for (int d = 0; d < DIM; ++d) qg[c][d] += sU[le][B][c] * g[d];
    }
#pragma unroll
    for (int c = 0; c < DPN; ++c)
#pragma unroll
      for (int d = 0; d < DIM; ++d) sQ[le][q][c * DIM + d] = qg[c][d];
    sQ[le][q][DPN * DIM] = det * tabw[q];
  }
  __syncthreads();
  for (int i = threadIdx.x; i < EPB * NEN; i += NT) {
    const int le = i / NEN, a = i - le * NEN;
    const int64_t e = e0 + le;
    if (e >= nel) continue;
    double v = 0, vg[DPN * DIM];
#pragma unroll
    for (int k = 0; k < DPN * DIM; ++k) vg[k] = 0;
    for (int q = 0; q < NQP; ++q) {
      const double Na = tabN[q * NEN + a], jxw = sQ[le][q][DPN * DIM];
      v += Na * jxw;
#pragma unroll
      for (int k = 0; k < DPN * DIM; ++k) vg[k] += Na * sQ[le][q][k] * jxw;
    }
    double *o = S + (e * NEN + a) * W;
    o[0] = v;
#pragma unroll
    for (int k = 0; k < DPN * DIM; ++k) o[1 + k] = vg[k];
  }
}

// Node pass: grad(node*DPN + c, i) = (sum over the node's incidences of the
// staged weighted gradient) / (sum of the staged volume), in element order.
template <int DIM, int DPN>
__global__ __launch_bounds__(NT) void k_smooth_node(int64_t nnodes, const int32_t *__restrict__ ptr,
                                                    const uint32_t *__restrict__ inc, const double *__restrict__ S,
                                                    double *__restrict__ G) {
  constexpr int W = 1 + DPN * DIM;
  const int64_t A = blockIdx.x * (int64_t)NT + threadIdx.x;
  if (A >= nnodes) return;
  double v = 0, vg[DPN * DIM];
#pragma unroll
  for (int k = 0; k < DPN * DIM; ++k) vg[k] = 0;
  for (int k = ptr[A]; k < ptr[A + 1]; ++k) {
    const double *s = S + (int64_t)inc[k] * W;
    v += s[0];
#pragma unroll
    for (int j = 0; j < DPN * DIM; ++j) vg[j] += s[1 + j];
  }
#pragma unroll
  for (int c = 0; c < DPN; ++c)
#pragma unroll
    for (int d = 0; d < DIM; ++d) G[(A * DPN + c) * DIM + d] = vg[c * DIM + d] / v;
}

// ------------------------------------------------------------- ZZ recovery
// Element pass (p = 1): a thread per (element, qp) writes the sample
//   S[(e*NQP + q)*2*DIM ..] = (x_qp, grad u(x_qp))
// with x_qp = sum_A N_A x_A and grad u = sum_A u_A grad N_A
// (ZZGradientRecovery.hpp:95-108).
template <int DIM>
__global__ __launch_bounds__(NT) void k_zz_samples(int64_t nel, const int32_t *__restrict__ el,
                                                   const double *__restrict__ X, const double *__restrict__ tabN,
                                                   const double *__restrict__ tabdN, const double *__restrict__ U,
                                                   double *__restrict__ S) {
  constexpr int NEN = DIM == 2 ? 4 : 8, NQP = NEN;
  const int64_t t = blockIdx.x * (int64_t)NT + threadIdx.x;
  if (t >= nel * NQP) return;
  const int64_t e = t / NQP;
  const int q = (int)(t - e * NQP);
  double sX[NEN][DIM], u[NEN];
  for (int A = 0; A < NEN; ++A) {
    const int64_t node = el[e * NEN + A];
#pragma unroll
    for (int d = 0; d < DIM; ++d) sX[A][d] = X[3 * node + d];
    u[A] = U[node];
  }
  double det, Ji[DIM * DIM];
  qp_geometry<DIM>(sX, tabdN, q, NEN, det, Ji);
  double xq[DIM], gq[DIM];
#pragma unroll
  for (int d = 0; d < DIM; ++d) xq[d] = gq[d] = 0;
  for (int A = 0; A < NEN; ++A) {
    const double Na = tabN[q * NEN + A];
    double g[DIM];
    shape_grad<DIM>(Ji, &tabdN[(q * NEN + A) * DIM], g);
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
      xq[d] += Na * sX[A][d];
      gq[d] += u[A] * g[d];
    }
  }
  double *o = S + t * 2 * DIM;
#pragma unroll
  for (int d = 0; d < DIM; ++d) {
    o[d] = xq[d];
    o[DIM + d] = gq[d];
  }
}

// Node pass: the least-squares fit of the patch's samples with basis
// (1, x - x_n) by its normal equations, solved by Gaussian elimination with
// partial pivoting (which a scaling of the offset columns would not change
// beyond rounding); the recovered gradient is the fit's constant term
// (ZZGradientRecovery.hpp:112-134).
template <int DIM>
__global__ __launch_bounds__(NT) void k_zz_node(int64_t nnodes, const double *__restrict__ X,
                                                const int32_t *__restrict__ ptr, const uint32_t *__restrict__ inc,
                                                const double *__restrict__ S, double *__restrict__ G) {
  constexpr int NEN = DIM == 2 ? 4 : 8, NQP = NEN, NB = DIM + 1;
  const int64_t A = blockIdx.x * (int64_t)NT + threadIdx.x;
  if (A >= nnodes) return;
  if (ptr[A] == ptr[A + 1]) {
    // a node of no element: an empty fit, whose minimum-norm solution is 0
#pragma unroll
    for (int d = 0; d < DIM; ++d) G[A * DIM + d] = 0.0;
    return;
  }
  double xn[DIM];
#pragma unroll
  for (int d = 0; d < DIM; ++d) xn[d] = X[3 * A + d];
  double M[NB][NB + DIM];  // [P^T P | P^T b]
#pragma unroll
  for (int r = 0; r < NB; ++r)
#pragma unroll
    for (int c = 0; c < NB + DIM; ++c) M[r][c] = 0;
  for (int k = ptr[A]; k < ptr[A + 1]; ++k) {
    const int64_t e = inc[k] / NEN;
    for (int q = 0; q < NQP; ++q) {
      const double *s = S + (e * NQP + q) * 2 * DIM;
      double pr[NB];
      pr[0] = 1.0;
#pragma unroll
      for (int d = 0; d < DIM; ++d) pr[1 + d] = s[d] - xn[d];
#pragma unroll
      for (int r = 0; r < NB; ++r) {
#pragma unroll
        for (int c = 0; c < NB; ++c) M[r][c] += pr[r] * pr[c];
#pragma unroll
        for (int d = 0; d < DIM; ++d) M[r][NB + d] += pr[r] * s[DIM + d];
      }
    }
  }
  // forward elimination with partial pivoting, then back substitution
#pragma unroll
  for (int c = 0; c < NB; ++c) {
    int piv = c;
#pragma unroll
    for (int r = c + 1; r < NB; ++r)
      if (fabs(M[r][c]) > fabs(M[piv][c])) piv = r;
    if (piv != c) {
#pragma unroll
      for (int j = 0; j < NB + DIM; ++j) {
        const double t = M[c][j];
        M[c][j] = M[piv][j];
        M[piv][j] = t;
      }
    }
#pragma unroll
    for (int r = c + 1; r < NB; ++r) {
      const double f = M[r][c] / M[c][c];
#pragma unroll
      for (int j = c; j < NB + DIM; ++j) M[r][j] -= f * M[c][j];
    }
  }
  double a[NB][DIM];
#pragma unroll
  for (int r = NB - 1; r >= 0; --r)
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
      double s = M[r][NB + d];
#pragma unroll
      for (int c = r + 1; c < NB; ++c) s -= M[r][c] * a[c][d];
      a[r][d] = s / M[r][r];
    }
#pragma unroll
  for (int d = 0; d < DIM; ++d) G[A * DIM + d] = a[0][d];
}

template <int DIM, int P, int DPN>
void launch_smooth(yafel_hip_mesh_s *m, const double *U, double *S, double *G) {
  using RS = RShape<DIM, P>;
  hipStream_t s = m->ctx->stream;
  k_smooth_elem<DIM, P, DPN><<<ceil_div(m->nel, RS::EPB), NT, 0, s>>>(
      m->nel, m->elements.as<int32_t>(), m->nodes.as<double>(), m->tabN.as<double>(), m->tabdN.as<double>(),
      m->tabw.as<double>(), U, S);
  HIPCHK(hipGetLastError());
  k_smooth_node<DIM, DPN><<<ceil_div(m->n_nodes, NT), NT, 0, s>>>(m->n_nodes, m->inc_ptr.as<int32_t>(),
                                                                  m->inc.as<uint32_t>(), S, G);
  HIPCHK(hipGetLastError());
}

template <int DIM, int P>
void launch_smooth_p(yafel_hip_mesh_s *m, const double *U, double *S, double *G) {
  switch (m->dpn) {
    case 1: launch_smooth<DIM, P, 1>(m, U, S, G); break;
    case 2: launch_smooth<DIM, P, 2>(m, U, S, G); break;
    default: launch_smooth<DIM, P, 3>(m, U, S, G); break;
  }
}

}  // namespace

extern "C" {

int yafel_hip_gradient_local_smoothing(yafel_hip_mesh m, const double *u_dev, double *grad_dev) {
  YH_TRY_BEGIN
  REQUIRE(m, "gradient_local_smoothing: null mesh");
  REQUIRE(m->dpn >= 1 && m->dpn <= 3, "gradient_local_smoothing: 1 to 3 dofs per node");
  if (m->n_nodes == 0) return YAFEL_HIP_OK;
  REQUIRE(u_dev && grad_dev, "gradient_local_smoothing: null vector");
  yafel_hip_ctx_s *ctx = m->ctx;
  HIPCHK(hipSetDevice(ctx->device));
  mesh_build_incidence(m);
  const size_t need = (size_t)m->nel * m->nen * (1 + m->dpn * m->dim) * 8 + 8;
  if (m->stageG.bytes < need) m->stageG.alloc(need);
  double *sp = m->stageG.as<double>();
  if (m->dim == 2) {
    switch (m->p) {
      case 1: launch_smooth_p<2, 1>(m, u_dev, sp, grad_dev); break;
      case 2: launch_smooth_p<2, 2>(m, u_dev, sp, grad_dev); break;
      case 3: launch_smooth_p<2, 3>(m, u_dev, sp, grad_dev); break;
      case 4: launch_smooth_p<2, 4>(m, u_dev, sp, grad_dev); break;
      default: fail(YAFEL_HIP_EINVAL, "gradient_local_smoothing: 2-D elements are supported for p = 1..4");
    }
  } else {
    switch (m->p) {
      case 1: launch_smooth_p<3, 1>(m, u_dev, sp, grad_dev); break;
      case 2: launch_smooth_p<3, 2>(m, u_dev, sp, grad_dev); break;
      default: fail(YAFEL_HIP_EINVAL, "gradient_local_smoothing: 3-D elements are supported for p = 1..2");
    }
  }
  return YAFEL_HIP_OK;
  YH_TRY_END(m ? m->ctx : nullptr)
}

int yafel_hip_gradient_zz(yafel_hip_mesh m, const double *u_dev, double *grad_dev) {
  YH_TRY_BEGIN
  REQUIRE(m, "gradient_zz: null mesh");
  REQUIRE(m->p == 1, "gradient_zz: ZZ recovery is for p = 1 elements");
  REQUIRE(m->dpn == 1, "gradient_zz: ZZ recovery fits a scalar field (1 dof per node)");
  if (m->n_nodes == 0) return YAFEL_HIP_OK;
  REQUIRE(u_dev && grad_dev, "gradient_zz: null vector");
  yafel_hip_ctx_s *ctx = m->ctx;
  HIPCHK(hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  mesh_build_incidence(m);
  const size_t need = (size_t)m->nel * m->nqp * 2 * m->dim * 8 + 8;
  if (m->stageG.bytes < need) m->stageG.alloc(need);
  double *sp = m->stageG.as<double>();
  const int64_t ns = m->nel * m->nqp;
  if (m->dim == 2) {
    k_zz_samples<2><<<ceil_div(ns, NT), NT, 0, s>>>(m->nel, m->elements.as<int32_t>(), m->nodes.as<double>(),
                                                    m->tabN.as<double>(), m->tabdN.as<double>(), u_dev, sp);
    HIPCHK(hipGetLastError());
    k_zz_node<2><<<ceil_div(m->n_nodes, NT), NT, 0, s>>>(m->n_nodes, m->nodes.as<double>(), m->inc_ptr.as<int32_t>(),
                                                         m->inc.as<uint32_t>(), sp, grad_dev);
  } else {
    k_zz_samples<3><<<ceil_div(ns, NT), NT, 0, s>>>(m->nel, m->elements.as<int32_t>(), m->nodes.as<double>(),
                                                    m->tabN.as<double>(), m->tabdN.as<double>(), u_dev, sp);
    HIPCHK(hipGetLastError());
    k_zz_node<3><<<ceil_div(m->n_nodes, NT), NT, 0, s>>>(m->n_nodes, m->nodes.as<double>(), m->inc_ptr.as<int32_t>(),
                                                         m->inc.as<uint32_t>(), sp, grad_dev);
  }
  HIPCHK(hipGetLastError());
  return YAFEL_HIP_OK;
  YH_TRY_END(m ? m->ctx : nullptr)
}

}  // extern "C"
