// CGAssembly on the device (include/assembly/CGAssembly.hpp:27-186) and
// DirichletBC::apply (src/boundary_conditions/DirichletBC.cpp:27-56).
//
// Assembly is two passes:
//  1. element pass: a workgroup takes EPB elements; their node coordinates go
//     to LDS, one thread per (element, quadrature point) forms J, det J, J^-1
//     and the physical gradients of every shape function into LDS
//     (Element::update, include/element/Element.hpp:109-149), then one thread
//     per (element, local node A) accumulates row block A of K_e over the
//     quadrature points with the physics formula, and R_e(A).  K_e / R_e go to
//     a staging buffer, row-major per element (the local_tangent layout).
//  2. gather pass: one thread per stored node pair sums its contributions in
//     element order from the plan built with the pattern -- deterministic, no
//     atomics.  The residual is gathered per node the same way and ADDED to R
//     (the reference accumulates GlobalResidual).
#include <algorithm>

#include "internal.hpp"
#include "kernels.hpp"
#include "mesh.hpp"
#include "sortprim.hpp"

using namespace yh;

namespace {

constexpr int NT = 256;

struct Phys {
  int kind;
  double f, lambda, mu;
  double body[3], U[3];
  double kdiff, h, alpha, Q;
  // derived SUPG terms
  double Kt[9];
  double Uh[3];
};

template <int DIM, int P>
struct ElemShape {
  static constexpr int N1 = P + 1;
  static constexpr int NEN = DIM == 2 ? N1 * N1 : N1 * N1 * N1;
  static constexpr int NQP = NEN;
  static constexpr int LDS_PER = (NQP * NEN * DIM + 1) * 8 + (NQP + 1) * 8 + NEN * DIM * 8;
  static constexpr int EPB0 = NT / NEN;
  static constexpr int EPB1 = 60000 / LDS_PER;
  static constexpr int EPB = (EPB0 < EPB1 ? EPB0 : EPB1) > 0 ? (EPB0 < EPB1 ? EPB0 : EPB1) : 1;
  // symmetric staging component-major across elements when a workgroup holds
  // enough elements for its stores to form long runs (Q1 hexes and quads,
  // Q2 quads); element-major otherwise (a Q4 element's 325 blocks are
  // contiguous there, and a workgroup holds 5 elements)
  static constexpr bool SOA = EPB >= 16;
  // per element: gradients and jxw, with component-major staging each padded
  // by one double so that threads of consecutive elements read different
  // LDS banks
  static constexpr int GP = NQP * NEN * DIM + (SOA ? 1 : 0), WP = NQP + (SOA ? 1 : 0);
};

// whether the symmetric staging of a (dim, p) mesh is component-major
inline bool sym_soa(int dim, int p) {
  if (dim == 2) {
    switch (p) {
      case 1: return ElemShape<2, 1>::SOA;
      case 2: return ElemShape<2, 2>::SOA;
      case 3: return ElemShape<2, 3>::SOA;
      default: return ElemShape<2, 4>::SOA;
    }
  }
  return p == 1 ? ElemShape<3, 1>::SOA : ElemShape<3, 2>::SOA;
}

template <int DIM>
__device__ __forceinline__ void jac_inv(const double *J, double &det, double *Ji) {
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

// Block (A, B) of the element tangent, k[a][b] = K_e[A*DPN + a][B*DPN + b],
// from the physical gradients and jxw staged in LDS.
template <int DIM, int DPN, int NEN, int NQP>
__device__ __forceinline__ void tangent_block(const Phys &ph, double ke, const double (*sG)[NEN][DIM],
                                              const double *sW, const double *__restrict__ tabN, int A, int B,
                                              double (&k)[DPN][DPN]) {
#pragma unroll
  for (int a = 0; a < DPN; ++a)
#pragma unroll
    for (int b = 0; b < DPN; ++b) k[a][b] = 0;
  for (int q = 0; q < NQP; ++q) {
    const double jxw = sW[q];
    const double *gA = sG[q][A];
    const double *gB = sG[q][B];
    if (ph.kind == YAFEL_HIP_PHYS_POISSON) {
      // apps/BasicPoisson/main.cpp:51-58
      double dt = 0;
#pragma unroll
      for (int d = 0; d < DIM; ++d) dt += gA[d] * gB[d];
      k[0][0] -= ke * dt * jxw;
    } else if (ph.kind == YAFEL_HIP_PHYS_ELASTICITY) {
      // apps/BasicLinearElasticity/main.cpp:70-110: sum_jl C_ijkl dN_A,j
      // dN_B,l with C_ijkl = lam d_ij d_kl + mu (d_ik d_jl + d_il d_jk),
      // contracted in closed form (the tensor's zero terms are not formed)
      const double lam = ke * ph.lambda, mu = ke * ph.mu;
      double dab = 0;
#pragma unroll
      for (int d = 0; d < DIM; ++d) dab += gA[d] * gB[d];
#pragma unroll
      for (int ii = 0; ii < DPN; ++ii)
#pragma unroll
        for (int kk = 0; kk < DPN; ++kk) {
          const double s = lam * (gA[ii] * gB[kk]) + mu * (gA[kk] * gB[ii] + (ii == kk ? dab : 0.0));
          k[ii][kk] -= s * jxw;
        }
    } else {
      // apps/supg_learning/supg_advection.hpp:90-103 (advection added once per A component)
      const double NA = tabN[q * NEN + A];
      double s = 0;
#pragma unroll
      for (int Ac = 0; Ac < DIM; ++Ac)
#pragma unroll
        for (int Bc = 0; Bc < DIM; ++Bc)
          s += (NA * ph.U[Bc] * gB[Bc] + gA[Ac] * gB[Bc] * ph.Kt[Ac * DIM + Bc]) * jxw;
      k[0][0] += s;
    }
  }
}

// Element tangents and residuals, EPB elements per workgroup: coordinates to
// LDS; a thread per (element, qp) forms J, det, J^-1 and the physical
// gradients into LDS; then the tangent blocks and the residual rows.
// SYM (Poisson, elasticity, whose tangents the reference fills for B >= A and
// mirrors): only the blocks B >= A are formed, a thread per (element, block),
// and staged in the packed upper triangle, NB = NEN (NEN + 1) / 2 blocks,
// either component-major (S::SOA), consecutive threads taking consecutive
// elements of one block,
//   Ke[((idx(A, B) * DPN + a) * DPN + b) * nel + e],
// so that both these stores and the gather's loads (a wavefront takes
// neighbouring nodes, whose entries come from neighbouring elements) are
// contiguous, or element-major, consecutive threads taking consecutive
// blocks of one element,
//   Ke[(e * NB + idx(A, B)) * DPN^2 + a * DPN + b];
// the gather reads block (B, A) as the transpose.  Otherwise
// every block is formed, a thread per (element, row block A), into the
// row-major element matrix.
template <int DIM, int P, int DPN, bool SYM>
__global__ __launch_bounds__(NT) void k_element(int64_t nel, const int32_t *__restrict__ el,
                                                const double *__restrict__ X,
                                                const double *__restrict__ tabN,
                                                const double *__restrict__ tabdN,
                                                const double *__restrict__ tabw, Phys ph,
                                                const double *__restrict__ coef, double *__restrict__ Ke,
                                                double *__restrict__ Re, int want_k, int want_r) {
  using S = ElemShape<DIM, P>;
  constexpr int NEN = S::NEN, NQP = S::NQP, EPB = S::EPB, NLD = NEN * DPN;
  __shared__ double sGp[EPB][S::GP];  // [q][A][d] per element, padded
  __shared__ double sW[EPB][S::WP];
  __shared__ double sX[EPB][NEN][DIM];
  auto sG = [&](int le) { return reinterpret_cast<double(*)[NEN][DIM]>(&sGp[le][0]); };
  const int64_t e0 = (int64_t)blockIdx.x * EPB;
  for (int i = threadIdx.x; i < EPB * NEN; i += NT) {
    const int le = i / NEN, a = i - (i / NEN) * NEN;
    const int64_t e = e0 + le;
    if (e < nel) {
      const int64_t node = el[e * NEN + a];
#pragma unroll
      for (int d = 0; d < DIM; ++d) sX[le][a][d] = X[3 * node + d];
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < EPB * NQP; i += NT) {
    const int le = i / NQP, q = i - (i / NQP) * NQP;
    if (e0 + le >= nel) continue;
    // J(i,j) = sum_A x_A(i) dN_A/dxi_j  (Element.hpp:116-124)
    double J[DIM * DIM];
#pragma unroll
    for (int k = 0; k < DIM * DIM; ++k) J[k] = 0;
    for (int A = 0; A < NEN; ++A) {
      const double *dn = &tabdN[(q * NEN + A) * DIM];
#pragma unroll
      for (int r = 0; r < DIM; ++r)
#pragma unroll
        for (int c = 0; c < DIM; ++c) J[r * DIM + c] += sX[le][A][r] * dn[c];
    }
    double det, Ji[DIM * DIM];
    jac_inv<DIM>(J, det, Ji);
    sW[le][q] = det * tabw[q];
    // grad_x N_A = J^-T grad_xi N_A  (Element.hpp:131-147)
    for (int A = 0; A < NEN; ++A) {
      const double *dn = &tabdN[(q * NEN + A) * DIM];
#pragma unroll
      for (int d = 0; d < DIM; ++d) {
        double s = 0;
#pragma unroll
        for (int k = 0; k < DIM; ++k) s += Ji[k * DIM + d] * dn[k];
        sG(le)[q][A][d] = s;
      }
    }
  }
  __syncthreads();
  if (want_k && SYM) {
    constexpr int NB = NEN * (NEN + 1) / 2;
    for (int i = threadIdx.x; i < EPB * NB; i += NT) {
      const int bi = S::SOA ? i / EPB : i - (i / NB) * NB, le = S::SOA ? i - bi * EPB : i / NB;
      const int64_t e = e0 + le;
      if (e >= nel) continue;
      int A = 0, rem = bi;
      while (rem >= NEN - A) {
        rem -= NEN - A;
        ++A;
      }
      double k[DPN][DPN];
      tangent_block<DIM, DPN, NEN, NQP>(ph, coef ? coef[e] : 1.0, sG(le), sW[le], tabN, A, A + rem, k);
#pragma unroll
      for (int a = 0; a < DPN; ++a)
#pragma unroll
        for (int b = 0; b < DPN; ++b)
          Ke[S::SOA ? (((int64_t)bi * DPN + a) * DPN + b) * nel + e : (e * NB + bi) * (DPN * DPN) + a * DPN + b] =
              k[a][b];
    }
  }
  for (int i = threadIdx.x; i < EPB * NEN; i += NT) {
    const int le = i / NEN, A = i - (i / NEN) * NEN;
    const int64_t e = e0 + le;
    if (e >= nel) continue;
    const double ke = coef ? coef[e] : 1.0;
    if (want_k && !SYM) {
      double *Krow = Ke + e * (int64_t)NLD * NLD;
      for (int B = 0; B < NEN; ++B) {
        double k[DPN][DPN];
        tangent_block<DIM, DPN, NEN, NQP>(ph, ke, sG(le), sW[le], tabN, A, B, k);
#pragma unroll
        for (int a = 0; a < DPN; ++a)
#pragma unroll
          for (int b = 0; b < DPN; ++b) Krow[(A * DPN + a) * NLD + B * DPN + b] = k[a][b];
      }
    }
    if (want_r) {
      double r[DPN];
#pragma unroll
      for (int a = 0; a < DPN; ++a) r[a] = 0;
      for (int q = 0; q < NQP; ++q) {
        const double jxw = sW[le][q];
        const double NA = tabN[q * NEN + A];
        if (ph.kind == YAFEL_HIP_PHYS_POISSON) {
          r[0] += (ph.f * jxw) * NA;
        } else if (ph.kind == YAFEL_HIP_PHYS_ELASTICITY) {
#pragma unroll
          for (int a = 0; a < DPN; ++a) r[a] += ph.body[a] * NA * jxw;
        } else {
          const double *gA = sG(le)[q][A];
#pragma unroll
          for (int Ac = 0; Ac < DIM; ++Ac) r[0] -= (NA * ph.Q + ph.Uh[Ac] * gA[Ac]) * jxw;
        }
      }
#pragma unroll
      for (int a = 0; a < DPN; ++a) Re[(e * NEN + A) * DPN + a] = r[a];
    }
  }
}

// FP32 element matrices of the Poisson tangent for 2-D elements of order
// p = 3, 4 on the matrix cores.  With G[q*2+d][a] = sqrt(jxw_q) (J^-T grad N_a)_d
// the element matrix is K_e = -k_e G^T G (apps/BasicPoisson/main.cpp:51-58): a
// (2 nqp) x nen -> nen x nen contraction, run as 32x32x2 FP32 MFMA steps with
// nen padded to 32.  G feeds both operands: A[m][k] = G[k][m] and B[k][n] =
// G[k][n] are the same register on every lane.  One wavefront per element,
// four per workgroup; the residual stays FP64.
typedef float v16f __attribute__((ext_vector_type(16)));

template <int P>
__global__ __launch_bounds__(256) void k_element_mfma(int64_t nel, const int32_t *__restrict__ el,
                                                      const double *__restrict__ X,
                                                      const double *__restrict__ tabN,
                                                      const double *__restrict__ tabdN,
                                                      const double *__restrict__ tabw, double f,
                                                      const double *__restrict__ coef, float *__restrict__ Ke,
                                                      double *__restrict__ Re, int want_k, int want_r,
                                                      const int32_t *__restrict__ dmap = nullptr,
                                                      double *__restrict__ val = nullptr,
                                                      float *__restrict__ mstage = nullptr) {
  constexpr int N1 = P + 1, NEN = N1 * N1, NQP = NEN, KD = 2 * NQP;
  static_assert(NEN <= 32 && KD % 2 == 0, "one 32x32 tile per element");
  __shared__ float sG[4][KD][32];
  __shared__ double sX[4][NEN][2];
  __shared__ double sW[4][NQP];       // jxw
  __shared__ float sS[4][NQP][4];     // sqrt(jxw) J^-1, the factor of G at each qp
  __shared__ double sdN[NQP][NEN][2]; // gradient table (FP64, for J)
  __shared__ float sdNf[NQP][32][2];  // FP32 copy, zero past NEN (for G)
  // the tables once per workgroup; the workgroup then loops over groups of
  // four elements, one per wavefront
  for (int t = threadIdx.x; t < NQP * 32; t += 256) {
    const int q = t >> 5, a = t & 31;
    const double g0 = a < NEN ? tabdN[(q * NEN + a) * 2] : 0.0, g1 = a < NEN ? tabdN[(q * NEN + a) * 2 + 1] : 0.0;
    if (a < NEN) {
      sdN[q][a][0] = g0;
      sdN[q][a][1] = g1;
    }
    sdNf[q][a][0] = (float)g0;
    sdNf[q][a][1] = (float)g1;
  }
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int64_t ngroups = (nel + 3) / 4;
  // node coordinates of the next group are loaded while this one computes
  double nx0 = 0, nx1 = 0;
  auto load_x = [&](int64_t grp) {
    const int64_t e = grp * 4 + w;
    if (grp < ngroups && e < nel && l < NEN) {
      const int64_t node = el[e * NEN + l];
      nx0 = X[3 * node];
      nx1 = X[3 * node + 1];
    }
  };
  load_x(blockIdx.x);
  for (int64_t grp = blockIdx.x; grp < ngroups; grp += gridDim.x) {
    const int64_t e = grp * 4 + w;
    const bool live = e < nel;  // uniform over the wavefront
    if (live && l < NEN) {
      sX[w][l][0] = nx0;
      sX[w][l][1] = nx1;
    }
    __syncthreads();
    load_x(grp + gridDim.x);
    if (live && l < NQP) {
      // Element::update (Element.hpp:116-130), in FP64
      double J[4] = {0, 0, 0, 0};
#pragma unroll 5
      for (int A = 0; A < NEN; ++A) {
        const double g0 = sdN[l][A][0], g1 = sdN[l][A][1];
        J[0] += sX[w][A][0] * g0;
        J[1] += sX[w][A][0] * g1;
        J[2] += sX[w][A][1] * g0;
        J[3] += sX[w][A][1] * g1;
      }
      double det, Ji[4];
      jac_inv<2>(J, det, Ji);
      const double jxw = det * tabw[l], sq = sqrt(jxw);
#pragma unroll
      for (int k = 0; k < 4; ++k) sS[w][l][k] = (float)(sq * Ji[k]);
      sW[w][l] = jxw;
    }
    __syncthreads();
    if (live) {
      // G[2q + d][a] = sqrt(jxw_q) (J^-T grad_xi N_a)_d, formed in FP32 (the
      // products are FP32 anyway); a lane fills column a of every other row pair
      const int a = l & 31, half = l >> 5;
#pragma unroll 5
      for (int q = half; q < NQP; q += 2) {
        const float g0 = sdNf[q][a][0], g1 = sdNf[q][a][1];
        sG[w][2 * q][a] = sS[w][q][0] * g0 + sS[w][q][2] * g1;
        sG[w][2 * q + 1][a] = sS[w][q][1] * g0 + sS[w][q][3] * g1;
      }
    }
    __syncthreads();
    if (live && want_k) {
      v16f acc = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll 5
      for (int kk = 0; kk < KD / 2; ++kk) {
        const float v = sG[w][2 * kk + (l >> 5)][l & 31];
        acc = __builtin_amdgcn_mfma_f32_32x32x2f32(v, v, acc, 0, 0, 0);
      }
      const float ke = (float)(coef ? coef[e] : 1.0);
      const int col = l & 31;
      if (dmap) {
        // straight into the matrix: an entry with one contributing element
        // goes to its sliced position as FP64 (what the gather's FP64 sum of
        // one FP32 term gives), the others to their place in the compact
        // staging of shared entries (code -1 - index)
        const int32_t *dm = dmap + e * (int64_t)NEN * NEN;
        int32_t d[16];
#pragma unroll
        for (int i = 0; i < 16; ++i) {
          const int row = 8 * (i >> 2) + 4 * (l >> 5) + (i & 3);
          d[i] = row < NEN && col < NEN ? dm[row * NEN + col] : 0;
        }
#pragma unroll
        for (int i = 0; i < 16; ++i) {
          const int row = 8 * (i >> 2) + 4 * (l >> 5) + (i & 3);
          if (row < NEN && col < NEN) {
            const float v = -ke * acc[i];
            if (d[i] >= 0) val[d[i]] = (double)v;
            else mstage[-1 - (int64_t)d[i]] = v;
          }
        }
      } else {
        float *Kout = Ke + e * (int64_t)NEN * NEN;
#pragma unroll
        for (int i = 0; i < 16; ++i) {
          // accumulator i of lane l holds C[8 (i/4) + 4 (l/32) + i%4][l%32]
          const int row = 8 * (i >> 2) + 4 * (l >> 5) + (i & 3);
          if (row < NEN && col < NEN) Kout[row * NEN + col] = -ke * acc[i];
        }
      }
    }
    if (live && want_r && l < NEN) {
      double r = 0;
      for (int q = 0; q < NQP; ++q) r += (f * sW[w][q]) * tabN[q * NEN + l];
      Re[e * NEN + l] = r;
    }
  }
}

// one thread per stored node pair: sum the element contributions in order
// (SYM: packed upper-triangle staging, block (a, b) with a > b read as the
// transpose of block (b, a))
template <int DPN, int FMT, class KT = double, bool SYM = false>
__global__ __launch_bounds__(NT) void k_gather(int64_t nruns, int nen, const int32_t *__restrict__ cptr,
                                               const uint32_t *__restrict__ contrib,
                                               const KT *__restrict__ Ke,
                                               const int32_t *__restrict__ run_row,
                                               const int32_t *__restrict__ node_ptr,
                                               const int32_t *__restrict__ sptr,
                                               const int32_t *__restrict__ iperm, int64_t ntot,
                                               double *__restrict__ val) {
  const int64_t j = blockIdx.x * (int64_t)NT + threadIdx.x;
  if (j >= nruns) return;
  const int nld = nen * DPN;
  const int64_t nen2 = (int64_t)nen * nen;
  double acc[DPN][DPN];
#pragma unroll
  for (int a = 0; a < DPN; ++a)
#pragma unroll
    for (int b = 0; b < DPN; ++b) acc[a][b] = 0;
  const int c0 = cptr[j], c1 = cptr[j + 1];
  for (int c = c0; c < c1; ++c) {
    const int64_t t = contrib[c];
    const int64_t e = t / nen2;
    const int r = (int)(t - e * nen2);
    const int a = r / nen, b = r - (r / nen) * nen;
    if (SYM) {
      const int lo = a <= b ? a : b, hi = a <= b ? b : a;
      const int64_t nb = (int64_t)nen * (nen + 1) / 2;
      const int64_t bi = (int64_t)lo * nen - (int64_t)lo * (lo - 1) / 2 + (hi - lo);
      const KT *k = Ke + (e * nb + bi) * (DPN * DPN);
#pragma unroll
      for (int i = 0; i < DPN; ++i)
#pragma unroll
        for (int jj = 0; jj < DPN; ++jj) acc[i][jj] += a <= b ? k[i * DPN + jj] : k[jj * DPN + i];
    } else {
      const KT *k = Ke + e * (int64_t)nld * nld + (int64_t)(a * DPN) * nld + b * DPN;
#pragma unroll
      for (int i = 0; i < DPN; ++i)
#pragma unroll
        for (int jj = 0; jj < DPN; ++jj) acc[i][jj] += k[i * nld + jj];
    }
  }
  // sliced position of the run's entries: run j is entry t of node row A
  const int A = run_row[j];
  const int64_t t = j - node_ptr[A];
  if (FMT == YAFEL_HIP_BCSR3) {
    const int64_t k = dev::sell_pos(sptr, iperm, A, t);
#pragma unroll
    for (int i = 0; i < DPN; ++i)
#pragma unroll
      for (int jj = 0; jj < DPN; ++jj) val[(int64_t)(i * 3 + jj) * ntot + k] = acc[i][jj];
  } else {
    // scalar row A * DPN + i holds component jj of run j at entry t * DPN + jj
#pragma unroll
    for (int i = 0; i < DPN; ++i)
#pragma unroll
      for (int jj = 0; jj < DPN; ++jj)
        val[dev::sell_pos(sptr, iperm, (int64_t)A * DPN + i, t * DPN + jj)] = acc[i][jj];
  }
}

// The symmetric (component-major, upper-triangle) staging gathered in the
// order of the sliced storage: SG wavefronts per slice, lane l taking stored
// row s * 64 + l at every SG-th slot (an entry is a chain of dependent loads
// -- pointer, contribution, values -- so the slots are spread over waves),
// so every store of a slot is 64 consecutive values, and neighbouring lanes
// (neighbouring nodes) read the same block of neighbouring elements.  Each entry sums its contributions in element
// order, as k_gather does; padding slots are left as they are (zero).
// Rows: block rows (BCSR3) or scalar rows (CSR, dof ci of node r / DPN, whose
// slot t holds dof t % DPN of the node entry t / DPN).
constexpr int kGatherSG = 16;  // wavefronts per slice

template <int DPN, int FMT>
__global__ __launch_bounds__(NT) void k_gather_sym(int64_t nsl, int64_t rows, int64_t nel, int nen,
                                                   const int32_t *__restrict__ cptr,
                                                   const uint32_t *__restrict__ contrib,
                                                   const double *__restrict__ Ke,
                                                   const int32_t *__restrict__ node_ptr,
                                                   const int32_t *__restrict__ sptr,
                                                   const int32_t *__restrict__ perm, int64_t ntot,
                                                   double *__restrict__ val) {
  constexpr int NV = FMT == YAFEL_HIP_BCSR3 ? DPN : 1;  // values per entry: NV x NV
  const int64_t wid = (int64_t)blockIdx.x * (NT / 64) + (threadIdx.x >> 6);
  const int64_t s = wid / kGatherSG;
  if (s >= nsl) return;
  const int l = threadIdx.x & 63;
  const int64_t i = s * 64 + l;
  const int64_t r = i < rows ? (perm ? (int64_t)perm[i] : i) : 0;
  const int64_t A = FMT == YAFEL_HIP_BCSR3 ? r : r / DPN;
  const int ci = FMT == YAFEL_HIP_BCSR3 ? 0 : (int)(r - A * DPN);
  const int n0 = node_ptr[A], nlen = i < rows ? node_ptr[A + 1] - n0 : 0;
  const int p0 = sptr[s], ns = sptr[s + 1] - p0;
  const int64_t nen2 = (int64_t)nen * nen;
  for (int t = (int)(wid - s * kGatherSG); t < ns; t += kGatherSG) {
    const int tn = FMT == YAFEL_HIP_BCSR3 ? t : t / DPN, cj = FMT == YAFEL_HIP_BCSR3 ? 0 : t - (t / DPN) * DPN;
    if (tn >= nlen) continue;
    const int64_t j = n0 + tn;
    double acc[NV][NV];
#pragma unroll
    for (int a = 0; a < NV; ++a)
#pragma unroll
      for (int b = 0; b < NV; ++b) acc[a][b] = 0;
    const int c0 = cptr[j], c1 = cptr[j + 1];
    for (int c = c0; c < c1; ++c) {
      const int64_t tc = contrib[c];
      const int64_t e = tc / nen2;
      const int rr = (int)(tc - e * nen2);
      const int a = rr / nen, b = rr - (rr / nen) * nen;
      const int lo = a <= b ? a : b, hi = a <= b ? b : a;
      const int64_t bi = (int64_t)lo * nen - (int64_t)lo * (lo - 1) / 2 + (hi - lo);
      const double *Kb = Ke + bi * (DPN * DPN) * nel + e;  // component m at Kb[m * nel]
      if (FMT == YAFEL_HIP_BCSR3) {
#pragma unroll
        for (int ii = 0; ii < NV; ++ii)
#pragma unroll
          for (int jj = 0; jj < NV; ++jj) acc[ii][jj] += Kb[(int64_t)(a <= b ? ii * DPN + jj : jj * DPN + ii) * nel];
      } else {
        acc[0][0] += Kb[(int64_t)(a <= b ? ci * DPN + cj : cj * DPN + ci) * nel];
      }
    }
    const int64_t k = (int64_t)(p0 + t) * 64 + l;
#pragma unroll
    for (int ii = 0; ii < NV; ++ii)
#pragma unroll
      for (int jj = 0; jj < NV; ++jj) val[(int64_t)(ii * NV + jj) * ntot + k] = acc[ii][jj];
  }
}

// Direct-write map of the FP32 matrix-core assembly (CSR, one dof per node):
// for every element entry t = e nen^2 + a nen + b, the sliced position of its
// stored entry when t is that entry's only contribution, else -1 - (its index
// in the compact staging of shared entries, which keeps each shared entry's
// contributions together and in element order).  off: exclusive scan of
// (1 << 32 | count) over the runs with more than one contribution.
__global__ void k_dmap_count(int64_t nruns, const int32_t *__restrict__ cptr, int64_t *__restrict__ cnt) {
  const int64_t j = blockIdx.x * (int64_t)NT + threadIdx.x;
  if (j > nruns) return;
  const int c = j < nruns ? cptr[j + 1] - cptr[j] : 0;
  cnt[j] = c > 1 ? ((int64_t)1 << 32) | c : 0;
}

__global__ void k_dmap_fill(int64_t nruns, const int32_t *__restrict__ cptr, const uint32_t *__restrict__ contrib,
                            const int32_t *__restrict__ run_row, const int32_t *__restrict__ node_ptr,
                            const int32_t *__restrict__ sptr, const int32_t *__restrict__ iperm,
                            const int64_t *__restrict__ off, int32_t *__restrict__ dmap, int32_t *__restrict__ mpos,
                            int32_t *__restrict__ mbeg) {
  const int64_t j = blockIdx.x * (int64_t)NT + threadIdx.x;
  if (j >= nruns) return;
  const int c0 = cptr[j], c1 = cptr[j + 1];
  const int A = run_row[j];
  const int64_t pos = dev::sell_pos(sptr, iperm, A, j - node_ptr[A]);
  if (c1 - c0 == 1) {
    dmap[contrib[c0]] = (int32_t)pos;
  } else {
    const int64_t o = off[j];
    const int m = (int)(o >> 32), b = (int)(o & 0xffffffffLL);
    mpos[m] = (int32_t)pos;
    mbeg[m] = b;
    for (int c = c0; c < c1; ++c) dmap[contrib[c]] = -1 - (b + (c - c0));
  }
}

// the shared entries: each one's contributions summed in element order, as
// k_gather sums them
__global__ void k_gather_shared(int64_t nm, const int32_t *__restrict__ mpos, const int32_t *__restrict__ mbeg,
                                const float *__restrict__ mstage, double *__restrict__ val) {
  const int64_t m = blockIdx.x * (int64_t)NT + threadIdx.x;
  if (m >= nm) return;
  double acc = 0;
  for (int c = mbeg[m]; c < mbeg[m + 1]; ++c) acc += (double)mstage[c];
  val[mpos[m]] = acc;
}

__global__ void k_gather_res(int64_t nnodes, int dpn, const int32_t *__restrict__ ptr,
                             const uint32_t *__restrict__ inc, const double *__restrict__ Re,
                             double *__restrict__ R) {
  const int64_t A = blockIdx.x * (int64_t)NT + threadIdx.x;
  if (A >= nnodes) return;
  for (int c = 0; c < dpn; ++c) {
    double acc = 0;
    for (int k = ptr[A]; k < ptr[A + 1]; ++k) acc += Re[(int64_t)inc[k] * dpn + c];
    R[A * dpn + c] += acc;
  }
}

// ------------------------------------------------------------ Dirichlet
__global__ void k_masked_copy(int64_t n, const uint8_t *mask, const double *g, double *gm) {
  int64_t i = blockIdx.x * (int64_t)NT + threadIdx.x;
  if (i < n) gm[i] = mask[i] ? g[i] : 0.0;
}
__global__ void k_sub(int64_t n, const double *a, double *rhs) {
  int64_t i = blockIdx.x * (int64_t)NT + threadIdx.x;
  if (i < n) rhs[i] -= a[i];
}

// DirichletBC::apply on the sliced storage: a thread per (block) row walks its
// entries DU at a time, every column and mask load of a step issued before the
// first test, so that a row costs len / DU dependent round trips instead of
// 2 len; entries past the row's length are never touched (padding repeats the
// last column, which may be the diagonal).
constexpr int DU = 8;

__global__ void k_dirichlet_csr(int64_t rows, const int32_t *ptr, const int32_t *sptr, const int32_t *iperm,
                                const int32_t *col, double *val, const uint8_t *mask, const double *g, double *rhs) {
  int64_t r = blockIdx.x * (int64_t)NT + threadIdx.x;
  if (r >= rows) return;
  const bool mr = mask[r] != 0;
  const int len = ptr[r + 1] - ptr[r];
  const int64_t k0 = dev::sell_pos(sptr, iperm, r, 0);
  for (int t0 = 0; t0 < len; t0 += DU) {
    int c[DU];
    uint8_t mc[DU];
#pragma unroll
    for (int u = 0; u < DU; ++u) c[u] = col[k0 + (int64_t)(t0 + u < len ? t0 + u : len - 1) * 64];
#pragma unroll
    for (int u = 0; u < DU; ++u) mc[u] = mask[c[u]];
#pragma unroll
    for (int u = 0; u < DU; ++u) {
      if (t0 + u < len && (mr || mc[u])) {
        const int64_t k = k0 + (int64_t)(t0 + u) * 64;
        val[k] = r == c[u] ? 1.0 : 0.0;
        if (r == c[u]) rhs[r] = g ? g[c[u]] : 0.0;
      }
    }
  }
}
__global__ void k_dirichlet_bcsr(int64_t brows, int64_t ntot, const int32_t *ptr, const int32_t *sptr,
                                 const int32_t *iperm, const int32_t *col, double *val, const uint8_t *mask,
                                 const double *g, double *rhs) {
  int64_t br = blockIdx.x * (int64_t)NT + threadIdx.x;
  if (br >= brows) return;
  const int len = ptr[br + 1] - ptr[br];
  const int64_t k0 = dev::sell_pos(sptr, iperm, br, 0);
  bool mr[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) mr[i] = mask[3 * br + i] != 0;
  for (int t0 = 0; t0 < len; t0 += DU) {
    int bcs[DU];
    uint8_t mc[DU][3];
#pragma unroll
    for (int u = 0; u < DU; ++u) bcs[u] = col[k0 + (int64_t)(t0 + u < len ? t0 + u : len - 1) * 64];
#pragma unroll
    for (int u = 0; u < DU; ++u)
#pragma unroll
      for (int j = 0; j < 3; ++j) mc[u][j] = mask[3 * (int64_t)bcs[u] + j];
#pragma unroll
    for (int u = 0; u < DU; ++u) {
      if (t0 + u >= len) continue;
      const int64_t k = k0 + (int64_t)(t0 + u) * 64;
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
          const int64_t r = 3 * br + i, c = 3 * (int64_t)bcs[u] + j;
          if (mr[i] || mc[u][j]) {
            val[(int64_t)(i * 3 + j) * ntot + k] = r == c ? 1.0 : 0.0;
            if (r == c) rhs[r] = g ? g[c] : 0.0;
          }
        }
    }
  }
}

Phys make_phys(const yafel_hip_physics *p) {
  Phys ph{};
  ph.kind = p->kind;
  ph.f = p->f;
  ph.lambda = p->lambda;
  ph.mu = p->mu;
  for (int d = 0; d < 3; ++d) {
    ph.body[d] = p->body[d];
    ph.U[d] = p->U[d];
  }
  ph.kdiff = p->kdiff;
  ph.h = p->h;
  ph.alpha = p->alpha;
  ph.Q = p->Q;
  const int nsd = p->nsd;
  double nU = 0;
  for (int d = 0; d < nsd; ++d) nU += p->U[d] * p->U[d];
  nU = std::sqrt(nU);
  for (int a = 0; a < nsd; ++a)
    for (int b = 0; b < nsd; ++b) {
      double v = (a == b) ? p->kdiff : 0.0;
      if (nU > 0) v += (p->alpha * p->h / (nU * 2)) * p->U[a] * p->U[b];
      ph.Kt[a * nsd + b] = v;
    }
  for (int d = 0; d < 3; ++d) ph.Uh[d] = (nU > 0 && d < nsd) ? (p->alpha * p->h / 2) * p->U[d] / nU : 0.0;
  return ph;
}

template <int DIM, int P, int DPN>
void launch_elem_dpn(yafel_hip_mesh_s *m, const Phys &ph, const double *coef, double *Ke, double *Re, int wk,
                     int wr, bool sym) {
  using S = ElemShape<DIM, P>;
  const int grid = ceil_div(m->nel, S::EPB);
  if (sym)
    k_element<DIM, P, DPN, true><<<grid, NT, 0, m->ctx->stream>>>(
        m->nel, m->elements.as<int32_t>(), m->nodes.as<double>(), m->tabN.as<double>(), m->tabdN.as<double>(),
        m->tabw.as<double>(), ph, coef, Ke, Re, wk, wr);
  else
    k_element<DIM, P, DPN, false><<<grid, NT, 0, m->ctx->stream>>>(
        m->nel, m->elements.as<int32_t>(), m->nodes.as<double>(), m->tabN.as<double>(), m->tabdN.as<double>(),
        m->tabw.as<double>(), ph, coef, Ke, Re, wk, wr);
  HIPCHK(hipGetLastError());
}

template <int DIM, int P>
void launch_elem_p(yafel_hip_mesh_s *m, const Phys &ph, const double *coef, double *Ke, double *Re, int wk, int wr,
                   bool sym) {
  switch (m->dpn) {
    case 1: launch_elem_dpn<DIM, P, 1>(m, ph, coef, Ke, Re, wk, wr, sym); break;
    case 2: launch_elem_dpn<DIM, P, 2>(m, ph, coef, Ke, Re, wk, wr, sym); break;
    default: launch_elem_dpn<DIM, P, 3>(m, ph, coef, Ke, Re, wk, wr, sym); break;
  }
}

// sym: stage only the blocks B >= A (packed upper triangle); the public
// element-matrix entry point keeps full row-major matrices
void launch_element(yafel_hip_mesh_s *m, const Phys &ph, const double *coef, double *Ke, double *Re, int wk, int wr,
                    bool sym = false) {
  if (m->dim == 2) {
    switch (m->p) {
      case 1: launch_elem_p<2, 1>(m, ph, coef, Ke, Re, wk, wr, sym); break;
      case 2: launch_elem_p<2, 2>(m, ph, coef, Ke, Re, wk, wr, sym); break;
      case 3: launch_elem_p<2, 3>(m, ph, coef, Ke, Re, wk, wr, sym); break;
      case 4: launch_elem_p<2, 4>(m, ph, coef, Ke, Re, wk, wr, sym); break;
      default: fail(YAFEL_HIP_EINVAL, "assemble: 2-D elements are supported for p = 1..4");
    }
  } else {
    switch (m->p) {
      case 1: launch_elem_p<3, 1>(m, ph, coef, Ke, Re, wk, wr, sym); break;
      case 2: launch_elem_p<3, 2>(m, ph, coef, Ke, Re, wk, wr, sym); break;
      default: fail(YAFEL_HIP_EINVAL, "assemble: 3-D elements are supported for p = 1..2");
    }
  }
}

void check_phys(yafel_hip_mesh_s *m, const yafel_hip_physics *p) {
  REQUIRE(p, "assemble: null physics");
  REQUIRE(p->nsd == m->dim, "assemble: physics nsd must equal the mesh dimension");
  REQUIRE(p->kind >= 0 && p->kind <= 2, "assemble: unknown physics kind");
  if (p->kind == YAFEL_HIP_PHYS_ELASTICITY) REQUIRE(m->dpn == m->dim, "assemble: elasticity needs dim dofs per node");
  else REQUIRE(m->dpn == 1, "assemble: scalar physics needs 1 dof per node");
}

}  // namespace

extern "C" {

int yafel_hip_element_matrices(yafel_hip_mesh m, const yafel_hip_physics *phys, double *Ke_dev) {
  YH_TRY_BEGIN
  REQUIRE(m && Ke_dev, "element_matrices: null argument");
  check_phys(m, phys);
  HIPCHK(hipSetDevice(m->ctx->device));
  DBuf coefd;
  const double *coef = nullptr;
  if (phys->coef) {
    if (phys->coef_on_device) coef = phys->coef;
    else {
      coefd.alloc(m->nel * 8);
      HIPCHK(hipMemcpyAsync(coefd.p, phys->coef, m->nel * 8, hipMemcpyHostToDevice, m->ctx->stream));
      coef = coefd.as<double>();
    }
  }
  Phys ph = make_phys(phys);
  launch_element(m, ph, coef, Ke_dev, nullptr, 1, 0);
  HIPCHK(hipStreamSynchronize(m->ctx->stream));
  return YAFEL_HIP_OK;
  YH_TRY_END(m ? m->ctx : nullptr)
}

int yafel_hip_element_matrices_fp32(yafel_hip_mesh m, const yafel_hip_physics *phys, float *Ke_dev) {
  YH_TRY_BEGIN
  REQUIRE(m && Ke_dev, "element_matrices_fp32: null argument");
  check_phys(m, phys);
  REQUIRE(m->dim == 2 && (m->p == 3 || m->p == 4) && m->dpn == 1 && phys->kind == YAFEL_HIP_PHYS_POISSON,
          "element_matrices_fp32: matrix-core element matrices cover 2-D Poisson with p = 3 or 4");
  HIPCHK(hipSetDevice(m->ctx->device));
  REQUIRE(!phys->coef || phys->coef_on_device, "element_matrices_fp32: coefficients must be device data");
  const unsigned g = (unsigned)std::min<int64_t>((m->nel + 3) / 4, (int64_t)m->ctx->num_cus * 3);
  hipStream_t s = m->ctx->stream;
  if (m->p == 3)
    k_element_mfma<3><<<g, 256, 0, s>>>(m->nel, m->elements.as<int32_t>(), m->nodes.as<double>(), m->tabN.as<double>(),
                                         m->tabdN.as<double>(), m->tabw.as<double>(), phys->f, phys->coef, Ke_dev,
                                         nullptr, 1, 0);
  else
    k_element_mfma<4><<<g, 256, 0, s>>>(m->nel, m->elements.as<int32_t>(), m->nodes.as<double>(), m->tabN.as<double>(),
                                         m->tabdN.as<double>(), m->tabw.as<double>(), phys->f, phys->coef, Ke_dev,
                                         nullptr, 1, 0);
  HIPCHK(hipGetLastError());
  return YAFEL_HIP_OK;
  YH_TRY_END(m ? m->ctx : nullptr)
}

// The direct-write map of A's pattern (CSR, one dof per node), see k_dmap_fill.
void build_direct_map(yafel_hip_mesh_s *m, yafel_hip_matrix_s *A) {
  if (A->asm_nshared >= 0) return;
  YhPlan *pl = A->plan;
  hipStream_t s = m->ctx->stream;
  const int64_t nruns = pl->nruns, nent = m->nel * (int64_t)m->nen * m->nen;
  REQUIRE(A->ntot < INT32_MAX && nent < INT32_MAX, "assemble: matrix too large for the direct-write map");
  DBuf cnt, off;
  cnt.alloc((nruns + 1) * 8);
  off.alloc((nruns + 1) * 8);
  k_dmap_count<<<ceil_div(nruns + 1, NT), NT, 0, s>>>(nruns, pl->cptr.as<int32_t>(), cnt.as<int64_t>());
  HIPCHK(hipGetLastError());
  exclusive_scan_i64(cnt.as<int64_t>(), off.as<int64_t>(), nruns + 1, s);
  int64_t tot = 0;
  HIPCHK(hipMemcpyAsync(&tot, off.as<int64_t>() + nruns, 8, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  const int64_t nm = tot >> 32;
  const int32_t ncon = (int32_t)(tot & 0xffffffffLL);
  A->asm_dmap.alloc(std::max<int64_t>(nent, 1) * 4);
  A->asm_mpos.alloc(std::max<int64_t>(nm, 1) * 4);
  A->asm_mbeg.alloc((nm + 1) * 4);
  A->asm_mstage.alloc(std::max<int64_t>(ncon, 1) * 4);
  if (nruns > 0) {
    k_dmap_fill<<<ceil_div(nruns, NT), NT, 0, s>>>(nruns, pl->cptr.as<int32_t>(), pl->contrib.as<uint32_t>(),
                                                   pl->run_row.as<int32_t>(), pl->node_ptr.as<int32_t>(),
                                                   A->sptr.as<int32_t>(),
                                                   A->iperm.p ? A->iperm.as<int32_t>() : nullptr,
                                                   off.as<int64_t>(), A->asm_dmap.as<int32_t>(),
                                                   A->asm_mpos.as<int32_t>(), A->asm_mbeg.as<int32_t>());
    HIPCHK(hipGetLastError());
  }
  HIPCHK(hipMemcpyAsync(A->asm_mbeg.as<int32_t>() + nm, &ncon, 4, hipMemcpyHostToDevice, s));
  HIPCHK(hipStreamSynchronize(s));
  A->asm_nshared = nm;
}

int yafel_hip_assemble(yafel_hip_mesh m, const yafel_hip_physics *phys, int requirements,
                       yafel_hip_matrix A, double *R_dev) {
  YH_TRY_BEGIN
  REQUIRE(m, "assemble: null mesh");
  check_phys(m, phys);
  const bool wk = (requirements & YAFEL_HIP_ASSEMBLE_TANGENT) != 0;
  const bool wr = (requirements & YAFEL_HIP_ASSEMBLE_RESIDUAL) != 0;
  REQUIRE(!wk || A, "assemble: tangent requested without a matrix");
  REQUIRE(!wr || R_dev, "assemble: residual requested without R");
  REQUIRE(!wk || (A->mesh == m && A->plan), "assemble: matrix was not created from this mesh's pattern");
  yafel_hip_ctx_s *ctx = m->ctx;
  HIPCHK(hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  const int nld = m->nen * m->dpn;
  DBuf coefd;
  const double *coef = nullptr;
  if (phys->coef) {
    if (phys->coef_on_device) coef = phys->coef;
    else {
      coefd.alloc(m->nel * 8);
      HIPCHK(hipMemcpyAsync(coefd.p, phys->coef, m->nel * 8, hipMemcpyHostToDevice, s));
      coef = coefd.as<double>();
    }
  }
  const bool f32 = (requirements & YAFEL_HIP_ASSEMBLE_FP32_ELEMENTS) != 0;
  REQUIRE(!f32 || (m->dim == 2 && (m->p == 3 || m->p == 4) && m->dpn == 1 && phys->kind == YAFEL_HIP_PHYS_POISSON),
          "assemble: FP32 matrix-core element matrices cover 2-D Poisson with p = 3 or 4");
  if (wk && !f32 && m->stageK.bytes < (size_t)m->nel * nld * nld * 8) m->stageK.alloc((size_t)m->nel * nld * nld * 8);
  // the FP32 matrix-core tangent of a CSR pattern is written straight into
  // the matrix (entries shared by several elements through a compact staging)
  const bool direct = wk && f32 && A->format == YAFEL_HIP_CSR;
  if (direct) build_direct_map(m, A);
  if (wk && f32 && !direct && m->stageK32.bytes < (size_t)m->nel * nld * nld * 4)
    m->stageK32.alloc((size_t)m->nel * nld * nld * 4);
  if (wr && m->stageR.bytes < (size_t)m->nel * nld * 8) m->stageR.alloc((size_t)m->nel * nld * 8);
  Phys ph = make_phys(phys);
  // symmetric tangents (the reference mirrors B < A) stage half the blocks
  const bool sym = !f32 && phys->kind != YAFEL_HIP_PHYS_SUPG;
  if (f32) {
    const unsigned g = (unsigned)std::min<int64_t>((m->nel + 3) / 4, (int64_t)m->ctx->num_cus * 3);
    float *k32 = wk && !direct ? m->stageK32.as<float>() : nullptr;
    double *r64 = wr ? m->stageR.as<double>() : nullptr;
    const int32_t *dmap = direct ? A->asm_dmap.as<int32_t>() : nullptr;
    double *dval = direct ? A->val64.as<double>() : nullptr;
    float *mst = direct ? A->asm_mstage.as<float>() : nullptr;
    if (m->p == 3)
      k_element_mfma<3><<<g, 256, 0, s>>>(m->nel, m->elements.as<int32_t>(), m->nodes.as<double>(),
                                           m->tabN.as<double>(), m->tabdN.as<double>(), m->tabw.as<double>(), ph.f,
                                           coef, k32, r64, wk, wr, dmap, dval, mst);
    else
      k_element_mfma<4><<<g, 256, 0, s>>>(m->nel, m->elements.as<int32_t>(), m->nodes.as<double>(),
                                           m->tabN.as<double>(), m->tabdN.as<double>(), m->tabw.as<double>(), ph.f,
                                           coef, k32, r64, wk, wr, dmap, dval, mst);
    HIPCHK(hipGetLastError());
    if (direct && A->asm_nshared > 0) {
      k_gather_shared<<<ceil_div(A->asm_nshared, NT), NT, 0, s>>>(A->asm_nshared, A->asm_mpos.as<int32_t>(),
                                                                  A->asm_mbeg.as<int32_t>(),
                                                                  A->asm_mstage.as<float>(), A->val64.as<double>());
      HIPCHK(hipGetLastError());
    }
  } else {
    launch_element(m, ph, coef, wk ? m->stageK.as<double>() : nullptr, wr ? m->stageR.as<double>() : nullptr, wk,
                   wr, sym);
  }
  if (wk) {
    YhPlan *pl = A->plan;
    const int g = ceil_div(pl->nruns, NT);
    const int32_t *sp = A->sptr.as<int32_t>();
    const int32_t *ip = A->iperm.p ? A->iperm.as<int32_t>() : nullptr;
    double *val = A->val64.as<double>();
    if (pl->nruns > 0 && !direct) {
#define GATHER(DPN_, FMT_, KT_, SYM_, K_)                                                                        \
  k_gather<DPN_, FMT_, KT_, SYM_><<<g, NT, 0, s>>>(pl->nruns, m->nen, pl->cptr.as<int32_t>(),                     \
                                                   pl->contrib.as<uint32_t>(), K_, pl->run_row.as<int32_t>(),     \
                                                   pl->node_ptr.as<int32_t>(), sp, ip, A->ntot, val)
      const double *k64 = m->stageK.as<double>();
      bool sym_aos = false;  // element-major symmetric staging: the thread-per-node-pair gather
      const int gs = ceil_div(A->nsl * kGatherSG, NT / 64);
      const int32_t *pm = A->perm.p ? A->perm.as<int32_t>() : nullptr;
      const int64_t srows = A->format == YAFEL_HIP_BCSR3 ? A->brows : A->rows;
#define GATHER_SYM(DPN_, FMT_)                                                                                   \
  k_gather_sym<DPN_, FMT_><<<gs, NT, 0, s>>>(A->nsl, srows, m->nel, m->nen, pl->cptr.as<int32_t>(),               \
                                             pl->contrib.as<uint32_t>(), k64, pl->node_ptr.as<int32_t>(), sp, pm, \
                                             A->ntot, val)
      if (sym && !sym_soa(m->dim, m->p)) sym_aos = true;
      if (f32) GATHER(1, YAFEL_HIP_CSR, float, false, m->stageK32.as<float>());
      else if (A->format == YAFEL_HIP_BCSR3 && sym && sym_aos) GATHER(3, YAFEL_HIP_BCSR3, double, true, k64);
      else if (m->dpn == 1 && sym && sym_aos) GATHER(1, YAFEL_HIP_CSR, double, true, k64);
      else if (m->dpn == 2 && sym && sym_aos) GATHER(2, YAFEL_HIP_CSR, double, true, k64);
      else if (sym && sym_aos) GATHER(3, YAFEL_HIP_CSR, double, true, k64);
      else if (A->format == YAFEL_HIP_BCSR3 && sym) GATHER_SYM(3, YAFEL_HIP_BCSR3);
      else if (A->format == YAFEL_HIP_BCSR3) GATHER(3, YAFEL_HIP_BCSR3, double, false, k64);
      else if (m->dpn == 1 && sym) GATHER_SYM(1, YAFEL_HIP_CSR);
      else if (m->dpn == 1) GATHER(1, YAFEL_HIP_CSR, double, false, k64);
      else if (m->dpn == 2 && sym) GATHER_SYM(2, YAFEL_HIP_CSR);
      else if (m->dpn == 2) GATHER(2, YAFEL_HIP_CSR, double, false, k64);
      else if (sym) GATHER_SYM(3, YAFEL_HIP_CSR);
      else GATHER(3, YAFEL_HIP_CSR, double, false, k64);
#undef GATHER_SYM
#undef GATHER
      HIPCHK(hipGetLastError());
    }
    A->has32 = false;
    A->v64_gen++;
    A->diag_ok = false;
  }
  if (wr) {
    mesh_build_incidence(m);
    k_gather_res<<<ceil_div(m->n_nodes, NT), NT, 0, s>>>(m->n_nodes, m->dpn, m->inc_ptr.as<int32_t>(),
                                                         m->inc.as<uint32_t>(), m->stageR.as<double>(), R_dev);
    HIPCHK(hipGetLastError());
  }
  if (coefd.p) HIPCHK(hipStreamSynchronize(s));
  return YAFEL_HIP_OK;
  YH_TRY_END(m ? m->ctx : nullptr)
}

int yafel_hip_dirichlet_apply(yafel_hip_matrix A, const uint8_t *mask_dev, const double *g_dev,
                              double *rhs_dev) {
  YH_TRY_BEGIN
  REQUIRE(A && mask_dev && rhs_dev, "dirichlet_apply: null argument");
  REQUIRE(A->has64, "dirichlet_apply: matrix has no FP64 values");
  REQUIRE(A->rows == A->cols, "dirichlet_apply: matrix must be square");
  yafel_hip_ctx_s *ctx = A->ctx;
  HIPCHK(hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  const int64_t n = A->rows;
  if (n == 0) return YAFEL_HIP_OK;
  DBuf gm, ag;
  if (g_dev) {
    // rhs -= A * bc_values, bc_values zero outside the BC dofs (DirichletBC.cpp:34-41)
    gm.alloc(n * 8);
    ag.alloc(n * 8);
    k_masked_copy<<<ceil_div(n, NT), NT, 0, s>>>(n, mask_dev, g_dev, gm.as<double>());
    HIPCHK(hipGetLastError());
    spmv_launch(A, YAFEL_HIP_FP64, gm.p, ag.p, s);
    k_sub<<<ceil_div(n, NT), NT, 0, s>>>(n, ag.as<double>(), rhs_dev);
    HIPCHK(hipGetLastError());
  }
  const double *gv = g_dev ? gm.as<double>() : nullptr;
  const int32_t *ip = A->iperm.p ? A->iperm.as<int32_t>() : nullptr;
  if (A->format == YAFEL_HIP_CSR)
    k_dirichlet_csr<<<ceil_div(A->brows, NT), NT, 0, s>>>(A->brows, A->ptr.as<int32_t>(), A->sptr.as<int32_t>(), ip,
                                                          A->col.as<int32_t>(), A->val64.as<double>(), mask_dev, gv,
                                                          rhs_dev);
  else
    k_dirichlet_bcsr<<<ceil_div(A->brows, NT), NT, 0, s>>>(A->brows, A->ntot, A->ptr.as<int32_t>(),
                                                           A->sptr.as<int32_t>(), ip, A->col.as<int32_t>(),
                                                           A->val64.as<double>(), mask_dev, gv, rhs_dev);
  HIPCHK(hipGetLastError());
  A->has32 = false;
  A->v64_gen++;
  A->diag_ok = false;
  if (g_dev) HIPCHK(hipStreamSynchronize(s));
  return YAFEL_HIP_OK;
  YH_TRY_END(A ? A->ctx : nullptr)
}

}  // extern "C"
