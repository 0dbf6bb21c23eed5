// Discontinuous Galerkin residual and its explicit RK4 driver on the device
// (DGAssembly<Physics> with {Residual, DtMass}, include/assembly/DGAssembly.hpp:32-305;
// DGRK4::step, include/time_integration/DGRK4.hpp:33-69) for Quad4 meshes with
// tensor-product elements of order p and the upwind advection physics of
// apps/BasicAdvection_DG/main.cpp:14-98.
//
// Set-up (once per mesh): element-private GLL nodes (DoFManager::make_raw_dofs),
// the mesh faces sorted by their corner pair (Mesh::buildInternalFaces, a
// radix sort of (min, max) corner keys), the matched face node lists
// (DoFManager::match_face_nodes), the face geometry at the face quadrature
// points (Element::face_update) and one inverse mass matrix per element
// (PartialPivLU of LocalMass, replaced by a Gauss-Jordan inverse).
//
// One residual evaluation is two kernels:
//   face kernel, a thread per (face, face point): upwind or outflow flux;
//   element kernel, a lane per (element, dof): the face terms of the dof in
//   face order, the volume term, then R = R_face + M^-1 (R_face + R_vol) --
//   the reference seeds the element's local residual with the face terms and
//   adds its M^-1 solve on top of them.
// An RK4 step is four evaluations plus vector updates, captured once into a
// HIP graph and replayed.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "internal.hpp"
#include "mesh.hpp"
#include "sortprim.hpp"

using namespace yh;

namespace {

constexpr int NT = 256;
constexpr int PMAX = 4;

struct DGFaceRec {
  int left, right, lf, rf;
};

}  // namespace

struct yafel_hip_dg_s {
  yafel_hip_ctx_s *ctx = nullptr;
  int p = 1, n1 = 2, nen = 4, nqp = 4;
  int64_t nel = 0, nfaces = 0, nbnd = 0;
  DBuf X;                 // dof node coordinates, nel*nen*3
  DBuf tabN, tabdN, tabw; // volume tables (GL points x GLL nodes)
  DBuf bN, bdN, bw;       // face tables (1-D)
  DBuf fnodes;            // int32 [4][n1] local face nodes
  DBuf faces;             // DGFaceRec [nfaces]
  DBuf lnodes, rnodes;    // int32 [nfaces][n1] matched local node lists (-1: no side)
  DBuf geo;               // double [nfaces][n1][4]: unit normal (x, y), jxw left list, jxw right list
  DBuf ef;                // int32 [nel][4]: (face << 1) | side (0 left, 1 right)
  DBuf minv;              // double [nel][nen][nen], transposed: minv[e][j][i] = (M_e^-1)_ij
  DBuf ginfo;             // double [nel][8]: affine flag, det J, J^-T (4) of affine cells
  DBuf mrefT;             // double [nen][nen]: reference mass inverse, transposed
  int64_t naffine = 0;    // cells on the affine path
  DBuf flux;              // double [nel][4][n1]: face-node flux sums of one evaluation (k_dg_flux)
  DBuf rk;                // U0, k1..k4, W (6 * ndofs)
  DBuf tdev;              // device time
  hipGraphExec_t graph = nullptr;
  double graph_key[4] = {0, 0, 0, 0};  // dt, velocity, source the graph was captured with
  double *graph_U = nullptr;
  ~yafel_hip_dg_s() {
    if (graph) (void)hipGraphExecDestroy(graph);
    yh::ctx_release(ctx);
  }
};

namespace {

// --------------------------------------------------------------- set-up
__global__ void k_dg_nodes(int64_t nel, int nen, const double *__restrict__ corners,
                           const int32_t *__restrict__ cells, const double *__restrict__ xl,
                           double *__restrict__ X) {
  const int64_t i = blockIdx.x * (int64_t)NT + threadIdx.x;
  if (i >= nel * nen) return;
  const int64_t e = i / nen;
  const int a = (int)(i - e * nen);
  const double x = xl[3 * a], y = xl[3 * a + 1];
  // interpolate_from_corners (DoFManager.cpp:296-338), gmsh corner order
  const double s0 = (1 - x) * (1 - y), s1 = (1 + x) * (1 - y), s2 = (1 + x) * (1 + y), s3 = (1 - x) * (1 + y);
  for (int d = 0; d < 3; ++d) {
    const double c0 = corners[3 * cells[4 * e] + d], c1 = corners[3 * cells[4 * e + 1] + d];
    const double c2 = corners[3 * cells[4 * e + 2] + d], c3 = corners[3 * cells[4 * e + 3] + d];
    X[3 * i + d] = (1.0 / 4.0) * (s0 * c0 + s1 * c1 + s2 * c2 + s3 * c3);
  }
}

// CellFace::canonicalCellFace + orient for Quad4: face f of cell c joins
// corners f and (f+1)%4; the cell is the face's left when that pair is
// increasing, its right otherwise.  Key = (min, max) corner pair.
__global__ void k_dg_face_keys(int64_t ncell, int64_t ncorner, const int32_t *__restrict__ cells,
                               uint64_t *__restrict__ key, uint32_t *__restrict__ val) {
  const int64_t i = blockIdx.x * (int64_t)NT + threadIdx.x;
  if (i >= 4 * ncell) return;
  const int64_t c = i >> 2;
  const int f = (int)(i & 3);
  const int64_t a = cells[4 * c + f], b = cells[4 * c + (f + 1) % 4];
  const int64_t lo = a < b ? a : b, hi = a < b ? b : a;
  key[i] = (uint64_t)(lo * ncorner + hi);
  val[i] = (uint32_t)((i << 1) | (b < a ? 1 : 0));
}

__global__ void k_dg_heads(int64_t n, const uint64_t *__restrict__ key, int32_t *__restrict__ head) {
  const int64_t i = blockIdx.x * (int64_t)NT + threadIdx.x;
  if (i < n) head[i] = (i == 0 || key[i] != key[i - 1]) ? 1 : 0;
}

__global__ void k_dg_face_init(int64_t nf, DGFaceRec *F) {
  const int64_t i = blockIdx.x * (int64_t)NT + threadIdx.x;
  if (i < nf) F[i] = DGFaceRec{-1, -1, -1, -1};
}

// Merge duplicates as buildInternalFaces does (build_faces.cpp:63-75: fields
// combined with max), so the result does not depend on the sort's tie order.
__global__ void k_dg_face_fill(int64_t n, const uint32_t *__restrict__ val, const int32_t *__restrict__ fid,
                               DGFaceRec *F) {
  const int64_t i = blockIdx.x * (int64_t)NT + threadIdx.x;
  if (i >= n) return;
  const uint32_t v = val[i];
  const int side = (int)(v & 1);
  const int64_t cf = v >> 1;
  const int c = (int)(cf >> 2), f = (int)(cf & 3);
  DGFaceRec *r = &F[fid[i] - 1];
  if (side == 0) {
    atomicMax(&r->left, c);
    atomicMax(&r->lf, f);
  } else {
    atomicMax(&r->right, c);
    atomicMax(&r->rf, f);
  }
}

// match_face_nodes (DoFManager.cpp:231-293): the left list is the left
// cell's face nodes, the right list the right cell's nodes at the same
// positions (first within 1e-6); a boundary face keeps its one side's list.
__global__ void k_dg_match(int64_t nf, int n1, int nen, const DGFaceRec *__restrict__ F,
                           const int32_t *__restrict__ fn, const double *__restrict__ X, int32_t *__restrict__ L,
                           int32_t *__restrict__ R, int *bad) {
  const int64_t f = blockIdx.x * (int64_t)NT + threadIdx.x;
  if (f >= nf) return;
  const DGFaceRec r = F[f];
  for (int i = 0; i < n1; ++i) {
    L[f * n1 + i] = -1;
    R[f * n1 + i] = -1;
  }
  if (r.left < 0 && r.right < 0) {
    atomicAdd(bad, 1);
    return;
  }
  if (r.left < 0) {
    for (int i = 0; i < n1; ++i) R[f * n1 + i] = fn[r.rf * n1 + i];
    return;
  }
  if (r.right < 0) {
    for (int i = 0; i < n1; ++i) L[f * n1 + i] = fn[r.lf * n1 + i];
    return;
  }
  for (int i = 0; i < n1; ++i) {
    const int nl = fn[r.lf * n1 + i];
    L[f * n1 + i] = nl;
    const double *xl = &X[3 * ((int64_t)r.left * nen + nl)];
    int found = -1;
    for (int j = 0; j < n1 && found < 0; ++j) {
      const int nr = fn[r.rf * n1 + j];
      const double *xr = &X[3 * ((int64_t)r.right * nen + nr)];
      const double d0 = xl[0] - xr[0], d1 = xl[1] - xr[1], d2 = xl[2] - xr[2];
      if (sqrt(d0 * d0 + d1 * d1 + d2 * d2) < 1.0e-6) found = nr;
    }
    R[f * n1 + i] = found;
    if (found < 0) atomicAdd(bad, 1);
  }
}

// Element::face_update<2> (Element.hpp:151-199) on a node list:
// n = (J(1,0), -J(0,0)) / |.|, jxw = |.| w.
__device__ void face_geom(int n1, int nen, int64_t el, const int32_t *nodes, const double *X, const double *bdN,
                          const double *bw, int fq, double &nx, double &ny, double &jxw) {
  double J0 = 0, J1 = 0;
  for (int A = 0; A < n1; ++A) {
    const double *x = &X[3 * (el * nen + nodes[A])];
    const double g = bdN[fq * n1 + A];
    J0 += x[0] * g;
    J1 += x[1] * g;
  }
  const double det = sqrt(J1 * J1 + J0 * J0);
  nx = J1 / det;
  ny = -J0 / det;
  jxw = det * bw[fq];
}

__global__ void k_dg_geom(int64_t nf, int n1, int nen, const DGFaceRec *__restrict__ F, const int32_t *__restrict__ L,
                          const int32_t *__restrict__ R, const double *__restrict__ X, const double *__restrict__ bdN,
                          const double *__restrict__ bw, double *__restrict__ geo) {
  const int64_t i = blockIdx.x * (int64_t)NT + threadIdx.x;
  if (i >= nf * n1) return;
  const int64_t f = i / n1;
  const int fq = (int)(i - f * n1);
  const DGFaceRec r = F[f];
  // unmatched node lists are reported by k_dg_match; leave their geometry NaN
  for (int k = 0; k < n1; ++k)
    if ((r.left >= 0 && L[f * n1 + k] < 0) || (r.right >= 0 && R[f * n1 + k] < 0) || (r.left < 0 && r.right < 0)) {
      geo[4 * i] = geo[4 * i + 1] = geo[4 * i + 2] = geo[4 * i + 3] = __builtin_nan("");
      return;
    }
  double nx, ny, jl = 0, jr = 0;
  if (r.left >= 0) {
    face_geom(n1, nen, r.left, &L[f * n1], X, bdN, bw, fq, nx, ny, jl);
    if (r.right >= 0) {
      double mx, my;
      face_geom(n1, nen, r.right, &R[f * n1], X, bdN, bw, fq, mx, my, jr);
    }
  } else {
    face_geom(n1, nen, r.right, &R[f * n1], X, bdN, bw, fq, nx, ny, jr);
  }
  geo[4 * i] = nx;
  geo[4 * i + 1] = ny;
  geo[4 * i + 2] = jl;
  geo[4 * i + 3] = jr;
}

__global__ void k_dg_elem_faces(int64_t nf, const DGFaceRec *__restrict__ F, int32_t *__restrict__ ef) {
  const int64_t f = blockIdx.x * (int64_t)NT + threadIdx.x;
  if (f >= nf) return;
  const DGFaceRec r = F[f];
  if (r.left >= 0) ef[4 * (int64_t)r.left + r.lf] = (int32_t)(f << 1);
  if (r.right >= 0) ef[4 * (int64_t)r.right + r.rf] = (int32_t)((f << 1) | 1);
}

// Volume geometry of element e at point q: jxw and J^-T.
__device__ void vol_geom(int nen, int64_t e, int q, const double *X, const double *dN, const double *w, double &jxw,
                         double Ji[4]) {
  double J[4] = {0, 0, 0, 0};
  for (int A = 0; A < nen; ++A) {
    const double *x = &X[3 * (e * nen + A)];
    const double g0 = dN[(q * nen + A) * 2], g1 = dN[(q * nen + A) * 2 + 1];
    J[0] += x[0] * g0;
    J[1] += x[0] * g1;
    J[2] += x[1] * g0;
    J[3] += x[1] * g1;
  }
  const double det = J[0] * J[3] - J[1] * J[2];
  Ji[0] = J[3] / det;
  Ji[1] = -J[1] / det;
  Ji[2] = -J[2] / det;
  Ji[3] = J[0] / det;
  jxw = det * w[q];
}

// Affine cells (parallelograms: x00 - x10 + x11 - x01 = 0 at the corners)
// have a constant Jacobian, so M_e = det J * M_ref and M_e^-1 = M_ref^-1 / det J:
// such a cell needs its det J and J^-T (one 64-byte record) instead of its
// 16x16 inverse and node coordinates.  The record holds J from quadrature
// point 0, as the element pass would form it.
__global__ void k_dg_affine(int64_t nel, int n1, int nen, const double *__restrict__ X, const double *__restrict__ dN,
                            const double *__restrict__ w, double *__restrict__ ginfo) {
  const int64_t e = blockIdx.x * (int64_t)NT + threadIdx.x;
  if (e >= nel) return;
  const double *x00 = &X[3 * (e * nen)], *x10 = &X[3 * (e * nen + n1 - 1)];
  const double *x11 = &X[3 * (e * nen + nen - 1)], *x01 = &X[3 * (e * nen + nen - n1)];
  double h = 0, b = 0;
  for (int d = 0; d < 2; ++d) {
    h = fmax(h, fmax(fabs(x10[d] - x00[d]), fabs(x01[d] - x00[d])));
    b = fmax(b, fabs(x00[d] - x10[d] + x11[d] - x01[d]));
  }
  double *g = &ginfo[8 * e];
  const bool aff = b <= 1e-13 * h;
  double jxw, Ji[4];
  vol_geom(nen, e, 0, X, dN, w, jxw, Ji);
  g[0] = aff ? 1.0 : 0.0;
  g[1] = jxw / w[0];
  for (int k = 0; k < 4; ++k) g[2 + k] = Ji[k];
  g[6] = g[7] = 0.0;
}

// Inverse of the element mass matrix M = sum_q jxw N N^T (LocalMass,
// main.cpp:73-78) by Gauss-Jordan with partial pivoting; one workgroup of
// nen threads per element, thread i owns row i of [M | I] in LDS.
template <int NEN>
__global__ __launch_bounds__(64) void k_dg_minv(int64_t nel, int nqp, const double *__restrict__ X,
                                                const double *__restrict__ N, const double *__restrict__ dN,
                                                const double *__restrict__ w, double *__restrict__ minv,
                                                int *bad) {
  __shared__ double A[NEN][2 * NEN];
  __shared__ double jx[64];
  __shared__ int piv;
  const int64_t e = blockIdx.x;
  const int t = threadIdx.x;
  for (int q0 = 0; q0 < nqp; q0 += 64) {
    __syncthreads();
    if (t < 64 && q0 + t < nqp) {
      double Ji[4], j;
      vol_geom(NEN, e, q0 + t, X, dN, w, j, Ji);
      jx[t] = j;
    }
    __syncthreads();
    if (t < NEN) {
      if (q0 == 0)
        for (int j = 0; j < 2 * NEN; ++j) A[t][j] = (j - NEN == t) ? 1.0 : 0.0;
      for (int q = q0; q < nqp && q < q0 + 64; ++q) {
        const double ni = jx[q - q0] * N[q * NEN + t];
        for (int j = 0; j < NEN; ++j) A[t][j] += ni * N[q * NEN + j];
      }
    }
  }
  __syncthreads();
  for (int k = 0; k < NEN; ++k) {
    if (t == 0) {
      int pr = k;
      for (int i = k + 1; i < NEN; ++i)
        if (fabs(A[i][k]) > fabs(A[pr][k])) pr = i;
      piv = pr;
      if (A[pr][k] == 0.0) atomicAdd(bad, 1);
    }
    __syncthreads();
    const int pr = piv;
    if (t < 2 * NEN && pr != k) {
      const double tmp = A[k][t];
      A[k][t] = A[pr][t];
      A[pr][t] = tmp;
    }
    __syncthreads();
    const double d = A[k][k];
    __syncthreads();
    if (t < 2 * NEN) A[k][t] /= d;
    __syncthreads();
    if (t < NEN && t != k) {
      const double l = A[t][k];
      for (int j = 0; j < 2 * NEN; ++j) A[t][j] -= l * A[k][j];
    }
    __syncthreads();
  }
  // stored transposed: lane i of the residual kernel reads (M^-1)_ij at
  // minv[e][j][i], so the lanes of one element read consecutive addresses
  if (t < NEN)
    for (int j = 0; j < NEN; ++j) minv[(e * NEN + j) * NEN + t] = A[t][NEN + j];
}

// ------------------------------------------------------------ residual
// AdvectionPhysics::Flux (upwind) on interior faces, BoundaryFlux (outflow
// only) on boundary faces, at every face point.  Each side of the face gets
// its face-node sums
//   slot[e][lf][k] = -+ sum_fq bN_i(fq) flux(fq) jxw_side(fq)
// at the position k of face node i on its cell's local face lf (+ for the
// right cell of an interior face, - otherwise, DGAssembly.hpp:106-189), so
// that the element pass reads its face terms from its own slots.  A face is
// handled by n1 adjacent lanes (64 / n1 faces per wavefront): lane k loads
// face node k's values and computes face point k's flux, and the lanes
// exchange them by shuffles for the interpolations and the sums.
template <int P>
__global__ __launch_bounds__(NT) void k_dg_flux(int64_t nf, const DGFaceRec *__restrict__ F,
                                               const int32_t *__restrict__ L, const int32_t *__restrict__ R,
                                               const int32_t *__restrict__ fn, const double *__restrict__ bN,
                                               const double *__restrict__ geo, double c0, double c1,
                                               const double *__restrict__ U, double *__restrict__ slot) {
  constexpr int n1 = P + 1, nen = n1 * n1, FPW = 64 / n1;
  const int lane = threadIdx.x & 63, grp = lane / n1, k = lane - grp * n1;
  const int64_t f = ((int64_t)blockIdx.x * (NT / 64) + (threadIdx.x >> 6)) * FPW + grp;
  const bool on = grp < FPW && f < nf;
  const int base = grp * n1;  // first lane of this face's group
  DGFaceRec r{-1, -1, 0, 0};
  double ul = 0, ur = 0, gx = 0, gy = 0, jl = 0, jr = 0;
  int nl = -1, nr = -1;
  if (on) {
    r = F[f];
    if (r.left >= 0) {
      nl = L[f * n1 + k];
      ul = U[(int64_t)r.left * nen + nl];
    }
    if (r.right >= 0) {
      nr = R[f * n1 + k];
      ur = U[(int64_t)r.right * nen + nr];
    }
    const double *g = &geo[4 * (f * n1 + k)];
    gx = g[0];
    gy = g[1];
    jl = g[2];
    jr = g[3];
  }
  const bool interior = r.left >= 0 && r.right >= 0;
  // u of each side at face point k (lane k), from the n1 node values
  double a = 0, b = 0;
#pragma unroll
  for (int j = 0; j < n1; ++j) {
    const double sv = bN[k * n1 + j];
    a += __shfl(ul, base + j, 64) * sv;
    b += __shfl(ur, base + j, 64) * sv;
  }
  const double vn = gx * c0 + gy * c1;  // dot(n, convection_velocity)
  double fl;
  if (interior) {
    fl = 0.5 * (vn * (a + b) + fabs(vn) * (a - b));
  } else {
    const double u = r.left >= 0 ? a : b;
    fl = vn > 0 ? vn * u : 0 * vn * u;
  }
  const double fjl = fl * jl, fjr = fl * jr;
  // lane i: the sums of face node i for each side
  double cl = 0, cr = 0;
#pragma unroll
  for (int q = 0; q < n1; ++q) {
    const double sv = bN[q * n1 + k];
    cl += sv * __shfl(fjl, base + q, 64);
    cr += sv * __shfl(fjr, base + q, 64);
  }
  if (!on) return;
  if (r.left >= 0) {
    // the left list is the left cell's own face order: position k
    slot[((int64_t)r.left * 4 + r.lf) * n1 + k] = -cl;
  }
  if (r.right >= 0) {
    int pos = 0;
#pragma unroll
    for (int j = 0; j < n1; ++j)
      if (fn[r.rf * n1 + j] == nr) pos = j;
    slot[((int64_t)r.right * 4 + r.rf) * n1 + pos] = interior ? cr : -cr;
  }
}

// Optional epilogue of the element pass for the RK4 stages (DGRK4.hpp:33-69):
// mode 1 also writes the next stage's input w = u0 + a r, mode 2 writes the
// step's result u = u0 + (dt/6) k1 + (dt/3) k2 + (dt/3) k3 + (dt/6) r (summed
// left to right) and advances the device time.  Each lane reads u0 and writes
// w or u at its own dof only, so w and u may be updated in place.
struct RkOut {
  int mode;  // 0: R only
  const double *u0;
  double *w;
  double a;
  const double *k1, *k2, *k3;
  double dt;
  double *u, *t;
};

// Element pass, a lane per (element, dof): the 64 / nen elements of a
// wavefront are handled by nen lanes each (p = 3: 4 elements of 16 dofs),
// and a workgroup loops over element groups so that it loads the shape
// tables into LDS once.  Lane (element, q) first forms u at volume point q
// (and, for a cell that is not a parallelogram, the geometry there) into
// LDS; lane (element, a) then adds the face terms of dof a from the cell's
// face slots in face order (DGAssembly.hpp:106-189), the volume term
// (LocalResidual, main.cpp:53-71) and row a of M^-1 (R_face + R_vol).
// On a parallelogram J^-T and det J are constant, so the volume term is
//   det J [ alpha sum_q w_q u_q dN_a/dxi(q) + beta sum_q w_q u_q dN_a/deta(q)
//           + src sum_q w_q N_a(q) ],
// alpha, beta = J^-T applied to the velocity, and M_e^-1 = M_ref^-1 / det J:
// about a third of the LDS reads of the general form.
template <int P>
__global__ __launch_bounds__(NT) void k_dg_elem(int64_t nel, const int32_t *__restrict__ fn,
                                               const int32_t *__restrict__ ef, const double *__restrict__ slot,
                                               const double *__restrict__ X, const double *__restrict__ N,
                                               const double *__restrict__ dN, const double *__restrict__ w,
                                               const double *__restrict__ minvT, const double *__restrict__ ginfo,
                                               const double *__restrict__ mrefT, double c0, double c1, double src,
                                               const double *__restrict__ U, double *__restrict__ Rout,
                                               double *__restrict__ Rface, RkOut rk) {
  constexpr int n1 = P + 1, nen = n1 * n1, nqp = nen, EPW = 64 / nen, WPB = NT / 64, EPB = WPB * EPW;
  constexpr int LD = nen + 1;  // padded rows: lanes reading a row or a column hit distinct banks
  __shared__ double sN[nqp][LD];
  __shared__ double sdN[nqp][LD][2];
  __shared__ double sMr[nen][LD];  // sMr[j][i] = (M_ref^-1)_ij
  __shared__ double sNw[nen];      // sum_q w_q N_a(q)
  __shared__ int sFn[4][n1];       // local face nodes
  __shared__ double sU[WPB][EPW][nen];
  __shared__ double sX[WPB][EPW][nen][2];
  __shared__ double sG[WPB][EPW][nqp][6];  // jxw, J^-T (4), u at each volume point (general cells)
  __shared__ double sWU[WPB][EPW][nqp];    // w_q u_q (parallelograms)
  __shared__ double sL[WPB][EPW][nen];
  for (int t = threadIdx.x; t < nqp * nen; t += NT) {
    const int q = t / nen, A = t - q * nen;
    sN[q][A] = N[t];
    sdN[q][A][0] = dN[2 * t];
    sdN[q][A][1] = dN[2 * t + 1];
    sMr[q][A] = mrefT[t];
  }
  for (int t = threadIdx.x; t < 4 * n1; t += NT) sFn[t / n1][t % n1] = fn[t];
  if (threadIdx.x < nen) {
    double sw = 0;
    for (int q = 0; q < nqp; ++q) sw += w[q] * N[q * nen + threadIdx.x];
    sNw[threadIdx.x] = sw;
  }
  __syncthreads();
  const int wv = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int el = l / nen, a = l - el * nen;
  // the local faces through node a and its position on each (loop-invariant)
  int lfk[2] = {0, 0}, nlf = 0;
#pragma unroll
  for (int lf = 0; lf < 4; ++lf)
#pragma unroll
    for (int j = 0; j < n1; ++j)
      if (sFn[lf][j] == a && nlf < 2) lfk[nlf++] = lf * n1 + j;
  const int64_t ngroups = (nel + EPB - 1) / EPB;
  for (int64_t grp = blockIdx.x; grp < ngroups; grp += gridDim.x) {
    const int64_t e = (grp * WPB + wv) * EPW + el;
    const bool act = el < EPW && e < nel;
    // parallelograms: det J and J^-T from the record; the others: row a of
    // their own M^-1, loaded early so its latency hides behind the face and
    // volume work, and the node coordinates
    double gi[6] = {0, 0, 0, 0, 0, 0};
    if (act)
#pragma unroll
      for (int k = 0; k < 6; ++k) gi[k] = ginfo[8 * e + k];
    const bool aff = gi[0] != 0.0;
    double mrow[nen];
    int efl[4] = {-1, -1, -1, -1};
    if (act) {
      if (!aff) {
        const double *Mt = &minvT[e * nen * nen];
#pragma unroll
        for (int j = 0; j < nen; ++j) mrow[j] = Mt[j * nen + a];
        sX[wv][el][a][0] = X[3 * (e * nen + a)];
        sX[wv][el][a][1] = X[3 * (e * nen + a) + 1];
      }
#pragma unroll
      for (int k = 0; k < 4; ++k) efl[k] = ef[4 * e + k];
      sU[wv][el][a] = U[e * nen + a];
    }
    __syncthreads();
    if (act) {
      const int q = a;
      double u = 0;
#pragma unroll
      for (int A = 0; A < nen; ++A) u += sU[wv][el][A] * sN[q][A];
      if (aff) {
        sWU[wv][el][q] = w[q] * u;
      } else {
        double J[4] = {0, 0, 0, 0};
#pragma unroll
        for (int A = 0; A < nen; ++A) {
          const double x0 = sX[wv][el][A][0], x1 = sX[wv][el][A][1];
          const double g0 = sdN[q][A][0], g1 = sdN[q][A][1];
          J[0] += x0 * g0;
          J[1] += x0 * g1;
          J[2] += x1 * g0;
          J[3] += x1 * g1;
        }
        const double det = J[0] * J[3] - J[1] * J[2];
        double *g = sG[wv][el][q];
        g[0] = det * w[q];
        g[1] = J[3] / det;
        g[2] = -J[1] / det;
        g[3] = -J[2] / det;
        g[4] = J[0] / det;
        g[5] = u;
      }
    }
    __syncthreads();
    double loc = 0, rf = 0;
    if (act) {
      // face terms of dof a from its cell's face slots, faces in the
      // reference's (global face) order; a face merged away by a
      // non-conforming mesh has no record (ef < 0)
      int fk[2], fid[2], nfa = 0;
#pragma unroll
      for (int k = 0; k < 2; ++k) {
        if (k >= nlf) continue;
        const int g = efl[lfk[k] / n1];
        if (g >= 0) {
          fk[nfa] = lfk[k];
          fid[nfa] = g >> 1;
          ++nfa;
        }
      }
      if (nfa == 2 && fid[1] < fid[0]) {
        const int t = fk[0];
        fk[0] = fk[1];
        fk[1] = t;
      }
      for (int k = 0; k < nfa; ++k) loc += slot[e * 4 * n1 + fk[k]];
      rf = loc;
      // volume term
      if (aff) {
        const double al = gi[2] * c0 + gi[3] * c1, be = gi[4] * c0 + gi[5] * c1;
        double s0 = 0, s1 = 0;
#pragma unroll
        for (int q = 0; q < nqp; ++q) {
          const double wu = sWU[wv][el][q];
          s0 += wu * sdN[q][a][0];
          s1 += wu * sdN[q][a][1];
        }
        loc += gi[1] * (al * s0 + be * s1 + src * sNw[a]);
#pragma unroll
        for (int j = 0; j < nen; ++j) mrow[j] = sMr[j][a];
      } else {
#pragma unroll 4
        for (int q = 0; q < nqp; ++q) {
          const double *g = sG[wv][el][q];
          const double g0 = sdN[q][a][0], g1 = sdN[q][a][1];
          const double gx = g[1] * g0 + g[3] * g1, gy = g[2] * g0 + g[4] * g1;
          loc += (g[5] * (gx * c0 + gy * c1) + src * sN[q][a]) * g[0];
        }
      }
      sL[wv][el][a] = loc;
    }
    __syncthreads();
    if (act) {
      // R = R_face + M^-1 (R_face + R_vol), row a
      double s = 0;
#pragma unroll
      for (int j = 0; j < nen; ++j) s += mrow[j] * sL[wv][el][j];
      if (aff) s /= gi[1];
      const double r = rf + s;
      const int64_t o = e * nen + a;
      if (rk.mode == 2) {
        rk.u[o] = rk.u0[o] + (rk.dt / 6) * rk.k1[o] + (rk.dt / 3) * rk.k2[o] + (rk.dt / 3) * rk.k3[o] + (rk.dt / 6) * r;
        if (o == 0) *rk.t += rk.dt;
      } else {
        Rout[o] = r;
        if (rk.mode == 1) rk.w[o] = rk.u0[o] + rk.a * r;
      }
      if (Rface) Rface[o] = rf;
    }
  }
}

// ------------------------------------------------------------------ RK4
struct DGPhys {
  double c0, c1, src;
};

DGPhys dg_phys(const yafel_hip_physics *phys) {
  REQUIRE(phys, "dg: null physics");
  REQUIRE(phys->kind == YAFEL_HIP_PHYS_DG_ADVECTION && phys->nsd == 2,
          "dg: physics must be the 2-D DG advection descriptor");
  return DGPhys{phys->U[0], phys->U[1], phys->f};
}

void dg_eval(yafel_hip_dg_s *D, const DGPhys &ph, const double *U, double *R, double *Rface, hipStream_t s,
             RkOut rk = RkOut{}) {
#define DG_ELEM(P)                                                                                               \
  k_dg_flux<P><<<ceil_div(D->nfaces, (NT / 64) * (64 / (P + 1))), NT, 0, s>>>(D->nfaces, D->faces.as<DGFaceRec>(), D->lnodes.as<int32_t>(), \
                                                      D->rnodes.as<int32_t>(), D->fnodes.as<int32_t>(),           \
                                                      D->bN.as<double>(), D->geo.as<double>(), ph.c0, ph.c1, U,   \
                                                      D->flux.as<double>());                                      \
  k_dg_elem<P><<<(unsigned)std::min<int64_t>(ceil_div(D->nel, (NT / 64) * (64 / ((P + 1) * (P + 1)))), \
                                               (int64_t)D->ctx->num_cus * 4),                                   \
                 NT, 0, s>>>(D->nel, D->fnodes.as<int32_t>(),                                                  \
                                D->ef.as<int32_t>(), D->flux.as<double>(), \
                                D->X.as<double>(), D->tabN.as<double>(), D->tabdN.as<double>(), D->tabw.as<double>(), \
                                D->minv.as<double>(), D->ginfo.as<double>(), D->mrefT.as<double>(), ph.c0, ph.c1, \
                                ph.src, U, R, Rface, rk)
  switch (D->p) {
    case 1: DG_ELEM(1); break;
    case 2: DG_ELEM(2); break;
    case 3: DG_ELEM(3); break;
    default: DG_ELEM(4); break;
  }
#undef DG_ELEM
  HIPCHK(hipGetLastError());
}

// One DGRK4::step on U (stream-ordered kernels only, so it can be captured).
void dg_rk4_step(yafel_hip_dg_s *D, const DGPhys &ph, double dt, double *U, hipStream_t s) {
  const int64_t n = D->nel * D->nen;
  double *U0 = D->rk.as<double>(), *k1 = U0 + n, *k2 = k1 + n, *k3 = k2 + n, *k4 = k3 + n, *W = k4 + n;
  (void)U0;
  (void)k4;
  // the stage inputs and the result come from the element passes' epilogues:
  // W = U + (dt/2) k1, W = U + (dt/2) k2, W = U + dt k3, then U itself
  RkOut st{1, U, W, dt / 2, nullptr, nullptr, nullptr, 0.0, nullptr, nullptr};
  dg_eval(D, ph, U, k1, nullptr, s, st);
  dg_eval(D, ph, W, k2, nullptr, s, st);
  st.a = dt;
  dg_eval(D, ph, W, k3, nullptr, s, st);
  RkOut fin{2, U, nullptr, 0.0, k1, k2, k3, dt, U, D->tdev.as<double>()};
  dg_eval(D, ph, W, nullptr, nullptr, s, fin);
  HIPCHK(hipGetLastError());
}

template <class T>
void upload(DBuf &b, const std::vector<T> &v, hipStream_t s) {
  b.alloc(std::max<size_t>(v.size(), 1) * sizeof(T));
  if (!v.empty()) HIPCHK(hipMemcpyAsync(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s));
}

}  // namespace

extern "C" {

int yafel_hip_dg_create(yafel_hip_ctx ctx, int p, int64_t n_corner, const double *corner_nodes, int64_t ncell,
                        const int32_t *cells, yafel_hip_dg *out) {
  YH_TRY_BEGIN
  REQUIRE(ctx && out && corner_nodes && cells, "dg_create: null argument");
  REQUIRE(p >= 1 && p <= PMAX, "dg_create: polynomial order must be 1..4");
  REQUIRE(n_corner > 0 && ncell > 0, "dg_create: empty mesh");
  REQUIRE(ncell * 4 < (int64_t)INT32_MAX / 2 && n_corner < (int64_t)INT32_MAX, "dg_create: mesh too large");
  for (int64_t i = 0; i < 4 * ncell; ++i)
    REQUIRE(cells[i] >= 0 && cells[i] < n_corner, "dg_create: cell references a node outside corner_nodes");
  HIPCHK(hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  std::unique_ptr<yafel_hip_dg_s> D(new yafel_hip_dg_s());
  ctx_retain(ctx);
  D->ctx = ctx;
  D->p = p;
  D->n1 = p + 1;
  D->nen = D->n1 * D->n1;
  D->nqp = D->nen;
  D->nel = ncell;
  const int n1 = D->n1, nen = D->nen;
  // element and face tables
  std::vector<double> N, dN, w, xl, bN, bdN, bw, bxl;
  element_tables(2, p, N, dN, w, xl);
  element_tables(1, p, bN, bdN, bw, bxl);
  upload(D->tabN, N, s);
  upload(D->tabdN, dN, s);
  upload(D->tabw, w, s);
  upload(D->bN, bN, s);
  upload(D->bdN, bdN, s);
  upload(D->bw, bw, s);
  // Element::build_element_faces (element_boundary.cpp:82-93): node of face f
  // whose face coordinate is boundary node i; faces run counter-clockwise
  std::vector<int32_t> fn(4 * n1, -1);
  for (int a = 0; a < nen; ++a) {
    const double x = xl[3 * a], y = xl[3 * a + 1];
    for (int f = 0; f < 4; ++f) {
      bool on;
      double sc;
      if (f == 0) { on = std::fabs(y + 1) < 1e-6; sc = x; }
      else if (f == 1) { on = std::fabs(x - 1) < 1e-6; sc = y; }
      else if (f == 2) { on = std::fabs(y - 1) < 1e-6; sc = -x; }
      else { on = std::fabs(x + 1) < 1e-6; sc = -y; }
      if (!on) continue;
      for (int i = 0; i < n1; ++i)
        if (std::fabs(sc - bxl[3 * i]) < 1e-6) fn[f * n1 + i] = a;
    }
  }
  for (int v : fn) REQUIRE(v >= 0, "dg_create: face node table incomplete");
  upload(D->fnodes, fn, s);
  // dof nodes
  DBuf dcorner, dcells, dxl;
  dcorner.alloc(n_corner * 24);
  dcells.alloc(ncell * 16);
  HIPCHK(hipMemcpyAsync(dcorner.p, corner_nodes, n_corner * 24, hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(dcells.p, cells, ncell * 16, hipMemcpyHostToDevice, s));
  upload(dxl, xl, s);
  const int64_t nd = ncell * nen;
  D->X.alloc(nd * 24);
  k_dg_nodes<<<ceil_div(nd, NT), NT, 0, s>>>(ncell, nen, dcorner.as<double>(), dcells.as<int32_t>(),
                                             dxl.as<double>(), D->X.as<double>());
  HIPCHK(hipGetLastError());
  // faces: sort the 4 ncell (cell, local face) entries by corner pair
  const int64_t T = 4 * ncell;
  DBuf key, val, head, fid, bad;
  key.alloc(T * 8);
  val.alloc(T * 4);
  head.alloc(T * 4);
  fid.alloc(T * 4);
  bad.alloc(8);
  HIPCHK(hipMemsetAsync(bad.p, 0, 8, s));
  k_dg_face_keys<<<ceil_div(T, NT), NT, 0, s>>>(ncell, n_corner, dcells.as<int32_t>(), key.as<uint64_t>(),
                                                val.as<uint32_t>());
  HIPCHK(hipGetLastError());
  sort_pairs_u64_u32(key.as<uint64_t>(), val.as<uint32_t>(), T,
                     bits_for((uint64_t)(n_corner - 1) * (uint64_t)n_corner + (uint64_t)(n_corner - 1)), s);
  k_dg_heads<<<ceil_div(T, NT), NT, 0, s>>>(T, key.as<uint64_t>(), head.as<int32_t>());
  HIPCHK(hipGetLastError());
  inclusive_scan_i32(head.as<int32_t>(), fid.as<int32_t>(), T, s);
  int32_t nf = 0;
  HIPCHK(hipMemcpyAsync(&nf, fid.as<int32_t>() + T - 1, 4, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  D->nfaces = nf;
  D->faces.alloc((int64_t)nf * sizeof(DGFaceRec));
  k_dg_face_init<<<ceil_div(nf, NT), NT, 0, s>>>(nf, D->faces.as<DGFaceRec>());
  k_dg_face_fill<<<ceil_div(T, NT), NT, 0, s>>>(T, val.as<uint32_t>(), fid.as<int32_t>(), D->faces.as<DGFaceRec>());
  HIPCHK(hipGetLastError());
  D->lnodes.alloc((int64_t)nf * n1 * 4);
  D->rnodes.alloc((int64_t)nf * n1 * 4);
  k_dg_match<<<ceil_div(nf, NT), NT, 0, s>>>(nf, n1, nen, D->faces.as<DGFaceRec>(), D->fnodes.as<int32_t>(),
                                             D->X.as<double>(), D->lnodes.as<int32_t>(), D->rnodes.as<int32_t>(),
                                             bad.as<int>());
  HIPCHK(hipGetLastError());
  D->geo.alloc((int64_t)nf * n1 * 32);
  k_dg_geom<<<ceil_div((int64_t)nf * n1, NT), NT, 0, s>>>(nf, n1, nen, D->faces.as<DGFaceRec>(),
                                                           D->lnodes.as<int32_t>(), D->rnodes.as<int32_t>(),
                                                           D->X.as<double>(), D->bdN.as<double>(), D->bw.as<double>(),
                                                           D->geo.as<double>());
  HIPCHK(hipGetLastError());
  D->ef.alloc(ncell * 16);
  HIPCHK(hipMemsetAsync(D->ef.p, 0xff, ncell * 16, s));
  k_dg_elem_faces<<<ceil_div(nf, NT), NT, 0, s>>>(nf, D->faces.as<DGFaceRec>(), D->ef.as<int32_t>());
  HIPCHK(hipGetLastError());
  D->minv.alloc(ncell * nen * nen * 8);
#define DG_MINV(NEN)                                                                                              \
  k_dg_minv<NEN><<<(unsigned)ncell, 64, 0, s>>>(ncell, D->nqp, D->X.as<double>(), D->tabN.as<double>(),           \
                                                D->tabdN.as<double>(), D->tabw.as<double>(), D->minv.as<double>(), \
                                                bad.as<int>() + 1)
  switch (p) {
    case 1: DG_MINV(4); break;
    case 2: DG_MINV(9); break;
    case 3: DG_MINV(16); break;
    default: DG_MINV(25); break;
  }
#undef DG_MINV
  HIPCHK(hipGetLastError());
  // reference mass matrix sum_q w_q N_i N_j and its inverse (Gauss-Jordan with
  // partial pivoting), for the affine cells
  {
    std::vector<double> A((size_t)nen * 2 * nen, 0.0), mt((size_t)nen * nen);
    for (int i = 0; i < nen; ++i) {
      for (int j = 0; j < nen; ++j) {
        double m = 0;
        for (int q = 0; q < D->nqp; ++q) m += w[q] * N[q * nen + i] * N[q * nen + j];
        A[(size_t)i * 2 * nen + j] = m;
      }
      A[(size_t)i * 2 * nen + nen + i] = 1.0;
    }
    for (int k = 0; k < nen; ++k) {
      int pr = k;
      for (int i = k + 1; i < nen; ++i)
        if (std::fabs(A[(size_t)i * 2 * nen + k]) > std::fabs(A[(size_t)pr * 2 * nen + k])) pr = i;
      REQUIRE(A[(size_t)pr * 2 * nen + k] != 0.0, "dg_create: singular reference mass matrix");
      if (pr != k)
        for (int j = 0; j < 2 * nen; ++j) std::swap(A[(size_t)k * 2 * nen + j], A[(size_t)pr * 2 * nen + j]);
      const double d = A[(size_t)k * 2 * nen + k];
      for (int j = 0; j < 2 * nen; ++j) A[(size_t)k * 2 * nen + j] /= d;
      for (int i = 0; i < nen; ++i) {
        if (i == k) continue;
        const double l = A[(size_t)i * 2 * nen + k];
        for (int j = 0; j < 2 * nen; ++j) A[(size_t)i * 2 * nen + j] -= l * A[(size_t)k * 2 * nen + j];
      }
    }
    for (int i = 0; i < nen; ++i)
      for (int j = 0; j < nen; ++j) mt[(size_t)j * nen + i] = A[(size_t)i * 2 * nen + nen + j];
    upload(D->mrefT, mt, s);
  }
  D->ginfo.alloc(ncell * 64);
  k_dg_affine<<<ceil_div(ncell, NT), NT, 0, s>>>(ncell, n1, nen, D->X.as<double>(), D->tabdN.as<double>(),
                                                 D->tabw.as<double>(), D->ginfo.as<double>());
  HIPCHK(hipGetLastError());
  int hb[2] = {0, 0};
  HIPCHK(hipMemcpyAsync(hb, bad.p, 8, hipMemcpyDeviceToHost, s));
  // boundary face count, for the info call
  std::vector<DGFaceRec> hf(nf);
  HIPCHK(hipMemcpyAsync(hf.data(), D->faces.p, (size_t)nf * sizeof(DGFaceRec), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  REQUIRE(hb[0] == 0, "dg_create: faces whose nodes do not match across the two cells (mesh not conforming)");
  REQUIRE(hb[1] == 0, "dg_create: singular element mass matrix");
  for (const auto &f : hf) D->nbnd += (f.left < 0 || f.right < 0) ? 1 : 0;
  {
    std::vector<double> gh((size_t)ncell * 8);
    HIPCHK(hipMemcpyAsync(gh.data(), D->ginfo.p, gh.size() * 8, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    for (int64_t e = 0; e < ncell; ++e) D->naffine += gh[(size_t)e * 8] != 0.0 ? 1 : 0;
  }
  D->flux.alloc(ncell * 4 * n1 * 8);
  D->rk.alloc(6 * nd * 8);
  D->tdev.alloc(8);
  HIPCHK(hipMemsetAsync(D->tdev.p, 0, 8, s));
  HIPCHK(hipStreamSynchronize(s));
  *out = D.release();
  return YAFEL_HIP_OK;
  YH_TRY_END(ctx)
}

int yafel_hip_dg_destroy(yafel_hip_dg D) {
  if (!D) return YAFEL_HIP_OK;
  (void)hipStreamSynchronize(D->ctx->stream);
  delete D;
  return YAFEL_HIP_OK;
}

int yafel_hip_dg_info(yafel_hip_dg D, int64_t *info) {
  YH_TRY_BEGIN
  REQUIRE(D && info, "dg_info: null argument");
  info[0] = D->nel * D->nen;
  info[1] = D->nfaces;
  info[2] = D->nbnd;
  info[3] = D->nen;
  return YAFEL_HIP_OK;
  YH_TRY_END(nullptr)
}

int yafel_hip_dg_affine_cells(yafel_hip_dg D, int64_t *count) {
  YH_TRY_BEGIN
  REQUIRE(D && count, "dg_affine_cells: null argument");
  *count = D->naffine;
  return YAFEL_HIP_OK;
  YH_TRY_END(nullptr)
}

int yafel_hip_dg_nodes(yafel_hip_dg D, double *nodes_dev) {
  YH_TRY_BEGIN
  REQUIRE(D && nodes_dev, "dg_nodes: null argument");
  HIPCHK(hipMemcpyAsync(nodes_dev, D->X.p, D->nel * D->nen * 24, hipMemcpyDeviceToDevice, D->ctx->stream));
  return YAFEL_HIP_OK;
  YH_TRY_END(D ? D->ctx : nullptr)
}

int yafel_hip_dg_faces(yafel_hip_dg D, int32_t *left, int32_t *right, int32_t *left_flocal, int32_t *right_flocal) {
  YH_TRY_BEGIN
  REQUIRE(D, "dg_faces: null handle");
  std::vector<DGFaceRec> hf(D->nfaces);
  HIPCHK(hipMemcpyAsync(hf.data(), D->faces.p, hf.size() * sizeof(DGFaceRec), hipMemcpyDeviceToHost, D->ctx->stream));
  HIPCHK(hipStreamSynchronize(D->ctx->stream));
  for (size_t i = 0; i < hf.size(); ++i) {
    if (left) left[i] = hf[i].left;
    if (right) right[i] = hf[i].right;
    if (left_flocal) left_flocal[i] = hf[i].lf;
    if (right_flocal) right_flocal[i] = hf[i].rf;
  }
  return YAFEL_HIP_OK;
  YH_TRY_END(D ? D->ctx : nullptr)
}

int yafel_hip_dg_residual(yafel_hip_dg D, const yafel_hip_physics *phys, double time, const double *U_dev,
                          double *R_dev, double *Rface_dev) {
  YH_TRY_BEGIN
  REQUIRE(D && U_dev && R_dev, "dg_residual: null argument");
  REQUIRE(U_dev != R_dev && (void *)U_dev != (void *)Rface_dev, "dg_residual: U must not alias R");
  const DGPhys ph = dg_phys(phys);
  (void)time;  // the advection physics is time-independent
  HIPCHK(hipSetDevice(D->ctx->device));
  dg_eval(D, ph, U_dev, R_dev, Rface_dev, D->ctx->stream);
  return YAFEL_HIP_OK;
  YH_TRY_END(D ? D->ctx : nullptr)
}

int yafel_hip_dg_rk4(yafel_hip_dg D, const yafel_hip_physics *phys, double dt, int64_t nsteps, double *time,
                     double *U_dev) {
  YH_TRY_BEGIN
  REQUIRE(D && U_dev && nsteps >= 0, "dg_rk4: bad argument");
  const DGPhys ph = dg_phys(phys);
  const double key[4] = {dt, ph.c0, ph.c1, ph.src};
  HIPCHK(hipSetDevice(D->ctx->device));
  hipStream_t s = D->ctx->stream;
  const double t0 = time ? *time : 0.0;
  HIPCHK(hipMemcpyAsync(D->tdev.p, &t0, 8, hipMemcpyHostToDevice, s));
  if (nsteps > 0 && (!D->graph || std::memcmp(D->graph_key, key, sizeof(key)) != 0 || D->graph_U != U_dev)) {
    if (D->graph) {
      HIPCHK(hipGraphExecDestroy(D->graph));
      D->graph = nullptr;
    }
    hipGraph_t g = nullptr;
    HIPCHK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    try {
      dg_rk4_step(D, ph, dt, U_dev, s);
    } catch (...) {
      (void)hipStreamEndCapture(s, &g);
      if (g) (void)hipGraphDestroy(g);
      throw;
    }
    HIPCHK(hipStreamEndCapture(s, &g));
    hipError_t ie = hipGraphInstantiate(&D->graph, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    HIPCHK(ie);
    std::memcpy(D->graph_key, key, sizeof(key));
    D->graph_U = U_dev;
  }
  for (int64_t k = 0; k < nsteps; ++k) HIPCHK(hipGraphLaunch(D->graph, s));
  double t1 = t0;
  HIPCHK(hipMemcpyAsync(&t1, D->tdev.p, 8, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  if (time) *time = t1;
  return YAFEL_HIP_OK;
  YH_TRY_END(D ? D->ctx : nullptr)
}

}  // extern "C"
