// DoFManager on the device (src/utils/DoFManager.cpp:143-229, 296-414), element
// tables, and the sparsity pattern + deterministic assembly gather plan.
#include <algorithm>
#include <cmath>

#include "internal.hpp"
#include "mesh.hpp"
#include "sortprim.hpp"

using namespace yh;

namespace {

constexpr int NT = 256;

// ------------------------------------------------------- element tables
// 1-D Gauss-Lobatto points (ascending): endpoints and the roots of P'_{N-1},
// by Newton's method on the Legendre three-term recurrence.
void gll_points(int N, std::vector<double> &x) {
  x.assign(N, 0.0);
  const double PI = 3.14159265358979323846;
  const int n = N - 1;
  for (int i = 0; i < N; ++i) {
    double t = -std::cos(PI * i / n);
    if (i > 0 && i < n) {
      for (int it = 0; it < 100; ++it) {
        // P_n and P_{n-1}
        double p0 = 1, p1 = t;
        for (int k = 2; k <= n; ++k) {
          double p2 = ((2 * k - 1) * t * p1 - (k - 1) * p0) / k;
          p0 = p1;
          p1 = p2;
        }
        // f = P'_n, f' = P''_n from the Legendre ODE
        const double dP = n * (t * p1 - p0) / (t * t - 1);
        const double d2P = (2 * t * dP - n * (n + 1) * p1) / (1 - t * t);
        const double dt = dP / d2P;
        t -= dt;
        if (std::abs(dt) < 1e-16) break;
      }
    }
    x[i] = t;
  }
}

// 1-D Gauss-Legendre points and weights (ascending).
void gl_points(int N, std::vector<double> &x, std::vector<double> &w) {
  x.assign(N, 0.0);
  w.assign(N, 0.0);
  const double PI = 3.14159265358979323846;
  for (int i = 0; i < N; ++i) {
    double t = -std::cos(PI * (i + 0.75) / (N + 0.5));
    double dP = 1;
    for (int it = 0; it < 100; ++it) {
      double p0 = 1, p1 = t;
      for (int k = 2; k <= N; ++k) {
        double p2 = ((2 * k - 1) * t * p1 - (k - 1) * p0) / k;
        p0 = p1;
        p1 = p2;
      }
      dP = N * (t * p1 - p0) / (t * t - 1);
      const double dt = p1 / dP;
      t -= dt;
      if (std::abs(dt) < 1e-16) break;
    }
    double p0 = 1, p1 = t;
    for (int k = 2; k <= N; ++k) {
      double p2 = ((2 * k - 1) * t * p1 - (k - 1) * p0) / k;
      p0 = p1;
      p1 = p2;
    }
    dP = N * (t * p1 - p0) / (t * t - 1);
    x[i] = t;
    w[i] = 2.0 / ((1 - t * t) * dP * dP);
  }
}

// Lagrange basis l_i on nodes xs and its derivative at t.
void lagrange(const std::vector<double> &xs, int i, double t, double &v, double &dv) {
  const int n = (int)xs.size();
  v = 1;
  for (int m = 0; m < n; ++m)
    if (m != i) v *= (t - xs[m]) / (xs[i] - xs[m]);
  dv = 0;
  for (int k = 0; k < n; ++k) {
    if (k == i) continue;
    double term = 1.0 / (xs[i] - xs[k]);
    for (int m = 0; m < n; ++m)
      if (m != i && m != k) term *= (t - xs[m]) / (xs[i] - xs[m]);
    dv += term;
  }
}

// ---------------------------------------------------------------- kernels
// raw nodes: every cell's local GLL nodes mapped from its corners
// (interpolate_from_corners, DoFManager.cpp:313-331; gmsh corner order).
__global__ void k_raw_nodes(int64_t ncell, int dim, int nen, const double *__restrict__ xl,
                            const int32_t *__restrict__ cells, const double *__restrict__ corners,
                            double *__restrict__ raw) {
  int64_t i = blockIdx.x * (int64_t)NT + threadIdx.x;
  if (i >= ncell * nen) return;
  const int64_t c = i / nen;
  const int a = (int)(i - c * nen);
  const double x0 = xl[3 * a], x1 = xl[3 * a + 1], x2 = xl[3 * a + 2];
  const int nc = dim == 2 ? 4 : 8;
  const int32_t *cc = &cells[c * nc];
  for (int d = 0; d < 3; ++d) {
    double o;
    if (dim == 2) {
      const double s0 = (1 - x0) * (1 - x1), s1 = (1 + x0) * (1 - x1);
      const double s2 = (1 + x0) * (1 + x1), s3 = (1 - x0) * (1 + x1);
      o = (1.0 / 4.0) * (s0 * corners[3 * (int64_t)cc[0] + d] + s1 * corners[3 * (int64_t)cc[1] + d] +
                         s2 * corners[3 * (int64_t)cc[2] + d] + s3 * corners[3 * (int64_t)cc[3] + d]);
    } else {
      double s[8];
      s[0] = (1 - x0) * (1 - x1) * (1 - x2);
      s[1] = (1 + x0) * (1 - x1) * (1 - x2);
      s[2] = (1 + x0) * (1 + x1) * (1 - x2);
      s[3] = (1 - x0) * (1 + x1) * (1 - x2);
      s[4] = (1 - x0) * (1 - x1) * (1 + x2);
      s[5] = (1 + x0) * (1 - x1) * (1 + x2);
      s[6] = (1 + x0) * (1 + x1) * (1 + x2);
      s[7] = (1 - x0) * (1 + x1) * (1 + x2);
      double acc = s[0] * corners[3 * (int64_t)cc[0] + d];
      for (int k = 1; k < 8; ++k) acc += s[k] * corners[3 * (int64_t)cc[k] + d];
      o = (1.0 / 8.0) * acc;
    }
    raw[3 * i + d] = o;
  }
}

__global__ void k_bbox(int64_t n, const double *__restrict__ raw, double *part) {
  __shared__ double smin[3][NT], smax[3][NT];
  double mn[3] = {1e300, 1e300, 1e300}, mx[3] = {-1e300, -1e300, -1e300};
  for (int64_t i = blockIdx.x * (int64_t)NT + threadIdx.x; i < n; i += (int64_t)gridDim.x * NT)
    for (int d = 0; d < 3; ++d) {
      mn[d] = fmin(mn[d], raw[3 * i + d]);
      mx[d] = fmax(mx[d], raw[3 * i + d]);
    }
  for (int d = 0; d < 3; ++d) {
    smin[d][threadIdx.x] = mn[d];
    smax[d][threadIdx.x] = mx[d];
  }
  __syncthreads();
  if (threadIdx.x < 3) {
    const int d = threadIdx.x;
    double a = 1e300, b = -1e300;
    for (int t = 0; t < NT; ++t) {
      a = fmin(a, smin[d][t]);
      b = fmax(b, smax[d][t]);
    }
    part[6 * blockIdx.x + d] = a;
    part[6 * blockIdx.x + 3 + d] = b;
  }
}

// snapped (z, y, x) sort key (recombine_all_duplicates, DoFManager.cpp:353-366):
// round(x/snap) per axis, offset by the bounding box, packed z-major.
__global__ void k_snap_keys(int64_t n, const double *__restrict__ raw, double snap, long long k0x,
                            long long k0y, long long k0z, int bx, int by, uint64_t *key, uint32_t *idx) {
  int64_t i = blockIdx.x * (int64_t)NT + threadIdx.x;
  if (i >= n) return;
  const uint64_t kx = (uint64_t)((long long)round(raw[3 * i] / snap) - k0x);
  const uint64_t ky = (uint64_t)((long long)round(raw[3 * i + 1] / snap) - k0y);
  const uint64_t kz = (uint64_t)((long long)round(raw[3 * i + 2] / snap) - k0z);
  key[i] = (((kz << by) | ky) << bx) | kx;
  idx[i] = (uint32_t)i;
}

// a new dof node starts where the sorted neighbour is not within 1e-6
// (uniq_func_x, DoFManager.cpp:368, 383-390)
__global__ void k_uniq_flags(int64_t n, const uint32_t *idx, const double *raw, int32_t *flag) {
  int64_t i = blockIdx.x * (int64_t)NT + threadIdx.x;
  if (i >= n) return;
  if (i == 0) {
    flag[0] = 1;
    return;
  }
  const uint32_t a = idx[i], b = idx[i - 1];
  double s = 0;
  for (int d = 0; d < 3; ++d) {
    const double t = raw[3 * (int64_t)a + d] - raw[3 * (int64_t)b + d];
    s += t * t;
  }
  flag[i] = sqrt(s) < 1.0e-6 ? 0 : 1;
}

__global__ void k_assign_nodes(int64_t n, const uint32_t *idx, const int32_t *flag, const int32_t *seg,
                               const double *raw, int32_t *elements, double *nodes) {
  int64_t i = blockIdx.x * (int64_t)NT + threadIdx.x;
  if (i >= n) return;
  const int32_t id = seg[i] - 1;
  elements[idx[i]] = id;
  if (flag[i])
    for (int d = 0; d < 3; ++d) nodes[3 * (int64_t)id + d] = raw[3 * (int64_t)idx[i] + d];
}

__global__ void k_check_cells(int64_t n, const int32_t *cells, int64_t nnode, int32_t *bad) {
  int64_t i = blockIdx.x * (int64_t)NT + threadIdx.x;
  if (i < n && (cells[i] < 0 || cells[i] >= nnode)) *bad = 1;
}

// node-pair keys of every element pair: key = A*n + B, value t = e*nen^2 + a*nen + b
__global__ void k_pair_keys(int64_t nel, int nen, const int32_t *__restrict__ el, uint64_t nn,
                            uint64_t *key, uint32_t *val) {
  int64_t t = blockIdx.x * (int64_t)NT + threadIdx.x;
  const int64_t nen2 = (int64_t)nen * nen;
  if (t >= nel * nen2) return;
  const int64_t e = t / nen2;
  const int r = (int)(t - e * nen2);
  const int a = r / nen, b = r - (r / nen) * nen;
  key[t] = (uint64_t)el[e * nen + a] * nn + (uint64_t)el[e * nen + b];
  val[t] = (uint32_t)t;
}

__global__ void k_node_keys(int64_t nel, int nen, const int32_t *__restrict__ el, uint64_t *key,
                            uint32_t *val) {
  int64_t t = blockIdx.x * (int64_t)NT + threadIdx.x;
  if (t >= nel * nen) return;
  key[t] = (uint64_t)el[t];
  val[t] = (uint32_t)t;
}

__global__ void k_flags(int64_t n, const uint64_t *key, int32_t *flag) {
  int64_t i = blockIdx.x * (int64_t)NT + threadIdx.x;
  if (i < n) flag[i] = (i == 0 || key[i] != key[i - 1]) ? 1 : 0;
}

__global__ void k_run_heads(int64_t n, const int32_t *flag, const int32_t *seg, int32_t *cptr) {
  int64_t i = blockIdx.x * (int64_t)NT + threadIdx.x;
  if (i < n && flag[i]) cptr[seg[i] - 1] = (int32_t)i;
}

// per run: (row, col) of the node pair and the node row pointer
__global__ void k_runs(int64_t nruns, const int32_t *cptr, const uint64_t *key, uint64_t nn,
                       int32_t *run_row, int32_t *run_col, int32_t *node_ptr) {
  int64_t j = blockIdx.x * (int64_t)NT + threadIdx.x;
  if (j >= nruns) return;
  const uint64_t k = key[cptr[j]];
  const int64_t A = (int64_t)(k / nn);
  run_row[j] = (int32_t)A;
  run_col[j] = (int32_t)(k % nn);
  const int64_t prev = j ? (int64_t)(key[cptr[j - 1]] / nn) : -1;
  for (int64_t a = prev + 1; a <= A; ++a) node_ptr[a] = (int32_t)j;
}

// node incidence row pointer from sorted node keys
__global__ void k_inc_ptr(int64_t n, const uint64_t *key, int32_t *ptr) {
  int64_t i = blockIdx.x * (int64_t)NT + threadIdx.x;
  if (i >= n) return;
  const int64_t A = (int64_t)key[i];
  const int64_t prev = i ? (int64_t)key[i - 1] : -1;
  for (int64_t a = prev + 1; a <= A; ++a) ptr[a] = (int32_t)i;
}

__global__ void k_fill_tail(int32_t *ptr, int64_t from, int64_t to, int32_t v) {
  int64_t i = from + blockIdx.x * (int64_t)NT + threadIdx.x;
  if (i <= to) ptr[i] = v;
}

// scalar CSR (dpn > 1) structure from the node pattern: row (A,c) holds
// (B,c') for every neighbour B in order, c' fastest.
__global__ void k_scalar_csr(int64_t nruns, int dpn, const int32_t *run_row, const int32_t *run_col,
                             const int32_t *node_ptr, int32_t *col, int32_t *spos) {
  int64_t j = blockIdx.x * (int64_t)NT + threadIdx.x;
  if (j >= nruns) return;
  const int A = run_row[j], B = run_col[j];
  const int64_t s0 = node_ptr[A], len = node_ptr[A + 1] - s0;
  const int64_t k = j - s0;
  const int64_t base = (int64_t)dpn * dpn * s0;
  spos[j] = (int32_t)(base + k * dpn);
  for (int c = 0; c < dpn; ++c)
    for (int c2 = 0; c2 < dpn; ++c2) col[base + c * dpn * len + k * dpn + c2] = B * dpn + c2;
}

__global__ void k_scalar_ptr(int64_t nnodes, int dpn, const int32_t *node_ptr, int32_t *ptr) {
  int64_t A = blockIdx.x * (int64_t)NT + threadIdx.x;
  if (A >= nnodes) return;
  const int64_t s0 = node_ptr[A], len = node_ptr[A + 1] - s0;
  for (int c = 0; c < dpn; ++c) ptr[A * dpn + c] = (int32_t)((int64_t)dpn * dpn * s0 + c * dpn * len);
  if (A == nnodes - 1) ptr[nnodes * dpn] = (int32_t)((int64_t)dpn * dpn * node_ptr[nnodes]);
}

__global__ void k_bc_plane(int64_t nnodes, int dpn, const double *nodes, int axis, double value,
                           double tol, int comp, uint8_t *mask) {
  int64_t A = blockIdx.x * (int64_t)NT + threadIdx.x;
  if (A >= nnodes) return;
  if (fabs(nodes[3 * A + axis] - value) < tol) {
    if (comp >= 0) mask[A * dpn + comp] = 1;
    else
      for (int c = 0; c < dpn; ++c) mask[A * dpn + c] = 1;
  }
}

}  // namespace

namespace yh {

void element_tables(int dim, int p, std::vector<double> &N, std::vector<double> &dN,
                    std::vector<double> &w, std::vector<double> &xl) {
  const int n1 = p + 1;
  std::vector<double> gll, gx, gw;
  gll_points(n1, gll);
  gl_points(n1, gx, gw);  // (p+1) Gauss points per axis integrate degree 2p exactly
  int nen = 1;
  for (int d = 0; d < dim; ++d) nen *= n1;
  const int nqp = nen;
  N.assign((size_t)nqp * nen, 0.0);
  dN.assign((size_t)nqp * nen * dim, 0.0);
  w.assign(nqp, 0.0);
  xl.assign((size_t)nen * 3, 0.0);
  auto split = [&](int a, int idx[3]) {
    idx[0] = a % n1;
    idx[1] = (a / n1) % n1;
    idx[2] = a / (n1 * n1);
  };
  for (int a = 0; a < nen; ++a) {
    int ia[3];
    split(a, ia);
    for (int d = 0; d < dim; ++d) xl[3 * a + d] = gll[ia[d]];
  }
  for (int q = 0; q < nqp; ++q) {
    int iq[3];
    split(q, iq);
    double wq = 1;
    for (int d = 0; d < dim; ++d) wq *= gw[iq[d]];
    w[q] = wq;
    for (int a = 0; a < nen; ++a) {
      int ia[3];
      split(a, ia);
      double v[3], dv[3];
      for (int d = 0; d < dim; ++d) lagrange(gll, ia[d], gx[iq[d]], v[d], dv[d]);
      double pv = 1;
      for (int d = 0; d < dim; ++d) pv *= v[d];
      N[(size_t)q * nen + a] = pv;
      for (int g = 0; g < dim; ++g) {
        double pg = 1;
        for (int d = 0; d < dim; ++d) pg *= (d == g) ? dv[d] : v[d];
        dN[((size_t)q * nen + a) * dim + g] = pg;
      }
    }
  }
}

void mesh_build_incidence(yafel_hip_mesh_s *m) {
  if (m->inc_ok) return;
  hipStream_t s = m->ctx->stream;
  const int64_t T = m->nel * m->nen;
  DBuf key, val;
  key.alloc(std::max<int64_t>(T, 1) * 8);
  m->inc.alloc(std::max<int64_t>(T, 1) * 4);
  m->inc_ptr.alloc((m->n_nodes + 1) * 4);
  k_node_keys<<<ceil_div(T, NT), NT, 0, s>>>(m->nel, m->nen, m->elements.as<int32_t>(), key.as<uint64_t>(),
                                             m->inc.as<uint32_t>());
  HIPCHK(hipGetLastError());
  sort_pairs_u64_u32(key.as<uint64_t>(), m->inc.as<uint32_t>(), T, bits_for((uint64_t)m->n_nodes), s);
  // nodes that appear in no element get empty ranges
  HIPCHK(hipMemsetAsync(m->inc_ptr.p, 0, (m->n_nodes + 1) * 4, s));
  k_inc_ptr<<<ceil_div(T, NT), NT, 0, s>>>(T, key.as<uint64_t>(), m->inc_ptr.as<int32_t>());
  HIPCHK(hipGetLastError());
  int32_t lastA = 0;
  uint64_t lk = 0;
  HIPCHK(hipMemcpyAsync(&lk, key.as<uint64_t>() + T - 1, 8, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  lastA = (int32_t)lk;
  if (lastA + 1 <= m->n_nodes)
    k_fill_tail<<<ceil_div(m->n_nodes - lastA, NT), NT, 0, s>>>(m->inc_ptr.as<int32_t>(), lastA + 1, m->n_nodes,
                                                                (int32_t)T);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(s));
  m->inc_ok = true;
}

}  // namespace yh

extern "C" {

int yafel_hip_mesh_create_cg(yafel_hip_ctx ctx, int dim, int p, int dpn, int64_t n_corner,
                             const double *corner_nodes, int64_t ncell, const int32_t *cells,
                             yafel_hip_mesh *out) {
  YH_TRY_BEGIN
  REQUIRE(ctx && out && corner_nodes && cells, "mesh_create_cg: null argument");
  REQUIRE(dim == 2 || dim == 3, "mesh_create_cg: dim must be 2 (Quad4) or 3 (Hex8)");
  REQUIRE(p >= 1 && p <= (dim == 2 ? 6 : 3), "mesh_create_cg: unsupported polynomial order");
  REQUIRE(dpn >= 1 && dpn <= 3, "mesh_create_cg: dof_per_node must be 1..3");
  REQUIRE(n_corner > 0 && ncell > 0, "mesh_create_cg: empty mesh");
  HIPCHK(hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  std::unique_ptr<yafel_hip_mesh_s> m(new yafel_hip_mesh_s());
  ctx_retain(ctx);
  m->ctx = ctx;
  m->dim = dim;
  m->p = p;
  m->dpn = dpn;
  m->nen = 1;
  for (int d = 0; d < dim; ++d) m->nen *= (p + 1);
  m->nqp = m->nen;
  m->nel = ncell;
  const int nc = dim == 2 ? 4 : 8;
  REQUIRE(ncell * (int64_t)m->nen < (int64_t)INT32_MAX, "mesh_create_cg: too many element nodes");
  std::vector<double> xl;
  element_tables(dim, p, m->hN, m->hdN, m->hw, xl);
  m->tabN.alloc(m->hN.size() * 8);
  m->tabdN.alloc(m->hdN.size() * 8);
  m->tabw.alloc(m->hw.size() * 8);
  HIPCHK(hipMemcpyAsync(m->tabN.p, m->hN.data(), m->hN.size() * 8, hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(m->tabdN.p, m->hdN.data(), m->hdN.size() * 8, hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(m->tabw.p, m->hw.data(), m->hw.size() * 8, hipMemcpyHostToDevice, s));
  DBuf dcells, dcorner, bad;
  dcells.alloc(ncell * nc * 4);
  dcorner.alloc(n_corner * 24);
  bad.alloc(4);
  HIPCHK(hipMemcpyAsync(dcells.p, cells, ncell * nc * 4, hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(dcorner.p, corner_nodes, n_corner * 24, hipMemcpyHostToDevice, s));
  HIPCHK(hipMemsetAsync(bad.p, 0, 4, s));
  k_check_cells<<<ceil_div(ncell * nc, NT), NT, 0, s>>>(ncell * nc, dcells.as<int32_t>(), n_corner, bad.as<int32_t>());
  HIPCHK(hipGetLastError());
  int32_t hbad = 0;
  HIPCHK(hipMemcpyAsync(&hbad, bad.p, 4, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  REQUIRE(hbad == 0, "mesh_create_cg: cell references a node outside corner_nodes");
  if (p == 1) {
    // make_cg_dofs p == 1 (DoFManager.cpp:153-160): the mesh passes through.
    m->n_nodes = n_corner;
    m->elements.swap(dcells);
    m->nodes.swap(dcorner);
  } else {
    // make_raw_dofs + recombine_all_duplicates (DoFManager.cpp:176-229, 340-414)
    const int64_t N = ncell * m->nen;
    DBuf dxl, raw, part, key, idx, flag, seg;
    dxl.alloc(xl.size() * 8);
    HIPCHK(hipMemcpyAsync(dxl.p, xl.data(), xl.size() * 8, hipMemcpyHostToDevice, s));
    raw.alloc(N * 24);
    k_raw_nodes<<<ceil_div(N, NT), NT, 0, s>>>(ncell, dim, m->nen, dxl.as<double>(), dcells.as<int32_t>(),
                                               dcorner.as<double>(), raw.as<double>());
    HIPCHK(hipGetLastError());
    const int gb = std::min<int64_t>(ceil_div(N, NT), 1024);
    part.alloc((size_t)gb * 6 * 8);
    k_bbox<<<gb, NT, 0, s>>>(N, raw.as<double>(), part.as<double>());
    HIPCHK(hipGetLastError());
    std::vector<double> hp((size_t)gb * 6);
    HIPCHK(hipMemcpyAsync(hp.data(), part.p, hp.size() * 8, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    double bmin[3] = {1e300, 1e300, 1e300}, bmax[3] = {-1e300, -1e300, -1e300};
    for (int b = 0; b < gb; ++b)
      for (int d = 0; d < 3; ++d) {
        bmin[d] = std::min(bmin[d], hp[6 * b + d]);
        bmax[d] = std::max(bmax[d], hp[6 * b + 3 + d]);
      }
    const double maxdim = std::max(bmax[0] - bmin[0], std::max(bmax[1] - bmin[1], bmax[2] - bmin[2]));
    REQUIRE(maxdim > 0, "mesh_create_cg: degenerate mesh");
    const double snap = maxdim * 1.0e-6;
    long long k0[3], k1[3];
    int bits[3];
    for (int d = 0; d < 3; ++d) {
      k0[d] = (long long)std::round(bmin[d] / snap);
      k1[d] = (long long)std::round(bmax[d] / snap);
      bits[d] = bits_for((uint64_t)(k1[d] - k0[d]));
    }
    REQUIRE(bits[0] + bits[1] + bits[2] <= 64, "mesh_create_cg: bounding box too elongated for the snap key");
    key.alloc(N * 8);
    idx.alloc(N * 4);
    k_snap_keys<<<ceil_div(N, NT), NT, 0, s>>>(N, raw.as<double>(), snap, k0[0], k0[1], k0[2], bits[0], bits[1],
                                               key.as<uint64_t>(), idx.as<uint32_t>());
    HIPCHK(hipGetLastError());
    sort_pairs_u64_u32(key.as<uint64_t>(), idx.as<uint32_t>(), N, bits[0] + bits[1] + bits[2], s);
    flag.alloc(N * 4);
    seg.alloc(N * 4);
    k_uniq_flags<<<ceil_div(N, NT), NT, 0, s>>>(N, idx.as<uint32_t>(), raw.as<double>(), flag.as<int32_t>());
    HIPCHK(hipGetLastError());
    inclusive_scan_i32(flag.as<int32_t>(), seg.as<int32_t>(), N, s);
    int32_t nu = 0;
    HIPCHK(hipMemcpyAsync(&nu, seg.as<int32_t>() + N - 1, 4, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    m->n_nodes = nu;
    m->elements.alloc(N * 4);
    m->nodes.alloc((int64_t)nu * 24);
    k_assign_nodes<<<ceil_div(N, NT), NT, 0, s>>>(N, idx.as<uint32_t>(), flag.as<int32_t>(), seg.as<int32_t>(),
                                                  raw.as<double>(), m->elements.as<int32_t>(), m->nodes.as<double>());
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(s));
  }
  REQUIRE(m->n_nodes * (int64_t)dpn < (int64_t)INT32_MAX, "mesh_create_cg: too many dofs");
  *out = m.release();
  return YAFEL_HIP_OK;
  YH_TRY_END(ctx)
}

int yafel_hip_mesh_destroy(yafel_hip_mesh m) {
  if (!m) return YAFEL_HIP_OK;
  (void)hipStreamSynchronize(m->ctx->stream);
  delete m;
  return YAFEL_HIP_OK;
}

int yafel_hip_mesh_info(yafel_hip_mesh m, int64_t *info) {
  YH_TRY_BEGIN
  REQUIRE(m && info, "mesh_info: null argument");
  info[0] = m->n_nodes;
  info[1] = m->nel;
  info[2] = m->nen;
  info[3] = m->n_nodes * m->dpn;
  info[4] = m->dim;
  info[5] = m->p;
  info[6] = m->dpn;
  return YAFEL_HIP_OK;
  YH_TRY_END(nullptr)
}

int yafel_hip_mesh_download(yafel_hip_mesh m, int32_t *elements, double *dof_nodes) {
  YH_TRY_BEGIN
  REQUIRE(m, "mesh_download: null mesh");
  hipStream_t s = m->ctx->stream;
  if (elements) HIPCHK(hipMemcpyAsync(elements, m->elements.p, m->nel * m->nen * 4, hipMemcpyDeviceToHost, s));
  if (dof_nodes) HIPCHK(hipMemcpyAsync(dof_nodes, m->nodes.p, m->n_nodes * 24, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  return YAFEL_HIP_OK;
  YH_TRY_END(m ? m->ctx : nullptr)
}

int yafel_hip_mesh_nodes_dev(yafel_hip_mesh m, const double **nodes_dev) {
  YH_TRY_BEGIN
  REQUIRE(m && nodes_dev, "mesh_nodes_dev: null argument");
  *nodes_dev = m->nodes.as<double>();
  return YAFEL_HIP_OK;
  YH_TRY_END(nullptr)
}

int yafel_hip_mesh_pattern(yafel_hip_mesh m, int format, int precisions, yafel_hip_matrix *out) {
  YH_TRY_BEGIN
  REQUIRE(m && out, "mesh_pattern: null argument");
  REQUIRE(format == YAFEL_HIP_CSR || format == YAFEL_HIP_BCSR3, "mesh_pattern: unknown format");
  REQUIRE(format == YAFEL_HIP_CSR || m->dpn == 3, "mesh_pattern: BCSR3 needs 3 dofs per node");
  yafel_hip_ctx_s *ctx = m->ctx;
  HIPCHK(hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  const int nen = m->nen, dpn = m->dpn;
  const int64_t T = m->nel * (int64_t)nen * nen;
  REQUIRE(T < (int64_t)INT32_MAX, "mesh_pattern: too many element pairs");
  const uint64_t nn = (uint64_t)m->n_nodes;
  std::unique_ptr<YhPlan> &slot = format == YAFEL_HIP_CSR ? m->plan_csr : m->plan_bcsr;
  if (!slot) {
    auto plan = std::make_unique<YhPlan>();
    plan->format = format;
    DBuf key, flag, seg;
    key.alloc(T * 8);
    plan->contrib.alloc(T * 4);
    k_pair_keys<<<ceil_div(T, NT), NT, 0, s>>>(m->nel, nen, m->elements.as<int32_t>(), nn, key.as<uint64_t>(),
                                               plan->contrib.as<uint32_t>());
    HIPCHK(hipGetLastError());
    sort_pairs_u64_u32(key.as<uint64_t>(), plan->contrib.as<uint32_t>(), T, bits_for(nn * nn - 1), s);
    flag.alloc(T * 4);
    seg.alloc(T * 4);
    k_flags<<<ceil_div(T, NT), NT, 0, s>>>(T, key.as<uint64_t>(), flag.as<int32_t>());
    HIPCHK(hipGetLastError());
    inclusive_scan_i32(flag.as<int32_t>(), seg.as<int32_t>(), T, s);
    int32_t nruns = 0;
    HIPCHK(hipMemcpyAsync(&nruns, seg.as<int32_t>() + T - 1, 4, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    plan->nruns = nruns;
    plan->cptr.alloc(((int64_t)nruns + 1) * 4);
    k_run_heads<<<ceil_div(T, NT), NT, 0, s>>>(T, flag.as<int32_t>(), seg.as<int32_t>(), plan->cptr.as<int32_t>());
    HIPCHK(hipGetLastError());
    const int32_t Ti = (int32_t)T;
    HIPCHK(hipMemcpyAsync(plan->cptr.as<int32_t>() + nruns, &Ti, 4, hipMemcpyHostToDevice, s));
    plan->run_row.alloc((int64_t)nruns * 4);
    plan->run_col.alloc((int64_t)nruns * 4);
    plan->node_ptr.alloc((m->n_nodes + 1) * 4);
    HIPCHK(hipMemsetAsync(plan->node_ptr.p, 0, (m->n_nodes + 1) * 4, s));
    k_runs<<<ceil_div(nruns, NT), NT, 0, s>>>(nruns, plan->cptr.as<int32_t>(), key.as<uint64_t>(), nn,
                                              plan->run_row.as<int32_t>(), plan->run_col.as<int32_t>(),
                                              plan->node_ptr.as<int32_t>());
    HIPCHK(hipGetLastError());
    int32_t lastA = 0;
    HIPCHK(hipMemcpyAsync(&lastA, plan->run_row.as<int32_t>() + nruns - 1, 4, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    k_fill_tail<<<ceil_div(m->n_nodes - lastA, NT), NT, 0, s>>>(plan->node_ptr.as<int32_t>(), lastA + 1, m->n_nodes,
                                                                nruns);
    HIPCHK(hipGetLastError());
    plan->nrow_nodes = (int64_t)lastA + 1;
    slot = std::move(plan);
  }
  YhPlan *plan = slot.get();
  const int32_t nruns = (int32_t)plan->nruns;
  // The reference's row count is max_row + 1 over the triplets
  // (sparse_csr.hpp:107); trailing nodes that no element touches are dropped.
  const int64_t nrow_nodes = plan->nrow_nodes;
  DBuf ptr, col, val;
  yafel_hip_matrix_s *A = nullptr;
  if (format == YAFEL_HIP_BCSR3) {
    ptr.alloc((nrow_nodes + 1) * 4);
    HIPCHK(hipMemcpyAsync(ptr.p, plan->node_ptr.p, (nrow_nodes + 1) * 4, hipMemcpyDeviceToDevice, s));
    col.alloc((int64_t)nruns * 4);
    HIPCHK(hipMemcpyAsync(col.p, plan->run_col.p, (int64_t)nruns * 4, hipMemcpyDeviceToDevice, s));
    val.alloc((int64_t)nruns * 72);
    HIPCHK(hipMemsetAsync(val.p, 0, (int64_t)nruns * 72, s));
    A = matrix_from_device(ctx, format, nrow_nodes, nrow_nodes, nruns, nrow_nodes + 1, ptr, col, val, precisions);
  } else if (dpn == 1) {
    ptr.alloc((nrow_nodes + 1) * 4);
    HIPCHK(hipMemcpyAsync(ptr.p, plan->node_ptr.p, (nrow_nodes + 1) * 4, hipMemcpyDeviceToDevice, s));
    col.alloc((int64_t)nruns * 4);
    HIPCHK(hipMemcpyAsync(col.p, plan->run_col.p, (int64_t)nruns * 4, hipMemcpyDeviceToDevice, s));
    val.alloc((int64_t)nruns * 8);
    HIPCHK(hipMemsetAsync(val.p, 0, (int64_t)nruns * 8, s));
    A = matrix_from_device(ctx, format, nrow_nodes, nrow_nodes, nruns, nrow_nodes + 1, ptr, col, val, precisions);
  } else {
    const int64_t nnz = (int64_t)nruns * dpn * dpn;
    REQUIRE(nnz < (int64_t)INT32_MAX, "mesh_pattern: scalar CSR exceeds int32 indices");
    const int64_t nrows = nrow_nodes * dpn;
    ptr.alloc((nrows + 1) * 4);
    col.alloc(nnz * 4);
    val.alloc(nnz * 8);
    plan->scal_pos.alloc((int64_t)nruns * 4);
    HIPCHK(hipMemsetAsync(val.p, 0, nnz * 8, s));
    k_scalar_ptr<<<ceil_div(nrow_nodes, NT), NT, 0, s>>>(nrow_nodes, dpn, plan->node_ptr.as<int32_t>(), ptr.as<int32_t>());
    k_scalar_csr<<<ceil_div(nruns, NT), NT, 0, s>>>(nruns, dpn, plan->run_row.as<int32_t>(), plan->run_col.as<int32_t>(),
                                                    plan->node_ptr.as<int32_t>(), col.as<int32_t>(),
                                                    plan->scal_pos.as<int32_t>());
    HIPCHK(hipGetLastError());
    A = matrix_from_device(ctx, format, nrows, nrows, nnz, nrows + 1, ptr, col, val, precisions);
  }
  HIPCHK(hipStreamSynchronize(s));
  A->mesh = m;
  A->plan = plan;
  *out = A;
  return YAFEL_HIP_OK;
  YH_TRY_END(m ? m->ctx : nullptr)
}

int yafel_hip_bc_mask_plane(yafel_hip_mesh m, int axis, double value, double tol, int comp,
                            uint8_t *mask_dev) {
  YH_TRY_BEGIN
  REQUIRE(m && mask_dev, "bc_mask_plane: null argument");
  REQUIRE(axis >= 0 && axis < 3, "bc_mask_plane: axis must be 0..2");
  REQUIRE(comp < m->dpn, "bc_mask_plane: component out of range");
  if (m->n_nodes)
    k_bc_plane<<<ceil_div(m->n_nodes, NT), NT, 0, m->ctx->stream>>>(m->n_nodes, m->dpn, m->nodes.as<double>(), axis,
                                                                    value, tol, comp, mask_dev);
  HIPCHK(hipGetLastError());
  return YAFEL_HIP_OK;
  YH_TRY_END(m ? m->ctx : nullptr)
}

}  // extern "C"
