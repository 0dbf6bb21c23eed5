// Device-resident Krylov solvers: CG / Jacobi-PCG / BiCGSTAB (+ right Jacobi),
// in FP64 or as FP32 inner iterations inside FP64 refinement.
//
// Every scalar recurrence (alpha, beta, omega, rho, the stopping test) runs
// on the device in FP64 inside the last workgroup of the kernel that produced
// the reduction (deterministic: partials are summed in block order), so a
// solve needs no host round trip per iteration.  The host launches chunks of
// iterations and polls a status word once per chunk; kernels after
// convergence return immediately.
//
// One CG iteration is two kernels:
//   A: p_new = z + beta p_old (reconstructed in the SpMV gather, written by the
//      row owner), q = A p_new, partial p_new.q  -> last block: alpha
//   B: x += alpha p_new, r -= alpha q, z = D^-1 r, partials r.z, r.r
//      -> last block: beta, stopping test
// MIXED: vectors and matrix are FP32 (dots accumulate in FP64); when the
// recurrence residual has dropped by refine_delta (or claims convergence) the
// inner loop pauses, x_hi += x_lo and r = b - A x_hi are formed in FP64 (a
// residual replacement that keeps the search direction), and the FP64 true
// residual decides convergence.
#include <chrono>
#include <cmath>

#include "internal.hpp"
#include "kernels.hpp"

using namespace yh;

namespace {

constexpr int NT = 256;

struct KState {
  double rho;      // CG: r.z (r.r unpreconditioned); BiCGSTAB: rhat.r
  double rho0;     // initial rho (reference stopping ratio)
  double rho_prev; // BiCGSTAB: rho of the previous iteration
  double rr;       // recurrence r.r
  double rr_ref;   // r.r right after the last refinement
  double bb;       // b.b
  double alpha, beta, omega;
  double crit;     // last stopping quantity
  double true_rr;  // last FP64 true residual squared
  double tol;
  double delta2;
  long long iter, max_iter, refinements;
  int done, status, paused, first;
  int method, stop, mixed, special;
  int pidx, pad0;
  unsigned int counter[8];
};

enum { M_CG = 0, M_PCG = 1, M_BICG = 2, M_BICGJ = 3 };

__device__ __forceinline__ bool bad(double v) { return !(fabs(v) <= 1.7976931348623157e308); }

// ------------------------------------------------------------------ drivers
template <class T> struct MatView {
  const int *ptr;
  const int *col;
  const T *val;
  int64_t brows;
  int64_t nb;
};

// SpMV with a fused per-row epilogue and up to 4 block-reduced sums; the last
// block to finish folds the partials (fixed order) and runs op.last().
template <int FMT, int L, class T, class Op>
__global__ __launch_bounds__(NT) void k_spmv_fused(MatView<T> A, Op op0) {
  Op op = op0;
  if (op.skip()) return;
  op.begin();
  const int lane = threadIdx.x & (L - 1);
  constexpr int NP = Op::NP;
  double acc[NP > 0 ? NP : 1];
#pragma unroll
  for (int j = 0; j < (NP > 0 ? NP : 1); ++j) acc[j] = 0;
  const int64_t rpb = NT / L;
  const int64_t chunk = (A.brows + gridDim.x - 1) / gridDim.x;
  const int64_t r0 = (int64_t)dev::xcd_block() * chunk;
  const int64_t r1 = r0 + chunk < A.brows ? r0 + chunk : A.brows;
  for (int64_t r = r0 + threadIdx.x / L; r < r1; r += rpb) {
    if (FMT == YAFEL_HIP_CSR) {
      T y = dev::csr_row<L, T>(A.ptr, A.col, A.val, r, lane, [&](int c) { return op.x(c); });
      if (lane == 0) op.row(r, y, acc);
    } else {
      T a[3];
      dev::bcsr_row<L, T>(A.ptr, A.col, A.val, A.nb, r, lane, a, [&](int64_t c) { return op.x(c); });
      if (lane == 0) {
        op.row(3 * r, a[0], acc);
        op.row(3 * r + 1, a[1], acc);
        op.row(3 * r + 2, a[2], acc);
      }
    }
  }
  if constexpr (NP > 0) {
    __shared__ double lds[NT / 64];
    __shared__ double sums[NP];
    const int blk = dev::xcd_block();
#pragma unroll
    for (int j = 0; j < NP; ++j) {
      double s = dev::block_sum<NT>(acc[j], lds);
      if (threadIdx.x == 0) op.part[(size_t)j * gridDim.x + blk] = s;
    }
    if (dev::last_block_done(op.counter)) {
#pragma unroll
      for (int j = 0; j < NP; ++j) {
        double s = dev::reduce_partials<NT>(op.part + (size_t)j * gridDim.x, gridDim.x, 1, lds);
        if (threadIdx.x == 0) sums[j] = s;
      }
      __syncthreads();
      if (threadIdx.x == 0) {
        *op.counter = 0;
        op.last(sums);
      }
    }
  }
}

// Elementwise pass with up to 4 block-reduced sums.
template <class Op>
__global__ __launch_bounds__(NT) void k_elem(int64_t n, Op op0) {
  Op op = op0;
  if (op.skip()) return;
  op.begin();
  constexpr int NP = Op::NP;
  double acc[NP > 0 ? NP : 1];
#pragma unroll
  for (int j = 0; j < (NP > 0 ? NP : 1); ++j) acc[j] = 0;
  const int64_t chunk = ((n + gridDim.x - 1) / gridDim.x + 63) & ~(int64_t)63;
  const int64_t i0 = (int64_t)dev::xcd_block() * chunk;
  const int64_t i1 = i0 + chunk < n ? i0 + chunk : n;
  for (int64_t i = i0 + threadIdx.x; i < i1; i += NT) op.elem(i, acc);
  if constexpr (NP > 0) {
    __shared__ double lds[NT / 64];
    __shared__ double sums[NP];
    const int blk = dev::xcd_block();
#pragma unroll
    for (int j = 0; j < NP; ++j) {
      double s = dev::block_sum<NT>(acc[j], lds);
      if (threadIdx.x == 0) op.part[(size_t)j * gridDim.x + blk] = s;
    }
    if (dev::last_block_done(op.counter)) {
#pragma unroll
      for (int j = 0; j < NP; ++j) {
        double s = dev::reduce_partials<NT>(op.part + (size_t)j * gridDim.x, gridDim.x, 1, lds);
        if (threadIdx.x == 0) sums[j] = s;
      }
      __syncthreads();
      if (threadIdx.x == 0) {
        *op.counter = 0;
        op.last(sums);
      }
    }
  }
}

// --------------------------------------------------------------------- ops
struct OpBase {
  KState *st;
  double *part;
  unsigned int *counter;
  __device__ bool running() const { return !st->done && !st->paused; }
};

// Stopping test after an iteration (FP64 path) or pause request (MIXED).
__device__ void after_iteration(KState *st, double rho_new, double rr) {
  st->iter += 1;
  st->rr = rr;
  const bool cgfam = st->method == M_CG || st->method == M_PCG;
  double crit;
  if (st->stop == YAFEL_HIP_STOP_RELRES) crit = st->bb > 0 ? sqrt(rr / st->bb) : 0.0;
  else crit = cgfam ? rho_new / st->rho0 : rr / st->bb;
  st->crit = crit;
  if (bad(rho_new) || bad(rr)) {
    st->status = YAFEL_HIP_STATUS_NAN;
    st->done = 1;
    return;
  }
  const bool conv = crit < st->tol;
  if (st->mixed || !cgfam) {
    // verify / refine against the FP64 true residual
    if (conv || rr < st->delta2 * st->rr_ref || st->iter >= st->max_iter) st->paused = 1;
  } else {
    if (conv) {
      st->status = YAFEL_HIP_STATUS_CONVERGED;
      st->done = 1;
      return;
    }
    if (st->iter >= st->max_iter) {
      st->status = YAFEL_HIP_STATUS_MAXITER;
      st->done = 1;
      return;
    }
  }
  if (!st->paused) {
    if (cgfam) {
      st->beta = rho_new / st->rho;
      st->rho = rho_new;
    } else {
      st->beta = (rho_new / st->rho) * (st->alpha / st->omega);
      st->rho_prev = st->rho;
      st->rho = rho_new;
    }
  } else if (!cgfam) {
    st->rho_prev = st->rho;
  }
}

// CG kernel A: q = A p_new with p_new = z + beta p_old.  The two p buffers
// alternate; st->pidx names the one holding the current p, so iterations
// skipped after a pause or convergence never desynchronise them.
template <class T>
struct OpCgA : OpBase {
  static constexpr int NP = 1;
  const T *z;
  T *pbuf0, *pbuf1;
  T *q;
  const T *pold;
  T *pnew;
  int cur;
  T beta;
  __device__ bool skip() const { return !running(); }
  __device__ void begin() {
    beta = st->first ? T(0) : (T)st->beta;
    cur = st->pidx;
    pold = cur ? pbuf1 : pbuf0;
    pnew = cur ? pbuf0 : pbuf1;
  }
  __device__ T x(int64_t c) const { return fma(beta, pold[c], z[c]); }
  __device__ void row(int64_t i, T y, double *acc) const {
    const T pn = fma(beta, pold[i], z[i]);
    pnew[i] = pn;
    q[i] = y;
    acc[0] += (double)pn * (double)y;
  }
  __device__ void last(const double *s) const {
    const double a = st->rho / s[0];
    st->first = 0;
    st->alpha = a;
    st->pidx = cur ^ 1;
    if (s[0] == 0.0 || bad(a)) {
      st->status = YAFEL_HIP_STATUS_BREAKDOWN;
      st->done = 1;
    }
  }
};

// CG kernel B: x += alpha p, r -= alpha q, z = D^-1 r.
template <class T, bool PREC>
struct OpCgB : OpBase {
  static constexpr int NP = 2;
  T *x, *r, *z;
  const T *pbuf0, *pbuf1, *q, *dinv;
  const T *p;
  T alpha;
  __device__ bool skip() const { return !running(); }
  __device__ void begin() {
    alpha = (T)st->alpha;
    p = st->pidx ? pbuf1 : pbuf0;
  }
  __device__ void elem(int64_t i, double *acc) const {
    x[i] = fma(alpha, p[i], x[i]);
    const T ri = fma(-alpha, q[i], r[i]);
    r[i] = ri;
    if (PREC) {
      const T zi = dinv[i] * ri;
      z[i] = zi;
      acc[0] += (double)ri * (double)zi;
    } else {
      acc[0] += (double)ri * (double)ri;
    }
    acc[1] += (double)ri * (double)ri;
  }
  __device__ void last(const double *s) const { after_iteration(st, s[0], s[1]); }
};

// FP64 residual r = b - A x_hi; stores r (in the inner precision TL) and z,
// and folds r.w (w = z, r or rhat), r.r, the FP64 true r.r and b.b.
// MODE 0 = initialisation, 1 = refinement / verification.
template <class TL, int MODE>
struct OpResidual : OpBase {
  static constexpr int NP = 4;
  const double *xh, *b;
  TL *r, *z;
  const TL *dinv, *rhat;
  int wkind;  // 0: r.r, 1: r.(D^-1 r) stored in z, 2: r.rhat
  __device__ bool skip() const {
    if (MODE == 0) return false;
    return st->done || !st->paused;
  }
  __device__ void begin() {}
  __device__ double x(int64_t c) const { return xh[c]; }
  __device__ void row(int64_t i, double y, double *acc) const {
    const double rh = b[i] - y;
    const TL rl = (TL)rh;
    r[i] = rl;
    double w;
    if (wkind == 1) {
      const TL zi = dinv[i] * rl;
      z[i] = zi;
      w = (double)zi;
    } else if (wkind == 2) {
      w = MODE == 0 ? (double)rl : (double)rhat[i];  // rhat = r0 at the start
    } else {
      w = (double)rl;
    }
    acc[0] += (double)rl * w;
    acc[1] += (double)rl * (double)rl;
    acc[2] += rh * rh;
    acc[3] += b[i] * b[i];
  }
  __device__ void last(const double *s) const {
    const bool cgfam = st->method == M_CG || st->method == M_PCG;
    st->true_rr = s[2];
    if (MODE == 0) {
      st->bb = s[3];
      st->rho = s[0];
      st->rho0 = s[0];
      st->rho_prev = 1.0;
      st->rr = s[1];
      st->rr_ref = s[1];
      st->first = 1;
      st->iter = 0;
      st->alpha = 1.0;
      st->omega = 1.0;
      st->beta = 0.0;
      st->paused = 0;
      st->pidx = 0;
      // early exits of the reference
      if (cgfam && st->stop == YAFEL_HIP_STOP_REFERENCE && s[0] == 0.0) {
        st->done = 1;  // cg_solve.hpp:52-54 / pcg_solve.hpp:148-150: x0 solves it
        st->crit = 0;
      } else if (!cgfam && st->stop == YAFEL_HIP_STOP_REFERENCE && s[3] < st->tol) {
        st->done = 1;  // bicgstab_solve.hpp:43-45 returns rhs
        st->special = 1;
        st->crit = 0;
      } else if (s[3] == 0.0) {
        st->done = 1;  // b = 0: x = 0 is exact
        st->special = 2;
        st->crit = 0;
      } else if (st->stop == YAFEL_HIP_STOP_RELRES && sqrt(s[2] / s[3]) < st->tol) {
        st->done = 1;
        st->crit = sqrt(s[2] / s[3]);
      }
      return;
    }
    // refinement / verification against the FP64 residual
    st->refinements += 1;
    double crit;
    if (st->stop == YAFEL_HIP_STOP_RELRES) crit = sqrt(s[2] / st->bb);
    else if (cgfam) crit = (st->method == M_PCG ? s[0] : s[2]) / st->rho0;
    else crit = s[2] / st->bb;
    st->crit = crit;
    if (bad(s[2])) {
      st->status = YAFEL_HIP_STATUS_NAN;
      st->done = 1;
      return;
    }
    if (crit < st->tol) {
      st->status = YAFEL_HIP_STATUS_CONVERGED;
      st->done = 1;
      return;
    }
    if (st->iter >= st->max_iter) {
      st->status = YAFEL_HIP_STATUS_MAXITER;
      st->done = 1;
      return;
    }
    if (cgfam) {
      st->beta = s[0] / st->rho;
      st->rho = s[0];
    } else {
      // rho_prev holds rho of the finished iteration's predecessor
      st->beta = (s[0] / st->rho_prev) * (st->alpha / st->omega);
      st->rho = s[0];
    }
    st->rr = s[1];
    st->rr_ref = s[1];
    st->paused = 0;
  }
};

// MIXED: x_hi += x_lo, x_lo = 0 (runs only while paused).
template <class TL>
struct OpFold : OpBase {
  static constexpr int NP = 0;
  double *xh;
  TL *xl;
  __device__ bool skip() const { return st->done || !st->paused; }
  __device__ void begin() {}
  __device__ void elem(int64_t i, double *) const {
    xh[i] += (double)xl[i];
    xl[i] = TL(0);
  }
  __device__ void last(const double *) const {}
};

// BiCGSTAB (bicgstab_solve.hpp:55-91): p = r + beta (p - omega v), ph = D^-1 p.
template <class T, bool PREC>
struct OpBiP : OpBase {
  static constexpr int NP = 0;
  const T *r, *v, *dinv;
  T *p, *ph;
  T beta, omega;
  __device__ bool skip() const { return !running(); }
  __device__ void begin() {
    beta = st->first ? T(0) : (T)st->beta;
    omega = (T)st->omega;
  }
  __device__ void elem(int64_t i, double *) const {
    const T pi = r[i] + beta * (p[i] - omega * v[i]);
    p[i] = pi;
    if (PREC) ph[i] = dinv[i] * pi;
  }
  __device__ void last(const double *) const {}
};

// v = A ph, rhat.v -> alpha
template <class T>
struct OpBiV : OpBase {
  static constexpr int NP = 1;
  const T *ph, *rhat;
  T *v;
  __device__ bool skip() const { return !running(); }
  __device__ void begin() {}
  __device__ T x(int64_t c) const { return ph[c]; }
  __device__ void row(int64_t i, T y, double *acc) const {
    v[i] = y;
    acc[0] += (double)rhat[i] * (double)y;
  }
  __device__ void last(const double *s) const {
    const double a = st->rho / s[0];
    st->first = 0;
    st->alpha = a;
    if (s[0] == 0.0 || bad(a)) {
      st->status = YAFEL_HIP_STATUS_BREAKDOWN;
      st->done = 1;
    }
  }
};

// s = r - alpha v, sh = D^-1 s
template <class T, bool PREC>
struct OpBiS : OpBase {
  static constexpr int NP = 0;
  const T *r, *v, *dinv;
  T *s, *sh;
  T alpha;
  __device__ bool skip() const { return !running(); }
  __device__ void begin() { alpha = (T)st->alpha; }
  __device__ void elem(int64_t i, double *) const {
    const T si = r[i] - v[i] * alpha;
    s[i] = si;
    if (PREC) sh[i] = dinv[i] * si;
  }
  __device__ void last(const double *) const {}
};

// t = A sh, t.t and t.s -> omega
template <class T>
struct OpBiT : OpBase {
  static constexpr int NP = 2;
  const T *sh, *s;
  T *t;
  __device__ bool skip() const { return !running(); }
  __device__ void begin() {}
  __device__ T x(int64_t c) const { return sh[c]; }
  __device__ void row(int64_t i, T y, double *acc) const {
    t[i] = y;
    acc[0] += (double)y * (double)y;
    acc[1] += (double)y * (double)s[i];
  }
  __device__ void last(const double *sm) const {
    const double w = sm[0] == 0.0 ? 0.0 : sm[1] / sm[0];
    st->omega = w;
    if (bad(w)) {
      st->status = YAFEL_HIP_STATUS_NAN;
      st->done = 1;
    }
  }
};

// x += alpha ph + omega sh, r = s - omega t, rhat.r, r.r
template <class T>
struct OpBiX : OpBase {
  static constexpr int NP = 2;
  T *x, *r;
  const T *ph, *sh, *s, *t, *rhat;
  T alpha, omega;
  __device__ bool skip() const { return !running(); }
  __device__ void begin() {
    alpha = (T)st->alpha;
    omega = (T)st->omega;
  }
  __device__ void elem(int64_t i, double *acc) const {
    x[i] += ph[i] * alpha + sh[i] * omega;
    const T ri = s[i] - t[i] * omega;
    r[i] = ri;
    acc[0] += (double)rhat[i] * (double)ri;
    acc[1] += (double)ri * (double)ri;
  }
  __device__ void last(const double *s) const {
    if (st->omega == 0.0) {
      // t = 0: the reference keeps going with w = 0, which makes beta infinite
      st->status = YAFEL_HIP_STATUS_BREAKDOWN;
      st->done = 1;
      st->iter += 1;
      return;
    }
    after_iteration(st, s[0], s[1]);
  }
};

// final report: r = b - A x (FP64), r.r and b.b into out[0..1]
struct OpReport : OpBase {
  static constexpr int NP = 2;
  const double *xh, *b;
  double *out;
  __device__ bool skip() const { return false; }
  __device__ void begin() {}
  __device__ double x(int64_t c) const { return xh[c]; }
  __device__ void row(int64_t i, double y, double *acc) const {
    const double r = b[i] - y;
    acc[0] += r * r;
    acc[1] += b[i] * b[i];
  }
  __device__ void last(const double *s) const {
    out[0] = s[0];
    out[1] = s[1];
  }
};

// ---------------------------------------------------------------- launchers
template <class T>
MatView<T> view(yafel_hip_matrix_s *A) {
  return MatView<T>{A->ptr.as<int>(), A->col.as<int>(),
                    (const T *)(sizeof(T) == 8 ? A->val64.p : A->val32.p), A->brows, A->nblk};
}

template <class T, class Op>
void launch_spmv_op(yafel_hip_matrix_s *A, const Op &op, hipStream_t s) {
  const int L = A->lanes;
  const int grid = grid_for(A->ctx, A->brows * (int64_t)L, NT);
  const MatView<T> v = view<T>(A);
  if (A->format == YAFEL_HIP_CSR) {
    switch (L) {
      case 2: k_spmv_fused<YAFEL_HIP_CSR, 2, T, Op><<<grid, NT, 0, s>>>(v, op); break;
      case 4: k_spmv_fused<YAFEL_HIP_CSR, 4, T, Op><<<grid, NT, 0, s>>>(v, op); break;
      case 8: k_spmv_fused<YAFEL_HIP_CSR, 8, T, Op><<<grid, NT, 0, s>>>(v, op); break;
      case 16: k_spmv_fused<YAFEL_HIP_CSR, 16, T, Op><<<grid, NT, 0, s>>>(v, op); break;
      case 32: k_spmv_fused<YAFEL_HIP_CSR, 32, T, Op><<<grid, NT, 0, s>>>(v, op); break;
      default: k_spmv_fused<YAFEL_HIP_CSR, 64, T, Op><<<grid, NT, 0, s>>>(v, op); break;
    }
  } else {
    switch (L) {
      case 8: k_spmv_fused<YAFEL_HIP_BCSR3, 8, T, Op><<<grid, NT, 0, s>>>(v, op); break;
      case 16: k_spmv_fused<YAFEL_HIP_BCSR3, 16, T, Op><<<grid, NT, 0, s>>>(v, op); break;
      case 32: k_spmv_fused<YAFEL_HIP_BCSR3, 32, T, Op><<<grid, NT, 0, s>>>(v, op); break;
      default: k_spmv_fused<YAFEL_HIP_BCSR3, 64, T, Op><<<grid, NT, 0, s>>>(v, op); break;
    }
  }
  HIPCHK(hipGetLastError());
}

template <class Op>
void launch_elem(yafel_hip_ctx_s *ctx, int64_t n, const Op &op, hipStream_t s) {
  const int grid = grid_for(ctx, n, NT);
  k_elem<Op><<<grid, NT, 0, s>>>(n, op);
  HIPCHK(hipGetLastError());
}

// ------------------------------------------------------------- solve driver
struct Vecs {
  // inner-precision vectors
  void *r, *z, *p0, *p1, *q, *xl, *rhat, *v, *s, *sh, *t, *ph;
};

void ensure_ws(yafel_hip_ctx_s *ctx, int64_t n) {
  YhWorkspace &w = ctx->ws;
  const size_t need = (size_t)std::max<int64_t>(n, 1);
  if (w.n64 < need) {
    for (auto &b : w.vec64) b.alloc(need * 8);
    w.n64 = need;
  }
  if (!w.partials.p) w.partials.alloc((size_t)ctx->num_cus * 8 * 8 * sizeof(double) + 4096);
  if (!w.state.p) {
    w.state.alloc(4096);
    HIPCHK(hipMemsetAsync(w.state.p, 0, 4096, ctx->stream));
  }
}

struct Solver {
  yafel_hip_ctx_s *ctx;
  yafel_hip_matrix_s *A;
  yafel_hip_solve_opts o;
  int64_t n;
  bool prec, mixed, cgfam;
  KState *st;
  double *part;
  unsigned int *cnt;
  const double *b;
  double *x;
  hipStream_t s;

  template <class O>
  O base(O op, int counter) const {
    op.st = st;
    op.part = part;
    op.counter = cnt + counter;
    return op;
  }

  template <class TL>
  void residual(int mode, const Vecs &v) {
    OpResidual<TL, 0> o0;
    OpResidual<TL, 1> o1;
    auto fill = [&](auto &op) {
      op = base(op, 0);
      op.xh = x;
      op.b = b;
      op.r = (TL *)v.r;
      op.z = (TL *)v.z;
      op.dinv = (const TL *)(sizeof(TL) == 8 ? A->dinv64.p : A->dinv32.p);
      op.rhat = (const TL *)v.rhat;
      op.wkind = cgfam ? (prec ? 1 : 0) : 2;
    };
    if (mode == 0) {
      fill(o0);
      launch_spmv_op<double>(A, o0, s);
    } else {
      fill(o1);
      launch_spmv_op<double>(A, o1, s);
    }
  }

  template <class T>
  void cg_iteration(const Vecs &v) {
    OpCgA<T> a{};
    a = base(a, 1);
    a.z = (const T *)(prec ? v.z : v.r);
    a.pbuf0 = (T *)v.p0;
    a.pbuf1 = (T *)v.p1;
    a.q = (T *)v.q;
    launch_spmv_op<T>(A, a, s);
    const T *dinv = (const T *)(sizeof(T) == 8 ? A->dinv64.p : A->dinv32.p);
    T *xx = mixed ? (T *)v.xl : (T *)x;
    if (prec) {
      OpCgB<T, true> bb{};
      bb = base(bb, 2);
      bb.x = xx; bb.r = (T *)v.r; bb.z = (T *)v.z; bb.pbuf0 = (const T *)v.p0; bb.pbuf1 = (const T *)v.p1;
      bb.q = (const T *)v.q; bb.dinv = dinv;
      launch_elem(ctx, n, bb, s);
    } else {
      OpCgB<T, false> bb{};
      bb = base(bb, 2);
      bb.x = xx; bb.r = (T *)v.r; bb.z = (T *)v.z; bb.pbuf0 = (const T *)v.p0; bb.pbuf1 = (const T *)v.p1;
      bb.q = (const T *)v.q; bb.dinv = dinv;
      launch_elem(ctx, n, bb, s);
    }
  }

  template <class T>
  void bicg_iteration(const Vecs &v) {
    const T *dinv = (const T *)(sizeof(T) == 8 ? A->dinv64.p : A->dinv32.p);
    if (prec) {
      OpBiP<T, true> p{};
      p = base(p, 1);
      p.r = (const T *)v.r; p.v = (const T *)v.v; p.dinv = dinv; p.p = (T *)v.p0; p.ph = (T *)v.ph;
      launch_elem(ctx, n, p, s);
    } else {
      OpBiP<T, false> p{};
      p = base(p, 1);
      p.r = (const T *)v.r; p.v = (const T *)v.v; p.dinv = dinv; p.p = (T *)v.p0; p.ph = (T *)v.p0;
      launch_elem(ctx, n, p, s);
    }
    const T *ph = (const T *)(prec ? v.ph : v.p0);
    OpBiV<T> vv{};
    vv = base(vv, 2);
    vv.ph = ph; vv.rhat = (const T *)v.rhat; vv.v = (T *)v.v;
    launch_spmv_op<T>(A, vv, s);
    if (prec) {
      OpBiS<T, true> so{};
      so = base(so, 3);
      so.r = (const T *)v.r; so.v = (const T *)v.v; so.dinv = dinv; so.s = (T *)v.s; so.sh = (T *)v.sh;
      launch_elem(ctx, n, so, s);
    } else {
      OpBiS<T, false> so{};
      so = base(so, 3);
      so.r = (const T *)v.r; so.v = (const T *)v.v; so.dinv = dinv; so.s = (T *)v.s; so.sh = (T *)v.s;
      launch_elem(ctx, n, so, s);
    }
    const T *sh = (const T *)(prec ? v.sh : v.s);
    OpBiT<T> to{};
    to = base(to, 4);
    to.sh = sh; to.s = (const T *)v.s; to.t = (T *)v.t;
    launch_spmv_op<T>(A, to, s);
    OpBiX<T> xo{};
    xo = base(xo, 5);
    xo.x = mixed ? (T *)v.xl : (T *)x; xo.r = (T *)v.r; xo.ph = ph; xo.sh = sh; xo.s = (const T *)v.s;
    xo.t = (const T *)v.t; xo.rhat = (const T *)v.rhat;
    launch_elem(ctx, n, xo, s);
  }

  template <class T>
  void refine(const Vecs &v) {
    if (mixed) {
      OpFold<T> f{};
      f = base(f, 6);
      f.xh = x;
      f.xl = (T *)v.xl;
      launch_elem(ctx, n, f, s);
    }
    residual<T>(1, v);
  }

  template <class T>
  void run(yafel_hip_solve_info *info) {
    YhWorkspace &w = ctx->ws;
    const size_t es = sizeof(T);
    // inner vectors carved from the FP64 workspace (FP32 uses half of each)
    auto vec = [&](int i) { return w.vec64[i].p; };
    Vecs v{vec(0), prec ? vec(1) : vec(0), vec(2), vec(3), vec(4), vec(5), vec(6), vec(7), vec(8), vec(9), vec(1), vec(3)};
    (void)es;
    if (!cgfam) {
      // BiCGSTAB: v.z unused; ph uses vec(3) (p1 slot), q unused
      v.z = vec(0);
    }
    if (mixed) HIPCHK(hipMemsetAsync(v.xl, 0, n * sizeof(T), s));
    if (cgfam) {
      HIPCHK(hipMemsetAsync(v.p0, 0, n * sizeof(T), s));
    } else {
      HIPCHK(hipMemsetAsync(v.p0, 0, n * sizeof(T), s));
      HIPCHK(hipMemsetAsync(v.v, 0, n * sizeof(T), s));
    }
    // state
    KState h{};
    h.tol = o.tol;
    h.stop = o.stop;
    h.delta2 = o.refine_delta * o.refine_delta;
    h.max_iter = o.max_iter;
    h.method = o.method;
    h.mixed = mixed ? 1 : 0;
    std::memcpy(ctx->pinned, &h, sizeof(KState));
    HIPCHK(hipMemcpyAsync(st, ctx->pinned, sizeof(KState), hipMemcpyHostToDevice, s));
    HIPCHK(hipStreamSynchronize(s));
    residual<T>(0, v);
    if (!cgfam) HIPCHK(hipMemcpyAsync(v.rhat, v.r, n * sizeof(T), hipMemcpyDeviceToDevice, s));
    const int K = o.check_every > 0 ? o.check_every : 16;
    const bool with_refine = mixed || !cgfam;
    char *pin = (char *)ctx->pinned;
    const size_t slot = 256;
    hipEvent_t ev[2];
    HIPCHK(hipEventCreateWithFlags(&ev[0], hipEventDisableTiming));
    HIPCHK(hipEventCreateWithFlags(&ev[1], hipEventDisableTiming));
    auto chunk = [&](int slot_i) {
      for (int k = 0; k < K; ++k) {
        if (cgfam) {
          cg_iteration<T>(v);
        } else {
          bicg_iteration<T>(v);
        }
      }
      if (with_refine) refine<T>(v);
      HIPCHK(hipMemcpyAsync(pin + slot * (1 + slot_i), st, sizeof(KState), hipMemcpyDeviceToHost, s));
      HIPCHK(hipEventRecord(ev[slot_i], s));
    };
    chunk(0);
    chunk(1);
    int polled = 0;
    int64_t guard = 0;
    const auto t_start = std::chrono::steady_clock::now();
    while (true) {
      HIPCHK(hipEventSynchronize(ev[polled & 1]));
      KState hs;
      std::memcpy(&hs, pin + slot * (1 + (polled & 1)), sizeof(KState));
      if (hs.done) break;
      chunk(polled & 1);
      ++polled;
      // every chunk advances at least one iteration or one refinement
      if (++guard > 2 * o.max_iter + 1000) fail(YAFEL_HIP_EDEVICE, "solver: no progress");
      if ((guard & 63) == 0 &&
          std::chrono::steady_clock::now() - t_start > std::chrono::seconds(900))
        fail(YAFEL_HIP_EDEVICE, "solver: exceeded 900 s");
    }
    HIPCHK(hipStreamSynchronize(s));
    (void)hipEventDestroy(ev[0]);
    (void)hipEventDestroy(ev[1]);
    KState fs;
    HIPCHK(hipMemcpy(&fs, st, sizeof(KState), hipMemcpyDeviceToHost));
    if (fs.special == 1) HIPCHK(hipMemcpyAsync(x, b, n * 8, hipMemcpyDeviceToDevice, s));
    if (fs.special == 2) HIPCHK(hipMemsetAsync(x, 0, n * 8, s));
    // report the FP64 true residual
    OpReport rep{};
    rep = base(rep, 7);
    rep.xh = x;
    rep.b = b;
    rep.out = part + 4 * (size_t)ctx->num_cus * 8;  // past the NP x grid partials
    launch_spmv_op<double>(A, rep, s);
    double rb[2];
    HIPCHK(hipMemcpyAsync(rb, rep.out, 16, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    info->status = fs.status;
    info->iterations = fs.iter;
    info->refinements = fs.refinements;
    info->criterion = fs.crit;
    info->rel_residual = rb[1] > 0 ? std::sqrt(rb[0] / rb[1]) : 0.0;
  }
};

}  // namespace

extern "C" {

int yafel_hip_solve_opts_default(yafel_hip_solve_opts *o) {
  if (!o) return YAFEL_HIP_EINVAL;
  o->method = YAFEL_HIP_CG;
  o->precision = YAFEL_HIP_FP64;
  o->stop = YAFEL_HIP_STOP_REFERENCE;
  o->tol = 1.0e-14;
  o->max_iter = 0;
  o->check_every = 16;
  o->refine_delta = 0.1;
  return YAFEL_HIP_OK;
}

int yafel_hip_solve(yafel_hip_matrix A, const yafel_hip_solve_opts *opts, const double *b_dev,
                    double *x_dev, yafel_hip_solve_info *info) {
  YH_TRY_BEGIN
  REQUIRE(A && opts && info, "solve: null argument");
  REQUIRE(A->rows == A->cols, "solve: matrix must be square");
  REQUIRE(A->rows == 0 || (b_dev && x_dev), "solve: null vector");
  REQUIRE(opts->method >= 0 && opts->method <= 3, "solve: unknown method");
  REQUIRE(opts->precision == YAFEL_HIP_FP64 || opts->precision == YAFEL_HIP_MIXED, "solve: precision must be FP64 or MIXED");
  REQUIRE(opts->stop == 0 || opts->stop == 1, "solve: unknown stopping rule");
  yafel_hip_ctx_s *ctx = A->ctx;
  HIPCHK(hipSetDevice(ctx->device));
  std::memset(info, 0, sizeof(*info));
  const int64_t n = A->rows;
  if (n == 0) return YAFEL_HIP_OK;
  REQUIRE(A->has64, "solve: matrix has no FP64 values");
  Solver S{};
  S.ctx = ctx;
  S.A = A;
  S.o = *opts;
  S.n = n;
  S.cgfam = opts->method == YAFEL_HIP_CG || opts->method == YAFEL_HIP_PCG_JACOBI;
  S.prec = opts->method == YAFEL_HIP_PCG_JACOBI || opts->method == YAFEL_HIP_BICGSTAB_JACOBI;
  S.mixed = opts->precision == YAFEL_HIP_MIXED;
  if (S.o.max_iter <= 0) S.o.max_iter = S.cgfam ? n : 10 * n;
  if (S.o.refine_delta <= 0) S.o.refine_delta = 0.1;
  if (S.o.check_every <= 0) S.o.check_every = 16;
  S.s = ctx->stream;
  if (S.prec || true) {
    if (!A->diag_ok) matrix_build_diag(A);
  }
  if (S.mixed && !A->has32) matrix_build_fp32(A);
  ensure_ws(ctx, n);
  S.st = ctx->ws.state.as<KState>();
  S.cnt = S.st->counter;
  S.part = ctx->ws.partials.as<double>();
  S.b = b_dev;
  S.x = x_dev;
  HIPCHK(hipEventRecord(ctx->t0, S.s));
  if (S.mixed) S.run<float>(info);
  else S.run<double>(info);
  HIPCHK(hipEventRecord(ctx->t1, S.s));
  HIPCHK(hipEventSynchronize(ctx->t1));
  float ms = 0;
  HIPCHK(hipEventElapsedTime(&ms, ctx->t0, ctx->t1));
  info->ms = ms;
  return YAFEL_HIP_OK;
  YH_TRY_END(A ? A->ctx : nullptr)
}

int yafel_hip_solve_host(yafel_hip_matrix A, const yafel_hip_solve_opts *opts, const double *b,
                         const double *x0, double *x, yafel_hip_solve_info *info) {
  YH_TRY_BEGIN
  REQUIRE(A && opts && info && (A->rows == 0 || (b && x)), "solve_host: null argument");
  yafel_hip_ctx_s *ctx = A->ctx;
  HIPCHK(hipSetDevice(ctx->device));
  const int64_t n = A->rows;
  DBuf db, dx;
  db.alloc(std::max<int64_t>(n, 1) * 8);
  dx.alloc(std::max<int64_t>(n, 1) * 8);
  if (n) {
    HIPCHK(hipMemcpyAsync(db.p, b, n * 8, hipMemcpyHostToDevice, ctx->stream));
    if (x0) HIPCHK(hipMemcpyAsync(dx.p, x0, n * 8, hipMemcpyHostToDevice, ctx->stream));
    else HIPCHK(hipMemsetAsync(dx.p, 0, n * 8, ctx->stream));
  }
  int rc = yafel_hip_solve(A, opts, db.as<double>(), dx.as<double>(), info);
  if (rc != YAFEL_HIP_OK) return rc;
  if (n) HIPCHK(hipMemcpyAsync(x, dx.p, n * 8, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return YAFEL_HIP_OK;
  YH_TRY_END(A ? A->ctx : nullptr)
}

}  // extern "C"
