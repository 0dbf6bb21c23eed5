// Krylov building blocks shared by the single-solve driver (krylov.hip) and
// the batched serving driver (serve.hip): the device solver state, the fused
// SpMV / elementwise kernel bodies with deterministic reductions (fixed-order
// folds of per-workgroup partials), and the per-method operations.
#pragma once
#include <cmath>

#include "internal.hpp"
#include "kernels.hpp"

namespace yh {
namespace kr {
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
  int upd;        // single-solve CG, two-pass form: U ran since the last QP
                  // (x += alpha p pending, U's partials valid)
  int pupd;       // single-solve CG: p = z + beta p pending (after a residual)
  int pidx;       // single-solve CG, two-pass form: p is p[pidx]
  int pad1;
  unsigned int counter[8];
};

enum { M_CG = 0, M_PCG = 1, M_BICG = 2, M_BICGJ = 3 };

__device__ __forceinline__ bool bad(double v) { return !(fabs(v) <= 1.7976931348623157e308); }

// ------------------------------------------------------------------ drivers
// Block-reduce NP sums into op.part[j * nblk + blk].  The fold of the
// partials and the scalar recurrence run in k_finish, a one-workgroup kernel
// launched right after on the same stream: kernel boundaries make the
// partials visible across XCDs without a device-scope fence in every
// workgroup (on gfx950 such a fence writes back and invalidates the whole L2,
// which would throw away the gathered vector's cache reuse).
template <int NP, class Op>
__device__ __forceinline__ void finish_partials(Op &op, double (&acc)[NP > 0 ? NP : 1], int blk, int nblk) {
  if constexpr (NP > 0) {
    __shared__ double lds[NT / 64];
#pragma unroll
    for (int j = 0; j < NP; ++j) {
      double s = dev::block_sum<NT>(acc[j], lds);
      if (threadIdx.x == 0) op.part[(size_t)j * nblk + blk] = s;
    }
  }
}

// Fold the nblk partials of each sum in block order (deterministic) and run
// op.last().  Evaluates op.skip() exactly as the producing kernel did: only
// op.last() changes the state, so both see the same flags.  Runs in one
// workgroup of NTF threads: the partial loads of all sums are independent, so
// a wide workgroup keeps them in flight together.
constexpr int NTF = 1024;

template <class Op>
__device__ __forceinline__ void finish_body(Op &op, int nblk) {
  if (op.skip()) return;
  op.begin();
  constexpr int NP = Op::NP;
  __shared__ double lds[NP > 0 ? NP : 1][NTF / 64];
  __shared__ double sums[NP > 0 ? NP : 1];
  double v[NP > 0 ? NP : 1];
#pragma unroll
  for (int j = 0; j < NP; ++j) v[j] = 0;
  for (int i = threadIdx.x; i < nblk; i += NTF)
#pragma unroll
    for (int j = 0; j < NP; ++j) v[j] += op.part[(size_t)j * nblk + i];
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
#pragma unroll
  for (int j = 0; j < NP; ++j) {
    const double t = dev::wave_sum(v[j]);
    if (l == 0) lds[j][w] = t;
  }
  __syncthreads();
  if (threadIdx.x < NP) {
    double t = 0;
#pragma unroll
    for (int q = 0; q < NTF / 64; ++q) t += lds[threadIdx.x][q];
    sums[threadIdx.x] = t;
  }
  __syncthreads();
  if (threadIdx.x == 0) op.last(sums);
}

template <class Op>
__global__ __launch_bounds__(NTF) void k_finish(Op op0, int nblk) {
  Op op = op0;
  finish_body(op, nblk);
}

// Sliced SpMV (one slice of 64 rows per wavefront, a row per lane) with a
// fused per-row epilogue and up to 4 block-reduced sums folded afterwards by
// k_finish.  The body takes its block id and block count explicitly so that a
// batched launch can run several independent solves, each on its own block
// range; the row -> block mapping depends on nothing else, so a solve's
// results do not depend on what runs beside it.
// T is the stored matrix type, Op::value_type the vector type.
// The wave's first slice's slot range is loaded before the state test, so
// that its latency overlaps the state read instead of following it.
template <int FMT, int U, class T, class Op>
__device__ __forceinline__ void spmv_fused_body(const SellView<T> &A, Op &op, int blk, int nblk) {
  using TV = typename Op::value_type;
  constexpr int RPL = FMT == YAFEL_HIP_CSR ? 1 : 3;  // rows per lane
  const int lane = threadIdx.x & 63;
  const int64_t nsl = A.nsl(), nw = (int64_t)nblk * (NT / 64);
  int64_t s = (int64_t)blk * (NT / 64) + (threadIdx.x >> 6);
  // unconditional loads at a clamped, in-bounds slice (sptr has nsl + 1
  // entries): no branch, so nothing waits on them before the state test
  const int64_t sc = s < nsl ? s : (nsl > 0 ? nsl - 1 : 0);
  int p0 = dev::load_in_place(&A.sptr[sc]), p1 = dev::load_in_place(&A.sptr[sc + (nsl > 0 ? 1 : 0)]);
  if (op.skip()) return;
  op.begin();
  constexpr int NP = Op::NP;
  double acc[NP > 0 ? NP : 1];
#pragma unroll
  for (int j = 0; j < (NP > 0 ? NP : 1); ++j) acc[j] = 0;
  bool first = true;
  for (; s < nsl; s += nw) {
    if (!first) {
      p0 = A.sptr[s];
      p1 = A.sptr[s + 1];
    }
    const int64_t i = s * 64 + lane;
    // row sums accumulate in FP64 whatever the storage and vector precision
    double a[RPL];
    if (FMT == YAFEL_HIP_CSR)
      a[0] = dev::sell_row_csr_at<U, T, double>(A, s, p0, p1, lane, [&](int c) { return (double)op.x(c); });
    else
      dev::sell_row_bcsr_at<U, T, double>(A, s, p0, p1, lane, a, [&](int64_t c) { return (double)op.x(c); });
    if (i < A.brows) {
      const int64_t r = A.row_of(i);
#pragma unroll
      for (int k = 0; k < RPL; ++k) op.row(RPL * r + k, (TV)a[k], acc);
    }
    first = false;
  }
  if constexpr (Op::kTimed) op.mark_work();
  finish_partials<NP>(op, acc, blk, nblk);
}

// at least four waves per SIMD (at most 128 VGPRs): a c2-sized grid of about
// a thousand workgroups then runs in one round of residency
template <int FMT, int U, class T, class Op>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(4))) void k_spmv_fused(SellView<T> A, Op op0) {
  Op op = op0;
  spmv_fused_body<FMT, U, T, Op>(A, op, dev::xcd_block(), gridDim.x);
}

// Elementwise pass with up to 4 block-reduced sums.
template <class Op>
__device__ __forceinline__ void elem_body(int64_t n, Op &op, int blk, int nblk) {
  if (op.skip()) return;
  op.begin();
  constexpr int NP = Op::NP;
  double acc[NP > 0 ? NP : 1];
#pragma unroll
  for (int j = 0; j < (NP > 0 ? NP : 1); ++j) acc[j] = 0;
  const int64_t chunk = ((n + nblk - 1) / nblk + 63) & ~(int64_t)63;
  const int64_t i0 = (int64_t)blk * chunk;
  const int64_t i1 = i0 + chunk < n ? i0 + chunk : n;
  for (int64_t i = i0 + threadIdx.x; i < i1; i += NT) op.elem(i, acc);
  finish_partials<NP>(op, acc, blk, nblk);
}

template <class Op>
__global__ __launch_bounds__(NT) void k_elem(int64_t n, Op op0) {
  Op op = op0;
  elem_body<Op>(n, op, dev::xcd_block(), gridDim.x);
}

// --------------------------------------------------------------------- ops
struct OpBase {
  static constexpr bool kTimed = false;  // recorded by the optional kernel timing
  KState *st;
  double *part;
  unsigned int *counter;
  // both flags read unconditionally (one scalar round trip, not two)
  __device__ bool running() const { return (st->done | st->paused) == 0; }
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

// FP64 residual r = b - A x_hi; stores r (in the inner precision TL) and z,
// and folds r.w (w = z, r or rhat), r.r, the FP64 true r.r and b.b.
// MODE 0 = initialisation, 1 = refinement / verification.
template <class TL, int MODE>
struct OpResidual : OpBase {
  using value_type = double;
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
      st->pupd = cgfam ? 1 : 0;
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
      st->pupd = 1;
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
  using value_type = T;
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
  using value_type = T;
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
  using value_type = double;
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

// --------------------------------------------- CG with folds in the consumer
// The single-solve CG runs three kernels per iteration and no one-workgroup
// fold:
//   Q: q = A p, partial p.q                                (sliced SpMV)
//   U: alpha from the p.q partials; r -= alpha q, z = D^-1 r, partials r.z,
//      r.r
//   P: beta and the stopping test from the r.z, r.r partials; x += alpha p,
//      p = z + beta p
// Every workgroup of U and P folds the producer's partials itself, in one
// fixed order, so all workgroups derive bit-identical scalars and take the
// same branch.  The state is versioned: U and P read KState `in` and their
// workgroup 0 writes the updated copy to `out` (the other slot of a pair), so
// no workgroup reads a field that another workgroup of the same kernel
// writes.  The kernel boundary makes `out` and the partials visible to the
// next kernel.

// Sum of each of NPIN partial arrays part[j * nblk + i], i < nblk, in a fixed
// order; the result is the same in every thread of every workgroup.  Split in
// two so that the loads (refold_load, this thread's share) can be issued
// before other waits: refold_reduce combines the shares across the workgroup.
// Up to RM of the thread's partials per sum are loaded before the first add
// (a round trip per RM partials instead of one per few); the adds keep the
// order, and the zeros added past nblk change nothing (v never becomes -0).
constexpr int RM = 8;
template <int NPIN>
__device__ __forceinline__ void refold_load(const double *__restrict__ part, int nblk, double (&v)[NPIN]) {
#pragma unroll
  for (int j = 0; j < NPIN; ++j) v[j] = 0;
  for (int i0 = threadIdx.x; i0 < nblk; i0 += RM * NT) {
    double t[NPIN][RM];
#pragma unroll
    for (int m = 0; m < RM; ++m)
#pragma unroll
      for (int j = 0; j < NPIN; ++j) t[j][m] = i0 + m * NT < nblk ? part[(size_t)j * nblk + i0 + m * NT] : 0.0;
#pragma unroll
    for (int m = 0; m < RM; ++m)
#pragma unroll
      for (int j = 0; j < NPIN; ++j) v[j] += t[j][m];
  }
}

template <int NPIN>
__device__ __forceinline__ void refold_reduce(const double (&v)[NPIN], double (&s)[NPIN]) {
  __shared__ double lds[NPIN][NT / 64];
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
#pragma unroll
  for (int j = 0; j < NPIN; ++j) {
    const double t = dev::wave_sum(v[j]);
    if (l == 0) lds[j][w] = t;
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < NPIN; ++j) {
    double t = 0;
#pragma unroll
    for (int q = 0; q < NT / 64; ++q) t += lds[j][q];
    s[j] = t;
  }
}

// Versioned elementwise pass: fold the producer's NPIN partials, let the op
// update the state (decide), publish it from workgroup 0, then finish the
// elementwise body with up to NP partial sums of its own.  Each thread loads
// its first VE = ve_of<value_type>() elements' inputs before the fold, so
// their latency overlaps the fold's; the host sizes the grid so that VE
// elements per thread cover the vector.  Thread 0 of each workgroup keeps the state in LDS and
// runs decide(); the other threads read the outcome from there (no
// per-thread copy of the state, which would land in scratch memory).
// Elements per thread, by vector type: on the 64^3 elasticity mixed solve
// (FP32 vectors) 8 beat 4 and 16 (62.0 ms per step against 62.9 and 63.6);
// on the Q4 256^2 Poisson FP64 solve 4 beat 8 (435 against 439 ms).
template <class T>
constexpr int ve_of() {
  return sizeof(T) == 4 ? 8 : 4;
}

// the state moves between global memory and LDS as 8-byte words, a word per
// thread
constexpr int KSW = (int)(sizeof(KState) / 8);
static_assert(sizeof(KState) % 8 == 0 && KSW <= NT, "KState words");

// The body takes its block id and block count explicitly, so that a batched
// launch can run several solves on their own block ranges; block 0 of the
// range publishes the state.
template <class Op>
__device__ __forceinline__ void velem_body(int64_t n, Op &op, int blk, int nblk) {
  constexpr int VE = ve_of<typename Op::value_type>();
  __shared__ KState ss;
  __shared__ int sgo;
  const int t = threadIdx.x;
  const bool b0 = blk == 0;
  const int64_t chunk = ((n + nblk - 1) / nblk + 63) & ~(int64_t)63;
  const int64_t i0 = (int64_t)blk * chunk + t;
  const int64_t i1 = (int64_t)blk * chunk + chunk < n ? (int64_t)blk * chunk + chunk : n;
  typename Op::In v[VE];
#pragma unroll
  for (int k = 0; k < VE; ++k)
    if (i0 + k * NT < i1) op.load(i0 + k * NT, v[k]);
  // the producer's partials are loaded before the state, so that the vector,
  // partial and state loads are in flight together
  constexpr int NPIN = Op::NPIN;
  double f[NPIN > 0 ? NPIN : 1];
  if constexpr (NPIN > 0) refold_load<NPIN>(op.pin, op.nin, f);
  if (t < KSW) reinterpret_cast<long long *>(&ss)[t] = reinterpret_cast<const long long *>(op.in)[t];
  __syncthreads();
  if (!op.active(ss)) {
    if (b0 && t < KSW) reinterpret_cast<long long *>(op.out)[t] = reinterpret_cast<const long long *>(&ss)[t];
    return;
  }
  if constexpr (NPIN > 0) refold_reduce<NPIN>(f, f);
  __syncthreads();  // every wave has read ss before thread 0 changes it
  if (t == 0) sgo = op.decide(ss, f) ? 1 : 0;
  __syncthreads();
  if (b0 && t < KSW) reinterpret_cast<long long *>(op.out)[t] = reinterpret_cast<const long long *>(&ss)[t];
  const bool go = sgo != 0;
  if (!go && !Op::kOnStop) return;
  op.begin(ss);
  constexpr int NP = Op::NP;
  double acc[NP > 0 ? NP : 1];
#pragma unroll
  for (int j = 0; j < (NP > 0 ? NP : 1); ++j) acc[j] = 0;
#pragma unroll
  for (int k = 0; k < VE; ++k)
    if (i0 + k * NT < i1) op.compute(i0 + k * NT, v[k], acc, go);
  for (int64_t i = i0 + VE * NT; i < i1; i += NT) {
    typename Op::In tmp;
    op.load(i, tmp);
    op.compute(i, tmp, acc, go);
  }
  finish_partials<NP>(op, acc, blk, nblk);
}

template <class Op>
__global__ __launch_bounds__(NT) void k_velem(int64_t n, Op op0) {
  Op op = op0;
  velem_body(n, op, dev::xcd_block(), gridDim.x);
}

struct VOpBase {
  static constexpr bool kOnStop = false;  // compute() also runs when decide() says stop
  const KState *in;
  KState *out;
  const double *pin;  // the producer's partials
  int nin;            // the producer's workgroup count
  double *part;       // this pass's partials
  static __device__ bool running(const KState &s) { return !s.done && !s.paused; }
};

// Q: q = A p and the p.q partials; reads the state, never writes it.
template <class T>
struct OpCgQ : OpBase {
  using value_type = T;
  static constexpr bool kTimed = true;  // the CG iteration's SpMV kernel
  static constexpr int NP = 1;
  const T *p;
  T *q;
  unsigned *work = nullptr;  // timed launches: set once the product ran
  __device__ bool skip() const { return !running(); }
  __device__ void begin() {}
  __device__ void mark_work() const {
    if (work && blockIdx.x == 0 && threadIdx.x == 0) *work = 1u;
  }
  __device__ T x(int64_t c) const { return p[c]; }
  __device__ void row(int64_t i, T y, double *acc) const {
    q[i] = y;
    acc[0] += (double)p[i] * (double)y;
  }
  __device__ void last(const double *) const {}
};

// U: alpha = rho / p.q; r -= alpha q, z = D^-1 r; r.z, r.r.
template <class T, bool PREC>
struct OpCgU : VOpBase {
  using value_type = T;
  static constexpr int NPIN = 1, NP = 2;
  T *r, *z;
  const T *q, *dinv;
  T alpha;
  struct In {
    T q, r, d;
  };
  __device__ bool active(const KState &s) const { return running(s); }
  __device__ bool decide(KState &s, const double *f) const {
    const double a = s.rho / f[0];
    if (f[0] == 0.0 || bad(a)) {
      s.status = YAFEL_HIP_STATUS_BREAKDOWN;
      s.done = 1;
      return false;
    }
    s.alpha = a;
    s.first = 0;
    s.upd = 1;
    return true;
  }
  __device__ void begin(const KState &s) { alpha = (T)s.alpha; }
  __device__ void load(int64_t i, In &v) const {
    v.q = q[i];
    v.r = r[i];
    if (PREC) v.d = dinv[i];
  }
  __device__ void compute(int64_t i, const In &v, double *acc, bool) const {
    const T ri = fma(-alpha, v.q, v.r);
    r[i] = ri;
    if (PREC) {
      const T zi = v.d * ri;
      z[i] = zi;
      acc[0] += (double)ri * (double)zi;
    } else {
      acc[0] += (double)ri * (double)ri;
    }
    acc[1] += (double)ri * (double)ri;
  }
};

// P: the stopping test and beta from r.z, r.r; x += alpha p (the iteration
// U finished) and, unless the solve stopped or paused for a refinement,
// p = z + beta p.
template <class T>
struct OpCgP : VOpBase {
  using value_type = T;
  static constexpr int NPIN = 2, NP = 0;
  static constexpr bool kOnStop = true;
  T *p, *x;
  const T *z;
  T alpha, beta;
  struct In {
    T z, p, x;
  };
  __device__ bool active(const KState &s) const { return running(s); }
  __device__ bool decide(KState &s, const double *f) const {
    after_iteration(&s, f[0], f[1]);
    return running(s);
  }
  __device__ void begin(const KState &s) {
    alpha = (T)s.alpha;
    beta = (T)s.beta;
  }
  __device__ void load(int64_t i, In &v) const {
    v.z = z[i];
    v.p = p[i];
    v.x = x[i];
  }
  __device__ void compute(int64_t i, const In &v, double *, bool go) const {
    x[i] = fma(alpha, v.p, v.x);
    if (go) p[i] = fma(beta, v.p, v.z);
  }
};

// p = z + beta p after the initial or a refinement residual (KState::pupd).
template <class T>
struct OpCgP0 : VOpBase {
  using value_type = T;
  static constexpr int NPIN = 0, NP = 0;
  T *p;
  const T *z;
  T beta;
  struct In {
    T z, p;
  };
  __device__ bool active(const KState &s) const { return running(s) && s.pupd; }
  __device__ bool decide(KState &s, const double *) const {
    s.pupd = 0;
    return true;
  }
  __device__ void begin(const KState &s) { beta = (T)s.beta; }
  __device__ void load(int64_t i, In &v) const {
    v.z = z[i];
    v.p = p[i];
  }
  __device__ void compute(int64_t i, const In &v, double *, bool) const { p[i] = fma(beta, v.p, v.z); }
};

// ----------------------------------------------------- CG in two passes
// On smaller matrices the launch and the fixed cost of the P pass weigh more
// than gathering two vectors, so P is folded into the next SpMV and an
// iteration is two kernels:
//   QP: fold U's r.z, r.r partials, beta and the stopping test; x += alpha p
//       for every row; q = A (z + beta p), with p_new = z + beta p formed in
//       the gather and written by each row's owner into the other p buffer;
//       partial p_new.q
//   U:  as in the three-pass form
// The arithmetic is that of Q, U, P: the same fma forms p_new and x, so the
// iterates are bit-identical.  After the initial or a refinement residual
// (no U since), QP only forms p = z + beta p.  When the test stops or pauses
// the iteration, QP applies x += alpha p over the vector and skips the SpMV.
template <int FMT, int U, class T, class Op>
__device__ __forceinline__ void spmv_qp_body(const SellView<T> &A, Op &op, int blk, int nblk, int64_t n) {
  __shared__ KState ss;
  __shared__ int sflags;
  const int t = threadIdx.x;
  const bool b0 = blk == 0;
  double f[2];
  refold_load<2>(op.pin, op.nin, f);
  if (t < KSW) reinterpret_cast<long long *>(&ss)[t] = reinterpret_cast<const long long *>(op.in)[t];
  __syncthreads();
  if (!op.active(ss)) {
    if (b0 && t < KSW) reinterpret_cast<long long *>(op.out)[t] = reinterpret_cast<const long long *>(&ss)[t];
    return;
  }
  refold_reduce<2>(f, f);
  __syncthreads();  // every wave has read ss before thread 0 changes it
  if (t == 0) sflags = op.decide(ss, f);
  __syncthreads();
  if (b0 && t < KSW) reinterpret_cast<long long *>(op.out)[t] = reinterpret_cast<const long long *>(&ss)[t];
  const int flags = sflags;
  op.begin(ss, flags);
  if (!(flags & Op::kGo)) {
    if (flags & Op::kXupd) {
      const int64_t chunk = (n + nblk - 1) / nblk;
      const int64_t i1 = (int64_t)blk * chunk + chunk < n ? (int64_t)blk * chunk + chunk : n;
      for (int64_t i = (int64_t)blk * chunk + t; i < i1; i += NT) op.xonly(i);
    }
    return;
  }
  const int lane = t & 63;
  double acc[1] = {0.0};
  const int64_t nsl = A.nsl(), nw = (int64_t)nblk * (NT / 64);
  for (int64_t sl = (int64_t)blk * (NT / 64) + (t >> 6); sl < nsl; sl += nw) {
    const int64_t i = sl * 64 + lane;
    if (FMT == YAFEL_HIP_CSR) {
      const double y = dev::sell_row_csr<U, T, double>(A, sl, lane, [&](int c) { return (double)op.x(c); });
      if (i < A.brows) op.row(A.row_of(i), y, acc);
    } else {
      double a[3];
      dev::sell_row_bcsr<U, T, double>(A, sl, lane, a, [&](int64_t c) { return (double)op.x(c); });
      if (i < A.brows) {
        const int64_t r = A.row_of(i);
        op.row(3 * r, a[0], acc);
        op.row(3 * r + 1, a[1], acc);
        op.row(3 * r + 2, a[2], acc);
      }
    }
  }
  finish_partials<1>(op, acc, blk, nblk);
}

template <int FMT, int U, class T, class Op>
__global__ __launch_bounds__(NT) void k_spmv_qp(SellView<T> A, Op op0, int64_t n) {
  Op op = op0;
  spmv_qp_body<FMT, U, T, Op>(A, op, dev::xcd_block(), gridDim.x, n);
}

template <class T>
struct OpCgQP : VOpBase {
  using value_type = T;
  static constexpr bool kTimed = true;  // the CG iteration's SpMV kernel
  static constexpr int kGo = 1, kXupd = 2;
  T *p0, *p1, *q, *xv;
  const T *z;
  const T *pold;
  T *pnew;
  T alpha, beta;
  int xupd;
  unsigned *work = nullptr;  // timed launches: set when the launch runs the product
  __device__ bool active(const KState &s) const { return running(s); }
  // flags: kGo (form p_new, q = A p_new), kXupd (x += alpha p first)
  __device__ int decide(KState &s, const double *f) const {
    int fl = 0;
    if (s.upd) {
      s.upd = 0;
      fl |= kXupd;
      after_iteration(&s, f[0], f[1]);
    }
    s.pupd = 0;
    if (running(s)) {
      fl |= kGo;
      s.pidx ^= 1;
    }
    return fl;
  }
  __device__ void begin(const KState &s, int fl) {
    alpha = (T)s.alpha;
    beta = (T)s.beta;
    xupd = fl & kXupd;
    if (work && (fl & kGo) && blockIdx.x == 0 && threadIdx.x == 0) *work = 1u;
    // s is the published state: with kGo its pidx already names p_new
    const int pi = (fl & kGo) ? (s.pidx ^ 1) : s.pidx;
    pold = pi ? p1 : p0;
    pnew = pi ? p0 : p1;
  }
  __device__ T x(int64_t c) const { return fma(beta, pold[c], z[c]); }
  __device__ void row(int64_t i, double y, double *acc) const {
    const T po = pold[i];
    const T pn = fma(beta, po, z[i]);
    const T yq = (T)y;
    pnew[i] = pn;
    q[i] = yq;
    if (xupd) xv[i] = fma(alpha, po, xv[i]);
    acc[0] += (double)pn * (double)yq;
  }
  __device__ void xonly(int64_t i) const { xv[i] = fma(alpha, pold[i], xv[i]); }
};

}  // namespace kr
}  // namespace yh
