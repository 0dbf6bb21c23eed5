// Device-resident Krylov solvers: CG / Jacobi-PCG / BiCGSTAB (+ right Jacobi),
// in FP64 or with FP32 matrix streams inside FP64 refinement.
//
// Every scalar recurrence (alpha, beta, omega, rho, the stopping test) runs
// on the device in FP64 from per-workgroup partials folded in a fixed order
// (deterministic), so a solve needs no host round trip per iteration.  The
// host launches chunks of iterations and polls a status word once per chunk;
// kernels after convergence return immediately.
//
// One CG iteration is three kernels and no separate fold:
//   Q: q = A p, partials p.q
//   U: alpha = rho / p.q folded by every workgroup; r -= alpha q,
//      z = D^-1 r, partials r.z, r.r
//   P: beta and the stopping test folded by every workgroup; x += alpha p,
//      p = z + beta p
// or, by default on smaller matrices, two: P moves into the next SpMV, which
// forms p = z + beta p in its gather (k_spmv_qp, OpCgQP).
// BiCGSTAB folds each reduction in a one-workgroup kernel launched right
// after its producer.
// MIXED: vectors and matrix are FP32 (row sums and dots accumulate in FP64);
// MIXED_MATRIX: only the matrix is streamed in FP32, vectors stay FP64.  In
// both, when the recurrence residual has dropped by refine_delta (or claims
// convergence) the inner loop pauses, x_hi += x_lo and r = b - A x_hi are
// formed with the FP64 matrix (a residual replacement that keeps the search
// direction), and the FP64 true residual decides convergence.
#include <chrono>
#include <mutex>
#include <string>
#include <cmath>
#include <hip/hip_ext.h>

#include "internal.hpp"
#include "kernels.hpp"

using namespace yh;

#include "krylov_ops.hpp"

using namespace yh::kr;

namespace {

// ---------------------------------------------------------------- launchers
template <class T>
SellView<T> view(yafel_hip_matrix_s *A) {
  return SellView<T>{A->sptr.as<int>(), A->perm.p ? A->perm.as<int>() : nullptr, A->col.as<int>(),
                     (const T *)(sizeof(T) == 8 ? A->val64.p : A->val32.p), A->brows, A->ntot,
                     A->col16.p ? A->col16.as<short>() : nullptr, keep_slots(A, (int)sizeof(T))};
}

// Fused SpMV grids: one slice per wavefront up to num_cus * 16 workgroups
// (a matrix of 16 k slices, c3's, runs one slice per wavefront instead of
// two), so that NP <= 4 partial sums per workgroup fit the partials
// workspace of 8 * fused_cap doubles.
inline int fused_cap(const yafel_hip_ctx_s *ctx) { return ctx->num_cus * 16; }

template <class T, class Op>
void launch_spmv_op(yafel_hip_matrix_s *A, const Op &op_in, hipStream_t s) {
  const int grid = sell_grid(A->nsl, fused_cap(A->ctx));
  const SellView<T> v = view<T>(A);
  hipEvent_t e0 = nullptr, e1 = nullptr;
  Op op = op_in;
  if constexpr (Op::kTimed) ktime_pair(A->ctx, &e0, &e1, &op.work);
  auto kern = A->format == YAFEL_HIP_CSR ? k_spmv_fused<YAFEL_HIP_CSR, kSellUnroll, T, Op>
                                         : k_spmv_fused<YAFEL_HIP_BCSR3, kSellUnroll, T, Op>;
  if (e0) hipExtLaunchKernelGGL(kern, dim3(grid), dim3(NT), 0, s, e0, e1, 0, v, op);
  else kern<<<grid, NT, 0, s>>>(v, op);
  HIPCHK(hipGetLastError());
  if constexpr (Op::NP > 0) {
    k_finish<Op><<<1, NTF, 0, s>>>(op, grid);
    HIPCHK(hipGetLastError());
  }
}

// The CG iteration's SpMV: partials only, folded by the next kernel.
template <class T, class Op>
int launch_spmv_nofold(yafel_hip_matrix_s *A, const Op &op_in, hipStream_t s) {
  const int grid = sell_grid(A->nsl, fused_cap(A->ctx));
  const SellView<T> v = view<T>(A);
  hipEvent_t e0 = nullptr, e1 = nullptr;
  Op op = op_in;
  if constexpr (Op::kTimed) ktime_pair(A->ctx, &e0, &e1, &op.work);
  auto kern = A->format == YAFEL_HIP_CSR && A->unroll == 8 ? k_spmv_fused<YAFEL_HIP_CSR, 8, T, Op>
              : A->format == YAFEL_HIP_CSR               ? k_spmv_fused<YAFEL_HIP_CSR, kSellUnroll, T, Op>
                                                         : k_spmv_fused<YAFEL_HIP_BCSR3, kSellUnroll, T, Op>;
  // timed launches carry their own start/stop events: the events then hold the
  // dispatch's begin and end time stamps, not those of separate markers
  if (e0) hipExtLaunchKernelGGL(kern, dim3(grid), dim3(NT), 0, s, e0, e1, 0, v, op);
  else kern<<<grid, NT, 0, s>>>(v, op);
  HIPCHK(hipGetLastError());
  return grid;
}

// The two-pass CG's SpMV (k_spmv_qp): versioned like the update passes, with
// the SpMV grid.
template <class T, class Op>
int launch_qp(yafel_hip_matrix_s *A, const Op &op_in, int64_t n, hipStream_t s) {
  const int grid = sell_grid(A->nsl, fused_cap(A->ctx));
  const SellView<T> v = view<T>(A);
  hipEvent_t e0 = nullptr, e1 = nullptr;
  Op op = op_in;
  if constexpr (Op::kTimed) ktime_pair(A->ctx, &e0, &e1, &op.work);
  auto kern = A->format == YAFEL_HIP_CSR && A->unroll == 8 ? k_spmv_qp<YAFEL_HIP_CSR, 8, T, Op>
              : A->format == YAFEL_HIP_CSR               ? k_spmv_qp<YAFEL_HIP_CSR, kSellUnroll, T, Op>
                                                         : k_spmv_qp<YAFEL_HIP_BCSR3, kSellUnroll, T, Op>;
  if (e0) hipExtLaunchKernelGGL(kern, dim3(grid), dim3(NT), 0, s, e0, e1, 0, v, op, n);
  else kern<<<grid, NT, 0, s>>>(v, op, n);
  HIPCHK(hipGetLastError());
  return grid;
}

template <class Op>
void launch_elem(yafel_hip_ctx_s *ctx, int64_t n, const Op &op, hipStream_t s) {
  const int grid = grid_for(ctx, n, NT);
  k_elem<Op><<<grid, NT, 0, s>>>(n, op);
  HIPCHK(hipGetLastError());
  if constexpr (Op::NP > 0) {
    k_finish<Op><<<1, NTF, 0, s>>>(op, grid);
    HIPCHK(hipGetLastError());
  }
}

// Default number of CG passes.  The two-pass form saves a launch and the P
// pass (about 7 us an iteration) but gathers z and p instead of p, which
// costs in proportion to the stored entries and more with FP64 vectors.
// Measured per iteration (MI355X): FP32 vectors gain 14-21 % up to 8 M stored
// scalars (2-D Poisson, 3-D elasticity 32^3 in CSR and BCSR3) and lose 8 %
// at 65 M (elasticity 64^3); FP64 vectors gain on CSR up to about 1.5 M
// entries and lose on BCSR3 from 16^3 on.
bool two_pass_pays(const yafel_hip_matrix_s *A, int precision) {
  const int64_t stored = A->format == YAFEL_HIP_BCSR3 ? 9 * A->nblk : A->nblk;
  if (precision == YAFEL_HIP_MIXED) return stored <= ((int64_t)16 << 20);
  return A->format == YAFEL_HIP_CSR && stored <= ((int64_t)3 << 19);
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
  if (!w.partials.p) w.partials.alloc((size_t)fused_cap(ctx) * 8 * sizeof(double) + 4096);
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
  KState *st;        // two versions: st[0], st[1]
  int64_t ver = 0;   // st[ver & 1] is current in stream order
  double *part;
  void *wv[10];      // ten n-vector slots of 8 bytes an element (FP32 vectors use half)
  const double *b;
  double *x;
  hipStream_t s;

  bool two_pass = false;  // CG iterations as QP, U (k_spmv_qp)
  KState *cur() const { return st + (ver & 1); }
  // partial regions: the CG SpMV's at 0, the update pass's at 2 * cap
  int cap() const { return fused_cap(ctx); }

  template <class O>
  O base(O op, int counter) const {
    op.st = cur();
    op.part = part;
    op.counter = cur()->counter + counter;
    return op;
  }

  // launch a versioned elementwise pass: reads cur(), publishes the next
  // version; returns its workgroup count (its partials' stride)
  template <class O>
  int vlaunch(O op, const double *pin, int nin, double *ppart) {
    op.in = cur();
    op.out = st + ((ver + 1) & 1);
    op.pin = pin;
    op.nin = nin;
    op.part = ppart;
    const int g = vgrid<typename O::value_type>();
    k_velem<O><<<g, NT, 0, s>>>(n, op);
    HIPCHK(hipGetLastError());
    ++ver;
    return g;
  }
  // ve_of<T>() elements per thread; a multiple of 8 workgroups keeps the
  // XCD-aware block order, and rounding up keeps each workgroup's chunk
  // <= that many times NT
  template <class T>
  int vgrid() const {
    constexpr int VE = ve_of<T>();
    int64_t g = (n + (int64_t)NT * VE - 1) / ((int64_t)NT * VE);
    if (g >= 8) g = (g + 7) / 8 * 8;
    if (g > cap()) g = cap();
    if (g < 1) g = 1;
    return (int)g;
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

  // TM: stored matrix type of the inner SpMV; T: inner vector type
  template <class TM, class T>
  void cg_iteration(const Vecs &v) {
    OpCgQ<T> a{};
    a = base(a, 1);
    a.p = (const T *)v.p0;
    a.q = (T *)v.q;
    const int ga = launch_spmv_nofold<TM>(A, a, s);
    const T *dinv = (const T *)(sizeof(T) == 8 ? A->dinv64.p : A->dinv32.p);
    T *xx = mixed ? (T *)v.xl : (T *)x;
    double *pu = part + 2 * (size_t)cap();
    int gu;
    if (prec) {
      OpCgU<T, true> u{};
      u.r = (T *)v.r; u.z = (T *)v.z; u.q = (const T *)v.q; u.dinv = dinv;
      gu = vlaunch(u, part, ga, pu);
    } else {
      OpCgU<T, false> u{};
      u.r = (T *)v.r; u.z = (T *)v.z; u.q = (const T *)v.q; u.dinv = dinv;
      gu = vlaunch(u, part, ga, pu);
    }
    OpCgP<T> pp{};
    pp.p = (T *)v.p0;
    pp.x = xx;
    pp.z = (const T *)(prec ? v.z : v.r);
    vlaunch(pp, pu, gu, nullptr);
  }

  // QP (versioned SpMV: beta, x += alpha p, q = A (z + beta p)), then U
  template <class TM, class T>
  void cg_iteration2(const Vecs &v) {
    double *pu = part + 2 * (size_t)cap();
    OpCgQP<T> a{};
    a.in = cur();
    a.out = st + ((ver + 1) & 1);
    a.pin = pu;
    a.nin = vgrid<T>();
    a.part = part;
    a.p0 = (T *)v.p0;
    a.p1 = (T *)v.p1;
    a.q = (T *)v.q;
    a.xv = mixed ? (T *)v.xl : (T *)x;
    a.z = (const T *)(prec ? v.z : v.r);
    const int ga = launch_qp<TM>(A, a, n, s);
    ++ver;
    const T *dinv = (const T *)(sizeof(T) == 8 ? A->dinv64.p : A->dinv32.p);
    if (prec) {
      OpCgU<T, true> u{};
      u.r = (T *)v.r; u.z = (T *)v.z; u.q = (const T *)v.q; u.dinv = dinv;
      vlaunch(u, part, ga, pu);
    } else {
      OpCgU<T, false> u{};
      u.r = (T *)v.r; u.z = (T *)v.z; u.q = (const T *)v.q; u.dinv = dinv;
      vlaunch(u, part, ga, pu);
    }
  }

  // p = z + beta p after the initial residual or a refinement
  template <class T>
  void cg_restart(const Vecs &v) {
    OpCgP0<T> p0{};
    p0.p = (T *)v.p0;
    p0.z = (const T *)(prec ? v.z : v.r);
    vlaunch(p0, nullptr, 0, nullptr);
  }

  template <class TM, class T>
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
    launch_spmv_op<TM>(A, vv, s);
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
    launch_spmv_op<TM>(A, to, s);
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
    if (cgfam && !two_pass) cg_restart<T>(v);
  }

  template <class TM, class T>
  void run(yafel_hip_solve_info *info) {
    // inner vectors in the workspace slots (FP32 uses half of each)
    auto vec = [&](int i) { return wv[i]; };
    Vecs v{vec(0), prec ? vec(1) : vec(0), vec(2), vec(3), vec(4), vec(5), vec(6), vec(7), vec(8), vec(9), vec(1), vec(3)};
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
    ver = 0;
    HIPCHK(hipMemcpyAsync(cur(), ctx->pinned, sizeof(KState), hipMemcpyHostToDevice, s));
    HIPCHK(hipStreamSynchronize(s));
    residual<T>(0, v);
    if (!cgfam) HIPCHK(hipMemcpyAsync(v.rhat, v.r, n * sizeof(T), hipMemcpyDeviceToDevice, s));
    if (cgfam && !two_pass) cg_restart<T>(v);
    const int K = o.check_every > 0 ? o.check_every : 16;
    const bool with_refine = mixed || !cgfam;
    char *pin = (char *)ctx->pinned;
    const size_t slot = 256;
    static_assert(sizeof(KState) <= slot && 2 * sizeof(KState) <= 4096, "KState slots");
    hipEvent_t *ev = ctx->sev;
    auto chunk = [&](int slot_i) {
      for (int k = 0; k < K; ++k) {
        if (cgfam && two_pass) {
          cg_iteration2<TM, T>(v);
        } else if (cgfam) {
          cg_iteration<TM, T>(v);
        } else {
          bicg_iteration<TM, T>(v);
        }
      }
      if (with_refine) refine<T>(v);
      HIPCHK(hipMemcpyAsync(pin + slot * (1 + slot_i), cur(), sizeof(KState), hipMemcpyDeviceToHost, s));
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
    KState fs;
    HIPCHK(hipMemcpy(&fs, cur(), sizeof(KState), hipMemcpyDeviceToHost));
    if (fs.special == 1) HIPCHK(hipMemcpyAsync(x, b, n * 8, hipMemcpyDeviceToDevice, s));
    if (fs.special == 2) HIPCHK(hipMemsetAsync(x, 0, n * 8, s));
    // report the FP64 true residual
    OpReport rep{};
    rep = base(rep, 7);
    rep.xh = x;
    rep.b = b;
    rep.out = part + 4 * (size_t)cap();  // past the NP x grid partials
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

}  // extern "C"

namespace {

// Workspace of one solve: ten n-vector slots of 8-byte elements, the
// partials area (8 doubles per fused-SpMV workgroup plus 4 KiB) and two
// solver states (4 KiB).
struct SolveWs {
  void *vec[10];
  double *part;
  KState *st;
};

size_t partials_bytes(const yafel_hip_ctx_s *ctx) { return (size_t)fused_cap(ctx) * 8 * sizeof(double) + 4096; }

// Checks the arguments and returns the solver, or nullptr when there is
// nothing to solve (n == 0).
bool solve_setup(Solver &S, yafel_hip_matrix_s *A, const yafel_hip_solve_opts *opts, const double *b_dev,
                 double *x_dev, yafel_hip_solve_info *info, const char *who) {
  const std::string w(who);
  REQUIRE(A && opts && info, w + ": null argument");
  REQUIRE(A->rows == A->cols, w + ": matrix must be square");
  REQUIRE(A->rows == 0 || (b_dev && x_dev), w + ": null vector");
  REQUIRE(opts->method >= 0 && opts->method <= 3, w + ": unknown method");
  REQUIRE(opts->precision == YAFEL_HIP_FP64 || opts->precision == YAFEL_HIP_MIXED ||
              opts->precision == YAFEL_HIP_MIXED_MATRIX,
          w + ": precision must be FP64, MIXED or MIXED_MATRIX");
  REQUIRE(opts->stop == 0 || opts->stop == 1, w + ": unknown stopping rule");
  HIPCHK(hipSetDevice(A->ctx->device));
  std::memset(info, 0, sizeof(*info));
  const int64_t n = A->rows;
  if (n == 0) return false;
  REQUIRE(A->has64, w + ": matrix has no FP64 values");
  S.ctx = A->ctx;
  S.A = A;
  S.o = *opts;
  S.n = n;
  S.cgfam = opts->method == YAFEL_HIP_CG || opts->method == YAFEL_HIP_PCG_JACOBI;
  S.prec = opts->method == YAFEL_HIP_PCG_JACOBI || opts->method == YAFEL_HIP_BICGSTAB_JACOBI;
  S.mixed = opts->precision == YAFEL_HIP_MIXED || opts->precision == YAFEL_HIP_MIXED_MATRIX;
  if (S.o.max_iter <= 0) S.o.max_iter = S.cgfam ? n : 10 * n;
  if (S.o.refine_delta <= 0) S.o.refine_delta = 0.1;
  if (S.o.check_every <= 0) S.o.check_every = 16;
  S.two_pass = S.cgfam && (A->cg_passes == 2 || (A->cg_passes == 0 && two_pass_pays(A, opts->precision)));
  S.s = S.ctx->stream;
  S.b = b_dev;
  S.x = x_dev;
  return true;
}

void solve_run(Solver &S, const SolveWs &ws, yafel_hip_solve_info *info) {
  yafel_hip_ctx_s *ctx = S.ctx;
  for (int i = 0; i < 10; ++i) S.wv[i] = ws.vec[i];
  S.st = ws.st;
  S.part = ws.part;
  HIPCHK(hipEventRecord(ctx->t0, S.s));
  if (S.o.precision == YAFEL_HIP_MIXED) S.run<float, float>(info);
  else if (S.o.precision == YAFEL_HIP_MIXED_MATRIX) S.run<float, double>(info);
  else S.run<double, double>(info);
  HIPCHK(hipEventRecord(ctx->t1, S.s));
  HIPCHK(hipEventSynchronize(ctx->t1));
  float ms = 0;
  HIPCHK(hipEventElapsedTime(&ms, ctx->t0, ctx->t1));
  info->ms = ms;
}

// ------------------------------------------------- allocation-free solve
// The scratch of yafel_hip_solve_noalloc: a file-scope device array, one
// copy per GPU, held until the process ends.  In host code `scratch` names
// a host-side stand-in, not device memory: its device address is looked up
// once per device (noalloc_scratch_init, from yafel_hip_ctx_create) and only
// that address is handed to kernels and copies.
constexpr size_t kNoallocVec = (((size_t)YAFEL_HIP_NOALLOC_MAX_ROWS * 8) + 255) & ~(size_t)255;
constexpr int kNoallocMaxCus = 512;  // partials sized for up to 512 compute units
constexpr size_t kNoallocPart = (size_t)kNoallocMaxCus * 16 * 8 * sizeof(double) + 4096;
constexpr size_t kNoallocState = 4096;
constexpr size_t kNoallocBytes = 10 * kNoallocVec + kNoallocPart + kNoallocState;
static_assert(2 * sizeof(KState) <= kNoallocState, "two solver states fit the state area");
constexpr int kNoallocDevices = 64;

__device__ __align__(256) unsigned char scratch[kNoallocBytes];

std::mutex g_scratch_mu;
void *g_scratch[kNoallocDevices];  // device address of `scratch` per device (nullptr: not looked up)

}  // namespace

namespace yh {
void noalloc_scratch_init(int device) {
  REQUIRE(device >= 0 && device < kNoallocDevices, "noalloc scratch: device index out of range");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch[device]) return;
  int cur = -1;
  HIPCHK(hipGetDevice(&cur));
  REQUIRE(cur == device, "noalloc scratch: the device must be current");
  void *p = nullptr;
  HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  REQUIRE(p != nullptr, "noalloc scratch: no device address");
  g_scratch[device] = p;
}
}  // namespace yh

extern "C" {

int yafel_hip_solve(yafel_hip_matrix A, const yafel_hip_solve_opts *opts, const double *b_dev,
                    double *x_dev, yafel_hip_solve_info *info) {
  YH_TRY_BEGIN
  Solver S{};
  if (!solve_setup(S, A, opts, b_dev, x_dev, info, "solve")) return YAFEL_HIP_OK;
  yafel_hip_ctx_s *ctx = A->ctx;
  if (!A->diag_ok) matrix_build_diag(A);
  if (S.mixed && !A->has32) matrix_build_fp32(A);
  ensure_ws(ctx, S.n);
  SolveWs ws;
  for (int i = 0; i < 10; ++i) ws.vec[i] = ctx->ws.vec64[i].p;
  ws.part = ctx->ws.partials.as<double>();
  ws.st = ctx->ws.state.as<KState>();
  solve_run(S, ws, info);
  return YAFEL_HIP_OK;
  YH_TRY_END(A ? A->ctx : nullptr)
}

int yafel_hip_solve_noalloc(yafel_hip_matrix A, const yafel_hip_solve_opts *opts, const double *b_dev,
                            double *x_dev, yafel_hip_solve_info *info) {
  YH_TRY_BEGIN
  Solver S{};
  if (!solve_setup(S, A, opts, b_dev, x_dev, info, "solve_noalloc")) return YAFEL_HIP_OK;
  yafel_hip_ctx_s *ctx = A->ctx;
  REQUIRE(S.n <= YAFEL_HIP_NOALLOC_MAX_ROWS,
          "solve_noalloc: " + std::to_string(S.n) + " rows, more than YAFEL_HIP_NOALLOC_MAX_ROWS (" +
              std::to_string((long long)YAFEL_HIP_NOALLOC_MAX_ROWS) + ")");
  REQUIRE(partials_bytes(ctx) <= kNoallocPart, "solve_noalloc: device has more compute units than the scratch serves");
  void *base = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    base = ctx->device < kNoallocDevices ? g_scratch[ctx->device] : nullptr;
  }
  REQUIRE(base, "solve_noalloc: scratch not set up on this device");
  // the diagonal and the FP32 copy are refreshed in place; their buffers
  // must exist already
  if (!A->diag_ok) {
    REQUIRE(A->diag64.p, "solve_noalloc: the matrix has no diagonal yet (call yafel_hip_matrix_diagonal once)");
    matrix_build_diag(A);
  }
  if (S.mixed && !A->has32) {
    const size_t need = (size_t)std::max<int64_t>(A->ntot * (A->format == YAFEL_HIP_CSR ? 1 : 9), 1) * 4;
    REQUIRE(A->val32.p && A->val32.bytes >= need,
            "solve_noalloc: the matrix has no FP32 copy yet (call yafel_hip_matrix_sync_fp32 once)");
    matrix_build_fp32(A);
  }
  unsigned char *p = (unsigned char *)base;
  SolveWs ws;
  for (int i = 0; i < 10; ++i) ws.vec[i] = p + i * kNoallocVec;
  ws.part = (double *)(p + 10 * kNoallocVec);
  ws.st = (KState *)(p + 10 * kNoallocVec + kNoallocPart);
  solve_run(S, ws, info);
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
