// k right-hand sides advanced together (k <= 64 per batch; larger k runs in
// batches of 64).  Vectors are row-major n x k, so the 64 lanes of a
// wavefront are the right-hand sides: in the products a wavefront takes a
// slice's rows, reads the slice's columns and values once per slot
// (coalesced) and broadcasts each row's entry to all lanes; the gathered rows
// of X are 64-wide coalesced reads, and every per-RHS dot product is a
// per-lane accumulation.  Each right-hand side keeps its own FP64 scalars and
// its own stopping decision; a right-hand side that has converged is frozen.
//
// Methods: BiCGSTAB with right Jacobi preconditioning (ph = D^-1 p and
// sh = D^-1 s are formed on the fly, never stored) and CG / Jacobi-PCG, in
// FP64 or as FP32 inner iterations with FP64 residual replacement.
//
// Residual replacement is one FP64 product for the whole batch, so the
// right-hand sides share it: one whose recurrence residual has dropped by
// refine_delta marks itself and keeps iterating, and the batch pauses for
// the replacement once half of the active right-hand sides are marked, when
// one has dropped by refine_delta^2, claims convergence or reaches max_iter.
#include <algorithm>
#include <chrono>
#include <cmath>

#include "internal.hpp"
#include "kernels.hpp"

using namespace yh;

namespace {

constexpr int NT = 256;
constexpr int KMAX = 64;

// what a product op's row epilogue reads of its own row (pre4), loaded before
// the row's sum so that the loads overlap it
struct Pre4 {
  float4 a, b;
  bool own;  // a is the row's own staged value (own4), b unused
};

struct MState {
  int done, paused, first, k;
  int method, stop, mixed, pidx;
  int nact;  // active right-hand sides after the last fold (within the 64-byte header the host polls)
  long long iter, max_iter, refinements;
  double tol, delta2;
  unsigned int counter[8];
  double rho[KMAX], rho_prev[KMAX], rho0[KMAX], alpha[KMAX], omega[KMAX], beta[KMAX];
  double rr[KMAX], rr_ref[KMAX], bb[KMAX], true_rr[KMAX], crit[KMAX];
  int active[KMAX], status[KMAX], want[KMAX];
  long long it[KMAX];
};

__device__ __forceinline__ bool bad(double v) { return !(fabs(v) <= 1.7976931348623157e308); }

// Sliced CSR storage (kernels.hpp): slices of 64 stored rows; perm maps a
// stored position to its row (nullptr: identity).
template <class T> struct CsrView {
  const int *sptr;
  const int *perm;
  const int *col;
  const T *val;
  int64_t rows;
};

// Per-lane block reduction of NP sums into part[(j*grid + blk)*64 + lane];
// k_mfinish folds them (fixed order) in a one-workgroup kernel launched next
// on the stream and calls op.last_lane per RHS, then op.last_all once with
// the counts of active and pause-wanting right-hand sides.
template <int NP, int NB = NT, class Op>
__device__ __forceinline__ void finish(Op &op, double (&acc)[NP > 0 ? NP : 1]) {
  if constexpr (NP > 0) {
    __shared__ double lds[NP][NB / 64][64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int blk = dev::xcd_block();
#pragma unroll
    for (int j = 0; j < NP; ++j) lds[j][w][lane] = acc[j];
    __syncthreads();
    if (w == 0) {
#pragma unroll
      for (int j = 0; j < NP; ++j) {
        double s = 0;
#pragma unroll
        for (int q = 0; q < NB / 64; ++q) s += lds[j][q][lane];
        op.part[((size_t)j * gridDim.x + blk) * 64 + lane] = s;
      }
    }
  }
}

constexpr int NTM = 1024;  // threads of the fold: 16 wavefronts share the blocks

template <class Op>
__global__ __launch_bounds__(NTM) void k_mfinish(Op op0, int nblk) {
  Op op = op0;
  if (op.skip()) return;
  op.begin();
  constexpr int NP = Op::NP, NW = NTM / 64;
  __shared__ double lds[NP][NW][64];
  __shared__ double sums[NP][64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  // thread (w, lane) sums blocks w, w+16, ... in order; then the 16 parts in
  // order.  Up to 32 of a thread's loads are issued before the first add, so
  // a fold over 512 blocks is one round trip per sum instead of four.
#pragma unroll
  for (int j = 0; j < NP; ++j) {
    double s = 0;
    for (int b0 = w; b0 < nblk; b0 += 32 * NW) {
      double v[32];
#pragma unroll
      for (int m = 0; m < 32; ++m) {
        const int b = b0 + m * NW;
        v[m] = b < nblk ? op.part[((size_t)j * nblk + b) * 64 + lane] : 0.0;
      }
#pragma unroll
      for (int m = 0; m < 32; ++m) s += v[m];
    }
    lds[j][w][lane] = s;
  }
  __syncthreads();
  if (w == 0) {
#pragma unroll
    for (int j = 0; j < NP; ++j) {
      double s = 0;
#pragma unroll
      for (int q = 0; q < NW; ++q) s += lds[j][q][lane];
      sums[j][lane] = s;
    }
  }
  __syncthreads();
  if (w == 0 && lane < op.st->k) {
    double v[NP];
#pragma unroll
    for (int j = 0; j < NP; ++j) v[j] = sums[j][lane];
    op.last_lane(lane, v);
  }
  __syncthreads();
  if (w == 0) {
    // the batch's active and pause-wanting columns counted by one wave, a
    // column per lane, for the batch-wide decisions of last_all
    const int kk = op.st->k;
    const bool a = lane < kk && op.st->active[lane] != 0;
    const bool wt = a && op.st->want[lane] != 0;
    const int na = (int)__popcll(__ballot(a)), nw = (int)__popcll(__ballot(wt));
    if (lane == 0) op.last_all(na, nw);
  }
}

// Y-producing kernel: a wavefront takes RB rows of a slice (32 for FP64, whose
// accumulators take twice the registers), lanes = right-hand sides; the
// slice's columns and values are read once per slot, coalesced, and each
// row's gather is a 64-wide read of X (sell_rows_multi).
template <class T, class Op>
__global__ __launch_bounds__(NT) void k_mspmm(CsrView<T> A, Op op0) {
  constexpr int RB = sizeof(T) == 4 ? 64 : 32, H = 64 / RB;
  Op op = op0;
  if (op.skip()) return;
  op.begin();
  constexpr int NP = Op::NP;
  double acc[NP > 0 ? NP : 1];
#pragma unroll
  for (int j = 0; j < (NP > 0 ? NP : 1); ++j) acc[j] = 0;
  const int lane = threadIdx.x & 63;
  const int k = op.st->k;
  const bool on = lane < k;
  op.bind1(on ? lane : 0);
  const int64_t units = ((A.rows + 63) >> 6) * H;
  const int64_t nw = (int64_t)gridDim.x * (NT / 64);
  for (int64_t u = (int64_t)dev::xcd_block() * (NT / 64) + (threadIdx.x >> 6); u < units; u += nw) {
    const int64_t s = u / H;
    const int h = (int)(u - s * H);
    T y[RB];
    const int jl = on ? lane : 0;
    dev::sell_rows_multi<RB, T>(A.sptr, A.col, A.val, (const T *)op.cscale(), s, h, lane, on,
                                [&](int c) { return op.x(c, jl); }, y);
#pragma unroll
    for (int r = 0; r < RB; ++r) {
      const int64_t i = s * 64 + h * RB + r;
      if (i < A.rows && on) op.row(A.perm ? (int64_t)A.perm[i] : i, lane, y[r], acc);
    }
  }
  finish<NP>(op, acc);
}

// FP32 with k a multiple of 4: a wavefront takes 32 rows of a slice, lane
// (rs = lane >> 4, g = lane & 15) right-hand sides 4g .. 4g+3 of rows
// 4i + rs, so each gather is a 16-byte read per lane (sell_rows_multi4).  The
// per-lane sums of four right-hand sides are regrouped by right-hand side in
// LDS before the same per-block partials as finish() writes.
template <int NP, class Op>
__device__ __forceinline__ void finish4(Op &op, double (&acc)[NP][4]) {
  __shared__ double lds[NP][NT / 64][4][64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, g = lane & 15, rs = lane >> 4;
  const int blk = dev::xcd_block();
#pragma unroll
  for (int j = 0; j < NP; ++j)
#pragma unroll
    for (int c = 0; c < 4; ++c) lds[j][w][rs][4 * g + c] = acc[j][c];
  __syncthreads();
  if (w == 0) {
#pragma unroll
    for (int j = 0; j < NP; ++j) {
      double s = 0;
#pragma unroll
      for (int q = 0; q < NT / 64; ++q)
#pragma unroll
        for (int r = 0; r < 4; ++r) s += lds[j][q][r][lane];
      op.part[((size_t)j * gridDim.x + blk) * 64 + lane] = s;
    }
  }
}

template <class Op>
__global__ __launch_bounds__(NT) void k_mspmm4(CsrView<float> A, Op op0) {
  constexpr int RB = 32, H = 64 / RB, NP = Op::NP;
  static_assert(NP > 0, "product ops reduce at least one sum");
  Op op = op0;
  if (op.skip()) return;
  op.begin();
  double acc[NP][4];
#pragma unroll
  for (int j = 0; j < NP; ++j)
#pragma unroll
    for (int c = 0; c < 4; ++c) acc[j][c] = 0;
  const int lane = threadIdx.x & 63, g = lane & 15, rs = lane >> 4;
  const int k4 = op.st->k >> 2;
  const bool on = g < k4;
  const int gl = on ? g : 0;
  op.bind4(gl);
  const int64_t units = ((A.rows + 63) >> 6) * H;
  const int64_t nw = (int64_t)gridDim.x * (NT / 64);
  for (int64_t u = (int64_t)dev::xcd_block() * (NT / 64) + (threadIdx.x >> 6); u < units; u += nw) {
    const int64_t s = u / H;
    const int h = (int)(u - s * H);
    float4 y[RB / 4];
    dev::sell_rows_multi4<RB>(A.sptr, A.col, A.val, (const float *)op.cscale(), s, h, lane, on,
                              [&](int c) { return op.x4(c, gl); }, y);
#pragma unroll
    for (int i = 0; i < RB / 4; ++i) {
      const int64_t r = s * 64 + h * RB + 4 * i + rs;
      if (r < A.rows && on) op.row4(A.perm ? (int64_t)A.perm[r] : r, 4 * gl, y[i], acc);
    }
  }
  finish4<NP>(op, acc);
}

// FP32 products with rows of X staged in LDS (k a multiple of 4).  Rows go
// in groups of up to 64, each grown breadth-first over the pattern from the
// lowest unassigned row, so that a group is a compact patch of the mesh: it
// touches far fewer distinct columns than it has entries (about 130 against
// 1,000 on a Q2 mesh, where 64 consecutive rows of a 1025-node-wide grid
// touch 340), so a workgroup takes one group and one pass of 32 right-hand
// sides at a time.  A group's 64 positions hold its rows longest first (empty
// positions: row -1, no entries):
//  1. stage: the X rows of the group's distinct columns (mx_dcol), 128 bytes
//     each, transformed as the op transforms a gathered row (op.stage4: the
//     column's Jacobi scale, p = z + beta p, ...), every load of a thread
//     issued before its first LDS store;
//  2. rows: lane (rho = lane >> 3, q = lane & 7) takes right-hand sides
//     4q .. 4q+3 of the pass for position rho of a set of 8 consecutive
//     positions (rows of about one length); wave w takes sets w and 7 - w, a
//     long and a short one.  The lane reads its row's values and local
//     indices (mx_lidx) from the group's copy in CSR order and 16 bytes of
//     staged x per entry, so one LDS read feeds four FMAs.  A row is summed
//     in column order.
// A group with more distinct columns than the LDS holds reads x directly.
constexpr int kStageRows = 384;  // distinct columns staged per group: 48 KiB at 32 right-hand sides
constexpr int KC = 32;           // right-hand sides per pass
constexpr int kStageIter = kStageRows / (NT / 8);

struct StagedCsr {
  const int *ptr;  // entry pointer per position (64 per group)
  const float *val;
  const unsigned short *lidx;
  const int *dptr, *dcol;
  const int *grow;  // row of each position, -1 when empty
  const unsigned short *dl;  // each position's own column in its group's list (0 when empty)
  int64_t ngr;      // groups
  int dcap;         // rows of X the launch's LDS holds
};

// The groups a workgroup of the staged products takes: the hardware deals
// workgroups to the 8 XCDs round-robin (block b runs on XCD b % 8), so each
// XCD takes a contiguous eighth of the groups and its blocks walk that range
// side by side.  The groups in flight on one XCD are then neighbours in the
// mesh, and the X rows that consecutive strips of groups share are still in
// that XCD's L2 when the next strip stages them (taking every gridDim.x-th
// group instead gave each XCD strips 8 apart, whose shared rows were fetched
// past L2 once per strip).
struct GroupWalk {
  int64_t g, end, step;
};
__device__ __forceinline__ GroupWalk group_walk(int64_t ngr) {
  const int G = gridDim.x, b = blockIdx.x;
  if ((G & 7) != 0) return GroupWalk{b, ngr, G};
  const int64_t chunk = (ngr + 7) / 8, lo = (int64_t)(b & 7) * chunk;
  return GroupWalk{lo + (b >> 3), lo + chunk < ngr ? lo + chunk : ngr, G >> 3};
}

// the row set of sigma = 2w + m: sets w and 7 - w of wave w
__device__ __forceinline__ int staged_set(int w, int m) { return m == 0 ? w : 7 - w; }

// Per-RHS partials of the staged products.  Lane (rho, q) holds right-hand
// sides 4q .. 4q+3 of each pass (a0: first pass, a1: second); the 8 rows of
// a wave are folded by shuffles (fixed order), then the waves through LDS.
template <int NP, class Op>
__device__ __forceinline__ void finish_staged(Op &op, double (&a0)[NP][4], double (&a1)[NP][4]) {
  __shared__ double lds[NP][NT / 64][64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, q = lane & 7;
  const int blk = dev::xcd_block();
#pragma unroll
  for (int j = 0; j < NP; ++j)
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      double s0 = a0[j][c], s1 = a1[j][c];
#pragma unroll
      for (int o = 8; o < 64; o <<= 1) {
        s0 += __shfl_xor(s0, o, 64);
        s1 += __shfl_xor(s1, o, 64);
      }
      if (lane < 8) {
        lds[j][w][4 * q + c] = s0;
        lds[j][w][32 + 4 * q + c] = s1;
      }
    }
  __syncthreads();
  if (w == 0) {
#pragma unroll
    for (int j = 0; j < NP; ++j) {
      double s = 0;
#pragma unroll
      for (int ww = 0; ww < NT / 64; ++ww) s += lds[j][ww][lane];
      op.part[((size_t)j * gridDim.x + blk) * 64 + lane] = s;
    }
  }
}

constexpr int kStageEnt = 2048;  // entries of a group staged in LDS

template <class Op>
__global__ __launch_bounds__(NT) void k_mspmm_lds(StagedCsr A, Op op0) {
  extern __shared__ float4 xs4[];  // A.dcap rows of KC / 4
  __shared__ __align__(16) float ev[kStageEnt];
  __shared__ __align__(16) unsigned short el[kStageEnt];
  __shared__ int ep[65];
  constexpr int NP = Op::NP;
  Op op = op0;
  if (op.skip()) return;
  op.begin();
  constexpr int NPA = NP > 0 ? NP : 1;
  double a0[NPA][4], a1[NPA][4];
#pragma unroll
  for (int j = 0; j < NPA; ++j)
#pragma unroll
    for (int c = 0; c < 4; ++c) a0[j][c] = a1[j][c] = 0;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, q = lane & 7, rho = lane >> 3;
  const int k = op.k;
  const int nh = (k + KC - 1) / KC;
  const GroupWalk gw = group_walk(A.ngr);
  for (int64_t g = gw.g; g < gw.end; g += gw.step) {
    const int d0 = A.dptr[g], d = A.dptr[g + 1] - d0;
    const int64_t r0 = g * 64;  // first position of the group
    const int e0 = A.ptr[r0], ne = A.ptr[r0 + 64] - e0;
    const bool staged = d <= A.dcap && ne <= kStageEnt;
    if (staged) {
      __syncthreads();  // every wave is done with the previous group
      if (tid < 65) ep[tid] = A.ptr[r0 + tid] - e0;
      float vv[kStageEnt / NT];
      unsigned short ll[kStageEnt / NT];
#pragma unroll
      for (int m = 0; m < kStageEnt / NT; ++m) {
        const int i = tid + m * NT;
        vv[m] = i < ne ? __builtin_nontemporal_load(&A.val[e0 + i]) : 0.f;
        ll[m] = i < ne ? __builtin_nontemporal_load(&A.lidx[e0 + i]) : (unsigned short)0;
      }
#pragma unroll
      for (int m = 0; m < kStageEnt / NT; ++m) {
        const int i = tid + m * NT;
        if (i < ne) {
          ev[i] = vv[m];
          el[i] = ll[m];
        }
      }
    }
    for (int h = 0; h < nh; ++h) {
      const int j0 = h * KC, jq = j0 / 4 + q;  // this lane's quad of right-hand sides
      const bool on = 4 * jq < k;
      op.bind4(on ? jq : j0 / 4);
      op.bind4s(on ? jq : j0 / 4);
      if (staged) {
        if (h > 0) __syncthreads();  // every wave is done with the previous pass's x
        if (on) {
          // thread (i0 = tid >> 3, q) stages quad q of rows i0, i0 + 32, ...
          const int i0 = tid >> 3;
          int c[kStageIter];
          float4 v[kStageIter];
#pragma unroll
          for (int m = 0; m < kStageIter; ++m) {
            const int i = i0 + m * (NT / 8);
            c[m] = i < d ? A.dcol[d0 + i] : -1;
          }
#pragma unroll
          for (int m = 0; m < kStageIter; ++m)
            if (c[m] >= 0) v[m] = op.stage4(c[m], jq);
#pragma unroll
          for (int m = 0; m < kStageIter; ++m)
            if (c[m] >= 0) xs4[(i0 + m * (NT / 8)) * (KC / 4) + q] = v[m];
        }
        __syncthreads();
#pragma unroll 1
        for (int m = 0; m < 2; ++m) {
          const int rl = staged_set(w, m) * 8 + rho;  // position within the group
          const int64_t r = A.grow[r0 + rl];
          // rows are padded to a multiple of 4 entries (zero values), so a
          // lane reads its row four entries at a time: 16 bytes of values
          // and 8 bytes of positions
          const int b = ep[rl], n4 = r >= 0 ? (ep[rl + 1] - b) >> 2 : 0;
          const Pre4 pr = op.pre4(r >= 0 ? r : 0, 4 * (on ? jq : j0 / 4));
          float4 y = make_float4(0.f, 0.f, 0.f, 0.f);
          const float4 *ev4 = reinterpret_cast<const float4 *>(ev + b);
          const uint2 *el4 = reinterpret_cast<const uint2 *>(el + b);
          for (int t = 0; t < n4; ++t) {
            const float4 v4 = ev4[t];
            const uint2 l4 = el4[t];
            const float vv[4] = {v4.x, v4.y, v4.z, v4.w};
            const int li[4] = {(int)(l4.x & 0xffffu), (int)(l4.x >> 16), (int)(l4.y & 0xffffu), (int)(l4.y >> 16)};
#pragma unroll
            for (int e = 0; e < 4; ++e) {
              const float4 x = xs4[li[e] * (KC / 4) + q];
              y.x = fmaf(vv[e], x.x, y.x);
              y.y = fmaf(vv[e], x.y, y.y);
              y.z = fmaf(vv[e], x.z, y.z);
              y.w = fmaf(vv[e], x.w, y.w);
            }
          }
          if (r >= 0 && on) {
            if (h == 0) op.row4p(r, 4 * jq, y, pr, a0);
            else op.row4p(r, 4 * jq, y, pr, a1);
          }
        }
      } else {
        // direct reads of x for groups too large to stage
#pragma unroll 1
        for (int m = 0; m < 2; ++m) {
          const int64_t pos = r0 + staged_set(w, m) * 8 + rho;
          const int64_t r = A.grow[pos];
          float4 y = make_float4(0.f, 0.f, 0.f, 0.f);
          if (r >= 0) {
            const int b = A.ptr[pos], n = A.ptr[pos + 1] - b;
            for (int t = 0; t < n; ++t) {
              const float vv = A.val[b + t];
              const float4 x = op.stage4(A.dcol[d0 + A.lidx[b + t]], on ? jq : j0 / 4);
              y.x = fmaf(vv, x.x, y.x);
              y.y = fmaf(vv, x.y, y.y);
              y.z = fmaf(vv, x.z, y.z);
              y.w = fmaf(vv, x.w, y.w);
            }
          }
          if (r >= 0 && on) {
            if (h == 0) op.row4p(r, 4 * jq, y, op.pre4(r, 4 * jq), a0);
            else op.row4p(r, 4 * jq, y, op.pre4(r, 4 * jq), a1);
          }
        }
      }
    }
  }
  if constexpr (NP > 0) finish_staged<NP>(op, a0, a1);
}

// The staged product software-pipelined, for matrices whose every group fits
// the LDS budgets: while the waves compute one unit (a group and a pass of 32
// right-hand sides) from LDS, every thread's loads of the next unit's rows of
// X -- and at a group's second pass, of the next group's entries and list --
// are already in flight into registers, and are stored to LDS after the next
// barrier.  Same rows, sums and order as k_mspmm_lds.
// KI staging loads per thread and unit (lists of up to 32 KI rows).  With
// KI = 6 and one 16-byte register per staged quad the kernel fits three waves
// per SIMD (at most 168 registers), four when the op keeps no partial sums
// (the plain product), else two.  Residency is what bounds these kernels:
// each unit waits on its staging loads and on LDS round trips, and more
// resident workgroups overlap those waits.
template <class Op, int KI>
constexpr int pipe_waves() {
  return KI <= 6 && sizeof(typename Op::Raw) <= 16 ? (Op::NP == 0 ? 4 : 3) : 2;
}

template <class Op, int KI>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(pipe_waves<Op, KI>(), pipe_waves<Op, KI>())))
void k_mspmm_pipe(StagedCsr A, Op op0) {
  extern __shared__ float4 xs4[];  // A.dcap rows of KC / 4
  __shared__ __align__(16) float ev[kStageEnt];
  __shared__ __align__(16) unsigned short el[kStageEnt];
  __shared__ int ep[65];
  __shared__ int grw[64];  // row of each position of the group
  __shared__ int gdl[64];  // its own column's place in the group's list
  constexpr int NP = Op::NP, SE = kStageEnt / NT;
  Op op = op0;
  if (op.skip()) return;
  op.begin();
  constexpr int NPA = NP > 0 ? NP : 1;
  double a0[NPA][4], a1[NPA][4];
#pragma unroll
  for (int j = 0; j < NPA; ++j)
#pragma unroll
    for (int c = 0; c < 4; ++c) a0[j][c] = a1[j][c] = 0;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, q = lane & 7, rho = lane >> 3, i0 = tid >> 3;
  const int k = op.k;
  const int nh = (k + KC - 1) / KC;
  const GroupWalk gw = group_walk(A.ngr);
  const int64_t ngr = gw.end;  // this workgroup's groups end here
  int64_t g = gw.g;
  if (g >= ngr) {
    if constexpr (NP > 0) finish_staged<NP>(op, a0, a1);
    return;
  }
  // registers of the next unit.  Every load is unconditional, at a clamped
  // in-bounds position, and nothing computes on its result before the store
  // to LDS: a conditional load, or arithmetic on its result, makes the
  // compiler wait for it at once instead of during the rows.
  using Raw = typename Op::Raw;
  int pc[KI];
  Raw px[KI];
  float pv[SE];
  unsigned short pl[SE];
  int pp = 0, prow = 0, pdl = 0, d0 = 0, d = 0, ne = 0;
  auto load_meta = [&](int64_t gg) {  // the group's list, entries and rows
    d0 = A.dptr[gg];
    d = A.dptr[gg + 1] - d0;
    const int64_t r0 = gg * 64;
    const int e0 = A.ptr[r0];
    ne = A.ptr[r0 + 64] - e0;
    pp = A.ptr[r0 + (tid < 64 ? tid : 64)] - e0;  // used by threads 0..64
    prow = A.grow[r0 + (tid & 63)];                // used by threads 0..63
    if constexpr (Op::kOwnStaged) pdl = A.dl[r0 + (tid & 63)];
#pragma unroll
    for (int m = 0; m < SE; ++m) {
      const int i = tid + m * NT;
      const int64_t e = ne > 0 ? e0 + (i < ne ? i : ne - 1) : 0;
      pv[m] = __builtin_nontemporal_load(&A.val[e]);
      pl[m] = __builtin_nontemporal_load(&A.lidx[e]);
    }
#pragma unroll
    for (int m = 0; m < KI; ++m) {
      const int i = i0 + m * (NT / 8);
      pc[m] = A.dcol[d > 0 ? d0 + (i < d ? i : 0) : 0];
    }
  };
  auto load_x = [&](int h) {  // quad q of the pass's rows i0, i0 + 32, ... (raw)
    const int jq = h * (KC / 4) + q;
    const int jl = 4 * jq < k ? jq : h * (KC / 4);
#pragma unroll
    for (int m = 0; m < KI; ++m) px[m] = op.load4(pc[m], jl);
  };
  load_meta(g);
  load_x(0);
  // the rows' op for each pass, bound once: the active flags and betas of a
  // lane's four right-hand sides do not change during the launch
  Op opb[2] = {op, op};
#pragma unroll
  for (int hh = 0; hh < 2; ++hh) {
    const int jq = hh * (KC / 4) + q;
    opb[hh].bind4(4 * jq < k ? jq : 0);
  }
  int h = 0;
  while (true) {
    __syncthreads();  // every wave is done with the LDS of the previous unit
    // every prefetched value is stored below: wait for all of them here, in
    // one place, so that no load is still counted as pending after the
    // (partly conditional) stores -- a later wait for the epilogue operands
    // would otherwise wait again (vmcnt(0); expcnt and lgkmcnt not waited)
    __builtin_amdgcn_s_waitcnt(0x0f70);
    if (h == 0) {
      if (tid < 65) ep[tid] = pp;
      if (tid < 64) grw[tid] = prow;
      if constexpr (Op::kOwnStaged)
        if (tid < 64) gdl[tid] = pdl;
#pragma unroll
      for (int m = 0; m < SE; ++m) {
        const bool in = tid + m * NT < ne;
        ev[tid + m * NT] = in ? pv[m] : 0.f;
        el[tid + m * NT] = in ? pl[m] : (unsigned short)0;
      }
    }
    {
      const int jq = h * (KC / 4) + q;
      const bool on = 4 * jq < k;
      op.bind4s(on ? jq : h * (KC / 4));
#pragma unroll
      for (int m = 0; m < KI; ++m)
        if (i0 + m * (NT / 8) < d && on) xs4[(i0 + m * (NT / 8)) * (KC / 4) + q] = op.make4(px[m]);
    }
    __syncthreads();
    const int j0 = h * KC, jq = j0 / 4 + q;
    const bool on = 4 * jq < k;
    // the two row sets of the wave (rows a, a + 2, ..., a + 14) and their
    // epilogue operands, loaded before the next unit's loads: vector loads
    // complete in issue order, so a wait for these must not wait for those
    int rls[2], rws[2];
    Pre4 prs[2];
#pragma unroll
    for (int m = 0; m < 2; ++m) {
      rls[m] = staged_set(w, m) * 8 + rho;
      rws[m] = grw[rls[m]];
      if constexpr (Op::kOwnStaged) {
        // the row's own column is staged: its operand from LDS, no load
        prs[m] = Op::own4(xs4[gdl[rls[m]] * (KC / 4) + q]);
      } else {
        prs[m] = op.pre4(rws[m] >= 0 ? rws[m] : 0, 4 * (on ? jq : j0 / 4));
      }
    }
    // the next unit's loads, in flight during this unit's rows
    const int64_t gn = g + gw.step;
    const bool second = h + 1 < nh;
    if (second) load_x(h + 1);
    else if (gn < ngr) load_meta(gn);
#pragma unroll
    for (int m = 0; m < 2; ++m) {
      const int rl = rls[m];
      const int64_t r = rws[m];
      const int b = ep[rl], n4 = r >= 0 ? (ep[rl + 1] - b) >> 2 : 0;
      float4 y = make_float4(0.f, 0.f, 0.f, 0.f);
      const float4 *ev4 = reinterpret_cast<const float4 *>(ev + b);
      const uint2 *el4 = reinterpret_cast<const uint2 *>(el + b);
      for (int t = 0; t < n4; ++t) {
        const float4 v4 = ev4[t];
        const uint2 l4 = el4[t];
        const float vv[4] = {v4.x, v4.y, v4.z, v4.w};
        const int li[4] = {(int)(l4.x & 0xffffu), (int)(l4.x >> 16), (int)(l4.y & 0xffffu), (int)(l4.y >> 16)};
        float4 xq[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) xq[e] = xs4[li[e] * (KC / 4) + q];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          y.x = fmaf(vv[e], xq[e].x, y.x);
          y.y = fmaf(vv[e], xq[e].y, y.y);
          y.z = fmaf(vv[e], xq[e].z, y.z);
          y.w = fmaf(vv[e], xq[e].w, y.w);
        }
        // the four staged x reads issued together, then the FMAs: one LDS
        // round trip per four entries instead of one per entry
        __builtin_amdgcn_sched_group_barrier(0x100, 2, 0);  // entries (LDS read)
        __builtin_amdgcn_sched_group_barrier(0x002, 8, 0);  // positions -> addresses
        __builtin_amdgcn_sched_group_barrier(0x100, 4, 0);  // staged x (LDS read)
        __builtin_amdgcn_sched_group_barrier(0x002, 16, 0); // FMAs
      }
      if (r >= 0 && on) {
        if (h == 0) opb[0].row4p(r, 4 * jq, y, prs[m], a0);
        else opb[1].row4p(r, 4 * jq, y, prs[m], a1);
      }
    }
    if (second) {
      ++h;
    } else {
      if (gn >= ngr) break;
      g = gn;
      h = 0;
      load_x(0);  // its list arrived during the rows above
    }
  }
  if constexpr (NP > 0) finish_staged<NP>(op, a0, a1);
}

// FP64 products with rows of X staged in LDS: the row groups, lists and
// padded CSR-order entries of the FP32 staged products, with FP64 values (the
// CSR-order copy of the FP64 matrix, or of A D^-1 for ops whose product
// scales the columns).  A pass stages 32 right-hand sides of the group's
// distinct columns, 256 bytes per row, through the op's per-right-hand-side
// gather (op.x, the same transform the sliced FP64 product applies to a
// gathered value).  Lane (rho = lane >> 4, q = lane & 15) takes right-hand
// sides 2q and 2q + 1 of the pass for position rho of a set of four
// consecutive positions, so the 16 lanes of a position read one whole staged
// row per entry: each 16-byte LDS read is free of bank conflicts.  Eight
// waves share a group (the LDS, not the registers, caps the workgroups per
// CU); wave w takes sets w and 15 - w (positions are longest first).  A row
// is summed entry by entry in column order with one FMA per entry, as the
// sliced FP64 product sums it, so the products are the same bit for bit; the
// row epilogue is the op's per-right-hand-side row().  Every group must fit
// the launch's LDS (staged64_ok).
constexpr int KC64 = 32;         // right-hand sides per pass
constexpr int NT64 = 512;        // threads per workgroup: 8 waves share a group's LDS
constexpr int kStageRows64 = 192;   // staged rows: 48 KiB
constexpr int kStageEnt64 = 1536;   // staged entries: 15 KiB (values and positions)

struct StagedCsr64 {
  const int *ptr;
  const double *val;
  const unsigned short *lidx;
  const int *dptr, *dcol;
  const int *grow;
  int64_t ngr;
  int dcap, ecap;  // rows of X and entries the launch's LDS holds
  int pair;        // ops with row2: a lane's two right-hand sides stored as one 16-byte pair
};

__device__ __forceinline__ int staged_set64(int w, int m) { return m == 0 ? w : 15 - w; }

// Per-RHS partials: lane (rho, q) holds right-hand sides 32h + 2q + c in
// acc[h][c]; the four positions of a wave are folded by shuffles (fixed
// order), then the 8 waves through LDS (the staged rows' space, free by now).
template <int NP, class Op>
__device__ __forceinline__ void finish_staged64(Op &op, double (&acc)[2][2][NP], double *lds) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, q = lane & 15;
  const int blk = dev::xcd_block();
  __syncthreads();
#pragma unroll
  for (int j = 0; j < NP; ++j)
#pragma unroll
    for (int h = 0; h < 2; ++h)
#pragma unroll
      for (int c = 0; c < 2; ++c) {
        double s = acc[h][c][j];
        s += __shfl_xor(s, 16, 64);
        s += __shfl_xor(s, 32, 64);
        if (lane < 16) lds[(j * (NT64 / 64) + w) * 64 + 32 * h + 2 * q + c] = s;
      }
  __syncthreads();
  if (w == 0) {
#pragma unroll
    for (int j = 0; j < NP; ++j) {
      double s = 0;
#pragma unroll
      for (int ww = 0; ww < NT64 / 64; ++ww) s += lds[(j * (NT64 / 64) + ww) * 64 + lane];
      op.part[((size_t)j * gridDim.x + blk) * 64 + lane] = s;
    }
  }
}

// Six waves per SIMD (at most 80 registers) for ops with up to two sums:
// three workgroups per CU, as many as the LDS of c5-sized groups holds.
template <class Op>
constexpr int lds64_waves() {
  return Op::NP <= 2 ? 6 : 4;
}

template <class Op>
__global__ __launch_bounds__(NT64) __attribute__((amdgpu_waves_per_eu(lds64_waves<Op>(), lds64_waves<Op>())))
void k_mspmm_lds64(StagedCsr64 A, Op op0) {
  // A.dcap rows of KC64 / 2 double2, then A.ecap values, then A.ecap positions
  extern __shared__ double2 xs2[];
  __shared__ int ep[65];
  __shared__ int grw[64];  // row of each position of the group
  double *ev = reinterpret_cast<double *>(xs2 + (size_t)A.dcap * (KC64 / 2));
  unsigned short *el = reinterpret_cast<unsigned short *>(ev + A.ecap);
  constexpr int NP = Op::NP, NPA = NP > 0 ? NP : 1;
  Op op = op0;
  if (op.skip()) return;
  op.begin();
  double acc[2][2][NPA];
#pragma unroll
  for (int h = 0; h < 2; ++h)
#pragma unroll
    for (int c = 0; c < 2; ++c)
#pragma unroll
      for (int j = 0; j < NPA; ++j) acc[h][c][j] = 0;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, q = lane & 15, rho = lane >> 4, i0 = tid >> 4;
  const int k = op.k;
  const int nh = (k + KC64 - 1) / KC64;
  const GroupWalk gw = group_walk(A.ngr);
  for (int64_t g = gw.g; g < gw.end; g += gw.step) {
    const int d0 = A.dptr[g], d = A.dptr[g + 1] - d0;
    const int64_t r0 = g * 64;
    const int e0 = A.ptr[r0], ne = A.ptr[r0 + 64] - e0;  // every group fits: d <= dcap, ne <= ecap
    // the group's entry pointers and rows, loaded before the barrier
    const int pp = A.ptr[r0 + (tid < 64 ? tid : 64)] - e0;  // threads 0..64
    const int prow = A.grow[r0 + (tid & 63)];               // threads 0..63
    __syncthreads();  // every wave is done with the previous group's LDS
    if (tid < 65) ep[tid] = pp;
    if (tid < 64) grw[tid] = prow;
    for (int i = tid; i < ne; i += NT64) {
      ev[i] = __builtin_nontemporal_load(&A.val[e0 + i]);
      el[i] = __builtin_nontemporal_load(&A.lidx[e0 + i]);
    }
#pragma unroll 1
    for (int h = 0; h < nh; ++h) {
      const int jb = h * KC64 + 2 * q;  // this lane's first right-hand side of the pass
      Op oc[2] = {op, op};
#pragma unroll
      for (int c = 0; c < 2; ++c) oc[c].bind1(jb + c < k ? jb + c : 0);
      if (h > 0) __syncthreads();  // every wave is done with the previous pass's rows
      // thread (i0, q) stages pair q of rows i0, i0 + 32, ...
#pragma unroll 4
      for (int i = i0; i < d; i += NT64 / 16) {
        const int c = A.dcol[d0 + i];
        xs2[i * (KC64 / 2) + q] = make_double2(jb < k ? oc[0].x(c, jb) : 0.0, jb + 1 < k ? oc[1].x(c, jb + 1) : 0.0);
      }
      __syncthreads();  // the group's entries (first pass) and this pass's rows are in LDS
#pragma unroll 1
      for (int m = 0; m < 2; ++m) {
        const int p = staged_set64(w, m) * 4 + rho;
        const int64_t r = grw[p];
        const int b = ep[p], n4 = r >= 0 ? (ep[p + 1] - b) >> 2 : 0;
        double y0 = 0, y1 = 0;
        const double2 *ev2 = reinterpret_cast<const double2 *>(ev + b);
        const uint2 *el4 = reinterpret_cast<const uint2 *>(el + b);
        for (int t = 0; t < n4; ++t) {
          const double2 va = ev2[2 * t], vb = ev2[2 * t + 1];
          const uint2 l4 = el4[t];
          const double vv[4] = {va.x, va.y, vb.x, vb.y};
          const int li[4] = {(int)(l4.x & 0xffffu), (int)(l4.x >> 16), (int)(l4.y & 0xffffu), (int)(l4.y >> 16)};
          double2 xq[4];
#pragma unroll
          for (int e = 0; e < 4; ++e) xq[e] = xs2[li[e] * (KC64 / 2) + q];
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            y0 = fma(vv[e], xq[e].x, y0);
            y1 = fma(vv[e], xq[e].y, y1);
          }
        }
        // constant indexes keep the sums in registers (h > 1, k > 64: plain
        // products only, which keep none)
        bool done = false;
        if constexpr (Op::kRow2) {
          if (r >= 0 && A.pair && jb + 1 < k) {
            if (h == 0) oc[0].row2(r, jb, y0, y1, oc[1], acc[0][0], acc[0][1]);
            else oc[0].row2(r, jb, y0, y1, oc[1], acc[1][0], acc[1][1]);
            done = true;
          }
        }
        if (r >= 0 && !done && h == 0) {
          if (jb < k) oc[0].row(r, jb, y0, acc[0][0]);
          if (jb + 1 < k) oc[1].row(r, jb + 1, y1, acc[0][1]);
        } else if (r >= 0 && !done) {
          if (jb < k) oc[0].row(r, jb, y0, acc[1][0]);
          if (jb + 1 < k) oc[1].row(r, jb + 1, y1, acc[1][1]);
        }
      }
    }
  }
  if constexpr (NP > 0) finish_staged64<NP>(op, acc, reinterpret_cast<double *>(xs2));
}

// Vector passes with a lane per right-hand side: workgroups of 1024 threads,
// two per CU, so that the one-workgroup fold after the pass reads a quarter
// of the partials that 256-thread workgroups at eight per CU would leave.
constexpr int NTE = 1024;  // threads of k_melem

template <class Op>
__global__ __launch_bounds__(NTE) void k_melem(int64_t n, Op op0) {
  Op op = op0;
  if (op.skip()) return;
  op.begin();
  constexpr int NP = Op::NP;
  double acc[NP > 0 ? NP : 1];
#pragma unroll
  for (int j = 0; j < (NP > 0 ? NP : 1); ++j) acc[j] = 0;
  const int lane = threadIdx.x & 63;
  const int k = op.st->k;
  const int64_t chunk = (n + gridDim.x - 1) / gridDim.x;
  const int64_t r0 = (int64_t)dev::xcd_block() * chunk;
  const int64_t r1 = r0 + chunk < n ? r0 + chunk : n;
  for (int64_t r = r0 + (threadIdx.x >> 6); r < r1; r += NTE / 64)
    if (lane < k) op.elem(r, lane, acc);
  finish<NP, NTE>(op, acc);
}

// FP32 vector passes with k a multiple of 4.  Lane (rs = lane >> 4,
// g = lane & 15) takes right-hand sides 4g .. 4g+3 of rows 4i + rs, so every
// load and store is 16 bytes per lane.  A workgroup owns `per` consecutive
// units of 64 rows and walks them ME4 rows per wave at a time: every load of
// those rows (op.load4) is issued before the first store (op.apply4), which
// keeps 16 rows' operands in flight per wave where a row-at-a-time loop had
// one.  The lane's active flags and step scalars are read once (op.bind4):
// they change only in the fold after the pass.  Workgroups of 1024 threads,
// two per CU, so that the fold after the pass reads a quarter of the
// partials that 256-thread workgroups would leave.
constexpr int NT4 = 1024;   // threads of k_melem4

template <class Op>
__global__ __launch_bounds__(NT4) void k_melem4(int64_t n, int64_t per, Op op0) {
  // ME4 row sets of 64 rows (16 waves x 4 rows) per step: 4, or 2 for the
  // passes with five operands per row, within 128 registers
  constexpr int NP = Op::NP, NPA = NP > 0 ? NP : 1, RS = (NT4 / 64) * 4, ME4 = Op::ME4;
  Op op = op0;
  if (op.skip()) return;
  op.begin();
  double acc[NPA][4];
#pragma unroll
  for (int j = 0; j < NPA; ++j)
#pragma unroll
    for (int c = 0; c < 4; ++c) acc[j][c] = 0;
  const int lane = threadIdx.x & 63, g = lane & 15, rs = lane >> 4, w = threadIdx.x >> 6;
  const bool on = g < (op.k >> 2);
  const int gl = on ? g : 0;
  op.bind4(gl);
  const int64_t r0 = (int64_t)dev::xcd_block() * per * 64;
  const int64_t r1 = r0 + per * 64 < n ? r0 + per * 64 : n;
  for (int64_t rb = r0 + w * 4 + rs; rb < r1; rb += RS * ME4) {
    typename Op::In in[ME4];
#pragma unroll
    for (int u = 0; u < ME4; ++u) {
      const int64_t r = rb + u * RS < r1 ? rb + u * RS : rb;  // in bounds; not applied
      in[u] = op.load4(r, r * op.k + 4 * gl);
    }
#pragma unroll
    for (int u = 0; u < ME4; ++u) {
      const int64_t r = rb + u * RS;
      if (r < r1 && on) op.apply4(r, r * op.k + 4 * gl, in[u], acc);
    }
  }
  if constexpr (NP > 0) {
    // the four rows of a wave folded by shuffles (fixed order), then the
    // waves through LDS, into the same per-block partials as finish()
    __shared__ double lds[NP][NT4 / 64][64];
#pragma unroll
    for (int j = 0; j < NP; ++j)
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        double v = acc[j][c];
        v += __shfl_xor(v, 16, 64);
        v += __shfl_xor(v, 32, 64);
        if (rs == 0) lds[j][w][4 * g + c] = v;
      }
    __syncthreads();
    if (w == 0) {
      const int blk = dev::xcd_block();
#pragma unroll
      for (int j = 0; j < NP; ++j) {
        double v = 0;
#pragma unroll
        for (int q = 0; q < NT4 / 64; ++q) v += lds[j][q][lane];
        op.part[((size_t)j * gridDim.x + blk) * 64 + lane] = v;
      }
    }
  }
}

struct MBase {
  static constexpr bool V4 = false;           // has load4 / apply4 for k_melem4
  static constexpr int ME4 = 4;               // k_melem4 row sets per step
  static constexpr bool kScaledVals = false;  // staged products read A D^-1
  // the row epilogue's operand is the row's own staged value (own4): the
  // pipelined product reads it from LDS instead of pre4's global loads
  static constexpr bool kOwnStaged = false;
  MState *st;
  double *part;
  unsigned int *counter;
  int k;
  __device__ bool running() const { return !st->done && !st->paused; }
  __device__ void last_all(int, int) const {}
  // product ops: no column scale unless the op supplies one
  __device__ const void *cscale() const { return nullptr; }
  // the four right-hand sides 4g .. 4g+3 of one lane (FP32 products): their
  // active flags, read once (they change only in last_lane, after every row
  // of the launch); and of the rows one thread stages (staged products)
  bool a4c[4];
  __device__ void bind4(int g) {
#pragma unroll
    for (int c = 0; c < 4; ++c) a4c[c] = st->active[4 * g + c] != 0;
  }
  __device__ void bind4s(int) {}
  // the lane's right-hand side in k_mspmm: its active flag read once (it
  // changes only in last_lane, after every row of the launch).  Only the FP32
  // CG product uses it: there the compiler then keeps 64 rows' gathers in
  // flight and the k = 63 mixed batch on 3-D elasticity 32^3 ran 3.3x faster,
  // while the FP64 products it was tried on came out 15-40 % slower.
  bool act1 = true;
  __device__ void bind1(int j) { act1 = st->active[j] != 0; }
  // FP64 staged products: the row epilogue of right-hand sides j, j + 1 as
  // one 16-byte pair (row2; o1 is the op bound to j + 1), where the op has one
  static constexpr bool kRow2 = false;
  template <class T>
  __device__ static void store2nt(T *dst, int64_t o, T y0, T y1, bool a0, bool a1) {
    if constexpr (sizeof(T) == 8) {
      typedef double v2d __attribute__((ext_vector_type(2)));
      if (a0 && a1) {
        const v2d w = {y0, y1};
        __builtin_nontemporal_store(w, reinterpret_cast<v2d *>(dst + o));
        return;
      }
    }
    if (a0) __builtin_nontemporal_store(y0, dst + o);
    if (a1) __builtin_nontemporal_store(y1, dst + o + 1);
  }
  template <class T>
  __device__ static void store4(T *dst, int64_t o, const float4 &v, const bool (&a)[4]) {
    if (a[0] && a[1] && a[2] && a[3]) {
      *reinterpret_cast<float4 *>(dst + o) = v;
    } else {
      if (a[0]) dst[o] = v.x;
      if (a[1]) dst[o + 1] = v.y;
      if (a[2]) dst[o + 2] = v.z;
      if (a[3]) dst[o + 3] = v.w;
    }
  }
  // The products' n x k outputs, written once and read by the next kernel
  // (hundreds of MB: never from L2): non-temporal, so that they do not evict
  // the staged rows of X that the next strip of row groups re-reads from L2.
  template <class T>
  __device__ static void store4nt(T *dst, int64_t o, const float4 &v, const bool (&a)[4]) {
    if constexpr (sizeof(T) == 4) {
      typedef float v4f __attribute__((ext_vector_type(4)));
      if (a[0] && a[1] && a[2] && a[3]) {
        const v4f w = {v.x, v.y, v.z, v.w};
        __builtin_nontemporal_store(w, reinterpret_cast<v4f *>(dst + o));
        return;
      }
    }
    store4(dst, o, v, a);
  }
};

__device__ __forceinline__ float4 ld4(const float *p, int64_t o) { return *reinterpret_cast<const float4 *>(p + o); }
__device__ __forceinline__ float comp(const float4 &v, int c) { return c == 0 ? v.x : c == 1 ? v.y : c == 2 ? v.z : v.w; }


__device__ void request_pause_or_advance(MState *st, int j, double rho_new, double rr, bool cgfam) {
  st->it[j] += 1;
  const double crit = st->bb[j] > 0 ? sqrt(rr / st->bb[j]) : 0.0;
  st->crit[j] = crit;
  st->rr[j] = rr;
  if (bad(rho_new) || bad(rr)) {
    st->status[j] = YAFEL_HIP_STATUS_NAN;
    st->active[j] = 0;
    return;
  }
  if (crit < st->tol || rr < st->delta2 * st->delta2 * st->rr_ref[j] || st->it[j] >= st->max_iter) {
    st->paused = 1;  // plain store from several lanes, same value
    st->rho_prev[j] = st->rho[j];
    return;
  }
  if (rr < st->delta2 * st->rr_ref[j]) st->want[j] = 1;
  if (cgfam) {
    st->beta[j] = rho_new / st->rho[j];
  } else {
    st->beta[j] = (rho_new / st->rho[j]) * (st->alpha[j] / st->omega[j]);
  }
  st->rho_prev[j] = st->rho[j];
  st->rho[j] = rho_new;
}

// after an iteration: pause for the shared residual replacement once half of
// the na active right-hand sides (nw of them wanting one) want one
__device__ void pause_when_half_want(MState *st, int na, int nw) {
  if (nw > 0 && 2 * nw >= na) st->paused = 1;
}

__device__ void all_inactive_done(MState *st, int na) {
  st->nact = na;
  if (!na) st->done = 1;
}

// ---- BiCGSTAB (right Jacobi): p = r + beta (p - omega v)
template <class T>
struct MP : MBase {
  static constexpr int NP = 0;
  const T *r, *v;
  T *p;
  __device__ bool skip() const { return !running(); }
  __device__ void begin() {}
  __device__ void elem(int64_t i, int j, double *) const {
    if (!st->active[j]) return;
    const int64_t o = i * k + j;
    const T beta = st->first ? T(0) : (T)st->beta[j], om = (T)st->omega[j];
    p[o] = r[o] + beta * (p[o] - om * v[o]);
  }
  static constexpr bool V4 = sizeof(T) == 4;
  float be4[4], om4[4];
  __device__ void bind4(int g) {
    MBase::bind4(g);
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      be4[c] = st->first ? 0.f : (float)st->beta[4 * g + c];
      om4[c] = (float)st->omega[4 * g + c];
    }
  }
  struct In {
    float4 r, p, v;
  };
  __device__ In load4(int64_t, int64_t o) const {
    return In{ld4((const float *)r, o), ld4((const float *)p, o), ld4((const float *)v, o)};
  }
  __device__ void apply4(int64_t, int64_t o, const In &a, double (&)[1][4]) const {
    const float4 pn = make_float4(a.r.x + be4[0] * (a.p.x - om4[0] * a.v.x), a.r.y + be4[1] * (a.p.y - om4[1] * a.v.y),
                                  a.r.z + be4[2] * (a.p.z - om4[2] * a.v.z), a.r.w + be4[3] * (a.p.w - om4[3] * a.v.w));
    store4((float *)p, o, pn, a4c);
  }
  __device__ void last_lane(int, const double *) const {}
};

// v = A (D^-1 p), rhat.v -> alpha; D^-1 scales the matrix columns
template <class T>
struct MV : MBase {
  static constexpr int NP = 1;
  const T *p, *rhat, *dinv;
  T *v;
  __device__ bool skip() const { return !running(); }
  __device__ void begin() {}
  __device__ const T *cscale() const { return dinv; }
  __device__ T x(int c, int j) const { return p[(int64_t)c * k + j]; }
  // the staged products read A D^-1 (the values scaled by their column's
  // D^-1 at set-up), so the staged rows of p are p itself
  static constexpr bool kScaledVals = true;
  using Raw = float4;
  __device__ Raw load4(int c, int g) const { return ld4((const float *)p, (int64_t)c * k + 4 * g); }
  __device__ static float4 make4(const Raw &r) { return r; }
  __device__ float4 stage4(int c, int g) const { return load4(c, g); }
  __device__ void row(int64_t i, int j, T y, double *acc) const {
    if (!st->active[j]) return;
    v[i * k + j] = y;
    acc[0] += (double)rhat[i * k + j] * (double)y;
  }
  static constexpr bool kRow2 = sizeof(T) == 8;
  __host__ bool pair_ok() const { return ((uintptr_t)v | (uintptr_t)rhat) % 16 == 0; }
  __device__ void row2(int64_t i, int j, T y0, T y1, const MV &o1, double *acc0, double *acc1) const {
    const int64_t o = i * k + j;
    store2nt(v, o, y0, y1, act1, o1.act1);
    const double2 rh = *reinterpret_cast<const double2 *>(rhat + o);
    if (act1) acc0[0] += rh.x * (double)y0;
    if (o1.act1) acc1[0] += rh.y * (double)y1;
  }
  __device__ float4 x4(int c, int g) const { return ld4((const float *)p, (int64_t)c * k + 4 * g); }
  __device__ Pre4 pre4(int64_t i, int j0) const { return Pre4{ld4((const float *)rhat, i * k + j0), {}, false}; }
  __device__ void row4p(int64_t i, int j0, const float4 &y, const Pre4 &pr, double (&acc)[NP][4]) const {
    store4nt(v, i * k + j0, y, a4c);
#pragma unroll
    for (int c = 0; c < 4; ++c)
      if (a4c[c]) acc[0][c] += (double)comp(pr.a, c) * (double)comp(y, c);
  }
  __device__ void row4(int64_t i, int j0, const float4 &y, double (&acc)[NP][4]) const {
    row4p(i, j0, y, pre4(i, j0), acc);
  }
  __device__ void last_lane(int j, const double *s) const {
    if (!st->active[j]) return;
    const double a = st->rho[j] / s[0];
    st->alpha[j] = a;
    if (s[0] == 0.0 || bad(a)) {
      st->status[j] = YAFEL_HIP_STATUS_BREAKDOWN;
      st->active[j] = 0;
    }
  }
  __device__ void last_all(int na, int nw) const {
    st->first = 0;
    all_inactive_done(st, na);
  }
};

// s = r - alpha v
template <class T>
struct MS : MBase {
  static constexpr int NP = 0;
  const T *r, *v;
  T *s;
  __device__ bool skip() const { return !running(); }
  __device__ void begin() {}
  __device__ void elem(int64_t i, int j, double *) const {
    if (!st->active[j]) return;
    const int64_t o = i * k + j;
    s[o] = r[o] - v[o] * (T)st->alpha[j];
  }
  static constexpr bool V4 = sizeof(T) == 4;
  float al4[4];
  __device__ void bind4(int g) {
    MBase::bind4(g);
#pragma unroll
    for (int c = 0; c < 4; ++c) al4[c] = (float)st->alpha[4 * g + c];
  }
  struct In {
    float4 r, v;
  };
  __device__ In load4(int64_t, int64_t o) const { return In{ld4((const float *)r, o), ld4((const float *)v, o)}; }
  __device__ void apply4(int64_t, int64_t o, const In &a, double (&)[1][4]) const {
    const float4 sn = make_float4(a.r.x - a.v.x * al4[0], a.r.y - a.v.y * al4[1], a.r.z - a.v.z * al4[2],
                                  a.r.w - a.v.w * al4[3]);
    store4((float *)s, o, sn, a4c);
  }
  __device__ void last_lane(int, const double *) const {}
};

// t = A (D^-1 s), t.t, t.s -> omega; D^-1 scales the matrix columns
template <class T>
struct MT : MBase {
  static constexpr int NP = 2;
  const T *s, *dinv;
  T *t;
  __device__ bool skip() const { return !running(); }
  __device__ void begin() {}
  __device__ const T *cscale() const { return dinv; }
  __device__ T x(int c, int j) const { return s[(int64_t)c * k + j]; }
  // staged products: values of A D^-1, rows of s as they are
  static constexpr bool kScaledVals = true;
  using Raw = float4;
  __device__ Raw load4(int c, int g) const { return ld4((const float *)s, (int64_t)c * k + 4 * g); }
  __device__ static float4 make4(const Raw &r) { return r; }
  __device__ float4 stage4(int c, int g) const { return load4(c, g); }
  __device__ void row(int64_t i, int j, T y, double *acc) const {
    if (!st->active[j]) return;
    t[i * k + j] = y;
    acc[0] += (double)y * (double)y;
    acc[1] += (double)y * (double)s[i * k + j];
  }
  static constexpr bool kRow2 = sizeof(T) == 8;
  __host__ bool pair_ok() const { return ((uintptr_t)t | (uintptr_t)s) % 16 == 0; }
  __device__ void row2(int64_t i, int j, T y0, T y1, const MT &o1, double *acc0, double *acc1) const {
    const int64_t o = i * k + j;
    store2nt(t, o, y0, y1, act1, o1.act1);
    const double2 sv = *reinterpret_cast<const double2 *>(s + o);
    if (act1) {
      acc0[0] += (double)y0 * (double)y0;
      acc0[1] += (double)y0 * sv.x;
    }
    if (o1.act1) {
      acc1[0] += (double)y1 * (double)y1;
      acc1[1] += (double)y1 * sv.y;
    }
  }
  __device__ float4 x4(int c, int g) const { return ld4((const float *)s, (int64_t)c * k + 4 * g); }
  __device__ Pre4 pre4(int64_t i, int j0) const { return Pre4{ld4((const float *)s, i * k + j0), {}, false}; }
  // the staged rows are rows of s, so the own row's staged quad is s itself
  static constexpr bool kOwnStaged = true;
  __device__ static Pre4 own4(const float4 &staged) { return Pre4{staged, {}, true}; }
  __device__ void row4p(int64_t i, int j0, const float4 &y, const Pre4 &pr, double (&acc)[NP][4]) const {
    store4nt(t, i * k + j0, y, a4c);
#pragma unroll
    for (int c = 0; c < 4; ++c)
      if (a4c[c]) {
        const double yc = comp(y, c);
        acc[0][c] += yc * yc;
        acc[1][c] += yc * (double)comp(pr.a, c);
      }
  }
  __device__ void row4(int64_t i, int j0, const float4 &y, double (&acc)[NP][4]) const {
    row4p(i, j0, y, pre4(i, j0), acc);
  }
  __device__ void last_lane(int j, const double *sm) const {
    if (!st->active[j]) return;
    const double w = sm[0] == 0.0 ? 0.0 : sm[1] / sm[0];
    st->omega[j] = w;
    if (w == 0.0 || bad(w)) {
      st->status[j] = YAFEL_HIP_STATUS_BREAKDOWN;
      st->active[j] = 0;
    }
  }
  __device__ void last_all(int na, int nw) const { all_inactive_done(st, na); }
};

// x += alpha D^-1 p + omega D^-1 s, r = s - omega t; rhat.r, r.r
template <class T>
struct MX : MBase {
  static constexpr int NP = 2;
  T *x, *r;
  const T *p, *s, *t, *rhat, *dinv;
  __device__ bool skip() const { return !running(); }
  __device__ void begin() {}
  __device__ void elem(int64_t i, int j, double *acc) const {
    if (!st->active[j]) return;
    const int64_t o = i * k + j;
    const T a = (T)st->alpha[j], w = (T)st->omega[j], d = dinv[i];
    x[o] += (d * p[o]) * a + (d * s[o]) * w;
    const T ri = s[o] - t[o] * w;
    r[o] = ri;
    acc[0] += (double)rhat[o] * (double)ri;
    acc[1] += (double)ri * (double)ri;
  }
  static constexpr bool V4 = sizeof(T) == 4;
  static constexpr int ME4 = 2;
  float al4[4], om4[4];
  __device__ void bind4(int g) {
    MBase::bind4(g);
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      al4[c] = (float)st->alpha[4 * g + c];
      om4[c] = (float)st->omega[4 * g + c];
    }
  }
  struct In {
    float4 x, p, s, t, rh;
    float d;
  };
  __device__ In load4(int64_t i, int64_t o) const {
    return In{ld4((const float *)x, o), ld4((const float *)p, o), ld4((const float *)s, o), ld4((const float *)t, o),
              ld4((const float *)rhat, o), (float)dinv[i]};
  }
  __device__ void apply4(int64_t, int64_t o, const In &a, double (&acc)[2][4]) const {
    const float xv[4] = {a.x.x, a.x.y, a.x.z, a.x.w}, pv[4] = {a.p.x, a.p.y, a.p.z, a.p.w};
    const float sv[4] = {a.s.x, a.s.y, a.s.z, a.s.w}, tv[4] = {a.t.x, a.t.y, a.t.z, a.t.w};
    const float hv[4] = {a.rh.x, a.rh.y, a.rh.z, a.rh.w};
    float xn[4], rn[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      xn[c] = xv[c] + ((a.d * pv[c]) * al4[c] + (a.d * sv[c]) * om4[c]);
      rn[c] = sv[c] - tv[c] * om4[c];
      if (a4c[c]) {
        acc[0][c] += (double)hv[c] * (double)rn[c];
        acc[1][c] += (double)rn[c] * (double)rn[c];
      }
    }
    store4((float *)x, o, make_float4(xn[0], xn[1], xn[2], xn[3]), a4c);
    store4((float *)r, o, make_float4(rn[0], rn[1], rn[2], rn[3]), a4c);
  }
  __device__ void last_lane(int j, const double *sm) const {
    if (!st->active[j]) return;
    request_pause_or_advance(st, j, sm[0], sm[1], false);
  }
  __device__ void last_all(int na, int nw) const {
    st->iter += 1;
    pause_when_half_want(st, na, nw);
    all_inactive_done(st, na);
  }
};

// ---- CG / Jacobi-PCG: A kernel q = A p_new, p_new = z + beta p_old
template <class T>
struct MCgA : MBase {
  static constexpr int NP = 1;
  const T *z;
  T *p0, *p1, *q;
  const T *pold;
  T *pnew;
  int cur;
  __device__ bool skip() const { return !running(); }
  __device__ void begin() {
    cur = st->pidx;
    pold = cur ? p1 : p0;
    pnew = cur ? p0 : p1;
  }
  __device__ T bet(int j) const { return st->first ? T(0) : (T)st->beta[j]; }
  // beta of this lane's right-hand side, read from the state once (bind1)
  // instead of at every gather, where the loads of the state would stay in
  // flight beside the gathers
  T bl;
  __device__ void bind1(int j) {
    MBase::bind1(j);
    bl = bet(j);
  }
  __device__ T x(int c, int j) const {
    const int64_t o = (int64_t)c * k + j;
    return fma(bl, pold[o], z[o]);
  }
  __device__ void row(int64_t i, int j, T y, double *acc) const {
    if (!(sizeof(T) == 4 ? act1 : st->active[j] != 0)) return;
    const int64_t o = i * k + j;
    const T pn = fma(bl, pold[o], z[o]);
    pnew[o] = pn;
    q[o] = y;
    acc[0] += (double)pn * (double)y;
  }
  float b4[4];  // beta of this lane's four right-hand sides (bind4)
  __device__ void bind4(int g) {
    MBase::bind4(g);
#pragma unroll
    for (int c = 0; c < 4; ++c) b4[c] = (float)bet(4 * g + c);
  }
  __device__ float4 pn4(int64_t o) const {
    const float4 po = ld4((const float *)pold, o), zz = ld4((const float *)z, o);
    return make_float4(fmaf(b4[0], po.x, zz.x), fmaf(b4[1], po.y, zz.y), fmaf(b4[2], po.z, zz.z),
                       fmaf(b4[3], po.w, zz.w));
  }
  __device__ float4 x4(int c, int g) const { return pn4((int64_t)c * k + 4 * g); }
  // staged rows: p_new = z + beta p_old of the column (betas of bind4s)
  float b4s[4];  // beta of the four right-hand sides this thread stages (bind4s)
  __device__ void bind4s(int g) {
#pragma unroll
    for (int c = 0; c < 4; ++c) b4s[c] = (float)bet(4 * g + c);
  }
  struct Raw {
    float4 po, z;
  };
  __device__ Raw load4(int c, int g) const {
    const int64_t o = (int64_t)c * k + 4 * g;
    return Raw{ld4((const float *)pold, o), ld4((const float *)z, o)};
  }
  __device__ float4 make4(const Raw &r) const {
    return make_float4(fmaf(b4s[0], r.po.x, r.z.x), fmaf(b4s[1], r.po.y, r.z.y), fmaf(b4s[2], r.po.z, r.z.z),
                       fmaf(b4s[3], r.po.w, r.z.w));
  }
  __device__ float4 stage4(int c, int g) const { return make4(load4(c, g)); }
  __device__ Pre4 pre4(int64_t i, int j0) const {
    const int64_t o = i * k + j0;
    return Pre4{ld4((const float *)pold, o), ld4((const float *)z, o), false};
  }
  // the staged rows are p_new = z + beta p_old (make4 with this pass's betas,
  // the betas row4p uses), so the own row's staged quad is p_new itself
  static constexpr bool kOwnStaged = true;
  __device__ static Pre4 own4(const float4 &staged) { return Pre4{staged, {}, true}; }
  __device__ void row4p(int64_t i, int j0, const float4 &y, const Pre4 &pr, double (&acc)[NP][4]) const {
    const int64_t o = i * k + j0;
    const float4 pn = pr.own ? pr.a : make_float4(fmaf(b4[0], pr.a.x, pr.b.x), fmaf(b4[1], pr.a.y, pr.b.y),
                                  fmaf(b4[2], pr.a.z, pr.b.z), fmaf(b4[3], pr.a.w, pr.b.w));
    store4nt(pnew, o, pn, a4c);
    store4nt(q, o, y, a4c);
#pragma unroll
    for (int c = 0; c < 4; ++c)
      if (a4c[c]) acc[0][c] += (double)comp(pn, c) * (double)comp(y, c);
  }
  __device__ void row4(int64_t i, int j0, const float4 &y, double (&acc)[NP][4]) const {
    row4p(i, j0, y, pre4(i, j0), acc);
  }
  __device__ void last_lane(int j, const double *s) const {
    if (!st->active[j]) return;
    const double a = st->rho[j] / s[0];
    st->alpha[j] = a;
    if (s[0] == 0.0 || bad(a)) {
      st->status[j] = YAFEL_HIP_STATUS_BREAKDOWN;
      st->active[j] = 0;
    }
  }
  __device__ void last_all(int na, int nw) const {
    st->first = 0;
    st->pidx = cur ^ 1;
    all_inactive_done(st, na);
  }
};

// x += alpha p, r -= alpha q, z = D^-1 r; r.z, r.r
template <class T>
struct MCgB : MBase {
  static constexpr int NP = 2;
  T *x, *r, *z;
  const T *p0, *p1, *q, *dinv;
  int prec;
  const T *p;
  __device__ bool skip() const { return !running(); }
  __device__ void begin() { p = st->pidx ? p1 : p0; }
  __device__ void elem(int64_t i, int j, double *acc) const {
    if (!st->active[j]) return;
    const int64_t o = i * k + j;
    const T a = (T)st->alpha[j];
    x[o] = fma(a, p[o], x[o]);
    const T ri = fma(-a, q[o], r[o]);
    r[o] = ri;
    T zi = ri;
    if (prec) {
      zi = dinv[i] * ri;
      z[o] = zi;
    }
    acc[0] += (double)ri * (double)zi;
    acc[1] += (double)ri * (double)ri;
  }
  static constexpr bool V4 = sizeof(T) == 4;
  static constexpr int ME4 = 2;
  float al4[4];
  __device__ void bind4(int g) {
    MBase::bind4(g);
#pragma unroll
    for (int c = 0; c < 4; ++c) al4[c] = (float)st->alpha[4 * g + c];
  }
  struct In {
    float4 x, p, q, r;
    float d;
  };
  __device__ In load4(int64_t i, int64_t o) const {
    return In{ld4((const float *)x, o), ld4((const float *)p, o), ld4((const float *)q, o), ld4((const float *)r, o),
              prec ? (float)dinv[i] : 1.f};
  }
  __device__ void apply4(int64_t, int64_t o, const In &a, double (&acc)[2][4]) const {
    const float xv[4] = {a.x.x, a.x.y, a.x.z, a.x.w}, pv[4] = {a.p.x, a.p.y, a.p.z, a.p.w};
    const float qv[4] = {a.q.x, a.q.y, a.q.z, a.q.w}, rv[4] = {a.r.x, a.r.y, a.r.z, a.r.w};
    float xn[4], rn[4], zn[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      xn[c] = fmaf(al4[c], pv[c], xv[c]);
      rn[c] = fmaf(-al4[c], qv[c], rv[c]);
      zn[c] = prec ? a.d * rn[c] : rn[c];
      if (a4c[c]) {
        acc[0][c] += (double)rn[c] * (double)zn[c];
        acc[1][c] += (double)rn[c] * (double)rn[c];
      }
    }
    store4((float *)x, o, make_float4(xn[0], xn[1], xn[2], xn[3]), a4c);
    store4((float *)r, o, make_float4(rn[0], rn[1], rn[2], rn[3]), a4c);
    if (prec) store4((float *)z, o, make_float4(zn[0], zn[1], zn[2], zn[3]), a4c);
  }
  __device__ void last_lane(int j, const double *sm) const {
    if (!st->active[j]) return;
    request_pause_or_advance(st, j, sm[0], sm[1], true);
  }
  __device__ void last_all(int na, int nw) const {
    st->iter += 1;
    pause_when_half_want(st, na, nw);
    all_inactive_done(st, na);
  }
};

// FP64 residual R = B - A X_hi -> r (inner precision), z (CG), rhat (init,
// BiCGSTAB); per RHS: r.w, r.r, true r.r, b.b.  MODE 0 init, 1 refinement.
template <class TL, int MODE>
struct MRes : MBase {
  static constexpr int NP = 4;
  const double *X, *B;
  TL *r, *z, *rhat;
  const TL *dinvl;
  int cgfam, prec;
  __device__ bool skip() const { return MODE == 0 ? false : (st->done || !st->paused); }
  __device__ void begin() {}
  __device__ double x(int c, int j) const { return X[(int64_t)c * k + j]; }
  __device__ void row(int64_t i, int j, double y, double *acc) const {
    const int64_t o = i * k + j;
    const bool act = MODE == 0 || st->active[j];
    const double rh = B[o] - y;
    const TL rl = (TL)rh;
    double w = (double)rl;
    if (act) {
      r[o] = rl;
      if (cgfam && prec) {
        const TL zi = dinvl[i] * rl;
        z[o] = zi;
        w = (double)zi;
      } else if (!cgfam) {
        if (MODE == 0) rhat[o] = rl;
        else w = (double)rhat[o];
      }
    }
    acc[0] += (double)rl * w;
    acc[1] += (double)rl * (double)rl;
    acc[2] += rh * rh;
    acc[3] += B[o] * B[o];
  }
  __device__ void last_lane(int j, const double *s) const {
    st->true_rr[j] = s[2];
    st->want[j] = 0;
    if (MODE == 0) {
      st->bb[j] = s[3];
      st->rho[j] = s[0];
      st->rho0[j] = s[0];
      st->rho_prev[j] = 1.0;
      st->rr[j] = s[1];
      st->rr_ref[j] = s[1];
      st->alpha[j] = 1.0;
      st->omega[j] = 1.0;
      st->beta[j] = 0.0;
      st->it[j] = 0;
      st->status[j] = YAFEL_HIP_STATUS_MAXITER;
      st->active[j] = 1;
      if (s[3] == 0.0 || sqrt(s[2] / s[3]) < st->tol) {
        st->status[j] = YAFEL_HIP_STATUS_CONVERGED;
        st->active[j] = 0;
        st->crit[j] = s[3] == 0.0 ? 0.0 : sqrt(s[2] / s[3]);
      }
      return;
    }
    if (!st->active[j]) return;
    const double crit = sqrt(s[2] / st->bb[j]);
    st->crit[j] = crit;
    if (bad(s[2])) {
      st->status[j] = YAFEL_HIP_STATUS_NAN;
      st->active[j] = 0;
      return;
    }
    if (crit < st->tol) {
      st->status[j] = YAFEL_HIP_STATUS_CONVERGED;
      st->active[j] = 0;
      return;
    }
    if (st->it[j] >= st->max_iter) {
      st->status[j] = YAFEL_HIP_STATUS_MAXITER;
      st->active[j] = 0;
      return;
    }
    // rho_prev[j] is rho before the last iteration for every lane -- whether the
    // lane asked for this refinement (rho not advanced) or advanced normally
    if (cgfam) st->beta[j] = s[0] / st->rho_prev[j];
    else st->beta[j] = (s[0] / st->rho_prev[j]) * (st->alpha[j] / st->omega[j]);
    st->rho[j] = s[0];
    st->rr[j] = s[1];
    st->rr_ref[j] = s[1];
  }
  __device__ void last_all(int na, int nw) const {
    if (MODE == 0) {
      st->first = 1;
      st->paused = 0;
      st->pidx = 0;
    } else {
      st->refinements += 1;
      st->paused = 0;
    }
    all_inactive_done(st, na);
  }
};

// x_hi += x_lo, x_lo = 0 (MIXED, while paused)
template <class TL>
struct MFold : MBase {
  static constexpr int NP = 0;
  double *X;
  TL *xl;
  __device__ bool skip() const { return st->done || !st->paused; }
  __device__ void begin() {}
  __device__ void elem(int64_t i, int j, double *) const {
    const int64_t o = i * k + j;
    X[o] += (double)xl[o];
    xl[o] = TL(0);
  }
  // k_melem4: every column, active or not, as elem does
  static constexpr bool V4 = sizeof(TL) == 4;
  struct In {
    double2 x01, x23;
    float4 l;
  };
  __device__ In load4(int64_t, int64_t o) const {
    const double2 *x2 = reinterpret_cast<const double2 *>(X + o);
    return In{x2[0], x2[1], ld4((const float *)xl, o)};
  }
  __device__ void apply4(int64_t, int64_t o, const In &a, double (&)[1][4]) const {
    double2 *x2 = reinterpret_cast<double2 *>(X + o);
    x2[0] = make_double2(a.x01.x + (double)a.l.x, a.x01.y + (double)a.l.y);
    x2[1] = make_double2(a.x23.x + (double)a.l.z, a.x23.y + (double)a.l.w);
    *reinterpret_cast<float4 *>((float *)xl + o) = make_float4(0.f, 0.f, 0.f, 0.f);
  }
  __device__ void last_lane(int, const double *) const {}
};

// final per-RHS true residual
struct MReport : MBase {
  static constexpr int NP = 2;
  const double *X, *B;
  __device__ bool skip() const { return false; }
  __device__ void begin() {}
  __device__ double x(int c, int j) const { return X[(int64_t)c * k + j]; }
  __device__ void row(int64_t i, int j, double y, double *acc) const {
    const double r = B[i * k + j] - y;
    acc[0] += r * r;
    acc[1] += B[i * k + j] * B[i * k + j];
  }
  __device__ void last_lane(int j, const double *s) const {
    st->true_rr[j] = s[0];
    st->bb[j] = s[1];
  }
};

// Y = A X for the plain product (yafel_hip_spmm): no state, no sums.
struct MPlain {
  static constexpr int NP = 0;
  static constexpr bool kScaledVals = false;
  static constexpr bool kOwnStaged = false;
  const float *X;
  float *Y;
  int k;
  double *part;
  __device__ bool skip() const { return false; }
  __device__ void begin() {}
  __device__ void bind4(int) {}
  __device__ void bind4s(int) {}
  struct Raw {
    float4 x;
  };
  __device__ Raw load4(int c, int g) const { return Raw{ld4(X, (int64_t)c * k + 4 * g)}; }
  __device__ static float4 make4(const Raw &r) { return r.x; }
  __device__ float4 stage4(int c, int g) const { return ld4(X, (int64_t)c * k + 4 * g); }
  __device__ Pre4 pre4(int64_t, int) const { return Pre4{}; }
  __device__ void row4p(int64_t i, int j0, const float4 &y, const Pre4 &, double (&)[1][4]) const {
    typedef float v4f __attribute__((ext_vector_type(4)));
    const v4f v = {y.x, y.y, y.z, y.w};  // written once: non-temporal, as MBase::store4nt
    __builtin_nontemporal_store(v, reinterpret_cast<v4f *>(Y + i * k + j0));
  }
};

// Y = A X in FP64 for the plain product through the staged kernel.
struct MPlain64 {
  static constexpr int NP = 0;
  static constexpr bool kScaledVals = false;
  const double *X;
  double *Y;
  int k;
  double *part;
  __device__ bool skip() const { return false; }
  __device__ void begin() {}
  __device__ void bind1(int) {}
  __device__ double x(int c, int j) const { return X[(int64_t)c * k + j]; }
  __device__ void row(int64_t i, int j, double y, double *) const {
    __builtin_nontemporal_store(y, Y + i * k + j);  // written once, as MBase::store4nt
  }
  static constexpr bool kRow2 = true;
  __host__ bool pair_ok() const { return (uintptr_t)Y % 16 == 0; }
  __device__ void row2(int64_t i, int j, double y0, double y1, const MPlain64 &, double *, double *) const {
    MBase::store2nt(Y, i * k + j, y0, y1, true, true);
  }
};

// The CSR-order copy of the staged products (pattern part, once per
// pattern): each entry's sliced position, the distinct columns of each group
// of 64 consecutive rows and each entry's position in its group's list.
// mx_dmax = 0 when a list is too long for 16-bit positions (never staged).
template <class T>
__global__ void k_take(int64_t n, const int *__restrict__ src, const T *__restrict__ v, T *__restrict__ out) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) out[i] = src[i] >= 0 ? v[src[i]] : T(0);
}

// the same, each value times its column's D^-1 (the product the FP32
// sliced kernels form as they load a value and its column's scale)
template <class T>
__global__ void k_take_scaled(int64_t n, const int *__restrict__ src, const int *__restrict__ col,
                              const T *__restrict__ v, const T *__restrict__ dinv, T *__restrict__ out) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) {
    const int e = src[i];
    out[i] = e >= 0 ? v[e] * dinv[col[e]] : T(0);
  }
}

void matrix_build_staged(yafel_hip_matrix_s *A) {
  if (A->mx_dmax >= 0) return;
  const int64_t rows = A->brows, nsl = A->nsl, ntot = A->ntot, nnz = A->nblk;
  hipStream_t st = A->ctx->stream;
  std::vector<int> ptr(rows + 1), sptr(nsl + 1), col(std::max<int64_t>(ntot, 1)), ip;
  HIPCHK(hipMemcpyAsync(ptr.data(), A->ptr.p, (rows + 1) * 4, hipMemcpyDeviceToHost, st));
  HIPCHK(hipMemcpyAsync(sptr.data(), A->sptr.p, (nsl + 1) * 4, hipMemcpyDeviceToHost, st));
  if (ntot) HIPCHK(hipMemcpyAsync(col.data(), A->col.p, ntot * 4, hipMemcpyDeviceToHost, st));
  if (A->iperm.p) {
    ip.resize(rows);
    HIPCHK(hipMemcpyAsync(ip.data(), A->iperm.p, rows * 4, hipMemcpyDeviceToHost, st));
  }
  HIPCHK(hipStreamSynchronize(st));
  // each row's columns and their positions in the sliced storage, CSR order
  std::vector<int> rcol(std::max<int64_t>(nnz, 1)), rsrc(std::max<int64_t>(nnz, 1));
  for (int64_t r = 0; r < rows; ++r) {
    const int64_t i = ip.empty() ? r : ip[r];
    const int64_t base = (int64_t)sptr[i >> 6] * 64 + (i & 63);
    for (int t = 0; t < ptr[r + 1] - ptr[r]; ++t) {
      const int64_t k = base + (int64_t)t * 64;
      rsrc[ptr[r] + t] = (int)k;
      rcol[ptr[r] + t] = col[k];
    }
  }
  col.clear();
  col.shrink_to_fit();
  // groups of up to 64 rows, each grown breadth-first over the pattern from
  // the lowest row not yet in a group; `order` is the queue of every group
  std::vector<int> grp(rows, -1), order, gstart;
  order.reserve(rows);
  for (int64_t seed = 0; seed < rows; ++seed) {
    if (grp[seed] >= 0) continue;
    const int g = (int)gstart.size();
    const size_t b = order.size();
    gstart.push_back((int)b);
    grp[seed] = g;
    order.push_back((int)seed);
    for (size_t head = b; head < order.size() && order.size() - b < 64; ++head) {
      const int u = order[head];
      for (int e = ptr[u]; e < ptr[u + 1] && order.size() - b < 64; ++e) {
        const int v = rcol[e];
        if (v >= 0 && v < rows && grp[v] < 0) {
          grp[v] = g;
          order.push_back(v);
        }
      }
    }
  }
  const int64_t ngr = (int64_t)gstart.size();
  gstart.push_back((int)rows);
  // 64 positions per group: its rows longest first (ties in queue order),
  // then empty positions
  const int64_t npos = ngr * 64;
  REQUIRE(npos < INT32_MAX, "staged products: too many rows");
  std::vector<int> grow(std::max<int64_t>(npos, 1), -1);
  for (int64_t g = 0; g < ngr; ++g) {
    auto first = order.begin() + gstart[g], last = order.begin() + gstart[g + 1];
    std::stable_sort(first, last, [&](int x, int y) { return ptr[x + 1] - ptr[x] > ptr[y + 1] - ptr[y]; });
    std::copy(first, last, grow.begin() + 64 * g);
  }
  order.clear();
  order.shrink_to_fit();
  // rows padded to a multiple of 4 entries (zero values, position 0)
  std::vector<int> pp(npos + 1, 0);
  for (int64_t p = 0; p < npos; ++p) {
    const int r = grow[p];
    pp[p + 1] = pp[p] + (r >= 0 ? ((ptr[r + 1] - ptr[r] + 3) & ~3) : 0);
  }
  const int64_t npad = pp[npos];
  REQUIRE(npad < INT32_MAX, "staged products: too many entries");
  std::vector<int> src(std::max<int64_t>(npad, 1), -1), dptr(ngr + 1), dcol, tmp, ecol(std::max<int64_t>(npad, 1), 0);
  std::vector<unsigned short> lidx(std::max<int64_t>(npad, 1), 0), dl(std::max<int64_t>(npos, 1), 0);
  dcol.reserve(nnz / 3 + 1);
  int dmax = 0;
  int64_t emax = 0, nbig = 0;
  bool fits = true;
  for (int64_t g = 0; g < ngr; ++g) {
    const int64_t p0 = 64 * g, p1 = p0 + 64;
    tmp.clear();
    for (int64_t p = p0; p < p1; ++p) {
      const int r = grow[p];
      if (r < 0) continue;
      for (int t = 0; t < ptr[r + 1] - ptr[r]; ++t) {
        src[pp[p] + t] = rsrc[ptr[r] + t];
        ecol[pp[p] + t] = rcol[ptr[r] + t];
        tmp.push_back(rcol[ptr[r] + t]);
      }
      // every row's own column is staged (a product's row epilogue reads it
      // from LDS), also where the row stores no diagonal entry
      if (r < A->bcols) tmp.push_back(r);
    }
    std::sort(tmp.begin(), tmp.end());
    tmp.erase(std::unique(tmp.begin(), tmp.end()), tmp.end());
    dptr[g] = (int)dcol.size();
    dcol.insert(dcol.end(), tmp.begin(), tmp.end());
    dmax = std::max(dmax, (int)tmp.size());
    emax = std::max<int64_t>(emax, pp[p1] - pp[p0]);
    if ((int64_t)tmp.size() > kStageRows || pp[p1] - pp[p0] > kStageEnt) ++nbig;
    if (tmp.size() > 65535) {
      fits = false;
      continue;
    }
    for (int64_t e = pp[p0]; e < pp[p1]; ++e)
      if (src[e] >= 0) lidx[e] = (unsigned short)(std::lower_bound(tmp.begin(), tmp.end(), ecol[e]) - tmp.begin());
    for (int64_t p = p0; p < p1; ++p) {
      const int r = grow[p];
      if (r < 0 || r >= A->bcols) continue;
      dl[p] = (unsigned short)(std::lower_bound(tmp.begin(), tmp.end(), r) - tmp.begin());
    }
  }
  dptr[ngr] = (int)dcol.size();
  A->mx_ptr.alloc((npos + 1) * 4);
  HIPCHK(hipMemcpy(A->mx_ptr.p, pp.data(), (npos + 1) * 4, hipMemcpyHostToDevice));
  A->mx_grow.alloc(std::max<int64_t>(npos, 1) * 4);
  HIPCHK(hipMemcpy(A->mx_grow.p, grow.data(), std::max<int64_t>(npos, 1) * 4, hipMemcpyHostToDevice));
  A->mx_dl.alloc(std::max<int64_t>(npos, 1) * 2);
  HIPCHK(hipMemcpy(A->mx_dl.p, dl.data(), std::max<int64_t>(npos, 1) * 2, hipMemcpyHostToDevice));
  A->mx_src.alloc(src.size() * 4);
  A->mx_lidx.alloc(lidx.size() * 2);
  A->mx_dptr.alloc((ngr + 1) * 4);
  A->mx_dcol.alloc(std::max<size_t>(dcol.size(), 1) * 4);
  HIPCHK(hipMemcpy(A->mx_src.p, src.data(), src.size() * 4, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(A->mx_lidx.p, lidx.data(), lidx.size() * 2, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(A->mx_dptr.p, dptr.data(), (ngr + 1) * 4, hipMemcpyHostToDevice));
  if (!dcol.empty()) HIPCHK(hipMemcpy(A->mx_dcol.p, dcol.data(), dcol.size() * 4, hipMemcpyHostToDevice));
  A->mx_ngr = ngr;
  A->mx_dmax = fits ? dmax : 0;
  A->mx_emax = emax;
  A->mx_nbig = nbig;
}

// The staged products' FP32 values, refreshed when the FP32 copy changed.
void matrix_refresh_staged(yafel_hip_matrix_s *A) {
  if (A->mx_gen == A->v32_gen) return;
  const int64_t nnz = (int64_t)(A->mx_src.bytes / 4);
  if (A->mx_val32.bytes < (size_t)std::max<int64_t>(nnz, 1) * 4) A->mx_val32.alloc(std::max<int64_t>(nnz, 1) * 4);
  if (nnz) {
    k_take<float><<<ceil_div(nnz, 256), 256, 0, A->ctx->stream>>>(nnz, A->mx_src.as<int>(), A->val32.as<float>(),
                                                          A->mx_val32.as<float>());
    HIPCHK(hipGetLastError());
  }
  A->mx_gen = A->v32_gen;
}

// The staged values of A D^-1, refreshed when the FP32 copy or the diagonal
// changed.
const float *matrix_staged_scaled(yafel_hip_matrix_s *A) {
  if (A->mx_sgen == A->v32_gen && A->mx_dgen == A->diag_gen) return A->mx_val32s.as<float>();
  REQUIRE(A->diag_ok && A->dinv32.p, "staged products: no diagonal");
  const int64_t nnz = (int64_t)(A->mx_src.bytes / 4);
  if (A->mx_val32s.bytes < (size_t)std::max<int64_t>(nnz, 1) * 4) A->mx_val32s.alloc(std::max<int64_t>(nnz, 1) * 4);
  if (nnz) {
    k_take_scaled<float><<<ceil_div(nnz, 256), 256, 0, A->ctx->stream>>>(nnz, A->mx_src.as<int>(), A->col.as<int>(),
                                                                 A->val32.as<float>(), A->dinv32.as<float>(),
                                                                 A->mx_val32s.as<float>());
    HIPCHK(hipGetLastError());
  }
  A->mx_sgen = A->v32_gen;
  A->mx_dgen = A->diag_gen;
  return A->mx_val32s.as<float>();
}

// The FP64 staged products' values: the CSR-order copy of the FP64 matrix,
// or of A D^-1 (scaled), refreshed when the values or the diagonal changed.
const double *matrix_staged_val64(yafel_hip_matrix_s *A, bool scaled) {
  const int64_t nnz = (int64_t)(A->mx_src.bytes / 4);
  DBuf &dst = scaled ? A->mx_val64s : A->mx_val64;
  if (scaled ? (A->mx_sgen64 == A->v64_gen && A->mx_dgen64 == A->diag_gen) : A->mx_gen64 == A->v64_gen)
    return dst.as<double>();
  REQUIRE(!scaled || (A->diag_ok && A->dinv64.p), "staged products: no diagonal");
  if (dst.bytes < (size_t)std::max<int64_t>(nnz, 1) * 8) dst.alloc(std::max<int64_t>(nnz, 1) * 8);
  if (nnz) {
    if (scaled)
      k_take_scaled<double><<<ceil_div(nnz, 256), 256, 0, A->ctx->stream>>>(
          nnz, A->mx_src.as<int>(), A->col.as<int>(), A->val64.as<double>(), A->dinv64.as<double>(),
          dst.as<double>());
    else
      k_take<double><<<ceil_div(nnz, 256), 256, 0, A->ctx->stream>>>(nnz, A->mx_src.as<int>(),
                                                                      A->val64.as<double>(), dst.as<double>());
    HIPCHK(hipGetLastError());
  }
  if (scaled) {
    A->mx_sgen64 = A->v64_gen;
    A->mx_dgen64 = A->diag_gen;
  } else {
    A->mx_gen64 = A->v64_gen;
  }
  return dst.as<double>();
}

// True when the staged products run on A (CSR, lists with 16-bit positions).
bool staged_ok(yafel_hip_matrix_s *A) {
  if (A->format != YAFEL_HIP_CSR || A->brows == 0) return false;
  matrix_build_staged(A);
  return A->mx_dmax > 0;
}

// True when the FP32 staged products run on A: at most one group in eight
// is over their LDS budget.  Those groups read X directly, one dependent
// gather per entry; when most are (3-D elasticity in CSR: about 5,000 entries
// per group of 64 rows) the sliced products with 16-byte gathers are faster.
bool staged32_ok(yafel_hip_matrix_s *A) { return staged_ok(A) && A->mx_nbig * 8 <= A->mx_ngr; }

// True when the FP64 staged products run on A: every group's list and
// entries fit their LDS budget.  Otherwise (3-D elasticity in CSR: about
// 5,000 entries per group of 64 rows) the sliced FP64 product, which gathers
// X directly, is the faster one.
bool staged64_ok(yafel_hip_matrix_s *A) {
  return staged_ok(A) && A->mx_dmax <= kStageRows64 && A->mx_emax <= kStageEnt64;
}

// grid: at most this many workgroups; the kernel's residency per CU caps it
// further.  Returns the grid launched (the partials' block count).
template <class Op>
int launch_mspmm_lds(yafel_hip_matrix_s *A, const Op &op, int grid, hipStream_t s) {
  matrix_refresh_staged(A);
  const float *val = Op::kScaledVals ? matrix_staged_scaled(A) : A->mx_val32.as<float>();
  const StagedCsr v{A->mx_ptr.as<int>(), val, A->mx_lidx.as<unsigned short>(), A->mx_dptr.as<int>(),
                    A->mx_dcol.as<int>(), A->mx_grow.as<int>(), A->mx_dl.as<unsigned short>(), A->mx_ngr,
                    std::min(A->mx_dmax, kStageRows)};
  const size_t lds = (size_t)v.dcap * KC * sizeof(float);
  auto cap = [&](int waves) {
    return (int)std::max<int64_t>(1, std::min<int64_t>({(int64_t)grid, A->mx_ngr, (int64_t)waves * A->ctx->num_cus}));
  };
  int g;
  if (A->mx_dmax <= kStageRows / 2 && A->mx_emax <= kStageEnt) {
    g = cap(pipe_waves<Op, kStageIter / 2>());
    k_mspmm_pipe<Op, kStageIter / 2><<<g, NT, lds, s>>>(v, op);
  } else if (A->mx_dmax <= kStageRows && A->mx_emax <= kStageEnt) {
    g = cap(pipe_waves<Op, kStageIter>());
    k_mspmm_pipe<Op, kStageIter><<<g, NT, lds, s>>>(v, op);
  } else {
    g = cap(2);
    k_mspmm_lds<Op><<<g, NT, lds, s>>>(v, op);
  }
  HIPCHK(hipGetLastError());
  return g;
}

// FP64 staged product; same grid rule as launch_mspmm_lds (residency per CU
// from the launch's LDS).
template <class Op>
int launch_mspmm_lds64(yafel_hip_matrix_s *A, const Op &op, int grid, hipStream_t s) {
  const double *val = matrix_staged_val64(A, Op::kScaledVals);
  // at least 64 staged rows: the partials' fold reuses their space
  const int dcap = std::max(64, std::min(A->mx_dmax, kStageRows64));
  const int ecap = (int)std::min<int64_t>((A->mx_emax + 3) & ~(int64_t)3, kStageEnt64);
  REQUIRE(A->mx_dmax <= dcap && A->mx_emax <= ecap, "staged FP64 product: a row group exceeds the LDS budget");
  int pair = 0;  // k even and 16-byte aligned outputs: a lane's two right-hand sides are adjacent and aligned
  if constexpr (Op::kRow2) pair = op.k % 2 == 0 && op.pair_ok();
  const StagedCsr64 v{A->mx_ptr.as<int>(), val, A->mx_lidx.as<unsigned short>(), A->mx_dptr.as<int>(),
                      A->mx_dcol.as<int>(), A->mx_grow.as<int>(), A->mx_ngr, dcap, ecap, pair};
  const size_t lds = (size_t)dcap * KC64 * 8 + (size_t)ecap * 10;
  const int per_cu = (int)std::max<size_t>(1, std::min<size_t>(4, (160 * 1024) / (lds + 512)));
  const int g = (int)std::max<int64_t>(
      1, std::min<int64_t>({(int64_t)grid, A->mx_ngr, (int64_t)per_cu * A->ctx->num_cus}));
  k_mspmm_lds64<Op><<<g, NT64, lds, s>>>(v, op);
  HIPCHK(hipGetLastError());
  return g;
}

template <class T>
CsrView<T> cview(yafel_hip_matrix_s *A) {
  return CsrView<T>{A->sptr.as<int>(), A->perm.p ? A->perm.as<int>() : nullptr, A->col.as<int>(),
                    (const T *)(sizeof(T) == 8 ? A->val64.p : A->val32.p), A->rows};
}

// dst (n x knew) = the columns cols[] of src (n x kold); cols[j] < 0: zero
template <class T>
__global__ void k_take_cols(int64_t n, int kold, int knew, const int *__restrict__ cols, const T *__restrict__ src,
                            T *__restrict__ dst) {
  const int64_t t = blockIdx.x * (int64_t)NT + threadIdx.x;
  if (t >= n * knew) return;
  const int64_t i = t / knew;
  const int c = cols[t - i * knew];
  dst[t] = c >= 0 ? src[i * kold + c] : T(0);
}

// dst column cols_out[j] (n x kout) = src column j (n x kin), for cols_out[j] >= 0
__global__ void k_put_cols(int64_t n, int kin, int kout, const int *__restrict__ cols_out,
                           const double *__restrict__ src, double *__restrict__ dst) {
  const int64_t t = blockIdx.x * (int64_t)NT + threadIdx.x;
  if (t >= n * kin) return;
  const int64_t i = t / kin;
  const int c = cols_out[t - i * kin];
  if (c >= 0) dst[i * kout + c] = src[t];
}

struct Multi {
  yafel_hip_ctx_s *ctx;
  yafel_hip_matrix_s *A;
  yafel_hip_solve_opts o;
  int64_t n;
  int k;
  bool mixed, cgfam, prec;
  MState *st;
  double *part;
  hipStream_t s;
  int grid;
  DBuf vecs[9];
  DBuf partbuf, stbuf;
  bool staged = false;  // FP32 products stage X in LDS
  bool staged64 = false;  // FP64 products stage X in LDS

  template <class O> O base(O op, int c) const {
    op.st = st;
    op.part = part;
    op.counter = st->counter + c;
    op.k = k;
    return op;
  }
  template <class T, class O> void spmm(const O &op) {
    // RB-row units of the slices, four per workgroup, within the partials'
    // grid
    const bool v4 = sizeof(T) == 4 && k % 4 == 0;
    const int64_t units = A->nsl * (sizeof(T) == 4 && !v4 ? 1 : 2);
    int g = (int)std::max<int64_t>(1, std::min<int64_t>((units + 3) / 4, grid));
    if constexpr (sizeof(T) == 4) {
      if (staged && v4) {
        // one group of 64 rows and one pass of 32 right-hand sides per
        // workgroup at a time, as many workgroups as are resident at once:
        // a larger grid would only add partials for the fold to read
        g = launch_mspmm_lds(A, op, grid, s);
      } else if (v4) {
        k_mspmm4<O><<<g, NT, 0, s>>>(cview<T>(A), op);
      } else {
        k_mspmm<T, O><<<g, NT, 0, s>>>(cview<T>(A), op);
      }
    } else if (staged64) {
      g = launch_mspmm_lds64(A, op, grid, s);
    } else {
      k_mspmm<T, O><<<g, NT, 0, s>>>(cview<T>(A), op);
    }
    HIPCHK(hipGetLastError());
    if constexpr (O::NP > 0) {
      k_mfinish<O><<<1, NTM, 0, s>>>(op, g);
      HIPCHK(hipGetLastError());
    }
  }
  template <class O> void elem(const O &op) {
    int g = grid;
    if (O::V4 && k % 4 == 0) {
      // units of 64 rows, spread evenly over two workgroups per CU
      const int64_t cap = std::max<int64_t>(1, std::min<int64_t>(grid, 2 * (int64_t)ctx->num_cus));
      const int64_t units = (n + 63) / 64, per = std::max<int64_t>(1, (units + cap - 1) / cap);
      g = (int)((units + per - 1) / per);
      if constexpr (O::V4) k_melem4<O><<<g, NT4, 0, s>>>(n, per, op);
    } else {
      // at most two workgroups per CU, each with at least 64 rows
      const int64_t cap = std::max<int64_t>(1, std::min<int64_t>(grid, 2 * (int64_t)ctx->num_cus));
      g = (int)std::max<int64_t>(1, std::min<int64_t>(cap, (n + 63) / 64));
      k_melem<O><<<g, NTE, 0, s>>>(n, op);
    }
    HIPCHK(hipGetLastError());
    if constexpr (O::NP > 0) {
      k_mfinish<O><<<1, NTM, 0, s>>>(op, g);
      HIPCHK(hipGetLastError());
    }
  }

  template <class T>
  void run(const double *B0, double *X0, yafel_hip_solve_info *info) {
    // B0, X0: the caller's n x k0 columns.  Once the batch has been compacted
    // (below), B and X are owned n x k copies of the columns still in it, and
    // map[j] is the caller's column of batch column j (-1: padding).
    const int k0 = k;
    const double *B = B0;
    double *X = X0;
    std::vector<int> map(k0);
    for (int j = 0; j < k0; ++j) map[j] = j;
    DBuf Bc, Xc, tmp, colsd;
    const size_t nk = (size_t)n * k;
    for (auto &v : vecs) v.alloc(nk * sizeof(T));
    T *r = vecs[0].as<T>(), *rhat = vecs[1].as<T>(), *p0 = vecs[2].as<T>(), *p1 = vecs[3].as<T>();
    T *v = vecs[4].as<T>(), *sv = vecs[5].as<T>(), *t = vecs[6].as<T>(), *xl = vecs[7].as<T>();
    T *z = vecs[8].as<T>();
    const T *dinv = (const T *)(sizeof(T) == 8 ? A->dinv64.p : A->dinv32.p);
    HIPCHK(hipMemsetAsync(p0, 0, nk * sizeof(T), s));
    HIPCHK(hipMemsetAsync(v, 0, nk * sizeof(T), s));
    if (mixed) HIPCHK(hipMemsetAsync(xl, 0, nk * sizeof(T), s));
    MState h{};
    h.k = k;
    h.method = o.method;
    h.stop = o.stop;
    h.mixed = mixed;
    h.tol = o.tol;
    h.delta2 = o.refine_delta * o.refine_delta;
    h.max_iter = o.max_iter;
    static_assert(sizeof(MState) + 3 * 1024 <= 65536, "pinned staging too small");
    std::memcpy(ctx->pinned, &h, sizeof(MState));
    HIPCHK(hipMemcpyAsync(st, ctx->pinned, sizeof(MState), hipMemcpyHostToDevice, s));
    HIPCHK(hipStreamSynchronize(s));
    const T *zz = prec ? z : r;
    auto residual = [&](int mode) {
      auto fill = [&](auto op) {
        op.X = X; op.B = B; op.r = r; op.z = z; op.rhat = rhat; op.dinvl = dinv;
        op.cgfam = cgfam; op.prec = prec;
        return op;
      };
      if (mode == 0) spmm<double>(fill(base(MRes<T, 0>{}, 0)));
      else spmm<double>(fill(base(MRes<T, 1>{}, 0)));
    };
    residual(0);
    const int K = o.check_every > 0 ? o.check_every : 16;
    auto iteration = [&]() {
      if (cgfam) {
        MCgA<T> a = base(MCgA<T>{}, 1);
        a.z = zz; a.p0 = p0; a.p1 = p1; a.q = v;
        spmm<T>(a);
        MCgB<T> bb = base(MCgB<T>{}, 2);
        bb.x = mixed ? xl : (T *)X; bb.r = r; bb.z = z; bb.p0 = p0; bb.p1 = p1; bb.q = v; bb.dinv = dinv;
        bb.prec = prec;
        elem(bb);
      } else {
        MP<T> pp = base(MP<T>{}, 1);
        pp.r = r; pp.v = v; pp.p = p0;
        elem(pp);
        MV<T> vv = base(MV<T>{}, 2);
        vv.p = p0; vv.rhat = rhat; vv.dinv = dinv; vv.v = v;
        spmm<T>(vv);
        MS<T> ss = base(MS<T>{}, 3);
        ss.r = r; ss.v = v; ss.s = sv;
        elem(ss);
        MT<T> tt = base(MT<T>{}, 4);
        tt.s = sv; tt.dinv = dinv; tt.t = t;
        spmm<T>(tt);
        MX<T> xx = base(MX<T>{}, 5);
        xx.x = mixed ? xl : (T *)X; xx.r = r; xx.p = p0; xx.s = sv; xx.t = t; xx.rhat = rhat; xx.dinv = dinv;
        elem(xx);
      }
    };
    auto refine = [&]() {
      if (mixed) {
        MFold<T> f = base(MFold<T>{}, 6);
        f.X = X;
        f.xl = xl;
        elem(f);
      }
      residual(1);
    };
    char *pin = (char *)ctx->pinned;
    hipEvent_t ev[2];
    HIPCHK(hipEventCreateWithFlags(&ev[0], hipEventDisableTiming));
    HIPCHK(hipEventCreateWithFlags(&ev[1], hipEventDisableTiming));
    const size_t slot = 1024;  // status headers after the state image
    auto chunk = [&](int si) {
      for (int i = 0; i < K; ++i) iteration();
      refine();
      HIPCHK(hipMemcpyAsync(pin + 16384 + slot * si, st, 64, hipMemcpyDeviceToHost, s));
      HIPCHK(hipEventRecord(ev[si], s));
    };
    auto final_info = [&](const MState &f, int j, yafel_hip_solve_info &out) {
      out.status = f.status[j];
      out.iterations = f.it[j];
      out.refinements = f.refinements;
      out.criterion = f.crit[j];
      out.rel_residual = f.bb[j] > 0 ? std::sqrt(f.true_rr[j] / f.bb[j]) : 0.0;
    };
    // Compaction: the batch keeps paying for every column until its last
    // right-hand side converges.  Once the active ones fit in three quarters
    // of the batch (rounded up to a multiple of 4), the columns that left are
    // finished (their info and x go to the caller: a right-hand side leaves
    // only at a refinement, after its x was folded into X), and every vector
    // that lives across iterations, X, B and the per-column state move to
    // the first columns of a narrower batch.  Each column's arithmetic is
    // unchanged; only the shared pause decisions see fewer columns.
    // Returns true when the state already says done.
    auto compact = [&]() -> bool {
      MState fs;
      HIPCHK(hipMemcpyAsync(&fs, st, sizeof(MState), hipMemcpyDeviceToHost, s));
      HIPCHK(hipStreamSynchronize(s));
      if (fs.done) return true;
      std::vector<int> cols;
      for (int j = 0; j < k; ++j)
        if (fs.active[j]) cols.push_back(j);
      const int kn = std::max(4, ((int)cols.size() + 3) & ~3);
      if (kn >= k) return false;
      const int nkeep = (int)cols.size();
      cols.resize(kn, -1);
      std::vector<int> out(k, -1), nmap(kn, -1);
      for (int j = 0; j < k; ++j)
        if (map[j] >= 0 && !fs.active[j]) {
          final_info(fs, j, info[map[j]]);
          out[j] = map[j];
        }
      for (int jj = 0; jj < kn; ++jj) nmap[jj] = cols[jj] >= 0 ? map[cols[jj]] : -1;
      if (colsd.bytes < (size_t)2 * KMAX * 4) colsd.alloc((size_t)2 * KMAX * 4);
      int *cd = colsd.as<int>(), *od = cd + KMAX;
      HIPCHK(hipMemcpyAsync(cd, cols.data(), kn * 4, hipMemcpyHostToDevice, s));
      HIPCHK(hipMemcpyAsync(od, out.data(), k * 4, hipMemcpyHostToDevice, s));
      // the finished columns' x to the caller (in place while X is the caller's)
      if (X != X0)
        k_put_cols<<<ceil_div((int64_t)n * k, NT), NT, 0, s>>>(n, k, k0, od, X, X0);
      if (tmp.bytes < (size_t)n * k * 8) tmp.alloc((size_t)n * k * 8);
      for (T *buf : {r, rhat, p0, p1, v, xl, z}) {
        k_take_cols<T><<<ceil_div((int64_t)n * kn, NT), NT, 0, s>>>(n, k, kn, cd, buf, tmp.as<T>());
        HIPCHK(hipMemcpyAsync(buf, tmp.p, (size_t)n * kn * sizeof(T), hipMemcpyDeviceToDevice, s));
      }
      DBuf nx, nb;
      nx.alloc((size_t)n * kn * 8);
      nb.alloc((size_t)n * kn * 8);
      k_take_cols<double><<<ceil_div((int64_t)n * kn, NT), NT, 0, s>>>(n, k, kn, cd, X, nx.as<double>());
      k_take_cols<double><<<ceil_div((int64_t)n * kn, NT), NT, 0, s>>>(n, k, kn, cd, B, nb.as<double>());
      HIPCHK(hipGetLastError());
      MState ns = fs;
      ns.k = kn;
      ns.nact = nkeep;
      for (int jj = 0; jj < kn; ++jj) {
        const int c = cols[jj];
        if (c >= 0) {
          ns.rho[jj] = fs.rho[c]; ns.rho_prev[jj] = fs.rho_prev[c]; ns.rho0[jj] = fs.rho0[c];
          ns.alpha[jj] = fs.alpha[c]; ns.omega[jj] = fs.omega[c]; ns.beta[jj] = fs.beta[c];
          ns.rr[jj] = fs.rr[c]; ns.rr_ref[jj] = fs.rr_ref[c]; ns.bb[jj] = fs.bb[c];
          ns.true_rr[jj] = fs.true_rr[c]; ns.crit[jj] = fs.crit[c];
          ns.active[jj] = fs.active[c]; ns.status[jj] = fs.status[c]; ns.want[jj] = fs.want[c];
          ns.it[jj] = fs.it[c];
        } else {
          ns.rho[jj] = ns.rho_prev[jj] = ns.rho0[jj] = ns.alpha[jj] = ns.omega[jj] = ns.beta[jj] = 0;
          ns.rr[jj] = ns.rr_ref[jj] = ns.bb[jj] = ns.true_rr[jj] = ns.crit[jj] = 0;
          ns.active[jj] = 0; ns.status[jj] = YAFEL_HIP_STATUS_CONVERGED; ns.want[jj] = 0; ns.it[jj] = 0;
        }
      }
      std::memcpy(ctx->pinned, &ns, sizeof(MState));
      HIPCHK(hipMemcpyAsync(st, ctx->pinned, sizeof(MState), hipMemcpyHostToDevice, s));
      HIPCHK(hipStreamSynchronize(s));
      Xc.swap(nx);
      Bc.swap(nb);
      X = Xc.as<double>();
      B = Bc.as<double>();
      map.swap(nmap);
      k = kn;
      return false;
    };
    chunk(0);
    chunk(1);
    int polled = 0;
    int64_t guard = 0;
    const auto t0 = std::chrono::steady_clock::now();
    while (true) {
      HIPCHK(hipEventSynchronize(ev[polled & 1]));
      const int *hs = (const int *)(pin + 16384 + slot * (polled & 1));
      if (hs[0]) break;
      const int nact = hs[8];  // MState::nact
      if (nact > 0 && k >= 8 && std::max(4, (nact + 3) & ~3) <= (3 * k) / 4) {
        HIPCHK(hipStreamSynchronize(s));  // the other chunk in flight has run as well
        if (compact()) break;
        chunk(0);
        chunk(1);
        polled = 0;
        continue;
      }
      chunk(polled & 1);
      ++polled;
      if (++guard > 2 * o.max_iter + 1000) fail(YAFEL_HIP_EDEVICE, "solve_multi: no progress");
      if ((guard & 63) == 0 && std::chrono::steady_clock::now() - t0 > std::chrono::seconds(900))
        fail(YAFEL_HIP_EDEVICE, "solve_multi: exceeded 900 s");
    }
    HIPCHK(hipStreamSynchronize(s));
    (void)hipEventDestroy(ev[0]);
    (void)hipEventDestroy(ev[1]);
    MReport rep = base(MReport{}, 7);
    rep.X = X;
    rep.B = B;
    spmm<double>(rep);
    MState fs;
    HIPCHK(hipMemcpyAsync(&fs, st, sizeof(MState), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    for (int j = 0; j < k; ++j)
      if (map[j] >= 0) final_info(fs, j, info[map[j]]);
    if (X != X0) {
      if (colsd.bytes < (size_t)2 * KMAX * 4) colsd.alloc((size_t)2 * KMAX * 4);
      HIPCHK(hipMemcpyAsync(colsd.p, map.data(), k * 4, hipMemcpyHostToDevice, s));
      k_put_cols<<<ceil_div((int64_t)n * k, NT), NT, 0, s>>>(n, k, k0, colsd.as<int>(), X, X0);
      HIPCHK(hipGetLastError());
      HIPCHK(hipStreamSynchronize(s));
    }
    k = k0;
  }
};

}  // namespace

namespace yh {
bool spmm_lds_launch(yafel_hip_matrix_s *A, int k, const float *X, float *Y, hipStream_t s) {
  // float4 staging reads need 16-byte aligned rows of X
  if (k < 1 || k > KMAX || k % 4 != 0 || ((uintptr_t)X % 16) != 0 || ((uintptr_t)Y % 16) != 0 || !staged32_ok(A))
    return false;
  MPlain op{X, Y, k, nullptr};
  launch_mspmm_lds(A, op, A->ctx->num_cus * 8, s);
  return true;
}

bool spmm_lds64_launch(yafel_hip_matrix_s *A, int k, const double *X, double *Y, hipStream_t s) {
  if (k < 1 || !staged64_ok(A)) return false;
  MPlain64 op{X, Y, k, nullptr};  // the kernel walks passes of 32 right-hand sides up to k
  launch_mspmm_lds64(A, op, A->ctx->num_cus * 8, s);
  return true;
}
}  // namespace yh

extern "C" int yafel_hip_solve_multi(yafel_hip_matrix A, const yafel_hip_solve_opts *opts, int k,
                                     const double *B_dev, double *X_dev, yafel_hip_solve_info *info) {
  YH_TRY_BEGIN
  REQUIRE(A && opts && info && k > 0, "solve_multi: bad argument");
  REQUIRE(A->format == YAFEL_HIP_CSR, "solve_multi: CSR matrices only");
  REQUIRE(A->rows == A->cols, "solve_multi: matrix must be square");
  REQUIRE(A->rows == 0 || (B_dev && X_dev), "solve_multi: null vector");
  REQUIRE(opts->method >= 0 && opts->method <= 3, "solve_multi: unknown method");
  REQUIRE(opts->precision == YAFEL_HIP_FP64 || opts->precision == YAFEL_HIP_MIXED,
          "solve_multi: precision must be FP64 or MIXED");
  REQUIRE(opts->stop == YAFEL_HIP_STOP_RELRES, "solve_multi: the multi-RHS solver stops on the relative residual");
  yafel_hip_ctx_s *ctx = A->ctx;
  HIPCHK(hipSetDevice(ctx->device));
  const int64_t n = A->rows;
  std::memset(info, 0, sizeof(*info) * k);
  if (n == 0) return YAFEL_HIP_OK;
  REQUIRE(A->has64, "solve_multi: matrix has no FP64 values");
  if (!A->diag_ok) matrix_build_diag(A);
  const bool mixed = opts->precision == YAFEL_HIP_MIXED;
  if (mixed && !A->has32) matrix_build_fp32(A);
  const bool staged = mixed && staged32_ok(A);  // the FP32 products stage X in LDS
  const bool staged64 = staged64_ok(A);        // and the FP64 ones
  HIPCHK(hipEventRecord(ctx->t0, ctx->stream));
  // batches of up to 64 right-hand sides, each a contiguous column block;
  // the batch is copied into a packed n x kb buffer and back
  for (int j0 = 0; j0 < k; j0 += KMAX) {
    const int kb = k - j0 < KMAX ? k - j0 : KMAX;
    Multi M{};
    M.ctx = ctx;
    M.A = A;
    M.o = *opts;
    M.n = n;
    M.k = kb;
    M.mixed = mixed;
    M.staged = staged;
    M.staged64 = staged64;
    M.cgfam = opts->method == YAFEL_HIP_CG || opts->method == YAFEL_HIP_PCG_JACOBI;
    M.prec = opts->method == YAFEL_HIP_PCG_JACOBI || opts->method == YAFEL_HIP_BICGSTAB_JACOBI;
    if (!M.prec && !M.cgfam) fail(YAFEL_HIP_EINVAL, "solve_multi: BiCGSTAB runs with right Jacobi preconditioning only");
    if (M.o.max_iter <= 0) M.o.max_iter = M.cgfam ? n : 10 * n;
    if (M.o.refine_delta <= 0) M.o.refine_delta = 0.1;
    M.s = ctx->stream;
    M.grid = grid_for(ctx, n * 64, NT);
    M.partbuf.alloc((size_t)4 * M.grid * 64 * sizeof(double) + 4096);
    M.stbuf.alloc(sizeof(MState) + 256);
    M.part = M.partbuf.as<double>();
    M.st = M.stbuf.as<MState>();
    DBuf Bb, Xb;
    const double *B = B_dev;
    double *X = X_dev;
    if (kb != k) {
      Bb.alloc((size_t)n * kb * 8);
      Xb.alloc((size_t)n * kb * 8);
      HIPCHK(hipMemcpy2DAsync(Bb.p, kb * 8, B_dev + j0, (size_t)k * 8, kb * 8, n, hipMemcpyDeviceToDevice, ctx->stream));
      HIPCHK(hipMemcpy2DAsync(Xb.p, kb * 8, X_dev + j0, (size_t)k * 8, kb * 8, n, hipMemcpyDeviceToDevice, ctx->stream));
      B = Bb.as<double>();
      X = Xb.as<double>();
    }
    if (mixed) M.run<float>(B, X, info + j0);
    else M.run<double>(B, X, info + j0);
    if (kb != k)
      HIPCHK(hipMemcpy2DAsync(X_dev + j0, (size_t)k * 8, Xb.p, kb * 8, kb * 8, n, hipMemcpyDeviceToDevice, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
  }
  HIPCHK(hipEventRecord(ctx->t1, ctx->stream));
  HIPCHK(hipEventSynchronize(ctx->t1));
  float ms = 0;
  HIPCHK(hipEventElapsedTime(&ms, ctx->t0, ctx->t1));
  for (int j = 0; j < k; ++j) info[j].ms = ms;
  return YAFEL_HIP_OK;
  YH_TRY_END(A ? A->ctx : nullptr)
}
