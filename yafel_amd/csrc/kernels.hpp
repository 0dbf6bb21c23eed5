// Device building blocks shared by the SpMV and Krylov kernels: the sliced
// (SELL-64) matrix view and its per-row products.
#pragma once
#include "internal.hpp"

namespace yh {

// Slots per step of the sliced row loops inside the fused Krylov kernels.
constexpr int kSellUnroll = 4;

template <class T>
struct SellView {
  const int *sptr;   // slot offset of each slice (nsl + 1)
  const int *perm;   // stored position -> row, nullptr for the identity
  const int *col;    // sliced block columns
  const T *val;      // sliced values (BCSR3: 9 arrays of ntot)
  int64_t brows;     // stored rows (block rows for BCSR3)
  int64_t ntot;      // sliced entries per value array
  const short *col16;  // column - stored position, 16-bit (nullptr: use col)
  int keep;            // slices starting below this slot read with the caching policy
  __device__ __forceinline__ int64_t nsl() const { return (brows + 63) >> 6; }
  __device__ __forceinline__ int64_t row_of(int64_t i) const { return perm ? (int64_t)perm[i] : i; }
};

namespace dev {

// A load the compiler issues where it is written: a relaxed wavefront-scope
// atomic load is a plain global load on the device, but an ordered memory
// access is not sunk past a later branch (a load issued before a test keeps
// its latency overlapped with the test's own load).
template <class T>
__device__ __forceinline__ T load_in_place(const T *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
}

// Matrix streams.  Past a byte budget (SellView::keep, in slots) they are
// loaded non-temporally: such loads do not allocate in the Infinity Cache,
// so a stream larger than the cache does not evict itself and everything else
// on every pass.  The leading slots within the budget are loaded with the
// default policy and stay cache-resident across the products of a solve.
template <bool KEEP, class T>
__device__ __forceinline__ T ld_stream(const T *p) {
  if constexpr (KEEP) return *p;
  else return __builtin_nontemporal_load(p);
}
template <class T>
__device__ __forceinline__ bool keep_slice(const SellView<T> &A, int p0) {
  return p0 < A.keep;
}

// Sliced position of entry t of row r (iperm may be nullptr: identity).
__device__ __forceinline__ int64_t sell_pos(const int *__restrict__ sptr, const int *__restrict__ iperm,
                                            int64_t r, int64_t t) {
  const int64_t i = iperm ? (int64_t)iperm[r] : r;
  return ((int64_t)sptr[i >> 6] + t) * 64 + (i & 63);
}

// One CSR row per lane: the lane at `lane` of slice s sums its row's slots in
// order (padding contributes exact zeros), U slots per step with every load
// of the step issued before the first FMA.  Matrix values and columns are
// streamed once, so they are loaded non-temporally; the gathered vector stays
// in cache.  xf(c) returns x[c].  T is the stored value type, TV the
// accumulation type.  Columns come from the 16-bit offsets to the stored
// position when the matrix has them (2 bytes an entry instead of 4), else
// from the int32 array; cf(k) is the column of entry k.
// The steps are software-pipelined: a step's gathers are issued, then the
// next step's columns and values, and only then does the lane wait for the
// gathers, so the gathers' cache latency overlaps the next step's HBM loads
// (vector loads complete in issue order, so the order of issue matters).
// p0, p1: the slice's slot range (sptr[s], sptr[s + 1]).
template <int U, class T, class TV, bool KEEP, class XF, class CF>
__device__ __forceinline__ TV sell_row_csr_cols(const SellView<T> &A, int p0, int p1, int lane, XF xf, CF cf) {
  const int64_t base = (int64_t)p0 * 64 + lane;
  const int len = p1 - p0;
  TV acc = 0;
  int t = 0;
  if (U <= len) {
    int c[U];
    T v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int64_t k = base + (int64_t)u * 64;
      c[u] = cf(k);
      v[u] = ld_stream<KEEP>(&A.val[k]);
    }
    for (;;) {
      TV xg[U];
#pragma unroll
      for (int u = 0; u < U; ++u) xg[u] = xf(c[u]);
      const bool more = t + 2 * U <= len;  // uniform over the slice
      T vn[U];
      if (more) {
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const int64_t k = base + (int64_t)(t + U + u) * 64;
          c[u] = cf(k);
          vn[u] = ld_stream<KEEP>(&A.val[k]);
        }
      }
#pragma unroll
      for (int u = 0; u < U; ++u) acc = fma((TV)v[u], xg[u], acc);
      t += U;
      if (!more) break;
#pragma unroll
      for (int u = 0; u < U; ++u) v[u] = vn[u];
    }
  }
  for (; t < len; ++t) {
    const int64_t k = base + (int64_t)t * 64;
    acc = fma((TV)ld_stream<KEEP>(&A.val[k]), xf(cf(k)), acc);
  }
  return acc;
}

// Row sum of slice s's row at `lane`, its slot range already loaded.
template <int U, class T, class TV, bool KEEP, class XF>
__device__ __forceinline__ TV sell_row_csr_keep(const SellView<T> &A, int64_t s, int p0, int p1, int lane, XF xf) {
  if (A.col16) {
    const int i = (int)(s * 64 + lane);
    return sell_row_csr_cols<U, T, TV, KEEP>(A, p0, p1, lane, xf,
                                             [&](int64_t k) { return i + (int)ld_stream<KEEP>(&A.col16[k]); });
  }
  return sell_row_csr_cols<U, T, TV, KEEP>(A, p0, p1, lane, xf,
                                           [&](int64_t k) { return ld_stream<KEEP>(&A.col[k]); });
}

template <int U, class T, class TV, class XF>
__device__ __forceinline__ TV sell_row_csr_at(const SellView<T> &A, int64_t s, int p0, int p1, int lane, XF xf) {
  if (keep_slice(A, p0)) return sell_row_csr_keep<U, T, TV, true>(A, s, p0, p1, lane, xf);
  return sell_row_csr_keep<U, T, TV, false>(A, s, p0, p1, lane, xf);
}

template <int U, class T, class TV, class XF>
__device__ __forceinline__ TV sell_row_csr(const SellView<T> &A, int64_t s, int lane, XF xf) {
  return sell_row_csr_at<U, T, TV>(A, s, A.sptr[s], A.sptr[s + 1], lane, xf);
}

// One 3x3-block row per lane; a[0..2] receive the three row sums.  Blocks are
// summed in order, entries of a block row-wise (j = 0, 1, 2).  xf(i) returns
// x[i] for a scalar index; cf(k) is the block column of entry k.
template <int U, class T, class TV, bool KEEP, class XF, class CF>
__device__ __forceinline__ void sell_row_bcsr_cols(const SellView<T> &A, int p0, int p1, int lane, TV a[3], XF xf,
                                                   CF cf) {
  const int64_t base = (int64_t)p0 * 64 + lane;
  const int len = p1 - p0;
  const int64_t nt = A.ntot;
  TV a0 = 0, a1 = 0, a2 = 0;
  int t = 0;
  for (; t + U <= len; t += U) {
    int64_t c[U];
    T v[U][9];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int64_t k = base + (int64_t)(t + u) * 64;
      c[u] = 3 * (int64_t)cf(k);
#pragma unroll
      for (int e = 0; e < 9; ++e) v[u][e] = ld_stream<KEEP>(&A.val[e * nt + k]);
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const TV x0 = xf(c[u]), x1 = xf(c[u] + 1), x2 = xf(c[u] + 2);
      a0 = fma((TV)v[u][0], x0, a0); a0 = fma((TV)v[u][1], x1, a0); a0 = fma((TV)v[u][2], x2, a0);
      a1 = fma((TV)v[u][3], x0, a1); a1 = fma((TV)v[u][4], x1, a1); a1 = fma((TV)v[u][5], x2, a1);
      a2 = fma((TV)v[u][6], x0, a2); a2 = fma((TV)v[u][7], x1, a2); a2 = fma((TV)v[u][8], x2, a2);
    }
  }
  for (; t < len; ++t) {
    const int64_t k = base + (int64_t)t * 64;
    const int64_t c = 3 * (int64_t)cf(k);
    const TV x0 = xf(c), x1 = xf(c + 1), x2 = xf(c + 2);
    a0 = fma((TV)ld_stream<KEEP>(&A.val[0 * nt + k]), x0, a0);
    a0 = fma((TV)ld_stream<KEEP>(&A.val[1 * nt + k]), x1, a0);
    a0 = fma((TV)ld_stream<KEEP>(&A.val[2 * nt + k]), x2, a0);
    a1 = fma((TV)ld_stream<KEEP>(&A.val[3 * nt + k]), x0, a1);
    a1 = fma((TV)ld_stream<KEEP>(&A.val[4 * nt + k]), x1, a1);
    a1 = fma((TV)ld_stream<KEEP>(&A.val[5 * nt + k]), x2, a1);
    a2 = fma((TV)ld_stream<KEEP>(&A.val[6 * nt + k]), x0, a2);
    a2 = fma((TV)ld_stream<KEEP>(&A.val[7 * nt + k]), x1, a2);
    a2 = fma((TV)ld_stream<KEEP>(&A.val[8 * nt + k]), x2, a2);
  }
  a[0] = a0;
  a[1] = a1;
  a[2] = a2;
}

template <int U, class T, class TV, bool KEEP, class XF>
__device__ __forceinline__ void sell_row_bcsr_keep(const SellView<T> &A, int64_t s, int p0, int p1, int lane,
                                                   TV a[3], XF xf) {
  if (A.col16) {
    const int i = (int)(s * 64 + lane);
    sell_row_bcsr_cols<U, T, TV, KEEP>(A, p0, p1, lane, a, xf,
                                       [&](int64_t k) { return i + (int)ld_stream<KEEP>(&A.col16[k]); });
  } else {
    sell_row_bcsr_cols<U, T, TV, KEEP>(A, p0, p1, lane, a, xf,
                                       [&](int64_t k) { return ld_stream<KEEP>(&A.col[k]); });
  }
}

template <int U, class T, class TV, class XF>
__device__ __forceinline__ void sell_row_bcsr_at(const SellView<T> &A, int64_t s, int p0, int p1, int lane, TV a[3],
                                                 XF xf) {
  if (keep_slice(A, p0)) sell_row_bcsr_keep<U, T, TV, true>(A, s, p0, p1, lane, a, xf);
  else sell_row_bcsr_keep<U, T, TV, false>(A, s, p0, p1, lane, a, xf);
}

template <int U, class T, class TV, class XF>
__device__ __forceinline__ void sell_row_bcsr(const SellView<T> &A, int64_t s, int lane, TV a[3], XF xf) {
  sell_row_bcsr_at<U, T, TV>(A, s, A.sptr[s], A.sptr[s + 1], lane, a, xf);
}

// Value of `v` in lane l (l wave-uniform), for 32- and 64-bit types.
__device__ __forceinline__ int lane_bcast(int v, int l) { return __builtin_amdgcn_readlane(v, l); }
__device__ __forceinline__ float lane_bcast(float v, int l) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}
__device__ __forceinline__ double lane_bcast(double v, int l) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)b, l), hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Multi-RHS rows of one slice: rows h*RB .. h*RB+RB-1 of slice s, with the
// wavefront's lanes as right-hand sides.  For each slot the 64 lanes load the
// slice's 64 (column, value) pairs in one coalesced step (the next slot's
// pairs are loaded before this slot's gathers are issued, so only the gathers
// wait on memory); each of the RB rows' pair is then broadcast (readlane) and
// every lane gathers its right-hand side's x at that column, so one slot
// issues RB independent 64-wide gathers.  y[r] accumulates row h*RB + r in
// slot order (padding adds exact zeros).  xf(c) returns this lane's x at
// column c and must be safe to call in every lane (callers clamp the lanes
// beyond their right-hand sides to a valid one); those lanes, `on` false,
// accumulate zeros.
// With a column scale cs (nullptr: none) the product is A diag(cs) x, the
// scale applied to each value as it is loaded (one gather per slot, not per
// row).
template <int RB, class T, class XF>
__device__ __forceinline__ void sell_rows_multi(const int *__restrict__ sptr, const int *__restrict__ col,
                                                const T *__restrict__ val, const T *__restrict__ cs, int64_t s,
                                                int h, int lane, bool on, XF xf, T (&y)[RB]) {
  const int p0 = sptr[s], ns = sptr[s + 1] - p0;
#pragma unroll
  for (int r = 0; r < RB; ++r) y[r] = 0;
  if (ns == 0) return;
  int64_t q = (int64_t)p0 * 64 + lane;
  int cl = __builtin_nontemporal_load(&col[q]);
  T vl = __builtin_nontemporal_load(&val[q]);
  if (cs) vl *= cs[cl];
  for (int t = 0; t < ns; ++t) {
    int cn = 0;
    T vn = 0;
    if (t + 1 < ns) {
      q += 64;
      cn = __builtin_nontemporal_load(&col[q]);
      vn = __builtin_nontemporal_load(&val[q]);
      if (cs) vn *= cs[cn];
    }
#pragma unroll
    for (int r = 0; r < RB; ++r) {
      const int c = lane_bcast(cl, h * RB + r);
      const T v = lane_bcast(vl, h * RB + r);
      const T x = xf(c);
      y[r] = fma(v, on ? x : T(0), y[r]);
    }
    cl = cn;
    vl = vn;
  }
}
// FP32 multi-RHS rows of one slice, four rows per gather instruction: lane
// (rs = lane >> 4, g = lane & 15) takes right-hand sides 4g .. 4g+3 of rows
// h*RB + 4i + rs, so a 16-byte gather per lane covers four rows' 256-byte X
// rows at once.  Entries and their order as in sell_rows_multi; xf4(c) returns
// this lane's four x values at column c and must be safe in every lane.
template <int RB, class XF4>
__device__ __forceinline__ void sell_rows_multi4(const int *__restrict__ sptr, const int *__restrict__ col,
                                                 const float *__restrict__ val, const float *__restrict__ cs,
                                                 int64_t s, int h, int lane, bool on, XF4 xf4,
                                                 float4 (&y)[RB / 4]) {
  const int rs = lane >> 4;
  const int p0 = sptr[s], ns = sptr[s + 1] - p0;
#pragma unroll
  for (int i = 0; i < RB / 4; ++i) y[i] = make_float4(0.f, 0.f, 0.f, 0.f);
  if (ns == 0) return;
  int64_t q = (int64_t)p0 * 64 + lane;
  int cl = __builtin_nontemporal_load(&col[q]);
  float vl = __builtin_nontemporal_load(&val[q]);
  if (cs) vl *= cs[cl];
  for (int t = 0; t < ns; ++t) {
    int cn = 0;
    float vn = 0;
    if (t + 1 < ns) {
      q += 64;
      cn = __builtin_nontemporal_load(&col[q]);
      vn = __builtin_nontemporal_load(&val[q]);
      if (cs) vn *= cs[cn];
    }
#pragma unroll
    for (int i = 0; i < RB / 4; ++i) {
      const int src = h * RB + 4 * i + rs;
      const int c = __shfl(cl, src, 64);
      const float v = __shfl(vl, src, 64);
      float4 x = xf4(c);
      if (!on) x = make_float4(0.f, 0.f, 0.f, 0.f);
      y[i].x = fmaf(v, x.x, y[i].x);
      y[i].y = fmaf(v, x.y, y[i].y);
      y[i].z = fmaf(v, x.z, y[i].z);
      y[i].w = fmaf(v, x.w, y[i].w);
    }
    cl = cn;
    vl = vn;
  }
}

}  // namespace dev
}  // namespace yh
