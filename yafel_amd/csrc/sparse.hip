// Device sparse matrices (CSR and 3x3-block BCSR), SpMV and multi-RHS SpMM.
//
// Device layout: both formats are stored sliced (SELL-64, kernels.hpp): the
// stored rows (block rows for BCSR3) are cut into slices of 64, one per
// wavefront, each lane owns a row and walks its entries slot by slot, so the
// matrix streams are read with unit stride across the lanes and the gathered
// vector entries of neighbouring rows are neighbours too.  BCSR3 values are
// entry-major, v[e * ntot + k] for e = 3 i + j.  The reference's host layout
// (row_ptr, col_index, data[b * 9 + i * 3 + j]) is converted on upload and
// download.
#include <algorithm>
#include <cmath>

#include "internal.hpp"
#include "kernels.hpp"
#include "sortprim.hpp"

using namespace yh;

namespace {

constexpr int kSortWindow = 256;  // rows sorted by length within windows of this size

__global__ void k_aos_to_soa(const double *__restrict__ aos, double *__restrict__ soa, int64_t nb) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < nb * 9; i += stride) {
    int64_t b = i / 9;
    int e = (int)(i - b * 9);
    soa[(int64_t)e * nb + b] = aos[i];
  }
}

__global__ void k_to_fp32(const double *__restrict__ a, float *__restrict__ b, int64_t n) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) b[i] = (float)a[i];
}

// ------------------------------------------------------------ slicing
// Longest row of each slice with rows in their natural order.
__global__ void k_slice_len(int64_t brows, int64_t nsl, const int *__restrict__ ptr, int64_t *slen) {
  const int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (s > nsl) return;
  int64_t m = 0;
  if (s < nsl) {
    const int64_t r1 = (s + 1) * 64 < brows ? (s + 1) * 64 : brows;
    for (int64_t r = s * 64; r < r1; ++r) m = max(m, (int64_t)(ptr[r + 1] - ptr[r]));
  }
  slen[s] = m;  // slen[nsl] = 0 closes the scan
}

// Stable sort of the rows of each window of kSortWindow rows by decreasing
// length (rank counting, one workgroup per window); writes perm and the
// longest row of each slice in sorted order.
__global__ __launch_bounds__(kSortWindow) void k_window_sort(int64_t brows, int64_t nsl, const int *__restrict__ ptr,
                                                             int *__restrict__ perm, int64_t *__restrict__ slen) {
  __shared__ int len[kSortWindow];
  __shared__ int slen_l[kSortWindow];
  const int64_t w0 = (int64_t)blockIdx.x * kSortWindow;
  const int n = (int)min((int64_t)kSortWindow, brows - w0);
  const int t = threadIdx.x;
  const int mine = t < n ? ptr[w0 + t + 1] - ptr[w0 + t] : -1;
  len[t] = mine;
  __syncthreads();
  if (t < n) {
    int rank = 0;
    for (int j = 0; j < n; ++j) rank += (len[j] > mine) || (len[j] == mine && j < t);
    perm[w0 + rank] = (int)(w0 + t);
    slen_l[rank] = mine;
  }
  __syncthreads();
  if (t < kSortWindow / 64) {
    const int64_t s = w0 / 64 + t;
    if (s < nsl) slen[s] = t * 64 < n ? slen_l[t * 64] : 0;
  }
  if (blockIdx.x == 0 && t == 0) slen[nsl] = 0;
}

__global__ void k_slots_to_i32(int64_t n, const int64_t *__restrict__ a, int *__restrict__ b) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) b[i] = (int)a[i];
}

__global__ void k_inverse(int64_t n, const int *__restrict__ perm, int *__restrict__ iperm) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) iperm[perm[i]] = (int)i;
}

// Scatter CSR-ordered columns/values into the sliced arrays; one thread per
// stored position writes every slot of its lane (real entries, then padding).
__global__ void k_sell_fill(int64_t brows, int64_t nsl, int64_t nblk, int64_t ntot, int nval,
                            const int *__restrict__ ptr, const int *__restrict__ col,
                            const double *__restrict__ val, const int *__restrict__ perm,
                            const int *__restrict__ sptr, int *__restrict__ scol, double *__restrict__ sval) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= nsl * 64) return;
  const int64_t s = i >> 6;
  const int lane = (int)(i & 63);
  const int p0 = sptr[s], ns = sptr[s + 1] - p0;
  int b = 0, len = 0, padc = 0;
  if (i < brows) {
    const int64_t r = perm ? perm[i] : i;
    b = ptr[r];
    len = ptr[r + 1] - b;
    padc = len > 0 ? col[b + len - 1] : 0;
  }
  for (int t = 0; t < ns; ++t) {
    const int64_t k = ((int64_t)p0 + t) * 64 + lane;
    const bool real = t < len;
    scol[k] = real ? col[b + t] : padc;
    for (int e = 0; e < nval; ++e) sval[e * ntot + k] = real ? val[e * nblk + b + t] : 0.0;
  }
}

// 16-bit column offsets: entry k of stored position i holds col - i.  The
// rows of a window of 256 are sorted among themselves only, so a banded
// pattern (a mesh numbering) keeps every offset small.  Lanes past the last
// row point at the last column (their values are zeros).  *bad is set when an
// offset does not fit; the matrix then keeps using the int32 columns.
__global__ void k_sell_col16(int64_t brows, int64_t bcols, int64_t nsl, const int *__restrict__ sptr,
                             const int *__restrict__ scol, short *__restrict__ c16, int *bad) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= nsl * 64) return;
  const int64_t s = i >> 6;
  const int lane = (int)(i & 63);
  const int p0 = sptr[s], ns = sptr[s + 1] - p0;
  for (int t = 0; t < ns; ++t) {
    const int64_t k = ((int64_t)p0 + t) * 64 + lane;
    const int64_t c = i < brows ? (int64_t)scol[k] : bcols - 1;
    int64_t d = c - i;
    if (d < -32768 || d > 32767) {
      atomicOr(bad, 1);
      d = 0;
    }
    c16[k] = (short)d;
  }
}

// Sliced storage back to the reference's host layout (CSR order; BCSR3
// blocks as data[b * 9 + e]).
__global__ void k_sell_to_csr(int64_t brows, int64_t ntot, int nval, const int *__restrict__ ptr,
                              const int *__restrict__ sptr, const int *__restrict__ iperm,
                              const int *__restrict__ scol, const double *__restrict__ sval,
                              int *__restrict__ col, double *__restrict__ val) {
  const int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (r >= brows) return;
  const int b = ptr[r], len = ptr[r + 1] - b;
  for (int t = 0; t < len; ++t) {
    const int64_t k = dev::sell_pos(sptr, iperm, r, t);
    if (col) col[b + t] = scol[k];
    if (val)
      for (int e = 0; e < nval; ++e) val[(int64_t)(b + t) * nval + e] = sval[e * ntot + k];
  }
}

// Diagonal and its inverse (JacobiPreconditioner.hpp:124-137: 1/a_ii, kept
// even when a_ii == 0 -- the reference only warns).
__global__ void k_diag_csr(int64_t rows, const int *__restrict__ ptr, const int *__restrict__ sptr,
                           const int *__restrict__ iperm, const int *__restrict__ col,
                           const double *__restrict__ val, double *diag, double *dinv, float *dinv32) {
  int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (r >= rows) return;
  double d = 0;
  const int len = ptr[r + 1] - ptr[r];
  for (int t = 0; t < len; ++t) {
    const int64_t k = dev::sell_pos(sptr, iperm, r, t);
    if (col[k] == r) { d = val[k]; break; }
  }
  diag[r] = d;
  dinv[r] = 1.0 / d;
  dinv32[r] = (float)(1.0 / d);
}

__global__ void k_diag_bcsr(int64_t brows, int64_t ntot, const int *__restrict__ ptr, const int *__restrict__ sptr,
                            const int *__restrict__ iperm, const int *__restrict__ col,
                            const double *__restrict__ val, double *diag, double *dinv, float *dinv32) {
  int64_t br = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (br >= brows) return;
  double d[3] = {0, 0, 0};
  const int len = ptr[br + 1] - ptr[br];
  for (int t = 0; t < len; ++t) {
    const int64_t k = dev::sell_pos(sptr, iperm, br, t);
    if (col[k] == br) {
      for (int i = 0; i < 3; ++i) d[i] = val[(int64_t)(4 * i) * ntot + k];
      break;
    }
  }
  for (int i = 0; i < 3; ++i) {
    diag[3 * br + i] = d[i];
    dinv[3 * br + i] = 1.0 / d[i];
    dinv32[3 * br + i] = (float)(1.0 / d[i]);
  }
}

// The FP32 copy with each row's rounding errors folded into its diagonal
// (block): per column component, row r of the copy then sums to the FP64
// row's sum up to the rounding of one diagonal entry.  For a smooth vector v,
// (A v)_i = v_i sum_j a_ij + terms in the variation of v, so the rounding
// error of A acts on the smooth (low) modes, which set the iteration count,
// mainly through the row sums; keeping those keeps the FP32 operator's low
// modes close to the FP64 one's.  It pays on dense rows (high-order
// elements, 3x3 blocks), where large entries of both signs cancel.  B = 3:
// BCSR3 blocks; B = 1: CSR.
template <int B>
__global__ void k_fp32_rowsum(int64_t rows, int64_t nsl, int64_t ntot, const int *__restrict__ ptr,
                              const int *__restrict__ sptr, const int *__restrict__ perm,
                              const int *__restrict__ col, const double *__restrict__ val64,
                              float *__restrict__ val32) {
  // the conversion and the fold in one pass: a thread per stored row (slice
  // s, lane l) converts every slot of its lane, padding included, summing
  // its row's rounding errors in slot order
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t s = i >> 6;
  if (s >= nsl) return;
  const int64_t r = i < rows ? (perm ? (int64_t)perm[i] : i) : -1;
  const int len = r >= 0 ? ptr[r + 1] - ptr[r] : 0;
  const int p0 = sptr[s], ns = sptr[s + 1] - p0;
  double d[B * B];
#pragma unroll
  for (int e = 0; e < B * B; ++e) d[e] = 0;
  int64_t kd = -1;
  for (int t = 0; t < ns; ++t) {
    const int64_t k = (int64_t)(p0 + t) * 64 + (i & 63);
    const bool live = t < len;
    if (live && col[k] == r) kd = k;
#pragma unroll
    for (int e = 0; e < B * B; ++e) {
      const double a = val64[(int64_t)e * ntot + k];
      const float a32 = (float)a;
      val32[(int64_t)e * ntot + k] = a32;
      if (live) d[e] += a - (double)a32;  // exact: a32 is a rounded
    }
  }
  if (kd < 0) return;
#pragma unroll
  for (int e = 0; e < B * B; ++e) val32[(int64_t)e * ntot + kd] = (float)((double)val32[(int64_t)e * ntot + kd] + d[e]);
}

// ---------------------------------------------------------------- SpMV
// One slice per wavefront (grid-stride over slices when the grid is capped).
// Row sums accumulate in FP64 for either storage precision; the FP32 product
// is rounded once, at the store.
template <int U, class T>
__global__ __launch_bounds__(256) void k_sell_spmv_csr(SellView<T> A, const T *__restrict__ x, T *__restrict__ y) {
  const int lane = threadIdx.x & 63;
  const int64_t nsl = A.nsl(), nw = (int64_t)gridDim.x * 4;
  for (int64_t s = (int64_t)dev::xcd_block() * 4 + (threadIdx.x >> 6); s < nsl; s += nw) {
    const double acc = dev::sell_row_csr<U, T, double>(A, s, lane, [&](int c) { return (double)x[c]; });
    const int64_t i = s * 64 + lane;
    if (i < A.brows) y[A.row_of(i)] = (T)acc;
  }
}

template <int U, class T>
__global__ __launch_bounds__(256) void k_sell_spmv_bcsr(SellView<T> A, const T *__restrict__ x, T *__restrict__ y) {
  const int lane = threadIdx.x & 63;
  const int64_t nsl = A.nsl(), nw = (int64_t)gridDim.x * 4;
  for (int64_t s = (int64_t)dev::xcd_block() * 4 + (threadIdx.x >> 6); s < nsl; s += nw) {
    double a[3];
    dev::sell_row_bcsr<U, T, double>(A, s, lane, a, [&](int64_t c) { return (double)x[c]; });
    const int64_t i = s * 64 + lane;
    if (i < A.brows) {
      const int64_t r = A.row_of(i);
      y[3 * r] = (T)a[0];
      y[3 * r + 1] = (T)a[1];
      y[3 * r + 2] = (T)a[2];
    }
  }
}

// Y = A X for k right-hand sides, row-major X (n x k).  A wavefront takes RB
// rows of a slice (two passes of 32 rows for FP64, whose accumulators take
// twice the registers) and 64 right-hand sides at a time: the slice's
// columns and values are read once per slot, coalesced, and each row's X
// gather is a 64-wide coalesced read.
template <class T>
__global__ __launch_bounds__(256) void k_csr_spmm(int64_t rows, int k, const int *__restrict__ sptr,
                                                  const int *__restrict__ perm, const int *__restrict__ col,
                                                  const T *__restrict__ val, const T *__restrict__ X,
                                                  T *__restrict__ Y) {
  constexpr int RB = sizeof(T) == 4 ? 64 : 32, H = 64 / RB;
  const int lane = threadIdx.x & 63;
  const int64_t units = ((rows + 63) >> 6) * H;
  const int64_t nw = (int64_t)gridDim.x * 4;
  for (int64_t u = (int64_t)dev::xcd_block() * 4 + (threadIdx.x >> 6); u < units; u += nw) {
    const int64_t s = u / H;
    const int h = (int)(u - s * H);
    for (int j0 = 0; j0 < k; j0 += 64) {
      const int j = j0 + lane, jj = j < k ? j : j0;
      T y[RB];
      dev::sell_rows_multi<RB, T>(sptr, col, val, nullptr, s, h, lane, j < k,
                                  [&](int c) { return X[(int64_t)c * k + jj]; }, y);
#pragma unroll
      for (int r = 0; r < RB; ++r) {
        const int64_t i = s * 64 + h * RB + r;
        if (i < rows && j < k) Y[(perm ? (int64_t)perm[i] : i) * k + j] = y[r];
      }
    }
  }
}

// FP32 with k a multiple of 4: RB = 32 rows of a slice per wavefront, four
// rows per 16-byte gather (sell_rows_multi4).
__global__ __launch_bounds__(256) void k_csr_spmm4(int64_t rows, int k, const int *__restrict__ sptr,
                                                   const int *__restrict__ perm, const int *__restrict__ col,
                                                   const float *__restrict__ val, const float *__restrict__ X,
                                                   float *__restrict__ Y) {
  constexpr int RB = 32, H = 64 / RB;
  const int lane = threadIdx.x & 63, g = lane & 15, rs = lane >> 4;
  const int k4 = k >> 2;
  const float4 *X4 = reinterpret_cast<const float4 *>(X);
  float4 *Y4 = reinterpret_cast<float4 *>(Y);
  const int64_t units = ((rows + 63) >> 6) * H;
  const int64_t nw = (int64_t)gridDim.x * 4;
  for (int64_t u = (int64_t)dev::xcd_block() * 4 + (threadIdx.x >> 6); u < units; u += nw) {
    const int64_t s = u / H;
    const int h = (int)(u - s * H);
    for (int j0 = 0; j0 < k4; j0 += 16) {
      const bool on = j0 + g < k4;
      const int jj = on ? j0 + g : j0;
      float4 y[RB / 4];
      dev::sell_rows_multi4<RB>(sptr, col, val, nullptr, s, h, lane, on,
                                [&](int c) { return X4[(int64_t)c * k4 + jj]; }, y);
#pragma unroll
      for (int i = 0; i < RB / 4; ++i) {
        const int64_t r = s * 64 + h * RB + 4 * i + rs;
        if (r < rows && on) Y4[(perm ? (int64_t)perm[r] : r) * k4 + jj] = y[i];
      }
    }
  }
}

template <class T>
__global__ __launch_bounds__(256) void k_bcsr_spmm(int64_t brows, int64_t ntot, int k, const int *__restrict__ ptr,
                                                   const int *__restrict__ sptr, const int *__restrict__ iperm,
                                                   const int *__restrict__ col, const T *__restrict__ val,
                                                   const T *__restrict__ X, T *__restrict__ Y) {
  const int lane = threadIdx.x & 63;
  const int64_t chunk = (brows + gridDim.x - 1) / gridDim.x;
  const int64_t r0 = (int64_t)dev::xcd_block() * chunk;
  const int64_t r1 = r0 + chunk < brows ? r0 + chunk : brows;
  for (int64_t rr = r0 + threadIdx.x / 64; rr < r1; rr += 4) {
    const int br = __builtin_amdgcn_readfirstlane((int)rr);
    const int len = ptr[br + 1] - ptr[br];
    const int64_t base = dev::sell_pos(sptr, iperm, br, 0);
    for (int j0 = 0; j0 < k; j0 += 64) {
      const int j = j0 + lane;
      T a0 = 0, a1 = 0, a2 = 0;
      if (j < k)
        for (int t = 0; t < len; ++t) {
          const int64_t q = base + (int64_t)t * 64;
          const int64_t c = col[q];
          const T x0 = X[(3 * c + 0) * k + j], x1 = X[(3 * c + 1) * k + j], x2 = X[(3 * c + 2) * k + j];
          a0 = fma(val[0 * ntot + q], x0, a0); a0 = fma(val[1 * ntot + q], x1, a0); a0 = fma(val[2 * ntot + q], x2, a0);
          a1 = fma(val[3 * ntot + q], x0, a1); a1 = fma(val[4 * ntot + q], x1, a1); a1 = fma(val[5 * ntot + q], x2, a1);
          a2 = fma(val[6 * ntot + q], x0, a2); a2 = fma(val[7 * ntot + q], x1, a2); a2 = fma(val[8 * ntot + q], x2, a2);
        }
      if (j < k) {
        Y[(3 * (int64_t)br + 0) * k + j] = a0;
        Y[(3 * (int64_t)br + 1) * k + j] = a1;
        Y[(3 * (int64_t)br + 2) * k + j] = a2;
      }
    }
  }
}

std::vector<int32_t> to_i32(const int64_t *a, int64_t n, int64_t limit, const char *what) {
  std::vector<int32_t> o(n);
  for (int64_t i = 0; i < n; ++i) {
    if (a[i] < 0 || a[i] > limit)
      fail(YAFEL_HIP_EINVAL, std::string(what) + ": index out of range");
    o[i] = (int32_t)a[i];
  }
  return o;
}

template <class T>
SellView<T> sell_view(yafel_hip_matrix_s *A) {
  return SellView<T>{A->sptr.as<int>(), A->perm.p ? A->perm.as<int>() : nullptr, A->col.as<int>(),
                     (const T *)(sizeof(T) == 8 ? A->val64.p : A->val32.p), A->brows, A->ntot,
                     A->col16.p ? A->col16.as<short>() : nullptr, keep_slots(A, (int)sizeof(T))};
}

// Cap on the plain SpMV grid: past one slice per wavefront the waves loop.
inline int spmv_cap(yafel_hip_ctx_s *ctx) { return ctx->num_cus * 16; }

template <class T>
void launch_csr(yafel_hip_matrix_s *A, const T *x, T *y, hipStream_t s) {
  const SellView<T> v = sell_view<T>(A);
  const int grid = sell_grid(A->nsl, spmv_cap(A->ctx));
  switch (A->unroll) {
    case 1: k_sell_spmv_csr<1, T><<<grid, 256, 0, s>>>(v, x, y); break;
    case 2: k_sell_spmv_csr<2, T><<<grid, 256, 0, s>>>(v, x, y); break;
    case 8: k_sell_spmv_csr<8, T><<<grid, 256, 0, s>>>(v, x, y); break;
    default: k_sell_spmv_csr<4, T><<<grid, 256, 0, s>>>(v, x, y); break;
  }
  HIPCHK(hipGetLastError());
}

template <class T>
void launch_bcsr(yafel_hip_matrix_s *A, const T *x, T *y, hipStream_t s) {
  const SellView<T> v = sell_view<T>(A);
  const int grid = sell_grid(A->nsl, spmv_cap(A->ctx));
  switch (A->unroll) {
    case 1: k_sell_spmv_bcsr<1, T><<<grid, 256, 0, s>>>(v, x, y); break;
    case 2: k_sell_spmv_bcsr<2, T><<<grid, 256, 0, s>>>(v, x, y); break;
    case 8: k_sell_spmv_bcsr<8, T><<<grid, 256, 0, s>>>(v, x, y); break;
    default: k_sell_spmv_bcsr<4, T><<<grid, 256, 0, s>>>(v, x, y); break;
  }
  HIPCHK(hipGetLastError());
}

yafel_hip_matrix_s *new_matrix(yafel_hip_ctx_s *ctx, int format) {
  auto *A = new yafel_hip_matrix_s();
  ctx_retain(ctx);
  A->ctx = ctx;
  A->format = format;
  return A;
}

}  // namespace

namespace yh {

void matrix_make_sliced(yafel_hip_matrix_s *A, DBuf &ptr, DBuf &col, DBuf &val64) {
  hipStream_t s = A->ctx->stream;
  const int64_t brows = A->brows, nblk = A->nblk;
  const int nval = A->format == YAFEL_HIP_CSR ? 1 : 9;
  const int64_t nsl = (brows + 63) / 64;
  A->nsl = nsl;
  A->ptr.swap(ptr);
  A->perm.release();
  A->iperm.release();
  A->sptr.alloc((nsl + 1) * 4);
  int64_t slots = 0;
  if (nsl > 0) {
    // slot counts with natural row order and with rows sorted inside windows
    DBuf su, ss, perm, off;
    su.alloc((nsl + 1) * 8);
    ss.alloc((nsl + 1) * 8);
    off.alloc((nsl + 1) * 8);
    perm.alloc(brows * 4);
    const int *p = A->ptr.as<int>();
    k_slice_len<<<ceil_div(nsl + 1, 256), 256, 0, s>>>(brows, nsl, p, su.as<int64_t>());
    k_window_sort<<<ceil_div(brows, kSortWindow), kSortWindow, 0, s>>>(brows, nsl, p, perm.as<int>(), ss.as<int64_t>());
    HIPCHK(hipGetLastError());
    int64_t tot[2];
    exclusive_scan_i64(su.as<int64_t>(), off.as<int64_t>(), nsl + 1, s);
    HIPCHK(hipMemcpyAsync(&tot[0], off.as<int64_t>() + nsl, 8, hipMemcpyDeviceToHost, s));
    exclusive_scan_i64(ss.as<int64_t>(), su.as<int64_t>(), nsl + 1, s);
    HIPCHK(hipMemcpyAsync(&tot[1], su.as<int64_t>() + nsl, 8, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    // sorting moves the output rows around; take it only when it saves padding
    const bool sorted = tot[1] < tot[0] - tot[0] / 50;
    slots = sorted ? tot[1] : tot[0];
    REQUIRE(slots < (int64_t)INT32_MAX, "matrix: sliced storage exceeds int32 slot offsets");
    k_slots_to_i32<<<ceil_div(nsl + 1, 256), 256, 0, s>>>(nsl + 1, sorted ? su.as<int64_t>() : off.as<int64_t>(),
                                                         A->sptr.as<int>());
    HIPCHK(hipGetLastError());
    if (sorted) {
      A->perm.swap(perm);
      A->iperm.alloc(brows * 4);
      k_inverse<<<ceil_div(brows, 256), 256, 0, s>>>(brows, A->perm.as<int>(), A->iperm.as<int>());
      HIPCHK(hipGetLastError());
    }
    A->ntot = slots * 64;
    // long scalar rows (Q4 CSR: about 36 slots per slice) keep more loads in
    // flight with 8 slots per step; shorter rows and 3x3 blocks (ten loads
    // per slot already) run best at 4
    A->unroll = A->format == YAFEL_HIP_CSR && slots >= 24 * nsl ? 8 : 4;
    A->col.alloc(std::max<int64_t>(A->ntot, 1) * 4);
    A->val64.alloc(std::max<int64_t>(A->ntot * nval, 1) * 8);
    k_sell_fill<<<ceil_div(nsl * 64, 256), 256, 0, s>>>(brows, nsl, nblk, A->ntot, nval, A->ptr.as<int>(),
                                                         col.as<int>(), val64.as<double>(),
                                                         A->perm.p ? A->perm.as<int>() : nullptr, A->sptr.as<int>(),
                                                         A->col.as<int>(), A->val64.as<double>());
    HIPCHK(hipGetLastError());
    {
      DBuf bad;
      bad.alloc(8);
      HIPCHK(hipMemsetAsync(bad.p, 0, 4, s));
      A->col16.alloc(std::max<int64_t>(A->ntot, 1) * 2);
      k_sell_col16<<<ceil_div(nsl * 64, 256), 256, 0, s>>>(brows, A->bcols, nsl, A->sptr.as<int>(), A->col.as<int>(),
                                                           A->col16.as<short>(), bad.as<int>());
      HIPCHK(hipGetLastError());
      int hb = 0;
      HIPCHK(hipMemcpyAsync(&hb, bad.p, 4, hipMemcpyDeviceToHost, s));
      HIPCHK(hipStreamSynchronize(s));
      if (hb || A->bcols < 1) A->col16.release();
    }
  } else {
    HIPCHK(hipMemsetAsync(A->sptr.p, 0, 4, s));
    A->ntot = 0;
    A->col.alloc(4);
    A->val64.alloc(8 * nval);
    HIPCHK(hipStreamSynchronize(s));
  }
  col.release();
  val64.release();
  A->has64 = true;
  A->has32 = false;
  A->v64_gen++;
  A->diag_ok = false;
}

void spmv_launch(yafel_hip_matrix_s *A, int precision, const void *x, void *y, hipStream_t s) {
  if (A->brows == 0) return;
  if (precision == YAFEL_HIP_FP32) {
    REQUIRE(A->has32, "spmv: matrix has no FP32 copy");
    if (A->format == YAFEL_HIP_CSR) launch_csr<float>(A, (const float *)x, (float *)y, s);
    else launch_bcsr<float>(A, (const float *)x, (float *)y, s);
  } else {
    REQUIRE(A->has64, "spmv: matrix has no FP64 copy");
    if (A->format == YAFEL_HIP_CSR) launch_csr<double>(A, (const double *)x, (double *)y, s);
    else launch_bcsr<double>(A, (const double *)x, (double *)y, s);
  }
}

void spmm_launch(yafel_hip_matrix_s *A, int precision, int k, const void *X, void *Y, hipStream_t s) {
  if (A->brows == 0) return;
  const int grid = grid_for(A->ctx, A->brows * 64, 256);
  const int *p = A->ptr.as<int>(), *sp = A->sptr.as<int>(), *c = A->col.as<int>();
  const int *ip = A->iperm.p ? A->iperm.as<int>() : nullptr;
  const int *pm = A->perm.p ? A->perm.as<int>() : nullptr;
  // CSR: RB-row units of the slices, four per workgroup
  const int64_t units = A->nsl * (precision == YAFEL_HIP_FP32 ? 1 : 2);
  const int gcsr = (int)std::min<int64_t>((units + 3) / 4, (int64_t)A->ctx->num_cus * 8);
  if (precision == YAFEL_HIP_FP32) {
    REQUIRE(A->has32, "spmm: matrix has no FP32 copy");
    if (spmm_lds_launch(A, k, (const float *)X, (float *)Y, s)) return;
    if (A->format == YAFEL_HIP_CSR && k % 4 == 0 && ((uintptr_t)X % 16) == 0 && ((uintptr_t)Y % 16) == 0)
      k_csr_spmm4<<<(int)std::min<int64_t>((A->nsl * 2 + 3) / 4, (int64_t)A->ctx->num_cus * 8), 256, 0, s>>>(
          A->rows, k, sp, pm, c, A->val32.as<float>(), (const float *)X, (float *)Y);
    else if (A->format == YAFEL_HIP_CSR)
      k_csr_spmm<float><<<gcsr, 256, 0, s>>>(A->rows, k, sp, pm, c, A->val32.as<float>(), (const float *)X, (float *)Y);
    else
      k_bcsr_spmm<float><<<grid, 256, 0, s>>>(A->brows, A->ntot, k, p, sp, ip, c, A->val32.as<float>(),
                                              (const float *)X, (float *)Y);
  } else {
    REQUIRE(A->has64, "spmm: matrix has no FP64 copy");
    if (spmm_lds64_launch(A, k, (const double *)X, (double *)Y, s)) return;
    if (A->format == YAFEL_HIP_CSR)
      k_csr_spmm<double><<<gcsr, 256, 0, s>>>(A->rows, k, sp, pm, c, A->val64.as<double>(), (const double *)X,
                                              (double *)Y);
    else
      k_bcsr_spmm<double><<<grid, 256, 0, s>>>(A->brows, A->ntot, k, p, sp, ip, c, A->val64.as<double>(),
                                               (const double *)X, (double *)Y);
  }
  HIPCHK(hipGetLastError());
}

void matrix_build_fp32(yafel_hip_matrix_s *A) {
  const int64_t n = A->ntot * (A->format == YAFEL_HIP_CSR ? 1 : 9);
  if (!A->val32.p || A->val32.bytes < (size_t)std::max<int64_t>(n, 1) * 4) A->val32.alloc((size_t)std::max<int64_t>(n, 1) * 4);
  if (n > 0) {
    // dense rows only (at least 32 stored scalars a row on average): on Q4
    // Poisson (36) the mixed solve took 3887 inner iterations with the fold
    // against 4408 without, on 3-D elasticity in BCSR3 (78) 862 against 868;
    // on Q1 and Q2 rows (9-16) it added 2-3 % to the served solves' times
    const bool bcsr = A->format == YAFEL_HIP_BCSR3;
    const bool fold = (double)A->nblk * (bcsr ? 9 : 1) >= 32.0 * (double)A->rows;
    const int *pm = A->perm.p ? A->perm.as<int>() : nullptr;
    const int g = (int)ceil_div(A->nsl * 64, 256);
    if (fold && bcsr)
      k_fp32_rowsum<3><<<g, 256, 0, A->ctx->stream>>>(A->brows, A->nsl, A->ntot, A->ptr.as<int>(), A->sptr.as<int>(),
                                                     pm, A->col.as<int>(), A->val64.as<double>(), A->val32.as<float>());
    else if (fold)
      k_fp32_rowsum<1><<<g, 256, 0, A->ctx->stream>>>(A->brows, A->nsl, A->ntot, A->ptr.as<int>(), A->sptr.as<int>(),
                                                     pm, A->col.as<int>(), A->val64.as<double>(), A->val32.as<float>());
    else
      k_to_fp32<<<grid_for(A->ctx, n, 256), 256, 0, A->ctx->stream>>>(A->val64.as<double>(), A->val32.as<float>(), n);
    HIPCHK(hipGetLastError());
  }
  A->has32 = true;
  A->v32_gen++;
}

void matrix_build_diag(yafel_hip_matrix_s *A) {
  const int64_t n = A->rows;
  if (!A->diag64.p) {
    A->diag64.alloc(std::max<int64_t>(n, 1) * 8);
    A->dinv64.alloc(std::max<int64_t>(n, 1) * 8);
    A->dinv32.alloc(std::max<int64_t>(n, 1) * 4);
  }
  const int *ip = A->iperm.p ? A->iperm.as<int>() : nullptr;
  if (A->brows > 0) {
    if (A->format == YAFEL_HIP_CSR)
      k_diag_csr<<<ceil_div(A->brows, 256), 256, 0, A->ctx->stream>>>(
          A->brows, A->ptr.as<int>(), A->sptr.as<int>(), ip, A->col.as<int>(), A->val64.as<double>(),
          A->diag64.as<double>(), A->dinv64.as<double>(), A->dinv32.as<float>());
    else
      k_diag_bcsr<<<ceil_div(A->brows, 256), 256, 0, A->ctx->stream>>>(
          A->brows, A->ntot, A->ptr.as<int>(), A->sptr.as<int>(), ip, A->col.as<int>(), A->val64.as<double>(),
          A->diag64.as<double>(), A->dinv64.as<double>(), A->dinv32.as<float>());
    HIPCHK(hipGetLastError());
  }
  A->diag_ok = true;
  A->diag_gen++;
}

// Shared by the triplet and pattern constructors: takes the CSR-ordered device
// arrays and converts them to sliced storage.
yafel_hip_matrix_s *matrix_from_device(yafel_hip_ctx_s *ctx, int format, int64_t brows, int64_t bcols,
                                       int64_t nblk, int64_t ptr_len, DBuf &ptr, DBuf &col, DBuf &val64,
                                       int precisions) {
  std::unique_ptr<yafel_hip_matrix_s> A(new_matrix(ctx, format));
  const int B = format == YAFEL_HIP_CSR ? 1 : 3;
  A->brows = brows;
  A->bcols = bcols;
  A->rows = brows * B;
  A->cols = bcols * B;
  A->nblk = nblk;
  A->ptr_len = ptr_len;
  matrix_make_sliced(A.get(), ptr, col, val64);
  if (precisions & YAFEL_HIP_FP32) matrix_build_fp32(A.get());
  return A.release();
}

}  // namespace yh

extern "C" {

int yafel_hip_csr_create(yafel_hip_ctx ctx, int64_t rows, int64_t cols, const int64_t *row_ptr,
                         const int64_t *col_index, const double *data, int precisions,
                         yafel_hip_matrix *out) {
  YH_TRY_BEGIN
  REQUIRE(ctx && out && row_ptr, "csr_create: null argument");
  REQUIRE(rows >= 0 && cols >= 0 && rows < INT32_MAX && cols < INT32_MAX, "csr_create: bad dimensions");
  const int64_t nnz = row_ptr[rows];
  REQUIRE(nnz >= 0 && nnz < INT32_MAX, "csr_create: nnz out of int32 range");
  REQUIRE(nnz == 0 || (col_index && data), "csr_create: null col_index/data");
  for (int64_t r = 0; r < rows; ++r) REQUIRE(row_ptr[r] <= row_ptr[r + 1], "csr_create: row_ptr not monotone");
  HIPCHK(hipSetDevice(ctx->device));
  auto p32 = to_i32(row_ptr, rows + 1, nnz, "csr_create row_ptr");
  auto c32 = to_i32(col_index, nnz, cols - 1, "csr_create col_index");
  hipStream_t s = ctx->stream;
  DBuf ptr, col, val;
  ptr.alloc((rows + 1) * 4);
  col.alloc(std::max<int64_t>(nnz, 1) * 4);
  val.alloc(std::max<int64_t>(nnz, 1) * 8);
  HIPCHK(hipMemcpyAsync(ptr.p, p32.data(), (rows + 1) * 4, hipMemcpyHostToDevice, s));
  if (nnz) {
    HIPCHK(hipMemcpyAsync(col.p, c32.data(), nnz * 4, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(val.p, data, nnz * 8, hipMemcpyHostToDevice, s));
  }
  std::unique_ptr<yafel_hip_matrix_s> A(matrix_from_device(ctx, YAFEL_HIP_CSR, rows, cols, nnz, rows + 1, ptr, col,
                                                           val, precisions));
  HIPCHK(hipStreamSynchronize(s));
  *out = A.release();
  return YAFEL_HIP_OK;
  YH_TRY_END(ctx)
}

int yafel_hip_bcsr3_create(yafel_hip_ctx ctx, int64_t brows, int64_t bcols, const int64_t *brow_ptr,
                           const int64_t *bcol_index, const double *data, int precisions,
                           yafel_hip_matrix *out) {
  YH_TRY_BEGIN
  REQUIRE(ctx && out && brow_ptr, "bcsr3_create: null argument");
  REQUIRE(brows >= 0 && bcols >= 0 && 3 * brows < INT32_MAX && 3 * bcols < INT32_MAX,
          "bcsr3_create: bad dimensions");
  const int64_t nb = brow_ptr[brows];
  REQUIRE(nb >= 0 && 9 * nb < INT32_MAX * 4ll, "bcsr3_create: too many blocks");
  REQUIRE(nb == 0 || (bcol_index && data), "bcsr3_create: null bcol_index/data");
  for (int64_t r = 0; r < brows; ++r) REQUIRE(brow_ptr[r] <= brow_ptr[r + 1], "bcsr3_create: brow_ptr not monotone");
  HIPCHK(hipSetDevice(ctx->device));
  auto p32 = to_i32(brow_ptr, brows + 1, nb, "bcsr3_create brow_ptr");
  auto c32 = to_i32(bcol_index, nb, bcols - 1, "bcsr3_create bcol_index");
  hipStream_t s = ctx->stream;
  DBuf ptr, col, val;
  ptr.alloc((brows + 1) * 4);
  col.alloc(std::max<int64_t>(nb, 1) * 4);
  val.alloc(std::max<int64_t>(nb, 1) * 72);
  HIPCHK(hipMemcpyAsync(ptr.p, p32.data(), (brows + 1) * 4, hipMemcpyHostToDevice, s));
  if (nb) {
    DBuf tmp;
    tmp.alloc(nb * 72);
    HIPCHK(hipMemcpyAsync(col.p, c32.data(), nb * 4, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(tmp.p, data, nb * 72, hipMemcpyHostToDevice, s));
    k_aos_to_soa<<<grid_for(ctx, nb * 9, 256), 256, 0, s>>>(tmp.as<double>(), val.as<double>(), nb);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(s));
  }
  std::unique_ptr<yafel_hip_matrix_s> A(matrix_from_device(ctx, YAFEL_HIP_BCSR3, brows, bcols, nb, brows + 1, ptr,
                                                           col, val, precisions));
  HIPCHK(hipStreamSynchronize(s));
  *out = A.release();
  return YAFEL_HIP_OK;
  YH_TRY_END(ctx)
}

int yafel_hip_matrix_destroy(yafel_hip_matrix A) {
  if (!A) return YAFEL_HIP_OK;
  (void)hipStreamSynchronize(A->ctx->stream);
  delete A;
  return YAFEL_HIP_OK;
}

int yafel_hip_matrix_info_get(yafel_hip_matrix A, yafel_hip_matrix_info *info) {
  YH_TRY_BEGIN
  REQUIRE(A && info, "matrix_info: null argument");
  info->format = A->format;
  info->precisions = (A->has64 ? YAFEL_HIP_FP64 : 0) | (A->has32 ? YAFEL_HIP_FP32 : 0);
  info->rows = A->rows;
  info->cols = A->cols;
  info->nblocks = A->nblk;
  info->nnz = A->nblk * (A->format == YAFEL_HIP_CSR ? 1 : 9);
  info->ptr_len = A->ptr_len;
  return YAFEL_HIP_OK;
  YH_TRY_END(nullptr)
}

int yafel_hip_matrix_download(yafel_hip_matrix A, int64_t *row_ptr, int64_t *col_index, double *data) {
  YH_TRY_BEGIN
  REQUIRE(A, "matrix_download: null matrix");
  hipStream_t s = A->ctx->stream;
  HIPCHK(hipSetDevice(A->ctx->device));
  if (row_ptr) {
    std::vector<int32_t> p(A->brows + 1);
    HIPCHK(hipMemcpyAsync(p.data(), A->ptr.p, p.size() * 4, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    for (size_t i = 0; i < p.size(); ++i) row_ptr[i] = p[i];
    for (int64_t i = (int64_t)p.size(); i < A->ptr_len; ++i) row_ptr[i] = A->ptr_tail.empty() ? p.back() : A->ptr_tail[i - p.size()];
  }
  if ((col_index || data) && A->nblk) {
    REQUIRE(!data || A->has64, "matrix_download: no FP64 values");
    const int nval = A->format == YAFEL_HIP_CSR ? 1 : 9;
    DBuf dc, dv;
    if (col_index) dc.alloc(A->nblk * 4);
    if (data) dv.alloc(A->nblk * nval * 8);
    k_sell_to_csr<<<ceil_div(A->brows, 256), 256, 0, s>>>(
        A->brows, A->ntot, nval, A->ptr.as<int>(), A->sptr.as<int>(), A->iperm.p ? A->iperm.as<int>() : nullptr,
        A->col.as<int>(), A->val64.as<double>(), col_index ? dc.as<int>() : nullptr,
        data ? dv.as<double>() : nullptr);
    HIPCHK(hipGetLastError());
    std::vector<int32_t> c(col_index ? A->nblk : 0);
    if (col_index) HIPCHK(hipMemcpyAsync(c.data(), dc.p, A->nblk * 4, hipMemcpyDeviceToHost, s));
    if (data) HIPCHK(hipMemcpyAsync(data, dv.p, A->nblk * nval * 8, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    for (size_t i = 0; i < c.size(); ++i) col_index[i] = c[i];
  }
  return YAFEL_HIP_OK;
  YH_TRY_END(A->ctx)
}

int yafel_hip_matrix_sync_fp32(yafel_hip_matrix A) {
  YH_TRY_BEGIN
  REQUIRE(A && A->has64, "matrix_sync_fp32: matrix has no FP64 values");
  HIPCHK(hipSetDevice(A->ctx->device));
  matrix_build_fp32(A);
  return YAFEL_HIP_OK;
  YH_TRY_END(A ? A->ctx : nullptr)
}

int yafel_hip_matrix_set_unroll(yafel_hip_matrix A, int unroll) {
  YH_TRY_BEGIN
  REQUIRE(A, "matrix_set_unroll: null matrix");
  REQUIRE(unroll == 1 || unroll == 2 || unroll == 4 || unroll == 8, "matrix_set_unroll: unroll must be 1, 2, 4 or 8");
  A->unroll = unroll;
  return YAFEL_HIP_OK;
  YH_TRY_END(A ? A->ctx : nullptr)
}

int yafel_hip_matrix_set_cache_bytes(yafel_hip_matrix A, int64_t bytes) {
  YH_TRY_BEGIN
  REQUIRE(A, "matrix_set_cache_bytes: null matrix");
  REQUIRE(bytes >= -1, "matrix_set_cache_bytes: bytes must be -1 (default) or >= 0");
  A->cache_bytes = bytes;
  return YAFEL_HIP_OK;
  YH_TRY_END(A ? A->ctx : nullptr)
}

int yafel_hip_matrix_index_bytes(yafel_hip_matrix A, int *bytes) {
  YH_TRY_BEGIN
  REQUIRE(A && bytes, "matrix_index_bytes: null argument");
  *bytes = A->col16.p ? 2 : 4;
  return YAFEL_HIP_OK;
  YH_TRY_END(A ? A->ctx : nullptr)
}

int yafel_hip_matrix_set_cg_passes(yafel_hip_matrix A, int passes) {
  YH_TRY_BEGIN
  REQUIRE(A, "matrix_set_cg_passes: null matrix");
  REQUIRE(passes == 0 || passes == 2 || passes == 3, "matrix_set_cg_passes: passes must be 0, 2 or 3");
  A->cg_passes = passes;
  return YAFEL_HIP_OK;
  YH_TRY_END(A ? A->ctx : nullptr)
}

int yafel_hip_matrix_diagonal(yafel_hip_matrix A, double *diag_dev) {
  YH_TRY_BEGIN
  REQUIRE(A && diag_dev, "matrix_diagonal: null argument");
  HIPCHK(hipSetDevice(A->ctx->device));
  matrix_build_diag(A);
  HIPCHK(hipMemcpyAsync(diag_dev, A->diag64.p, A->rows * 8, hipMemcpyDeviceToDevice, A->ctx->stream));
  return YAFEL_HIP_OK;
  YH_TRY_END(A ? A->ctx : nullptr)
}

int yafel_hip_spmv(yafel_hip_matrix A, int precision, const void *x_dev, void *y_dev) {
  YH_TRY_BEGIN
  REQUIRE(A && (A->rows == 0 || (x_dev && y_dev)), "spmv: null argument");
  REQUIRE(x_dev != y_dev, "spmv: x and y must not alias");
  spmv_launch(A, precision, x_dev, y_dev, A->ctx->stream);
  return YAFEL_HIP_OK;
  YH_TRY_END(A ? A->ctx : nullptr)
}

int yafel_hip_spmv_host(yafel_hip_matrix A, const double *x, double *y) {
  YH_TRY_BEGIN
  REQUIRE(A && (A->rows == 0 || (x && y)), "spmv_host: null argument");
  HIPCHK(hipSetDevice(A->ctx->device));
  hipStream_t s = A->ctx->stream;
  DBuf dx, dy;
  dx.alloc(std::max<int64_t>(A->cols, 1) * 8);
  dy.alloc(std::max<int64_t>(A->rows, 1) * 8);
  if (A->cols) HIPCHK(hipMemcpyAsync(dx.p, x, A->cols * 8, hipMemcpyHostToDevice, s));
  spmv_launch(A, YAFEL_HIP_FP64, dx.p, dy.p, s);
  if (A->rows) HIPCHK(hipMemcpyAsync(y, dy.p, A->rows * 8, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  return YAFEL_HIP_OK;
  YH_TRY_END(A ? A->ctx : nullptr)
}

int yafel_hip_spmm(yafel_hip_matrix A, int precision, int k, const void *X_dev, void *Y_dev) {
  YH_TRY_BEGIN
  REQUIRE(A && k > 0 && (A->rows == 0 || (X_dev && Y_dev)), "spmm: bad argument");
  REQUIRE(X_dev != Y_dev, "spmm: X and Y must not alias");
  spmm_launch(A, precision, k, X_dev, Y_dev, A->ctx->stream);
  return YAFEL_HIP_OK;
  YH_TRY_END(A ? A->ctx : nullptr)
}

}  // extern "C"
