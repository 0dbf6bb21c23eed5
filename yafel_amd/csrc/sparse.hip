// Device sparse matrices (CSR and 3x3-block BCSR), SpMV and multi-RHS SpMM.
//
// Device layout
//  CSR   : int32 row_ptr[rows+1], int32 col[nnz], values[nnz] (FP64 and/or FP32).
//  BCSR3 : int32 brow_ptr[brows+1], int32 bcol[nb], values entry-major
//          v[e * nb + b] for e = 3 i + j (structure-of-arrays), so that a lane
//          per block reads every entry stream with unit stride across lanes.
//          The reference keeps data[b * 9 + i * 3 + j]; conversion happens on
//          upload/download.
#include <algorithm>
#include <cmath>

#include "internal.hpp"
#include "kernels.hpp"

using namespace yh;

namespace {

__global__ void k_aos_to_soa(const double *__restrict__ aos, double *__restrict__ soa, int64_t nb) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < nb * 9; i += stride) {
    int64_t b = i / 9;
    int e = (int)(i - b * 9);
    soa[(int64_t)e * nb + b] = aos[i];
  }
}

__global__ void k_soa_to_aos(const double *__restrict__ soa, double *__restrict__ aos, int64_t nb) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < nb * 9; i += stride) {
    int64_t b = i / 9;
    int e = (int)(i - b * 9);
    aos[i] = soa[(int64_t)e * nb + b];
  }
}

__global__ void k_to_fp32(const double *__restrict__ a, float *__restrict__ b, int64_t n) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) b[i] = (float)a[i];
}

// Diagonal and its inverse (JacobiPreconditioner.hpp:124-137: 1/a_ii, kept
// even when a_ii == 0 -- the reference only warns).
__global__ void k_diag_csr(const int *__restrict__ ptr, const int *__restrict__ col,
                           const double *__restrict__ val, int64_t rows, double *diag,
                           double *dinv, float *dinv32) {
  int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (r >= rows) return;
  double d = 0;
  for (int k = ptr[r]; k < ptr[r + 1]; ++k)
    if (col[k] == r) { d = val[k]; break; }
  diag[r] = d;
  dinv[r] = 1.0 / d;
  dinv32[r] = (float)(1.0 / d);
}

__global__ void k_diag_bcsr(const int *__restrict__ ptr, const int *__restrict__ col,
                            const double *__restrict__ val, int64_t brows, int64_t nb, double *diag,
                            double *dinv, float *dinv32) {
  int64_t br = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (br >= brows) return;
  double d[3] = {0, 0, 0};
  for (int k = ptr[br]; k < ptr[br + 1]; ++k)
    if (col[k] == br) {
      for (int i = 0; i < 3; ++i) d[i] = val[(int64_t)(4 * i) * nb + k];
      break;
    }
  for (int i = 0; i < 3; ++i) {
    diag[3 * br + i] = d[i];
    dinv[3 * br + i] = 1.0 / d[i];
    dinv32[3 * br + i] = (float)(1.0 / d[i]);
  }
}

template <int L, class T>
__global__ __launch_bounds__(256) void k_csr_spmv(int64_t rows, const int *__restrict__ ptr,
                                                  const int *__restrict__ col,
                                                  const T *__restrict__ val,
                                                  const T *__restrict__ x, T *__restrict__ y) {
  const int lane = threadIdx.x & (L - 1);
  const int64_t rpb = 256 / L;
  const int64_t chunk = (rows + gridDim.x - 1) / gridDim.x;
  const int64_t r0 = (int64_t)dev::xcd_block() * chunk;
  const int64_t r1 = r0 + chunk < rows ? r0 + chunk : rows;
  for (int64_t r = r0 + threadIdx.x / L; r < r1 + 0; r += rpb) {
    T acc = dev::csr_row<L, T, T>(ptr, col, val, r, lane, [&](int c) { return x[c]; });
    if (lane == 0) y[r] = acc;
  }
}

template <int G, class T>
__global__ __launch_bounds__(256) void k_bcsr_spmv(int64_t brows, int64_t nb,
                                                   const int *__restrict__ ptr,
                                                   const int *__restrict__ col,
                                                   const T *__restrict__ val,
                                                   const T *__restrict__ x, T *__restrict__ y) {
  const int lane = threadIdx.x & (G - 1);
  const int64_t rpb = 256 / G;
  const int64_t chunk = (brows + gridDim.x - 1) / gridDim.x;
  const int64_t r0 = (int64_t)dev::xcd_block() * chunk;
  const int64_t r1 = r0 + chunk < brows ? r0 + chunk : brows;
  for (int64_t br = r0 + threadIdx.x / G; br < r1; br += rpb) {
    T a[3];
    dev::bcsr_row<G, T, T>(ptr, col, val, nb, br, lane, a, [&](int64_t c) { return x[c]; });
    if (lane < 3) y[3 * br + lane] = lane == 0 ? a[0] : (lane == 1 ? a[1] : a[2]);
  }
}

// Y = A X for k right-hand sides, row-major X (n x k).  One wavefront per row;
// lanes run over right-hand sides, so the X/Y rows are read/written coalesced
// and each matrix entry is one broadcast load.
template <class T>
__global__ __launch_bounds__(256) void k_csr_spmm(int64_t rows, int k, const int *__restrict__ ptr,
                                                  const int *__restrict__ col,
                                                  const T *__restrict__ val,
                                                  const T *__restrict__ X, T *__restrict__ Y) {
  const int lane = threadIdx.x & 63;
  const int64_t wpb = 4;
  const int64_t chunk = (rows + gridDim.x - 1) / gridDim.x;
  const int64_t r0 = (int64_t)dev::xcd_block() * chunk;
  const int64_t r1 = r0 + chunk < rows ? r0 + chunk : rows;
  for (int64_t r = r0 + threadIdx.x / 64; r < r1; r += wpb) {
    const int b = ptr[r], e = ptr[r + 1];
    for (int j0 = 0; j0 < k; j0 += 64) {
      const int j = j0 + lane;
      T acc = 0;
      if (j < k)
        for (int q = b; q < e; ++q) acc = fma(val[q], X[(int64_t)col[q] * k + j], acc);
      if (j < k) Y[r * (int64_t)k + j] = acc;
    }
  }
}

template <class T>
__global__ __launch_bounds__(256) void k_bcsr_spmm(int64_t brows, int64_t nb, int k,
                                                   const int *__restrict__ ptr,
                                                   const int *__restrict__ col,
                                                   const T *__restrict__ val,
                                                   const T *__restrict__ X, T *__restrict__ Y) {
  const int lane = threadIdx.x & 63;
  const int64_t chunk = (brows + gridDim.x - 1) / gridDim.x;
  const int64_t r0 = (int64_t)dev::xcd_block() * chunk;
  const int64_t r1 = r0 + chunk < brows ? r0 + chunk : brows;
  for (int64_t br = r0 + threadIdx.x / 64; br < r1; br += 4) {
    const int b = ptr[br], e = ptr[br + 1];
    for (int j0 = 0; j0 < k; j0 += 64) {
      const int j = j0 + lane;
      T a0 = 0, a1 = 0, a2 = 0;
      if (j < k)
        for (int q = b; q < e; ++q) {
          const int64_t c = col[q];
          const T x0 = X[(3 * c + 0) * k + j], x1 = X[(3 * c + 1) * k + j], x2 = X[(3 * c + 2) * k + j];
          a0 = fma(val[0 * nb + q], x0, fma(val[1 * nb + q], x1, fma(val[2 * nb + q], x2, a0)));
          a1 = fma(val[3 * nb + q], x0, fma(val[4 * nb + q], x1, fma(val[5 * nb + q], x2, a1)));
          a2 = fma(val[6 * nb + q], x0, fma(val[7 * nb + q], x1, fma(val[8 * nb + q], x2, a2)));
        }
      if (j < k) {
        Y[(3 * br + 0) * k + j] = a0;
        Y[(3 * br + 1) * k + j] = a1;
        Y[(3 * br + 2) * k + j] = a2;
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
void launch_csr(yafel_hip_matrix_s *A, const T *x, T *y, hipStream_t s) {
  const int L = A->lanes;
  const int grid = grid_for(A->ctx, A->rows * (int64_t)L, 256);
  const int *p = A->ptr.as<int>(), *c = A->col.as<int>();
  const T *v = (const T *)(sizeof(T) == 8 ? A->val64.p : A->val32.p);
  switch (L) {
    case 2: k_csr_spmv<2, T><<<grid, 256, 0, s>>>(A->rows, p, c, v, x, y); break;
    case 4: k_csr_spmv<4, T><<<grid, 256, 0, s>>>(A->rows, p, c, v, x, y); break;
    case 8: k_csr_spmv<8, T><<<grid, 256, 0, s>>>(A->rows, p, c, v, x, y); break;
    case 16: k_csr_spmv<16, T><<<grid, 256, 0, s>>>(A->rows, p, c, v, x, y); break;
    case 32: k_csr_spmv<32, T><<<grid, 256, 0, s>>>(A->rows, p, c, v, x, y); break;
    default: k_csr_spmv<64, T><<<grid, 256, 0, s>>>(A->rows, p, c, v, x, y); break;
  }
  HIPCHK(hipGetLastError());
}

template <class T>
void launch_bcsr(yafel_hip_matrix_s *A, const T *x, T *y, hipStream_t s) {
  const int G = A->lanes;
  const int grid = grid_for(A->ctx, A->brows * (int64_t)G, 256);
  const int *p = A->ptr.as<int>(), *c = A->col.as<int>();
  const T *v = (const T *)(sizeof(T) == 8 ? A->val64.p : A->val32.p);
  switch (G) {
    case 8: k_bcsr_spmv<8, T><<<grid, 256, 0, s>>>(A->brows, A->nblk, p, c, v, x, y); break;
    case 16: k_bcsr_spmv<16, T><<<grid, 256, 0, s>>>(A->brows, A->nblk, p, c, v, x, y); break;
    case 32: k_bcsr_spmv<32, T><<<grid, 256, 0, s>>>(A->brows, A->nblk, p, c, v, x, y); break;
    default: k_bcsr_spmv<64, T><<<grid, 256, 0, s>>>(A->brows, A->nblk, p, c, v, x, y); break;
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

void matrix_choose_lanes(yafel_hip_matrix_s *A) {
  const double avg = A->brows > 0 ? (double)A->nblk / (double)A->brows : 0.0;
  int L;
  if (A->format == YAFEL_HIP_CSR) {
    L = avg < 4 ? 2 : avg < 8 ? 4 : avg < 16 ? 8 : avg < 32 ? 16 : avg < 64 ? 32 : 64;
  } else {
    L = avg <= 8 ? 8 : avg <= 16 ? 16 : avg <= 32 ? 32 : 64;
  }
  A->lanes = L;
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
  const int *p = A->ptr.as<int>(), *c = A->col.as<int>();
  if (precision == YAFEL_HIP_FP32) {
    REQUIRE(A->has32, "spmm: matrix has no FP32 copy");
    if (A->format == YAFEL_HIP_CSR)
      k_csr_spmm<float><<<grid, 256, 0, s>>>(A->rows, k, p, c, A->val32.as<float>(), (const float *)X, (float *)Y);
    else
      k_bcsr_spmm<float><<<grid, 256, 0, s>>>(A->brows, A->nblk, k, p, c, A->val32.as<float>(), (const float *)X, (float *)Y);
  } else {
    REQUIRE(A->has64, "spmm: matrix has no FP64 copy");
    if (A->format == YAFEL_HIP_CSR)
      k_csr_spmm<double><<<grid, 256, 0, s>>>(A->rows, k, p, c, A->val64.as<double>(), (const double *)X, (double *)Y);
    else
      k_bcsr_spmm<double><<<grid, 256, 0, s>>>(A->brows, A->nblk, k, p, c, A->val64.as<double>(), (const double *)X, (double *)Y);
  }
  HIPCHK(hipGetLastError());
}

void matrix_build_fp32(yafel_hip_matrix_s *A) {
  const int64_t n = A->nblk * (A->format == YAFEL_HIP_CSR ? 1 : 9);
  if (!A->val32.p || A->val32.bytes < (size_t)n * 4) A->val32.alloc((size_t)std::max<int64_t>(n, 1) * 4);
  if (n > 0) {
    k_to_fp32<<<grid_for(A->ctx, n, 256), 256, 0, A->ctx->stream>>>(A->val64.as<double>(), A->val32.as<float>(), n);
    HIPCHK(hipGetLastError());
  }
  A->has32 = true;
}

void matrix_build_diag(yafel_hip_matrix_s *A) {
  const int64_t n = A->rows;
  if (!A->diag64.p) {
    A->diag64.alloc(std::max<int64_t>(n, 1) * 8);
    A->dinv64.alloc(std::max<int64_t>(n, 1) * 8);
    A->dinv32.alloc(std::max<int64_t>(n, 1) * 4);
  }
  if (A->brows > 0) {
    if (A->format == YAFEL_HIP_CSR)
      k_diag_csr<<<ceil_div(A->brows, 256), 256, 0, A->ctx->stream>>>(
          A->ptr.as<int>(), A->col.as<int>(), A->val64.as<double>(), A->brows, A->diag64.as<double>(),
          A->dinv64.as<double>(), A->dinv32.as<float>());
    else
      k_diag_bcsr<<<ceil_div(A->brows, 256), 256, 0, A->ctx->stream>>>(
          A->ptr.as<int>(), A->col.as<int>(), A->val64.as<double>(), A->brows, A->nblk,
          A->diag64.as<double>(), A->dinv64.as<double>(), A->dinv32.as<float>());
    HIPCHK(hipGetLastError());
  }
  A->diag_ok = true;
}

// Shared by the triplet constructors: takes ownership of device arrays.
yafel_hip_matrix_s *matrix_from_device(yafel_hip_ctx_s *ctx, int format, int64_t brows, int64_t bcols,
                                       int64_t nblk, int64_t ptr_len, DBuf &ptr, DBuf &col, DBuf &val64,
                                       int precisions) {
  auto *A = new_matrix(ctx, format);
  const int B = format == YAFEL_HIP_CSR ? 1 : 3;
  A->brows = brows;
  A->bcols = bcols;
  A->rows = brows * B;
  A->cols = bcols * B;
  A->nblk = nblk;
  A->ptr_len = ptr_len;
  std::swap(A->ptr.p, ptr.p); std::swap(A->ptr.bytes, ptr.bytes);
  std::swap(A->col.p, col.p); std::swap(A->col.bytes, col.bytes);
  std::swap(A->val64.p, val64.p); std::swap(A->val64.bytes, val64.bytes);
  A->has64 = true;
  matrix_choose_lanes(A);
  if (precisions & YAFEL_HIP_FP32) matrix_build_fp32(A);
  return A;
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
  std::unique_ptr<yafel_hip_matrix_s> A(new_matrix(ctx, YAFEL_HIP_CSR));
  A->rows = A->brows = rows;
  A->cols = A->bcols = cols;
  A->nblk = nnz;
  A->ptr_len = rows + 1;
  A->ptr.alloc((rows + 1) * 4);
  A->col.alloc(std::max<int64_t>(nnz, 1) * 4);
  A->val64.alloc(std::max<int64_t>(nnz, 1) * 8);
  hipStream_t s = ctx->stream;
  HIPCHK(hipMemcpyAsync(A->ptr.p, p32.data(), (rows + 1) * 4, hipMemcpyHostToDevice, s));
  if (nnz) {
    HIPCHK(hipMemcpyAsync(A->col.p, c32.data(), nnz * 4, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(A->val64.p, data, nnz * 8, hipMemcpyHostToDevice, s));
  }
  A->has64 = true;
  matrix_choose_lanes(A.get());
  if (precisions & YAFEL_HIP_FP32) matrix_build_fp32(A.get());
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
  std::unique_ptr<yafel_hip_matrix_s> A(new_matrix(ctx, YAFEL_HIP_BCSR3));
  A->brows = brows;
  A->bcols = bcols;
  A->rows = 3 * brows;
  A->cols = 3 * bcols;
  A->nblk = nb;
  A->ptr_len = brows + 1;
  A->ptr.alloc((brows + 1) * 4);
  A->col.alloc(std::max<int64_t>(nb, 1) * 4);
  A->val64.alloc(std::max<int64_t>(nb, 1) * 72);
  hipStream_t s = ctx->stream;
  HIPCHK(hipMemcpyAsync(A->ptr.p, p32.data(), (brows + 1) * 4, hipMemcpyHostToDevice, s));
  if (nb) {
    DBuf tmp;
    tmp.alloc(nb * 72);
    HIPCHK(hipMemcpyAsync(A->col.p, c32.data(), nb * 4, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(tmp.p, data, nb * 72, hipMemcpyHostToDevice, s));
    k_aos_to_soa<<<grid_for(ctx, nb * 9, 256), 256, 0, s>>>(tmp.as<double>(), A->val64.as<double>(), nb);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(s));
  }
  A->has64 = true;
  matrix_choose_lanes(A.get());
  if (precisions & YAFEL_HIP_FP32) matrix_build_fp32(A.get());
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
  if (col_index && A->nblk) {
    std::vector<int32_t> c(A->nblk);
    HIPCHK(hipMemcpyAsync(c.data(), A->col.p, c.size() * 4, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    for (size_t i = 0; i < c.size(); ++i) col_index[i] = c[i];
  }
  if (data && A->nblk) {
    REQUIRE(A->has64, "matrix_download: no FP64 values");
    if (A->format == YAFEL_HIP_CSR) {
      HIPCHK(hipMemcpyAsync(data, A->val64.p, A->nblk * 8, hipMemcpyDeviceToHost, s));
    } else {
      DBuf tmp;
      tmp.alloc(A->nblk * 72);
      k_soa_to_aos<<<grid_for(A->ctx, A->nblk * 9, 256), 256, 0, s>>>(A->val64.as<double>(), tmp.as<double>(), A->nblk);
      HIPCHK(hipGetLastError());
      HIPCHK(hipMemcpyAsync(data, tmp.p, A->nblk * 72, hipMemcpyDeviceToHost, s));
      HIPCHK(hipStreamSynchronize(s));
    }
    HIPCHK(hipStreamSynchronize(s));
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

int yafel_hip_matrix_set_lanes(yafel_hip_matrix A, int lanes) {
  YH_TRY_BEGIN
  REQUIRE(A, "matrix_set_lanes: null matrix");
  const bool pow2 = lanes > 0 && (lanes & (lanes - 1)) == 0;
  if (A->format == YAFEL_HIP_CSR) REQUIRE(pow2 && lanes >= 2 && lanes <= 64, "matrix_set_lanes: CSR lanes must be 2..64, a power of two");
  else REQUIRE(pow2 && lanes >= 8 && lanes <= 64, "matrix_set_lanes: BCSR3 lanes must be 8..64, a power of two");
  A->lanes = lanes;
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
