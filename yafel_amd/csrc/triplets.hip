// sparse_csr(triplets) and sparse_bcsr<3>(triplets) on the device.
//
// CSR (old_handmade_linalg/sparse_csr.hpp:90-149): triplets ordered by
// (row, col, value) -- two stable radix passes, value first -- and each run of
// equal (row, col) summed left to right in that order, which is exactly the
// reference's summation order.  BCSR (old_handmade_linalg/sparse_bcsr.hpp:93-176):
// ordered by (I, J, i, j) with ties kept in input order, summed into
// zero-filled 3x3 blocks.
#include <algorithm>

#include "internal.hpp"
#include "sortprim.hpp"

using namespace yh;

namespace {

constexpr int NT = 256;

__device__ __forceinline__ uint64_t orderable(double v) {
  uint64_t u = __double_as_longlong(v);
  return (u >> 63) ? ~u : (u | 0x8000000000000000ull);
}

__global__ void k_iota_vkey(int64_t n, const double *__restrict__ v, uint64_t *key, uint32_t *idx) {
  int64_t i = blockIdx.x * (int64_t)NT + threadIdx.x;
  if (i >= n) return;
  key[i] = orderable(v[i]);
  idx[i] = (uint32_t)i;
}

__global__ void k_iota(int64_t n, uint32_t *idx) {
  int64_t i = blockIdx.x * (int64_t)NT + threadIdx.x;
  if (i < n) idx[i] = (uint32_t)i;
}

// key of the (row, col) pair of the triplet idx[i]
__global__ void k_rc_key(int64_t n, const uint32_t *idx, const int64_t *r, const int64_t *c,
                         uint64_t ncols, uint64_t *key) {
  int64_t i = blockIdx.x * (int64_t)NT + threadIdx.x;
  if (i >= n) return;
  const uint32_t t = idx[i];
  key[i] = (uint64_t)r[t] * ncols + (uint64_t)c[t];
}

__global__ void k_bcsr_key(int64_t n, const int64_t *r, const int64_t *c, uint64_t nbc, uint64_t *key) {
  int64_t i = blockIdx.x * (int64_t)NT + threadIdx.x;
  if (i >= n) return;
  const uint64_t R = r[i], C = c[i];
  key[i] = (((R / 3) * nbc + C / 3) * 3 + R % 3) * 3 + C % 3;
}

// head flags of runs of equal key (>> shift)
__global__ void k_heads(int64_t n, const uint64_t *key, int shift_div9, int32_t *flag) {
  int64_t i = blockIdx.x * (int64_t)NT + threadIdx.x;
  if (i >= n) return;
  uint64_t a = key[i], b = i ? key[i - 1] : 0;
  if (shift_div9) {
    a /= 9;
    b /= 9;
  }
  flag[i] = (i == 0 || a != b) ? 1 : 0;
}

// head position of every run: pos[seg-1] = i
__global__ void k_head_pos(int64_t n, const int32_t *flag, const int32_t *seg, int64_t *pos) {
  int64_t i = blockIdx.x * (int64_t)NT + threadIdx.x;
  if (i >= n) return;
  if (flag[i]) pos[seg[i] - 1] = i;
}

// CSR: one thread per (row, col) run, sequential left-to-right sum.
__global__ void k_csr_runs(int64_t nrun, int64_t n, const int64_t *pos, const uint64_t *key,
                           const uint32_t *idx, const double *v, uint64_t ncols, int32_t *col,
                           double *val, int32_t *ptr) {
  int64_t s = blockIdx.x * (int64_t)NT + threadIdx.x;
  if (s >= nrun) return;
  const int64_t a = pos[s], e = s + 1 < nrun ? pos[s + 1] : n;
  double acc = v[idx[a]];
  for (int64_t i = a + 1; i < e; ++i) acc += v[idx[i]];
  val[s] = acc;
  const uint64_t k = key[a];
  const int64_t row = (int64_t)(k / ncols);
  col[s] = (int32_t)(k % ncols);
  const int64_t prev = s ? (int64_t)(key[pos[s - 1]] / ncols) : -1;
  for (int64_t rr = prev + 1; rr <= row; ++rr) ptr[rr] = (int32_t)s;
}

// BCSR: one thread per (I,J,i,j) run; blocks are runs of key/9.
__global__ void k_bcsr_runs(int64_t nrun, int64_t n, int64_t nb, const int64_t *pos,
                            const uint64_t *key, const uint32_t *idx, const double *v,
                            const int32_t *bseg, double *val) {
  int64_t s = blockIdx.x * (int64_t)NT + threadIdx.x;
  if (s >= nrun) return;
  const int64_t a = pos[s], e = s + 1 < nrun ? pos[s + 1] : n;
  double acc = v[idx[a]];
  for (int64_t i = a + 1; i < e; ++i) acc += v[idx[i]];
  const uint64_t k = key[a];
  const int ent = (int)(k % 9);
  const int64_t blk = bseg[a] - 1;
  val[(int64_t)ent * nb + blk] = acc;
}

__global__ void k_bcsr_blocks(int64_t nblk, int64_t n, const int64_t *bpos, const uint64_t *key,
                              uint64_t nbc, int32_t *col, int32_t *ptr) {
  int64_t s = blockIdx.x * (int64_t)NT + threadIdx.x;
  if (s >= nblk) return;
  const uint64_t bk = key[bpos[s]] / 9;
  const int64_t I = (int64_t)(bk / nbc);
  col[s] = (int32_t)(bk % nbc);
  const int64_t prev = s ? (int64_t)((key[bpos[s - 1]] / 9) / nbc) : -1;
  for (int64_t rr = prev + 1; rr <= I; ++rr) ptr[rr] = (int32_t)s;
}

__global__ void k_set_i32(int32_t *p, int32_t v) { *p = v; }

struct Upload {
  DBuf r, c, v;
};

void upload(yafel_hip_ctx_s *ctx, int64_t nt, const int64_t *rows, const int64_t *cols,
            const double *vals, Upload &u) {
  u.r.alloc(nt * 8);
  u.c.alloc(nt * 8);
  u.v.alloc(nt * 8);
  HIPCHK(hipMemcpyAsync(u.r.p, rows, nt * 8, hipMemcpyHostToDevice, ctx->stream));
  HIPCHK(hipMemcpyAsync(u.c.p, cols, nt * 8, hipMemcpyHostToDevice, ctx->stream));
  HIPCHK(hipMemcpyAsync(u.v.p, vals, nt * 8, hipMemcpyHostToDevice, ctx->stream));
}

}  // namespace

extern "C" {

int yafel_hip_csr_from_triplets(yafel_hip_ctx ctx, int64_t nt, const int64_t *rows,
                                const int64_t *cols, const double *vals, int precisions,
                                yafel_hip_matrix *out) {
  YH_TRY_BEGIN
  REQUIRE(ctx && out && (nt == 0 || (rows && cols && vals)), "csr_from_triplets: null argument");
  REQUIRE(nt >= 0 && nt < (int64_t)UINT32_MAX, "csr_from_triplets: too many triplets");
  HIPCHK(hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  int64_t maxr = 0, maxc = 0;
  for (int64_t i = 0; i < nt; ++i) {
    REQUIRE(rows[i] >= 0 && cols[i] >= 0, "csr_from_triplets: negative index");
    maxr = std::max(maxr, rows[i]);
    maxc = std::max(maxc, cols[i]);
  }
  REQUIRE(maxr + 1 < INT32_MAX && maxc + 1 < INT32_MAX, "csr_from_triplets: index exceeds int32");
  DBuf ptr, col, val;
  if (nt == 0) {
    ptr.alloc(8);
    HIPCHK(hipMemsetAsync(ptr.p, 0, 8, s));
    *out = matrix_from_device(ctx, YAFEL_HIP_CSR, 0, 0, 0, 0, ptr, col, val, precisions);
    return YAFEL_HIP_OK;
  }
  Upload u;
  upload(ctx, nt, rows, cols, vals, u);
  const uint64_t ncols = (uint64_t)maxc + 1;
  DBuf key, idx, flag, seg, pos;
  key.alloc(nt * 8);
  idx.alloc(nt * 4);
  const int g = ceil_div(nt, NT);
  k_iota_vkey<<<g, NT, 0, s>>>(nt, u.v.as<double>(), key.as<uint64_t>(), idx.as<uint32_t>());
  HIPCHK(hipGetLastError());
  sort_pairs_u64_u32(key.as<uint64_t>(), idx.as<uint32_t>(), nt, 64, s);
  k_rc_key<<<g, NT, 0, s>>>(nt, idx.as<uint32_t>(), u.r.as<int64_t>(), u.c.as<int64_t>(), ncols, key.as<uint64_t>());
  HIPCHK(hipGetLastError());
  sort_pairs_u64_u32(key.as<uint64_t>(), idx.as<uint32_t>(), nt, bits_for((uint64_t)maxr * ncols + maxc), s);
  flag.alloc(nt * 4);
  seg.alloc(nt * 4);
  k_heads<<<g, NT, 0, s>>>(nt, key.as<uint64_t>(), 0, flag.as<int32_t>());
  HIPCHK(hipGetLastError());
  inclusive_scan_i32(flag.as<int32_t>(), seg.as<int32_t>(), nt, s);
  int32_t nnz = 0;
  HIPCHK(hipMemcpy(&nnz, seg.as<int32_t>() + nt - 1, 4, hipMemcpyDeviceToHost));
  pos.alloc((int64_t)nnz * 8);
  k_head_pos<<<g, NT, 0, s>>>(nt, flag.as<int32_t>(), seg.as<int32_t>(), pos.as<int64_t>());
  HIPCHK(hipGetLastError());
  const int64_t nrows = maxr + 1;
  ptr.alloc((nrows + 1) * 4);
  col.alloc((int64_t)nnz * 4);
  val.alloc((int64_t)nnz * 8);
  k_csr_runs<<<ceil_div(nnz, NT), NT, 0, s>>>(nnz, nt, pos.as<int64_t>(), key.as<uint64_t>(), idx.as<uint32_t>(),
                                              u.v.as<double>(), ncols, col.as<int32_t>(), val.as<double>(),
                                              ptr.as<int32_t>());
  HIPCHK(hipGetLastError());
  k_set_i32<<<1, 1, 0, s>>>(ptr.as<int32_t>() + nrows, nnz);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(s));
  auto *A = matrix_from_device(ctx, YAFEL_HIP_CSR, nrows, (int64_t)ncols, nnz, nrows + 1, ptr, col, val, precisions);
  HIPCHK(hipStreamSynchronize(s));
  *out = A;
  return YAFEL_HIP_OK;
  YH_TRY_END(ctx)
}

int yafel_hip_bcsr3_from_triplets(yafel_hip_ctx ctx, int64_t nt, const int64_t *rows,
                                  const int64_t *cols, const double *vals, int precisions,
                                  yafel_hip_matrix *out) {
  YH_TRY_BEGIN
  REQUIRE(ctx && out && (nt == 0 || (rows && cols && vals)), "bcsr3_from_triplets: null argument");
  REQUIRE(nt > 0, "bcsr3_from_triplets: empty matrices are not supported (as in sparse_bcsr.hpp:97)");
  REQUIRE(nt < (int64_t)UINT32_MAX, "bcsr3_from_triplets: too many triplets");
  HIPCHK(hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  // block extents.  The reference updates _brows/_bcols from the second sorted
  // triplet on (sparse_bcsr.hpp:140-152), i.e. over all but the smallest key.
  int64_t maxI = 0, maxJ = 0;
  int64_t imin = 0;
  auto bkey = [&](int64_t i) {
    return std::make_tuple(rows[i] / 3, cols[i] / 3, rows[i] % 3, cols[i] % 3);
  };
  for (int64_t i = 0; i < nt; ++i) {
    REQUIRE(rows[i] >= 0 && cols[i] >= 0, "bcsr3_from_triplets: negative index");
    maxI = std::max(maxI, rows[i] / 3);
    maxJ = std::max(maxJ, cols[i] / 3);
    if (bkey(i) < bkey(imin)) imin = i;
  }
  int64_t brows_ref = 0, bcols_ref = 0;
  for (int64_t i = 0; i < nt; ++i) {
    if (i == imin) continue;
    brows_ref = std::max(brows_ref, rows[i] / 3 + 1);
    bcols_ref = std::max(bcols_ref, cols[i] / 3 + 1);
  }
  REQUIRE(3 * (maxI + 1) < INT32_MAX && 3 * (maxJ + 1) < INT32_MAX, "bcsr3_from_triplets: index exceeds int32");
  const uint64_t nbc = (uint64_t)maxJ + 1;
  Upload u;
  upload(ctx, nt, rows, cols, vals, u);
  DBuf key, idx, flag, seg, pos, bflag, bseg, bpos, ptr, col, val;
  key.alloc(nt * 8);
  idx.alloc(nt * 4);
  const int g = ceil_div(nt, NT);
  k_bcsr_key<<<g, NT, 0, s>>>(nt, u.r.as<int64_t>(), u.c.as<int64_t>(), nbc, key.as<uint64_t>());
  k_iota<<<g, NT, 0, s>>>(nt, idx.as<uint32_t>());
  HIPCHK(hipGetLastError());
  sort_pairs_u64_u32(key.as<uint64_t>(), idx.as<uint32_t>(), nt, bits_for(((uint64_t)maxI * nbc + maxJ) * 9 + 8), s);
  flag.alloc(nt * 4);
  seg.alloc(nt * 4);
  bflag.alloc(nt * 4);
  bseg.alloc(nt * 4);
  k_heads<<<g, NT, 0, s>>>(nt, key.as<uint64_t>(), 0, flag.as<int32_t>());
  k_heads<<<g, NT, 0, s>>>(nt, key.as<uint64_t>(), 1, bflag.as<int32_t>());
  HIPCHK(hipGetLastError());
  inclusive_scan_i32(flag.as<int32_t>(), seg.as<int32_t>(), nt, s);
  inclusive_scan_i32(bflag.as<int32_t>(), bseg.as<int32_t>(), nt, s);
  int32_t nrun = 0, nb = 0;
  HIPCHK(hipMemcpy(&nrun, seg.as<int32_t>() + nt - 1, 4, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(&nb, bseg.as<int32_t>() + nt - 1, 4, hipMemcpyDeviceToHost));
  pos.alloc((int64_t)nrun * 8);
  bpos.alloc((int64_t)nb * 8);
  k_head_pos<<<g, NT, 0, s>>>(nt, flag.as<int32_t>(), seg.as<int32_t>(), pos.as<int64_t>());
  k_head_pos<<<g, NT, 0, s>>>(nt, bflag.as<int32_t>(), bseg.as<int32_t>(), bpos.as<int64_t>());
  HIPCHK(hipGetLastError());
  const int64_t nbr = maxI + 1;
  ptr.alloc((nbr + 1) * 4);
  col.alloc((int64_t)nb * 4);
  val.alloc((int64_t)nb * 72);
  HIPCHK(hipMemsetAsync(val.p, 0, (int64_t)nb * 72, s));
  k_bcsr_blocks<<<ceil_div(nb, NT), NT, 0, s>>>(nb, nt, bpos.as<int64_t>(), key.as<uint64_t>(), nbc,
                                                col.as<int32_t>(), ptr.as<int32_t>());
  k_bcsr_runs<<<ceil_div(nrun, NT), NT, 0, s>>>(nrun, nt, nb, pos.as<int64_t>(), key.as<uint64_t>(),
                                                idx.as<uint32_t>(), u.v.as<double>(), bseg.as<int32_t>(),
                                                val.as<double>());
  k_set_i32<<<1, 1, 0, s>>>(ptr.as<int32_t>() + nbr, nb);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(s));
  auto *A = matrix_from_device(ctx, YAFEL_HIP_BCSR3, nbr, (int64_t)nbc, nb, nbr + 1, ptr, col, val, precisions);
  // rows()/cols() as the reference reports them (zero when only one triplet)
  if (nt == 1) {
    A->rows = A->cols = 0;
  } else {
    A->rows = 3 * brows_ref;
    A->cols = 3 * bcols_ref;
  }
  HIPCHK(hipStreamSynchronize(s));
  *out = A;
  return YAFEL_HIP_OK;
  YH_TRY_END(ctx)
}

}  // extern "C"
