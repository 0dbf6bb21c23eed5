// FP64 dot product with a deterministic two-stage reduction
// (replaces VectorExpression::dot, old_handmade_linalg/VectorExpression.hpp:31-55).
#include "internal.hpp"

using namespace yh;

namespace {

constexpr int NT = 256;

__global__ __launch_bounds__(NT) void k_dot_partial(int64_t n, const double *__restrict__ a,
                                                    const double *__restrict__ b, double *part) {
  __shared__ double lds[NT / 64];
  const int64_t chunk = ((n + gridDim.x - 1) / gridDim.x + 63) & ~(int64_t)63;
  const int64_t i0 = (int64_t)dev::xcd_block() * chunk;
  const int64_t i1 = i0 + chunk < n ? i0 + chunk : n;
  double s = 0;
  for (int64_t i = i0 + threadIdx.x; i < i1; i += NT) s = fma(a[i], b[i], s);
  s = dev::block_sum<NT>(s, lds);
  if (threadIdx.x == 0) part[dev::xcd_block()] = s;
}

__global__ __launch_bounds__(NT) void k_dot_final(int npart, const double *part, double *out) {
  __shared__ double lds[NT / 64];
  double s = dev::reduce_partials<NT>(part, npart, 1, lds);
  if (threadIdx.x == 0) *out = s;
}

// STREAM-style copy, 16-byte vector loads/stores, grid-stride
typedef int v4i __attribute__((ext_vector_type(4)));
__global__ __launch_bounds__(NT) void k_copy(int64_t n16, const v4i *__restrict__ a, v4i *__restrict__ b) {
  int64_t i = blockIdx.x * (int64_t)NT + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * NT;
  for (; i < n16; i += stride) b[i] = __builtin_nontemporal_load(&a[i]);
}

// Read-only stream: 16-byte non-temporal loads folded into one word per
// thread (written once, so the loads cannot be dropped).  The ceiling for
// kernels that read far more than they write, such as the SpMV.
__global__ __launch_bounds__(NT) void k_read(int64_t n16, const v4i *__restrict__ a, int *__restrict__ out) {
  int64_t i = blockIdx.x * (int64_t)NT + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * NT;
  v4i acc = {0, 0, 0, 0};
  for (; i < n16; i += stride) acc ^= __builtin_nontemporal_load(&a[i]);
  out[blockIdx.x * NT + threadIdx.x] = acc.x ^ acc.y ^ acc.z ^ acc.w;
}

}  // namespace

extern "C" int yafel_hip_read_bandwidth(yafel_hip_ctx ctx, size_t bytes, int reps, double *gbs) {
  YH_TRY_BEGIN
  REQUIRE(ctx && gbs && reps > 0 && bytes >= 16, "read_bandwidth: bad argument");
  HIPCHK(hipSetDevice(ctx->device));
  const int64_t n16 = (int64_t)(bytes / 16);
  const int grid = ctx->num_cus * 16;
  DBuf a, o;
  a.alloc(n16 * 16);
  o.alloc((size_t)grid * NT * sizeof(int));
  HIPCHK(hipMemsetAsync(a.p, 1, n16 * 16, ctx->stream));
  k_read<<<grid, NT, 0, ctx->stream>>>(n16, a.as<v4i>(), o.as<int>());
  HIPCHK(hipEventRecord(ctx->t0, ctx->stream));
  for (int r = 0; r < reps; ++r) k_read<<<grid, NT, 0, ctx->stream>>>(n16, a.as<v4i>(), o.as<int>());
  HIPCHK(hipGetLastError());
  HIPCHK(hipEventRecord(ctx->t1, ctx->stream));
  HIPCHK(hipEventSynchronize(ctx->t1));
  float ms = 0;
  HIPCHK(hipEventElapsedTime(&ms, ctx->t0, ctx->t1));
  *gbs = (double)n16 * 16 * reps / (ms * 1e-3) / 1e9;
  return YAFEL_HIP_OK;
  YH_TRY_END(ctx)
}

extern "C" int yafel_hip_copy_bandwidth(yafel_hip_ctx ctx, size_t bytes, int reps, double *gbs) {
  YH_TRY_BEGIN
  REQUIRE(ctx && gbs && reps > 0 && bytes >= 16, "copy_bandwidth: bad argument");
  HIPCHK(hipSetDevice(ctx->device));
  const int64_t n16 = (int64_t)(bytes / 16);
  DBuf a, b;
  a.alloc(n16 * 16);
  b.alloc(n16 * 16);
  HIPCHK(hipMemsetAsync(a.p, 1, n16 * 16, ctx->stream));
  const int grid = ctx->num_cus * 16;
  k_copy<<<grid, NT, 0, ctx->stream>>>(n16, a.as<v4i>(), b.as<v4i>());
  HIPCHK(hipEventRecord(ctx->t0, ctx->stream));
  for (int r = 0; r < reps; ++r) k_copy<<<grid, NT, 0, ctx->stream>>>(n16, a.as<v4i>(), b.as<v4i>());
  HIPCHK(hipGetLastError());
  HIPCHK(hipEventRecord(ctx->t1, ctx->stream));
  HIPCHK(hipEventSynchronize(ctx->t1));
  float ms = 0;
  HIPCHK(hipEventElapsedTime(&ms, ctx->t0, ctx->t1));
  *gbs = 2.0 * (double)n16 * 16 * reps / (ms * 1e-3) / 1e9;
  return YAFEL_HIP_OK;
  YH_TRY_END(ctx)
}

extern "C" int yafel_hip_dot(yafel_hip_ctx ctx, int64_t n, const double *a, const double *b,
                             double *result) {
  YH_TRY_BEGIN
  REQUIRE(ctx && result && (n == 0 || (a && b)), "dot: null argument");
  HIPCHK(hipSetDevice(ctx->device));
  if (n == 0) {
    *result = 0;
    return YAFEL_HIP_OK;
  }
  const int grid = grid_for(ctx, n, NT);
  DBuf part;
  part.alloc((grid + 1) * sizeof(double));
  k_dot_partial<<<grid, NT, 0, ctx->stream>>>(n, a, b, part.as<double>());
  HIPCHK(hipGetLastError());
  k_dot_final<<<1, NT, 0, ctx->stream>>>(grid, part.as<double>(), part.as<double>() + grid);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(result, part.as<double>() + grid, 8, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return YAFEL_HIP_OK;
  YH_TRY_END(ctx)
}
