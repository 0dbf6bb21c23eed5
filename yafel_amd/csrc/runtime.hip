// Context, errors, device memory helpers, timers and the synthetic-input
// generators of the C ABI.
#include <cmath>
#include <stdexcept>
#include <unordered_map>
#include <vector>

#include "internal.hpp"

namespace yh {

static thread_local std::string g_thread_err;

void set_thread_error(const std::string &msg) { g_thread_err = msg; }
const char *thread_error() { return g_thread_err.c_str(); }

void fail(int code, const std::string &msg) { throw Error{code, msg}; }

void hip_check(hipError_t e, const char *what, const char *file, int line) {
  if (e == hipSuccess) return;
  char buf[512];
  snprintf(buf, sizeof(buf), "HIP error %d (%s) in %s at %s:%d", (int)e, hipGetErrorString(e), what,
           file, line);
  fail(e == hipErrorOutOfMemory ? YAFEL_HIP_ENOMEM : YAFEL_HIP_EDEVICE, buf);
}

void record_error(yafel_hip_ctx_s *ctx, const std::string &msg) {
  set_thread_error(msg);
  if (ctx) ctx->err = msg;
}

static thread_local hipStream_t g_alloc_stream = nullptr;

// A thread with an allocation stream (a serving setup worker) keeps the
// buffers it releases itself in size classes and hands them out again: its
// work is ordered on that one stream, so a buffer freed by one request's
// set-up is ready for the next one's without a sync.  Buffers released on
// another thread (a finished request's matrix) are freed stream-ordered.
struct TempPool {
  hipStream_t s = nullptr;
  std::unordered_map<size_t, std::vector<void *>> free;
  size_t cached = 0;
  size_t max_cached = 0;  // bytes the pool may hold
  void *take(size_t sc) {
    auto it = free.find(sc);
    if (it == free.end() || it->second.empty()) return nullptr;
    void *p = it->second.back();
    it->second.pop_back();
    cached -= sc;
    return p;
  }
  bool put(void *p, size_t sc) {
    if (cached + sc > max_cached) return false;
    free[sc].push_back(p);
    cached += sc;
    return true;
  }
  ~TempPool() {
    for (auto &kv : free)
      for (void *p : kv.second) (void)hipFreeAsync(p, s);
  }
};
static thread_local TempPool *g_pool = nullptr;

static size_t size_class(size_t b) {
  size_t sc = 4096;
  while (sc < b) sc <<= 1;
  return sc;
}

void set_alloc_stream(hipStream_t s, size_t max_cached) {
  delete g_pool;
  g_pool = nullptr;
  g_alloc_stream = s;
  if (s) {
    g_pool = new TempPool();
    g_pool->s = s;
    g_pool->max_cached = max_cached;
  }
}

void DBuf::alloc(size_t b) {
  release();
  if (b == 0) return;
  if (g_alloc_stream) {
    const size_t sc = size_class(b);
    p = g_pool ? g_pool->take(sc) : nullptr;
    if (!p) HIPCHK(hipMallocAsync(&p, sc, g_alloc_stream));
    s = g_alloc_stream;
    cap = sc;
    pool = g_pool;
  } else {
    HIPCHK(hipMalloc(&p, b));
  }
  bytes = b;
}
void DBuf::release() {
  if (p) {
    if (pool && pool == g_pool && s == g_pool->s && g_pool->put(p, cap)) {
      // kept for this thread's next allocation of the same class
    } else if (s) {
      (void)hipFreeAsync(p, s);
    } else {
      (void)hipFree(p);
    }
  }
  p = nullptr;
  bytes = 0;
  s = nullptr;
  cap = 0;
  pool = nullptr;
}

int grid_for(yafel_hip_ctx_s *ctx, int64_t work_items, int block) {
  int64_t want = (work_items + block - 1) / block;
  int64_t cap = (int64_t)ctx->num_cus * 8;
  int64_t g = want < cap ? want : cap;
  if (g < 1) g = 1;
  if (g >= 8) g = (g / 8) * 8;
  return (int)g;
}

}  // namespace yh

namespace yh {

void ktime_pair(yafel_hip_ctx_s *ctx, hipEvent_t *start, hipEvent_t *stop, unsigned **work) {
  *start = *stop = nullptr;
  *work = nullptr;
  if (!ctx->ktime) return;
  const size_t idx = ctx->kused / 2;
  if (idx < yafel_hip_ctx_s::kKworkCap) *work = ctx->kwork.as<unsigned>() + idx;
  while (ctx->kev.size() < ctx->kused + 2) {
    hipEvent_t e;
    HIPCHK(hipEventCreate(&e));
    ctx->kev.push_back(e);
  }
  *start = ctx->kev[ctx->kused];
  *stop = ctx->kev[ctx->kused + 1];
  ctx->kused += 2;
}

static void ctx_free(yafel_hip_ctx_s *ctx) {
  (void)hipSetDevice(ctx->device);
  for (hipEvent_t e : ctx->kev) (void)hipEventDestroy(e);
  (void)hipStreamSynchronize(ctx->stream);
  for (auto &b : ctx->ws.vec64) b.release();
  for (auto &b : ctx->ws.vec32) b.release();
  ctx->ws.partials.release();
  ctx->ws.state.release();
  if (ctx->pinned) (void)hipHostFree(ctx->pinned);
  (void)hipEventDestroy(ctx->t0);
  (void)hipEventDestroy(ctx->t1);
  for (hipEvent_t e : ctx->sev)
    if (e) (void)hipEventDestroy(e);
  (void)hipStreamDestroy(ctx->stream);
  delete ctx;
}

void ctx_retain(yafel_hip_ctx_s *ctx) {
  if (ctx) ctx->refs.fetch_add(1, std::memory_order_relaxed);
}

void ctx_release(yafel_hip_ctx_s *ctx) {
  if (ctx && ctx->refs.fetch_sub(1, std::memory_order_acq_rel) == 1) ctx_free(ctx);
}

}  // namespace yh

using namespace yh;

extern "C" {

int yafel_hip_ctx_create(int device, yafel_hip_ctx *out) {
  YH_TRY_BEGIN
  REQUIRE(out, "yafel_hip_ctx_create: out is NULL");
  int ndev = 0;
  HIPCHK(hipGetDeviceCount(&ndev));
  REQUIRE(device >= 0 && device < ndev, "yafel_hip_ctx_create: no such device");
  HIPCHK(hipSetDevice(device));
  auto *c = new yafel_hip_ctx_s();
  c->device = device;
  hipDeviceProp_t prop;
  HIPCHK(hipGetDeviceProperties(&prop, device));
  c->num_cus = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
  HIPCHK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  HIPCHK(hipEventCreate(&c->t0));
  HIPCHK(hipEventCreate(&c->t1));
  HIPCHK(hipEventCreateWithFlags(&c->sev[0], hipEventDisableTiming));
  HIPCHK(hipEventCreateWithFlags(&c->sev[1], hipEventDisableTiming));
  c->pinned_bytes = 65536;
  HIPCHK(hipHostMalloc(&c->pinned, c->pinned_bytes, hipHostMallocDefault));
  noalloc_scratch_init(device);
  *out = c;
  return YAFEL_HIP_OK;
  YH_TRY_END(nullptr)
}

int yafel_hip_ctx_destroy(yafel_hip_ctx ctx) {
  yh::ctx_release(ctx);
  return YAFEL_HIP_OK;
}

const char *yafel_hip_last_error(yafel_hip_ctx ctx) {
  if (ctx) return ctx->err.c_str();
  return thread_error();
}

int yafel_hip_ctx_sync(yafel_hip_ctx ctx) {
  YH_TRY_BEGIN
  REQUIRE(ctx, "null context");
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return YAFEL_HIP_OK;
  YH_TRY_END(ctx)
}

void *yafel_hip_ctx_stream(yafel_hip_ctx ctx) { return ctx ? (void *)ctx->stream : nullptr; }

int yafel_hip_malloc(yafel_hip_ctx ctx, size_t bytes, void **dptr) {
  YH_TRY_BEGIN
  REQUIRE(ctx && dptr, "yafel_hip_malloc: null argument");
  *dptr = nullptr;
  if (bytes == 0) bytes = 256;
  HIPCHK(hipMalloc(dptr, bytes));
  return YAFEL_HIP_OK;
  YH_TRY_END(ctx)
}

int yafel_hip_free(yafel_hip_ctx ctx, void *dptr) {
  YH_TRY_BEGIN
  REQUIRE(ctx, "yafel_hip_free: null context");
  if (dptr) {
    HIPCHK(hipStreamSynchronize(ctx->stream));
    HIPCHK(hipFree(dptr));
  }
  return YAFEL_HIP_OK;
  YH_TRY_END(ctx)
}

int yafel_hip_memcpy_h2d(yafel_hip_ctx ctx, void *dst, const void *src, size_t bytes) {
  YH_TRY_BEGIN
  REQUIRE(ctx && (bytes == 0 || (dst && src)), "yafel_hip_memcpy_h2d: null argument");
  if (bytes) {
    HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
  }
  return YAFEL_HIP_OK;
  YH_TRY_END(ctx)
}

int yafel_hip_memcpy_d2h(yafel_hip_ctx ctx, void *dst, const void *src, size_t bytes) {
  YH_TRY_BEGIN
  REQUIRE(ctx && (bytes == 0 || (dst && src)), "yafel_hip_memcpy_d2h: null argument");
  if (bytes) {
    HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
  }
  return YAFEL_HIP_OK;
  YH_TRY_END(ctx)
}

int yafel_hip_memcpy_d2d(yafel_hip_ctx ctx, void *dst, const void *src, size_t bytes) {
  YH_TRY_BEGIN
  REQUIRE(ctx && (bytes == 0 || (dst && src)), "yafel_hip_memcpy_d2d: null argument");
  if (bytes) HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, ctx->stream));
  return YAFEL_HIP_OK;
  YH_TRY_END(ctx)
}

int yafel_hip_memzero(yafel_hip_ctx ctx, void *dst, size_t bytes) {
  YH_TRY_BEGIN
  REQUIRE(ctx && (bytes == 0 || dst), "yafel_hip_memzero: null argument");
  if (bytes) HIPCHK(hipMemsetAsync(dst, 0, bytes, ctx->stream));
  return YAFEL_HIP_OK;
  YH_TRY_END(ctx)
}

int yafel_hip_timer_begin(yafel_hip_ctx ctx) {
  YH_TRY_BEGIN
  REQUIRE(ctx, "null context");
  HIPCHK(hipEventRecord(ctx->t0, ctx->stream));
  return YAFEL_HIP_OK;
  YH_TRY_END(ctx)
}

int yafel_hip_timer_end(yafel_hip_ctx ctx, double *ms) {
  YH_TRY_BEGIN
  REQUIRE(ctx && ms, "null argument");
  HIPCHK(hipEventRecord(ctx->t1, ctx->stream));
  HIPCHK(hipEventSynchronize(ctx->t1));
  float f = 0;
  HIPCHK(hipEventElapsedTime(&f, ctx->t0, ctx->t1));
  *ms = f;
  return YAFEL_HIP_OK;
  YH_TRY_END(ctx)
}

int yafel_hip_kernel_timing(yafel_hip_ctx ctx, int enable) {
  YH_TRY_BEGIN
  REQUIRE(ctx, "kernel_timing: null context");
  HIPCHK(hipStreamSynchronize(ctx->stream));
  ctx->ktime = enable != 0;
  ctx->kused = 0;
  if (ctx->ktime) {
    const size_t b = yafel_hip_ctx_s::kKworkCap * sizeof(unsigned);
    if (ctx->kwork.bytes < b) ctx->kwork.alloc(b);
    HIPCHK(hipMemsetAsync(ctx->kwork.p, 0, b, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
  }
  return YAFEL_HIP_OK;
  YH_TRY_END(ctx)
}

int yafel_hip_kernel_time(yafel_hip_ctx ctx, double *ms_total, int64_t *launches) {
  YH_TRY_BEGIN
  REQUIRE(ctx && ms_total && launches, "kernel_time: null argument");
  HIPCHK(hipStreamSynchronize(ctx->stream));
  const size_t nl = ctx->kused / 2;
  std::vector<unsigned> work(std::min(nl, yafel_hip_ctx_s::kKworkCap));
  if (!work.empty())
    HIPCHK(hipMemcpy(work.data(), ctx->kwork.p, work.size() * sizeof(unsigned), hipMemcpyDeviceToHost));
  double t = 0;
  int64_t nw = 0;
  for (size_t i = 0; i < work.size(); ++i) {
    if (!work[i]) continue;  // returned at once: not a product
    float f = 0;
    HIPCHK(hipEventElapsedTime(&f, ctx->kev[2 * i], ctx->kev[2 * i + 1]));
    t += f;
    ++nw;
  }
  *ms_total = t;
  *launches = nw;
  return YAFEL_HIP_OK;
  YH_TRY_END(ctx)
}

// Structured unit-domain corner mesh.  Corner nodes x fastest; cells in
// lexicographic order; corner order lexicographic for p = 1 (the local node
// order the p = 1 shape functions use) and gmsh counter-clockwise for p >= 2
// (the order DoFManager's corner interpolation uses).
int yafel_hip_structured_mesh(int dim, int nx, int ny, int nz, int p, double *nodes, int32_t *cells) {
  YH_TRY_BEGIN
  REQUIRE(dim == 2 || dim == 3, "structured_mesh: dim must be 2 or 3");
  REQUIRE(nx > 0 && ny > 0 && (dim == 2 || nz > 0), "structured_mesh: cell counts must be > 0");
  REQUIRE(nodes && cells, "structured_mesh: null output");
  if (dim == 2) nz = 0;
  const int64_t sx = nx + 1, sy = ny + 1;
  for (int64_t k = 0; k <= nz; ++k)
    for (int64_t j = 0; j <= ny; ++j)
      for (int64_t i = 0; i <= nx; ++i) {
        int64_t n = i + sx * (j + sy * k);
        nodes[3 * n + 0] = (double)i / nx;
        nodes[3 * n + 1] = (double)j / ny;
        nodes[3 * n + 2] = dim == 3 ? (double)k / nz : 0.0;
      }
  const bool ccw = p >= 2;
  auto id = [&](int64_t i, int64_t j, int64_t k) { return (int32_t)(i + sx * (j + sy * k)); };
  int64_t c = 0;
  const int nl = dim == 3 ? 2 : 1;
  for (int64_t k = 0; k < (dim == 3 ? nz : 1); ++k)
    for (int64_t j = 0; j < ny; ++j)
      for (int64_t i = 0; i < nx; ++i, ++c) {
        int32_t *o = &cells[(dim == 3 ? 8 : 4) * c];
        for (int l = 0; l < nl; ++l) {
          o[4 * l + 0] = id(i, j, k + l);
          o[4 * l + 1] = id(i + 1, j, k + l);
          o[4 * l + 2] = ccw ? id(i + 1, j + 1, k + l) : id(i, j + 1, k + l);
          o[4 * l + 3] = ccw ? id(i, j + 1, k + l) : id(i + 1, j + 1, k + l);
        }
      }
  return YAFEL_HIP_OK;
  YH_TRY_END(nullptr)
}

// splitmix64 (Steele, Lea & Flood, 2014) feeding Box-Muller pairs.  No FMA
// contraction, so every host produces the same deviates for a seed.
#pragma clang fp contract(off)
static inline uint64_t splitmix64(uint64_t &s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

int yafel_hip_normal_fill(uint64_t seed, int64_t n, double mean, double sd, double *out) {
  YH_TRY_BEGIN
  REQUIRE(n == 0 || out, "normal_fill: null output");
  uint64_t s = seed;
  const double two_pi = 6.283185307179586476925286766559;
  for (int64_t i = 0; i < n; i += 2) {
    const double u1 = ((splitmix64(s) >> 11) + 1) * (1.0 / 9007199254740992.0);
    const double u2 = (splitmix64(s) >> 11) * (1.0 / 9007199254740992.0);
    const double r = std::sqrt(-2.0 * std::log(u1));
    out[i] = mean + sd * r * std::cos(two_pi * u2);
    if (i + 1 < n) out[i + 1] = mean + sd * r * std::sin(two_pi * u2);
  }
  return YAFEL_HIP_OK;
  YH_TRY_END(nullptr)
}

}  // extern "C"
