// Internal declarations shared by the HIP translation units of libyafel_hip.so.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <atomic>
#include <string>
#include <utility>
#include <vector>

#include "yafel_hip.h"

namespace yh {

// ---------------------------------------------------------------- errors
struct Error {
  int code;
  std::string msg;
};

void set_thread_error(const std::string &msg);
const char *thread_error();

#define YH_TRY_BEGIN try {
#define YH_TRY_END(ctx)                                   \
  }                                                       \
  catch (const yh::Error &e) {                            \
    yh::record_error(ctx, e.msg);                         \
    return e.code;                                        \
  }                                                       \
  catch (const std::exception &e) {                       \
    yh::record_error(ctx, e.what());                      \
    return YAFEL_HIP_EINVAL;                              \
  }

[[noreturn]] void fail(int code, const std::string &msg);
void hip_check(hipError_t e, const char *what, const char *file, int line);
#define HIPCHK(x) ::yh::hip_check((x), #x, __FILE__, __LINE__)
#define REQUIRE(cond, msg)                                     \
  do {                                                         \
    if (!(cond)) ::yh::fail(YAFEL_HIP_EINVAL, std::string(msg)); \
  } while (0)

// -------------------------------------------------------- device buffers
// RAII device allocation (stream-ordered free on the owning stream).
// Device buffer.  Allocated with hipMalloc, or -- on a thread that has set
// alloc_stream (the serving setup worker) -- stream-ordered from the device's
// memory pool on that stream, so that building one request never waits for
// the kernels other streams are running (hipFree synchronises the device).
struct TempPool;

struct DBuf {
  void *p = nullptr;
  size_t bytes = 0;
  hipStream_t s = nullptr;   // stream of a stream-ordered allocation
  size_t cap = 0;            // allocated bytes (size class) of a stream-ordered allocation
  TempPool *pool = nullptr;  // the allocating thread's pool (stream-ordered allocations)
  DBuf() = default;
  DBuf(const DBuf &) = delete;
  DBuf &operator=(const DBuf &) = delete;
  ~DBuf() { release(); }
  void alloc(size_t b);
  void release();
  void swap(DBuf &o) {
    std::swap(p, o.p);
    std::swap(bytes, o.bytes);
    std::swap(s, o.s);
    std::swap(cap, o.cap);
    std::swap(pool, o.pool);
  }
  template <class T> T *as() const { return reinterpret_cast<T *>(p); }
};

// Stream for the stream-ordered allocations of the calling thread (nullptr:
// plain hipMalloc/hipFree).  With a stream set, buffers the thread releases
// itself are kept in a per-thread pool of at most max_cached bytes for its
// next allocations; resetting frees the pool.
void set_alloc_stream(hipStream_t s, size_t max_cached = 0);

}  // namespace yh

struct yafel_hip_ctx_s;

namespace yh {
void record_error(yafel_hip_ctx_s *ctx, const std::string &msg);
// Contexts are reference counted: every matrix, mesh and server holds one
// reference, and yafel_hip_ctx_destroy drops the caller's.  The stream,
// workspace and pinned memory are freed when the last reference goes, so
// handles may be destroyed in any order.
void ctx_retain(yafel_hip_ctx_s *ctx);
void ctx_release(yafel_hip_ctx_s *ctx);
}  // namespace yh

// Solver workspace: grown on demand, owned by the context.
struct YhWorkspace {
  yh::DBuf vec64[10];   // FP64 n-vectors
  yh::DBuf vec32[10];   // FP32 n-vectors
  yh::DBuf partials;    // per-block partial sums
  yh::DBuf state;       // device solver state
  size_t n64 = 0, n32 = 0;
};

struct yafel_hip_ctx_s {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t t0 = nullptr, t1 = nullptr;
  hipEvent_t sev[2] = {nullptr, nullptr};  // the solvers' chunk events (created with the context)
  std::string err;
  int num_cus = 256;
  YhWorkspace ws;
  void *pinned = nullptr;  // host-pinned status words
  size_t pinned_bytes = 0;
  std::atomic<int> refs{1};
  // per-launch timing of the solvers' inner SpMV kernel (off unless asked for)
  bool ktime = false;
  std::vector<hipEvent_t> kev;  // start/stop pairs of the recorded launches
  size_t kused = 0;             // events in use
  // one word per recorded launch, set to 1 by the launch when it ran the
  // product (a launch after convergence or a refinement pause returns at once)
  yh::DBuf kwork;
  static constexpr size_t kKworkCap = (size_t)1 << 16;
};

namespace yh {
// Looks up, once per device and with that device current, the device address
// of the allocation-free solve's scratch (yafel_hip_solve_noalloc).
void noalloc_scratch_init(int device);
// When kernel timing is on, returns a (start, stop) event pair to record
// around one launch on ctx->stream and the launch's work word (nullptr past
// kKworkCap launches); nullptrs otherwise.
void ktime_pair(yafel_hip_ctx_s *ctx, hipEvent_t *start, hipEvent_t *stop, unsigned **work);
}  // namespace yh

struct YhPlan;  // assembly gather plan (per mesh & format)

// Device storage of both formats is sliced (SELL-64): the stored rows (block
// rows for BCSR3) are cut into slices of 64, one per wavefront, and a slice
// holds its rows' entries slot-major -- entry t of the row at stored position
// i (slice s = i / 64, lane l = i % 64) lives at (sptr[s] + t) * 64 + l.  Rows
// shorter than their slice's longest are padded with zero values whose column
// repeats the row's last column.  Stored positions are the rows themselves, or,
// when sorting rows by length inside windows of 256 saves padding, perm[i] is
// the row at position i and iperm its inverse.  The CSR row pointer is kept
// for the row lengths and the reference layout.
struct yafel_hip_matrix_s {
  yafel_hip_ctx_s *ctx = nullptr;
  int format = YAFEL_HIP_CSR;
  int64_t rows = 0, cols = 0;   // scalar dims
  int64_t brows = 0, bcols = 0; // rows of the stored structure (CSR: = rows)
  int64_t nblk = 0;             // stored blocks (CSR: scalars), padding excluded
  int64_t ptr_len = 0;          // reference host row_ptr length
  int64_t nsl = 0;              // slices of 64 rows
  int64_t ntot = 0;             // sliced entries per value array (64 * slots)
  yh::DBuf ptr;                 // int32 CSR row pointer (brows+1)
  yh::DBuf sptr;                // int32 slot offset of each slice (nsl+1)
  yh::DBuf perm, iperm;         // int32 position->row / row->position (empty: identity)
  yh::DBuf col;                 // int32 sliced block columns (ntot)
  yh::DBuf col16;               // int16 column - stored position (ntot), when every entry fits; else empty
  // CSR-order copy for the FP32 multi-RHS products that stage rows of X in
  // LDS (built on first use): rows in groups of up to 64 grown breadth-first
  // over the pattern, 64 positions per group (rows longest first, then
  // empty), each group's distinct columns, and every entry's position in its
  // group's list
  yh::DBuf mx_ptr;              // int32 (positions+1): entry pointer, every row padded to a multiple of 4 entries
  yh::DBuf mx_src;              // int32: sliced position of each entry in CSR order (-1: padding)
  yh::DBuf mx_lidx;             // uint16: position of the entry's column in its group's list (0: padding)
  yh::DBuf mx_dptr;             // int32 (groups+1): start of each group's list in mx_dcol
  yh::DBuf mx_dcol;             // int32: the groups' distinct columns, ascending within a group
  yh::DBuf mx_val32;            // FP32 values in CSR order, zero padding (refreshed from val32)
  int mx_dmax = -1;             // longest list (-1: not built; 0: not usable)
  int64_t mx_emax = 0;          // most (padded) entries in a group
  int64_t mx_nbig = 0;          // groups over the FP32 staged products' LDS budget (read X directly)
  int64_t v32_gen = 0;          // bumped whenever val32 is rebuilt
  int64_t mx_gen = -1;          // the val32 generation mx_val32 holds
  // direct-write map of the FP32 matrix-core assembly (built on first use):
  // per element entry its sliced position, or -1 - its index in the compact
  // staging of entries shared by several elements; per shared entry its
  // position and the start of its contributions
  yh::DBuf asm_dmap, asm_mpos, asm_mbeg, asm_mstage;
  int64_t asm_nshared = -1;     // shared entries (-1: map not built)
  yh::DBuf mx_grow;             // row of each staged position (64 per group), -1 when empty
  yh::DBuf mx_dl;               // uint16: each position's own column in its group's list (0: empty position)
  int64_t mx_ngr = 0;           // staged groups
  yh::DBuf mx_val32s;           // staged values times their column's D^-1
  int64_t mx_sgen = -1, mx_dgen = -1;  // the val32 and diagonal generations mx_val32s holds
  int64_t diag_gen = 0;         // bumped whenever the diagonal is rebuilt
  // the FP64 staged products' values (CSR order, as mx_val32 / mx_val32s)
  yh::DBuf mx_val64, mx_val64s;
  int64_t v64_gen = 0;          // bumped whenever val64 changes
  int64_t mx_gen64 = -1, mx_sgen64 = -1, mx_dgen64 = -1;  // generations mx_val64 / mx_val64s hold
  yh::DBuf val64, val32;        // CSR: ntot; BCSR3: 9 arrays of ntot (v[e * ntot + k], e = 3i+j)
  yh::DBuf diag64, dinv64, dinv32;
  bool has64 = false, has32 = false, diag_ok = false;
  void *mesh = nullptr;         // mesh this pattern came from (assembly)
  YhPlan *plan = nullptr;
  int unroll = 4;               // slots per step of the plain SpMV kernel
  int64_t cache_bytes = -1;     // bytes of each value precision's stream read with the caching policy (-1: default)
  int cg_passes = 0;            // kernels per single-solve CG iteration (2, 3; 0: by size)
  // reference-layout extras: the host row_ptr may be longer than brows+1
  std::vector<int64_t> ptr_tail;
  ~yafel_hip_matrix_s() { yh::ctx_release(ctx); }
};

namespace yh {

// Default share of a matrix stream (values and columns of one precision)
// read with the caching policy, so that it stays in the 256 MB Infinity Cache
// across the products of a solve while the vectors keep room beside it.
constexpr int64_t kCacheBytesDefault = (int64_t)224 << 20;

// Slots of A's sliced streams for values of vb bytes that fall within
// `budget` bytes (the matrix's own budget when budget < 0).
inline int keep_slots(const yafel_hip_matrix_s *A, int vb, int64_t budget = -1) {
  if (budget < 0) budget = A->cache_bytes < 0 ? kCacheBytesDefault : A->cache_bytes;
  const int64_t per = 64LL * ((A->format == YAFEL_HIP_BCSR3 ? 9 : 1) * vb + (A->col16.p ? 2 : 4));
  const int64_t slots = A->ntot / 64, k = budget / per;
  return (int)(k < slots ? k : slots);
}

// ------------------------------------------------------- launch helpers
inline int ceil_div(int64_t a, int64_t b) { return (int)((a + b - 1) / b); }
// grid size for grid-stride kernels: a multiple of 8 (one share per XCD)
int grid_for(yafel_hip_ctx_s *ctx, int64_t work_items, int block);

// Kernels and host launchers implemented across the .hip files.
// Grid of a sliced kernel with 256-thread blocks: one slice per wavefront
// (every block has work, so the XCD-aware block order keeps all XCDs busy),
// at most `cap` blocks, beyond which wavefronts loop over slices.
inline int sell_grid(int64_t nsl, int cap) {
  const int64_t g = (nsl + 3) / 4;
  return (int)(g < 1 ? 1 : (g < cap ? g : cap));
}
void spmv_launch(yafel_hip_matrix_s *A, int precision, const void *x, void *y, hipStream_t s);
void spmm_launch(yafel_hip_matrix_s *A, int precision, int k, const void *X, void *Y,
                 hipStream_t s);
// FP32 CSR Y = A X with the slices' X rows staged in LDS (k <= 64); false when
// the matrix does not qualify and the caller must use another kernel.
bool spmm_lds_launch(yafel_hip_matrix_s *A, int k, const float *X, float *Y, hipStream_t s);
bool spmm_lds64_launch(yafel_hip_matrix_s *A, int k, const double *X, double *Y, hipStream_t s);
void matrix_build_fp32(yafel_hip_matrix_s *A);
void matrix_build_diag(yafel_hip_matrix_s *A);
// Converts CSR-ordered device arrays (int32 ptr[brows+1], col[nblk], FP64
// values: CSR nblk, BCSR3 9 entry-major arrays of nblk) into A's sliced
// storage; ptr is kept, col and values are replaced.
void matrix_make_sliced(yafel_hip_matrix_s *A, DBuf &ptr, DBuf &col, DBuf &val64);

}  // namespace yh

// ----------------------------------------------------------------- device
namespace yh {
namespace dev {

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ __forceinline__ float wave_sumf(float v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
template <int L, class T>
__device__ __forceinline__ T group_sum(T v) {
#pragma unroll
  for (int o = L / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Block-wide sum (blockDim.x a multiple of 64, at most 1024); result valid in
// all threads.  Fixed combination order -> deterministic.
template <int NT>
__device__ __forceinline__ double block_sum(double v, double *lds /* >= NT/64 */) {
  v = wave_sum(v);
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  __syncthreads();
  if (l == 0) lds[w] = v;
  __syncthreads();
  double t = 0;
#pragma unroll
  for (int i = 0; i < NT / 64; ++i) t += lds[i];
  return t;
}

// XCD-aware logical block id: consecutive logical ids land on one XCD
// (hardware dispatches workgroups round-robin over the 8 XCDs).
__device__ __forceinline__ int xcd_block() {
  const int G = gridDim.x, b = blockIdx.x;
  if ((G & 7) != 0) return b;
  return (b & 7) * (G >> 3) + (b >> 3);
}

// Deterministic sum of npart partials (fixed order), executed by one block.
template <int NT>
__device__ __forceinline__ double reduce_partials(const double *part, int npart, int stride,
                                                  double *lds) {
  double s = 0;
  for (int i = threadIdx.x; i < npart; i += NT) s += part[(size_t)i * stride];
  return block_sum<NT>(s, lds);
}

}  // namespace dev
}  // namespace yh
