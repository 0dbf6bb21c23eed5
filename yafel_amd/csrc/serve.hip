// Device-resident request loop with continuous batching.
//
// Every admitted request owns a slot: its DoF numbering, pattern, assembled
// FP64 + FP32 matrix, Jacobi diagonal and FP32 Krylov vectors stay on the
// device.  A table of slot descriptors drives batched launches: one launch
// covers every active slot, each slot on its own contiguous block range
// (block counts fixed by the slot's size alone) with its own FP64 scalars and
// its own folds of its partial sums -- so a request's result is bit-identical
// whether it runs alone or in a mix.  A CG iteration is the single-solve
// three-pass iteration (Q: SpMV, U and P: folds in every workgroup of the
// slot) with each slot's state in a versioned pair; every slot of the table
// takes part in every versioned launch, so one parity serves them all.
//
// Two host threads drive the device.  A setup worker owned by the server takes
// queued requests in FIFO order as the HBM budget and the slot limit allow and
// builds each one (DoF numbering, pattern, assembly, Dirichlet, FP32 copy,
// diagonal) on its own setup stream, with stream-ordered allocations so that
// it never waits on the solve stream.  The caller's thread (progress) joins
// the slots the worker finished, keeps up to four chunks of batched iterations
// in flight on the solve stream, and retires slots whose solve finished.  A
// server is driven from one caller thread at a time.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <thread>

#include "krylov_ops.hpp"
#include "mesh.hpp"

using namespace yh;
using namespace yh::kr;

namespace {

constexpr int GCAP = 1024;  // block cap per slot and phase

struct SlotDesc {
  int fmt, pad0;
  int g_spmv, g_elem;
  int off_spmv, off_elem;
  int64_t n, brows, ntot;
  const int *sptr, *perm, *col;
  const short *col16;  // 16-bit column offsets (nullptr: col)
  int keep32;          // leading slots of the FP32 stream read with the caching policy
  int pad1;
  const float *v32;
  const double *v64;
  const float *dinv;
  float *r, *z, *p, *q, *xl;
  float *p1;  // the second p buffer of the two-pass iteration
  double *x;
  const double *b;
  KState *st;  // two versions: st[0], st[1]
  double *part;
  double tol;
  long long max_iter;
};

__device__ __forceinline__ int find_slot(const SlotDesc *d, int ns, int b, bool spmv) {
  int lo = 0, hi = ns - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    const int off = spmv ? d[mid].off_spmv : d[mid].off_elem;
    if (off <= b) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

enum { PH_INIT = 0, PH_Q = 1, PH_U = 2, PH_P = 3, PH_P0 = 4, PH_FOLD = 5, PH_REF = 6, PH_QP = 7, PH_PN = 8 };

template <int PH>
constexpr bool spmv_phase() {
  return PH == PH_INIT || PH == PH_Q || PH == PH_REF || PH == PH_QP;
}

template <class O>
__device__ __forceinline__ void bind(O &op, const SlotDesc &d, int par) {
  op.st = d.st + par;
  op.part = d.part;
  op.counter = d.st[par].counter;
}

// a versioned pass of the slot: state version par in, par ^ 1 out
template <class O>
__device__ __forceinline__ void vbind(O &op, const SlotDesc &d, int par, const double *pin, int nin,
                                      double *ppart) {
  op.in = d.st + par;
  op.out = d.st + (par ^ 1);
  op.pin = pin;
  op.nin = nin;
  op.part = ppart;
}

// partial regions of a slot: the SpMV's at 0, the update pass's at 2 gmax
__device__ __forceinline__ double *upart(const SlotDesc &d) {
  return d.part + 2 * (size_t)(d.g_spmv > d.g_elem ? d.g_spmv : d.g_elem);
}

// The FP64 stream is read once per refinement: non-temporal throughout.
template <class T, class O>
__device__ __forceinline__ void run_spmv(const SlotDesc &d, const T *val, O &op, int blk) {
  const SellView<T> A{d.sptr, d.perm, d.col, val, d.brows, d.ntot, d.col16, sizeof(T) == 4 ? d.keep32 : 0};
  if (d.fmt == YAFEL_HIP_BCSR3) spmv_fused_body<YAFEL_HIP_BCSR3, kSellUnroll, T, O>(A, op, blk, d.g_spmv);
  else spmv_fused_body<YAFEL_HIP_CSR, kSellUnroll, T, O>(A, op, blk, d.g_spmv);
}

// The closing pass of a two-pass chunk (versioned: state par in, par ^ 1
// out): the three-pass P of the chunk's last iteration -- U's partials
// folded, the stopping test and beta, x += alpha p and p = z + beta p -- with
// the p that the QP passes left in either buffer read and the new p written
// into the first, and pidx reset, so that the chunk ends in the state a
// three-pass chunk leaves (p in d.p).  A slot that stopped or paused in a QP
// pass (its U skipped) only has its p moved.  Its arithmetic is OpCgP's.
// Unlike velem_body it reads the state before the vectors: which buffer holds
// p is known only from the state.
__device__ __forceinline__ void cg_close_body(const SlotDesc &d, int blk, int nblk, int par) {
  __shared__ KState ss;
  __shared__ int sfl;
  constexpr int kX = 1, kP = 2, kCopy = 4;
  const int t = threadIdx.x;
  double f[2];
  refold_load<2>(upart(d), d.g_elem, f);
  const KState *in = d.st + par;
  KState *out = d.st + (par ^ 1);
  if (t < KSW) reinterpret_cast<long long *>(&ss)[t] = reinterpret_cast<const long long *>(in)[t];
  __syncthreads();
  const bool run = VOpBase::running(ss);
  if (run) refold_reduce<2>(f, f);
  __syncthreads();  // every wave has read ss before thread 0 changes it
  if (t == 0) {
    int fl = 0;
    const int pi = ss.pidx;
    if (run) {
      after_iteration(&ss, f[0], f[1]);
      fl |= kX;
      if (VOpBase::running(ss)) fl |= kP;
    }
    if (!(fl & kP) && pi) fl |= kCopy;
    ss.pidx = 0;
    sfl = fl | (pi << 4);
  }
  __syncthreads();
  if (blk == 0 && t < KSW) reinterpret_cast<long long *>(out)[t] = reinterpret_cast<const long long *>(&ss)[t];
  const int fl = sfl;
  if (!(fl & (kX | kCopy))) return;
  const float *pold = (fl >> 4) ? d.p1 : d.p;
  const float alpha = (float)ss.alpha, beta = (float)ss.beta;
  const int64_t n = d.n, chunk = (n + nblk - 1) / nblk;
  const int64_t i1 = (int64_t)blk * chunk + chunk < n ? (int64_t)blk * chunk + chunk : n;
  for (int64_t i = (int64_t)blk * chunk + t; i < i1; i += NT) {
    const float po = pold[i];
    if (fl & kX) d.xl[i] = fma(alpha, po, d.xl[i]);
    if (fl & kP) d.p[i] = fma(beta, po, d.z[i]);
    else if (fl & kCopy) d.p[i] = po;
  }
}

// Builds the phase's operation for slot d and either runs the phase body on
// block blk (FIN = false) or folds the slot's partials (FIN = true; the
// residual phases only).  par is the state version of the launch.
template <int PH, bool FIN>
__device__ __forceinline__ void slot_phase(const SlotDesc &d, int blk, int par) {
  constexpr bool SP = spmv_phase<PH>();
  const int nblk = SP ? d.g_spmv : d.g_elem;
  if (PH == PH_INIT || PH == PH_REF) {
    OpResidual<float, PH == PH_INIT ? 0 : 1> op{};
    bind(op, d, par);
    op.xh = d.x;
    op.b = d.b;
    op.r = d.r;
    op.z = d.z;
    op.dinv = d.dinv;
    op.rhat = nullptr;
    op.wkind = 1;
    if (FIN) finish_body(op, nblk);
    else run_spmv<double>(d, d.v64, op, blk);
  } else if (PH == PH_Q) {
    OpCgQ<float> op{};
    bind(op, d, par);
    op.p = d.p;
    op.q = d.q;
    if (!FIN) run_spmv<float>(d, d.v32, op, blk);
  } else if (PH == PH_QP) {
    // two-pass iteration: U's partials folded, beta and the stopping test,
    // x += alpha p, q = A (z + beta p) with p_new written into the other p
    OpCgQP<float> op{};
    vbind(op, d, par, upart(d), d.g_elem, d.part);
    op.p0 = d.p;
    op.p1 = d.p1;
    op.q = d.q;
    op.xv = d.xl;
    op.z = d.z;
    if (!FIN) {
      const SellView<float> A{d.sptr, d.perm, d.col, d.v32, d.brows, d.ntot, d.col16, d.keep32};
      if (d.fmt == YAFEL_HIP_BCSR3) spmv_qp_body<YAFEL_HIP_BCSR3, kSellUnroll, float>(A, op, blk, d.g_spmv, d.n);
      else spmv_qp_body<YAFEL_HIP_CSR, kSellUnroll, float>(A, op, blk, d.g_spmv, d.n);
    }
  } else if (PH == PH_PN) {
    if (!FIN) cg_close_body(d, blk, nblk, par);
  } else if (PH == PH_U) {
    OpCgU<float, true> op{};
    vbind(op, d, par, d.part, d.g_spmv, upart(d));
    op.r = d.r;
    op.z = d.z;
    op.q = d.q;
    op.dinv = d.dinv;
    if (!FIN) velem_body(d.n, op, blk, nblk);
  } else if (PH == PH_P) {
    OpCgP<float> op{};
    vbind(op, d, par, upart(d), d.g_elem, nullptr);
    op.p = d.p;
    op.x = d.xl;
    op.z = d.z;
    if (!FIN) velem_body(d.n, op, blk, nblk);
  } else if (PH == PH_P0) {
    OpCgP0<float> op{};
    vbind(op, d, par, nullptr, 0, nullptr);
    op.p = d.p;
    op.z = d.z;
    if (!FIN) velem_body(d.n, op, blk, nblk);
  } else if (!FIN) {
    OpFold<float> op{};
    bind(op, d, par);
    op.xh = d.x;
    op.xl = d.xl;
    elem_body(d.n, op, blk, nblk);
  }
}

// Replayed graphs take the table's slot count and block totals from here
// (uploaded with the table), so that one capture serves every slot set whose
// totals fit its grid; blocks past the totals return at once.
struct TableHdr {
  int ns, ts, te, pad;
};

template <int PH>
__global__ __launch_bounds__(NT) void kb_phase(const SlotDesc *__restrict__ slots, int ns, int par,
                                               const TableHdr *__restrict__ hdr) {
  constexpr bool SP = spmv_phase<PH>();
  if (hdr) {
    ns = hdr->ns;
    if ((int)blockIdx.x >= (SP ? hdr->ts : hdr->te)) return;
  }
  const int s = find_slot(slots, ns, blockIdx.x, SP);
  const SlotDesc d = slots[s];
  slot_phase<PH, false>(d, blockIdx.x - (SP ? d.off_spmv : d.off_elem), par);
}

// one workgroup per slot folds that slot's partials
template <int PH>
__global__ __launch_bounds__(NTF) void kb_finish(const SlotDesc *__restrict__ slots, int ns, int par,
                                                 const TableHdr *__restrict__ hdr) {
  if (hdr && (int)blockIdx.x >= hdr->ns) return;
  const SlotDesc d = slots[blockIdx.x];
  slot_phase<PH, true>(d, 0, par);
}

// zero a joining slot's solver state (version par) and set its configuration
__global__ void kb_init_state(const SlotDesc *slots, int ns, int par) {
  const int j = blockIdx.x;
  if (j >= ns) return;
  KState *st = slots[j].st + par;
  unsigned int *w = (unsigned int *)st;
  for (int i = threadIdx.x; i < (int)(sizeof(KState) / 4); i += blockDim.x) w[i] = 0u;
  __syncthreads();
  if (threadIdx.x == 0) {
    st->tol = slots[j].tol;
    st->stop = YAFEL_HIP_STOP_RELRES;
    st->delta2 = 0.01;
    st->max_iter = slots[j].max_iter;
    st->method = 1;  // Jacobi-PCG
    st->mixed = 1;
  }
}

struct SlotStatus {
  int done, status;
  long long iter, refinements;
  double crit;
};

__global__ void kb_status(const SlotDesc *slots, int ns, int par, SlotStatus *out, const TableHdr *hdr) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= (hdr ? hdr->ns : ns)) return;
  const KState *st = slots[j].st + par;
  out[j].done = st->done;
  out[j].status = st->status;
  out[j].iter = st->iter;
  out[j].refinements = st->refinements;
  out[j].crit = st->crit;
}

struct Req {
  int64_t id;
  int state = YAFEL_HIP_REQ_QUEUED;
  int dim, p, dpn, bc;
  double tol;
  int64_t max_iter;
  std::vector<double> nodes, coef, load;
  std::vector<int32_t> cells;
  yafel_hip_physics phys;
  size_t est_bytes = 0;
  int64_t ndofs = 0;
  std::vector<double> x;
  yafel_hip_solve_info info{};
  std::string error;
  // host time stamps: submitted, taken by a setup worker, set up, joined the
  // batch (its first chunk launched), found finished
  std::chrono::steady_clock::time_point t_submit, t_admit, t_ready, t_join, t_done;
};

struct Slot {
  Req *req = nullptr;
  yafel_hip_matrix A = nullptr;
  DBuf b, mask;
  void *buf = nullptr;  // one stream-ordered allocation for the solve vectors and state
  size_t buf_bytes = 0;
  SlotDesc d{};
  int64_t ndofs = 0;
  int64_t last_chunk = 0;  // sequence number of the last chunk whose table lists the slot
  bool joined = false, retired = false;
};

template <class T>
T *carve(char *&p, size_t count) {
  T *r = (T *)p;
  p += ((count * sizeof(T)) + 255) & ~(size_t)255;
  return r;
}

}  // namespace

struct yafel_hip_server_s {
  yafel_hip_ctx_s *ctx = nullptr;   // solve stream
  std::vector<yafel_hip_ctx_s *> sctx;  // one setup context (stream) per setup worker
  size_t budget = 0, in_use = 0, peak = 0;
  int max_slots = 0;
  int64_t next_id = 1;
  std::map<int64_t, std::unique_ptr<Req>> reqs;
  std::deque<Req *> queue;
  std::vector<Slot *> active;
  // Retired slots whose matrix, load and mask may still be read by a chunk in
  // flight: a slot found finished when chunk N is drained was already listed
  // in chunk N + 1.  They are freed once the last chunk listing them is drained.
  std::vector<Slot *> grave;
  int64_t chunk_seq = 0;    // chunks launched
  int64_t drained_seq = 0;  // the last chunk drained (chunks drain in launch order)
  yafel_hip_server_stats stats{};
  // descriptor / status rings: each progress call uses at most two entries and
  // at most kDepth calls' work is in flight, so an entry is never rewritten
  // while a launch that reads it (or a copy from its staging area) can still
  // run, nor a status area before the host has read it
  static constexpr int kDepth = 4;  // chunks in flight
  static constexpr int RING = 16;
  static_assert(RING > 2 * (kDepth + 1), "server: descriptor ring too short for the chunks in flight");
  static constexpr size_t STAGE = 131072;  // pinned bytes per ring entry (table / statuses)
  DBuf desc[RING], status[RING];
  int ring = 0;
  struct Inflight {
    hipEvent_t ev;
    int64_t seq;
    int ring;
    std::vector<Slot *> slots;
    char *pin;
    int gr;  // graph replay: the status slot it uses (-1: direct)
  };
  std::deque<Inflight> inflight;
  hipEvent_t evs[RING];
  // Graph replay (off unless yafel_hip_server_set_graphs turns it on): a
  // chunk's launches are replayed from an executable graph whose kernels read
  // the slot table and its totals from gdesc / ghdr.  The table is uploaded on
  // the stream when the slot set changes (a join or a retirement ends the
  // epoch), so that needs no new capture; the grids are captured with caps on
  // the table's totals, and a slot set over a cap (or far below it) is
  // captured again.  One graph per chunk form (three or two passes) and
  // status slot r: slot r's chunks start at state-version parity r & 1 (a
  // chunk advances the version by an odd count), and each slot has its own
  // status buffer, pinned status area and event, busy while its chunk is in
  // flight.
  bool graphs = false;
  static constexpr int GR = 4;
  static_assert(GR >= kDepth, "server: a status slot per chunk in flight");
  hipGraphExec_t gexec[2][GR] = {};
  bool gbusy[GR] = {};
  int gcap_s = 0, gcap_e = 0;  // the captured grids (block totals)
  int64_t epoch = 0;           // bumped at every join and retirement
  int64_t gepoch = -1;         // the epoch whose table gdesc / ghdr hold
  // graphs replaced by a capture with other caps, with the last chunk launched
  // while they were current: destroyed once that chunk has drained
  std::vector<std::pair<int64_t, hipGraphExec_t>> gold;
  DBuf gdesc, ghdr, gstatus[GR];
  char *gpin = nullptr;        // GR status areas
  hipEvent_t gev[GR] = {};
  int64_t captures = 0, replays = 0;
  int64_t stats_two = 0, stats_three = 0;  // chunks of each form
  char *pin = nullptr;  // RING status areas, then RING table staging areas
  int K = 8;
  // direct submission: CG chunks as two passes an iteration (QP, U) or three
  // (yafel_hip_server_set_cg_passes).  Two cost a launch less but gather two
  // vectors, which pays while the batch is small, so by default (passes 0) a
  // chunk takes two passes while its active slots store at most
  // two_pass_limit scalars in all.
  int passes = 0;
  int64_t two_pass_limit = kTwoPassLimit;
  static constexpr int64_t kTwoPassLimit = (int64_t)1 << 22;
  int64_t ver = 0;  // state version parity of the next launch (every slot's)
  std::vector<Slot *> joining;
  hipStream_t copy_stream = nullptr;  // result copies to the host
  // shared with the setup worker, guarded by mu
  std::mutex mu;
  std::condition_variable cv;        // wakes the worker
  std::condition_variable cv_ready;  // wakes an idle progress call
  std::vector<std::thread> workers;
  bool stop = false;
  int held = 0;                     // slots in setup, set up, joining or active
  std::vector<Slot *> ready;        // set up by the worker, not yet joined
  std::vector<int64_t> failed_ids;  // setups that failed, not yet reported
};

namespace {

size_t estimate_bytes(const Req &r) {
  const int64_t nen = (int64_t)std::pow(r.p + 1, r.dim);
  const int64_t ncell = (int64_t)r.cells.size() / (r.dim == 3 ? 8 : 4);
  const int64_t T = ncell * nen * nen;                     // element node pairs
  const int64_t nnodes_up = ncell * nen;
  const int64_t nld = nen * r.dpn;
  const int64_t nnz_up = T * r.dpn * r.dpn;
  size_t b = 0;
  b += (size_t)T * (8 + 4) * 2 + (size_t)T * 4 * 3;          // pattern sort (keys, values, temp) and plan
  b += (size_t)ncell * nld * nld * 8;                         // element staging
  b += (size_t)nnz_up * (8 + 4 + 4);                          // FP64 + FP32 values, column indices
  b += (size_t)nnodes_up * r.dpn * (8 * 6 + 4 * 8);           // vectors
  return b + (1 << 20);
}

void plane_bounds(const Req &r, double lo[3], double hi[3]) {
  for (int d = 0; d < 3; ++d) {
    lo[d] = 1e300;
    hi[d] = -1e300;
  }
  const size_t nn = r.nodes.size() / 3;
  for (size_t i = 0; i < nn; ++i)
    for (int d = 0; d < 3; ++d) {
      lo[d] = std::min(lo[d], r.nodes[3 * i + d]);
      hi[d] = std::max(hi[d], r.nodes[3 * i + d]);
    }
}

void check_rc(int rc, yafel_hip_ctx_s *c) {
  if (rc != YAFEL_HIP_OK) fail(rc, yafel_hip_last_error(c));
}

// Build the request's device system on setup context c.
Slot *setup_slot(yafel_hip_ctx_s *c, Req *r) {
  auto *sl = new Slot();
  sl->req = r;
  yafel_hip_mesh m = nullptr;
  try {
    check_rc(yafel_hip_mesh_create_cg(c, r->dim, r->p, r->dpn, (int64_t)r->nodes.size() / 3, r->nodes.data(),
                                      (int64_t)r->cells.size() / (r->dim == 3 ? 8 : 4), r->cells.data(), &m),
             c);
    const int fmt = r->dpn == 3 ? YAFEL_HIP_BCSR3 : YAFEL_HIP_CSR;
    check_rc(yafel_hip_mesh_pattern(m, fmt, YAFEL_HIP_FP64, &sl->A), c);
    const int64_t n = m->n_nodes * r->dpn;
    sl->ndofs = n;
    sl->b.alloc(n * 8);
    HIPCHK(hipMemsetAsync(sl->b.p, 0, n * 8, c->stream));
    yafel_hip_physics ph = r->phys;
    ph.coef = r->coef.empty() ? nullptr : r->coef.data();
    ph.coef_on_device = 0;
    int req = YAFEL_HIP_ASSEMBLE_TANGENT;
    if (r->load.empty()) {
      req |= YAFEL_HIP_ASSEMBLE_RESIDUAL;
    } else {
      REQUIRE((int64_t)r->load.size() == n, "server: load length differs from the number of dofs");
      HIPCHK(hipMemcpyAsync(sl->b.p, r->load.data(), n * 8, hipMemcpyHostToDevice, c->stream));
    }
    check_rc(yafel_hip_assemble(m, &ph, req, sl->A, sl->b.as<double>()), c);
    sl->mask.alloc(n);
    HIPCHK(hipMemsetAsync(sl->mask.p, 0, n, c->stream));
    double lo[3], hi[3];
    plane_bounds(*r, lo, hi);
    const double ext = std::max(hi[0] - lo[0], std::max(hi[1] - lo[1], hi[2] - lo[2]));
    const double tol = 1e-9 * (ext > 0 ? ext : 1.0);
    if (r->bc == YAFEL_HIP_BC_BOX) {
      for (int d = 0; d < r->dim; ++d) {
        check_rc(yafel_hip_bc_mask_plane(m, d, lo[d], tol, -1, sl->mask.as<uint8_t>()), c);
        check_rc(yafel_hip_bc_mask_plane(m, d, hi[d], tol, -1, sl->mask.as<uint8_t>()), c);
      }
    } else {
      check_rc(yafel_hip_bc_mask_plane(m, 0, lo[0], tol, -1, sl->mask.as<uint8_t>()), c);
    }
    check_rc(yafel_hip_dirichlet_apply(sl->A, sl->mask.as<uint8_t>(), nullptr, sl->b.as<double>()), c);
    matrix_build_fp32(sl->A);
    matrix_build_diag(sl->A);
    HIPCHK(hipStreamSynchronize(c->stream));
    // the plan and staging are not needed once assembled
    sl->A->mesh = nullptr;
    sl->A->plan = nullptr;
    yafel_hip_mesh_destroy(m);
    m = nullptr;
  } catch (...) {
    if (m) yafel_hip_mesh_destroy(m);
    if (sl->A) yafel_hip_matrix_destroy(sl->A);
    delete sl;
    throw;
  }
  return sl;
}

// Allocate the slot's solve buffers on the solve stream and fill its descriptor.
void join_slot(yafel_hip_server_s *S, Slot *sl) {
  yafel_hip_matrix_s *A = sl->A;
  const int64_t n = A->rows;
  Req *r = sl->req;
  const int g_spmv = sell_grid(A->nsl, GCAP);
  // elementwise passes: as many elements per thread as the single solve's
  // FP32 passes take -- every workgroup of a pass refolds the producer's
  // partials, so fewer, fuller workgroups read far fewer partials
  int64_t ge = ceil_div(n, (int64_t)NT * ve_of<float>());
  if (ge >= 8) ge = (ge + 7) / 8 * 8;
  const int g_elem = (int)std::max<int64_t>(1, std::min<int64_t>(GCAP, ge));
  const size_t nbytes = 6 * (((size_t)n * 4 + 255) & ~(size_t)255) + (((size_t)n * 8 + 255) & ~(size_t)255) +
                        ((2 * sizeof(KState) + 255) & ~(size_t)255) + (size_t)4 * std::max(g_spmv, g_elem) * 8 + 256;
  HIPCHK(hipMallocAsync(&sl->buf, nbytes, S->ctx->stream));
  sl->buf_bytes = nbytes;
  char *p = (char *)sl->buf;
  SlotDesc &d = sl->d;
  d.fmt = A->format;
  d.pad0 = 0;
  d.g_spmv = g_spmv;
  d.g_elem = g_elem;
  d.n = n;
  d.brows = A->brows;
  d.ntot = A->ntot;
  d.sptr = A->sptr.as<int>();
  d.perm = A->perm.p ? A->perm.as<int>() : nullptr;
  d.col = A->col.as<int>();
  d.col16 = A->col16.p ? A->col16.as<short>() : nullptr;
  d.v32 = A->val32.as<float>();
  d.v64 = A->val64.as<double>();
  d.dinv = A->dinv32.as<float>();
  d.r = carve<float>(p, n);
  d.z = carve<float>(p, n);
  d.p = carve<float>(p, n);
  d.p1 = carve<float>(p, n);
  d.q = carve<float>(p, n);
  d.xl = carve<float>(p, n);
  d.x = carve<double>(p, n);
  d.st = carve<KState>(p, 2);
  d.part = carve<double>(p, (size_t)4 * std::max(g_spmv, g_elem));
  d.b = sl->b.as<double>();
  hipStream_t s = S->ctx->stream;
  d.tol = r->tol;
  d.max_iter = r->max_iter > 0 ? r->max_iter : n;
  HIPCHK(hipMemsetAsync(d.p, 0, n * 4, s));
  HIPCHK(hipMemsetAsync(d.xl, 0, n * 4, s));
  HIPCHK(hipMemsetAsync(d.x, 0, n * 8, s));
}

// The slot's solve vectors and state: freed on the solve stream, so the free is
// ordered after every chunk already launched with the slot in its table.
void release_solve_buffers(yafel_hip_server_s *S, Slot *sl) {
  if (sl->buf) (void)hipFreeAsync(sl->buf, S->ctx->stream);
  sl->buf = nullptr;
}

// The slot's matrix, load and mask were allocated on a setup stream, which is
// not ordered after the solve stream.  Called only when no chunk that lists
// the slot can still run on the device (its last chunk was drained on the
// host, or the solve stream was synchronized), so freeing them on the setup
// stream cannot race a read.  Returns the slot's bytes to the HBM budget.
void release_matrix(yafel_hip_server_s *S, Slot *sl) {
  delete sl->A;
  sl->A = nullptr;
  sl->b.release();
  sl->mask.release();
  {
    std::lock_guard<std::mutex> lk(S->mu);
    S->in_use -= std::min(S->in_use, sl->req ? sl->req->est_bytes : 0);
    S->held -= 1;
  }
  S->cv.notify_one();
}

// A setup worker: FIFO admission under the HBM budget and the slot limit;
// builds on its own setup context.
void worker_main(yafel_hip_server_s *S, yafel_hip_ctx_s *c) {
  (void)hipSetDevice(S->ctx->device);
  // the workers' caches of released set-up buffers are not charged to the
  // budget, so together they are held to an eighth of it (4 GiB per worker at most)
  const size_t cache = std::min<size_t>((size_t)4 << 30, S->budget / (8 * std::max<size_t>(1, S->sctx.size())));
  set_alloc_stream(c->stream, cache);
  std::unique_lock<std::mutex> lk(S->mu);
  while (true) {
    S->cv.wait(lk, [S] {
      if (S->stop) return true;
      while (!S->queue.empty() && S->queue.front()->state == YAFEL_HIP_REQ_CANCELLED) S->queue.pop_front();
      return !S->queue.empty() && S->held < S->max_slots && S->in_use + S->queue.front()->est_bytes <= S->budget;
    });
    if (S->stop) break;
    Req *r = S->queue.front();
    S->queue.pop_front();
    r->state = YAFEL_HIP_REQ_RUNNING;
    r->t_admit = std::chrono::steady_clock::now();
    S->in_use += r->est_bytes;
    S->peak = std::max(S->peak, S->in_use);
    S->held += 1;
    lk.unlock();
    Slot *sl = nullptr;
    std::string err;
    const auto t0 = std::chrono::steady_clock::now();
    try {
      sl = setup_slot(c, r);
    } catch (const Error &e) {
      err = e.msg;
    } catch (const std::exception &e) {
      err = e.what();
    }
    const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    lk.lock();
    S->stats.setup_ms += ms;
    if (sl) {
      r->ndofs = sl->ndofs;
      r->t_ready = std::chrono::steady_clock::now();
      S->ready.push_back(sl);
      S->stats.admitted++;
      S->cv_ready.notify_one();
    } else {
      r->state = YAFEL_HIP_REQ_FAILED;
      r->error = err;
      S->in_use -= std::min(S->in_use, r->est_bytes);
      S->held -= 1;
      S->stats.failed++;
      S->failed_ids.push_back(r->id);
      S->cv_ready.notify_one();
    }
  }
  set_alloc_stream(nullptr);
}

// upload a descriptor table for `slots` (offsets recomputed) into ring entry
// e (returned), or into `dst` when given (the graph epoch's table)
int upload_table(yafel_hip_server_s *S, const std::vector<Slot *> &slots, int *tot_spmv, int *tot_elem,
                 DBuf *dst = nullptr) {
  std::vector<SlotDesc> h(slots.size());
  int os = 0, oe = 0;
  // the cache budget shared by the table's FP32 streams in proportion to their size
  double total = 0;
  for (Slot *sl : slots) total += (double)sl->A->ntot * (sl->A->format == YAFEL_HIP_BCSR3 ? 38 : 6);
  for (size_t i = 0; i < slots.size(); ++i) {
    h[i] = slots[i]->d;
    const yafel_hip_matrix_s *A = slots[i]->A;
    const double mine = (double)A->ntot * (A->format == YAFEL_HIP_BCSR3 ? 38 : 6);
    h[i].keep32 = keep_slots(A, 4, total > 0 ? (int64_t)(kCacheBytesDefault * (mine / total)) : 0);
    h[i].off_spmv = os;
    h[i].off_elem = oe;
    os += h[i].g_spmv;
    oe += h[i].g_elem;
  }
  *tot_spmv = os;
  *tot_elem = oe;
  const int e = S->ring;
  S->ring = (S->ring + 1) % yafel_hip_server_s::RING;
  const size_t bytes = std::max<size_t>(h.size(), 1) * sizeof(SlotDesc);
  REQUIRE(bytes <= yafel_hip_server_s::STAGE, "server: too many slots for the descriptor table");
  if (S->desc[e].bytes < yafel_hip_server_s::STAGE) S->desc[e].alloc(yafel_hip_server_s::STAGE);
  if (S->status[e].bytes < yafel_hip_server_s::STAGE) S->status[e].alloc(yafel_hip_server_s::STAGE);
  // stream-ordered copy from this entry's own pinned staging area
  char *stage = S->pin + (yafel_hip_server_s::RING + e) * yafel_hip_server_s::STAGE;
  std::memcpy(stage, h.data(), bytes);
  HIPCHK(hipMemcpyAsync(dst ? dst->p : S->desc[e].p, stage, bytes, hipMemcpyHostToDevice, S->ctx->stream));
  return e;
}

// One batched launch of phase PH over the table; the residual phases are
// followed by their per-slot folds, the versioned passes advance the state
// version of every slot in the table.
template <int PH>
void launch_phase(yafel_hip_server_s *S, const SlotDesc *desc, int ns, int blocks, const TableHdr *hdr = nullptr) {
  if (ns == 0 || blocks == 0) return;
  const int par = (int)(S->ver & 1);
  kb_phase<PH><<<blocks, NT, 0, S->ctx->stream>>>(desc, ns, par, hdr);
  HIPCHK(hipGetLastError());
  if (PH == PH_INIT || PH == PH_REF) {
    kb_finish<PH><<<ns, NTF, 0, S->ctx->stream>>>(desc, ns, par, hdr);
    HIPCHK(hipGetLastError());
  }
  if (PH == PH_U || PH == PH_P || PH == PH_P0 || PH == PH_QP || PH == PH_PN) ++S->ver;
  S->stats.batched_launches++;
}

// One chunk for the slots of table `desc`: p = z + beta p after a slot's
// initial or refinement residual, K iterations, the refinement, and every
// slot's status into `status`.  Three passes an iteration (Q, U, P) or, with
// two_pass, two after the first (Q, U, then QP, U: the QP pass applies the
// previous iteration's P in its gather) and a closing pass (PN) that applies
// the last one's P and moves p back into d.p.  Both forms leave the same
// state, advance the state version by an odd count (graph replay keeps
// graphs per starting parity) and give bit-identical iterates, so the form can change
// from chunk to chunk.  With hdr (graph capture) ns, ts and te are the caps
// of the grids and the kernels read the table's own counts from hdr.
void chunk_launches(yafel_hip_server_s *S, const SlotDesc *desc, int ns, int ts, int te, SlotStatus *status,
                    bool two_pass, const TableHdr *hdr = nullptr) {
  if (two_pass) {
    launch_phase<PH_P0>(S, desc, ns, te, hdr);
    launch_phase<PH_Q>(S, desc, ns, ts, hdr);
    launch_phase<PH_U>(S, desc, ns, te, hdr);
    for (int k = 1; k < S->K; ++k) {
      launch_phase<PH_QP>(S, desc, ns, ts, hdr);
      launch_phase<PH_U>(S, desc, ns, te, hdr);
    }
    launch_phase<PH_PN>(S, desc, ns, te, hdr);
  } else {
    launch_phase<PH_P0>(S, desc, ns, te, hdr);
    for (int k = 0; k < S->K; ++k) {
      launch_phase<PH_Q>(S, desc, ns, ts, hdr);
      launch_phase<PH_U>(S, desc, ns, te, hdr);
      launch_phase<PH_P>(S, desc, ns, te, hdr);
    }
  }
  launch_phase<PH_FOLD>(S, desc, ns, te, hdr);
  launch_phase<PH_REF>(S, desc, ns, ts, hdr);
  kb_status<<<ceil_div(ns, 64), 64, 0, S->ctx->stream>>>(desc, ns, (int)(S->ver & 1), status, hdr);
  HIPCHK(hipGetLastError());
}

// Retire the current graphs: a chunk replayed from them may still run, so
// they are destroyed after the last chunk launched so far has drained.
void drop_graphs(yafel_hip_server_s *S) {
  for (auto &row : S->gexec)
    for (auto &g : row)
      if (g) {
        S->gold.emplace_back(S->chunk_seq, g);
        g = nullptr;
      }
  S->gcap_s = S->gcap_e = 0;
  S->gepoch = -1;
}

void reap_graphs(yafel_hip_server_s *S, bool all) {
  std::vector<std::pair<int64_t, hipGraphExec_t>> keep;
  for (auto &o : S->gold) {
    if (all || o.first <= S->drained_seq) (void)hipGraphExecDestroy(o.second);
    else keep.push_back(o);
  }
  S->gold.swap(keep);
}

// a grid cap: the next power of two, at least 256 blocks
int grid_cap(int blocks) {
  int c = 256;
  while (c < blocks) c <<= 1;
  return c;
}

// The active slot set's table and totals into gdesc / ghdr (stream-ordered,
// after every chunk launched before), and new caps when the totals outgrow
// the captured grids or use less than an eighth of them.
void graph_table(yafel_hip_server_s *S) {
  using SS = yafel_hip_server_s;
  if (S->gdesc.bytes < SS::STAGE) S->gdesc.alloc(SS::STAGE);
  if (S->ghdr.bytes < 256) S->ghdr.alloc(256);
  int ts, te;
  const int e = upload_table(S, S->active, &ts, &te, &S->gdesc);
  // the header from the end of the same staging entry (the table fills at
  // most max_slots descriptors of it)
  TableHdr *h = reinterpret_cast<TableHdr *>(S->pin + (SS::RING + e + 1) * SS::STAGE - sizeof(TableHdr));
  *h = TableHdr{(int)S->active.size(), ts, te, 0};
  HIPCHK(hipMemcpyAsync(S->ghdr.p, h, sizeof(TableHdr), hipMemcpyHostToDevice, S->ctx->stream));
  const bool fits = ts <= S->gcap_s && te <= S->gcap_e;
  const bool slack = S->gcap_s > 8 * grid_cap(ts) || S->gcap_e > 8 * grid_cap(te);
  if (!fits || slack) {
    drop_graphs(S);
    S->gcap_s = grid_cap(ts);
    S->gcap_e = grid_cap(te);
  }
  S->gepoch = S->epoch;
}

// Capture the chunk of form `form` (1: two passes) for status slot r,
// starting at version parity r & 1; the launches only record (nothing runs,
// no counter moves).
void capture_graph(yafel_hip_server_s *S, int form, int r) {
  using SS = yafel_hip_server_s;
  if (S->gstatus[r].bytes < SS::STAGE) S->gstatus[r].alloc(SS::STAGE);
  const int64_t ver = S->ver, launches = S->stats.batched_launches;
  hipStream_t st = S->ctx->stream;
  S->ver = r & 1;
  hipGraph_t g = nullptr;
  HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed));
  try {
    chunk_launches(S, S->gdesc.as<SlotDesc>(), S->max_slots, S->gcap_s, S->gcap_e, S->gstatus[r].as<SlotStatus>(),
                   form == 1, S->ghdr.as<TableHdr>());
    HIPCHK(hipMemcpyAsync(S->gpin + r * SS::STAGE, S->gstatus[r].p, S->max_slots * sizeof(SlotStatus),
                          hipMemcpyDeviceToHost, st));
  } catch (...) {
    (void)hipStreamEndCapture(st, &g);
    if (g) (void)hipGraphDestroy(g);
    S->ver = ver;
    S->stats.batched_launches = launches;
    throw;
  }
  S->ver = ver;
  S->stats.batched_launches = launches;
  HIPCHK(hipStreamEndCapture(st, &g));
  const hipError_t rc = hipGraphInstantiate(&S->gexec[form][r], g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  HIPCHK(rc);
  S->captures++;
}

void drain_one(yafel_hip_server_s *S, std::vector<int64_t> &done_ids) {
  auto f = S->inflight.front();
  S->inflight.pop_front();
  HIPCHK(hipEventSynchronize(f.ev));
  S->drained_seq = f.seq;
  if (f.gr >= 0) S->gbusy[f.gr] = false;
  const SlotStatus *st = (const SlotStatus *)f.pin;
  for (size_t i = 0; i < f.slots.size(); ++i) {
    Slot *sl = f.slots[i];
    if (sl->retired || !st[i].done) continue;
    Req *r = sl->req;
    r->info.status = st[i].status;
    r->info.iterations = st[i].iter;
    r->info.refinements = st[i].refinements;
    r->info.criterion = st[i].crit;
    r->info.rel_residual = st[i].crit;  // the FP64 true residual at the last refinement
    r->x.resize(sl->d.n);
    // x is final: its chunk completed
    HIPCHK(hipMemcpyAsync(r->x.data(), sl->d.x, sl->d.n * 8, hipMemcpyDeviceToHost, S->copy_stream));
    HIPCHK(hipStreamSynchronize(S->copy_stream));
    {
      std::lock_guard<std::mutex> lk(S->mu);
      r->state = YAFEL_HIP_REQ_DONE;
      r->t_done = std::chrono::steady_clock::now();
      S->stats.completed++;
    }
    done_ids.push_back(r->id);
    sl->retired = true;
  }
}

}  // namespace

extern "C" {

int yafel_hip_server_create(yafel_hip_ctx ctx, size_t hbm_budget, int max_slots, yafel_hip_server *out) {
  return yafel_hip_server_create_ex(ctx, hbm_budget, max_slots, 0, out);
}

int yafel_hip_server_create_ex(yafel_hip_ctx ctx, size_t hbm_budget, int max_slots, int setup_threads,
                               yafel_hip_server *out) {
  YH_TRY_BEGIN
  REQUIRE(ctx && out && max_slots > 0 && hbm_budget > 0, "server_create: bad argument");
  REQUIRE(setup_threads >= 0 && setup_threads <= 16, "server_create: setup_threads must be 0..16");
  // a table of max_slots descriptors and, at the end of its staging area, the
  // graph replay header
  REQUIRE((size_t)max_slots * sizeof(SlotDesc) + sizeof(TableHdr) <= yafel_hip_server_s::STAGE,
          "server_create: max_slots too large");
  HIPCHK(hipSetDevice(ctx->device));
  if (setup_threads == 0) setup_threads = 3;
  auto *S = new yafel_hip_server_s();
  ctx_retain(ctx);
  S->ctx = ctx;
  for (int i = 0; i < setup_threads; ++i) {
    yafel_hip_ctx_s *c = nullptr;
    int rc = yafel_hip_ctx_create(ctx->device, &c);
    if (rc != YAFEL_HIP_OK) {
      for (auto *o : S->sctx) yafel_hip_ctx_destroy(o);
      delete S;
      ctx_release(ctx);
      fail(rc, yafel_hip_last_error(nullptr));
    }
    S->sctx.push_back(c);
  }
  S->budget = hbm_budget;
  S->max_slots = max_slots;
  S->stats.hbm_budget = hbm_budget;
  for (auto &e : S->evs) HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  HIPCHK(hipHostMalloc((void **)&S->pin, 2 * yafel_hip_server_s::RING * yafel_hip_server_s::STAGE,
                       hipHostMallocDefault));
  HIPCHK(hipHostMalloc((void **)&S->gpin, yafel_hip_server_s::GR * yafel_hip_server_s::STAGE, hipHostMallocDefault));
  for (auto &e : S->gev) HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  HIPCHK(hipStreamCreateWithFlags(&S->copy_stream, hipStreamNonBlocking));
  for (auto *c : S->sctx) S->workers.emplace_back(worker_main, S, c);
  *out = S;
  return YAFEL_HIP_OK;
  YH_TRY_END(ctx)
}

int yafel_hip_server_destroy(yafel_hip_server S) {
  if (!S) return YAFEL_HIP_OK;
  {
    std::lock_guard<std::mutex> lk(S->mu);
    S->stop = true;
  }
  S->cv.notify_all();
  for (auto &w : S->workers)
    if (w.joinable()) w.join();
  // nothing that lists a slot runs after this
  (void)hipStreamSynchronize(S->ctx->stream);
  for (Slot *sl : S->ready) S->joining.push_back(sl);
  S->ready.clear();
  for (auto *v : {&S->active, &S->joining, &S->grave})
    for (Slot *sl : *v) {
      release_solve_buffers(S, sl);
      release_matrix(S, sl);
      delete sl;
    }
  (void)hipStreamSynchronize(S->ctx->stream);
  drop_graphs(S);
  reap_graphs(S, true);
  for (auto &e : S->evs) (void)hipEventDestroy(e);
  for (auto &e : S->gev)
    if (e) (void)hipEventDestroy(e);
  if (S->pin) (void)hipHostFree(S->pin);
  if (S->gpin) (void)hipHostFree(S->gpin);
  for (auto *c : S->sctx) (void)hipStreamSynchronize(c->stream);
  if (S->copy_stream) (void)hipStreamDestroy(S->copy_stream);
  for (auto *c : S->sctx) yafel_hip_ctx_destroy(c);
  yafel_hip_ctx_s *ctx = S->ctx;
  delete S;
  ctx_release(ctx);
  return YAFEL_HIP_OK;
}

int yafel_hip_server_submit(yafel_hip_server S, const yafel_hip_request *q, int64_t *id) {
  YH_TRY_BEGIN
  REQUIRE(S && q && id, "server_submit: null argument");
  REQUIRE(q->dim == 2 || q->dim == 3, "server_submit: dim must be 2 or 3");
  REQUIRE(q->n_corner > 0 && q->ncell > 0 && q->corner_nodes && q->cells, "server_submit: empty mesh");
  REQUIRE(q->tol > 0, "server_submit: tol must be positive");
  auto r = std::make_unique<Req>();
  r->id = S->next_id++;
  r->dim = q->dim;
  r->p = q->p;
  r->dpn = q->dof_per_node;
  r->bc = q->bc;
  r->tol = q->tol;
  r->max_iter = q->max_iter;
  r->nodes.assign(q->corner_nodes, q->corner_nodes + q->n_corner * 3);
  r->cells.assign(q->cells, q->cells + q->ncell * (q->dim == 3 ? 8 : 4));
  r->phys = q->phys;
  if (q->phys.coef) {
    REQUIRE(!q->phys.coef_on_device, "server_submit: coefficients must be host data");
    r->coef.assign(q->phys.coef, q->phys.coef + q->ncell);
  }
  r->phys.coef = nullptr;
  r->est_bytes = estimate_bytes(*r);
  r->t_submit = std::chrono::steady_clock::now();
  if (r->est_bytes > S->budget) {
    std::lock_guard<std::mutex> lk(S->mu);
    S->stats.failed++;
    fail(YAFEL_HIP_ENOMEM, "server_submit: the request alone needs " + std::to_string(r->est_bytes) +
                                " bytes, more than the HBM budget");
  }
  if (q->load) {
    REQUIRE(q->load_len > 0, "server_submit: load given without its length");
    r->load.assign(q->load, q->load + q->load_len);
  }
  *id = r->id;
  {
    std::lock_guard<std::mutex> lk(S->mu);
    S->queue.push_back(r.get());
    S->reqs[r->id] = std::move(r);
    S->stats.submitted++;
    S->stats.queued = (int64_t)S->queue.size();
  }
  S->cv.notify_one();
  return YAFEL_HIP_OK;
  YH_TRY_END(S ? S->ctx : nullptr)
}

int yafel_hip_server_progress(yafel_hip_server S, int64_t *done_ids, int max_ids, int *n_done) {
  YH_TRY_BEGIN
  REQUIRE(S && n_done, "server_progress: null argument");
  HIPCHK(hipSetDevice(S->ctx->device));
  std::vector<int64_t> done;
  // 1. pick up what the setup worker finished; with nothing on the device,
  //    wait briefly for it rather than return at once
  {
    std::unique_lock<std::mutex> lk(S->mu);
    if (S->active.empty() && S->joining.empty() && S->inflight.empty() && S->ready.empty() &&
        S->failed_ids.empty() && (S->held > 0 || !S->queue.empty()))
      S->cv_ready.wait_for(lk, std::chrono::milliseconds(2),
                           [S] { return !S->ready.empty() || !S->failed_ids.empty(); });
    for (Slot *sl : S->ready) S->joining.push_back(sl);
    S->ready.clear();
    for (int64_t id : S->failed_ids) done.push_back(id);
    S->failed_ids.clear();
  }
  // 2. drain the oldest chunk when the pipeline is full: up to kDepth chunks
  //    in flight, so that the host builds and submits the next chunk while
  //    the device still runs the ones before it
  while (S->inflight.size() >= (size_t)yafel_hip_server_s::kDepth) drain_one(S, done);
  // 3. retire finished slots (their x is on the host): the solve buffers are
  //    freed on the solve stream now, the matrix once its last chunk drains
  std::vector<Slot *> keep;
  for (Slot *sl : S->active) {
    if (sl->retired) {
      release_solve_buffers(S, sl);
      S->grave.push_back(sl);
    } else {
      keep.push_back(sl);
    }
  }
  if (keep.size() != S->active.size()) S->epoch++;
  S->active.swap(keep);
  // 4. joining slots: allocate, initialise (FP64 residual) and add to the batch
  if (!S->joining.empty()) {
    S->epoch++;
    for (Slot *sl : S->joining) join_slot(S, sl);
    int ts, te;
    const int e = upload_table(S, S->joining, &ts, &te);
    kb_init_state<<<(int)S->joining.size(), 64, 0, S->ctx->stream>>>(S->desc[e].as<SlotDesc>(),
                                                                     (int)S->joining.size(), (int)(S->ver & 1));
    HIPCHK(hipGetLastError());
    launch_phase<PH_INIT>(S, S->desc[e].as<SlotDesc>(), (int)S->joining.size(), ts);
    const auto now = std::chrono::steady_clock::now();
    for (Slot *sl : S->joining) {
      sl->req->t_join = now;
      S->active.push_back(sl);
    }
    S->joining.clear();
  }
  // 5. one chunk for every active slot
  if (!S->active.empty()) {
    const int ns = (int)S->active.size();
    yafel_hip_server_s::Inflight f;
    bool two = S->passes == 2;
    if (S->passes == 0) {
      int64_t scalars = 0;
      for (Slot *sl : S->active) scalars += sl->A->ntot * (sl->A->format == YAFEL_HIP_BCSR3 ? 9 : 1);
      two = scalars <= S->two_pass_limit;
    }
    (two ? S->stats_two : S->stats_three)++;
    if (S->graphs) {
      if (S->gepoch != S->epoch) graph_table(S);
      // a free status slot of the chunk's starting parity
      const int p = (int)(S->ver & 1);
      while (S->gbusy[p] && S->gbusy[p + 2]) drain_one(S, done);
      const int r = S->gbusy[p] ? p + 2 : p;
      const int form = two ? 1 : 0;
      if (!S->gexec[form][r]) capture_graph(S, form, r);
      HIPCHK(hipGraphLaunch(S->gexec[form][r], S->ctx->stream));
      S->gbusy[r] = true;
      S->ver += 2 * S->K + 1;
      S->stats.batched_launches += two ? 2 * S->K + 4 : 3 * S->K + 3;
      S->replays++;
      f.ev = S->gev[r];
      f.ring = -1;
      f.gr = r;
      f.pin = S->gpin + r * yafel_hip_server_s::STAGE;
    } else {
      int ts, te;
      const int e = upload_table(S, S->active, &ts, &te);
      chunk_launches(S, S->desc[e].as<SlotDesc>(), ns, ts, te, S->status[e].as<SlotStatus>(), two);
      f.ev = S->evs[e];
      f.ring = e;
      f.gr = -1;
      f.pin = S->pin + e * yafel_hip_server_s::STAGE;
      HIPCHK(hipMemcpyAsync(f.pin, S->status[e].p, ns * sizeof(SlotStatus), hipMemcpyDeviceToHost, S->ctx->stream));
    }
    f.seq = ++S->chunk_seq;
    for (Slot *sl : S->active) sl->last_chunk = f.seq;
    f.slots = S->active;
    HIPCHK(hipEventRecord(f.ev, S->ctx->stream));
    S->inflight.push_back(f);
    S->stats.chunks++;
  } else {
    while (!S->inflight.empty()) drain_one(S, done);
  }
  reap_graphs(S, false);
  // retired slots whose last chunk has drained: nothing on the device reads
  // their matrix any more
  {
    std::vector<Slot *> wait;
    for (Slot *sl : S->grave) {
      if (sl->last_chunk <= S->drained_seq) {
        release_matrix(S, sl);
        delete sl;
      } else {
        wait.push_back(sl);
      }
    }
    S->grave.swap(wait);
  }
  {
    std::lock_guard<std::mutex> lk(S->mu);
    S->stats.queued = (int64_t)S->queue.size();
    S->stats.active = (int64_t)S->active.size();
    S->stats.hbm_in_use = S->in_use;
    S->stats.hbm_peak = S->peak;
  }
  *n_done = 0;
  for (int64_t id : done) {
    if (done_ids && *n_done < max_ids) done_ids[*n_done] = id;
    ++*n_done;
  }
  return YAFEL_HIP_OK;
  YH_TRY_END(S ? S->ctx : nullptr)
}

int yafel_hip_server_query(yafel_hip_server S, int64_t id, int *state, int64_t *ndofs) {
  YH_TRY_BEGIN
  REQUIRE(S, "server_query: null server");
  std::lock_guard<std::mutex> lk(S->mu);
  auto it = S->reqs.find(id);
  REQUIRE(it != S->reqs.end(), "server_query: unknown request id");
  if (state) *state = it->second->state;
  if (ndofs) *ndofs = it->second->ndofs;
  return YAFEL_HIP_OK;
  YH_TRY_END(S ? S->ctx : nullptr)
}

int yafel_hip_server_result(yafel_hip_server S, int64_t id, double *x, int64_t n, yafel_hip_solve_info *info) {
  YH_TRY_BEGIN
  REQUIRE(S, "server_result: null server");
  std::lock_guard<std::mutex> lk(S->mu);
  auto it = S->reqs.find(id);
  REQUIRE(it != S->reqs.end(), "server_result: unknown request id");
  Req *r = it->second.get();
  // a failed or cancelled request is reported once and then forgotten, as a
  // finished one is once its result has been read
  if (r->state == YAFEL_HIP_REQ_FAILED) {
    const std::string msg = "server_result: request failed: " + r->error;
    S->reqs.erase(it);
    fail(YAFEL_HIP_ESTATE, msg);
  }
  if (r->state == YAFEL_HIP_REQ_CANCELLED) {
    S->reqs.erase(it);
    fail(YAFEL_HIP_ESTATE, "server_result: request was cancelled");
  }
  REQUIRE(r->state == YAFEL_HIP_REQ_DONE, "server_result: request not finished");
  REQUIRE(n >= (int64_t)r->x.size(), "server_result: output too short");
  if (x) std::memcpy(x, r->x.data(), r->x.size() * 8);
  if (info) *info = r->info;
  S->reqs.erase(it);
  return YAFEL_HIP_OK;
  YH_TRY_END(S ? S->ctx : nullptr)
}

int yafel_hip_server_times(yafel_hip_server S, int64_t id, double *ms) {
  YH_TRY_BEGIN
  REQUIRE(S && ms, "server_times: null argument");
  std::lock_guard<std::mutex> lk(S->mu);
  auto it = S->reqs.find(id);
  REQUIRE(it != S->reqs.end(), "server_times: unknown request id");
  const Req *r = it->second.get();
  REQUIRE(r->state == YAFEL_HIP_REQ_DONE, "server_times: request not finished");
  auto d = [](std::chrono::steady_clock::time_point a, std::chrono::steady_clock::time_point b) {
    return std::chrono::duration<double, std::milli>(b - a).count();
  };
  ms[0] = d(r->t_submit, r->t_admit);
  ms[1] = d(r->t_admit, r->t_ready);
  ms[2] = d(r->t_ready, r->t_join);
  ms[3] = d(r->t_join, r->t_done);
  return YAFEL_HIP_OK;
  YH_TRY_END(S ? S->ctx : nullptr)
}

int yafel_hip_server_cancel(yafel_hip_server S, int64_t id) {
  YH_TRY_BEGIN
  REQUIRE(S, "server_cancel: null server");
  std::lock_guard<std::mutex> lk(S->mu);
  auto it = S->reqs.find(id);
  REQUIRE(it != S->reqs.end(), "server_cancel: unknown request id");
  Req *r = it->second.get();
  if (r->state == YAFEL_HIP_REQ_QUEUED) {
    r->state = YAFEL_HIP_REQ_CANCELLED;
    S->stats.cancelled++;
    // out of the queue, and its mesh and load freed; the small record stays
    // for server_query until server_result reports the cancellation
    S->queue.erase(std::remove(S->queue.begin(), S->queue.end(), r), S->queue.end());
    S->stats.queued = (int64_t)S->queue.size();
    std::vector<double>().swap(r->nodes);
    std::vector<double>().swap(r->coef);
    std::vector<double>().swap(r->load);
    std::vector<int32_t>().swap(r->cells);
  } else {
    fail(YAFEL_HIP_ESTATE, "server_cancel: only queued requests can be cancelled");
  }
  return YAFEL_HIP_OK;
  YH_TRY_END(S ? S->ctx : nullptr)
}

int yafel_hip_server_stats_get(yafel_hip_server S, yafel_hip_server_stats *st) {
  YH_TRY_BEGIN
  REQUIRE(S && st, "server_stats: null argument");
  std::lock_guard<std::mutex> lk(S->mu);
  *st = S->stats;
  st->graph_captures = S->captures;
  st->graph_replays = S->replays;
  st->two_pass_chunks = S->stats_two;
  st->three_pass_chunks = S->stats_three;
  return YAFEL_HIP_OK;
  YH_TRY_END(S ? S->ctx : nullptr)
}

int yafel_hip_server_set_cg_passes(yafel_hip_server S, int passes, int64_t two_pass_limit) {
  YH_TRY_BEGIN
  REQUIRE(S, "server_set_cg_passes: null server");
  REQUIRE(passes == 0 || passes == 2 || passes == 3, "server_set_cg_passes: passes must be 0, 2 or 3");
  S->passes = passes;
  S->two_pass_limit = two_pass_limit > 0 ? two_pass_limit : yafel_hip_server_s::kTwoPassLimit;
  return YAFEL_HIP_OK;
  YH_TRY_END(S ? S->ctx : nullptr)
}

int yafel_hip_server_set_graphs(yafel_hip_server S, int on) {
  YH_TRY_BEGIN
  REQUIRE(S, "server_set_graphs: null server");
  HIPCHK(hipSetDevice(S->ctx->device));
  if (!on) drop_graphs(S);
  S->graphs = on != 0;
  return YAFEL_HIP_OK;
  YH_TRY_END(S ? S->ctx : nullptr)
}

}  // extern "C"
