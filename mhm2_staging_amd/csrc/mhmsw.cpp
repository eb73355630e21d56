// Host runtime of the mhmsw C ABI (include/mhmsw.h).
//
// One context = one device, one HIP stream, device buffers that grow with the
// largest batch seen (as adept_sw::GPUDriver keeps its buffers for the life of
// the driver, src/adept-sw/driver.cpp:198-247, but sized by actual bytes
// instead of 2*rlen_limit*20000).  A batch is packed on the host to 2-bit codes
// plus an N bitmap, copied once, and then aligned entirely on the device:
// unpack -> forward pass -> 32-bit rerun of flagged pairs -> reverse pass ->
// finalize, with no host round trip between the passes (the reference needs
// one to size its reverse launch, driver.cpp:395).
#include <chrono>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <functional>
#include <atomic>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "../../include/mhmsw.h"
#include "sw_common.h"
#include "sw_launch.h"
#include "cigar_core.h"
#include "ctx_info.h"
#include "host_pinned.h"
#include "pack2bit.h"
#include "thread_pool.h"

using namespace mhmsw;

namespace {

template <class T>
struct DevBuf {
  T* ptr = nullptr;
  size_t cap = 0;  // elements
  hipError_t ensure(size_t n) {
    if (n <= cap) return hipSuccess;
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    cap = 0;
    size_t want = std::max<size_t>(n + n / 4, std::max<size_t>(1, 256 / sizeof(T)));
    hipError_t e = hipMalloc((void**)&ptr, want * sizeof(T));
    if (e == hipSuccess) cap = want;
    return e;
  }
  void release() {
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    cap = 0;
  }
};


constexpr int kEvents = 7;  // job start, 5 phase ends, CIGAR pass end

}  // namespace

struct mhmsw_contigs {
  int device = 0;
  int64_t n = 0;
  std::vector<int32_t> len;
  uint32_t* d2 = nullptr;     // 2-bit codes
  uint32_t* dN = nullptr;     // N bitmap
  int64_t* dbase = nullptr;   // first base of each contig (multiple of 32)
  int64_t bytes = 0;
};

struct mhmsw_ctx {
  int device = 0;
  int cus = 0;
  hipStream_t stream = nullptr;
  std::string err;
  mhmsw_scoring scoring{1, 3, 5, 2, 2};
  ScoreParams sp{};
  int mode = MHMSW_MODE_FAST;
  int read_is_ref = 1;
  int score2 = 0;               // compute SSW's second-best score
  bool last_job_score2 = false;
  int cigar = 0;                // compute SSW's CIGAR and mismatch count
  bool last_job_cigar = false;
  mhmsw_tuning tuning{8, 257, 5};
  // 256-thread blocks of each persistent kernel that fit one CU at once: every
  // packed kernel shape a class can run as (0: unknown), and the band classes
  struct ShapeOcc {
    int lanes = 0, regs = 0, occ = 0;
  };
  ShapeOcc occ_shapes[5 * kMaxP] = {};
  int occ_of(FastShape s) const {
    for (const ShapeOcc& o : occ_shapes)
      if (o.lanes == s.lanes && o.regs == s.regs) return o.occ;
    return 0;
  }
  int occ_band[7] = {};  // all, P = 1, 2, 4, 8, 16 (four-lane pairs), the diagonal class
  std::vector<int32_t> lens;  // (read, contig window) length of each staged pair, for re-routing

  // staged batch
  int64_t npairs = 0;
  int64_t total = 0;
  int64_t cells = 0;
  int max_plain_cols = 0;
  int max_read_len = 0, max_ctg_len = 0;
  bool fast_class_used[kNumClasses] = {};
  bool staged = false;
  bool in_flight = false;
  bool have_results = false;

  DevBuf<uint32_t> d_packed2, d_nmask, d_codes;
  DevBuf<int64_t> d_base;
  DevBuf<int32_t> d_len, d_bucket, d_rank, d_errors;
  DevBuf<unsigned long long> d_counters;
  DevBuf<int32_t> d_win_ctg, d_win_off;
  const mhmsw_contigs* res = nullptr;  // resident contigs of the staged batch (windowed batches)
  int64_t n_fast = 0, n_plain = 0;
  DevBuf<PairEnd> d_fwd, d_rev;
  DevBuf<int32_t> d_s2;
  DevBuf<SwTask> d_tasks;
  DevBuf<PassSched> d_sched;
  DevBuf<AlnResult> d_results;
  // CIGAR pass (cigar_kernels.hip)
  DevBuf<int32_t> d_cig_lists, d_cig_count, d_cig_band, d_cig_rows, d_cig_nops, d_cig_mism;
  DevBuf<int64_t> d_cig_off;
  DevBuf<uint32_t> d_cig_dirs, d_cig_gen, d_cig_ops;
  DevBuf<unsigned long long> d_cig_used;
  int64_t cig_ops_cap = 0, cig_gen_dwords = 0, cig_dirs_dwords = 0;
  unsigned long long h_cig_used[8] = {};
  // hipGraph of the whole job, replayed while the staged geometry is unchanged
  bool use_graphs = true;
  bool last_job_graph = false;
  // per-phase and per-class events of eager jobs (mhmsw_set_phase_timing);
  // last_job_phases: the last job recorded them
  bool phase_timing = true;
  bool last_job_phases = false;
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  std::vector<unsigned char> graph_key;
  std::vector<unsigned char> eager_key;  // geometry of the last eager job (captured if the next one repeats it)
  int64_t graph_captures = 0, graph_replays = 0;
  hipEvent_t ev[kEvents] = {};
  hipEvent_t cev[kNumClasses][2] = {};
  hipEvent_t wev[2] = {};  // around the window extraction of a windowed batch (eager jobs)
  bool last_job_windows = false;
  bool class_launched[kNumClasses] = {};
  int64_t class_pairs[kNumClasses] = {};
  int64_t class_cells[kNumClasses] = {};
  int32_t class_max_rows[kNumClasses] = {};  // longest row side of a class's pairs (route_rows)
  mhmsw_stats stats{};
  int32_t h_errors[4] = {};
  // An upload goes through pinned staging with asynchronous copies on the
  // context stream; the next upload first waits for that stream, so the
  // staging (and the device buffers the copies fill) are reused only after.
  HostPinned h_stage;
  // Words the job copies back into pinned memory at its end, before ev_done:
  // completion needs no further copy (a null-stream copy can wait behind
  // another context's work on a shared hardware queue).
  struct JobTail {
    int32_t errors[4];
    unsigned long long counters[2];
    unsigned long long cig_used[8];
  };
  JobTail* tail = nullptr;
  hipEvent_t ev_done = nullptr;
  // tail_results: the job also copies its results into pinned memory before
  // ev_done (serving slots), so taking them needs no device copy
  bool tail_results = false, last_job_tail_results = false;
  bool job_host_results = false;  // this job only (mhmsw_submit*): as tail_results
  // host times of the staged batch and of the job's launch (reported with its stats)
  double host_pack_ms = 0, host_upload_ms = 0, host_launch_ms = 0;
  HostPinned h_results;
};

static void reroute(mhmsw_ctx* c);

bool mhmsw::ctx_info(const mhmsw_ctx* c, CtxInfo* out) {
  if (!c || !out) return false;
  out->device = c->device;
  out->cus = c->cus;
  out->stream = c->stream;
  out->sc = c->sp;
  out->mode = c->mode;
  out->tuning = c->tuning;
  return true;
}

int mhmsw::ctx_fail(mhmsw_ctx* c, int code, const char* msg) {
  if (c) c->err = msg;
  return code;
}

static int fail(mhmsw_ctx* c, int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  if (c) c->err = buf;
  return code;
}

// Device-to-host copy on the context's own stream (never the null stream,
// which can share a hardware queue with another context's running job).
static hipError_t d2h_sync(mhmsw_ctx* c, void* dst, const void* src, size_t bytes) {
  hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  return e;
}

#define HIP_OR_FAIL(c, expr)                                                                        \
  do {                                                                                              \
    hipError_t _e = (expr);                                                                         \
    if (_e != hipSuccess) return fail((c), MHMSW_E_HIP, "%s: %s (%s:%d)", #expr, hipGetErrorString(_e), \
                                      __FILE__, __LINE__);                                          \
  } while (0)

static void compute_score_params(mhmsw_ctx* c) {
  const mhmsw_scoring& s = c->scoring;
  ScoreParams& p = c->sp;
  p.match = s.match;
  p.mismatch = s.mismatch;
  p.gap_open = s.gap_open;
  p.gap_ext = s.gap_ext;
  p.ambiguity = s.ambiguity;
  p.read_is_ref = c->read_is_ref;
  p.bias = std::max(0, std::max(s.mismatch, s.ambiguity));  // ssw_init, ssw_core.cpp:1076-1083
  // packed kernel: column and row codes select one byte of an 8-byte table
  // holding 16*(s + gap_open): bytes 0..3 = match, mismatch x3 (selector a^b),
  // bytes 4..7 = ambiguity (selector has the N bit).  Needs every entry in
  // [0, 255] and a 16*score range that fits 16 bits for realistic lengths.
  const int m16 = 16 * (s.match + s.gap_open);
  const int x16 = 16 * (s.gap_open - s.mismatch);
  const int n16 = 16 * (s.gap_open - s.ambiguity);
  p.fast_ok = (m16 <= 255 && x16 >= 0 && n16 >= 0) ? 1 : 0;
  if (p.fast_ok) {
    p.perm_lo = (uint32_t)m16 | ((uint32_t)x16 << 8) | ((uint32_t)x16 << 16) | ((uint32_t)x16 << 24);
    p.perm_hi = (uint32_t)n16 * 0x01010101u;
  } else {
    p.perm_lo = p.perm_hi = 0;
  }
}

extern "C" {

int mhmsw_device_count(int* count) {
  if (!count) return MHMSW_E_ARG;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
  *count = n;
  return MHMSW_OK;
}

int mhmsw_device_total_mem(int device_id, size_t* bytes) {
  if (!bytes) return MHMSW_E_ARG;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device_id) != hipSuccess) return MHMSW_E_NODEVICE;
  *bytes = prop.totalGlobalMem;
  return MHMSW_OK;
}

int mhmsw_device_name(int device_id, char* buf, size_t buflen) {
  if (!buf || buflen == 0) return MHMSW_E_ARG;
  if (hipDeviceGetPCIBusId(buf, (int)buflen, device_id) != hipSuccess) return MHMSW_E_NODEVICE;
  return MHMSW_OK;
}

int mhmsw_create(int device_id, mhmsw_ctx** out) {
  if (!out) return MHMSW_E_ARG;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return MHMSW_E_NODEVICE;
  if (device_id < 0 || device_id >= n) return MHMSW_E_ARG;
  mhmsw_ctx* c = new (std::nothrow) mhmsw_ctx();
  if (!c) return MHMSW_E_NOMEM;
  c->device = device_id;
  if (hipSetDevice(device_id) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return MHMSW_E_HIP;
  }
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device_id) != hipSuccess) {
    (void)hipStreamDestroy(c->stream);
    delete c;
    return MHMSW_E_HIP;
  }
  c->cus = prop.multiProcessorCount;
  for (int i = 0; i < kEvents; i++) {
    if (hipEventCreate(&c->ev[i]) != hipSuccess) {
      mhmsw_destroy(c);
      return MHMSW_E_HIP;
    }
  }
  for (int i = 0; i < kNumClasses; i++) {
    if (hipEventCreate(&c->cev[i][0]) != hipSuccess || hipEventCreate(&c->cev[i][1]) != hipSuccess) {
      mhmsw_destroy(c);
      return MHMSW_E_HIP;
    }
  }
  if (hipEventCreate(&c->wev[0]) != hipSuccess || hipEventCreate(&c->wev[1]) != hipSuccess ||
      hipEventCreateWithFlags(&c->ev_done, hipEventDisableTiming) != hipSuccess) {
    mhmsw_destroy(c);
    return MHMSW_E_HIP;
  }
  if (c->d_sched.ensure(kNumPasses) != hipSuccess || c->d_errors.ensure(4) != hipSuccess ||
      c->d_counters.ensure(2) != hipSuccess ||
      hipHostMalloc((void**)&c->tail, sizeof(mhmsw_ctx::JobTail), hipHostMallocDefault) != hipSuccess) {
    c->tail = nullptr;
    mhmsw_destroy(c);
    return MHMSW_E_NOMEM;
  }
  std::memset(c->tail, 0, sizeof(mhmsw_ctx::JobTail));
  // every kernel shape a class can run as: the class's own group size, the
  // narrow forward shape (a class of many pairs) and its trimmed form (a
  // class of short enough rows)
  {
    int n = 0;
    for (int P = 1; P <= kMaxP; P++) {
      const FastShape shapes[5] = {FastShape{16, P}, FastShape{64, P},
                                   fast_shape(16, P, true, INT64_MAX / 64, c->cus, INT32_MAX),
                                   fast_shape(16, P, true, INT64_MAX / 64, c->cus, 1),
                                   fast_shape(64, P, true, INT64_MAX / 64, c->cus, INT32_MAX)};
      for (const FastShape& s : shapes) {
        if (c->occ_of(s) > 0) continue;
        int occ = 0;
        if (fast_kernel_occupancy(s, &occ) != hipSuccess) occ = 0;
        c->occ_shapes[n++] = mhmsw_ctx::ShapeOcc{s.lanes, s.regs, occ};
      }
    }
  }
  for (int i = 0; i < 7; i++)
    if (band_kernel_occupancy(i == 6 ? kBandDiag : i ? 1 << (i - 1) : 0, &c->occ_band[i]) != hipSuccess)
      c->occ_band[i] = 0;
  compute_score_params(c);
  // the no-allocation entry point's scratch on this device (a failure here
  // is reported again, with its message, by that entry point's first call)
  (void)noalloc_setup(device_id);
  *out = c;
  return MHMSW_OK;
}

int mhmsw_destroy(mhmsw_ctx* c) {
  if (!c) return MHMSW_E_ARG;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  c->d_packed2.release();
  c->d_nmask.release();
  c->d_codes.release();
  c->d_base.release();
  c->d_len.release();
  c->d_bucket.release();
  c->d_rank.release();
  c->d_errors.release();
  c->d_counters.release();
  c->d_win_ctg.release();
  c->d_win_off.release();
  c->d_fwd.release();
  c->d_rev.release();
  c->d_tasks.release();
  c->d_sched.release();
  c->d_results.release();
  c->d_cig_lists.release();
  c->d_cig_count.release();
  c->d_cig_band.release();
  c->d_cig_rows.release();
  c->d_cig_nops.release();
  c->d_cig_mism.release();
  c->d_cig_off.release();
  c->d_cig_dirs.release();
  c->d_cig_gen.release();
  c->d_cig_ops.release();
  c->d_cig_used.release();
  for (int i = 0; i < kEvents; i++)
    if (c->ev[i]) (void)hipEventDestroy(c->ev[i]);
  if (c->ev_done) (void)hipEventDestroy(c->ev_done);
  for (int i = 0; i < 2; i++)
    if (c->wev[i]) (void)hipEventDestroy(c->wev[i]);
  if (c->tail) (void)hipHostFree(c->tail);
  c->h_stage.release();
  c->h_results.release();
  for (int i = 0; i < kNumClasses; i++)
    for (int k = 0; k < 2; k++)
      if (c->cev[i][k]) (void)hipEventDestroy(c->cev[i][k]);
  if (c->graph_exec) (void)hipGraphExecDestroy(c->graph_exec);
  if (c->graph) (void)hipGraphDestroy(c->graph);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return MHMSW_OK;
}

const char* mhmsw_last_error(const mhmsw_ctx* c) { return c ? c->err.c_str() : "null context"; }

int mhmsw_set_scoring(mhmsw_ctx* c, const mhmsw_scoring* s) {
  if (!c || !s) return MHMSW_E_ARG;
  if (c->in_flight) return fail(c, MHMSW_E_STATE, "scoring changed while a batch is in flight");
  if (s->match < 1 || s->match > 127 || s->mismatch < 0 || s->mismatch > 127 || s->gap_open < 1 ||
      s->gap_open > 127 || s->gap_ext < 1 || s->gap_ext > 127 || s->ambiguity < 0 || s->ambiguity > 127)
    return fail(c, MHMSW_E_SCORING, "scoring out of range (match 1..127, costs 0..127, gaps 1..127)");
  if (s->gap_ext > s->gap_open) return fail(c, MHMSW_E_SCORING, "gap_ext must not exceed gap_open");
  // SSW's lazy-F loop never lets a gap be followed directly by the opposite
  // gap (ssw_core.cpp:381).  Results equal textbook Gotoh only when such a
  // path can never beat mismatches plus one gap.
  const int worst = std::max(s->mismatch, s->ambiguity);
  if (s->gap_open + s->gap_ext < worst || 2 * s->gap_ext < worst)
    return fail(c, MHMSW_E_SCORING,
                "scoring where a gap followed by the opposite gap beats mismatches (gap_open+gap_ext < %d or "
                "2*gap_ext < %d) has no exact SSW equivalent",
                worst, worst);
  c->scoring = *s;
  compute_score_params(c);
  if (c->staged) reroute(c);
  return MHMSW_OK;
}

int mhmsw_set_orientation(mhmsw_ctx* c, int read_is_ssw_ref) {
  if (!c) return MHMSW_E_ARG;
  if (c->in_flight) return fail(c, MHMSW_E_STATE, "orientation changed while a batch is in flight");
  c->read_is_ref = read_is_ssw_ref ? 1 : 0;
  compute_score_params(c);
  if (c->staged) reroute(c);
  return MHMSW_OK;
}

int mhmsw_set_score2(mhmsw_ctx* c, int enable) {
  if (!c) return MHMSW_E_ARG;
  if (c->in_flight) return fail(c, MHMSW_E_STATE, "score2 changed while a batch is in flight");
  c->score2 = enable ? 1 : 0;
  if (c->staged) reroute(c);
  return MHMSW_OK;
}

int mhmsw_set_mode(mhmsw_ctx* c, int mode) {
  if (!c) return MHMSW_E_ARG;
  if (mode != MHMSW_MODE_FAST && mode != MHMSW_MODE_PLAIN) return fail(c, MHMSW_E_ARG, "unknown mode %d", mode);
  if (c->in_flight) return fail(c, MHMSW_E_STATE, "mode changed while a batch is in flight");
  c->mode = mode;
  if (c->staged) reroute(c);
  return MHMSW_OK;
}

}  // extern "C"

// A batch's contig windows come either as strings (ctgs/ctg_off/ctg_len) or
// as ranges of resident contigs (res/ctg_id/win_off/win_len).
struct StageInput {
  int64_t npairs = 0;
  const char* reads = nullptr;
  const int64_t* read_off = nullptr;
  const int32_t* read_len = nullptr;
  const char* ctgs = nullptr;
  const int64_t* ctg_off = nullptr;
  const int32_t* ctg_len = nullptr;
  // or one pointer per sequence (mhmsw_stage_seqs): read p at read_ptr[p],
  // window p at ctg_ptr[p]; then reads/read_off and ctgs/ctg_off are unused
  const char* const* read_ptr = nullptr;
  const char* const* ctg_ptr = nullptr;
  const mhmsw_contigs* res = nullptr;
  const int32_t* ctg_id = nullptr;
  const int32_t* win_off = nullptr;
  const int32_t* win_len = nullptr;
};

// What routing needs from a context: the forward work class of a pair.
struct PackParams {
  int mode = MHMSW_MODE_FAST;
  int fast_ok = 1;
  int read_is_ref = 1;
  int g64_min_rows = 257;
  int score2 = 0;
};

// rows of a pair's forward rectangle: the side route_class sizes its class by
static int32_t route_rows(const PackParams& pp, int32_t lref, int32_t lqry) {
  return pp.score2 ? lref : std::min(lref, lqry);
}

// host copy of make_task's forward routing (sw_kernels.hip)
static int route_class(const PackParams& pp, int32_t lref, int32_t lqry) {
  if (lref <= 0 || lqry <= 0) return -1;
  const bool wide = pp.mode == MHMSW_MODE_PLAIN || !pp.fast_ok;
  if (pp.score2) {  // rows are SSW-ref
    if (!wide && lref <= kFast16MaxRows && lref < pp.g64_min_rows && lqry + kMaxQueryPad <= kFastMaxCols)
      return kClassFast16 + fast16_class_of(lref);
    if (!wide && lref <= kFast64MaxRows && lqry + kMaxQueryPad <= kFastMaxCols)
      return kClassFast64 + (lref + 127) / 128 - 1;
    return kClassPlain;
  }
  const int32_t shorter = std::min(lref, lqry), longer = std::max(lref, lqry);
  if (!wide && shorter <= kFast16MaxRows && shorter < pp.g64_min_rows && longer <= kFastMaxCols)
    return kClassFast16 + fast16_class_of(shorter);
  if (!wide && shorter <= kFast64MaxRows && longer <= kFastMaxCols) return kClassFast64 + (shorter + 127) / 128 - 1;
  return kClassPlain;
}

// Allocator whose resize() leaves elements uninitialized: the packed code
// buffers are filled in parallel by the packing threads (each writes, and so
// first touches, its own sequences' dwords) instead of zeroed serially.
template <class T>
struct UninitAlloc : std::allocator<T> {
  template <class U>
  struct rebind {
    using other = UninitAlloc<U>;
  };
  UninitAlloc() = default;
  template <class U>
  UninitAlloc(const UninitAlloc<U>&) noexcept {}
  template <class U>
  void construct(U* p) noexcept {
    ::new ((void*)p) U;
  }
  template <class U, class... A>
  void construct(U* p, A&&... a) {
    ::new ((void*)p) U(std::forward<A>(a)...);
  }
};

// A batch laid out and packed on the host (2-bit codes + N bitmap), ready to
// be copied to a context.  Packing is pure CPU work, so the serving loop does
// it on the submitting thread.
struct PackedBatch {
  int64_t npairs = 0, total = 0, cells = 0;
  int max_plain_cols = 0;
  int max_read_len = 0, max_ctg_len = 0;
  bool used[kNumClasses] = {};
  int64_t cpairs[kNumClasses] = {}, ccells[kNumClasses] = {};
  int32_t crows[kNumClasses] = {};
  int64_t nfast = 0, nplain = 0;
  std::vector<int64_t, UninitAlloc<int64_t>> base;  // [2 * npairs] first base of each sequence
  std::vector<int32_t, UninitAlloc<int32_t>> len;   // [2 * npairs]
  std::vector<uint32_t, UninitAlloc<uint32_t>> two, nm;  // every dword is written by pack_batch
  // the packed codes and tables: in two/nm/base/len, or (packed in place) in
  // the context's pinned staging
  uint32_t* two_p = nullptr;
  uint32_t* nm_p = nullptr;
  int64_t* base_p = nullptr;
  int32_t* len_p = nullptr;
  size_t two_n = 0, nm_n = 0;  // dwords
  std::vector<int32_t> win_ctg, win_off;
  const mhmsw_contigs* res = nullptr;
  std::string err;
  PackedBatch() = default;
  PackedBatch(PackedBatch&&) = default;  // moving keeps two_p / nm_p valid (vector storage moves with them)
  PackedBatch& operator=(PackedBatch&&) = default;
  PackedBatch(const PackedBatch&) = delete;  // a copy's two_p / nm_p would point into the original
  PackedBatch& operator=(const PackedBatch&) = delete;
};

using pack2::pack_seq;

// Where pack_batch writes the 2-bit codes and N bits (two_n and nm_n dwords)
// and the per-sequence base and length tables (2 * npairs each): false when
// it cannot provide them.
using PackDest = std::function<bool(size_t two_n, size_t nm_n, uint32_t** two, uint32_t** nm, int64_t** base,
                                    int32_t** len)>;

// The first of pair p's length and layout checks that fails, as a message in
// m (m[0] == 0: the pair is valid).
static void pair_error(const PackParams& pp, const StageInput& in, int64_t p, char* m, size_t mn) {
  const bool windowed = in.res != nullptr;
  const int32_t rl = in.read_len[p], cl = windowed ? in.win_len[p] : in.ctg_len[p];
  m[0] = 0;
  if (rl < 0 || cl < 0) {
    snprintf(m, mn, "negative length at pair %lld", (long long)p);
  } else if ((in.read_ptr && rl > 0 && !in.read_ptr[p]) || (!windowed && in.ctg_ptr && cl > 0 && !in.ctg_ptr[p])) {
    snprintf(m, mn, "pair %lld: null sequence pointer", (long long)p);
  } else if (windowed && (in.ctg_id[p] < 0 || in.ctg_id[p] >= in.res->n || in.win_off[p] < 0 ||
                          (int64_t)in.win_off[p] + cl > in.res->len[(size_t)in.ctg_id[p]])) {
    snprintf(m, mn, "pair %lld: window outside its resident contig", (long long)p);
  } else if (std::min(rl, cl) > kPlainMaxCols) {
    snprintf(m, mn, "pair %lld: both sequences exceed %d bases (%d, %d)", (long long)p, kPlainMaxCols, rl, cl);
  } else if ((int64_t)rl >= (1 << 20) || (int64_t)cl >= (1 << 20)) {
    snprintf(m, mn, "pair %lld: sequence longer than 2^20 bases", (long long)p);
  } else if (pp.score2 && (pp.read_is_ref ? cl : rl) + kMaxQueryPad > kPlainMaxCols) {
    snprintf(m, mn, "pair %lld: score2 needs SSW-query sequences of at most %d bases", (long long)p,
             kPlainMaxCols - kMaxQueryPad);
  }
}

// Packs a batch: validates every pair, lays out its sequences (32-base
// blocks from kPadBases on; string batches interleave each pair's read and
// window, windowed batches put every read first and every window after them,
// and only the reads are packed on the host: the device extracts the windows
// from the resident contigs), routes each pair to its forward work class and
// packs the sequences to 2-bit codes plus N bits.  The layout is summed on the
// calling thread (one pass over the lengths); validation, routing and packing
// run in one parallel region over contiguous chunks of pairs.
static int pack_batch(const PackParams& pp, const StageInput& in, PackedBatch& pb, int max_threads,
                      const PackDest* dest = nullptr) {
  const int64_t npairs = in.npairs;
  const bool windowed = in.res != nullptr;
  const bool reads_ok = in.read_len && (in.read_ptr || (in.reads && in.read_off));
  const bool ctgs_ok = in.ctg_len && (in.ctg_ptr || (in.ctgs && in.ctg_off));
  if (npairs < 0 || (npairs > 0 && !reads_ok) || (npairs > 0 && !windowed && !ctgs_ok) ||
      (npairs > 0 && windowed && (!in.ctg_id || !in.win_off || !in.win_len))) {
    pb.err = "bad batch arguments";
    return MHMSW_E_ARG;
  }
  if (npairs > (int64_t)1 << 30) {
    pb.err = "batch too large (more than 2^30 pairs)";
    return MHMSW_E_ARG;
  }
  const int64_t nseq = 2 * npairs;
  pb = PackedBatch();
  pb.npairs = npairs;
  pb.res = in.res;
  int nth = std::max(1, std::min(max_threads, (int)std::max(1u, std::thread::hardware_concurrency())));
  if (npairs < 4096) nth = 1;
  struct Chunk {
    int64_t p0 = 0, p1 = 0, cells = 0, nfast = 0, nplain = 0;
    int max_plain_cols = 0, max_rl = 0, max_cl = 0;
    bool used[kNumClasses] = {};
    int64_t cpairs[kNumClasses] = {}, ccells[kNumClasses] = {};
    int32_t crows[kNumClasses] = {};
    int64_t bad = -1;  // first invalid pair of the chunk
    std::string err;
  };
  const int nchunks = nth == 1 ? 1 : 4 * nth;
  std::vector<Chunk> ch((size_t)nchunks);
  for (int k = 0; k < nchunks; k++) {
    ch[(size_t)k].p0 = npairs * k / nchunks;
    ch[(size_t)k].p1 = npairs * (k + 1) / nchunks;
  }
  // layout: each chunk's first base (reads region, and the windows region of
  // a windowed batch); lengths outside 0..2^20 count as 0 here and fail the
  // batch below
  auto span = [](int32_t len) -> int64_t {
    return (uint32_t)len < (1u << 20) ? ((int64_t)len + kSeqAlign - 1) / kSeqAlign * kSeqAlign : 0;
  };
  const int32_t* clen = windowed ? in.win_len : in.ctg_len;
  std::vector<int64_t> chunk_base((size_t)nchunks), chunk_wbase((size_t)nchunks);
  bool suspect = false;
  int64_t acc = kPadBases;
  for (int k = 0; k < nchunks; k++) {
    chunk_base[(size_t)k] = acc;
    for (int64_t p = ch[(size_t)k].p0; p < ch[(size_t)k].p1; p++) {
      const int32_t rl = in.read_len[p], cl = clen[p];
      suspect |= (uint32_t)rl >= (1u << 20) || (uint32_t)cl >= (1u << 20);
      acc += span(rl) + (windowed ? 0 : span(cl));
    }
  }
  const int64_t reads_end = acc;  // windowed: the packed region ends here
  if (windowed) {
    for (int k = 0; k < nchunks; k++) {
      chunk_wbase[(size_t)k] = acc;
      for (int64_t p = ch[(size_t)k].p0; p < ch[(size_t)k].p1; p++) acc += span(clen[p]);
    }
  }
  const int64_t body_end = acc;
  pb.total = (acc + kPadBases + 31) / 32 * 32;
  if (suspect || pb.total >= (int64_t)1 << 33) {
    // the first invalid pair fails the batch, as the parallel pass reports it
    char m[256];
    for (int64_t p = 0; p < npairs; p++) {
      pair_error(pp, in, p, m, sizeof(m));
      if (m[0]) {
        pb.err = m;
        return MHMSW_E_ARG;
      }
    }
    if (pb.total >= (int64_t)1 << 33) {  // the band kernels index code dwords with 32 bits
      pb.err = "batch too large (2^33 bases or more with padding)";
      return MHMSW_E_ARG;
    }
  }
  const int64_t packed_end = windowed ? reads_end : pb.total;
  pb.two_n = (size_t)(packed_end / 16);
  pb.nm_n = (size_t)(packed_end / 32);
  if (dest) {
    if (!(*dest)(pb.two_n, pb.nm_n, &pb.two_p, &pb.nm_p, &pb.base_p, &pb.len_p)) {
      pb.err = "cannot allocate the pinned staging buffer";
      return MHMSW_E_NOMEM;
    }
  } else {
    pb.two.resize(pb.two_n);
    pb.nm.resize(pb.nm_n);
    pb.base.resize((size_t)std::max<int64_t>(nseq, 1));
    pb.len.resize((size_t)std::max<int64_t>(nseq, 1));
    pb.two_p = pb.two.data();
    pb.nm_p = pb.nm.data();
    pb.base_p = pb.base.data();
    pb.len_p = pb.len.data();
  }
  // pads before the first and after the last packed sequence
  std::fill(pb.two_p, pb.two_p + kPadBases / 16, 0u);
  std::fill(pb.nm_p, pb.nm_p + kPadBases / 32, 0u);
  if (!windowed) {
    std::fill(pb.two_p + body_end / 16, pb.two_p + pb.two_n, 0u);
    std::fill(pb.nm_p + body_end / 32, pb.nm_p + pb.nm_n, 0u);
  }
  // one parallel region: validate, place, route and pack each chunk's pairs
  std::atomic<int> next_chunk(0);
  auto region = [&](int) {
    char m[256];
    for (int k; (k = next_chunk.fetch_add(1)) < nchunks;) {
      Chunk& c = ch[(size_t)k];
      int64_t at = chunk_base[(size_t)k], wat = windowed ? chunk_wbase[(size_t)k] : 0;
      for (int64_t p = c.p0; p < c.p1; p++) {
        pair_error(pp, in, p, m, sizeof(m));
        if (m[0]) {
          c.bad = p;
          c.err = m;
          break;
        }
        const int32_t rl = in.read_len[p], cl = clen[p];
        pb.len_p[2 * p] = rl;
        pb.len_p[2 * p + 1] = cl;
        // pack_seq writes every dword of a sequence's 32-base blocks
        static_assert(kSeqAlign == 32, "sequences start on 32-base blocks");
        pb.base_p[2 * p] = at;
        pack_seq(in.read_ptr ? in.read_ptr[p] : in.reads + in.read_off[p], rl, pb.two_p + at / 16, pb.nm_p + at / 32);
        at += span(rl);
        if (windowed) {
          pb.base_p[2 * p + 1] = wat;
          wat += span(cl);
        } else {
          pb.base_p[2 * p + 1] = at;
          pack_seq(in.ctg_ptr ? in.ctg_ptr[p] : in.ctgs + in.ctg_off[p], cl, pb.two_p + at / 16, pb.nm_p + at / 32);
          at += span(cl);
        }
        c.cells += (int64_t)rl * cl;
        c.max_plain_cols = std::max(c.max_plain_cols, (int)std::min(rl, cl));
        c.max_rl = std::max(c.max_rl, (int)rl);
        c.max_cl = std::max(c.max_cl, (int)cl);
        const int32_t lref = pp.read_is_ref ? rl : cl, lqry = pp.read_is_ref ? cl : rl;
        const int cls = route_class(pp, lref, lqry);
        if (cls >= 0) {
          c.used[cls] = true;
          c.cpairs[cls]++;
          c.ccells[cls] += (int64_t)rl * cl;
          c.crows[cls] = std::max(c.crows[cls], route_rows(pp, lref, lqry));
          if (cls == kClassPlain) {
            c.nplain++;
          } else {
            c.nfast++;
          }
        }
      }
    }
  };
  hostpool::pool().run_claimed(nth, std::function<void(int)>(region));
  // the first invalid pair fails the batch; otherwise the totals
  for (int k = 0; k < nchunks; k++) {
    const Chunk& c = ch[(size_t)k];
    if (c.bad >= 0) {
      pb.err = c.err;
      return MHMSW_E_ARG;
    }
    pb.cells += c.cells;
    pb.max_plain_cols = std::max(pb.max_plain_cols, c.max_plain_cols);
    pb.max_read_len = std::max(pb.max_read_len, c.max_rl);
    pb.max_ctg_len = std::max(pb.max_ctg_len, c.max_cl);
    for (int q = 0; q < kNumClasses; q++) {
      pb.used[q] = pb.used[q] || c.used[q];
      pb.cpairs[q] += c.cpairs[q];
      pb.ccells[q] += c.ccells[q];
      pb.crows[q] = std::max(pb.crows[q], c.crows[q]);
    }
    pb.nfast += c.nfast;
    pb.nplain += c.nplain;
  }
  if (windowed) {
    pb.win_ctg.assign(in.ctg_id, in.ctg_id + npairs);
    pb.win_off.assign(in.win_off, in.win_off + npairs);
  }
  return MHMSW_OK;
}

static PackParams pack_params(const mhmsw_ctx* c) {
  PackParams pp;
  pp.mode = c->mode;
  pp.fast_ok = c->sp.fast_ok;
  pp.read_is_ref = c->read_is_ref;
  pp.g64_min_rows = c->tuning.g64_min_rows;
  pp.score2 = c->score2;
  return pp;
}

// Recompute the forward class table of the staged batch (after a change of
// mode or tuning) from the stored lengths.
static void reroute(mhmsw_ctx* c) {
  const PackParams pp = pack_params(c);
  for (int k = 0; k < kNumClasses; k++) {
    c->fast_class_used[k] = false;
    c->class_pairs[k] = 0;
    c->class_cells[k] = 0;
    c->class_max_rows[k] = 0;
  }
  c->n_fast = c->n_plain = 0;
  for (int64_t p = 0; p < c->npairs; p++) {
    const int32_t rl = c->lens[(size_t)(2 * p)], cl = c->lens[(size_t)(2 * p + 1)];
    const int32_t lref = pp.read_is_ref ? rl : cl, lqry = pp.read_is_ref ? cl : rl;
    const int cls = route_class(pp, lref, lqry);
    if (cls < 0) continue;
    c->fast_class_used[cls] = true;
    c->class_pairs[cls]++;
    c->class_cells[cls] += (int64_t)lref * lqry;
    c->class_max_rows[cls] = std::max(c->class_max_rows[cls], route_rows(pp, lref, lqry));
    if (cls == kClassPlain) {
      c->n_plain++;
    } else {
      c->n_fast++;
    }
  }
}

// Copy a packed batch to the context's device buffers; returns when the host
// buffers may be released.
// Copy packed batches to the context as one job: the parts follow each other
// in the device layout (every part's total is a multiple of 32 bases, so
// each sequence keeps its 32-base alignment) and only each part's packed
// region crosses PCIe (windowed batches pack their reads only).  Returns when
// the host buffers may be released.
// memcpy split over up to 16 threads for large copies (host memory bandwidth
// is several times one core's copy rate)
static void par_memcpy(void* dst, const void* src, size_t bytes) {
  constexpr size_t kChunk = (size_t)256 << 10;
  const int nth = (int)std::min<size_t>(std::min<size_t>(16, std::max(1u, std::thread::hardware_concurrency())),
                                        (bytes + kChunk - 1) / kChunk);
  if (nth <= 1) {
    std::memcpy(dst, src, bytes);
    return;
  }
  hostpool::pool().run(nth, [&](int t) {
    const size_t a = bytes * (size_t)t / (size_t)nth, e = bytes * (size_t)(t + 1) / (size_t)nth;
    std::memcpy(static_cast<uint8_t*>(dst) + a, static_cast<const uint8_t*>(src) + a, e - a);
  });
}

// Device buffers of a job of npairs pairs over total padded bases (they only
// grow: a later, smaller job reuses them).
static hipError_t ensure_job_buffers(mhmsw_ctx* c, int64_t npairs, int64_t total) {
  const size_t np1 = (size_t)std::max<int64_t>(npairs, 1), ns1 = (size_t)std::max<int64_t>(2 * npairs, 1);
  hipError_t e = c->d_packed2.ensure((size_t)(total / 16 + 4));
  if (e == hipSuccess) e = c->d_nmask.ensure((size_t)(total / 32 + 4));
  if (e == hipSuccess) e = c->d_codes.ensure((size_t)(total / 8 + 64));
  if (e == hipSuccess) e = c->d_s2.ensure(np1);
  if (e == hipSuccess) e = c->d_base.ensure(ns1);
  if (e == hipSuccess) e = c->d_len.ensure(ns1);
  if (e == hipSuccess) e = c->d_bucket.ensure(np1);
  if (e == hipSuccess) e = c->d_rank.ensure(np1);
  if (e == hipSuccess) e = c->d_fwd.ensure(np1);
  if (e == hipSuccess) e = c->d_rev.ensure(np1);
  if (e == hipSuccess) e = c->d_results.ensure(np1);
  if (e == hipSuccess) e = c->d_tasks.ensure(np1 + 64 * kNumClasses + 64);
  return e;
}

// Device bytes the context holds (every device buffer at its capacity).
static int64_t device_bytes(const mhmsw_ctx* c) {
  int64_t b = 0;
  auto add = [&](const auto& buf) { b += (int64_t)(buf.cap * sizeof(*buf.ptr)); };
  add(c->d_packed2), add(c->d_nmask), add(c->d_codes), add(c->d_base), add(c->d_len), add(c->d_bucket);
  add(c->d_rank), add(c->d_errors), add(c->d_counters), add(c->d_win_ctg), add(c->d_win_off), add(c->d_fwd);
  add(c->d_rev), add(c->d_s2), add(c->d_tasks), add(c->d_sched), add(c->d_results), add(c->d_cig_lists);
  add(c->d_cig_count), add(c->d_cig_band), add(c->d_cig_rows), add(c->d_cig_nops), add(c->d_cig_mism);
  add(c->d_cig_off), add(c->d_cig_dirs), add(c->d_cig_gen), add(c->d_cig_ops), add(c->d_cig_used);
  return b;
}

constexpr int64_t kCigDirsDwords = 64ll << 20;

// Device bytes the buffers of a job of npairs pairs over total padded bases
// ask for: ensure_job_buffers, the window tables of a windowed batch, and
// prepare_cigar's first-pass pools when the CIGAR pass runs (a pass that
// outgrows them reruns with larger ones).  DevBuf allocates 1.25x what it is
// asked for (at least 64 elements).
static int64_t job_buffer_bytes(int64_t npairs, int64_t total, bool windowed, bool cigar) {
  const int64_t np = std::max<int64_t>(npairs, 1), ns = std::max<int64_t>(2 * npairs, 1);
  int64_t b = (total / 16 + 4) * 4 + (total / 32 + 4) * 4 + (total / 8 + 64) * 4;  // 2-bit, N bits, nibbles
  b += np * 4 + ns * 8 + ns * 4 + 2 * np * 4 + 2 * np * (int64_t)sizeof(PairEnd) + np * (int64_t)sizeof(AlnResult) +
       (np + 64 * kNumClasses + 64) * (int64_t)sizeof(SwTask);
  if (windowed) b += 2 * np * 4;
  if (cigar)
    b += kCigClasses * np * 4 + 2 * kCigClasses * 4 + np * 4 + 4 + 2 * np * 4 + np * 8 + 8 * 8 + (12 * np + 1024) * 4 +
         std::max<int64_t>(4ll << 20, 32 * np) * 4 + kCigDirsDwords * 4;
  return b;
}

// what such a job may allocate on a context that holds no buffers yet
static int64_t job_alloc_bound(int64_t npairs, int64_t total, bool windowed, bool cigar) {
  return job_buffer_bytes(npairs, total, windowed, cigar) * 5 / 4 + (1 << 20);
}

// Where a job's host data sit in the pinned staging: each part's 2-bit codes
// and N bits, then the per-sequence tables and the window tables of the
// whole job, each on a 256-byte boundary.
struct StageLayout {
  std::vector<size_t> at_two, at_nm;
  size_t at_base = 0, at_len = 0, at_wc = 0, at_wo = 0, bytes = 0;
};
static StageLayout stage_layout(const size_t* two_n, const size_t* nm_n, size_t nparts, int64_t npairs, bool windows) {
  StageLayout L;
  auto at = [&](size_t b) {
    const size_t a = L.bytes;
    L.bytes += (b + 255) & ~(size_t)255;
    return a;
  };
  L.at_two.resize(nparts);
  L.at_nm.resize(nparts);
  for (size_t k = 0; k < nparts; k++) {
    L.at_two[k] = at(two_n[k] * 4);
    L.at_nm[k] = at(nm_n[k] * 4);
  }
  L.at_base = at((size_t)(2 * npairs) * 8);
  L.at_len = at((size_t)(2 * npairs) * 4);
  L.at_wc = at(windows ? (size_t)npairs * 4 : 0);
  L.at_wo = at(windows ? (size_t)npairs * 4 : 0);
  return L;
}

static int upload_parts(mhmsw_ctx* c, const PackedBatch* const* parts, size_t nparts) {
  if (nparts == 0) return fail(c, MHMSW_E_ARG, "no batch");
  const mhmsw_contigs* res = parts[0]->res;
  for (size_t k = 0; k < nparts; k++) {
    if (parts[k]->res != res) return fail(c, MHMSW_E_ARG, "parts of one job use different resident contigs");
  }
  if (res && res->device != c->device)
    return fail(c, MHMSW_E_ARG, "resident contigs live on device %d, context on %d", res->device, c->device);
  if (c->in_flight) {
    int rc = mhmsw_wait(c);
    if (rc != MHMSW_OK) return rc;
  }
  HIP_OR_FAIL(c, hipSetDevice(c->device));
  HIP_OR_FAIL(c, hipStreamSynchronize(c->stream));  // copies of an earlier upload (a stage without a job)
  c->staged = false;
  c->have_results = false;
  int64_t npairs = 0, total = 0;
  c->cells = 0;
  c->max_plain_cols = c->max_read_len = c->max_ctg_len = 0;
  c->n_fast = c->n_plain = 0;
  std::memset(c->fast_class_used, 0, sizeof(c->fast_class_used));
  std::memset(c->class_pairs, 0, sizeof(c->class_pairs));
  std::memset(c->class_cells, 0, sizeof(c->class_cells));
  std::memset(c->class_max_rows, 0, sizeof(c->class_max_rows));
  for (size_t k = 0; k < nparts; k++) {
    const PackedBatch& pb = *parts[k];
    npairs += pb.npairs;
    total += pb.total;
    c->cells += pb.cells;
    c->max_plain_cols = std::max(c->max_plain_cols, pb.max_plain_cols);
    c->max_read_len = std::max(c->max_read_len, pb.max_read_len);
    c->max_ctg_len = std::max(c->max_ctg_len, pb.max_ctg_len);
    c->n_fast += pb.nfast;
    c->n_plain += pb.nplain;
    for (int q = 0; q < kNumClasses; q++) {
      c->fast_class_used[q] = c->fast_class_used[q] || pb.used[q];
      c->class_pairs[q] += pb.cpairs[q];
      c->class_cells[q] += pb.ccells[q];
      c->class_max_rows[q] = std::max(c->class_max_rows[q], pb.crows[q]);
    }
  }
  const int64_t nseq = 2 * npairs;
  c->npairs = npairs;
  c->total = total;
  HIP_OR_FAIL(c, ensure_job_buffers(c, npairs, total));
  // everything the job reads from the host, laid out in the pinned staging;
  // the per-sequence tables of several parts are joined there, with their
  // bases shifted by the parts before them
  const bool windows = res && npairs > 0;
  std::vector<size_t> two_n(nparts), nm_n(nparts);
  for (size_t k = 0; k < nparts; k++) {
    two_n[k] = parts[k]->two_n;
    nm_n[k] = parts[k]->nm_n;
  }
  const StageLayout L = stage_layout(two_n.data(), nm_n.data(), nparts, npairs, windows);
  const std::vector<size_t>& at_two = L.at_two;
  const std::vector<size_t>& at_nm = L.at_nm;
  const size_t at_base = L.at_base, at_len = L.at_len, at_wc = L.at_wc, at_wo = L.at_wo;
  if (windows) {
    HIP_OR_FAIL(c, c->d_win_ctg.ensure((size_t)npairs));
    HIP_OR_FAIL(c, c->d_win_off.ensure((size_t)npairs));
  }
  HIP_OR_FAIL(c, c->h_stage.ensure(L.bytes));  // no reallocation when a part was packed in place
  uint8_t* hs = c->h_stage.ptr;
  int64_t* sbase = reinterpret_cast<int64_t*>(hs + at_base);
  int32_t* slen = reinterpret_cast<int32_t*>(hs + at_len);
  int32_t* swc = reinterpret_cast<int32_t*>(hs + at_wc);
  int32_t* swo = reinterpret_cast<int32_t*>(hs + at_wo);
  // fill the staging (on up to 4 threads for large jobs), then copy it out
  auto fill_part = [&](size_t k, int64_t sh, int64_t ps) {
    const PackedBatch& pb = *parts[k];
    // one part alone fills the staging on its own up to 16 threads per array
    auto copy = [&](void* d, const void* src, size_t n) {
      if (nparts == 1) {
        par_memcpy(d, src, n);
      } else {
        std::memcpy(d, src, n);
      }
    };
    // a part packed in place already sits at its staging offset
    if (pb.two_n && pb.two_p != reinterpret_cast<uint32_t*>(hs + at_two[k]))
      copy(hs + at_two[k], pb.two_p, pb.two_n * 4);
    if (pb.nm_n && pb.nm_p != reinterpret_cast<uint32_t*>(hs + at_nm[k])) copy(hs + at_nm[k], pb.nm_p, pb.nm_n * 4);
    const size_t nq = (size_t)(2 * pb.npairs);
    if (sh == 0) {
      if (nq && pb.base_p != sbase + 2 * ps) copy(sbase + 2 * ps, pb.base_p, nq * 8);
    } else {
      for (size_t q = 0; q < nq; q++) sbase[2 * ps + (int64_t)q] = pb.base_p[q] + sh;
    }
    if (nq && pb.len_p != slen + 2 * ps) copy(slen + 2 * ps, pb.len_p, nq * 4);
    if (windows && pb.npairs > 0) {
      std::memcpy(swc + ps, pb.win_ctg.data(), (size_t)pb.npairs * 4);
      std::memcpy(swo + ps, pb.win_off.data(), (size_t)pb.npairs * 4);
    }
  };
  std::vector<int64_t> part_sh(nparts), part_ps(nparts);
  {
    int64_t sh = 0, ps = 0;
    for (size_t k = 0; k < nparts; k++) {
      part_sh[k] = sh;
      part_ps[k] = ps;
      sh += parts[k]->total;
      ps += parts[k]->npairs;
    }
  }
  const int nfill = nparts > 1 && npairs >= (1 << 16) ? (int)std::min<size_t>(4, nparts) : 1;
  if (nfill == 1) {
    for (size_t k = 0; k < nparts; k++) fill_part(k, part_sh[k], part_ps[k]);
  } else {
    std::atomic<size_t> next(0);
    auto worker = [&]() {
      for (size_t k; (k = next.fetch_add(1)) < nparts;) fill_part(k, part_sh[k], part_ps[k]);
    };
    std::vector<std::thread> th;
    for (int t = 1; t < nfill; t++) th.emplace_back(worker);
    worker();
    for (auto& t : th) t.join();
  }
  for (size_t k = 0; k < nparts; k++) {
    const PackedBatch& pb = *parts[k];
    if (pb.two_n)
      HIP_OR_FAIL(c, hipMemcpyAsync(c->d_packed2.ptr + part_sh[k] / 16, hs + at_two[k], pb.two_n * 4,
                                    hipMemcpyHostToDevice, c->stream));
    if (pb.nm_n)
      HIP_OR_FAIL(c, hipMemcpyAsync(c->d_nmask.ptr + part_sh[k] / 32, hs + at_nm[k], pb.nm_n * 4,
                                    hipMemcpyHostToDevice, c->stream));
  }
  c->lens.resize((size_t)nseq);
  if (nseq > 0) {
    std::memcpy(c->lens.data(), slen, (size_t)nseq * sizeof(int32_t));
    HIP_OR_FAIL(c, hipMemcpyAsync(c->d_base.ptr, sbase, (size_t)nseq * 8, hipMemcpyHostToDevice, c->stream));
    HIP_OR_FAIL(c, hipMemcpyAsync(c->d_len.ptr, slen, (size_t)nseq * 4, hipMemcpyHostToDevice, c->stream));
  }
  c->res = res;
  if (windows) {
    HIP_OR_FAIL(c, hipMemcpyAsync(c->d_win_ctg.ptr, swc, (size_t)npairs * 4, hipMemcpyHostToDevice, c->stream));
    HIP_OR_FAIL(c, hipMemcpyAsync(c->d_win_off.ptr, swo, (size_t)npairs * 4, hipMemcpyHostToDevice, c->stream));
  }
  c->staged = true;
  return MHMSW_OK;
}

static int upload_batch(mhmsw_ctx* c, const PackedBatch& pb) {
  const PackedBatch* one = &pb;
  return upload_parts(c, &one, 1);
}

static int stage_impl(mhmsw_ctx* c, const StageInput& in) {
  if (in.res && in.res->device != c->device)
    return fail(c, MHMSW_E_ARG, "resident contigs live on device %d, context on %d", in.res->device, c->device);
  // The batch is packed straight into the pinned staging, at the offsets the
  // upload copies from: first nothing may still read it (a job, or the
  // copies of an earlier stage).
  if (c->in_flight) {
    int rc = mhmsw_wait(c);
    if (rc != MHMSW_OK) return rc;
  }
  HIP_OR_FAIL(c, hipSetDevice(c->device));
  HIP_OR_FAIL(c, hipStreamSynchronize(c->stream));
  const bool windows = in.res != nullptr && in.npairs > 0;
  const PackDest dest = [&](size_t two_n, size_t nm_n, uint32_t** two, uint32_t** nm, int64_t** base,
                            int32_t** len) {
    const StageLayout L = stage_layout(&two_n, &nm_n, 1, in.npairs, windows);
    if (c->h_stage.ensure(L.bytes) != hipSuccess) return false;
    *two = reinterpret_cast<uint32_t*>(c->h_stage.ptr + L.at_two[0]);
    *nm = reinterpret_cast<uint32_t*>(c->h_stage.ptr + L.at_nm[0]);
    *base = reinterpret_cast<int64_t*>(c->h_stage.ptr + L.at_base);
    *len = reinterpret_cast<int32_t*>(c->h_stage.ptr + L.at_len);
    return true;
  };
  PackedBatch pb;
  const auto t0 = std::chrono::steady_clock::now();
  int rc = pack_batch(pack_params(c), in, pb, 16, &dest);
  if (rc != MHMSW_OK) return fail(c, rc, "%s", pb.err.c_str());
  const auto t1 = std::chrono::steady_clock::now();
  rc = upload_batch(c, pb);
  c->host_pack_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
  c->host_upload_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t1).count();
  return rc;
}

extern "C" {

int mhmsw_stage(mhmsw_ctx* c, int64_t npairs, const char* reads, const int64_t* read_off, const int32_t* read_len,
                const char* ctgs, const int64_t* ctg_off, const int32_t* ctg_len) {
  if (!c) return MHMSW_E_ARG;
  StageInput in;
  in.npairs = npairs;
  in.reads = reads;
  in.read_off = read_off;
  in.read_len = read_len;
  in.ctgs = ctgs;
  in.ctg_off = ctg_off;
  in.ctg_len = ctg_len;
  return stage_impl(c, in);
}

int mhmsw_stage_seqs(mhmsw_ctx* c, int64_t npairs, const char* const* reads, const int32_t* read_len,
                     const char* const* ctgs, const int32_t* ctg_len) {
  if (!c) return MHMSW_E_ARG;
  StageInput in;
  in.npairs = npairs;
  in.read_ptr = reads;
  in.read_len = read_len;
  in.ctg_ptr = ctgs;
  in.ctg_len = ctg_len;
  return stage_impl(c, in);
}

int mhmsw_stage_windows(mhmsw_ctx* c, const mhmsw_contigs* contigs, int64_t npairs, const char* reads,
                        const int64_t* read_off, const int32_t* read_len, const int32_t* ctg_id, const int32_t* win_off,
                        const int32_t* win_len) {
  if (!c || !contigs) return MHMSW_E_ARG;
  StageInput in;
  in.npairs = npairs;
  in.reads = reads;
  in.read_off = read_off;
  in.read_len = read_len;
  in.res = contigs;
  in.ctg_id = ctg_id;
  in.win_off = win_off;
  in.win_len = win_len;
  return stage_impl(c, in);
}

}  // extern "C"


// Columns the 32-bit kernel keeps in LDS: the shorter sequence, or with
// score2 the SSW-query plus its padding (any pair may be rerun there).
static int plain_lds_cols(const mhmsw_ctx* c) {
  if (c->score2) return (c->read_is_ref ? c->max_ctg_len : c->max_read_len) + kMaxQueryPad;
  return c->max_plain_cols;
}

// Jobs with at most this many packed-form pairs run every band class of the
// reverse pass in one launch, and their reverse rectangles in one class per
// group size (the largest the forward pass used: a reverse rectangle is never
// taller than its forward one).  An empty persistent launch costs about 4.5 us,
// a few times what running these jobs' few rectangles at the larger P costs.
constexpr int64_t kBandAllMaxPairs = 65536;
// A small job's widest band task: 16 diagonals (P <= 4).  Measured on 10k,
// 20k and 50k pairs of 150 x 350 bases, reverse pass at 1, 3, 7 and 15:
// 0.150 / 0.134 / 0.113 / 0.143 ms, 0.198 / 0.164 / 0.116 / 0.151 ms and
// 0.348 / 0.275 / 0.157 / 0.158 ms.
constexpr int32_t kSmallJobBandMaxD = 7;
// blocks per CU of the diagonal class's launch in large jobs
constexpr int kDiagBlocksPerCu = 8;

// Forward launches keep H as 32 H in the narrow shapes (at most 8 lanes per
// pair; a fold of the row maxima every 32 steps instead of 16) when the
// scoring table's bytes still fit ((match + gap open) * 32 <= 255) and no pair
// of class k can reach kFastOverflowScore32 (a score is at most match times
// the rows).
static bool fold32_ok(const mhmsw_ctx* c, int k) {
  return c->sp.fast_ok && c->sp.match + c->sp.gap_open <= 7 &&
         (int64_t)c->class_max_rows[k] * c->sp.match < kFastOverflowScore32;
}

static BatchView make_view(mhmsw_ctx* c) {
  BatchView b;
  std::memset(&b, 0, sizeof(b));  // the graph cache compares views bytewise
  b.band_max_d = kBandMaxD;
  b.npairs = c->npairs;
  b.total = c->total;
  b.seq_base = c->d_base.ptr;
  b.seq_len = c->d_len.ptr;
  b.packed2 = c->d_packed2.ptr;
  b.nmask = c->d_nmask.ptr;
  b.codes = c->d_codes.ptr;
  b.fwd = c->d_fwd.ptr;
  b.rev = c->d_rev.ptr;
  b.sched = c->d_sched.ptr;
  b.tasks = c->d_tasks.ptr;
  b.bucket = c->d_bucket.ptr;
  b.rank = c->d_rank.ptr;
  b.results = c->d_results.ptr;
  b.errors = c->d_errors.ptr;
  b.counters = c->d_counters.ptr;
  b.sc = c->sp;
  b.mode = c->mode;
  b.g64_min_rows = c->tuning.g64_min_rows;
  b.bin_shift = c->tuning.bin_shift;
  b.score2 = c->score2;
  b.s2 = c->d_s2.ptr;
  if (c->n_fast <= kBandAllMaxPairs) {
    int maxp16 = 0, maxp64 = 0;
    for (int P = 1; P <= kMaxP; P++) {
      if (c->fast_class_used[kClassFast16 + P - 1]) maxp16 = P;
      if (c->fast_class_used[kClassFast64 + P - 1]) maxp64 = P;
    }
    b.rev_one16 = maxp64 > 0 ? kMaxP : maxp16;
    b.rev_one64 = maxp64;
    // A small job's band pass is latency-bound (a wavefront or two per SIMD):
    // its time is the longest lane's chain, 2P cell updates per u-step.
    // Bands wider than 7 diagonals each side run as rectangles in 16-lane
    // groups, whose chain per column is P cell updates for up to 32P rows.
    b.band_max_d = kSmallJobBandMaxD;
  } else if (c->mode == MHMSW_MODE_FAST) {
    // wide-band rectangles probed first; the few full rectangles of pass
    // kPassRev2 run in the largest class of each group size
    int maxp16 = 0, maxp64 = 0;
    for (int P = 1; P <= kMaxP; P++) {
      if (c->fast_class_used[kClassFast16 + P - 1]) maxp16 = P;
      if (c->fast_class_used[kClassFast64 + P - 1]) maxp64 = P;
    }
    b.rev_probe = 1;
    b.rev2_one16 = maxp64 > 0 ? kMaxP : maxp16;
    b.rev2_one64 = maxp64;
  }
  if (c->res) {
    b.res2 = c->res->d2;
    b.resN = c->res->dN;
    b.res_base = c->res->dbase;
    b.win_ctg = c->d_win_ctg.ptr;
    b.win_off = c->d_win_off.ptr;
  }
  return b;
}

// Device buffers of the CIGAR pass for the staged batch: op pool and generic
// pool grow when a pass overflows them (complete() reruns the pass);
// direction scratch is a fixed 256 MiB, kCigDirsDwords (one slot per
// wavefront, rows of the tallest narrow rectangle; a slot of
// kCigNarrowMaxRows rows always fits).
static int prepare_cigar(mhmsw_ctx* c) {
  const size_t np = (size_t)std::max<int64_t>(c->npairs, 1);
  c->cig_ops_cap = std::max<int64_t>(c->cig_ops_cap, 12 * (int64_t)np + 1024);
  // wide-band pool: about 1% of pairs take a wide band, ~4 KB each for 250-row rectangles
  c->cig_gen_dwords = std::max<int64_t>(c->cig_gen_dwords, std::max<int64_t>(4ll << 20, 32 * (int64_t)np));
  c->cig_dirs_dwords = kCigDirsDwords;
  HIP_OR_FAIL(c, c->d_cig_lists.ensure((size_t)kCigClasses * np));
  HIP_OR_FAIL(c, c->d_cig_count.ensure(2 * kCigClasses));
  HIP_OR_FAIL(c, c->d_cig_band.ensure(np));
  HIP_OR_FAIL(c, c->d_cig_rows.ensure(1));
  HIP_OR_FAIL(c, c->d_cig_nops.ensure(np));
  HIP_OR_FAIL(c, c->d_cig_mism.ensure(np));
  HIP_OR_FAIL(c, c->d_cig_off.ensure(np));
  HIP_OR_FAIL(c, c->d_cig_used.ensure(8));
  HIP_OR_FAIL(c, c->d_cig_ops.ensure((size_t)c->cig_ops_cap));
  HIP_OR_FAIL(c, c->d_cig_gen.ensure((size_t)c->cig_gen_dwords));
  HIP_OR_FAIL(c, c->d_cig_dirs.ensure((size_t)c->cig_dirs_dwords));
  return MHMSW_OK;
}

static CigarView make_cigar_view(mhmsw_ctx* c) {
  CigarView v;
  std::memset(&v, 0, sizeof(v));  // part of the graph key, compared bytewise
  v.npairs = c->npairs;
  v.codes = c->d_codes.ptr;
  v.seq_base = c->d_base.ptr;
  v.seq_len = c->d_len.ptr;
  v.results = c->d_results.ptr;
  v.sc = c->sp;
  v.lists = c->d_cig_lists.ptr;
  v.list_count = c->d_cig_count.ptr;
  v.band = c->d_cig_band.ptr;
  v.max_rows = c->d_cig_rows.ptr;
  v.dirs = c->d_cig_dirs.ptr;
  v.dirs_dwords = (int64_t)c->d_cig_dirs.cap;
  v.gen = c->d_cig_gen.ptr;
  v.gen_dwords = (int64_t)c->d_cig_gen.cap;
  v.n_ops = c->d_cig_nops.ptr;
  v.mism = c->d_cig_mism.ptr;
  v.off = c->d_cig_off.ptr;
  v.ops = c->d_cig_ops.ptr;
  v.ops_cap = (int64_t)c->d_cig_ops.cap;
  v.used = c->d_cig_used.ptr;
  return v;
}

// wavefronts of the CIGAR launches: 16 per CU for the narrow bands, 12 per CU
// for the wide ones (one pair per wavefront, 8 KB of LDS each)
static hipError_t enqueue_cigar(mhmsw_ctx* c, hipStream_t st) {
  return launch_cigar(make_cigar_view(c), c->cus * 16, c->cus * 12, st);
}


// ev: record phase events (eager launches only; events recorded inside a
// captured graph cannot be timed)
static int enqueue_job(mhmsw_ctx* c, const BatchView& b, hipStream_t st, bool ev, int* maxp16_out, int* maxp64_out) {
  // persistent grids: enough wavefronts to fill the device, but no more
  // blocks than a class can use (4 wavefronts per 256-thread block, 64 / G
  // pairs per wavefront in G-lane groups, 64 in band classes) and no more per
  // CU than the kernel keeps resident (occ), so
  // that a small job's tasks are shared out by the work counter rather than
  // one per wavefront to blocks that wait for a free CU
  // A class with few wavefront tasks runs on fewer waves per CU: with at
  // least kMinTasksPerWave tasks for each wave, the work counter shares them
  // out evenly, where one task per wave leaves the spread to the order in
  // which the dispatcher places blocks (a 20k-pair job's 1,250 blocks at up
  // to 8 per CU: job 0.55 ms; at 2 per CU: 0.45 ms).
  constexpr int64_t kMinTasksPerWave = 2;
  auto grid_cap = [&](int64_t pairs, int pairs_per_block, int occ) {
    int per_cu = occ > 0 ? std::min(c->tuning.blocks_per_cu, occ) : c->tuning.blocks_per_cu;
    const int64_t blocks = (pairs + pairs_per_block - 1) / pairs_per_block;  // one task per wave
    const int64_t fit = blocks / ((int64_t)c->cus * kMinTasksPerWave);      // blocks per CU at that share
    if (fit < per_cu) per_cu = (int)std::max<int64_t>(std::min(2, per_cu), fit);
    const int64_t full = std::max(1, c->cus * per_cu);
    return (int)std::max<int64_t>(1, std::min<int64_t>(full, blocks));
  };
  // the 32-bit form runs one wavefront per 64-thread block; its grid does not
  // follow the packed form's tuning (16 wavefronts per CU: its fastest
  // measured shape, 1.7x the rate at 8 on the length-skewed 4M-pair batch)
  const int plain_grid = (int)std::max<int64_t>(1, std::min<int64_t>((int64_t)c->cus * 16, c->npairs));
  // a packed class's launch: the kernel shape of its pass, 4 wavefronts of
  // 64 / lanes pairs per block
  auto fast = [&](const BatchView& v, int group, int P, bool forward, int64_t pairs) {
    const int k = (group == 64 ? kClassFast64 : kClassFast16) + P - 1;
    const FastShape s = fast_shape(group, P, forward, pairs, c->cus, c->class_max_rows[k]);
    BatchView w = v;
    w.fold32 = forward && fold32_ok(c, k);
    return launch_fast(w, s, grid_cap(pairs, 256 / s.lanes, c->occ_of(s)), st);
  };
  // the band kernels' grids do not follow it either: 2 blocks (8 wavefronts)
  // per CU, or fewer where a kernel keeps fewer resident.  More waves per CU
  // only share the L2 among more code streams (the P = 4 class measured
  // 238 us at 2 blocks per CU and 322 us at 3 on the uneven 1M-pair batch;
  // no class was faster at 3 or more)
  auto band_grid = [&](int64_t tasks, int occ) {
    const int per_cu = occ > 0 ? std::min(2, occ) : 2;
    return (int)std::max<int64_t>(1, std::min<int64_t>((int64_t)c->cus * per_cu, (tasks + 255) / 256));
  };
  const int plain_cols = std::max(1, plain_lds_cols(c));
  // each pass's kernels see that pass's scheduling state
  const BatchView bf = with_pass(b, kPassFwd), br = with_pass(b, kPassRerun), bv = with_pass(b, kPassRev);
  // A pair's score is at most match * its shorter length, so when no pair can
  // reach the packed form's exact range no pair is flagged for a 32-bit rerun
  // (score2 reruns aside); and when besides no pair runs 32-bit and every
  // reverse rectangle fits a packed class, the reverse pass has no 32-bit task.
  const bool rerun_possible =
      c->score2 || (int64_t)c->sp.match * c->max_plain_cols >= kFastOverflowScore;
  const bool rev_plain_possible = c->mode != MHMSW_MODE_FAST || !c->sp.fast_ok || c->n_plain > 0 || rerun_possible ||
                                  c->max_plain_cols > kFast64MaxRows ||
                                  std::max(c->max_read_len, c->max_ctg_len) > kFastMaxCols;
  if (ev) HIP_OR_FAIL(c, hipEventRecord(c->ev[0], st));
  HIP_OR_FAIL(c, launch_clear_job(b, st));
  if (c->npairs > 0) {
    HIP_OR_FAIL(c, launch_unpack(b, st));
    if (b.win_ctg) {
      if (ev) HIP_OR_FAIL(c, hipEventRecord(c->wev[0], st));
      HIP_OR_FAIL(c, launch_extract_windows(b, st));
      if (ev) HIP_OR_FAIL(c, hipEventRecord(c->wev[1], st));
    }
    HIP_OR_FAIL(c, launch_schedule(b, kPassFwd, st));
  }
  if (ev) HIP_OR_FAIL(c, hipEventRecord(c->ev[1], st));
  int maxp16 = 0, maxp64 = 0;
  if (c->npairs > 0) {
    for (int P = kMaxP; P >= 1; P--) {
      const int k = kClassFast64 + P - 1;
      if (c->fast_class_used[k]) {
        if (ev) HIP_OR_FAIL(c, hipEventRecord(c->cev[k][0], st));
        HIP_OR_FAIL(c, fast(bf, 64, P, true, c->class_pairs[k]));
        if (ev) HIP_OR_FAIL(c, hipEventRecord(c->cev[k][1], st));
        maxp64 = std::max(maxp64, P);
      }
    }
    for (int P = kMaxP; P >= 1; P--) {
      const int k = kClassFast16 + P - 1;
      if (c->fast_class_used[k]) {
        if (ev) HIP_OR_FAIL(c, hipEventRecord(c->cev[k][0], st));
        HIP_OR_FAIL(c, fast(bf, 16, P, true, c->class_pairs[k]));
        if (ev) HIP_OR_FAIL(c, hipEventRecord(c->cev[k][1], st));
        maxp16 = std::max(maxp16, P);
      }
    }
  }
  if (ev) HIP_OR_FAIL(c, hipEventRecord(c->ev[2], st));
  const bool any_fast = maxp16 > 0 || maxp64 > 0;
  if (c->npairs > 0) {
    if (c->fast_class_used[kClassPlain]) {
      if (ev) HIP_OR_FAIL(c, hipEventRecord(c->cev[kClassPlain][0], st));
      HIP_OR_FAIL(c, launch_plain(bf, plain_grid, plain_cols, st));
      if (ev) HIP_OR_FAIL(c, hipEventRecord(c->cev[kClassPlain][1], st));
    }
    if (any_fast && rerun_possible) {
      HIP_OR_FAIL(c, launch_schedule(b, kPassRerun, st));
      HIP_OR_FAIL(c, launch_plain(br, plain_grid, plain_cols, st));
    }
  }
  if (ev) HIP_OR_FAIL(c, hipEventRecord(c->ev[3], st));
  if (c->npairs > 0) {
    HIP_OR_FAIL(c, launch_schedule(b, kPassRev, st));
    if (any_fast) {
      // a reverse rectangle is never taller than its forward one, so the
      // reverse classes are bounded by the forward ones
      const int rev16 = maxp64 > 0 ? kMaxP : maxp16;
      // a small job's rectangles sit in one class per group size (make_view)
      const int low64 = b.rev_one64 > 0 ? b.rev_one64 : 1, low16 = b.rev_one16 > 0 ? b.rev_one16 : 1;
      for (int P = maxp64; P >= low64; P--)
        HIP_OR_FAIL(c, fast(bv, 64, P, false, c->n_fast));
      for (int P = rev16; P >= low16; P--)
        HIP_OR_FAIL(c, fast(bv, 16, P, false, c->n_fast));
      // pairs whose begin lies within a narrow band of diagonals: one lane each
      if (c->n_fast <= kBandAllMaxPairs) {
        // one launch for every class: enough wavefronts for every task of all
        HIP_OR_FAIL(c, launch_band(bv, 0, band_grid(c->n_fast + 64 * kNumBandClasses, c->occ_band[0]), st));
      } else {
        for (int i = 5; i >= 1; i--)
          HIP_OR_FAIL(c, launch_band(bv, 1 << (i - 1), band_grid(c->n_fast, c->occ_band[i]), st));
        // the diagonal class's lanes each stream two short code runs and do
        // little arithmetic on them: more waves per CU hide their loads
        const int diag_cu = c->occ_band[6] > 0 ? std::min(kDiagBlocksPerCu, c->occ_band[6]) : kDiagBlocksPerCu;
        HIP_OR_FAIL(c, launch_band(bv, kBandDiag,
                                   (int)std::max<int64_t>(1, std::min<int64_t>((int64_t)c->cus * diag_cu,
                                                                               (c->n_fast + 255) / 256)),
                                   st));
      }
    }
    if (rev_plain_possible) HIP_OR_FAIL(c, launch_plain(bv, plain_grid, plain_cols, st));
    if (any_fast && b.rev_probe) {
      // full rectangles of the wide-band pairs whose probe did not reach the score
      const BatchView bv2 = with_pass(b, kPassRev2);
      HIP_OR_FAIL(c, launch_schedule(b, kPassRev2, st));
      if (b.rev2_one64 > 0)
        HIP_OR_FAIL(c, fast(bv2, 64, b.rev2_one64, false, c->n_fast));
      if (b.rev2_one16 > 0)
        HIP_OR_FAIL(c, fast(bv2, 16, b.rev2_one16, false, c->n_fast));
    }
  }
  if (ev) HIP_OR_FAIL(c, hipEventRecord(c->ev[4], st));
  if (c->npairs > 0) HIP_OR_FAIL(c, launch_finalize(b, st));
  if (ev) HIP_OR_FAIL(c, hipEventRecord(c->ev[5], st));
  if (c->cigar) HIP_OR_FAIL(c, enqueue_cigar(c, st));
  if (ev) HIP_OR_FAIL(c, hipEventRecord(c->ev[6], st));
  *maxp16_out = maxp16;
  *maxp64_out = maxp64;
  return MHMSW_OK;
}

static void drop_graph(mhmsw_ctx* c) {
  if (c->graph_exec) (void)hipGraphExecDestroy(c->graph_exec);
  if (c->graph) (void)hipGraphDestroy(c->graph);
  c->graph_exec = nullptr;
  c->graph = nullptr;
  c->graph_key.clear();
  c->eager_key.clear();
}

static int align_staged_impl(mhmsw_ctx* c);

int mhmsw_align_staged(mhmsw_ctx* c) {
  if (!c) return MHMSW_E_ARG;
  if (!c->staged) return fail(c, MHMSW_E_STATE, "no staged batch");
  if (c->in_flight) {
    int rc = mhmsw_wait(c);
    if (rc != MHMSW_OK) return rc;
  }
  const auto t0 = std::chrono::steady_clock::now();
  const int rc = align_staged_impl(c);
  c->host_launch_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return rc;
}

static int align_staged_impl(mhmsw_ctx* c) {
  if (c->score2 && plain_lds_cols(c) > kPlainMaxCols)
    return fail(c, MHMSW_E_ARG, "score2 needs SSW-query sequences of at most %d bases (longest: %d)",
                kPlainMaxCols - kMaxQueryPad, plain_lds_cols(c) - kMaxQueryPad);
  HIP_OR_FAIL(c, hipSetDevice(c->device));
  hipStream_t st = c->stream;
  BatchView b = make_view(c);
  c->last_job_score2 = c->score2 != 0;
  c->last_job_cigar = c->cigar != 0;
  if (c->cigar) {
    int rc = prepare_cigar(c);
    if (rc != MHMSW_OK) return rc;
  }
  for (int k = 0; k < kNumClasses; k++) c->class_launched[k] = c->npairs > 0 && c->fast_class_used[k];
  int maxp16 = 0, maxp64 = 0;
  // The job depends only on the view, the classes used, the mode and the
  // plain LDS size.  A geometry's first job runs eagerly; it is captured as a
  // graph when the next job has the same geometry, and replayed while that
  // lasts.  One-off geometries (a driver fed new batches, serving) never pay
  // for a capture.
  std::vector<unsigned char> key;
  bool graph = false;
  if (c->use_graphs) {
    // everything the launches depend on: the view, the classes used and
    // their pair counts (grid sizes), the reverse pass's launch choice (from
    // the packed pair count), the plain kernel's LDS size (the longest shorter
    // side, which with the scoring in the view also decides whether the rerun
    // pass runs), mode, grid tuning, the CIGAR pass and the longest sequence
    // (whether reverse rectangles can need the 32-bit form)
    const size_t kv = sizeof(BatchView), kc = sizeof(c->fast_class_used), kq = sizeof(c->class_pairs),
                 kr = sizeof(c->class_max_rows);
    const int64_t extra[5] = {plain_lds_cols(c), c->mode * 64 + c->tuning.blocks_per_cu, c->cigar, c->n_fast,
                              (int64_t)std::max(c->max_read_len, c->max_ctg_len)};
    key.resize(kv + kc + kq + kr + sizeof(extra) + sizeof(CigarView));
    unsigned char* kp = key.data();
    std::memcpy(kp, &b, kv);
    std::memcpy(kp + kv, c->fast_class_used, kc);
    std::memcpy(kp + kv + kc, c->class_pairs, kq);
    std::memcpy(kp + kv + kc + kq, c->class_max_rows, kr);
    std::memcpy(kp + kv + kc + kq + kr, extra, sizeof(extra));
    if (c->cigar) {
      const CigarView cv = make_cigar_view(c);
      std::memcpy(kp + kv + kc + kq + kr + sizeof(extra), &cv, sizeof(cv));
    }
    graph = (c->graph_exec && key == c->graph_key) || key == c->eager_key;
  }
  c->last_job_graph = graph;
  c->last_job_phases = !graph && c->phase_timing;
  c->last_job_windows = c->res != nullptr && c->npairs > 0;
  c->last_job_tail_results = (c->tail_results || c->job_host_results) && c->npairs > 0;
  if (c->last_job_tail_results) HIP_OR_FAIL(c, c->h_results.ensure((size_t)c->npairs * sizeof(AlnResult)));
  // after the job: its error words, counters and CIGAR pool use into pinned
  // memory, then ev_done (outside any captured graph)
  auto enqueue_tail = [&]() -> hipError_t {
    hipError_t e = hipMemcpyAsync(c->tail->errors, c->d_errors.ptr, sizeof(c->tail->errors), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess)
      e = hipMemcpyAsync(c->tail->counters, c->d_counters.ptr, sizeof(c->tail->counters), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess && c->cigar)
      e = hipMemcpyAsync(c->tail->cig_used, c->d_cig_used.ptr, sizeof(c->tail->cig_used), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess && c->last_job_tail_results)
      e = hipMemcpyAsync(c->h_results.ptr, c->d_results.ptr, (size_t)c->npairs * sizeof(AlnResult),
                         hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipEventRecord(c->ev_done, st);
    return e;
  };
  if (!graph) {
    // without phase timing only the job's two events are recorded, as for a
    // replayed graph (each event record is host time in every call)
    if (!c->phase_timing) HIP_OR_FAIL(c, hipEventRecord(c->ev[0], st));
    int rc = enqueue_job(c, b, st, c->phase_timing, &maxp16, &maxp64);
    if (rc != MHMSW_OK) return rc;
    if (!c->phase_timing) HIP_OR_FAIL(c, hipEventRecord(c->ev[kEvents - 1], st));
    HIP_OR_FAIL(c, enqueue_tail());
    if (c->use_graphs) c->eager_key = key;
  } else {
    if (!c->graph_exec || key != c->graph_key) {
      drop_graph(c);
      HIP_OR_FAIL(c, hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
      int rc = enqueue_job(c, b, st, false, &maxp16, &maxp64);
      hipGraph_t g = nullptr;
      hipError_t e = hipStreamEndCapture(st, &g);
      if (rc != MHMSW_OK) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
      }
      if (e != hipSuccess) return fail(c, MHMSW_E_HIP, "hipStreamEndCapture: %s", hipGetErrorString(e));
      c->graph = g;
      e = hipGraphInstantiate(&c->graph_exec, c->graph, nullptr, nullptr, 0);
      if (e != hipSuccess) {
        drop_graph(c);
        return fail(c, MHMSW_E_HIP, "hipGraphInstantiate: %s", hipGetErrorString(e));
      }
      c->graph_key = key;
      c->graph_captures++;
    } else {
      c->graph_replays++;
    }
    HIP_OR_FAIL(c, hipEventRecord(c->ev[0], st));
    HIP_OR_FAIL(c, hipGraphLaunch(c->graph_exec, st));
    HIP_OR_FAIL(c, hipEventRecord(c->ev[kEvents - 1], st));
    HIP_OR_FAIL(c, enqueue_tail());
  }
  c->in_flight = true;
  c->have_results = false;
  return MHMSW_OK;
}

int mhmsw_set_phase_timing(mhmsw_ctx* c, int enable) {
  if (!c) return MHMSW_E_ARG;
  if (c->in_flight) return fail(c, MHMSW_E_STATE, "phase timing changed while a batch is in flight");
  c->phase_timing = enable != 0;
  return MHMSW_OK;
}

int mhmsw_set_graphs(mhmsw_ctx* c, int enable) {
  if (!c) return MHMSW_E_ARG;
  if (c->in_flight) return fail(c, MHMSW_E_STATE, "graph mode changed while a batch is in flight");
  c->use_graphs = enable != 0;
  if (!c->use_graphs) drop_graph(c);
  return MHMSW_OK;
}

// A submitted job's results go to the host: the job copies them into pinned
// memory as its last step, so mhmsw_results is a host copy.
static int align_for_host(mhmsw_ctx* c) {
  c->job_host_results = true;
  const int rc = mhmsw_align_staged(c);
  c->job_host_results = false;
  return rc;
}

int mhmsw_submit(mhmsw_ctx* c, int64_t npairs, const char* reads, const int64_t* read_off, const int32_t* read_len,
                 const char* ctgs, const int64_t* ctg_off, const int32_t* ctg_len) {
  int rc = mhmsw_stage(c, npairs, reads, read_off, read_len, ctgs, ctg_off, ctg_len);
  if (rc != MHMSW_OK) return rc;
  return align_for_host(c);
}

int mhmsw_submit_seqs(mhmsw_ctx* c, int64_t npairs, const char* const* reads, const int32_t* read_len,
                      const char* const* ctgs, const int32_t* ctg_len) {
  int rc = mhmsw_stage_seqs(c, npairs, reads, read_len, ctgs, ctg_len);
  if (rc != MHMSW_OK) return rc;
  return align_for_host(c);
}

// After a job with the CIGAR pass: a pass whose op pool or generic pool was
// too small is rerun with pools of the size it asked for (the device counted
// what it needed), synchronously on the context stream.
static int settle_cigar(mhmsw_ctx* c) {
  for (int attempt = 0;; attempt++) {
    if (attempt == 0) {
      std::memcpy(c->h_cig_used, c->tail->cig_used, sizeof(c->h_cig_used));  // copied by the job
    } else {
      HIP_OR_FAIL(c, d2h_sync(c, c->h_cig_used, c->d_cig_used.ptr, sizeof(c->h_cig_used)));
    }
    const unsigned long long* u = c->h_cig_used;
    if (u[3] != 0)
      return fail(c, MHMSW_E_INTERNAL, "CIGAR pass: %llu tracebacks left their band", (unsigned long long)u[3]);
    const bool grow_ops = (int64_t)u[0] > (int64_t)c->d_cig_ops.cap;
    const bool grow_gen = u[2] != 0;
    if (!grow_ops && !grow_gen) return MHMSW_OK;
    if (attempt >= 8) return fail(c, MHMSW_E_INTERNAL, "CIGAR pass: pools did not settle");
    if (grow_ops) c->cig_ops_cap = (int64_t)u[0] + (int64_t)(u[0] / 4) + 1024;
    if (grow_gen) c->cig_gen_dwords = std::max<int64_t>(2 * (int64_t)u[1], 2 * c->cig_gen_dwords);
    int rc = prepare_cigar(c);
    if (rc != MHMSW_OK) return rc;
    HIP_OR_FAIL(c, enqueue_cigar(c, c->stream));
    HIP_OR_FAIL(c, hipStreamSynchronize(c->stream));
    c->stats.cigar_reruns++;
  }
}

static int complete(mhmsw_ctx* c) {
  c->in_flight = false;
  std::memcpy(c->h_errors, c->tail->errors, sizeof(c->h_errors));
  unsigned long long counters[2] = {0, 0};
  std::memcpy(counters, c->tail->counters, sizeof(counters));
  float ms[kEvents] = {};
  float job = 0.f;
  if (!c->last_job_phases) {
    HIP_OR_FAIL(c, hipEventElapsedTime(&job, c->ev[0], c->ev[kEvents - 1]));
  } else {
    for (int i = 1; i < kEvents; i++) HIP_OR_FAIL(c, hipEventElapsedTime(&ms[i], c->ev[i - 1], c->ev[i]));
    for (int i = 1; i < kEvents; i++) job += ms[i];
  }
  mhmsw_stats& s = c->stats;
  s.npairs = c->npairs;
  s.cells = c->cells;
  s.reverse_cells = (int64_t)counters[0];
  s.fast_pairs = c->n_fast;
  s.plain_pairs = c->n_plain;
  s.rerun_pairs = c->h_errors[1];
  static_assert(kNumFwdClasses == MHMSW_NUM_CLASSES, "forward classes are the public ones");
  for (int k = 0; k < kNumFwdClasses; k++) {
    s.class_pairs[k] = c->class_pairs[k];
    s.class_cells[k] = c->class_cells[k];
    float cm = 0.f;
    if (c->class_launched[k] && c->last_job_phases)
      HIP_OR_FAIL(c, hipEventElapsedTime(&cm, c->cev[k][0], c->cev[k][1]));
    s.class_ms[k] = cm;
    const FastShape sh = k == kClassPlain ? FastShape{0, 0}
                                          : fast_shape(k < kClassFast64 ? 16 : 64,
                                                       k < kClassFast64 ? k + 1 : k - kClassFast64 + 1, true,
                                                       c->class_pairs[k], c->cus, c->class_max_rows[k]);
    s.class_lanes[k] = sh.lanes;
    s.class_regs[k] = sh.regs;
    s.class_tag_bits[k] = k == kClassPlain ? 0 : sh.lanes <= 8 && fold32_ok(c, k) ? 5 : 4;
  }
  s.prep_ms = ms[1];
  float wm = 0.f;
  if (c->last_job_windows && c->last_job_phases) HIP_OR_FAIL(c, hipEventElapsedTime(&wm, c->wev[0], c->wev[1]));
  s.window_ms = wm;
  s.window_bases = 0;
  if (c->last_job_windows)
    for (int64_t p = 0; p < c->npairs; p++) s.window_bases += c->lens[(size_t)(2 * p + 1)];
  s.packed_scoring = c->sp.fast_ok;
  s.device_bytes = device_bytes(c);
  s.fwd_fast_ms = ms[2];
  s.fwd_plain_ms = ms[3];
  s.rev_ms = ms[4] + ms[5];
  s.cigar_ms = ms[6];
  s.job_ms = job;
  s.cigar_reruns = 0;
  s.host_pack_ms = c->host_pack_ms;
  s.host_upload_ms = c->host_upload_ms;
  s.host_launch_ms = c->host_launch_ms;
  if (c->last_job_cigar) {
    int rc = settle_cigar(c);
    if (rc != MHMSW_OK) return rc;
  }
  c->have_results = true;
  if (c->h_errors[0] != 0 || c->h_errors[2] != 0)
    return fail(c, MHMSW_E_INTERNAL, "%d pairs: reverse pass did not reproduce the forward score; %d end points "
                "outside their pair", c->h_errors[0], c->h_errors[2]);
  return MHMSW_OK;
}

int mhmsw_query(mhmsw_ctx* c) {
  if (!c) return MHMSW_E_ARG;
  if (!c->in_flight) return c->have_results ? 1 : fail(c, MHMSW_E_STATE, "no batch submitted");
  hipError_t e = hipEventQuery(c->ev_done);
  if (e == hipErrorNotReady) return 0;
  if (e != hipSuccess) return fail(c, MHMSW_E_HIP, "hipEventQuery: %s", hipGetErrorString(e));
  int rc = complete(c);
  return rc == MHMSW_OK ? 1 : rc;
}

int mhmsw_wait(mhmsw_ctx* c) {
  if (!c) return MHMSW_E_ARG;
  if (!c->in_flight) return c->have_results ? MHMSW_OK : fail(c, MHMSW_E_STATE, "no batch submitted");
  HIP_OR_FAIL(c, hipSetDevice(c->device));
  HIP_OR_FAIL(c, hipEventSynchronize(c->ev_done));
  return complete(c);
}

int mhmsw_results(mhmsw_ctx* c, mhmsw_result* out, int64_t npairs) {
  if (!c || (!out && npairs > 0)) return MHMSW_E_ARG;
  if (c->in_flight) {
    int rc = mhmsw_wait(c);
    if (rc != MHMSW_OK) return rc;
  }
  if (!c->have_results) return fail(c, MHMSW_E_STATE, "no results: submit a batch first");
  if (npairs != c->npairs) return fail(c, MHMSW_E_ARG, "results requested for %lld pairs, batch has %lld",
                                       (long long)npairs, (long long)c->npairs);
  static_assert(sizeof(mhmsw_result) == sizeof(AlnResult), "result layout");
  if (npairs > 0 && c->last_job_tail_results) {
    std::memcpy(out, c->h_results.ptr, (size_t)npairs * sizeof(AlnResult));  // copied by the job
  } else if (npairs > 0) {
    HIP_OR_FAIL(c, d2h_sync(c, out, c->d_results.ptr, (size_t)npairs * sizeof(AlnResult)));
  }
  if (c->h_errors[0] != 0 || c->h_errors[2] != 0)
    return fail(c, MHMSW_E_INTERNAL, "%d pairs: reverse pass did not reproduce the forward score; %d end points "
                "outside their pair", c->h_errors[0], c->h_errors[2]);
  return MHMSW_OK;
}

int mhmsw_results_score2(mhmsw_ctx* c, int32_t* out, int64_t npairs) {
  if (!c || (!out && npairs > 0)) return MHMSW_E_ARG;
  if (c->in_flight) {
    int rc = mhmsw_wait(c);
    if (rc != MHMSW_OK) return rc;
  }
  if (!c->have_results) return fail(c, MHMSW_E_STATE, "no results: submit a batch first");
  if (!c->last_job_score2) return fail(c, MHMSW_E_STATE, "the last batch ran without score2 (mhmsw_set_score2)");
  if (npairs != c->npairs) return fail(c, MHMSW_E_ARG, "score2 requested for %lld pairs, batch has %lld",
                                       (long long)npairs, (long long)c->npairs);
  if (npairs > 0) HIP_OR_FAIL(c, d2h_sync(c, out, c->d_s2.ptr, (size_t)npairs * sizeof(int32_t)));
  return MHMSW_OK;
}

int mhmsw_set_cigar(mhmsw_ctx* c, int enable) {
  if (!c) return MHMSW_E_ARG;
  if (c->in_flight) return fail(c, MHMSW_E_STATE, "cigar changed while a batch is in flight");
  c->cigar = enable ? 1 : 0;
  return MHMSW_OK;
}

int mhmsw_results_cigar(mhmsw_ctx* c, int64_t npairs, int32_t* mismatches, int64_t* op_off, uint32_t* ops,
                        int64_t ops_cap, int64_t* total_ops) {
  if (!c) return MHMSW_E_ARG;
  if (c->in_flight) {
    int rc = mhmsw_wait(c);
    if (rc != MHMSW_OK) return rc;
  }
  if (!c->have_results) return fail(c, MHMSW_E_STATE, "no results: submit a batch first");
  if (!c->last_job_cigar) return fail(c, MHMSW_E_STATE, "the last batch ran without the CIGAR pass (mhmsw_set_cigar)");
  if (npairs != c->npairs) return fail(c, MHMSW_E_ARG, "CIGARs requested for %lld pairs, batch has %lld",
                                       (long long)npairs, (long long)c->npairs);
  std::vector<int32_t> n((size_t)npairs), mm((size_t)npairs);
  std::vector<int64_t> off((size_t)npairs);
  if (npairs > 0) {
    HIP_OR_FAIL(c, d2h_sync(c, n.data(), c->d_cig_nops.ptr, (size_t)npairs * 4));
    HIP_OR_FAIL(c, d2h_sync(c, mm.data(), c->d_cig_mism.ptr, (size_t)npairs * 4));
    HIP_OR_FAIL(c, d2h_sync(c, off.data(), c->d_cig_off.ptr, (size_t)npairs * 8));
  }
  int64_t total = 0;
  for (int64_t p = 0; p < npairs; p++) {
    if (n[(size_t)p] < 0)
      return fail(c, MHMSW_E_INTERNAL, "pair %lld: CIGAR not computed (%d)", (long long)p, n[(size_t)p]);
    total += n[(size_t)p];
  }
  if (total_ops) *total_ops = total;
  if (mismatches && npairs > 0) std::memcpy(mismatches, mm.data(), (size_t)npairs * 4);
  if (op_off) {
    int64_t acc = 0;
    for (int64_t p = 0; p < npairs; p++) {
      op_off[p] = acc;
      acc += n[(size_t)p];
    }
    op_off[npairs] = acc;
  }
  if (!ops) return MHMSW_OK;
  if (ops_cap < total) return fail(c, MHMSW_E_ARG, "ops buffer holds %lld ops, batch has %lld", (long long)ops_cap,
                                   (long long)total);
  const int64_t used = (int64_t)c->h_cig_used[0];
  std::vector<uint32_t> pool((size_t)std::max<int64_t>(used, 1));
  if (used > 0)
    HIP_OR_FAIL(c, d2h_sync(c, pool.data(), c->d_cig_ops.ptr, (size_t)used * 4));
  int64_t acc = 0;
  for (int64_t p = 0; p < npairs; p++) {
    const int32_t k = n[(size_t)p];
    if (k > 0) {
      if (off[(size_t)p] < 0 || off[(size_t)p] + k > used)
        return fail(c, MHMSW_E_INTERNAL, "pair %lld: CIGAR outside the op pool", (long long)p);
      std::memcpy(ops + acc, pool.data() + off[(size_t)p], (size_t)k * 4);
    }
    acc += k;
  }
  return MHMSW_OK;
}

int mhmsw_get_stats(mhmsw_ctx* c, mhmsw_stats* out) {
  if (!c || !out) return MHMSW_E_ARG;
  if (!c->have_results) return fail(c, MHMSW_E_STATE, "no completed batch");
  *out = c->stats;
  return MHMSW_OK;
}


// ----------------------------------------------------------------------------
// tuning
// ----------------------------------------------------------------------------
extern "C" int mhmsw_set_tuning(mhmsw_ctx* c, const mhmsw_tuning* t) {
  if (!c || !t) return MHMSW_E_ARG;
  if (c->in_flight) return fail(c, MHMSW_E_STATE, "tuning changed while a batch is in flight");
  if (t->blocks_per_cu < 1 || t->blocks_per_cu > 16 || t->g64_min_rows < 33 || t->g64_min_rows > 257 ||
      t->bin_shift < 3 || t->bin_shift > 7)
    return fail(c, MHMSW_E_ARG, "tuning out of range (blocks_per_cu 1..16, g64_min_rows 33..257, bin_shift 3..7)");
  c->tuning = *t;
  if (c->staged) reroute(c);
  return MHMSW_OK;
}

extern "C" int mhmsw_get_tuning(const mhmsw_ctx* c, mhmsw_tuning* t) {
  if (!c || !t) return MHMSW_E_ARG;
  *t = c->tuning;
  return MHMSW_OK;
}

static int time_job(mhmsw_ctx* c, int repeats, double* best_ms) {
  double best = 1e30;
  for (int r = 0; r < repeats; r++) {
    int rc = mhmsw_align_staged(c);
    if (rc == MHMSW_OK) rc = mhmsw_wait(c);
    if (rc != MHMSW_OK) return rc;
    best = std::min(best, c->stats.job_ms);
  }
  *best_ms = best;
  return MHMSW_OK;
}

extern "C" int mhmsw_autotune(mhmsw_ctx* c, int repeats, mhmsw_tuning* best_out) {
  if (!c) return MHMSW_E_ARG;
  if (!c->staged) return fail(c, MHMSW_E_STATE, "autotune needs a staged batch");
  if (repeats < 1) repeats = 1;
  const bool graphs = c->use_graphs;
  c->use_graphs = false;  // per-phase events of eager launches
  mhmsw_tuning best = c->tuning;
  double best_ms = 0;
  int rc = time_job(c, repeats, &best_ms);
  static const int kBlocks[] = {2, 4, 6, 8, 12};
  static const int kG64[] = {257, 225, 193, 161, 129};
  static const int kBins[] = {3, 4, 5, 6};
  for (int round = 0; rc == MHMSW_OK && round < 3; round++) {
    const int* vals = round == 0 ? kBlocks : round == 1 ? kG64 : kBins;
    const int nvals = round == 0 ? 5 : round == 1 ? 5 : 4;
    for (int i = 0; i < nvals && rc == MHMSW_OK; i++) {
      mhmsw_tuning t = best;
      if (round == 0) t.blocks_per_cu = vals[i];
      if (round == 1) t.g64_min_rows = vals[i];
      if (round == 2) t.bin_shift = vals[i];
      if (std::memcmp(&t, &best, sizeof(t)) == 0) continue;
      rc = mhmsw_set_tuning(c, &t);
      double ms = 0;
      if (rc == MHMSW_OK) rc = time_job(c, repeats, &ms);
      if (rc == MHMSW_OK && ms < best_ms * 0.995) {
        best = t;
        best_ms = ms;
      }
    }
  }
  c->use_graphs = graphs;
  if (rc != MHMSW_OK) return rc;
  rc = mhmsw_set_tuning(c, &best);
  if (rc != MHMSW_OK) return rc;
  if (best_out) *best_out = best;
  return MHMSW_OK;
}

// ----------------------------------------------------------------------------
// resident contigs
// ----------------------------------------------------------------------------
extern "C" int mhmsw_contigs_load(int device_id, int64_t nctgs, const char* seqs, const int64_t* off, const int32_t* len,
                                  mhmsw_contigs** out) {
  if (!out || nctgs < 0 || (nctgs > 0 && (!seqs || !off || !len))) return MHMSW_E_ARG;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return MHMSW_E_NODEVICE;
  if (device_id < 0 || device_id >= n) return MHMSW_E_ARG;
  if (hipSetDevice(device_id) != hipSuccess) return MHMSW_E_HIP;
  mhmsw_contigs* r = new (std::nothrow) mhmsw_contigs();
  if (!r) return MHMSW_E_NOMEM;
  r->device = device_id;
  r->n = nctgs;
  r->len.assign(len, len + nctgs);
  std::vector<int64_t> base((size_t)std::max<int64_t>(nctgs, 1));
  int64_t acc = 0;
  for (int64_t i = 0; i < nctgs; i++) {
    if (len[i] < 0) {
      delete r;
      return MHMSW_E_ARG;
    }
    base[(size_t)i] = acc;
    acc += ((int64_t)len[i] + 31) / 32 * 32;
  }
  const size_t n2 = (size_t)(acc / 16 + 4), nm = (size_t)(acc / 32 + 4);
  std::vector<uint32_t> h2(n2, 0), hN(nm, 0);
  {
    int nth = (int)std::min<unsigned>(16u, std::max(1u, std::thread::hardware_concurrency()));
    auto work = [&](int tid) {
      for (int64_t i = tid; i < nctgs; i += nth) {
        pack_seq(seqs + off[i], len[i], h2.data() + base[(size_t)i] / 16, hN.data() + base[(size_t)i] / 32);
      }
    };
    std::vector<std::thread> th;
    for (int t = 0; t < nth; t++) th.emplace_back(work, t);
    for (auto& t : th) t.join();
  }
  bool ok = hipMalloc((void**)&r->d2, n2 * 4) == hipSuccess && hipMalloc((void**)&r->dN, nm * 4) == hipSuccess &&
            hipMalloc((void**)&r->dbase, base.size() * 8) == hipSuccess &&
            hipMemcpy(r->d2, h2.data(), n2 * 4, hipMemcpyHostToDevice) == hipSuccess &&
            hipMemcpy(r->dN, hN.data(), nm * 4, hipMemcpyHostToDevice) == hipSuccess &&
            hipMemcpy(r->dbase, base.data(), base.size() * 8, hipMemcpyHostToDevice) == hipSuccess;
  if (!ok) {
    mhmsw_contigs_free(r);
    return MHMSW_E_HIP;
  }
  r->bytes = (int64_t)(n2 * 4 + nm * 4 + base.size() * 8);
  *out = r;
  return MHMSW_OK;
}

extern "C" int mhmsw_contigs_free(mhmsw_contigs* r) {
  if (!r) return MHMSW_E_ARG;
  (void)hipSetDevice(r->device);
  if (r->d2) (void)hipFree(r->d2);
  if (r->dN) (void)hipFree(r->dN);
  if (r->dbase) (void)hipFree(r->dbase);
  delete r;
  return MHMSW_OK;
}

extern "C" int64_t mhmsw_contigs_bytes(const mhmsw_contigs* r) { return r ? r->bytes : 0; }

// ----------------------------------------------------------------------------
// serving loop: continuous batching over several contexts
// ----------------------------------------------------------------------------
#include <chrono>
#include <condition_variable>
#include <deque>
#include <memory>
#include <mutex>
#include <unordered_map>

namespace {

double steady_now() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

struct Request {
  uint64_t id = 0;
  int64_t npairs = 0;
  PackedBatch packed;  // packed on the submitting thread
  int state = 0;  // 0 queued, 1 running, 2 done, < 0 failed
  int slot = -1;
  int64_t est = 0;
  std::vector<mhmsw_result> results;
  // second-best scores and CIGARs, when the server computes them
  bool has_s2 = false, has_cigar = false;
  std::vector<int32_t> s2, mism;
  std::vector<int64_t> op_off;
  std::vector<uint32_t> ops;
  mhmsw_req_times times{};
  std::string err;
};


// Requests queued together run as one job: their packed batches are
// uploaded one after another (upload_parts) and the results split afterwards.
// Pairs are independent, so a request's results do not depend on its company.
struct Job {
  std::vector<std::shared_ptr<Request>> reqs;
  std::vector<int64_t> first;  // first pair of each request in the job
  int64_t npairs = 0;
  int64_t total = 0;  // padded bases of the job's layout
  int64_t est = 0;    // device bytes its buffers may take (job_alloc_bound)
  bool s2 = false, cigar = false;  // the outputs this job computes
  std::vector<mhmsw_result> results;
};

// the second-best scores and CIGARs of a finished job, split by request
int split_extras(mhmsw_ctx* c, Job& j) {
  std::vector<int32_t> s2, mism;
  std::vector<int64_t> off;
  std::vector<uint32_t> ops;
  const size_t n = (size_t)std::max<int64_t>(j.npairs, 1);
  if (j.s2) {
    s2.resize(n);
    const int rc = mhmsw_results_score2(c, s2.data(), j.npairs);
    if (rc != MHMSW_OK) return rc;
  }
  if (j.cigar) {
    mism.resize(n);
    off.resize(n + 1);
    int64_t total = 0;
    int rc = mhmsw_results_cigar(c, j.npairs, mism.data(), off.data(), nullptr, 0, &total);
    if (rc != MHMSW_OK) return rc;
    ops.resize((size_t)std::max<int64_t>(total, 1));
    rc = mhmsw_results_cigar(c, j.npairs, nullptr, nullptr, ops.data(), total, &total);
    if (rc != MHMSW_OK) return rc;
  }
  for (size_t i = 0; i < j.reqs.size(); i++) {
    Request& r = *j.reqs[i];
    const int64_t a = j.first[i], b = a + r.npairs;
    r.has_s2 = j.s2;
    r.has_cigar = j.cigar;
    if (j.s2) r.s2.assign(s2.begin() + a, s2.begin() + b);
    if (j.cigar) {
      r.mism.assign(mism.begin() + a, mism.begin() + b);
      r.op_off.resize((size_t)r.npairs + 1);
      for (int64_t p = 0; p <= r.npairs; p++) r.op_off[(size_t)p] = off[(size_t)(a + p)] - off[(size_t)a];
      r.ops.assign(ops.begin() + off[(size_t)a], ops.begin() + off[(size_t)b]);
    }
  }
  return MHMSW_OK;
}

constexpr int64_t kMergeMaxPairs = (int64_t)1 << 20;
// Buffers every slot holds from the start, so that jobs up to this size never
// reallocate while others run (hipFree waits for the whole device).
constexpr int64_t kSlotReservePairs = (int64_t)1 << 18;
constexpr int64_t kSlotReserveBases = kSlotReservePairs * 1280;

}  // namespace

struct mhmsw_server {
  int device = 0;
  PackParams pp;
  const mhmsw_contigs* contigs = nullptr;
  std::vector<mhmsw_ctx*> slots;
  std::vector<std::shared_ptr<Job>> running;  // per slot: the job handed to its worker
  std::vector<char> taken;                    // per slot: the worker has started that job
  std::vector<std::thread> workers;           // one per slot
  int64_t budget = 0, used = 0;
  std::mutex mu;
  std::condition_variable cv_done, cv_work, cv_slot;
  std::deque<std::shared_ptr<Request>> queue;
  std::unordered_map<uint64_t, std::shared_ptr<Request>> reqs;
  std::unordered_map<uint64_t, std::shared_ptr<Request>> extras;  // results taken, extras not yet
  int score2 = 0, cigar = 0;  // outputs of jobs admitted from now on
  uint64_t next_id = 1;
  bool stop = false;
  std::thread th;
  std::string err;
};

// One worker per slot runs its jobs start to end: upload, launch, wait for the
// device, copy the results out and split them by request.  A slow step on one
// slot (a large upload, a buffer that grows) never holds up another slot's
// completion.
static void slot_worker(mhmsw_server* s, size_t k) {
  (void)hipSetDevice(s->device);
  mhmsw_ctx* c = s->slots[k];
  std::unique_lock<std::mutex> lk(s->mu);
  for (;;) {
    s->cv_slot.wait(lk, [&] { return s->stop || (s->running[k] && !s->taken[k]); });
    if (s->stop) break;
    std::shared_ptr<Job> j = s->running[k];
    s->taken[k] = 1;
    lk.unlock();
    // the slot's context belongs to this thread while it runs the job
    mhmsw_set_score2(c, j->s2 ? 1 : 0);
    mhmsw_set_cigar(c, j->cigar ? 1 : 0);
    std::vector<const PackedBatch*> parts;
    parts.reserve(j->reqs.size());
    for (auto& r : j->reqs) parts.push_back(&r->packed);
    int rc = upload_parts(c, parts.data(), parts.size());
    if (rc == MHMSW_OK) rc = mhmsw_align_staged(c);
    const double launched = steady_now();
    // block on the job's completion event (this thread has nothing else to do)
    const int q = rc == MHMSW_OK ? (mhmsw_wait(c) == MHMSW_OK ? 1 : MHMSW_E_HIP) : rc;
    const double fin = steady_now();
    mhmsw_stats st{};
    if (q == 1) {
      j->results.resize((size_t)std::max<int64_t>(j->npairs, 1));
      rc = mhmsw_results(c, j->results.data(), j->npairs);
      if (rc == MHMSW_OK && (j->s2 || j->cigar)) rc = split_extras(c, *j);
      if (rc == MHMSW_OK) {
        mhmsw_get_stats(c, &st);
        for (size_t i = 0; i < j->reqs.size(); i++) {
          Request& r = *j->reqs[i];
          r.results.assign(j->results.begin() + j->first[i], j->results.begin() + j->first[i] + r.npairs);
        }
      }
    }
    lk.lock();
    const double now = steady_now();
    for (auto& rp : j->reqs) {
      Request& r = *rp;
      r.times.launched_s = launched;
      r.times.finished_s = fin;
      r.times.done_s = now;
      r.times.device_ms = st.job_ms;
      r.times.job_est_bytes = (double)j->est;
      if (q < 0 || rc != MHMSW_OK) {
        r.state = q < 0 ? q : rc;
        r.err = mhmsw_last_error(c);
      } else {
        r.state = 2;
      }
      r.packed = PackedBatch();  // release the host copy
    }
    s->running[k].reset();
    s->taken[k] = 0;
    s->used -= j->est;
    s->cv_done.notify_all();
    s->cv_work.notify_all();
  }
}

// The dispatcher admits queued requests, in arrival order, to free slots;
// requests waiting together share a job.
static void server_loop(mhmsw_server* s) {
  std::unique_lock<std::mutex> lk(s->mu);
  while (!s->stop) {
    bool admitted = false;
    while (!s->queue.empty()) {
      int freek = -1;
      for (size_t k = 0; k < s->slots.size() && freek < 0; k++)
        if (!s->running[k]) freek = (int)k;
      if (freek < 0) break;
      auto j = std::make_shared<Job>();
      j->s2 = s->score2 != 0;
      j->cigar = s->cigar != 0;
      while (!s->queue.empty()) {
        const std::shared_ptr<Request>& r = s->queue.front();
        const bool first = j->reqs.empty();
        if (!first && j->npairs + r->npairs > kMergeMaxPairs) break;
        // wait until the job with this request fits the memory not taken by running jobs
        const int64_t est = job_alloc_bound(j->npairs + r->npairs, j->total + r->packed.total, true, j->cigar);
        if (s->used + est > s->budget) break;
        j->first.push_back(j->npairs);
        j->npairs += r->npairs;
        j->total += r->packed.total;
        j->est = est;
        j->reqs.push_back(r);
        s->queue.pop_front();
      }
      if (j->reqs.empty()) break;
      const double now = steady_now();
      for (auto& r : j->reqs) {
        r->state = 1;
        r->slot = freek;
        r->times.admit_s = now;
      }
      s->running[(size_t)freek] = j;
      s->taken[(size_t)freek] = 0;
      s->used += j->est;
      admitted = true;
    }
    if (admitted) s->cv_slot.notify_all();
    // woken by a submission, a finished job or stop (each changes the state
    // checked above under this lock, so no wake-up is lost)
    s->cv_work.wait(lk);
  }
}

extern "C" {

double mhmsw_server_clock(void) { return steady_now(); }

const char* mhmsw_server_last_error(const mhmsw_server* s) { return s ? s->err.c_str() : "null server"; }

int mhmsw_server_create(int device_id, const mhmsw_scoring* scoring, const mhmsw_contigs* contigs, int slots,
                        int64_t hbm_budget, mhmsw_server** out) {
  if (!out || !contigs || slots < 1 || slots > 16) return MHMSW_E_ARG;
  *out = nullptr;
  if (contigs->device != device_id) return MHMSW_E_ARG;
  mhmsw_server* s = new (std::nothrow) mhmsw_server();
  if (!s) return MHMSW_E_NOMEM;
  s->device = device_id;
  s->contigs = contigs;
  for (int k = 0; k < slots; k++) {
    mhmsw_ctx* c = nullptr;
    int rc = mhmsw_create(device_id, &c);
    if (rc == MHMSW_OK && scoring) rc = mhmsw_set_scoring(c, scoring);
    if (rc == MHMSW_OK) rc = mhmsw_set_graphs(c, 0);  // every request has its own geometry
    if (rc == MHMSW_OK && (ensure_job_buffers(c, kSlotReservePairs, kSlotReserveBases) != hipSuccess ||
                           c->h_stage.ensure((size_t)kSlotReservePairs * 128, true) != hipSuccess ||
                           c->h_results.ensure((size_t)kSlotReservePairs * sizeof(AlnResult), true) != hipSuccess))
      rc = fail(c, MHMSW_E_NOMEM, "cannot reserve the slot's buffers");
    if (c) c->tail_results = true;
    // first use of the slot's stream now, not inside the first request's latency
    if (rc == MHMSW_OK && (hipMemsetAsync(c->d_errors.ptr, 0, 4 * sizeof(int32_t), c->stream) != hipSuccess ||
                           hipStreamSynchronize(c->stream) != hipSuccess))
      rc = fail(c, MHMSW_E_HIP, "slot stream set-up failed");
    if (rc != MHMSW_OK) {
      if (c) mhmsw_destroy(c);
      for (auto* x : s->slots) mhmsw_destroy(x);
      delete s;
      return rc;
    }
    s->slots.push_back(c);
  }
  s->running.resize((size_t)slots);
  s->taken.assign((size_t)slots, 0);
  s->pp = pack_params(s->slots[0]);
  if (hbm_budget <= 0) {
    size_t tot = 0;
    mhmsw_device_total_mem(device_id, &tot);
    hbm_budget = (int64_t)(tot * 0.8) - contigs->bytes;
  }
  s->budget = hbm_budget;
  s->th = std::thread(server_loop, s);
  for (size_t k = 0; k < s->slots.size(); k++) s->workers.emplace_back(slot_worker, s, k);
  *out = s;
  return MHMSW_OK;
}

int mhmsw_server_submit(mhmsw_server* s, int64_t npairs, const char* reads, const int64_t* read_off,
                        const int32_t* read_len, const int32_t* ctg_id, const int32_t* win_off, const int32_t* win_len,
                        uint64_t* req_id) {
  if (!s || !req_id || npairs < 0 || (npairs > 0 && (!reads || !read_off || !read_len || !ctg_id || !win_off || !win_len)))
    return MHMSW_E_ARG;
  auto r = std::make_shared<Request>();
  r->times.submit_s = steady_now();
  r->npairs = npairs;
  // pack here, on the submitting thread, off the dispatcher's path
  StageInput in;
  in.npairs = npairs;
  in.reads = reads;
  in.read_off = read_off;
  in.read_len = read_len;
  in.res = s->contigs;
  in.ctg_id = ctg_id;
  in.win_off = win_off;
  in.win_len = win_len;
  int rc = pack_batch(s->pp, in, r->packed, 4);
  if (rc != MHMSW_OK) {
    std::lock_guard<std::mutex> g(s->mu);
    s->err = r->packed.err;
    return rc;
  }
  {
    std::lock_guard<std::mutex> g(s->mu);
    // a request that alone would not fit the budget could never be admitted
    r->est = job_alloc_bound(npairs, r->packed.total, true, s->cigar != 0);
    if (r->est > s->budget) {
      char m[256];
      snprintf(m, sizeof(m), "request of %lld pairs may need %lld bytes of device memory, more than the server's HBM "
               "budget of %lld", (long long)npairs, (long long)r->est, (long long)s->budget);
      s->err = m;
      return MHMSW_E_ARG;
    }
    r->id = s->next_id++;
    s->reqs[r->id] = r;
    s->queue.push_back(r);
    *req_id = r->id;
  }
  s->cv_work.notify_one();
  return MHMSW_OK;
}

int mhmsw_server_status(mhmsw_server* s, uint64_t req_id) {
  if (!s) return MHMSW_E_ARG;
  std::lock_guard<std::mutex> g(s->mu);
  auto it = s->reqs.find(req_id);
  if (it == s->reqs.end()) return MHMSW_E_ARG;
  return it->second->state;
}

int mhmsw_server_wait(mhmsw_server* s, uint64_t req_id) {
  if (!s) return MHMSW_E_ARG;
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->reqs.find(req_id);
  if (it == s->reqs.end()) return MHMSW_E_ARG;
  std::shared_ptr<Request> r = it->second;
  s->cv_done.wait(lk, [&] { return r->state == 2 || r->state < 0; });
  if (r->state < 0) {
    s->err = r->err;
    return r->state;
  }
  return MHMSW_OK;
}

int mhmsw_server_results(mhmsw_server* s, uint64_t req_id, mhmsw_result* out, int64_t npairs, mhmsw_req_times* times) {
  if (!s) return MHMSW_E_ARG;
  int rc = mhmsw_server_wait(s, req_id);
  std::lock_guard<std::mutex> g(s->mu);
  auto it = s->reqs.find(req_id);
  if (it == s->reqs.end()) return MHMSW_E_ARG;
  std::shared_ptr<Request> r = it->second;
  if (times) *times = r->times;
  if (rc != MHMSW_OK) {
    s->reqs.erase(it);
    return rc;
  }
  if (npairs != r->npairs || (npairs > 0 && !out)) {
    s->err = "results requested for a different number of pairs";
    return MHMSW_E_ARG;
  }
  if (npairs > 0) std::memcpy(out, r->results.data(), (size_t)npairs * sizeof(mhmsw_result));
  r->results.clear();
  if (r->has_s2 || r->has_cigar) s->extras[req_id] = r;
  s->reqs.erase(it);
  return MHMSW_OK;
}

int mhmsw_server_set_outputs(mhmsw_server* s, int score2, int cigar) {
  if (!s) return MHMSW_E_ARG;
  std::lock_guard<std::mutex> g(s->mu);
  s->score2 = score2 != 0;
  s->cigar = cigar != 0;
  return MHMSW_OK;
}

int mhmsw_server_results_extra(mhmsw_server* s, uint64_t req_id, int64_t npairs, int32_t* score2,
                               int32_t* mismatches, int64_t* op_off, uint32_t* ops, int64_t ops_cap,
                               int64_t* total_ops) {
  if (!s) return MHMSW_E_ARG;
  std::lock_guard<std::mutex> g(s->mu);
  auto it = s->extras.find(req_id);
  if (it == s->extras.end()) {
    s->err = "no second-best scores or CIGARs held for this request (results not taken, or outputs off)";
    return MHMSW_E_STATE;
  }
  Request& r = *it->second;
  if (npairs != r.npairs) {
    s->err = "extras requested for a different number of pairs";
    return MHMSW_E_ARG;
  }
  if (score2) {
    if (!r.has_s2) {
      s->err = "the request ran without second-best scores";
      return MHMSW_E_STATE;
    }
    if (npairs > 0) std::memcpy(score2, r.s2.data(), (size_t)npairs * sizeof(int32_t));
  }
  if (mismatches || op_off || ops || total_ops) {
    if (!r.has_cigar) {
      s->err = "the request ran without CIGARs";
      return MHMSW_E_STATE;
    }
    const int64_t total = (int64_t)r.ops.size();
    if (total_ops) *total_ops = total;
    if (mismatches && npairs > 0) std::memcpy(mismatches, r.mism.data(), (size_t)npairs * sizeof(int32_t));
    if (op_off) std::memcpy(op_off, r.op_off.data(), (size_t)(npairs + 1) * sizeof(int64_t));
    if (ops) {
      if (ops_cap < total) {
        s->err = "ops buffer smaller than the request's total ops";
        return MHMSW_E_ARG;
      }
      if (total > 0) std::memcpy(ops, r.ops.data(), (size_t)total * sizeof(uint32_t));
    }
  }
  // released once the ops (or, without CIGARs, the scores) have been taken
  if (ops || !r.has_cigar) s->extras.erase(it);
  return MHMSW_OK;
}

int mhmsw_server_destroy(mhmsw_server* s) {
  if (!s) return MHMSW_E_ARG;
  {
    std::lock_guard<std::mutex> g(s->mu);
    s->stop = true;
  }
  s->cv_work.notify_all();
  s->cv_slot.notify_all();
  if (s->th.joinable()) s->th.join();
  for (auto& w : s->workers)
    if (w.joinable()) w.join();
  for (auto* c : s->slots) mhmsw_destroy(c);
  delete s;
  return MHMSW_OK;
}

}  // extern "C"
