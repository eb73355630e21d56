// mhmsw_submit_noalloc: the batch entry point for callers that cannot
// allocate device memory (code under HIP graph capture, tight serving loops,
// harnesses that forbid hipMalloc).  No counterpart in the reference.
//
// Inputs and outputs are device memory the caller owns.  Everything else the
// job needs (nibble codes, the per-sequence layout,
// work lists, scheduling state, per-pair end points, counters) lives in one
// file-scope device array, `scratch`, of fixed size: MHMSW_NOALLOC_MAX_PAIRS
// pairs and MHMSW_NOALLOC_MAX_BASES padded bases.  Each GPU holds its own copy
// of the array (about 28 MiB) from the load of this code object until the
// process ends.  Host code never passes the array's name to a kernel: the
// device address of each GPU's copy comes from hipGetSymbolAddress, taken
// once per device with that device current (noalloc_setup, called by
// mhmsw_create and again by the entry point's first call on a device).
//
// Every call on a GPU shares that GPU's array and leaves its contents behind;
// the job clears what must start at zero (error words, counters, scheduling
// state, status) with its own kernels, never with hipMemsetAsync, so the
// whole call may be captured into a graph.  Calls that overlap on one GPU (two
// streams, or two host threads) race on the array.
//
// The job is the one mhmsw_submit runs (sw_kernels.hip), except that the
// host does not know the pairs' lengths: the layout and packing run on the
// device, and every work class is launched (a class without work exits at
// once), about 55 launches.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <mutex>

#include "../../include/mhmsw.h"
#include "ctx_info.h"
#include "sw_common.h"
#include "sw_launch.h"

namespace mhmsw {
namespace {

constexpr int64_t kNaMaxPairs = MHMSW_NOALLOC_MAX_PAIRS;
constexpr int64_t kNaMaxBases = MHMSW_NOALLOC_MAX_BASES;
static_assert(kNaMaxBases % 32 == 0, "the forward region stays a multiple of 32 bases");
// code region: front pad, the sequences, tail pad (reverse tasks read it backwards)
constexpr int64_t kNaTotal = kPadBases + kNaMaxBases + kPadBases;

constexpr size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// byte offsets into the scratch array
constexpr size_t kOffCodes = 0;
constexpr size_t kCodesBytes = align256((size_t)(kNaTotal / 8 + 64) * 4);  // + 64 dwords of read slack
constexpr size_t kOffBase = kOffCodes + kCodesBytes;
constexpr size_t kOffLen = kOffBase + align256((size_t)(2 * kNaMaxPairs) * 8);
constexpr size_t kOffBucket = kOffLen + align256((size_t)(2 * kNaMaxPairs) * 4);
constexpr size_t kOffRank = kOffBucket + align256((size_t)kNaMaxPairs * 4);
constexpr size_t kOffFwd = kOffRank + align256((size_t)kNaMaxPairs * 4);
constexpr size_t kOffRev = kOffFwd + align256((size_t)kNaMaxPairs * sizeof(PairEnd));
constexpr size_t kOffTasks = kOffRev + align256((size_t)kNaMaxPairs * sizeof(PairEnd));
constexpr size_t kOffSched = kOffTasks + align256((size_t)(kNaMaxPairs + 64 * kNumClasses + 64) * sizeof(SwTask));
constexpr size_t kOffErrors = kOffSched + align256(kNumPasses * sizeof(PassSched));
constexpr size_t kOffCounters = kOffErrors + 256;
constexpr size_t kOffState = kOffCounters + 256;
constexpr size_t kScratchBytes = kOffState + 256;

// state words: [0] status (MHMSW_OK / MHMSW_E_ARG), [1] first invalid pair
// (-1: none), [2] reason (1 invalid lengths, 2 more than the maximum bases),
// [4..5] end of the last sequence (int64, bases)
constexpr int kStateWords = 8;

}  // namespace

// The scratch array.  Its name is a host-side stand-in in host code: only
// the address from hipGetSymbolAddress reaches a kernel.
static __device__ __align__(256) uint8_t scratch[kScratchBytes];

namespace {

// SSW's base translation (src/ssw/ssw.cpp:13-26): A/a/U/u 0, C/c 1, G/g 2,
// T/t 3, anything else N (4).  OR 0x20 folds upper case onto lower case and
// maps no other byte onto these five letters.
__device__ __forceinline__ uint32_t na_code(uint32_t ch) {
  const uint32_t c = ch | 0x20u;
  uint32_t k = (uint32_t)kCodeN;
  k = c == 't' ? 3u : k;
  k = c == 'g' ? 2u : k;
  k = c == 'c' ? 1u : k;
  k = (c == 'a' || c == 'u') ? 0u : k;
  return k;
}

__device__ __forceinline__ int64_t span32(int32_t len) { return ((int64_t)len + kSeqAlign - 1) / kSeqAlign * kSeqAlign; }

// One workgroup: validates every pair, lays the sequences out as the host
// path does (read of pair p = sequence 2p, window = 2p+1, each on a 32-base
// boundary after the front pad) and records the status.  A batch with an
// invalid pair, or more padded bases than the array holds, gets empty
// sequences everywhere (every result is then the score-0 sentinel).
constexpr int kLayoutThreads = 1024;

__global__ __launch_bounds__(kLayoutThreads) void na_layout_kernel(int64_t npairs, const int32_t* __restrict__ read_len,
                                                                   const int32_t* __restrict__ ctg_len,
                                                                   int64_t* __restrict__ seq_base,
                                                                   int32_t* __restrict__ seq_len, int32_t* state) {
  __shared__ int64_t sums[kLayoutThreads];
  __shared__ int32_t first_bad;
  const int t = threadIdx.x;
  if (t == 0) first_bad = 0x7FFFFFFF;
  __syncthreads();
  const int64_t chunk = (npairs + kLayoutThreads - 1) / kLayoutThreads;
  const int64_t p0 = std::min<int64_t>(npairs, chunk * t), p1 = std::min<int64_t>(npairs, p0 + chunk);
  int64_t acc = 0;
  int32_t bad = 0x7FFFFFFF;
  for (int64_t p = p0; p < p1; p++) {
    const int32_t rl = read_len[p], cl = ctg_len[p];
    const bool ok = rl >= 0 && cl >= 0 && rl < (1 << 20) && cl < (1 << 20) && (rl <= kPlainMaxCols || cl <= kPlainMaxCols);
    if (!ok && bad == 0x7FFFFFFF) bad = (int32_t)p;
    acc += span32(ok ? rl : 0) + span32(ok ? cl : 0);
  }
  if (bad != 0x7FFFFFFF) atomicMin(&first_bad, bad);
  sums[t] = acc;
  __syncthreads();
  // exclusive scan of the per-thread sums (Hillis-Steele over 1024 entries)
  for (int o = 1; o < kLayoutThreads; o <<= 1) {
    const int64_t v = t >= o ? sums[t - o] : 0;
    __syncthreads();
    sums[t] += v;
    __syncthreads();
  }
  const int64_t total = sums[kLayoutThreads - 1];
  const int64_t excl = sums[t] - acc;
  const bool too_many = total > kNaMaxBases;
  const bool fail = too_many || first_bad != 0x7FFFFFFF;
  int64_t b = kPadBases + excl;
  for (int64_t p = p0; p < p1; p++) {
    int32_t rl = read_len[p], cl = ctg_len[p];
    if (fail) {
      seq_base[2 * p] = kPadBases;
      seq_base[2 * p + 1] = kPadBases;
      seq_len[2 * p] = 0;
      seq_len[2 * p + 1] = 0;
      continue;
    }
    seq_base[2 * p] = b;
    seq_len[2 * p] = rl;
    b += span32(rl);
    seq_base[2 * p + 1] = b;
    seq_len[2 * p + 1] = cl;
    b += span32(cl);
  }
  if (t == 0) {
    state[0] = fail ? MHMSW_E_ARG : MHMSW_OK;
    state[1] = first_bad != 0x7FFFFFFF ? first_bad : -1;
    state[2] = first_bad != 0x7FFFFFFF ? 1 : (too_many ? 2 : 0);
    const int64_t end = kPadBases + (fail ? 0 : total);
    reinterpret_cast<int64_t*>(state + 4)[0] = end;
  }
}

// ASCII -> nibble codes of both sequences of every pair, one wavefront per
// pair; nibbles past a sequence's length, up to its 32-base boundary, are 0.
// Block 0 also zeroes the front pad and the pad after the last sequence.
__global__ __launch_bounds__(256) void na_pack_kernel(int64_t npairs, const char* __restrict__ reads,
                                                      const int64_t* __restrict__ read_off,
                                                      const char* __restrict__ ctgs, const int64_t* __restrict__ ctg_off,
                                                      const int64_t* __restrict__ seq_base,
                                                      const int32_t* __restrict__ seq_len, uint32_t* __restrict__ codes,
                                                      const int32_t* state) {
  const int lane = threadIdx.x & 63;
  if (blockIdx.x == 0) {
    const int64_t end = reinterpret_cast<const int64_t*>(state + 4)[0];  // a multiple of 32
    for (int i = threadIdx.x; i < (int)(kPadBases / 8); i += blockDim.x) {
      codes[i] = 0;
      codes[end / 8 + i] = 0;  // end + kPadBases <= kNaTotal
    }
  }
  const int64_t nwaves = (int64_t)gridDim.x * (blockDim.x / 64);
  for (int64_t p = (int64_t)blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64; p < npairs; p += nwaves) {
#pragma unroll
    for (int side = 0; side < 2; side++) {
      const int64_t s = 2 * p + side;
      const int32_t len = seq_len[s];
      if (len <= 0 || seq_base[s] + span32(len) > kPadBases + kNaMaxBases) continue;  // the layout bounds both
      const char* src = side ? ctgs + ctg_off[p] : reads + read_off[p];
      uint32_t* dst = codes + (seq_base[s] >> 3);
      const int32_t ndw = (int32_t)(span32(len) >> 3);
      for (int32_t w = lane; w < ndw; w += 64) {
        uint32_t out = 0;
#pragma unroll
        for (int k = 0; k < 8; k++) {
          const int32_t pos = 8 * w + k;
          const uint32_t code = pos < len ? na_code((uint32_t)(uint8_t)src[pos]) : 0u;
          out |= code << (4 * k);
        }
        dst[w] = out;
      }
    }
  }
}

// The batch's status for the caller: the layout's, or MHMSW_E_INTERNAL when
// the reverse pass did not reproduce a forward score (never expected).
__global__ void na_status_kernel(const int32_t* state, const int32_t* errors, int32_t* status) {
  if (threadIdx.x == 0) {
    int32_t st = state[0];
    if (st == MHMSW_OK && (errors[0] != 0 || errors[2] != 0)) st = MHMSW_E_INTERNAL;
    status[0] = st;
    status[1] = state[1];
  }
}

std::mutex g_mu;
constexpr int kMaxDevices = 64;
void* g_ptr[kMaxDevices] = {};

}  // namespace

int noalloc_setup(int device) {
  if (device < 0 || device >= kMaxDevices) return MHMSW_E_ARG;
  std::lock_guard<std::mutex> lk(g_mu);
  if (g_ptr[device]) return MHMSW_OK;
  int cur = 0;
  if (hipGetDevice(&cur) != hipSuccess) return MHMSW_E_HIP;
  if (cur != device && hipSetDevice(device) != hipSuccess) return MHMSW_E_HIP;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (cur != device) (void)hipSetDevice(cur);
  if (e != hipSuccess || !p) return MHMSW_E_HIP;
  g_ptr[device] = p;
  return MHMSW_OK;
}

}  // namespace mhmsw

using namespace mhmsw;

extern "C" int mhmsw_submit_noalloc(mhmsw_ctx* ctx, int64_t npairs, const char* reads, const int64_t* read_off,
                                    const int32_t* read_len, const char* ctgs, const int64_t* ctg_off,
                                    const int32_t* ctg_len, mhmsw_result* results, int32_t* status, void* stream) {
  CtxInfo ci;
  if (!ctx_info(ctx, &ci)) return MHMSW_E_ARG;
  char msg[256];
  if (npairs < 0 || npairs > kNaMaxPairs) {
    snprintf(msg, sizeof(msg), "mhmsw_submit_noalloc: %lld pairs; the scratch holds at most %lld",
             (long long)npairs, (long long)kNaMaxPairs);
    return ctx_fail(ctx, MHMSW_E_ARG, msg);
  }
  if (!status || (npairs > 0 && (!reads || !read_off || !read_len || !ctgs || !ctg_off || !ctg_len || !results)))
    return ctx_fail(ctx, MHMSW_E_ARG, "mhmsw_submit_noalloc: null pointer argument");
  if (hipSetDevice(ci.device) != hipSuccess) return ctx_fail(ctx, MHMSW_E_HIP, "mhmsw_submit_noalloc: hipSetDevice failed");
  if (noalloc_setup(ci.device) != MHMSW_OK)
    return ctx_fail(ctx, MHMSW_E_HIP, "mhmsw_submit_noalloc: hipGetSymbolAddress of the scratch array failed");
  uint8_t* base = static_cast<uint8_t*>(g_ptr[ci.device]);
  hipStream_t st = stream ? static_cast<hipStream_t>(stream) : ci.stream;

  BatchView b;
  std::memset(&b, 0, sizeof(b));
  b.npairs = npairs;
  b.total = kNaTotal;
  b.seq_base = reinterpret_cast<const int64_t*>(base + kOffBase);
  b.seq_len = reinterpret_cast<const int32_t*>(base + kOffLen);
  b.codes = reinterpret_cast<uint32_t*>(base + kOffCodes);
  b.fwd = reinterpret_cast<PairEnd*>(base + kOffFwd);
  b.rev = reinterpret_cast<PairEnd*>(base + kOffRev);
  b.sched = reinterpret_cast<PassSched*>(base + kOffSched);
  b.tasks = reinterpret_cast<SwTask*>(base + kOffTasks);
  b.bucket = reinterpret_cast<int32_t*>(base + kOffBucket);
  b.rank = reinterpret_cast<int32_t*>(base + kOffRank);
  b.results = reinterpret_cast<AlnResult*>(results);
  b.errors = reinterpret_cast<int32_t*>(base + kOffErrors);
  b.counters = reinterpret_cast<unsigned long long*>(base + kOffCounters);
  b.sc = ci.sc;
  b.mode = ci.mode;
  b.g64_min_rows = ci.tuning.g64_min_rows;
  b.bin_shift = ci.tuning.bin_shift;
  b.score2 = 0;
  // one launch runs every band class of a batch of up to 65536 pairs, and it
  // holds the classes up to P = 8 (|l - k| <= 15)
  b.band_max_d = npairs <= 65536 ? 15 : kBandMaxD;
  int32_t* state = reinterpret_cast<int32_t*>(base + kOffState);
  static_assert(kStateWords * 4 <= 256, "state words fit their slot");
  static_assert(sizeof(mhmsw_result) == sizeof(AlnResult), "result layout");

  auto launched = [&](const char* what) -> int {
    const hipError_t e = hipGetLastError();
    if (e == hipSuccess) return MHMSW_OK;
    snprintf(msg, sizeof(msg), "mhmsw_submit_noalloc: launch of %s failed: %s", what, hipGetErrorString(e));
    return ctx_fail(ctx, MHMSW_E_HIP, msg);
  };
#define NA_CHECK(expr, what)                                           \
  do {                                                                 \
    const hipError_t _e = (expr);                                      \
    if (_e != hipSuccess) {                                            \
      snprintf(msg, sizeof(msg), "mhmsw_submit_noalloc: %s failed: %s", what, hipGetErrorString(_e)); \
      return ctx_fail(ctx, MHMSW_E_HIP, msg);                          \
    }                                                                  \
  } while (0)

  NA_CHECK(launch_clear_job(b, st), "clear");
  hipLaunchKernelGGL(na_layout_kernel, dim3(1), dim3(kLayoutThreads), 0, st, npairs, read_len, ctg_len,
                     const_cast<int64_t*>(b.seq_base), const_cast<int32_t*>(b.seq_len), state);
  if (int rc = launched("layout")) return rc;
  {
    const int64_t blocks = std::max<int64_t>(1, std::min<int64_t>((npairs + 3) / 4, (int64_t)ci.cus * 16));
    hipLaunchKernelGGL(na_pack_kernel, dim3((unsigned)blocks), dim3(256), 0, st, npairs, reads, read_off, ctgs, ctg_off,
                       b.seq_base, b.seq_len, b.codes, state);
    if (int rc = launched("pack")) return rc;
  }
  if (npairs > 0) {
    const int full = std::max(1, ci.cus * ci.tuning.blocks_per_cu);
    const int plain_grid = (int)std::max<int64_t>(1, std::min<int64_t>((int64_t)ci.cus * 16, npairs));
    const BatchView bf = with_pass(b, kPassFwd), br = with_pass(b, kPassRerun), bv = with_pass(b, kPassRev);
    NA_CHECK(launch_schedule(b, kPassFwd, st), "forward schedule");
    for (int P = kMaxP; P >= 1; P--) NA_CHECK(launch_fast(bf, fast_shape(64, P, true, npairs, ci.cus, INT32_MAX), full, st), "forward pass");
    for (int P = kMaxP; P >= 1; P--) NA_CHECK(launch_fast(bf, fast_shape(16, P, true, npairs, ci.cus, INT32_MAX), full, st), "forward pass");
    NA_CHECK(launch_plain(bf, plain_grid, kPlainMaxCols, st), "forward pass (32-bit)");
    NA_CHECK(launch_schedule(b, kPassRerun, st), "rerun schedule");
    NA_CHECK(launch_plain(br, plain_grid, kPlainMaxCols, st), "32-bit rerun");
    NA_CHECK(launch_schedule(b, kPassRev, st), "reverse schedule");
    for (int P = kMaxP; P >= 1; P--) NA_CHECK(launch_fast(bv, fast_shape(64, P, false, npairs, ci.cus, INT32_MAX), full, st), "reverse pass");
    for (int P = kMaxP; P >= 1; P--) NA_CHECK(launch_fast(bv, fast_shape(16, P, false, npairs, ci.cus, INT32_MAX), full, st), "reverse pass");
    if (npairs <= 65536) {
      NA_CHECK(launch_band(bv, 0, (int)std::min<int64_t>(full, (npairs + 64 * kNumBandClasses + 255) / 256), st),
               "reverse band pass");
    } else {
      for (int P = 16; P >= 1; P /= 2) NA_CHECK(launch_band(bv, P, full, st), "reverse band pass");
      NA_CHECK(launch_band(bv, kBandDiag, full, st), "reverse band pass");
    }
    NA_CHECK(launch_plain(bv, plain_grid, kPlainMaxCols, st), "reverse pass (32-bit)");
    NA_CHECK(launch_finalize(b, st), "finalize");
  }
  hipLaunchKernelGGL(na_status_kernel, dim3(1), dim3(64), 0, st, state, b.errors, status);
  if (int rc = launched("status")) return rc;
#undef NA_CHECK
  return MHMSW_OK;
}
