// adept_sw::GPUDriver over the mhmsw C ABI.
//
// Keeps the reference's call protocol (src/klign.cpp:361-367,
// test/ssw-test.cpp:53-58): init once; per batch run_kernel_forwards ->
// kernel_block -> run_kernel_backwards -> kernel_block -> get_aln_results.
// run_kernel_forwards starts the whole job (both passes run on the device
// without a host round trip); run_kernel_backwards only checks the sequence.
// Errors throw std::runtime_error("GPU assert ...") like cudaErrchk
// (src/adept-sw/driver.cpp:17-26).
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <climits>
#include <cstdint>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/adept_sw/driver.hpp"
#include "../../include/mhmsw.h"
#include "thread_pool.h"

#ifndef ALN_AMBIGUITY_COST
#define ALN_AMBIGUITY_COST 1
#endif

namespace adept_sw {

struct DriverState {
  mhmsw_ctx *ctx = nullptr;
  mhmsw_scoring scoring{};
  size_t cap = 0;
  std::vector<short> store;  // 5 * cap
  std::vector<mhmsw_result> res;
  std::vector<const char *> rptr, cptr;  // each string's bytes, read in place
  std::vector<int32_t> rlen, clen;
  int64_t npairs = 0;
  bool submitted = false;
  bool fetched = false;
  bool backwards_called = false;
  bool score2 = false;        // batches compute score2 (set_score2)
  bool batch_score2 = false;  // the submitted batch did
  std::vector<int32_t> s2;
  bool cigar = false;         // batches compute CIGARs (set_cigar)
  bool batch_cigar = false;
  std::vector<int32_t> mism;
  std::vector<int64_t> op_off;
  std::vector<uint32_t> ops;
  std::vector<std::string> cigars;
};

static void throw_ctx(DriverState *s, int rc, const char *what, const char *file, int line) {
  std::ostringstream os;
  os << "GPU assert " << what << " (mhmsw " << rc << ": " << (s && s->ctx ? mhmsw_last_error(s->ctx) : "") << ") "
     << file << ":" << line << "\n";
  throw std::runtime_error(os.str());
}

#define MHMSW_CHECK(s, expr)                                    \
  do {                                                          \
    int _rc = (expr);                                           \
    if (_rc < 0) throw_ctx((s), _rc, #expr, __FILE__, __LINE__); \
  } while (0)

size_t get_tot_gpu_mem() {
  size_t b = 0;
  if (mhmsw_device_total_mem(0, &b) != MHMSW_OK) return 0;  // device 0, as driver.cpp:40-44
  return b;
}

size_t get_avail_gpu_mem_per_rank(int totRanks, int num_devices) {
  if (num_devices == 0) num_devices = get_num_node_gpus();
  if (!num_devices) return 0;
  int ranksPerDevice = std::max(1, totRanks / num_devices);
  return (size_t)(get_tot_gpu_mem() * 0.8) / ranksPerDevice;
}

std::string get_device_name(int device_id) {
  char buf[256];
  int rc = mhmsw_device_name(device_id, buf, sizeof(buf));
  if (rc != MHMSW_OK) {
    std::ostringstream os;
    os << "GPU assert mhmsw_device_name(" << device_id << ") failed: " << rc << "\n";
    throw std::runtime_error(os.str());
  }
  return std::string(buf);
}

int get_num_node_gpus() {
  int n = 0;
  if (mhmsw_device_count(&n) != MHMSW_OK) return 0;
  return n;
}

bool initialize_gpu(double &time_to_initialize, int &device_count, size_t &total_mem) {
  auto t = std::chrono::high_resolution_clock::now();
  device_count = get_num_node_gpus();
  if (device_count > 0) {
    total_mem = get_tot_gpu_mem();
    void *first_touch = nullptr;
    if (hipHostMalloc(&first_touch, sizeof(double), hipHostMallocDefault) == hipSuccess) (void)hipHostFree(first_touch);
  }
  std::chrono::duration<double> elapsed = std::chrono::high_resolution_clock::now() - t;
  time_to_initialize = elapsed.count();
  return device_count > 0;
}

bool initialize_gpu() {
  double t;
  int c;
  size_t m;
  return initialize_gpu(t, c, m);
}

GPUDriver::~GPUDriver() {
  if (driver_state) {
    if (driver_state->ctx) mhmsw_destroy(driver_state->ctx);
    delete driver_state;
    driver_state = nullptr;
  }
}

double GPUDriver::init(int upcxx_rank_me, int upcxx_rank_n, short match_score, short mismatch_score,
                       short gap_opening_score, short gap_extending_score, int rlen_limit) {
  (void)upcxx_rank_n;
  (void)rlen_limit;  // buffers are sized per batch by actual bytes
  auto t = std::chrono::high_resolution_clock::now();
  if (driver_state) {
    if (driver_state->ctx) mhmsw_destroy(driver_state->ctx);
    delete driver_state;
  }
  driver_state = new DriverState();
  int ndev = get_num_node_gpus();
  if (ndev <= 0) throw std::runtime_error("GPU assert no GPU device available\n");
  int device = upcxx_rank_me % ndev;  // as driver.cpp:220
  MHMSW_CHECK(driver_state, mhmsw_create(device, &driver_state->ctx));
  driver_state->scoring.match = match_score;
  driver_state->scoring.mismatch = -mismatch_score;
  driver_state->scoring.gap_open = -gap_opening_score;
  driver_state->scoring.gap_ext = -gap_extending_score;
  driver_state->scoring.ambiguity = ALN_AMBIGUITY_COST;
  MHMSW_CHECK(driver_state, mhmsw_set_scoring(driver_state->ctx, &driver_state->scoring));
  MHMSW_CHECK(driver_state, mhmsw_set_orientation(driver_state->ctx, 1));
  // the driver reports no per-phase times: skip their events (host time per call)
  MHMSW_CHECK(driver_state, mhmsw_set_phase_timing(driver_state->ctx, 0));
  std::chrono::duration<double> elapsed = std::chrono::high_resolution_clock::now() - t;
  return elapsed.count();
}

void GPUDriver::set_ambiguity(short ambiguity_cost) {
  if (!driver_state) throw std::runtime_error("GPU assert set_ambiguity before init\n");
  driver_state->scoring.ambiguity = ambiguity_cost;
  MHMSW_CHECK(driver_state, mhmsw_set_scoring(driver_state->ctx, &driver_state->scoring));
}

void GPUDriver::set_score2(bool enable) {
  if (!driver_state) throw std::runtime_error("GPU assert set_score2 before init\n");
  MHMSW_CHECK(driver_state, mhmsw_set_score2(driver_state->ctx, enable ? 1 : 0));
  driver_state->score2 = enable;
}

const int *GPUDriver::get_score2() {
  DriverState *s = driver_state;
  if (!s || !s->fetched || !s->batch_score2) return nullptr;
  return s->s2.data();
}

void GPUDriver::set_cigar(bool enable) {
  if (!driver_state) throw std::runtime_error("GPU assert set_cigar before init\n");
  MHMSW_CHECK(driver_state, mhmsw_set_cigar(driver_state->ctx, enable ? 1 : 0));
  driver_state->cigar = enable;
}

const std::string *GPUDriver::get_cigar_strings() {
  DriverState *s = driver_state;
  if (!s || !s->fetched || !s->batch_cigar) return nullptr;
  return s->cigars.data();
}

const int *GPUDriver::get_mismatches() {
  DriverState *s = driver_state;
  if (!s || !s->fetched || !s->batch_cigar) return nullptr;
  return s->mism.data();
}

void GPUDriver::set_read_is_ssw_ref(bool read_is_ref) {
  if (!driver_state) throw std::runtime_error("GPU assert set_read_is_ssw_ref before init\n");
  MHMSW_CHECK(driver_state, mhmsw_set_orientation(driver_state->ctx, read_is_ref ? 1 : 0));
}

// Each sequence by its own pointer and length: the strings are packed where
// they lie (the reference copies them into one buffer first, driver.cpp:300-339).
// A fresh batch's string objects are cache-cold, and one thread walks them at
// about 4 ns a string (160 us for 20k pairs), so large batches fill the
// arrays on the host pool, in chunks taken from a counter.
static void point_at(const std::vector<std::string> &reads, const std::vector<std::string> &contigs,
                     DriverState *s) {
  const size_t n = reads.size();
  s->rptr.resize(n);
  s->rlen.resize(n);
  s->cptr.resize(n);
  s->clen.resize(n);
  std::atomic<bool> too_long(false);
  auto fill = [&](size_t i0, size_t i1) {
    for (size_t i = i0; i < i1; i++) {
      const size_t lr = reads[i].size(), lc = contigs[i].size();
      if (lr > (size_t)INT32_MAX || lc > (size_t)INT32_MAX) too_long = true;
      s->rptr[i] = reads[i].data();
      s->rlen[i] = (int32_t)lr;
      s->cptr[i] = contigs[i].data();
      s->clen[i] = (int32_t)lc;
    }
  };
  constexpr size_t kChunk = 2048;
  if (n < 4 * kChunk) {
    fill(0, n);
  } else {
    const int nth = (int)std::min<unsigned>(16, std::max(1u, std::thread::hardware_concurrency()));
    std::atomic<size_t> next(0);
    hostpool::pool().run_claimed(nth, [&](int) {
      for (size_t i0; (i0 = next.fetch_add(kChunk)) < n;) fill(i0, std::min(n, i0 + kChunk));
    });
  }
  if (too_long) throw std::runtime_error("GPU assert sequence longer than 2^31 bytes\n");
}

void GPUDriver::run_kernel_forwards(std::vector<std::string> &reads, std::vector<std::string> &contigs,
                                    unsigned maxReadSize, unsigned maxContigSize) {
  (void)maxReadSize;
  (void)maxContigSize;
  DriverState *s = driver_state;
  if (!s) throw std::runtime_error("GPU assert run_kernel_forwards before init\n");
  if (reads.size() != contigs.size()) throw std::runtime_error("GPU assert reads and contigs differ in count\n");
  point_at(reads, contigs, s);
  s->npairs = (int64_t)reads.size();
  MHMSW_CHECK(s, mhmsw_submit_seqs(s->ctx, s->npairs, s->rptr.data(), s->rlen.data(), s->cptr.data(),
                                   s->clen.data()));
  s->submitted = true;
  s->fetched = false;
  s->backwards_called = false;
  s->batch_score2 = s->score2;
  s->batch_cigar = s->cigar;
}

void GPUDriver::run_kernel_backwards(std::vector<std::string> &reads, std::vector<std::string> &contigs,
                                     unsigned maxReadSize, unsigned maxContigSize) {
  (void)maxReadSize;
  (void)maxContigSize;
  DriverState *s = driver_state;
  if (!s || !s->submitted) throw std::runtime_error("GPU assert run_kernel_backwards before run_kernel_forwards\n");
  if ((int64_t)reads.size() != s->npairs || (int64_t)contigs.size() != s->npairs)
    throw std::runtime_error("GPU assert run_kernel_backwards with a different batch\n");
  s->backwards_called = true;  // the begin pass was queued with the forward pass
}

static void fetch(DriverState *s, AlignmentResults &a) {
  if (s->fetched) return;
  const size_t n = (size_t)s->npairs;
  s->res.resize(std::max<size_t>(n, 1));
  MHMSW_CHECK(s, mhmsw_results(s->ctx, s->res.data(), s->npairs));
  if (s->batch_score2) {
    s->s2.resize(std::max<size_t>(n, 1));
    MHMSW_CHECK(s, mhmsw_results_score2(s->ctx, s->s2.data(), s->npairs));
  }
  if (s->batch_cigar) {
    // SSW's cigar_string spelling: <length><op> per op (src/ssw/ssw.cpp:72-90)
    static const char kOps[] = "MIDNSHP=X";
    int64_t total = 0;
    s->mism.resize(std::max<size_t>(n, 1));
    s->op_off.resize(n + 1);
    MHMSW_CHECK(s, mhmsw_results_cigar(s->ctx, s->npairs, s->mism.data(), s->op_off.data(), nullptr, 0, &total));
    s->ops.resize((size_t)std::max<int64_t>(total, 1));
    MHMSW_CHECK(s, mhmsw_results_cigar(s->ctx, s->npairs, nullptr, nullptr, s->ops.data(), total, &total));
    s->cigars.resize(std::max<size_t>(n, 1));
    for (size_t i = 0; i < n; i++) {
      std::string &c = s->cigars[i];
      c.clear();
      for (int64_t k = s->op_off[i]; k < s->op_off[i + 1]; k++) {
        c += std::to_string(s->ops[(size_t)k] >> 4);
        c += kOps[s->ops[(size_t)k] & 0xFu];
      }
    }
  }
  if (n > s->cap) {
    s->cap = std::max<size_t>(n, KLIGN_GPU_BLOCK_SIZE);
    s->store.assign(5 * s->cap, 0);
  }
  short *base = s->store.data();
  a.ref_begin = base;
  a.query_begin = base + s->cap;
  a.ref_end = base + 2 * s->cap;
  a.query_end = base + 3 * s->cap;
  a.top_scores = base + 4 * s->cap;
  for (size_t i = 0; i < n; i++) {
    const mhmsw_result &r = s->res[i];
    // AlignmentResults holds short (driver.hpp:17-23); the reference kept
    // its values in range by its 1024-base limit, this engine takes longer
    // sequences, so a value past SHRT_MAX fails as the reference's errors do
    // (begins are at most their ends, and at least -1)
    if (r.score > SHRT_MAX || r.ctg_end > SHRT_MAX || r.read_end > SHRT_MAX) {
      s->fetched = false;
      std::ostringstream os;
      os << "GPU assert pair " << i << ": score " << r.score << ", read end " << r.read_end << ", contig end "
         << r.ctg_end << " do not fit the short fields of AlignmentResults " << __FILE__ << ":" << __LINE__ << "\n";
      throw std::runtime_error(os.str());
    }
    a.ref_begin[i] = (short)r.ctg_begin;
    a.ref_end[i] = (short)r.ctg_end;
    a.query_begin[i] = (short)r.read_begin;
    a.query_end[i] = (short)r.read_end;
    a.top_scores[i] = (short)r.score;
  }
  s->fetched = true;
}

bool GPUDriver::kernel_is_done() {
  DriverState *s = driver_state;
  if (!s || !s->submitted) return true;
  int rc = mhmsw_query(s->ctx);
  if (rc < 0) throw_ctx(s, rc, "mhmsw_query", __FILE__, __LINE__);
  if (rc == 0) return false;
  fetch(s, alignments);
  return true;
}

void GPUDriver::kernel_block() {
  DriverState *s = driver_state;
  if (!s || !s->submitted) return;
  MHMSW_CHECK(s, mhmsw_wait(s->ctx));
  fetch(s, alignments);
}

}  // namespace adept_sw
