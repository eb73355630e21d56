// Host runtime of the mhmla C ABI (include/mhmla.h): local assembly walks.
//
// The reference driver (src/gpu_loc_assem/driver.cpp:35-319) slices the
// contigs by a memory estimate, copies each slice as strings, launches the
// walk kernel for the right sides, reverse-complements every contig on the
// host, copies the contigs again and launches for the left sides.  Here a
// batch is copied once; the left sides read the contig reverse-complemented
// on the fly, so both sides of every contig are jobs of the same launches.
// Jobs are ordered by table size (largest first, so the longest jobs start
// first) and cut into launches whose scratch (tables sized by each side's own
// mer count, not by a batch-wide maximum) fits a budget.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "../../include/mhmla.h"
#include "host_pinned.h"
#include "locassm_launch.h"

using namespace mhmla;

namespace {

template <class T>
struct Buf {
  T* ptr = nullptr;
  size_t cap = 0;
  hipError_t ensure(size_t n) {
    if (n <= cap) return hipSuccess;
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    cap = 0;
    const size_t want = std::max<size_t>(n + n / 8, 64);
    const hipError_t e = hipMalloc((void**)&ptr, want * sizeof(T));
    if (e == hipSuccess) cap = want;
    return e;
  }
  void release() {
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    cap = 0;
  }
};

// fn(t, begin, end) over [0, n) split in contiguous parts, on up to 16
// threads (one below 'serial_below' items)
template <class Fn>
void parallel_parts(int64_t n, int64_t serial_below, Fn fn) {
  int nth = (int)std::min<unsigned>(16u, std::max(1u, std::thread::hardware_concurrency()));
  if (n < serial_below) nth = 1;
  if (nth == 1) {
    fn(0, int64_t(0), n);
    return;
  }
  std::vector<std::thread> th;
  for (int t = 0; t < nth; t++) th.emplace_back(fn, t, n * t / nth, n * (t + 1) / nth);
  for (auto& x : th) x.join();
}

// bytes the reverse complement maps (ACGTN and the IUPAC codes)
const bool* valid_ctg_byte() {
  static const auto table = [] {
    static bool t[256];
    for (int b = 0; b < 256; b++) t[b] = complement((uint8_t)b) != 0;
    return t;
  }();
  return table;
}

constexpr int64_t kDefaultBudget = int64_t(16) << 30;  // scratch bytes per launch (at most half the free memory)
constexpr int64_t kPad = 64;                           // bytes readable past every staged buffer
constexpr int64_t kMaxCap = int64_t(1) << 27;         // slots of one table
constexpr int kCounters = 9;                          // WalkView::counters
constexpr int32_t kLdsReadBytes = 12288;               // LDS per side for reads and a small table's info words

}  // namespace

struct mhmla_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  std::string err;
  // staged batch (host copies of what planning needs)
  bool staged = false;
  int64_t nctgs = 0;
  std::vector<int32_t> ctg_len;
  std::vector<int64_t> rfirst;  // reads of side j: [rfirst[j], rfirst[j+1]) in side order
  std::vector<int32_t> clen;    // their lengths
  Buf<uint8_t> d_ctg, d_seq, d_qual;
  Buf<int64_t> d_ctg_off, d_rfirst, d_read_off;
  Buf<int32_t> d_ctg_len, d_depth, d_read_len;
  HostPinned h_ctg, h_seq, h_qual;  // staging copies of the contigs and the side-ordered reads
  // last run
  bool ran = false;
  mhmla_params params{};
  std::vector<int32_t> host_terms;            // per job: walk terminations of sides answered on the host
  std::vector<uint8_t> on_device;             // per job
  Buf<int32_t> d_jobs, d_cap, d_out_len, d_out_terms;
  Buf<int64_t> d_soff;
  Buf<uint8_t> d_scratch, d_out_ext;
  Buf<unsigned long long> d_counters;
  mhmla_stats stats{};
  int64_t budget = kDefaultBudget;
};

namespace {

int fail(mhmla_ctx* c, int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  if (c) c->err = buf;
  return code;
}

int hip_fail(mhmla_ctx* c, hipError_t e, const char* what) {
  return fail(c, e == hipErrorOutOfMemory ? MHMLA_E_NOMEM : MHMLA_E_HIP, "%s: %s", what, hipGetErrorString(e));
}

#define LA_HIP(call, what)                                  \
  do {                                                      \
    const hipError_t e_ = (call);                           \
    if (e_ != hipSuccess) return hip_fail(c, e_, what);     \
  } while (0)

bool bad_params(const mhmla_params* p, std::string* why) {
  char b[256];
  if (p->min_kmer_len < 1 || p->max_kmer_len > 65535 || p->kmer_len < 1) {
    snprintf(b, sizeof b, "mer lengths must lie in 1..65535 (kmer_len %d, min %d, max %d)", p->kmer_len,
             p->min_kmer_len, p->max_kmer_len);
  } else if (p->shift <= 0) {
    snprintf(b, sizeof b, "shift must be positive (got %d)", p->shift);
  } else if (p->walk_len_limit < 0 || p->walk_len_limit > (1 << 24)) {
    snprintf(b, sizeof b, "walk_len_limit must lie in 0..16777216 (got %d)", p->walk_len_limit);
  } else if (p->max_reads < 0 || p->max_reads > 65535) {
    snprintf(b, sizeof b, "max_reads must lie in 0..65535 (got %d)", p->max_reads);
  } else {
    return false;
  }
  *why = b;
  return true;
}

// walks of a side whose table is empty: every walk is a dead end, so the
// schedule only shortens the mer (kernel.cpp:509-517)
int32_t empty_side_walks(const mhmla_params& p, int L) {
  const int max_len = std::min(p.max_kmer_len, L);
  int32_t n = 0;
  for (int k = p.kmer_len; k >= p.min_kmer_len && k <= max_len; k -= p.shift) n++;
  return n;
}

}  // namespace

extern "C" {

void mhmla_default_params(mhmla_params* p) {
  p->kmer_len = 33;
  p->min_kmer_len = 21;
  p->max_kmer_len = 121;
  p->shift = 8;
  p->walk_len_limit = 300;
  p->qual_offset = 33;
  p->min_qual = 10;
  p->min_hi_qual = 20;
  p->rating_thres = 0;
  p->max_reads = 3000;
  p->min_viable_depth = 0.2;
  p->min_expected_depth = 0.3;
}

int mhmla_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

int mhmla_create(int device, mhmla_ctx** out) {
  if (!out) return MHMLA_E_ARG;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return MHMLA_E_NODEVICE;
  if (device < 0 || device >= n) return MHMLA_E_ARG;
  mhmla_ctx* c = new mhmla_ctx();
  c->device = device;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess) {
    mhmla_destroy(c);
    return MHMLA_E_HIP;
  }
  *out = c;
  return MHMLA_OK;
}

void mhmla_destroy(mhmla_ctx* c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  c->d_ctg.release();
  c->d_seq.release();
  c->d_qual.release();
  c->h_ctg.release();
  c->h_seq.release();
  c->h_qual.release();
  c->d_ctg_off.release();
  c->d_rfirst.release();
  c->d_read_off.release();
  c->d_ctg_len.release();
  c->d_depth.release();
  c->d_read_len.release();
  c->d_jobs.release();
  c->d_cap.release();
  c->d_out_len.release();
  c->d_out_terms.release();
  c->d_soff.release();
  c->d_scratch.release();
  c->d_out_ext.release();
  c->d_counters.release();
  if (c->ev0) (void)hipEventDestroy(c->ev0);
  if (c->ev1) (void)hipEventDestroy(c->ev1);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

const char* mhmla_last_error(const mhmla_ctx* c) { return c ? c->err.c_str() : "no context"; }

int mhmla_stage(mhmla_ctx* c, int64_t nctgs, const char* ctg_seq, const int64_t* ctg_off, const int32_t* ctg_len,
                const double* depth, const int64_t* left_first, const int64_t* right_first, const int64_t* read_idx,
                int64_t nreads, const char* read_seq, const char* read_qual, const int64_t* read_off,
                const int32_t* read_len) {
  if (!c) return MHMLA_E_ARG;
  c->staged = false;
  c->ran = false;
  if (nctgs < 0 || nreads < 0) return fail(c, MHMLA_E_ARG, "negative count");
  if (nctgs > (int64_t(1) << 30)) return fail(c, MHMLA_E_ARG, "too many contigs (%lld)", (long long)nctgs);
  if (nctgs > 0 && (!ctg_seq || !ctg_off || !ctg_len || !depth || !left_first || !right_first))
    return fail(c, MHMLA_E_ARG, "null contig array");
  if (nreads > 0 && (!read_seq || !read_qual || !read_off || !read_len)) return fail(c, MHMLA_E_ARG, "null read array");
  // contigs: packed, checked for bytes the reverse complement maps
  int64_t total = 0;
  std::vector<int64_t> coff((size_t)nctgs);
  for (int64_t i = 0; i < nctgs; i++) {
    if (ctg_len[i] < 0 || ctg_off[i] < 0) return fail(c, MHMLA_E_ARG, "contig %lld: negative offset or length", (long long)i);
    coff[(size_t)i] = total;
    total += ctg_len[i];
  }
  LA_HIP(hipSetDevice(c->device), "hipSetDevice");
  LA_HIP(c->h_ctg.ensure((size_t)total + 1), "pinned contig staging");
  uint8_t* cbuf = c->h_ctg.ptr;
  cbuf[total] = 0;
  c->ctg_len.assign(ctg_len, ctg_len + nctgs);
  std::vector<int32_t> idepth((size_t)nctgs);
  {
    // per thread: the first contig holding a byte outside the alphabet
    std::vector<int64_t> bad_ctg(16, -1);
    std::vector<int32_t> bad_pos(16, 0);
    const bool* ok = valid_ctg_byte();
    parallel_parts(nctgs, 256, [&](int t, int64_t i0, int64_t i1) {
      for (int64_t i = i0; i < i1; i++) {
        const uint8_t* sq = (const uint8_t*)ctg_seq + ctg_off[i];
        bool good = true;
        for (int32_t j = 0; j < ctg_len[i]; j++) good &= ok[sq[j]];
        if (!good) {
          int32_t j = 0;
          while (ok[sq[j]]) j++;
          bad_ctg[(size_t)t] = i;
          bad_pos[(size_t)t] = j;
          return;
        }
        memcpy(cbuf + coff[(size_t)i], sq, (size_t)ctg_len[i]);
        idepth[(size_t)i] = (int32_t)depth[i];  // set_ext(int) receives the double depth (kernel.cpp:377)
      }
    });
    for (size_t t = 0; t < bad_ctg.size(); t++)
      if (bad_ctg[t] >= 0) {
        const int64_t i = bad_ctg[t];
        const int32_t j = bad_pos[t];
        return fail(c, MHMLA_E_ARG, "contig %lld holds byte 0x%02x at %d (not ACGTN or an IUPAC code)", (long long)i,
                    (unsigned)(uint8_t)ctg_seq[ctg_off[i] + j], j);
      }
  }
  // read lists, validated, then the reads copied in side order (side j =
  // contig * 2 + side, left 0, right 1) so each side's bytes are contiguous
  for (int64_t i = 0; i < nctgs; i++) {
    if (left_first[i + 1] < left_first[i] || right_first[i + 1] < right_first[i] || left_first[i] < 0 ||
        right_first[i] < 0)
      return fail(c, MHMLA_E_ARG, "contig %lld: read ranges decrease", (long long)i);
  }
  const int64_t nidx = nctgs > 0 ? std::max(left_first[nctgs], right_first[nctgs]) : 0;
  for (int64_t j = 0; j < nidx; j++)
    if (read_idx[j] < 0 || read_idx[j] >= nreads)
      return fail(c, MHMLA_E_ARG, "read_idx[%lld] = %lld is not a read number", (long long)j, (long long)read_idx[j]);
  for (int64_t r = 0; r < nreads; r++)
    if (read_len[r] < 0 || read_len[r] > 65535 || read_off[r] < 0)
      return fail(c, MHMLA_E_ARG, "read %lld: length %d outside 0..65535 or negative offset", (long long)r, read_len[r]);
  c->rfirst.assign((size_t)(2 * nctgs + 1), 0);
  c->clen.clear();
  std::vector<int64_t> croff;
  int64_t cbytes = 0;
  for (int64_t i = 0; i < nctgs; i++)
    for (int side = 0; side < 2; side++) {
      const int64_t* first = side ? right_first : left_first;
      c->rfirst[(size_t)(2 * i + side)] = (int64_t)c->clen.size();
      for (int64_t j = first[i]; j < first[i + 1]; j++) {
        croff.push_back(cbytes);
        c->clen.push_back(read_len[read_idx[j]]);
        cbytes += read_len[read_idx[j]];
      }
    }
  c->rfirst[(size_t)(2 * nctgs)] = (int64_t)c->clen.size();
  if (cbytes > int64_t(kOffMask) - kPad)
    return fail(c, MHMLA_E_ARG, "the sides' reads hold %lld bytes, above 2 GiB", (long long)cbytes);
  LA_HIP(c->h_seq.ensure((size_t)std::max<int64_t>(cbytes, 1)), "pinned read staging");
  LA_HIP(c->h_qual.ensure((size_t)std::max<int64_t>(cbytes, 1)), "pinned read staging");
  uint8_t* cseq = c->h_seq.ptr;
  uint8_t* cqual = c->h_qual.ptr;
  parallel_parts(nctgs, 256, [&](int, int64_t i0, int64_t i1) {
    for (int64_t i = i0; i < i1; i++)
      for (int side = 0; side < 2; side++) {
        const int64_t* first = side ? right_first : left_first;
        int64_t k = c->rfirst[(size_t)(2 * i + side)];
        for (int64_t j = first[i]; j < first[i + 1]; j++, k++) {
          const int64_t r = read_idx[j];
          memcpy(cseq + croff[(size_t)k], read_seq + read_off[r], (size_t)read_len[r]);
          memcpy(cqual + croff[(size_t)k], read_qual + read_off[r], (size_t)read_len[r]);
        }
      }
  });
  const int64_t ncr = (int64_t)c->clen.size();
  LA_HIP(c->d_ctg.ensure((size_t)(total + kPad)), "contig buffer");
  LA_HIP(c->d_ctg_off.ensure((size_t)nctgs + 1), "contig offsets");
  LA_HIP(c->d_ctg_len.ensure((size_t)nctgs + 1), "contig lengths");
  LA_HIP(c->d_depth.ensure((size_t)nctgs + 1), "depths");
  LA_HIP(c->d_rfirst.ensure((size_t)(2 * nctgs + 1)), "side read ranges");
  LA_HIP(c->d_read_off.ensure((size_t)ncr + 1), "read offsets");
  LA_HIP(c->d_read_len.ensure((size_t)ncr + 1), "read lengths");
  LA_HIP(c->d_seq.ensure((size_t)(cbytes + kPad)), "read bases");
  LA_HIP(c->d_qual.ensure((size_t)(cbytes + kPad)), "read qualities");
  hipStream_t st = c->stream;
  LA_HIP(hipMemcpyAsync(c->d_ctg.ptr, cbuf, (size_t)total + 1, hipMemcpyHostToDevice, st), "copy contigs");
  if (nctgs) {
    LA_HIP(hipMemcpyAsync(c->d_ctg_off.ptr, coff.data(), 8 * (size_t)nctgs, hipMemcpyHostToDevice, st), "copy");
    LA_HIP(hipMemcpyAsync(c->d_ctg_len.ptr, ctg_len, 4 * (size_t)nctgs, hipMemcpyHostToDevice, st), "copy");
    LA_HIP(hipMemcpyAsync(c->d_depth.ptr, idepth.data(), 4 * (size_t)nctgs, hipMemcpyHostToDevice, st), "copy");
  }
  LA_HIP(hipMemcpyAsync(c->d_rfirst.ptr, c->rfirst.data(), 8 * c->rfirst.size(), hipMemcpyHostToDevice, st), "copy");
  if (ncr) {
    LA_HIP(hipMemcpyAsync(c->d_read_off.ptr, croff.data(), 8 * (size_t)ncr, hipMemcpyHostToDevice, st), "copy");
    LA_HIP(hipMemcpyAsync(c->d_read_len.ptr, c->clen.data(), 4 * (size_t)ncr, hipMemcpyHostToDevice, st), "copy");
  }
  if (cbytes) {
    LA_HIP(hipMemcpyAsync(c->d_seq.ptr, cseq, (size_t)cbytes, hipMemcpyHostToDevice, st), "copy read bases");
    LA_HIP(hipMemcpyAsync(c->d_qual.ptr, cqual, (size_t)cbytes, hipMemcpyHostToDevice, st),
           "copy read qualities");
  }
  // the host vectors above go out of scope, and the pinned buffers are
  // refilled by the next stage: wait for the copies
  LA_HIP(hipStreamSynchronize(st), "staging copies");
  c->nctgs = nctgs;
  c->staged = true;
  c->stats = mhmla_stats{};
  c->stats.contigs = nctgs;
  return MHMLA_OK;
}

int mhmla_run(mhmla_ctx* c, const mhmla_params* p) {
  if (!c || !p) return MHMLA_E_ARG;
  if (!c->staged) return fail(c, MHMLA_E_STATE, "mhmla_run before mhmla_stage");
  std::string why;
  if (bad_params(p, &why)) return fail(c, MHMLA_E_ARG, "%s", why.c_str());
  c->ran = false;
  c->params = *p;
  const int64_t n = c->nctgs, njobs = 2 * n;
  const int W = p->walk_len_limit;
  c->host_terms.assign((size_t)njobs * 3, 0);
  c->on_device.assign((size_t)njobs, 0);
  // plan: one job per contig side with reads and at least one mer length to try
  struct Plan {
    int32_t job, cap;
    int64_t bytes;
  };
  std::vector<Plan> plan;
  plan.reserve((size_t)njobs);
  for (int64_t ci = 0; ci < n; ci++) {
    const int L = c->ctg_len[(size_t)ci];
    const int32_t iters = empty_side_walks(*p, L);
    for (int side = 0; side < 2; side++) {
      const int64_t job = 2 * ci + side;
      const int64_t f0 = c->rfirst[(size_t)job];
      const int64_t nr = std::min<int64_t>(c->rfirst[(size_t)job + 1] - f0, p->max_reads);
      int64_t occ = 0;
      for (int64_t i = 0; i < nr; i++) occ += std::max(0, c->clen[(size_t)(f0 + i)] - p->min_kmer_len);
      if (iters == 0 || occ == 0) {
        c->host_terms[(size_t)job * 3] = iters;  // every walk a dead end
        continue;
      }
      // at most 80% full; 16-bit vote counters while no count can reach 65536
      const int64_t cap = (occ + occ / 4 + 64 + 63) & ~int64_t(63);
      const bool wide = occ >= 65535;
      if (cap > kMaxCap)
        return fail(c, MHMLA_E_ARG, "contig %lld side %d: %lld mers exceed one table", (long long)ci, side,
                    (long long)occ);
      plan.push_back({(int32_t)job, (int32_t)cap | (wide ? kWideCounts : 0),
                      job_layout(cap, wide, nr, W, p->max_kmer_len).bytes});
      c->on_device[(size_t)job] = 1;
    }
  }
  std::stable_sort(plan.begin(), plan.end(), [](const Plan& a, const Plan& b) { return a.bytes > b.bytes; });
  const int64_t nd = (int64_t)plan.size();
  std::vector<int32_t> jobs((size_t)nd), caps((size_t)nd);
  std::vector<int64_t> soff((size_t)nd);
  std::vector<std::pair<int64_t, int64_t>> launches;  // [first, end) of plan
  size_t free_b = 0, total_b = 0;
  LA_HIP(hipSetDevice(c->device), "hipSetDevice");
  LA_HIP(hipMemGetInfo(&free_b, &total_b), "hipMemGetInfo");
  const int64_t budget = std::max<int64_t>(std::min<int64_t>(c->budget, (int64_t)(free_b / 2) + c->d_scratch.cap),
                                           nd ? plan[0].bytes : 0);
  int64_t used = 0, start = 0, biggest = 0;
  for (int64_t i = 0; i < nd; i++) {
    if (used + plan[(size_t)i].bytes > budget) {
      launches.push_back({start, i});
      biggest = std::max(biggest, used);
      start = i;
      used = 0;
    }
    jobs[(size_t)i] = plan[(size_t)i].job;
    caps[(size_t)i] = plan[(size_t)i].cap;
    soff[(size_t)i] = used;
    used += plan[(size_t)i].bytes;
  }
  if (nd > start) {
    launches.push_back({start, nd});
    biggest = std::max(biggest, used);
  }

  LA_HIP(hipSetDevice(c->device), "hipSetDevice");
  LA_HIP(c->d_jobs.ensure((size_t)nd + 1), "job list");
  LA_HIP(c->d_cap.ensure((size_t)nd + 1), "job tables");
  LA_HIP(c->d_soff.ensure((size_t)nd + 1), "job scratch");
  LA_HIP(c->d_scratch.ensure((size_t)biggest + 256), "scratch");
  LA_HIP(c->d_out_len.ensure((size_t)njobs + 1), "walk lengths");
  LA_HIP(c->d_out_terms.ensure((size_t)njobs * 3 + 1), "walk terminations");
  LA_HIP(c->d_out_ext.ensure((size_t)njobs * (size_t)W + 1), "walks");
  LA_HIP(c->d_counters.ensure(kCounters), "counters");
  hipStream_t st = c->stream;
  if (nd) {
    LA_HIP(hipMemcpyAsync(c->d_jobs.ptr, jobs.data(), 4 * (size_t)nd, hipMemcpyHostToDevice, st), "copy jobs");
    LA_HIP(hipMemcpyAsync(c->d_cap.ptr, caps.data(), 4 * (size_t)nd, hipMemcpyHostToDevice, st), "copy jobs");
    LA_HIP(hipMemcpyAsync(c->d_soff.ptr, soff.data(), 8 * (size_t)nd, hipMemcpyHostToDevice, st), "copy jobs");
  }
  LA_HIP(hipMemsetAsync(c->d_counters.ptr, 0, kCounters * sizeof(unsigned long long), st), "clear counters");
  LA_HIP(hipEventRecord(c->ev0, st), "event");
  WalkView v{};
  v.ctg_seq = c->d_ctg.ptr;
  v.ctg_off = c->d_ctg_off.ptr;
  v.ctg_len = c->d_ctg_len.ptr;
  v.depth = c->d_depth.ptr;
  v.rfirst = c->d_rfirst.ptr;
  v.read_seq = c->d_seq.ptr;
  v.read_qual = c->d_qual.ptr;
  v.read_off = c->d_read_off.ptr;
  v.read_len = c->d_read_len.ptr;
  v.scratch = c->d_scratch.ptr;
  v.out_len = c->d_out_len.ptr;
  v.out_ext = c->d_out_ext.ptr;
  v.out_terms = c->d_out_terms.ptr;
  v.counters = c->d_counters.ptr;
  v.kmer_len = p->kmer_len;
  v.min_kmer_len = p->min_kmer_len;
  v.max_kmer_len = p->max_kmer_len;
  v.shift = p->shift;
  v.walk_len_limit = W;
  v.qual_offset = p->qual_offset;
  v.min_qual = p->min_qual;
  v.min_hi_qual = p->min_hi_qual;
  v.max_reads = p->max_reads;
  v.lds_bytes = kLdsReadBytes;
  v.rules = Rules{p->min_viable_depth, p->min_expected_depth, p->rating_thres};
  for (const auto& lr : launches) {
    v.jobs = c->d_jobs.ptr + lr.first;
    v.cap = c->d_cap.ptr + lr.first;
    v.scratch_off = c->d_soff.ptr + lr.first;
    launch_walks(v, lr.second - lr.first, st);
  }
  LA_HIP(hipGetLastError(), "walk kernel launch");
  LA_HIP(hipEventRecord(c->ev1, st), "event");
  c->stats.sides = nd;
  c->stats.launches = (int64_t)launches.size();
  c->stats.scratch_bytes = biggest;
  c->ran = true;
  return MHMLA_OK;
}

int mhmla_wait(mhmla_ctx* c) {
  if (!c) return MHMLA_E_ARG;
  if (!c->ran) return fail(c, MHMLA_E_STATE, "mhmla_wait without a run");
  LA_HIP(hipSetDevice(c->device), "hipSetDevice");
  LA_HIP(hipStreamSynchronize(c->stream), "local assembly walks");
  float ms = 0.f;
  LA_HIP(hipEventElapsedTime(&ms, c->ev0, c->ev1), "event timing");
  unsigned long long cnt[kCounters];
  LA_HIP(hipMemcpy(cnt, c->d_counters.ptr, sizeof cnt, hipMemcpyDeviceToHost), "copy counters");
  if (cnt[3]) return fail(c, MHMLA_E_HIP, "%llu mer insertions found their table full", cnt[3]);
  c->stats.job_ms = ms;
  c->stats.mers = (int64_t)cnt[0];
  c->stats.walk_steps = (int64_t)cnt[1];
  c->stats.walks = (int64_t)cnt[2];
  for (int i = 0; i < 4; i++) c->stats.phase_cycles[i] = (int64_t)cnt[4 + i];
  c->stats.table_resizes = (int64_t)cnt[8];
  for (size_t j = 0; j < c->on_device.size(); j++)
    if (!c->on_device[j]) c->stats.walks += c->host_terms[3 * j];
  return MHMLA_OK;
}

int mhmla_results(mhmla_ctx* c, int32_t* left_len, char* left_ext, int32_t* right_len, char* right_ext,
                  int32_t* terms) {
  if (!c) return MHMLA_E_ARG;
  if (!c->ran) return fail(c, MHMLA_E_STATE, "mhmla_results without a run");
  const int64_t n = c->nctgs, njobs = 2 * n;
  const int64_t W = c->params.walk_len_limit;
  LA_HIP(hipSetDevice(c->device), "hipSetDevice");
  LA_HIP(hipStreamSynchronize(c->stream), "local assembly walks");
  std::vector<int32_t> len((size_t)njobs), tm((size_t)njobs * 3);
  std::vector<uint8_t> ext((size_t)(njobs * W));
  if (njobs) {
    LA_HIP(hipMemcpy(len.data(), c->d_out_len.ptr, 4 * (size_t)njobs, hipMemcpyDeviceToHost), "copy walk lengths");
    LA_HIP(hipMemcpy(tm.data(), c->d_out_terms.ptr, 12 * (size_t)njobs, hipMemcpyDeviceToHost), "copy terminations");
    if (W) LA_HIP(hipMemcpy(ext.data(), c->d_out_ext.ptr, (size_t)(njobs * W), hipMemcpyDeviceToHost), "copy walks");
  }
  for (int64_t j = 0; j < njobs; j++) {
    if (!c->on_device[(size_t)j]) {
      len[(size_t)j] = 0;
      for (int t = 0; t < 3; t++) tm[(size_t)(3 * j + t)] = c->host_terms[(size_t)(3 * j + t)];
    }
  }
  for (int64_t ci = 0; ci < n; ci++) {
    const int64_t jl = 2 * ci, jr = 2 * ci + 1;
    if (left_len) left_len[ci] = len[(size_t)jl];
    if (right_len) right_len[ci] = len[(size_t)jr];
    if (left_ext && W) memcpy(left_ext + ci * W, ext.data() + jl * W, (size_t)len[(size_t)jl]);
    if (right_ext && W) memcpy(right_ext + ci * W, ext.data() + jr * W, (size_t)len[(size_t)jr]);
    if (terms)
      for (int t = 0; t < 3; t++) {
        terms[6 * ci + t] = tm[(size_t)(3 * jl + t)];
        terms[6 * ci + 3 + t] = tm[(size_t)(3 * jr + t)];
      }
  }
  return MHMLA_OK;
}

int mhmla_extend(mhmla_ctx* c, const mhmla_params* p, int64_t nctgs, const char* ctg_seq, const int64_t* ctg_off,
                 const int32_t* ctg_len, const double* depth, const int64_t* left_first, const int64_t* right_first,
                 const int64_t* read_idx, int64_t nreads, const char* read_seq, const char* read_qual,
                 const int64_t* read_off, const int32_t* read_len, int32_t* left_len, char* left_ext,
                 int32_t* right_len, char* right_ext, int32_t* terms) {
  int rc = mhmla_stage(c, nctgs, ctg_seq, ctg_off, ctg_len, depth, left_first, right_first, read_idx, nreads, read_seq,
                       read_qual, read_off, read_len);
  if (rc == MHMLA_OK) rc = mhmla_run(c, p);
  if (rc == MHMLA_OK) rc = mhmla_wait(c);
  if (rc == MHMLA_OK) rc = mhmla_results(c, left_len, left_ext, right_len, right_ext, terms);
  return rc;
}

int mhmla_get_stats(const mhmla_ctx* c, mhmla_stats* out) {
  if (!c || !out) return MHMLA_E_ARG;
  *out = c->stats;
  return MHMLA_OK;
}

}  // extern "C"
