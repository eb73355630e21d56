// Device view of a staged local assembly batch and one launch of the walk
// kernel (locassm_kernels.hip), filled by the host runtime (locassm.cpp).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "locassm_core.h"

namespace mhmla {

// a table slot's key word: fingerprint (32 bits) | whether the slot's info
// word holds the slot of the next occurrence in the stored occurrence's read
// (1 bit) | byte offset of the stored occurrence in the read buffer (31
// bits); all ones = empty
constexpr uint64_t kEmpty = ~0ull;
constexpr uint64_t kHasNext = 0x80000000ull;
constexpr uint32_t kOffMask = 0x7FFFFFFFu;
// a slot's info word: during counting, the slot of the occurrence after the
// stored one (see kHasNext); then extension code (3 bits) | successor slot
// << 3, the walk marking a visited slot with the top bit
constexpr uint32_t kNoSlot = 0x0FFFFFFFu;
constexpr uint32_t kVisited = 0x80000000u;
// a job's table word: slots (any count up to 2^27) | wide-counter flag
constexpr int32_t kCapMask = 0x0FFFFFFF;
constexpr int32_t kWideCounts = 0x40000000;

struct WalkView {
  // staged batch
  const uint8_t* ctg_seq;
  const int64_t* ctg_off;
  const int32_t* ctg_len;
  const int32_t* depth;        // (int) of the double depth, as set_ext receives it
  // reads in side order: side j (contig * 2 + side) owns reads [rfirst[j], rfirst[j+1]),
  // whose bytes are contiguous in read_seq / read_qual
  const int64_t* rfirst;
  const uint8_t* read_seq;
  const uint8_t* read_qual;
  const int64_t* read_off;
  const int32_t* read_len;
  // this launch: jobs (contig * 2 + side, side 1 = right) and their scratch
  const int32_t* jobs;
  const int64_t* scratch_off;  // bytes into scratch, per launch slot
  const int32_t* cap;          // table slots | kWideCounts, per launch slot
  uint8_t* scratch;
  // outputs, indexed by job
  int32_t* out_len;
  uint8_t* out_ext;            // walk_len_limit bytes per job
  int32_t* out_terms;          // 3 per job: walks ending in X, F, R
  unsigned long long* counters;  // [0] mers counted, [1] walk steps, [2] walks, [3] full tables,
                                 // [4..7] shader cycles in clear, count, decide, walk (summed over sides),
                                 // [8] counts redone in a larger table
  // parameters
  int32_t kmer_len, min_kmer_len, max_kmer_len, shift, walk_len_limit;
  int32_t qual_offset, min_qual, min_hi_qual, max_reads;
  int32_t lds_bytes;  // dynamic LDS per side: sides whose reads (+64) fit it read them from LDS
  Rules rules;
};

// scratch layout of one job (byte offsets from its base), shared by the host
// planner and the kernel
struct JobLayout {
  int64_t key, cnt, info, roff, rlen, cum, walk0, walk1, smer, bytes;
};

inline __host__ __device__ int64_t la_align(int64_t x) { return (x + 255) & ~int64_t(255); }

inline __host__ __device__ JobLayout job_layout(int64_t cap, bool wide, int64_t nreads, int64_t walk_len_limit,
                                                int64_t max_kmer_len) {
  JobLayout l;
  l.key = 0;
  l.cnt = la_align(l.key + 8 * cap);
  l.info = la_align(l.cnt + (wide ? 32 : 16) * cap);
  l.roff = la_align(l.info + 4 * cap);
  l.rlen = la_align(l.roff + 8 * nreads);
  l.cum = la_align(l.rlen + 4 * nreads);
  l.walk0 = la_align(l.cum + 4 * (nreads + 1));
  l.walk1 = la_align(l.walk0 + walk_len_limit);
  l.smer = la_align(l.walk1 + walk_len_limit);
  l.bytes = la_align(l.smer + max_kmer_len + 64);
  return l;
}

void launch_walks(const WalkView& v, int64_t njobs, hipStream_t st);

}  // namespace mhmla
