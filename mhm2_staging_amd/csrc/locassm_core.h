// Local assembly: the per-mer extension rule and the hashing shared by the
// HIP kernels (locassm_kernels.hip) and the host (locassm.cpp, and the CPU
// check tests/cpp/locassm_core_check.cpp).
//
// The rule is MHM2's MerFreqs::set_ext with MerBase::get_base_rating
// (src/gpu_loc_assem/kernel.hpp:71-161): four bases, each rated 0..7 from its
// votes against depth-scaled thresholds, ranked by (rating, high-quality
// votes, votes), and the extension taken from the ranking.  The reference
// ranks with an exchange sort whose comparator also swaps full ties; only
// the keys of the first three entries (and the bases of the first two when
// their keys differ) reach the result, so a strict ranking gives the same
// extension.
#pragma once
#include <stdint.h>

#ifdef __HIPCC__
#include <hip/hip_runtime.h>
#define MHMLA_HD __host__ __device__ __forceinline__
#else
#define MHMLA_HD inline
#endif

namespace mhmla {

// extension codes: 0..3 = A C G T, then fork and none
constexpr int kExtF = 4;
constexpr int kExtX = 5;

struct Rules {
  double min_viable_depth;    // LASSM_MIN_VIABLE_DEPTH
  double min_expected_depth;  // LASSM_MIN_EXPECTED_DEPTH
  int32_t rating_thres;       // LASSM_RATING_THRES
};

// get_base_rating (kernel.hpp:71-82) with the thresholds already scaled.
MHMLA_HD int rating(uint32_t n, uint32_t hi, double min_viable, double min_expected) {
  const double dn = (double)n, dh = (double)hi;
  if (n == 0) return 0;
  if (n == 1) return 1;
  if (dn < min_viable) return 2;
  if (min_expected > dn) return dh < min_viable ? 3 : 4;
  if (dh < min_viable) return 5;
  if (min_viable < dh && dh < min_expected) return 6;
  return 7;
}

// one base's ranking key
struct BaseKey {
  int rating;
  uint32_t hi, low;
  int base;
};

// true when a ranks strictly before b: (rating, hi, low) larger, then the
// lower base index (an arbitrary strict order among full ties)
MHMLA_HD bool ranks_before(const BaseKey& a, const BaseKey& b) {
  if (a.rating != b.rating) return a.rating > b.rating;
  if (a.hi != b.hi) return a.hi > b.hi;
  if (a.low != b.low) return a.low > b.low;
  return a.base < b.base;
}

MHMLA_HD void order(BaseKey& x, BaseKey& y) {
  if (ranks_before(y, x)) {
    const BaseKey t = x;
    x = y;
    y = t;
  }
}

// set_ext (kernel.hpp:112-161): the extension code for votes low[b] (quality
// >= min_qual) and hi[b] (>= min_hi_qual), b = A C G T, at an integer depth.
MHMLA_HD int choose_ext(const uint32_t low[4], const uint32_t hi[4], int depth, const Rules& r) {
  double min_viable = r.min_viable_depth * depth;
  if (min_viable < 2.0) min_viable = 2.0;
  double min_expected = r.min_expected_depth * depth;
  if (min_expected < 2.0) min_expected = 2.0;
  BaseKey e0{rating(low[0], hi[0], min_viable, min_expected), hi[0], low[0], 0};
  BaseKey e1{rating(low[1], hi[1], min_viable, min_expected), hi[1], low[1], 1};
  BaseKey e2{rating(low[2], hi[2], min_viable, min_expected), hi[2], low[2], 2};
  BaseKey e3{rating(low[3], hi[3], min_viable, min_expected), hi[3], low[3], 3};
  // sorting network for four: e0 ranks first, e3 last
  order(e0, e1);
  order(e2, e3);
  order(e0, e2);
  order(e1, e3);
  order(e1, e2);
  const int top = e0.rating, runner_up = e1.rating;
  if (top <= r.rating_thres) return kExtX;
  if (top <= 3) return runner_up == 0 ? e0.base : kExtX;
  if (top < 6) return runner_up < 3 ? e0.base : kExtX;
  if (top == 6) return runner_up < 4 ? e0.base : kExtX;
  if (runner_up < 7) return e0.base;
  if (e2.rating == 7 || e0.low == e1.low) return kExtF;
  return e0.low > e1.low ? e0.base : e1.base;
}

// from a packed constant, not a string: on the GPU a string is a load from
// global memory, on the walk's critical path
MHMLA_HD char ext_char(int code) { return (char)((0x584654474341ull >> (8 * code)) & 0xFFu); }

// base index of a vote byte: A C G T -> 0..3, everything else -> -1
MHMLA_HD int vote_base(uint8_t c) {
  return c == 'A' ? 0 : c == 'C' ? 1 : c == 'G' ? 2 : c == 'T' ? 3 : -1;
}

// locassm_driver::revcomp (driver.hpp:19-36): ACGTN and the IUPAC codes
// U R Y K M S W B D H V (-> N); 0 for every other byte
MHMLA_HD uint8_t complement(uint8_t c) {
  switch (c) {
    case 'A': return 'T';
    case 'C': return 'G';
    case 'G': return 'C';
    case 'T': return 'A';
    case 'N': case 'U': case 'R': case 'Y': case 'K': case 'M': case 'S': case 'W': case 'B': case 'D':
    case 'H': case 'V': return 'N';
    default: return 0;
  }
}

// polynomial rolling hash of a mer's bytes modulo 2^64, mixed before use
constexpr uint64_t kHashBase = 0x9E3779B97F4A7C15ull | 1ull;

MHMLA_HD uint64_t mix64(uint64_t h) {
  h ^= h >> 33;
  h *= 0xff51afd7ed558ccdull;
  h ^= h >> 33;
  h *= 0xc4ceb9fe1a85ec53ull;
  h ^= h >> 33;
  return h;
}

MHMLA_HD uint64_t pow_base(int e) {
  uint64_t r = 1, b = kHashBase;
  while (e) {
    if (e & 1) r *= b;
    b *= b;
    e >>= 1;
  }
  return r;
}

}  // namespace mhmla
