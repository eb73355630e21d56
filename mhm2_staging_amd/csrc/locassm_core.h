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

// true when base a ranks strictly before base b: (rating, hi, low) larger,
// then the lower base index (an arbitrary strict order among full ties)
MHMLA_HD bool ranks_before(int ra, uint32_t ha, uint32_t la, int a, int rb, uint32_t hb, uint32_t lb, int b) {
  if (ra != rb) return ra > rb;
  if (ha != hb) return ha > hb;
  if (la != lb) return la > lb;
  return a < b;
}

// set_ext (kernel.hpp:112-161): the extension code for votes low[b] (quality
// >= min_qual) and hi[b] (>= min_hi_qual), b = A C G T, at an integer depth.
MHMLA_HD int choose_ext(const uint32_t low[4], const uint32_t hi[4], int depth, const Rules& r) {
  double min_viable = r.min_viable_depth * depth;
  if (min_viable < 2.0) min_viable = 2.0;
  double min_expected = r.min_expected_depth * depth;
  if (min_expected < 2.0) min_expected = 2.0;
  int rt[4];
  for (int b = 0; b < 4; b++) rt[b] = rating(low[b], hi[b], min_viable, min_expected);
  // the first three of the ranking
  int o[4] = {0, 1, 2, 3};
  for (int i = 0; i < 3; i++)
    for (int j = 3; j > i; j--)
      if (ranks_before(rt[o[j]], hi[o[j]], low[o[j]], o[j], rt[o[j - 1]], hi[o[j - 1]], low[o[j - 1]], o[j - 1])) {
        const int t = o[j];
        o[j] = o[j - 1];
        o[j - 1] = t;
      }
  const int top = rt[o[0]], runner_up = rt[o[1]];
  if (top <= r.rating_thres) return kExtX;
  if (top <= 3) return runner_up == 0 ? o[0] : kExtX;
  if (top < 6) return runner_up < 3 ? o[0] : kExtX;
  if (top == 6) return runner_up < 4 ? o[0] : kExtX;
  if (runner_up < 7) return o[0];
  if (rt[o[2]] == 7 || low[o[0]] == low[o[1]]) return kExtF;
  return low[o[0]] > low[o[1]] ? o[0] : o[1];
}

MHMLA_HD char ext_char(int code) { return "ACGTFX"[code]; }

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
