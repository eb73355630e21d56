// Layouts shared by the HIP kernels (sw_kernels.hip) and the host runtime (mhmsw.cpp).
#pragma once
#include <stdint.h>

namespace mhmsw {

// Sequence codes, one 4-bit nibble per base, 8 per dword, lowest nibble first:
// 0..3 = A,C,G,T, 4 = N (any non-ACGTU byte, as kBaseTranslation does in
// src/ssw/ssw.cpp:13-26).
constexpr int kCodeN = 4;

// Every sequence starts on a 32-base boundary of the unpacked buffers, after a
// front pad; kPadBases of slack sit before the first and after the last one so
// that kernels may read a few dwords outside a sequence.
constexpr int kSeqAlign = 32;
constexpr int64_t kPadBases = 1024;

// Classes of work (one persistent kernel launch each).
constexpr int kMaxP = 8;                 // packed registers per lane (2 rows each)
constexpr int kClassFast16 = 0;          // + (P-1): 16-lane groups, rows <= 32*P (see fast16_class_of)
constexpr int kClassFast64 = kMaxP;      // + (P-1): 64-lane groups, rows <= 128*P
constexpr int kClassPlain = 2 * kMaxP;   // 32-bit, one wavefront per pair
constexpr int kNumFwdClasses = 2 * kMaxP + 1;  // classes a forward pass uses
// Reverse pass only: banded form, one lane per pair, P = 1, 2, 4, 8 packed
// registers (4P diagonals) for |l - k| <= 1, 3, 7, 15: classes kClassBand + 0..3.
constexpr int kClassBand = kNumFwdClasses;
constexpr int kNumBandClasses = 6;
// |l - k| <= 31: four lanes of P = 4 per pair (64 diagonals)
constexpr int kClassBand2 = kClassBand + 4;
// |l - k| = 0 (D = 0, no gap can reach the score): one diagonal, one lane per pair
constexpr int kClassDiag = kClassBand + 5;
constexpr int kBandMaxD = 31;
constexpr int kNumClasses = kNumFwdClasses + kNumBandClasses;
constexpr int kNumBins = 128;            // column-length bins of 32 columns
constexpr int kNumBuckets = kNumClasses * kNumBins;

constexpr int kFast16MaxRows = 32 * kMaxP;   // 256

// The 16-lane class (0..7, as an offset from kClassFast16) of a pair with
// `rows` rows (1..256): rows in (32(k-1), 32k] for class k-1, except that the
// fifth class ends at 152 rows and the sixth holds 153..192.  The forward pass
// runs the fifth class as 4 lanes of 19 packed registers (152 rows), which
// 150/151-base reads fill; a 16-lane group of 5 registers (160 rows, the
// reverse pass's shape) still holds every row of it.
constexpr int fast16_class_of(int rows) {
  return rows <= 128 ? (rows + 31) / 32 - 1 : rows <= 152 ? 4 : rows <= 192 ? 5 : (rows + 31) / 32 - 1;
}
constexpr int kFast64MaxRows = 128 * kMaxP;  // 1024
constexpr int kFastMaxCols = 60000;          // step counter stays below 2^16
// A packed-form pair whose score reaches this is rerun in 32-bit (the packed
// kernels' values stay exact below 0x7C00 = 16 * 1984).
constexpr int32_t kFastOverflowScore = 1900;
// The same for the narrow forward shapes that keep 32 H (values exact below
// 0x7C00 = 32 * 992 while match + gap open <= 7).  The host gives them only
// classes whose rows times match stay below it, so none is ever rerun.
constexpr int32_t kFastOverflowScore32 = 960;
constexpr int kPlainMaxCols = 8192;          // LDS boundary row of the 32-bit kernel

// Task flags.
constexpr int32_t kTaskRowsAreRef = 1;  // rows hold SSW-ref (outer) positions
constexpr int32_t kTaskReverse = 2;     // result goes to the reverse (begin) table
constexpr int kTaskBandShift = 8;       // band tasks: (flags >> kTaskBandShift) & 0x1F = D
constexpr int kTaskScoreShift = 13;     // band tasks: flags >> kTaskScoreShift = the forward score

struct SwTask {
  int32_t pair;       // -1: empty slot
  int32_t flags;
  int64_t row_start;  // nibble index of row 0 in the code buffer
  int64_t col_start;  // nibble index of column 0
  int32_t row_len;
  int32_t col_len;
};

// Per-pair end point of one pass: the lexicographic optimum
// (max score, then least SSW-ref index i, then least SSW-query index j).
struct PairEnd {
  int32_t score;
  int32_t i;
  int32_t j;
  int32_t flags;  // kEndOverflow: the packed pass may have saturated
};
constexpr int32_t kEndOverflow = 1;

// Scheduling state of one pass (device memory).
struct PassSched {
  int32_t bucket_count[kNumBuckets];
  int32_t bucket_base[kNumBuckets];
  int32_t class_count[kNumClasses];   // real tasks
  int32_t class_base[kNumClasses];    // first slot
  int32_t class_slots[kNumClasses];   // padded slot count
  int32_t class_next[kNumClasses];    // persistent work counter
  int32_t npairs_routed;
};

// Final per-pair record, same layout as mhmsw_result in include/mhmsw.h.
struct AlnResult {
  int32_t score;
  int32_t read_begin, read_end;
  int32_t ctg_begin, ctg_end;
};

constexpr int32_t kEndWide = 2;  // the pair's forward pass ran in 32-bit (reverse must too)
constexpr int32_t kEndScore2 = 4;  // score2 must be recomputed in 32-bit (16-bit-pass padding differs)

// Per-batch scoring, pre-scaled for the packed kernel.
struct ScoreParams {
  int32_t match, mismatch, gap_open, gap_ext, ambiguity;
  uint32_t perm_lo, perm_hi;  // v_perm table of (s + gap_open) * 16 per class
  int32_t fast_ok;            // scoring fits the packed kernel
  int32_t read_is_ref;        // orientation
  int32_t bias;               // SSW's profile bias, max(0, mismatch, ambiguity)
};

// SSW pads its inner (query) sequence to whole vectors of 16 (8-bit pass) or
// 8 (16-bit pass, used once score + bias >= 255) positions that score 0
// against everything; its per-position maxima, and so score2, include them.
__host__ __device__ inline int ssw_query_pad(int n, int32_t score, int32_t bias) {
  const int v = score + bias >= 255 ? 8 : 16;
  return (v - n % v) % v;
}
constexpr int kMaxQueryPad = 15;

}  // namespace mhmsw
