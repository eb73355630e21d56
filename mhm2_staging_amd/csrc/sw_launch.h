// Host-callable launchers of the kernels in sw_kernels.hip.
#pragma once
#include <hip/hip_runtime_api.h>
#include <stdint.h>

#include "sw_common.h"

namespace mhmsw {

constexpr int kPassFwd = 0;    // forward pass: end points
constexpr int kPassRerun = 1;  // 32-bit forward rerun of pairs the packed pass flagged
constexpr int kPassRev = 2;    // reverse pass: begin points
constexpr int kPassRev2 = 3;   // large jobs: full wide-band rectangles whose probe did not reach the score
constexpr int kNumPasses = 4;
// Rows of a wide-band pair's probe rectangle in the reverse pass of a large job
constexpr int kProbeRows = 32;

struct BatchView {
  int64_t npairs;
  int64_t total;            // padded bases of the layout (codes at [0, total)); reverse tasks read them backwards
  const int64_t* seq_base;  // [2*npairs] first base of sequence s (read of pair p: 2p, contig window: 2p+1)
  const int32_t* seq_len;   // [2*npairs]
  const uint32_t* packed2;  // 2-bit codes, 16 per dword
  const uint32_t* nmask;    // N flags, 32 per dword
  uint32_t* codes;          // nibble codes, 8 per dword, total nibbles plus slack
  PairEnd* fwd;
  PairEnd* rev;
  PassSched* sched;         // kNumPasses of them, one per pass (with_pass selects a pass's)
  SwTask* tasks;
  int32_t* bucket;  // [npairs]
  int32_t* rank;    // [npairs]
  AlnResult* results;
  int32_t* errors;  // [0] reverse score mismatches, [1] pairs rerun in 32-bit
  unsigned long long* counters;  // [0] cells of the reverse pass
  ScoreParams sc;
  int32_t mode;     // 0 fast, 1 plain
  // windowed batches: contig window of pair p = resident contig win_ctg[p] at win_off[p]
  const uint32_t* res2;      // resident contigs, 2-bit codes (null for string batches)
  const uint32_t* resN;      // resident contigs, N bitmap
  const int64_t* res_base;   // first base of each resident contig
  const int32_t* win_ctg;
  const int32_t* win_off;
  int32_t g64_min_rows;  // shorter side at or above this: 64-lane groups
  int32_t bin_shift;     // column bucket width 1 << bin_shift
  // second-best score (SSW's sw_score_next_best with klign's maskLen): forward
  // rows are then always SSW-ref positions
  int32_t score2;
  int32_t* s2;           // [npairs]
  // reverse pass of a small job: every 16-lane (64-lane) rectangle task goes
  // to the class of P = rev_one16 (rev_one64), so that one launch per group
  // size runs them; 0: each task's class follows its rows
  int32_t rev_one16, rev_one64;
  // large jobs: a wide-band pair's reverse rectangle is first probed on its
  // first kProbeRows rows; pass kPassRev2 runs the full rectangle of the
  // pairs whose probe did not reach the score, all in the class of P =
  // rev2_one16 (rev2_one64 for 64-lane groups)
  int32_t rev_probe, rev2_one16, rev2_one64;
  // widest band (|l - k| <= band_max_d, at most kBandMaxD) a reverse task runs
  // in the one-lane band kernels; wider ones run as rectangles
  int32_t band_max_d;
  // forward launches of the packed kernel whose narrow shapes (at most 8
  // lanes per pair) keep H as 32 H and fold their row maxima every 32 steps:
  // set per launch when match + gap open <= 7 and the class's rows times
  // match stay below kFastOverflowScore32; 0 otherwise
  int32_t fold32;
};

// CIGAR pass over the results of a finished job (cigar_kernels.hip).
struct CigarView {
  int64_t npairs;
  const uint32_t* codes;    // nibble codes of the job (forward region)
  const int64_t* seq_base;  // [2*npairs] as BatchView
  const int32_t* seq_len;
  const AlnResult* results;
  ScoreParams sc;
  int32_t* lists;       // [kCigClasses][npairs] pairs of each band class
  int32_t* list_count;  // [kCigClasses] entries, then work counters
  int32_t* band;        // [npairs] band half-width of the pair's next attempt
  int32_t* max_rows;    // [1] tallest rectangle of the register classes
  uint32_t* dirs;       // direction scratch, one slot per wavefront
  int64_t dirs_dwords;
  uint32_t* gen;        // pool of the wide class
  int64_t gen_dwords;
  int32_t* n_ops;       // [npairs] op count (-1 op pool full, -2 deferred, -3 traceback error)
  int32_t* mism;        // [npairs]
  int64_t* off;         // [npairs] first op in ops
  uint32_t* ops;        // BAM-encoded ops
  int64_t ops_cap;
  unsigned long long* used;  // [0] ops reserved, [1] pool dwords reserved, [2] deferred pairs, [3] errors
};
// narrow_waves: wavefronts of each narrow-band launch; wide_waves: of the wide one
hipError_t launch_cigar(const CigarView& v, int narrow_waves, int wide_waves, hipStream_t st);

// The view of pass `pass`: the same batch with that pass's scheduling state.
// Kernels of a pass (classify, schedule, scatter and the work classes) take it.
inline BatchView with_pass(const BatchView& b, int pass) {
  BatchView v = b;
  v.sched = b.sched + pass;
  return v;
}

// error words, counters and the scheduling state of every pass (b: the view with the first PassSched)
hipError_t launch_clear_job(const BatchView& b, hipStream_t st);
hipError_t launch_unpack(const BatchView& b, hipStream_t st);           // 2-bit + N mask -> nibble codes
hipError_t launch_extract_windows(const BatchView& b, hipStream_t st);  // windows of resident contigs (if any)
hipError_t launch_schedule(const BatchView& b, int pass, hipStream_t st);
// The packed kernel that serves class (group, P) of a pass over `pairs`
// pairs on `cus` CUs whose rows are at most max_rows: `lanes` lanes per pair
// with `regs` packed registers each (lanes * regs = group * P, or fewer
// registers when max_rows allows)
struct FastShape {
  int lanes, regs;
};
FastShape fast_shape(int group, int P, bool forward, int64_t pairs, int cus, int max_rows);
hipError_t launch_fast(const BatchView& b, FastShape s, int grid, hipStream_t st);
// reverse pass, banded form: P = 1, 2, 4, 8 (|l - k| <= 2P - 1), 16 (four lanes
// of P = 4 per pair, |l - k| <= 31), kBandDiag (the single-diagonal class,
// |l - k| = 0) or 0 (every class up to P = 8 in one launch)
constexpr int kBandDiag = -1;
hipError_t launch_band(const BatchView& b, int P, int grid, hipStream_t st);
hipError_t launch_plain(const BatchView& b, int grid, int max_cols, hipStream_t st);
hipError_t launch_finalize(const BatchView& b, hipStream_t st);
// 256-thread blocks of a kernel that fit one CU at once (its registers decide)
hipError_t fast_kernel_occupancy(FastShape s, int* blocks_per_cu);
// the packed kernel's instantiations, built in parts (sw_fast_<G>_<P0>.hip:
// G lanes per pair, a list of P from P0); launch_fast and
// fast_kernel_occupancy dispatch to them
hipError_t launch_fast_2_8(const BatchView& b, int P, int grid, hipStream_t st);
hipError_t launch_fast_4_12(const BatchView& b, int P, int grid, hipStream_t st);
hipError_t launch_fast_4_16(const BatchView& b, int P, int grid, hipStream_t st);
hipError_t launch_fast_4_19(const BatchView& b, int P, int grid, hipStream_t st);
hipError_t launch_fast_8_12(const BatchView& b, int P, int grid, hipStream_t st);
hipError_t launch_fast_8_14(const BatchView& b, int P, int grid, hipStream_t st);
hipError_t launch_fast_16_1(const BatchView& b, int P, int grid, hipStream_t st);
hipError_t launch_fast_16_5(const BatchView& b, int P, int grid, hipStream_t st);
hipError_t launch_fast_16_12(const BatchView& b, int P, int grid, hipStream_t st);
hipError_t launch_fast_16_16(const BatchView& b, int P, int grid, hipStream_t st);
hipError_t launch_fast_16_20(const BatchView& b, int P, int grid, hipStream_t st);
hipError_t launch_fast_32_12(const BatchView& b, int P, int grid, hipStream_t st);
hipError_t launch_fast_32_14(const BatchView& b, int P, int grid, hipStream_t st);
hipError_t launch_fast_64_1(const BatchView& b, int P, int grid, hipStream_t st);
hipError_t launch_fast_64_5(const BatchView& b, int P, int grid, hipStream_t st);
hipError_t occupancy_fast_2_8(int P, int* blocks_per_cu);
hipError_t occupancy_fast_4_12(int P, int* blocks_per_cu);
hipError_t occupancy_fast_4_16(int P, int* blocks_per_cu);
hipError_t occupancy_fast_4_19(int P, int* blocks_per_cu);
hipError_t occupancy_fast_8_12(int P, int* blocks_per_cu);
hipError_t occupancy_fast_8_14(int P, int* blocks_per_cu);
hipError_t occupancy_fast_16_1(int P, int* blocks_per_cu);
hipError_t occupancy_fast_16_5(int P, int* blocks_per_cu);
hipError_t occupancy_fast_16_12(int P, int* blocks_per_cu);
hipError_t occupancy_fast_16_16(int P, int* blocks_per_cu);
hipError_t occupancy_fast_16_20(int P, int* blocks_per_cu);
hipError_t occupancy_fast_32_12(int P, int* blocks_per_cu);
hipError_t occupancy_fast_32_14(int P, int* blocks_per_cu);
hipError_t occupancy_fast_64_1(int P, int* blocks_per_cu);
hipError_t occupancy_fast_64_5(int P, int* blocks_per_cu);
hipError_t band_kernel_occupancy(int P, int* blocks_per_cu);  // P as launch_band

}  // namespace mhmsw
