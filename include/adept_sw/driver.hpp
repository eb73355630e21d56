// Drop-in declarations for mhm2's GPU Smith-Waterman driver.
//
// Declaration-compatible with src/adept-sw/driver.hpp of mhm2 (same namespace,
// types, member and free-function signatures), so klign.cpp (src/klign.cpp:356-388,
// 448-477) and test/ssw-test.cpp compile unchanged against it.  The
// implementation (mhm2_staging_amd/csrc/driver_shim.cpp) runs on the mhmsw
// C ABI (include/mhmsw.h) and returns SSW-exact results.
#pragma once

#include <chrono>
#include <cmath>
#include <string>
#include <vector>

#define NSTREAMS 2

#ifndef KLIGN_GPU_BLOCK_SIZE
#define KLIGN_GPU_BLOCK_SIZE 20000
#endif

namespace adept_sw {

// Results of the last batch, one entry per pair; "ref" is the contig window
// and "query" the read, 0-based inclusive (src/adept-sw/driver.hpp:17-23).
struct AlignmentResults {
  short *ref_begin = nullptr;
  short *query_begin = nullptr;
  short *ref_end = nullptr;
  short *query_end = nullptr;
  short *top_scores = nullptr;
};

size_t get_tot_gpu_mem();
size_t get_avail_gpu_mem_per_rank(int totRanks, int numDevices = 0);
std::string get_device_name(int device_id);
int get_num_node_gpus();

bool initialize_gpu();
bool initialize_gpu(double &time_to_initialize, int &device_count, size_t &total_mem);

struct DriverState;

class GPUDriver {
  DriverState *driver_state = nullptr;
  AlignmentResults alignments;

 public:
  ~GPUDriver();

  // returns the time to execute; penalties are passed negative, as klign does
  double init(int upcxx_rank_me, int upcxx_rank_n, short match_score, short mismatch_score, short gap_opening_score,
              short gap_extending_score, int rlen_limit);
  void run_kernel_forwards(std::vector<std::string> &reads, std::vector<std::string> &contigs, unsigned maxReadSize,
                           unsigned maxContigSize);
  void run_kernel_backwards(std::vector<std::string> &reads, std::vector<std::string> &contigs, unsigned maxReadSize,
                            unsigned maxContigSize);
  bool kernel_is_done();
  void kernel_block();

  AlignmentResults &get_aln_results() { return alignments; }

  // Additions (no counterpart in the reference header).  init() takes no
  // ambiguity cost; it defaults to ALN_AMBIGUITY_COST when that macro is
  // defined, else 1 (CMakeDefinitions.txt:117).  Call after init().
  void set_ambiguity(short ambiguity_cost);
  // Orientation of SSW's roles; default true = klign's CPU call (read is SSW-ref).
  void set_read_is_ssw_ref(bool read_is_ref);
  // SSW's second-best score (Alignment::sw_score_next_best) per pair, which
  // klign's GPU path sets to 0 (src/klign.cpp:378-379).  Enable before
  // run_kernel_forwards; get_score2() is valid after the second kernel_block
  // of a batch run with it (nullptr otherwise).
  void set_score2(bool enable);
  const int *get_score2();
  // SSW's CIGAR string and mismatch count per pair (Alignment::cigar_string
  // and ::mismatches of klign's call with report_cigar set,
  // src/klign.cpp:251-273), which klign cannot take from ADEPT (it keeps such
  // batches on the CPU, src/klign.cpp:589-599, and its GPU mapping writes
  // "*", :385).  Enable before run_kernel_forwards; the getters are valid
  // after the second kernel_block of a batch run with it (nullptr otherwise).
  void set_cigar(bool enable);
  const std::string *get_cigar_strings();
  const int *get_mismatches();
};

}  // namespace adept_sw
