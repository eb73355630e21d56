// klign_call_bench: the cost of one klign-sized call through the drop-in
// adept_sw::GPUDriver, as klign makes it (src/klign.cpp:356-388): a new batch
// of std::string reads and contig windows per call, run_kernel_forwards ->
// kernel_block -> run_kernel_backwards -> kernel_block -> get_aln_results.
//
// usage: klign_call_bench [--calls N] [--npairs P] [--batches B] [--warmup W]
//                         [--same-geometry 0|1]
// The batches are the seeded plain workload (150-base reads against 350-base
// windows, 1% substitutions, 0.1% indels; seeds 101, 102, ...), cycled so that
// every call hands the driver a batch other than the previous one.  Batch k
// holds P - k pairs, so that no two batches have the same geometry and every
// call launches its job afresh, as klign's blocks (of varying lengths) do;
// with --same-geometry 1 they all hold P pairs of equal lengths, and the
// driver replays the job it captured as a HIP graph.  Prints one
// JSON line: the median wall time per call and its parts, and the rate in
// GCUPS (forward cells per second of wall time).
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "adept_sw/driver.hpp"
#include "mhmsw_synth.h"

namespace {

double now_ms() {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

double median(std::vector<double> v) {
  std::sort(v.begin(), v.end());
  return v.empty() ? 0.0 : v[v.size() / 2];
}

struct Batch {
  std::vector<std::string> reads, ctgs;
  unsigned max_read = 0, max_ctg = 0;
  double cells = 0;
};

}  // namespace

int main(int argc, char** argv) {
  int calls = 40, batches = 4, warmup = 8, same = 0;
  long long npairs = 20000;
  for (int i = 1; i + 1 < argc; i += 2) {
    if (!std::strcmp(argv[i], "--calls")) {
      calls = std::atoi(argv[i + 1]);
    } else if (!std::strcmp(argv[i], "--npairs")) {
      npairs = std::atoll(argv[i + 1]);
    } else if (!std::strcmp(argv[i], "--batches")) {
      batches = std::atoi(argv[i + 1]);
    } else if (!std::strcmp(argv[i], "--warmup")) {
      warmup = std::atoi(argv[i + 1]);
    } else if (!std::strcmp(argv[i], "--same-geometry")) {
      same = std::atoi(argv[i + 1]);
    } else {
      std::fprintf(stderr, "unknown option %s\n", argv[i]);
      return 2;
    }
  }
  if (calls < 1 || batches < 1 || npairs < batches || warmup < 0) {
    std::fprintf(stderr, "bad arguments\n");
    return 2;
  }
  std::vector<Batch> bs((size_t)batches);
  for (int k = 0; k < batches; k++) {
    mhmsw_synth_batch sb;
    if (mhmsw_synth_generate(1, same ? npairs : npairs - k, (uint64_t)(101 + k), 0, &sb) != 0) {
      std::fprintf(stderr, "generator failed\n");
      return 1;
    }
    Batch& b = bs[(size_t)k];
    for (int64_t p = 0; p < sb.npairs; p++) {
      b.reads.emplace_back(sb.reads + sb.read_off[p], (size_t)sb.read_len[p]);
      b.ctgs.emplace_back(sb.ctgs + sb.ctg_off[p], (size_t)sb.ctg_len[p]);
      b.max_read = std::max(b.max_read, (unsigned)sb.read_len[p]);
      b.max_ctg = std::max(b.max_ctg, (unsigned)sb.ctg_len[p]);
      b.cells += (double)sb.read_len[p] * (double)sb.ctg_len[p];
    }
    mhmsw_synth_free(&sb);
  }
  try {
    adept_sw::GPUDriver drv;
    // klign's init: penalties negative (src/klign.cpp:463-465); mhm2's 1/3/5/2
    // with ambiguity 2
    drv.init(0, 1, 1, -3, -5, -2, (int)bs[0].max_read);
    drv.set_ambiguity(2);
    std::vector<short> first_scores;
    std::vector<double> wall, fwd, blk1, rest;
    long long checksum = 0;
    for (int i = 0; i < warmup + calls; i++) {
      Batch& b = bs[(size_t)(i % batches)];
      const double t0 = now_ms();
      drv.run_kernel_forwards(b.reads, b.ctgs, b.max_read, b.max_ctg);
      const double t1 = now_ms();
      drv.kernel_block();
      const double t2 = now_ms();
      drv.run_kernel_backwards(b.reads, b.ctgs, b.max_read, b.max_ctg);
      drv.kernel_block();
      adept_sw::AlignmentResults& r = drv.get_aln_results();
      // read every result, as klign's mapping loop does
      long long s = 0;
      for (size_t p = 0; p < b.reads.size(); p++)
        s += r.top_scores[p] + r.ref_begin[p] + r.ref_end[p] + r.query_begin[p] + r.query_end[p];
      const double t3 = now_ms();
      checksum += s;
      if (i == 0) first_scores.assign(r.top_scores, r.top_scores + b.reads.size());
      if (i % batches == 0 && i > 0 &&
          !std::equal(first_scores.begin(), first_scores.end(), r.top_scores)) {
        std::fprintf(stderr, "results of a repeated batch differ at call %d\n", i);
        return 1;
      }
      if (i >= warmup) {
        wall.push_back(t3 - t0);
        fwd.push_back(t1 - t0);
        blk1.push_back(t2 - t1);
        rest.push_back(t3 - t2);
      }
    }
    double cells = 0;
    for (const Batch& b : bs) cells += b.cells;
    cells /= batches;
    const double w = median(wall);
    std::printf("{\"workload\": \"%lld pairs per call, 150 bp reads vs 350 bp windows, a new batch every call%s, "
                "std::string in and short results out through adept_sw::GPUDriver\", \"calls\": %d, "
                "\"ms_per_call\": %.4f, \"value\": %.1f, \"unit\": \"GCUPS\", \"run_kernel_forwards_ms\": %.4f, "
                "\"first_kernel_block_ms\": %.4f, \"backwards_and_results_ms\": %.4f, \"checksum\": %lld}\n",
                npairs, same ? " (equal geometry: graph replays)" : "", calls, w, cells / (w * 1e-3) / 1e9,
                median(fwd), median(blk1), median(rest), checksum);
  } catch (const std::exception& e) {
    std::fprintf(stderr, "%s", e.what());
    return 1;
  }
  return 0;
}
