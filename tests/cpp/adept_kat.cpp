// The reference driver test's call sequence (test/ssw-test.cpp:240-431,
// TEST(MHMTest, AdeptSW)) compiled against the drop-in header: init with the
// tree's ALN_* scoring and rlen_limit 300, the 18 KAT pairs as (reads =
// queries, contigs = refs), forwards, block, backwards, block.  Prints
// "score read_begin read_end ctg_begin ctg_end" per pair.
#include <cstdio>
#include <string>
#include <vector>

#include "adept_sw/driver.hpp"

int main() {
  double t;
  int ndev;
  size_t mem;
  adept_sw::initialize_gpu(t, ndev, mem);
  if (ndev <= 0) {
    std::fprintf(stderr, "no GPU\n");
    return 2;
  }
  adept_sw::GPUDriver drv;
  drv.init(0, 1, (short)1, (short)-1, (short)-1, (short)-1, 300);
  const char *cases[18][2] = {
      {"ACGT", "ACGT"}, {"ACGT", "AACGT"}, {"ACGT", "ACGTT"}, {"TACGT", "ACGT"}, {"TTACGT", "ACGT"},
      {"ACGTT", "ACGT"}, {"ACGTTT", "ACGT"}, {"TACGTT", "ACGT"}, {"TTACGTT", "ACGT"}, {"TACGTTT", "ACGT"},
      {"TTACGTTT", "ACGT"}, {"AAAATTTTCCCCGGGG", "AAAATTTTCCCCGGGG"}, {"AAAATTTTACCCGGGG", "AAAATTTTCCCCGGGG"},
      {"AAAATTTTACCCCGGGG", "AAAATTTTCCCCGGGG"}, {"AAAATTTCCCCGGGG", "AAAATTTTCCCCGGGG"},
      {"GCTAGCTAGCTAGCTA", "AAAATTTTCCCCGGGG"}, {"GCTAAAATTTTCCCCGGGG", "AAAATTTTCCCCGGGG"},
      {"AAAATTTTCCCCGGGGACT", "AAAATTTTCCCCGGGG"}};
  std::vector<std::string> queries, refs;
  unsigned maxq = 0, maxr = 0;
  for (auto &c : cases) {
    queries.emplace_back(c[0]);
    refs.emplace_back(c[1]);
    if (queries.back().size() > maxq) maxq = queries.back().size();
    if (refs.back().size() > maxr) maxr = refs.back().size();
  }
  drv.run_kernel_forwards(queries, refs, maxq, maxr);
  drv.kernel_block();
  drv.run_kernel_backwards(queries, refs, maxq, maxr);
  drv.kernel_block();
  adept_sw::AlignmentResults &r = drv.get_aln_results();
  for (size_t i = 0; i < queries.size(); i++)
    std::printf("%d %d %d %d %d\n", r.top_scores[i], r.query_begin[i], r.query_end[i], r.ref_begin[i], r.ref_end[i]);
  return 0;
}
