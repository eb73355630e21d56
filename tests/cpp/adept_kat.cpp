// The reference driver test's call sequence (test/ssw-test.cpp:240-431,
// TEST(MHMTest, AdeptSW)) compiled against the drop-in header: init with the
// tree's ALN_* scoring and rlen_limit 300, the 18 KAT pairs as (reads =
// queries, contigs = refs), forwards, block, backwards, block.  Prints
// "score read_begin read_end ctg_begin ctg_end" per pair.
//
// Other modes: "--device-name" prints get_device_name(0) (driver.cpp:35-39:
// the PCI bus id); "--narrowing" aligns a 40000-base read with itself, whose
// end coordinates do not fit AlignmentResults' short fields, and prints the
// error the driver throws.
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "adept_sw/driver.hpp"

static int narrowing() {
  adept_sw::GPUDriver drv;
  drv.init(0, 1, (short)1, (short)-3, (short)-5, (short)-2, 300);
  std::string seq;  // i.i.d. bases: the window occurs once in the read
  unsigned x = 12345;
  for (int i = 0; i < 40000; i++) {
    x = x * 1103515245u + 12345u;
    seq += "ACGT"[(x >> 16) & 3];
  }
  std::vector<std::string> reads{seq, "ACGTACGT"}, ctgs{seq.substr(33000, 5000), "ACGTACGT"};
  try {
    drv.run_kernel_forwards(reads, ctgs, 40000, 5000);
    drv.kernel_block();
    drv.run_kernel_backwards(reads, ctgs, 40000, 5000);
    drv.kernel_block();
  } catch (const std::runtime_error &e) {
    std::printf("%s", e.what());
    return 0;
  }
  std::printf("no error\n");
  return 1;
}

int main(int argc, char **argv) {
  if (argc > 1 && !std::strcmp(argv[1], "--device-name")) {
    std::printf("%s\n", adept_sw::get_device_name(0).c_str());
    return 0;
  }
  if (argc > 1 && !std::strcmp(argv[1], "--narrowing")) return narrowing();
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
