// Calls locassm_driver::local_assem_driver the way mhm2's localassm.cpp does
// (localassm.cpp:882, :897), written against the reference's header, and
// prints every contig after extension.
//
// stdin: "n mer_len walk_len_limit qual_offset", then per contig
// "cid depth seq n_left n_right" and one "bases quals" line per read (left
// reads first); "-" stands for an empty string.  stdout: one line per
// contig, "cid extended_sequence".
#include <iostream>
#include <string>
#include <vector>

#include "gpu_loc_assem/driver.hpp"

static std::string field(const std::string& s) { return s == "-" ? std::string() : s; }

int main() {
  int64_t n;
  int mer_len, walk_len_limit, qual_offset;
  if (!(std::cin >> n >> mer_len >> walk_len_limit >> qual_offset)) return 2;
  std::vector<CtgWithReads> ctgs((size_t)n);
  uint32_t max_l = 0, max_r = 0, max_ctg = 0;
  for (auto& c : ctgs) {
    std::string seq;
    size_t nl, nr;
    std::cin >> c.cid >> c.depth >> seq >> nl >> nr;
    c.seq = field(seq);
    for (size_t side = 0; side < 2; side++)
      for (size_t i = 0; i < (side ? nr : nl); i++) {
        ReadSeq r;
        std::string b, q;
        std::cin >> b >> q;
        r.read_id = "r" + std::to_string(i);
        r.seq = field(b);
        r.quals = field(q);
        (side ? c.reads_right : c.reads_left).push_back(r);
      }
    c.max_reads = (unsigned)std::max(nl, nr);
    max_l = std::max<uint32_t>(max_l, (uint32_t)nl);
    max_r = std::max<uint32_t>(max_r, (uint32_t)nr);
    max_ctg = std::max<uint32_t>(max_ctg, (uint32_t)c.seq.size());
  }
  if (locassm_driver::get_gpu_per_node() <= 0) {
    std::cerr << "no GPU\n";
    return 3;
  }
  locassm_driver::accum_data sizes;
  try {
    locassm_driver::local_assem_driver(ctgs, max_ctg, 300, max_r, max_l, mer_len, 121, sizes, walk_len_limit,
                                       qual_offset, 1, 0, 0);
  } catch (const std::exception& e) {
    std::cerr << e.what() << "\n";
    return 4;
  }
  for (const auto& c : ctgs) std::cout << c.cid << " " << (c.seq.empty() ? "-" : c.seq) << "\n";
  return 0;
}
