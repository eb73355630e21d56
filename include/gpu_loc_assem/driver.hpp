// Declaration-compatible with mhm2's src/gpu_loc_assem/driver.hpp: the local
// assembly driver localassm.cpp calls (localassm.cpp:882, :897), implemented
// over the mhmla C ABI (include/mhmla.h) by libmhmsw.so.
#pragma once
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "locassem_struct.hpp"

namespace locassm_driver {

struct accum_data {
  std::vector<uint32_t> ht_sizes;
  std::vector<uint32_t> l_reads_count;
  std::vector<uint32_t> r_reads_count;
  std::vector<uint32_t> ctg_sizes;
};

struct ctg_bucket {
  std::vector<CtgWithReads> ctg_vec;
  accum_data sizes_vec;
  uint32_t l_max, r_max, max_contig_sz;
  ctg_bucket() : l_max{0}, r_max{0}, max_contig_sz{0} {}
  void clear();
};

// Reverse complement as the reference driver does it (driver.hpp:19-36):
// A C G T N map to T G C A N, the IUPAC codes U R Y K M S W B D H V to N;
// any other byte is reported and its output position left as it was.
inline void revcomp(char* str, char* str_rc, int size) {
  for (int i = 0; i < size; i++) {
    const char b = str[size - 1 - i];
    char o = 0;
    switch (b) {
      case 'A': o = 'T'; break;
      case 'C': o = 'G'; break;
      case 'G': o = 'C'; break;
      case 'T': o = 'A'; break;
      case 'N': case 'U': case 'R': case 'Y': case 'K': case 'M': case 'S': case 'W': case 'B': case 'D':
      case 'H': case 'V': o = 'N'; break;
      default: std::cout << "Illegal char:" << b << "\n"; break;
    }
    if (o) str_rc[i] = o;
  }
}

// String form (driver.hpp:38-58): a byte outside those is reported and dropped.
inline std::string revcomp(std::string instr) {
  std::string out;
  out.reserve(instr.size());
  for (size_t i = instr.size(); i-- > 0;) {
    switch (instr[i]) {
      case 'A': out += 'T'; break;
      case 'C': out += 'G'; break;
      case 'G': out += 'C'; break;
      case 'T': out += 'A'; break;
      case 'N': case 'U': case 'R': case 'Y': case 'K': case 'M': case 'S': case 'W': case 'B': case 'D':
      case 'H': case 'V': out += 'N'; break;
      default: std::cout << "Illegal char:" << instr[i] << "\n"; break;
    }
  }
  return out;
}

// Free device memory per rank sharing the device (driver.cpp:21-26).
size_t get_device_mem(int ranks_per_gpu, int gpu_id);

// Extends every contig of data_in in place: the left extension is prepended
// and the right one appended (driver.cpp:35-319).  mer_len is the first mer
// length; walk_len_limit, qual_offset as in the reference; the device is
// my_rank % devices (driver.cpp:39-41).  max_ctg_size, max_read_size,
// max_r_count, max_l_count, max_kmer_len and sizes_outliers only size the
// reference's buffers and are not needed here (buffers follow the data); the
// upper mer length is LASSM_MAX_KMER_LEN, as in the reference kernel.
// Errors throw std::runtime_error (the reference exits the process).
void local_assem_driver(std::vector<CtgWithReads>& data_in, uint32_t max_ctg_size, uint32_t max_read_size,
                        uint32_t max_r_count, uint32_t max_l_count, int mer_len, int max_kmer_len,
                        accum_data& sizes_outliers, int walk_len_limit, int qual_offset, int ranks, int my_rank,
                        int g_rank_me);

// Devices visible to this process, 0 on error (driver.cpp:28-33).
int get_gpu_per_node();

}  // namespace locassm_driver
