// Declaration-compatible with mhm2's src/gpu_loc_assem/locassem_struct.hpp:
// a contig with the reads that hang over its ends, as local_assem_driver
// (driver.hpp) takes them.
#pragma once
#ifndef UPCXX_SERIALIZED_FIELDS
#define UPCXX_SERIALIZED_FIELDS(...)
#endif
#include <cstdint>
#include <iostream>
#include <string>
#include <vector>

struct ReadSeq {
  std::string read_id;
  std::string seq;
  std::string quals;
  UPCXX_SERIALIZED_FIELDS(read_id, seq, quals);
};

struct CtgWithReads {
  int64_t cid;
  std::string seq;
  double depth;
  unsigned max_reads;
  std::vector<ReadSeq> reads_left;
  std::vector<ReadSeq> reads_right;
};
