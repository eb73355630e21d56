// locassm_driver (include/gpu_loc_assem/driver.hpp, declaration-compatible
// with mhm2's src/gpu_loc_assem/driver.hpp) over the mhmla C ABI.
//
// The reference driver packs the contigs of a slice, runs its kernel on the
// right sides, reverse-complements the contigs on the host, runs it again on
// the left sides and stitches the walks onto the contigs
// (src/gpu_loc_assem/driver.cpp:141-306).  Here the whole vector is packed
// once and both sides run in the same launches; the stitching is the same:
// the left extension (already reverse-complemented) goes in front, the right
// one at the end.
#include <hip/hip_runtime.h>

#include <map>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/gpu_loc_assem/driver.hpp"
#include "../../include/mhmla.h"

namespace {

std::mutex g_mu;
std::map<int, mhmla_ctx*> g_ctx;  // one context per device, kept for the process

mhmla_ctx* context_for(int device) {
  std::lock_guard<std::mutex> lk(g_mu);
  auto it = g_ctx.find(device);
  if (it != g_ctx.end()) return it->second;
  mhmla_ctx* c = nullptr;
  const int rc = mhmla_create(device, &c);
  if (rc != MHMLA_OK) throw std::runtime_error("GPUassert: local assembly context on device " + std::to_string(device) +
                                               " failed (" + std::to_string(rc) + ")");
  g_ctx[device] = c;
  return c;
}

}  // namespace

namespace locassm_driver {

void ctg_bucket::clear() {
  ctg_vec.clear();
  sizes_vec = accum_data{};
  l_max = r_max = max_contig_sz = 0;
}

size_t get_device_mem(int ranks_per_gpu, int gpu_id) {
  size_t free_b = 0, total_b = 0;
  if (hipSetDevice(gpu_id) != hipSuccess || hipMemGetInfo(&free_b, &total_b) != hipSuccess) return 0;
  return ranks_per_gpu > 0 ? free_b / (size_t)ranks_per_gpu : free_b;
}

int get_gpu_per_node() { return mhmla_device_count(); }

void local_assem_driver(std::vector<CtgWithReads>& data_in, uint32_t /*max_ctg_size*/, uint32_t /*max_read_size*/,
                        uint32_t /*max_r_count*/, uint32_t /*max_l_count*/, int mer_len, int /*max_kmer_len*/,
                        accum_data& /*sizes_outliers*/, int walk_len_limit, int qual_offset, int /*ranks*/,
                        int my_rank, int /*g_rank_me*/) {
  const int ngpu = mhmla_device_count();
  if (ngpu <= 0) throw std::runtime_error("GPUassert: no GPU visible for local assembly");
  mhmla_ctx* ctx = context_for(my_rank % ngpu);

  mhmla_params p;
  mhmla_default_params(&p);
  p.kmer_len = mer_len;
  p.walk_len_limit = walk_len_limit;
  p.qual_offset = qual_offset;
#ifdef LASSM_MIN_KMER_LEN
  p.min_kmer_len = LASSM_MIN_KMER_LEN;
#endif
#ifdef LASSM_MAX_KMER_LEN
  p.max_kmer_len = LASSM_MAX_KMER_LEN;
#endif
#ifdef LASSM_SHIFT_SIZE
  p.shift = LASSM_SHIFT_SIZE;
#endif
#ifdef LASSM_MIN_QUAL
  p.min_qual = LASSM_MIN_QUAL;
#endif
#ifdef LASSM_MIN_HI_QUAL
  p.min_hi_qual = LASSM_MIN_HI_QUAL;
#endif
#ifdef LASSM_MIN_VIABLE_DEPTH
  p.min_viable_depth = LASSM_MIN_VIABLE_DEPTH;
#endif
#ifdef LASSM_MIN_EXPECTED_DEPTH
  p.min_expected_depth = LASSM_MIN_EXPECTED_DEPTH;
#endif
#ifdef LASSM_RATING_THRES
  p.rating_thres = LASSM_RATING_THRES;
#endif
#ifdef LASSM_MAX_COUNT_MERS_READS
  p.max_reads = LASSM_MAX_COUNT_MERS_READS;
#endif

  const int64_t n = (int64_t)data_in.size();
  std::string ctgs, seqs, quals;
  std::vector<int64_t> ctg_off((size_t)n), left_first((size_t)n + 1), right_first((size_t)n + 1), read_off;
  std::vector<int32_t> ctg_len((size_t)n), read_len;
  std::vector<double> depth((size_t)n);
  std::vector<int64_t> left_idx, right_idx;
  auto add_read = [&](const ReadSeq& r) {
    if (r.quals.size() < r.seq.size())
      throw std::runtime_error("GPUassert: a read has fewer quality bytes than bases");
    read_off.push_back((int64_t)seqs.size());
    read_len.push_back((int32_t)r.seq.size());
    seqs += r.seq;
    quals.append(r.quals, 0, r.seq.size());
    return (int64_t)read_len.size() - 1;
  };
  for (int64_t c = 0; c < n; c++) {
    const CtgWithReads& ctg = data_in[(size_t)c];
    ctg_off[(size_t)c] = (int64_t)ctgs.size();
    ctg_len[(size_t)c] = (int32_t)ctg.seq.size();
    ctgs += ctg.seq;
    depth[(size_t)c] = ctg.depth;
    left_first[(size_t)c] = (int64_t)left_idx.size();
    for (const ReadSeq& r : ctg.reads_left) left_idx.push_back(add_read(r));
    right_first[(size_t)c] = (int64_t)right_idx.size();
    for (const ReadSeq& r : ctg.reads_right) right_idx.push_back(add_read(r));
  }
  left_first[(size_t)n] = (int64_t)left_idx.size();
  right_first[(size_t)n] = (int64_t)right_idx.size();
  // one read_idx array: left lists first, right lists after them
  std::vector<int64_t> read_idx(left_idx);
  read_idx.insert(read_idx.end(), right_idx.begin(), right_idx.end());
  for (auto& f : right_first) f += (int64_t)left_idx.size();

  const int64_t W = walk_len_limit > 0 ? walk_len_limit : 0;
  std::vector<int32_t> ll((size_t)n), rl((size_t)n);
  std::vector<char> le((size_t)(n * W + 1)), re((size_t)(n * W + 1));
  const int rc = mhmla_extend(ctx, &p, n, ctgs.data(), ctg_off.data(), ctg_len.data(), depth.data(), left_first.data(),
                              right_first.data(), read_idx.data(), (int64_t)read_len.size(), seqs.data(), quals.data(),
                              read_off.data(), read_len.data(), ll.data(), le.data(), rl.data(), re.data(), nullptr);
  if (rc != MHMLA_OK) throw std::runtime_error(std::string("GPUassert: local assembly: ") + mhmla_last_error(ctx));
  for (int64_t c = 0; c < n; c++) {
    std::string& s = data_in[(size_t)c].seq;
    if (ll[(size_t)c] > 0) s.insert(0, le.data() + c * W, (size_t)ll[(size_t)c]);
    if (rl[(size_t)c] > 0) s.append(re.data() + c * W, (size_t)rl[(size_t)c]);
  }
}

}  // namespace locassm_driver
