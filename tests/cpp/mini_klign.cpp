// mini-klign: mhm2's klign batching loop as a test harness over the drop-in
// adept_sw::GPUDriver.
//
// What it mirrors (src/klign.cpp of mhm2):
//   compute_alns_for_read window rule ........ :672-739 (seed -> window, orientation)
//   align_read (perfect-match shortcut, dispatch rule) ... :218-249
//   ssw_align_read (CPU SSW call and Aln mapping) ........ :251-273
//   kernel_align_block (CPU work-steal while the previous
//     GPU block runs, then dispatch) ..................... :552-606
//   flush_remaining ...................................... :608-621
//   _gpu_align_block_kernel (driver protocol, Aln mapping,
//     score2 = mismatches = 0) ........................... :356-388
// With --gpu-score2 the GPU side fills score2 from the driver's second-best
// scores (GPUDriver::set_score2 / get_score2) instead of klign's 0.
// With --report-cigar the run is klign's CIGAR-reporting one (its Filter's
// report_cigar, klign.cpp:435): every record gets set_sam_string's SAM line
// (:178-209), the perfect-match shortcut writes "<n>=" (:229) and the CPU
// side takes SSW's cigar_string and mismatches (:257-271).  klign then keeps
// blocks off the GPU (:589-599, the block goes to ssw_align_block); with
// --gpu-cigar as well, blocks go to the drop-in driver, which returns the
// same CIGARs, mismatches and score2 (set_cigar / set_score2).
//   switch_orient ........................................ src/utils.cpp:163-167
//   Aln fields and to_string ............................. include/alignments.hpp:58-72,
//                                                           src/alignments.cpp:96-103
// The k-mer seeding that feeds compute_alns_for_read (a distributed hash table)
// is replaced by the known source position of each synthetic read.
//
// The CPU side of the work-steal is the reference SSW (the library klign
// calls), loaded at run time from a path given on the command line; when
// that library is absent the scalar restatement of SSW's semantics is used
// and score2 is reported as -1 (unknown) for CPU-aligned pairs.
//
// usage: mini_klign --mode gpu|cpu --out FILE [--seed S] [--reads N]
//          [--block B] [--mem BYTES] [--no-steal] [--gpu-delay-ms D]
//          [--scoring m,x,o,e,a] [--ssw LIB] [--oracle LIB] [--gpu-score2]
//          [--report-cigar] [--gpu-cigar]
// Writes one line per alignment (Aln::to_string, then identity, mismatches,
// the side that aligned it: P perfect shortcut, G GPU, C CPU, and the SAM
// line when CIGARs are reported) and prints a JSON summary on stdout.
#include <dlfcn.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <future>
#include <mutex>
#include <sstream>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "adept_sw/driver.hpp"

namespace {

struct AlnScoring {
  int match, mismatch, gap_opening, gap_extending, ambiguity;
};

// Field set of mhm2's Aln (include/alignments.hpp:58-72) plus the side that
// produced it.
struct Aln {
  std::string read_id;
  int64_t cid = -1;
  int cstart = 0, cstop = 0, clen = 0;
  int rstart = 0, rstop = 0, rlen = 0;
  int score1 = 0, score2 = 0;
  int mismatches = 0;
  int16_t read_group_id = -1;
  char orient = 0;
  int8_t identity = 0;
  char side = '?';
  int64_t read_idx = 0;
  std::string sam_string;

  Aln() = default;
  Aln(const std::string &rid, int64_t cid_, int rstart_, int rstop_, int rlen_, int cstart_, int cstop_, int clen_,
      char orient_, int score1_ = 0, int score2_ = 0, int identity_ = 0, int mismatches_ = 0, int rg = -1)
      : read_id(rid), cid(cid_), cstart(cstart_), cstop(cstop_), clen(clen_), rstart(rstart_), rstop(rstop_),
        rlen(rlen_), score1(score1_), score2(score2_), mismatches(mismatches_), read_group_id((int16_t)rg),
        orient(orient_), identity((int8_t)identity_) {}

  std::string to_string() const {
    std::ostringstream os;
    os << read_id << "\t" << rstart + 1 << "\t" << rstop << "\t" << rlen << "\t"
       << "Contig" << cid << "\t" << cstart + 1 << "\t" << cstop << "\t" << clen << "\t"
       << (orient == '+' ? "Plus" : "Minus") << "\t" << score1 << "\t" << score2;
    return os.str();
  }
};

std::string revcomp(const std::string &s);

void switch_orient(int &start, int &stop, int &len) {
  int tmp = start;
  start = len - stop;
  stop = len - tmp;
}

// ---- CPU aligner (loaded at run time) ----
struct CScoring {
  int32_t match, mismatch, gap_open, gap_ext, ambiguity;
};
typedef int (*ssw_one_fn)(const char *, const char *, int32_t, const CScoring *, int32_t *);
typedef int (*ssw_one_cigar_fn)(const char *, const char *, int32_t, const CScoring *, int32_t *, char *, int32_t);
typedef int (*oracle_cigar_fn)(int64_t, const char *, const int64_t *, const int32_t *, const char *,
                               const int64_t *, const int32_t *, const CScoring *, int32_t *, int32_t *, int32_t *,
                               uint32_t *, int, int);
typedef int (*oracle_batch_fn)(int64_t, const char *, const int64_t *, const int32_t *, const char *,
                               const int64_t *, const int32_t *, const CScoring *, int32_t *, int);

struct CpuAligner {
  ssw_one_fn ssw_one = nullptr;
  ssw_one_cigar_fn ssw_one_cigar = nullptr;
  oracle_batch_fn oracle = nullptr;
  oracle_cigar_fn oracle_cigar = nullptr;
  CScoring sc{};

  // The same call with report_cigar set: out as align(), plus cigar_string.
  void align_cigar(const std::string &query, const std::string &ref, int32_t out[7], std::string &cigar) const {
    if (ssw_one_cigar) {
      std::vector<char> buf(2 * (query.size() + ref.size()) * 8 + 64);
      if (ssw_one_cigar(query.c_str(), ref.c_str(), (int32_t)ref.size(), &sc, out, buf.data(), (int32_t)buf.size()))
        throw std::runtime_error("SSW refused a pair");
      cigar = buf.data();
      return;
    }
    int64_t aoff = 0, boff = 0;
    int32_t alen = (int32_t)ref.size(), blen = (int32_t)query.size();
    int32_t r[5], mism = 0, nops = 0;
    const int maxops = (int)(2 * (query.size() + ref.size()) + 8);
    std::vector<uint32_t> ops((size_t)maxops);
    oracle_cigar(1, ref.data(), &aoff, &alen, query.data(), &boff, &blen, &sc, r, &mism, &nops, ops.data(), maxops, 1);
    out[0] = r[0];
    out[1] = -1;
    out[2] = r[1];
    out[3] = r[2];
    out[4] = r[3];
    out[5] = r[4];
    out[6] = mism;
    static const char kOps[] = "MIDNSHP=X";
    cigar.clear();
    for (int k = 0; k < nops && k < maxops; k++) {
      cigar += std::to_string(ops[(size_t)k] >> 4);
      cigar += kOps[ops[(size_t)k] & 0xFu];
    }
  }

  // SSW's Alignment fields klign reads: sw_score, sw_score_next_best,
  // ref_begin, ref_end, query_begin, query_end, mismatches.
  void align(const std::string &query, const std::string &ref, int32_t out[7]) const {
    if (ssw_one) {
      if (ssw_one(query.c_str(), ref.c_str(), (int32_t)ref.size(), &sc, out) != 0)
        throw std::runtime_error("SSW refused a pair");
      return;
    }
    // scalar restatement: SSW-ref = a = ref, SSW-query = b = query
    int64_t aoff = 0, boff = 0;
    int32_t alen = (int32_t)ref.size(), blen = (int32_t)query.size();
    int32_t r[5];
    oracle(1, ref.data(), &aoff, &alen, query.data(), &boff, &blen, &sc, r, 1);
    out[0] = r[0];
    out[1] = -1;
    out[2] = r[1];
    out[3] = r[2];
    out[4] = r[3];
    out[5] = r[4];
    out[6] = 0;
  }
};

// ---- klign's batching state ----
struct KlignSim {
  AlnScoring aln_scoring;
  CpuAligner cpu;
  int gpu_devices = 0;
  size_t gpu_mem_avail = 0;
  int64_t block_size = KLIGN_GPU_BLOCK_SIZE;
  bool steal = true;
  int gpu_delay_ms = 0;
  bool gpu_score2 = false;
  bool report_cigar = false;  // klign's ssw_filter.report_cigar
  bool gpu_cigar = false;     // CIGAR blocks go to the GPU driver too
  adept_sw::GPUDriver gpu_driver;

  std::vector<Aln> kernel_alns;
  std::vector<std::string> ctg_seqs, read_seqs;
  int64_t max_clen = 0, max_rlen = 0;
  std::future<void> active_kernel_fut;

  std::mutex alns_mu;
  std::vector<Aln> alns;
  int64_t num_perfect = 0, num_stolen = 0, num_gpu = 0, num_blocks = 0;
  double gpu_secs = 0;

  bool kernel_ready() {
    return !active_kernel_fut.valid() ||
           active_kernel_fut.wait_for(std::chrono::seconds(0)) == std::future_status::ready;
  }
  void add_aln(Aln &&a) {
    std::lock_guard<std::mutex> g(alns_mu);
    alns.push_back(std::move(a));
  }

  // klign.cpp:178-209
  static void set_sam_string(Aln &aln, std::string read_seq, const std::string &cigar) {
    aln.sam_string = aln.read_id + "\t";
    if (aln.orient == '-') {
      aln.sam_string += "16\t";
      if (read_seq != "*") read_seq = revcomp(read_seq);
    } else {
      aln.sam_string += "0\t";
    }
    aln.sam_string += "Contig" + std::to_string(aln.cid) + "\t" + std::to_string(aln.cstart + 1) + "\t";
    uint32_t mapq;
    if (aln.score2 == 0) {
      mapq = 60;
    } else {
      mapq = -4.343 * std::log(1 - (double)std::abs(aln.score1 - aln.score2) / (double)aln.score1);
      mapq = (uint32_t)(mapq + 4.99);
      mapq = mapq < 254 ? mapq : 254;
    }
    aln.sam_string += std::to_string(mapq) + "\t";
    aln.sam_string += cigar + "\t*\t0\t0\t*\t*\t";
    aln.sam_string += "AS:i:" + std::to_string(aln.score1) + "\tNM:i:" + std::to_string(aln.mismatches) +
                      "\tRG:Z:" + std::to_string(aln.read_group_id);
  }

  // klign.cpp:218-249
  void align_read(const std::string &rname, int64_t read_idx, int64_t cid, const std::string &rseq,
                  const std::string &cseq, int rstart, int rlen, int cstart, int clen, char orient, int overlap_len,
                  int read_group_id) {
    if (cseq.compare(0, overlap_len, rseq, rstart, overlap_len) == 0) {
      num_perfect++;
      int rstop = rstart + overlap_len;
      int cstop = cstart + overlap_len;
      if (orient == '-') switch_orient(rstart, rstop, rlen);
      int score1 = overlap_len * aln_scoring.match;
      int identity = 100 * score1 / aln_scoring.match / rlen;
      Aln aln(rname, cid, rstart, rstop, rlen, cstart, cstop, clen, orient, score1, 0, identity, 0, read_group_id);
      if (report_cigar) set_sam_string(aln, rseq, std::to_string(overlap_len) + "=");
      aln.side = 'P';
      aln.read_idx = read_idx;
      add_aln(std::move(aln));
    } else {
      max_clen = std::max((int64_t)cseq.size(), max_clen);
      max_rlen = std::max((int64_t)rseq.size(), max_rlen);
      int64_t num_alns = (int64_t)kernel_alns.size() + 1;
      int64_t tot_mem_est = num_alns * (max_clen + max_rlen + 2 * sizeof(int) + 5 * sizeof(short));
      kernel_alns.emplace_back(rname, cid, 0, 0, rlen, cstart, 0, clen, orient);
      kernel_alns.back().read_idx = read_idx;
      ctg_seqs.emplace_back(cseq);
      read_seqs.emplace_back(rseq);
      bool will_run_kernel = (tot_mem_est >= (int64_t)gpu_mem_avail) | (num_alns >= block_size);
      if (will_run_kernel) kernel_align_block(read_group_id);
    }
  }

  // klign.cpp:251-273
  void ssw_align_read(Aln &aln, const std::string &cseq, const std::string &rseq, int read_group_id) {
    int32_t r[7];
    std::string cigar;
    if (report_cigar) {
      cpu.align_cigar(cseq, rseq, r, cigar);
    } else {
      cpu.align(cseq, rseq, r);
    }
    aln.rstop = aln.rstart + r[3] + 1;
    aln.rstart += r[2];
    aln.cstop = aln.cstart + r[5] + 1;
    aln.cstart += r[4];
    if (aln.orient == '-') switch_orient(aln.rstart, aln.rstop, aln.rlen);
    aln.score1 = r[0];
    aln.score2 = r[1];
    aln.mismatches = r[6];
    aln.identity = (int8_t)(100 * r[0] / aln_scoring.match / aln.rlen);
    aln.read_group_id = (int16_t)read_group_id;
    if (report_cigar) set_sam_string(aln, rseq, cigar);
    aln.side = 'C';
    add_aln(std::move(aln));
  }

  // klign.cpp:356-388 and 390-413 (the block runs on its own thread)
  struct Block {
    std::vector<Aln> kernel_alns;
    std::vector<std::string> ctg_seqs, read_seqs;
    int64_t max_clen, max_rlen;
    int read_group_id;
  };
  void gpu_align_block_kernel(Block &b) {
    auto t0 = std::chrono::steady_clock::now();
    gpu_driver.run_kernel_forwards(b.read_seqs, b.ctg_seqs, (unsigned)b.max_rlen, (unsigned)b.max_clen);
    gpu_driver.kernel_block();
    gpu_driver.run_kernel_backwards(b.read_seqs, b.ctg_seqs, (unsigned)b.max_rlen, (unsigned)b.max_clen);
    gpu_driver.kernel_block();
    gpu_secs += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    auto aln_results = gpu_driver.get_aln_results();
    const int* s2 = gpu_score2 ? gpu_driver.get_score2() : nullptr;
    if (gpu_score2 && !s2) throw std::runtime_error("the driver returned no score2");
    const std::string *cig = report_cigar ? gpu_driver.get_cigar_strings() : nullptr;
    const int *mism = report_cigar ? gpu_driver.get_mismatches() : nullptr;
    if (report_cigar && (!cig || !mism)) throw std::runtime_error("the driver returned no CIGARs");
    std::vector<Aln> out;
    out.reserve(b.kernel_alns.size());
    for (size_t i = 0; i < b.kernel_alns.size(); i++) {
      Aln &aln = b.kernel_alns[i];
      aln.rstop = aln.rstart + aln_results.query_end[i] + 1;
      aln.rstart += aln_results.query_begin[i];
      aln.cstop = aln.cstart + aln_results.ref_end[i] + 1;
      aln.cstart += aln_results.ref_begin[i];
      if (aln.orient == '-') switch_orient(aln.rstart, aln.rstop, aln.rlen);
      aln.score1 = aln_results.top_scores[i];
      aln.score2 = s2 ? s2[i] : 0;
      aln.mismatches = mism ? mism[i] : 0;
      aln.identity = (int8_t)(100 * aln.score1 / aln_scoring.match / aln.rlen);
      aln.read_group_id = (int16_t)b.read_group_id;
      if (report_cigar) set_sam_string(aln, b.read_seqs[i], cig[i]);
      aln.side = 'G';
      out.push_back(std::move(aln));
    }
    if (gpu_delay_ms > 0) std::this_thread::sleep_for(std::chrono::milliseconds(gpu_delay_ms));
    std::lock_guard<std::mutex> g(alns_mu);
    num_gpu += (int64_t)out.size();
    for (auto &a : out) alns.push_back(std::move(a));
  }

  void clear_aln_bufs() {
    kernel_alns.clear();
    ctg_seqs.clear();
    read_seqs.clear();
    max_clen = 0;
    max_rlen = 0;
  }

  // klign.cpp:552-606
  void kernel_align_block(int read_group_id) {
    while ((gpu_devices == 0 || !kernel_ready()) && !kernel_alns.empty()) {
      if (!steal) {
        std::this_thread::yield();
        continue;
      }
      ssw_align_read(kernel_alns.back(), ctg_seqs.back(), read_seqs.back(), read_group_id);
      kernel_alns.pop_back();
      ctg_seqs.pop_back();
      read_seqs.pop_back();
      num_stolen++;
    }
    if (!kernel_alns.empty()) {
      if (active_kernel_fut.valid()) active_kernel_fut.get();
      auto blk = std::make_shared<Block>();
      blk->kernel_alns.swap(kernel_alns);
      blk->ctg_seqs.swap(ctg_seqs);
      blk->read_seqs.swap(read_seqs);
      blk->max_clen = max_clen;
      blk->max_rlen = max_rlen;
      blk->read_group_id = read_group_id;
      num_blocks++;
      if (!report_cigar || gpu_cigar) {
        active_kernel_fut = std::async(std::launch::async, [this, blk]() { gpu_align_block_kernel(*blk); });
      } else {
        // ssw_align_block: the block on the CPU, on its own thread
        active_kernel_fut = std::async(std::launch::async, [this, blk]() {
          for (size_t i = 0; i < blk->kernel_alns.size(); i++)
            ssw_align_read(blk->kernel_alns[i], blk->ctg_seqs[i], blk->read_seqs[i], blk->read_group_id);
        });
      }
    }
    clear_aln_bufs();
  }

  // klign.cpp:608-621
  void flush_remaining(int read_group_id) {
    if (!kernel_alns.empty()) kernel_align_block(read_group_id);
    if (active_kernel_fut.valid()) active_kernel_fut.get();
  }
};

// ---- seeded workload: contigs, reads with a known source, one seed per hit ----
struct Rng {
  uint64_t s;
  explicit Rng(uint64_t seed) : s(seed * 0x9E3779B97F4A7C15ull + 0x632BE59BD9B4E019ull) {}
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  int64_t uniform(int64_t lo, int64_t hi) { return lo + (int64_t)(next() % (uint64_t)(hi - lo + 1)); }
  double unit() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};

const char kBases[4] = {'A', 'C', 'G', 'T'};

std::string revcomp(const std::string &s) {
  std::string r(s.rbegin(), s.rend());
  for (auto &c : r) {
    switch (c) {
      case 'A': c = 'T'; break;
      case 'C': c = 'G'; break;
      case 'G': c = 'C'; break;
      case 'T': c = 'A'; break;
      default: c = 'N';
    }
  }
  return r;
}

struct Hit {
  int64_t cid;
  int pos_in_ctg;
  int pos_in_read;  // position of the seed k-mer in the read as stored
  bool rc;          // the stored read is the reverse complement of the contig strand
};

struct Workload {
  std::vector<std::string> ctgs;
  std::vector<std::string> reads;
  std::vector<std::vector<Hit>> hits;
};

Workload make_workload(uint64_t seed, int64_t nreads, int kmer_len) {
  Workload w;
  Rng rng(seed);
  const int nctgs = 300;
  for (int c = 0; c < nctgs; c++) {
    int len = (int)rng.uniform(120, 6000);
    std::string s(len, 'A');
    for (auto &ch : s) ch = rng.unit() < 0.002 ? 'N' : kBases[rng.next() & 3];
    w.ctgs.push_back(std::move(s));
  }
  w.reads.reserve(nreads);
  w.hits.reserve(nreads);
  for (int64_t r = 0; r < nreads; r++) {
    const int64_t cid = (int64_t)(rng.next() % nctgs);
    const std::string &ctg = w.ctgs[cid];
    const int clen = (int)ctg.size();
    const double u = rng.unit();
    const int rlen = u < 0.2 ? 100 : (u < 0.8 ? 150 : 250);
    const bool exact = rng.unit() < 0.15;
    int start = exact ? (int)rng.uniform(0, std::max(0, clen - rlen)) : (int)rng.uniform(-rlen / 3, clen - 2 * rlen / 3);
    std::string f;
    std::vector<int> src;  // contig position each read base copies, -1 otherwise
    int cpos = start;
    while ((int)f.size() < rlen) {
      const bool inside = cpos >= 0 && cpos < clen;
      const double e = exact ? 1.0 : rng.unit();
      if (!inside) {
        f.push_back(kBases[rng.next() & 3]);
        src.push_back(-1);
        cpos++;
      } else if (e < 0.0005) {  // insertion
        f.push_back(kBases[rng.next() & 3]);
        src.push_back(-1);
      } else if (e < 0.001) {  // deletion
        cpos++;
      } else if (e < 0.011) {  // substitution
        char b = kBases[rng.next() & 3];
        while (b == ctg[cpos]) b = kBases[rng.next() & 3];
        f.push_back(b);
        src.push_back(-1);
        cpos++;
      } else {
        f.push_back(ctg[cpos]);
        src.push_back(cpos);
        cpos++;
      }
    }
    // seed: the first k-mer copied without error from consecutive contig bases
    int seed_pos = -1;
    for (int p = 0; p + kmer_len <= rlen && seed_pos < 0; p++) {
      bool ok = src[p] >= 0;
      for (int k = 1; ok && k < kmer_len; k++) ok = src[p + k] == src[p] + k;
      if (ok) seed_pos = p;
    }
    const bool rc = (rng.next() & 1) != 0;
    std::vector<Hit> hv;
    if (seed_pos >= 0) {
      Hit h{cid, src[seed_pos], rc ? rlen - kmer_len - seed_pos : seed_pos, rc};
      hv.push_back(h);
    }
    if (rng.unit() < 0.05) {  // a spurious seed into another contig
      int64_t c2 = (int64_t)(rng.next() % nctgs);
      if (c2 != cid && (int)w.ctgs[c2].size() >= kmer_len) {
        Hit h{c2, (int)rng.uniform(0, (int64_t)w.ctgs[c2].size() - kmer_len), (int)rng.uniform(0, rlen - kmer_len),
              (rng.next() & 1) != 0};
        hv.push_back(h);
      }
    }
    w.reads.push_back(rc ? revcomp(f) : f);
    w.hits.push_back(std::move(hv));
  }
  return w;
}

void usage() {
  std::fprintf(stderr,
               "usage: mini_klign --mode gpu|cpu --out FILE [--seed S] [--reads N] [--block B] [--mem BYTES]\n"
               "                  [--no-steal] [--gpu-delay-ms D] [--scoring m,x,o,e,a] [--ssw LIB] [--oracle LIB]\n"
               "                  [--gpu-score2] [--report-cigar] [--gpu-cigar]\n");
}

}  // namespace

int main(int argc, char **argv) {
  std::string mode = "gpu", out_path, ssw_lib, oracle_lib;
  uint64_t seed = 11;
  int64_t nreads = 40000, block = KLIGN_GPU_BLOCK_SIZE, mem = 0;
  bool steal = true, k_gpu_score2 = false, k_report_cigar = false, k_gpu_cigar = false;
  int delay = 0;
  AlnScoring sc{1, 3, 5, 2, 2};
  const int kmer_len = 21;
  for (int i = 1; i < argc; i++) {
    std::string a = argv[i];
    auto val = [&]() -> std::string {
      if (i + 1 >= argc) {
        usage();
        std::exit(2);
      }
      return argv[++i];
    };
    if (a == "--mode") mode = val();
    else if (a == "--out") out_path = val();
    else if (a == "--seed") seed = std::strtoull(val().c_str(), nullptr, 10);
    else if (a == "--reads") nreads = std::strtoll(val().c_str(), nullptr, 10);
    else if (a == "--block") block = std::strtoll(val().c_str(), nullptr, 10);
    else if (a == "--mem") mem = std::strtoll(val().c_str(), nullptr, 10);
    else if (a == "--no-steal") steal = false;
    else if (a == "--gpu-score2") k_gpu_score2 = true;
    else if (a == "--report-cigar") k_report_cigar = true;
    else if (a == "--gpu-cigar") k_gpu_cigar = true;
    else if (a == "--gpu-delay-ms") delay = std::atoi(val().c_str());
    else if (a == "--ssw") ssw_lib = val();
    else if (a == "--oracle") oracle_lib = val();
    else if (a == "--scoring") {
      if (std::sscanf(val().c_str(), "%d,%d,%d,%d,%d", &sc.match, &sc.mismatch, &sc.gap_opening, &sc.gap_extending,
                      &sc.ambiguity) != 5) {
        usage();
        return 2;
      }
    } else {
      usage();
      return 2;
    }
  }
  if (out_path.empty() || (mode != "gpu" && mode != "cpu") || block < 1 || (mode == "cpu" && !steal)) {
    usage();
    return 2;
  }

  KlignSim k;
  k.report_cigar = k_report_cigar;
  k.gpu_cigar = k_gpu_cigar;
  if (k_gpu_cigar && !k_report_cigar) {
    usage();
    return 2;
  }
  k.aln_scoring = sc;
  k.block_size = block;
  k.steal = steal;
  k.gpu_delay_ms = delay;
  k.cpu.sc = CScoring{sc.match, sc.mismatch, sc.gap_opening, sc.gap_extending, sc.ambiguity};
  if (!ssw_lib.empty()) {
    void *h = dlopen(ssw_lib.c_str(), RTLD_NOW | RTLD_LOCAL);
    if (h) k.cpu.ssw_one = (ssw_one_fn)dlsym(h, "ssw_ref_align_one");
    if (h) k.cpu.ssw_one_cigar = (ssw_one_cigar_fn)dlsym(h, "ssw_ref_align_one_cigar");
  }
  if (!k.cpu.ssw_one && !oracle_lib.empty()) {
    void *h = dlopen(oracle_lib.c_str(), RTLD_NOW | RTLD_LOCAL);
    if (h) k.cpu.oracle = (oracle_batch_fn)dlsym(h, "oracle_align_batch");
    if (h) k.cpu.oracle_cigar = (oracle_cigar_fn)dlsym(h, "oracle_cigar_batch");
  }
  if (!k.cpu.ssw_one && !k.cpu.oracle && (steal || mode == "cpu")) {
    std::fprintf(stderr, "no CPU aligner: give --ssw or --oracle\n");
    return 2;
  }

  // klign.cpp:448-477: GPU count, memory budget, driver init
  if (mode == "gpu") {
    k.gpu_devices = adept_sw::get_num_node_gpus();
    if (k.gpu_devices <= 0) {
      std::fprintf(stderr, "no GPU\n");
      return 3;
    }
    k.gpu_mem_avail = mem > 0 ? (size_t)mem : adept_sw::get_avail_gpu_mem_per_rank(1, k.gpu_devices);
    k.gpu_driver.init(0, 1, (short)sc.match, (short)-sc.mismatch, (short)-sc.gap_opening, (short)-sc.gap_extending,
                      300);
    // klign's build passes N's cost through ALN_AMBIGUITY_COST; this harness sets it explicitly
    k.gpu_driver.set_ambiguity((short)sc.ambiguity);
    // CIGAR blocks on the GPU need score2 too: set_sam_string's mapq reads it
    if (k_gpu_cigar) k_gpu_score2 = true;
    if (k_gpu_score2) k.gpu_driver.set_score2(true);
    if (k_gpu_cigar) k.gpu_driver.set_cigar(true);
    k.gpu_score2 = k_gpu_score2;
  } else {
    k.gpu_devices = 0;
    k.gpu_mem_avail = mem > 0 ? (size_t)mem : 32 * 1024 * 1024;
  }

  Workload w = make_workload(seed, nreads, kmer_len);

  // compute_alns_for_read (klign.cpp:672-739) over every read's hits
  auto t0 = std::chrono::steady_clock::now();
  int64_t num_alns = 0;
  for (int64_t r = 0; r < nreads; r++) {
    const std::string &rseq = w.reads[r];
    const int rlen = (int)rseq.size();
    const std::string rseq_rc = revcomp(rseq);
    const std::string rname = "r" + std::to_string(r);
    for (const Hit &h : w.hits[r]) {
      const std::string &ctg = w.ctgs[h.cid];
      const int clen = (int)ctg.size();
      int pos_in_read = h.pos_in_read;
      char orient = '+';
      const std::string *rseq_ptr = &rseq;
      if (h.rc) {
        orient = '-';
        pos_in_read = rlen - (kmer_len + pos_in_read);
        rseq_ptr = &rseq_rc;
      }
      const int ctg_bases_left_of_kmer = h.pos_in_ctg;
      const int ctg_bases_right_of_kmer = clen - ctg_bases_left_of_kmer - kmer_len;
      const int read_bases_left_of_kmer = pos_in_read;
      const int read_bases_right_of_kmer = rlen - kmer_len - pos_in_read;
      const int left_of_kmer = std::min(read_bases_left_of_kmer, ctg_bases_left_of_kmer);
      const int right_of_kmer = std::min(read_bases_right_of_kmer, ctg_bases_right_of_kmer);
      const int cstart = h.pos_in_ctg - left_of_kmer;
      const int rstart = pos_in_read - left_of_kmer;
      const int overlap_len = left_of_kmer + kmer_len + right_of_kmer;
      const std::string read_subseq = rseq_ptr->substr(0, rlen);
      const std::string ctg_subseq = ctg.substr(cstart, overlap_len);
      k.align_read(rname, r, h.cid, read_subseq, ctg_subseq, rstart, rlen, cstart, clen, orient, overlap_len, 0);
      num_alns++;
    }
  }
  k.flush_remaining(0);
  const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();

  std::sort(k.alns.begin(), k.alns.end(), [](const Aln &a, const Aln &b) {
    return a.read_idx != b.read_idx ? a.read_idx < b.read_idx : a.cid < b.cid;
  });
  FILE *f = std::fopen(out_path.c_str(), "w");
  if (!f) {
    std::fprintf(stderr, "cannot write %s\n", out_path.c_str());
    return 2;
  }
  for (const Aln &a : k.alns)
    std::fprintf(f, "%s\t%d\t%d\t%c\t%s\n", a.to_string().c_str(), (int)a.identity, a.mismatches, a.side,
                 a.sam_string.c_str());
  std::fclose(f);
  std::printf(
      "{\"mode\": \"%s\", \"alns\": %lld, \"records\": %lld, \"perfect\": %lld, \"gpu\": %lld, \"stolen\": %lld, "
      "\"blocks\": %lld, \"secs\": %.4f, \"gpu_secs\": %.4f, \"cpu_aligner\": \"%s\"}\n",
      mode.c_str(), (long long)num_alns, (long long)k.alns.size(), (long long)k.num_perfect, (long long)k.num_gpu,
      (long long)k.num_stolen, (long long)k.num_blocks, secs, k.gpu_secs,
      k.cpu.ssw_one ? "ssw" : (k.cpu.oracle ? "oracle" : "none"));
  return 0;
}
