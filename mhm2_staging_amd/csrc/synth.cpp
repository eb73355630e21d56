// Seeded synthetic read/contig-window workloads for the klign alignment path.
//
// Each generator is deterministic in (config, seed, pair index) whatever the
// thread count: every pair draws from its own splitmix64 stream.  Sequences
// are ASCII (A, C, G, T, N) so they can go through the same entry points as
// klign's std::string batches (src/klign.cpp:239-241).
//
//   1  plain:       150 bp reads, 1% substitutions, 0.1% indels of 1-3 bp,
//                   window = source region +-100 bp (350 bp) of one 1 Mbp contig
//   2  uneven:      reads U{100..250}, windows U{150..1024}, contig pool with
//                   0.5% N, 2% substitutions, 0.2% indels, 10% unrelated pairs
//   3  skewed:      windows log-normal(ln 400, 0.5) clipped to [150, 1024],
//                   reads 150 (70%) / 250 (30%), 1% subs, 0.1% indels, 25% of
//                   reads with a random 20-60 bp tail
//   4  repeats:     1024 bp windows of tandem copies of a 16-64 bp unit with
//                   0.1% mutation, 250 bp reads at 0.1% error, 5% of reads
//                   4000 bp
//   5  serving:     64k contigs (log-uniform 500..20000 bp); pairs are 150 bp
//                   reads at 1% error against U[300,1024] windows of them
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "../../include/mhmsw_synth.h"

namespace {

struct Rng {
  uint64_t s;
  explicit Rng(uint64_t seed) : s(seed) {}
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double uniform() { return (next() >> 11) * (1.0 / 9007199254740992.0); }
  // uniform integer in [lo, hi]
  int64_t range(int64_t lo, int64_t hi) { return lo + (int64_t)(next() % (uint64_t)(hi - lo + 1)); }
  double normal() {
    double u1 = uniform(), u2 = uniform();
    if (u1 < 1e-300) u1 = 1e-300;
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
  }
};

uint64_t stream_seed(uint64_t seed, uint64_t stream, uint64_t idx) {
  Rng r(seed * 0x100000001B3ull ^ (stream << 56) ^ idx);
  r.next();
  return r.next();
}

const char kBases[4] = {'A', 'C', 'G', 'T'};

char random_base(Rng &r) { return kBases[r.next() & 3]; }

char other_base(Rng &r, char c) {
  for (;;) {
    char b = random_base(r);
    if (b != c) return b;
  }
}

void random_seq(Rng &r, std::string &out, int64_t n, double n_rate) {
  out.resize((size_t)n);
  for (int64_t i = 0; i < n; i++) out[(size_t)i] = (n_rate > 0 && r.uniform() < n_rate) ? 'N' : random_base(r);
}

// Copy src[pos...] with substitutions and indels until exactly `len` bases are
// emitted.  Reads past the end of src continue with random bases.
void mutate_copy(Rng &r, const char *src, int64_t src_len, int64_t pos, int64_t len, double sub, double indel,
                 std::string &out) {
  out.clear();
  out.reserve((size_t)len);
  int64_t i = pos;
  while ((int64_t)out.size() < len) {
    double x = r.uniform();
    if (x < indel * 0.5) {  // insertion of 1-3 random bases
      int k = (int)r.range(1, 3);
      for (int t = 0; t < k && (int64_t)out.size() < len; t++) out.push_back(random_base(r));
      continue;
    }
    if (x < indel) {  // deletion of 1-3 source bases
      i += r.range(1, 3);
      continue;
    }
    char c = (i >= 0 && i < src_len) ? src[i] : random_base(r);
    if (c != 'N' && r.uniform() < sub) c = other_base(r, c);
    out.push_back(c);
    i++;
  }
}

struct Pair {
  std::string read, ctg;
};

template <class F>
void parallel_pairs(int64_t n, int threads, F f) {
  if (threads <= 0) threads = (int)std::max(1u, std::thread::hardware_concurrency());
  threads = (int)std::min<int64_t>(threads, std::max<int64_t>(1, n / 256 + 1));
  std::vector<std::thread> th;
  for (int t = 0; t < threads; t++)
    th.emplace_back([=]() {
      for (int64_t p = t; p < n; p += threads) f(p);
    });
  for (auto &x : th) x.join();
}

void build_pool(uint64_t seed, int64_t len, double n_rate, std::string &pool, int threads) {
  pool.resize((size_t)len);
  const int64_t chunk = 1 << 20;
  int64_t nchunks = (len + chunk - 1) / chunk;
  parallel_pairs(nchunks, threads, [&](int64_t c) {
    Rng r(stream_seed(seed, 1, (uint64_t)c));
    int64_t b = c * chunk, e = std::min(len, b + chunk);
    for (int64_t i = b; i < e; i++) pool[(size_t)i] = (n_rate > 0 && r.uniform() < n_rate) ? 'N' : random_base(r);
  });
}

}  // namespace

extern "C" int mhmsw_synth_generate(int config, int64_t npairs, uint64_t seed, int threads, mhmsw_synth_batch *out) {
  if (!out) return -1;
  std::memset(out, 0, sizeof(*out));
  static const int64_t default_pairs[6] = {0, 10000, 1000000, 4000000, 500000, 100000};
  static const uint64_t default_seed[6] = {0, 1, 2, 3, 4, 5};
  if (config < 1 || config > 5) return -2;
  if (npairs <= 0) npairs = default_pairs[config];
  if (seed == 0) seed = default_seed[config];

  std::string pool;
  int64_t pool_len = 0;
  std::vector<int64_t> ctg_start, ctg_len5;  // config 5 contig table inside the pool
  if (config == 1) {
    pool_len = 1000000;
    build_pool(seed, pool_len, 0.0, pool, threads);
  } else if (config == 2) {
    pool_len = 64ll << 20;
    build_pool(seed, pool_len, 0.005, pool, threads);
  } else if (config == 3) {
    pool_len = 64ll << 20;
    build_pool(seed, pool_len, 0.0, pool, threads);
  } else if (config == 5) {
    const int64_t nctg = 64000;
    Rng r(stream_seed(seed, 2, 0));
    ctg_start.resize(nctg);
    ctg_len5.resize(nctg);
    int64_t acc = 0;
    for (int64_t c = 0; c < nctg; c++) {
      double l = std::exp(std::log(500.0) + r.uniform() * (std::log(20000.0) - std::log(500.0)));
      ctg_len5[c] = (int64_t)l;
      ctg_start[c] = acc;
      acc += ctg_len5[c];
    }
    pool_len = acc;
    build_pool(seed, pool_len, 0.0, pool, threads);
  }

  std::vector<Pair> pairs((size_t)npairs);
  const char *P = pool.data();
  parallel_pairs(npairs, threads, [&](int64_t p) {
    Rng r(stream_seed(seed, 3, (uint64_t)p));
    Pair &pr = pairs[(size_t)p];
    switch (config) {
      case 1: {
        int64_t s = r.range(100, pool_len - 250 - 1);
        pr.ctg.assign(P + s - 100, 350);
        mutate_copy(r, P, pool_len, s, 150, 0.01, 0.001, pr.read);
        break;
      }
      case 2: {
        int64_t rl = r.range(100, 250), wl = r.range(150, 1024);
        if (r.uniform() < 0.10) {
          random_seq(r, pr.read, rl, 0.0);
          random_seq(r, pr.ctg, wl, 0.005);
          break;
        }
        int64_t x = r.range(4096, pool_len - 4096 - wl);
        pr.ctg.assign(P + x, (size_t)wl);
        int64_t src = rl <= wl ? r.range(x, x + wl - rl) : x - r.range(0, rl - wl);
        mutate_copy(r, P, pool_len, src, rl, 0.02, 0.002, pr.read);
        break;
      }
      case 3: {
        double w = std::exp(std::log(400.0) + 0.5 * r.normal());
        int64_t wl = std::min<int64_t>(1024, std::max<int64_t>(150, (int64_t)std::llround(w)));
        int64_t rl = r.uniform() < 0.7 ? 150 : 250;
        int64_t x = r.range(4096, pool_len - 4096 - wl);
        pr.ctg.assign(P + x, (size_t)wl);
        int64_t src = rl <= wl ? r.range(x, x + wl - rl) : x - r.range(0, rl - wl);
        mutate_copy(r, P, pool_len, src, rl, 0.01, 0.001, pr.read);
        if (r.uniform() < 0.25) {
          int64_t tail = r.range(20, 60);
          for (int64_t i = rl - tail; i < rl; i++) pr.read[(size_t)i] = random_base(r);
        }
        break;
      }
      case 4: {
        int64_t ul = r.range(16, 64);
        std::string unit;
        random_seq(r, unit, ul, 0.0);
        const int64_t span = 8192;
        std::string tandem((size_t)span, 'A');
        for (int64_t i = 0; i < span; i++) tandem[(size_t)i] = unit[(size_t)(i % ul)];
        for (int64_t i = 0; i < span; i++)
          if (r.uniform() < 0.001) tandem[(size_t)i] = other_base(r, tandem[(size_t)i]);
        int64_t wx = r.range(0, span - 1024);
        pr.ctg.assign(tandem.data() + wx, 1024);
        int64_t rl = r.uniform() < 0.05 ? 4000 : 250;
        int64_t src = rl <= 1024 ? r.range(wx, wx + 1024 - rl) : r.range(0, span - rl);
        mutate_copy(r, tandem.data(), span, src, rl, 0.001, 0.0005, pr.read);
        break;
      }
      case 5: {
        int64_t c = r.range(0, (int64_t)ctg_start.size() - 1);
        int64_t cl = ctg_len5[(size_t)c];
        int64_t wl = std::min<int64_t>(cl, r.range(300, 1024));
        int64_t wx = ctg_start[(size_t)c] + r.range(0, cl - wl);
        pr.ctg.assign(P + wx, (size_t)wl);
        int64_t src = r.range(wx, wx + std::max<int64_t>(0, wl - 150));
        mutate_copy(r, P, pool_len, src, 150, 0.01, 0.001, pr.read);
        break;
      }
    }
  });

  int64_t rb = 0, cb = 0;
  for (auto &pr : pairs) {
    rb += (int64_t)pr.read.size();
    cb += (int64_t)pr.ctg.size();
  }
  out->npairs = npairs;
  out->reads = (char *)std::malloc((size_t)rb + 1);
  out->ctgs = (char *)std::malloc((size_t)cb + 1);
  out->read_off = (int64_t *)std::malloc(sizeof(int64_t) * (size_t)npairs);
  out->ctg_off = (int64_t *)std::malloc(sizeof(int64_t) * (size_t)npairs);
  out->read_len = (int32_t *)std::malloc(sizeof(int32_t) * (size_t)npairs);
  out->ctg_len = (int32_t *)std::malloc(sizeof(int32_t) * (size_t)npairs);
  if (!out->reads || !out->ctgs || !out->read_off || !out->ctg_off || !out->read_len || !out->ctg_len) {
    mhmsw_synth_free(out);
    return -3;
  }
  int64_t ro = 0, co = 0;
  for (int64_t p = 0; p < npairs; p++) {
    Pair &pr = pairs[(size_t)p];
    out->read_off[p] = ro;
    out->read_len[p] = (int32_t)pr.read.size();
    std::memcpy(out->reads + ro, pr.read.data(), pr.read.size());
    ro += (int64_t)pr.read.size();
    out->ctg_off[p] = co;
    out->ctg_len[p] = (int32_t)pr.ctg.size();
    std::memcpy(out->ctgs + co, pr.ctg.data(), pr.ctg.size());
    co += (int64_t)pr.ctg.size();
  }
  out->reads[ro] = 0;
  out->ctgs[co] = 0;
  out->reads_bytes = ro;
  out->ctgs_bytes = co;
  return 0;
}

extern "C" void mhmsw_synth_free(mhmsw_synth_batch *b) {
  if (!b) return;
  std::free(b->reads);
  std::free(b->ctgs);
  std::free(b->read_off);
  std::free(b->ctg_off);
  std::free(b->read_len);
  std::free(b->ctg_len);
  std::memset(b, 0, sizeof(*b));
}

extern "C" int mhmsw_synth_contigs(uint64_t seed, int64_t nctgs, int32_t min_len, int32_t max_len, int threads,
                                   mhmsw_synth_contig_set *out) {
  if (!out || nctgs <= 0 || min_len < 1 || max_len < min_len) return -1;
  std::memset(out, 0, sizeof(*out));
  std::vector<int32_t> len((size_t)nctgs);
  std::vector<int64_t> off((size_t)nctgs);
  Rng r(stream_seed(seed, 4, 0));
  int64_t acc = 0;
  for (int64_t c = 0; c < nctgs; c++) {
    double l = std::exp(std::log((double)min_len) + r.uniform() * (std::log((double)max_len) - std::log((double)min_len)));
    len[(size_t)c] = std::max<int32_t>(min_len, std::min<int32_t>(max_len, (int32_t)l));
    off[(size_t)c] = acc;
    acc += len[(size_t)c];
  }
  out->seqs = (char *)std::malloc((size_t)acc + 1);
  out->off = (int64_t *)std::malloc(sizeof(int64_t) * (size_t)nctgs);
  out->len = (int32_t *)std::malloc(sizeof(int32_t) * (size_t)nctgs);
  if (!out->seqs || !out->off || !out->len) {
    mhmsw_synth_contigs_free(out);
    return -3;
  }
  std::memcpy(out->off, off.data(), sizeof(int64_t) * (size_t)nctgs);
  std::memcpy(out->len, len.data(), sizeof(int32_t) * (size_t)nctgs);
  char *S = out->seqs;
  parallel_pairs(nctgs, threads, [&](int64_t c) {
    Rng rc(stream_seed(seed, 5, (uint64_t)c));
    for (int32_t i = 0; i < len[(size_t)c]; i++) S[off[(size_t)c] + i] = random_base(rc);
  });
  S[acc] = 0;
  out->nctgs = nctgs;
  out->bytes = acc;
  return 0;
}

extern "C" void mhmsw_synth_contigs_free(mhmsw_synth_contig_set *set) {
  if (!set) return;
  std::free(set->seqs);
  std::free(set->off);
  std::free(set->len);
  std::memset(set, 0, sizeof(*set));
}

extern "C" int mhmsw_synth_window_pairs(const mhmsw_synth_contig_set *ctgs, uint64_t seed, int64_t npairs, int threads,
                                        mhmsw_synth_window_batch *out) {
  if (!ctgs || !out || npairs < 0 || ctgs->nctgs <= 0) return -1;
  std::memset(out, 0, sizeof(*out));
  const int32_t rl = 150;
  out->reads = (char *)std::malloc((size_t)npairs * rl + 1);
  out->read_off = (int64_t *)std::malloc(sizeof(int64_t) * (size_t)std::max<int64_t>(npairs, 1));
  out->read_len = (int32_t *)std::malloc(sizeof(int32_t) * (size_t)std::max<int64_t>(npairs, 1));
  out->ctg_id = (int32_t *)std::malloc(sizeof(int32_t) * (size_t)std::max<int64_t>(npairs, 1));
  out->win_off = (int32_t *)std::malloc(sizeof(int32_t) * (size_t)std::max<int64_t>(npairs, 1));
  out->win_len = (int32_t *)std::malloc(sizeof(int32_t) * (size_t)std::max<int64_t>(npairs, 1));
  if (!out->reads || !out->read_off || !out->read_len || !out->ctg_id || !out->win_off || !out->win_len) {
    mhmsw_synth_window_pairs_free(out);
    return -3;
  }
  parallel_pairs(npairs, threads, [&](int64_t p) {
    Rng r(stream_seed(seed, 6, (uint64_t)p));
    int64_t c = r.range(0, ctgs->nctgs - 1);
    const int32_t cl = ctgs->len[c];
    const char *C = ctgs->seqs + ctgs->off[c];
    int32_t wl = (int32_t)std::min<int64_t>(cl, r.range(300, 1024));
    int32_t wo = (int32_t)r.range(0, cl - wl);
    int64_t src = wo + r.range(0, std::max<int32_t>(0, wl - rl));
    std::string read;
    mutate_copy(r, C, cl, src, rl, 0.01, 0.001, read);
    std::memcpy(out->reads + p * rl, read.data(), (size_t)rl);
    out->read_off[p] = p * rl;
    out->read_len[p] = rl;
    out->ctg_id[p] = (int32_t)c;
    out->win_off[p] = wo;
    out->win_len[p] = wl;
  });
  out->reads[npairs * rl] = 0;
  out->npairs = npairs;
  out->reads_bytes = npairs * rl;
  return 0;
}

extern "C" void mhmsw_synth_window_pairs_free(mhmsw_synth_window_batch *b) {
  if (!b) return;
  std::free(b->reads);
  std::free(b->read_off);
  std::free(b->read_len);
  std::free(b->ctg_id);
  std::free(b->win_off);
  std::free(b->win_len);
  std::memset(b, 0, sizeof(*b));
}
