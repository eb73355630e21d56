// CPU check of the per-pair CIGAR code the HIP kernels run
// (mhm2_staging_amd/csrc/cigar_core.h), compiled for the host: the same
// narrow band classes, band doubling, count / recount / write walks (CIGARs
// kept in registers up to 16 ops between the clips) and trivial cases as
// cigar_kernels.hip, with
// the direction bits in the blocked layout of a wavefront's scratch slot.  Bands the
// kernels give to their wavefront-per-pair kernel run here through the
// one-cell-at-a-time band DP of the same header (same cells, same bits).  Every pair is compared with
// the oracle (oracle/sw_oracle.c), which is what this test program links it
// against.
//
// stdin: "npairs match mismatch gap_open gap_ext ambiguity", then one line
// per pair: "<SSW-ref> <SSW-query>".  stdout: "ok <pairs> <class counts>" or
// the first differing pairs; exit status 1 on any difference.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <string>
#include <vector>

#include "../../mhm2_staging_amd/csrc/cigar_core.h"
#include "../../oracle/sw_oracle.h"

using namespace mhmsw;

static void pack(const std::string& s, std::vector<uint32_t>& codes, int64_t base) {
  for (size_t k = 0; k < s.size(); k++) {
    const uint32_t c = (uint32_t)oracle_translate_base((unsigned char)s[k]);
    const int64_t idx = base + (int64_t)k;
    codes[idx >> 3] |= c << ((idx & 7) * 4);
  }
}

// words per row of the slot of the kernel that runs half-width BW: 1 for
// half-width 1, 2 for the merged half-widths 2..7
constexpr int slot_words(int bw) { return bw == 1 ? 1 : 2; }

template <int BW>
static int32_t run_dp(const uint32_t* codes, const CigarGeo& g, const oracle_scoring& sc,
                      std::vector<uint32_t>& slot) {
  constexpr int DS = slot_words(BW);
  slot.assign((size_t)((g.ql() + 7) & ~7) * DS * 64, 0xDEADBEEFu);
  // lane 17 and lane 49 of a wavefront hold the same pair: both must give it intact
  cigar_band_dp<BW, DS>(codes, g, sc.match, sc.mismatch, sc.ambiguity, sc.gap_open, sc.gap_ext,
                        slot.data() + 49 * 8 * DS);
  return cigar_band_dp<BW, DS>(codes, g, sc.match, sc.mismatch, sc.ambiguity, sc.gap_open, sc.gap_ext,
                               slot.data() + 17 * 8 * DS);
}

int main() {
  int64_t np;
  oracle_scoring sc;
  if (!(std::cin >> np >> sc.match >> sc.mismatch >> sc.gap_open >> sc.gap_ext >> sc.ambiguity)) return 2;
  int64_t bad = 0, cls_count[kCigClasses + 1] = {0}, retries[kCigClasses] = {0}, final_wide = 0;
  std::vector<uint32_t> slot, gen;
  for (int64_t p = 0; p < np; p++) {
    std::string R, Q;
    std::cin >> R >> Q;
    if (R == "-") R.clear();
    // layout as in the batch buffers: 32-aligned bases after a front pad
    const int64_t rbase = 1024, qbase = 1024 + ((int64_t)R.size() + 31) / 32 * 32 + 32;
    std::vector<uint32_t> codes((size_t)(qbase + (int64_t)Q.size() + 2048) / 8, 0u);
    pack(R, codes, rbase);
    pack(Q, codes, qbase);
    std::vector<int8_t> rc(R.size() + 1), qc(Q.size() + 1);
    oracle_translate(R.data(), (int)R.size(), rc.data());
    oracle_translate(Q.data(), (int)Q.size(), qc.data());
    oracle_result res;
    oracle_align_codes(rc.data(), (int)R.size(), qc.data(), (int)Q.size(), &sc, &res);
    std::vector<uint32_t> want((size_t)(2 * (R.size() + Q.size()) + 8));
    int32_t want_mism = 0;
    const int want_n = oracle_cigar_codes(rc.data(), (int)R.size(), qc.data(), (int)Q.size(), &sc, &res, want.data(),
                                          (int)want.size(), &want_mism);

    CigarGeo g;
    g.rbase = rbase;
    g.qbase = qbase;
    g.rb = res.ref_begin;
    g.re = res.ref_end;
    g.qb = res.query_begin;
    g.qe = res.query_end;
    g.n = (int32_t)Q.size();
    g.score = res.score;
    std::vector<uint32_t> got;
    int32_t got_mism = 0, got_n = 0;
    CigarEmit em0;
    em0.out = nullptr;
    em0.nops = 0;
    if (cigar_trivial(codes.data(), g, em0)) {
      got.resize((size_t)em0.count);
      CigarEmit em;
      em.out = got.data();
      em.nops = em0.count;
      cigar_trivial(codes.data(), g, em);
      got_n = em0.count;
      got_mism = em0.mism;
      cls_count[kCigClasses]++;
    } else {
      int32_t bw = std::abs(g.rl() - g.ql()) + 1, layout = 0, dwn = 0;
      const uint32_t* dirs = nullptr;
      bool blocked = true;
      int first_cls = -1;
      for (;;) {
        const int cls = cigar_class(bw, g.ql());
        if (first_cls < 0) first_cls = cls;
        int32_t corner;
        switch (cls) {
          case 0: corner = run_dp<1>(codes.data(), g, sc, slot); break;
          case 1: corner = run_dp<2>(codes.data(), g, sc, slot); break;
          case 2: corner = run_dp<3>(codes.data(), g, sc, slot); break;
          case 3: corner = run_dp<4>(codes.data(), g, sc, slot); break;
          case 4: corner = run_dp<5>(codes.data(), g, sc, slot); break;
          case 5: corner = run_dp<6>(codes.data(), g, sc, slot); break;
          case 6: corner = run_dp<7>(codes.data(), g, sc, slot); break;
          default: {
            const int32_t W = 2 * bw + 1, DW = (4 * W + 31) / 32;
            gen.assign((size_t)2 * W + (size_t)g.ql() * DW, 0xDEADBEEFu);
            corner = cigar_band_dp_any(codes.data(), g, bw, sc.match, sc.mismatch, sc.ambiguity, sc.gap_open,
                                       sc.gap_ext, (int32_t*)gen.data(), (int32_t*)gen.data() + W, gen.data() + 2 * W);
          }
        }
        if (corner == g.score) {
          if (cls == kCigClassWide) {
            layout = bw;
            dwn = (4 * (2 * bw + 1) + 31) / 32;
            dirs = gen.data() + 2 * (2 * bw + 1);
            blocked = false;
          } else {
            layout = bw;
            dwn = slot_words(bw);
            dirs = slot.data() + 17 * 8 * dwn;
            blocked = true;
          }
          break;
        }
        retries[cls]++;
        if (bw > 4 * (g.rl() + g.ql()) + 4) {
          std::printf("pair %lld: band never reaches the score\n", (long long)p);
          return 1;
        }
        bw *= 2;
      }
      cls_count[first_cls]++;
      if (cigar_class(bw, g.ql()) == kCigClassWide) final_wide++;
      // as the kernels: band classes walk rows in groups first, the generic
      // class walks row by row; a traceback that stops off column 0 is
      // walked again comparing shifted codes
      const bool eq00 = nib_at(codes.data(), g.qbase + g.qb) == nib_at(codes.data(), g.rbase + g.rb);
      const bool band = blocked;
      auto fast_walk = [&](auto &e) -> int32_t {
        switch (dwn) {
          case 1: return cigar_walk_rows<1>(g, layout, dirs, eq00, e);
          case 2: return cigar_walk_rows<2>(g, layout, dirs, eq00, e);
          default: return -9;
        }
      };
      CigarEmitQ q1;
      CigarEmit c1;
      c1.out = nullptr;
      c1.nops = 0;
      int32_t jf;
      bool queued = false;
      if (band) {
        jf = fast_walk(q1);
        c1.count = q1.count;
        c1.mism = q1.mism;
        queued = jf == 0 && q1.fits();
      } else {
        jf = cigar_walk(codes.data(), g, layout, dirs, dwn, blocked, 0, eq00, c1);
      }
      CigarEmit c2 = c1;
      if (jf > 0) {
        c2 = CigarEmit();
        c2.out = nullptr;
        c2.nops = 0;
        jf = cigar_walk(codes.data(), g, layout, dirs, dwn, blocked, jf, eq00, c2);
      }
      if (jf < 0) {
        got_n = -1;
      } else {
        got.assign((size_t)c2.count, 0u);
        CigarEmit w;
        w.out = got.data();
        w.nops = c2.count;
        if (queued) {
          q1.write(got.data());
          w.count = c2.count;
          w.mism = c2.mism;
        } else if (band && jf == 0) {
          fast_walk(w);
        } else {
          cigar_walk(codes.data(), g, layout, dirs, dwn, blocked, jf, eq00, w);
        }
        got_n = c2.count;
        got_mism = c2.mism;
        if (w.count != c2.count || w.mism != c2.mism) got_n = -2;
      }
    }
    bool same = got_n == want_n && got_mism == want_mism;
    for (int k = 0; same && k < got_n; k++) same = got[(size_t)k] == want[(size_t)k];
    if (!same) {
      if (bad < 5) {
        std::printf("pair %lld (%s / %s): score %d rb %d re %d qb %d qe %d: got %d ops mism %d, want %d ops mism %d\n",
                    (long long)p, R.c_str(), Q.c_str(), res.score, res.ref_begin, res.ref_end, res.query_begin,
                    res.query_end, got_n, got_mism, want_n, want_mism);
      }
      bad++;
    }
  }
  if (bad) {
    std::printf("%lld of %lld pairs differ\n", (long long)bad, (long long)np);
    return 1;
  }
  std::printf("ok %lld", (long long)np);
  for (int c = 0; c <= kCigClasses; c++) std::printf(" %lld", (long long)cls_count[c]);
  std::fprintf(stderr, "retries by class:");
  for (int c = 0; c < kCigClasses; c++) std::fprintf(stderr, " %lld", (long long)retries[c]);
  std::fprintf(stderr, "; pairs ending in the wide class: %lld\n", (long long)final_wide);
  std::printf("\n");
  return 0;
}
