// ORACLE -- TEST INFRASTRUCTURE ONLY.
//
// A C entry point around the reference's own StripedSmithWaterman::Aligner,
// compiled together with the reference sources where they lie
// (src/ssw/ssw.cpp, src/ssw/ssw_core.cpp) by oracle/Makefile into
// oracle/_ref/libssw_ref.so.  It makes exactly klign's call
// (src/klign.cpp:251-273):
//   Align(query = SSW-query, ref = SSW-ref, |ref|,
//         Filter(report_begin_position=true, report_cigar=false, 0, 32767),
//         &aln, maskLen = max(|ref| / 2, 15))
// and returns score, ref_begin/end and query_begin/end per pair (and, from
// the *2 / _one entries, sw_score_next_best); ssw_ref_cigar_batch makes the
// same call with report_cigar = true, as klign does when it reports CIGARs.  Used to
// record golden vectors and, on the host cores of the GPU machine, as the
// "reference" CPU baseline.
#include <pthread.h>

#include <algorithm>
#include <atomic>
#include <string>
#include <thread>
#include <vector>

#include "ssw.hpp"

extern "C" {
typedef struct {
  int32_t match, mismatch, gap_open, gap_ext, ambiguity;
} ssw_ref_scoring;

typedef struct {
  int32_t score, ref_begin, ref_end, query_begin, query_end;
} ssw_ref_result;

int ssw_ref_align_batch(int64_t npairs, const char *a, const int64_t *a_off, const int32_t *a_len, const char *b,
                        const int64_t *b_off, const int32_t *b_len, const ssw_ref_scoring *sc, ssw_ref_result *out,
                        int nthreads);

/* As ssw_ref_align_batch, plus sw_score_next_best of each pair in score2[p]. */
int ssw_ref_align_batch2(int64_t npairs, const char *a, const int64_t *a_off, const int32_t *a_len, const char *b,
                         const int64_t *b_off, const int32_t *b_len, const ssw_ref_scoring *sc, ssw_ref_result *out,
                         int32_t *score2, int nthreads);

/* One call exactly as klign's ssw_align_read makes it (src/klign.cpp:251-273):
 * query = NUL-terminated contig window, ref = read of ref_len bases.
 * out[7] = sw_score, sw_score_next_best, ref_begin, ref_end, query_begin,
 * query_end, mismatches.  Returns 0, or -1 when Align refuses the input. */
int ssw_ref_align_one(const char *query, const char *ref, int32_t ref_len, const ssw_ref_scoring *sc, int32_t *out);

/* As ssw_ref_align_one with report_cigar = true: out[6] = mismatches and the
 * NUL-terminated cigar_string in cigar (truncated to cigar_cap bytes). */
int ssw_ref_align_one_cigar(const char *query, const char *ref, int32_t ref_len, const ssw_ref_scoring *sc,
                            int32_t *out, char *cigar, int32_t cigar_cap);

/* klign's call when it reports CIGARs (report_cigar = true, klign.cpp:435):
 * results, Alignment::mismatches, and Alignment::cigar (BAM-encoded ops) of
 * pair p at ops[p * maxops ..] with their count in nops[p] (only the first
 * maxops are written when there are more).  A pair that Align refuses gets
 * score -1 and nops -1. */
int ssw_ref_cigar_batch(int64_t npairs, const char *a, const int64_t *a_off, const int32_t *a_len, const char *b,
                        const int64_t *b_off, const int32_t *b_len, const ssw_ref_scoring *sc, ssw_ref_result *out,
                        int32_t *mismatches, int32_t *nops, uint32_t *ops, int maxops, int nthreads);
}

int ssw_ref_align_one_cigar(const char *query, const char *ref, int32_t ref_len, const ssw_ref_scoring *sc,
                            int32_t *out, char *cigar, int32_t cigar_cap) {
  StripedSmithWaterman::Aligner aligner((uint8_t)sc->match, (uint8_t)sc->mismatch, (uint8_t)sc->gap_open,
                                        (uint8_t)sc->gap_ext, (uint8_t)sc->ambiguity);
  StripedSmithWaterman::Filter filter(true, true, 0, 32767);
  StripedSmithWaterman::Alignment aln;
  aln.Clear();
  if (!aligner.Align(query, ref, ref_len, filter, &aln, std::max((int)(ref_len / 2), 15))) return -1;
  out[0] = aln.sw_score;
  out[1] = aln.sw_score_next_best;
  out[2] = aln.ref_begin;
  out[3] = aln.ref_end;
  out[4] = aln.query_begin;
  out[5] = aln.query_end;
  out[6] = aln.mismatches;
  if (cigar_cap > 0) {
    const size_t k = std::min(aln.cigar_string.size(), (size_t)cigar_cap - 1);
    aln.cigar_string.copy(cigar, k);
    cigar[k] = 0;
  }
  return 0;
}

int ssw_ref_cigar_batch(int64_t npairs, const char *a, const int64_t *a_off, const int32_t *a_len, const char *b,
                        const int64_t *b_off, const int32_t *b_len, const ssw_ref_scoring *sc, ssw_ref_result *out,
                        int32_t *mismatches, int32_t *nops, uint32_t *ops, int maxops, int nthreads) {
  if (nthreads <= 0) nthreads = (int)std::max(1u, std::thread::hardware_concurrency());
  std::atomic<int64_t> next(0);
  auto work = [&]() {
    StripedSmithWaterman::Aligner aligner((uint8_t)sc->match, (uint8_t)sc->mismatch, (uint8_t)sc->gap_open,
                                          (uint8_t)sc->gap_ext, (uint8_t)sc->ambiguity);
    StripedSmithWaterman::Filter filter(true, true, 0, 32767);
    std::string q, r;
    for (;;) {
      int64_t p0 = next.fetch_add(16);
      if (p0 >= npairs) break;
      int64_t p1 = std::min<int64_t>(p0 + 16, npairs);
      for (int64_t p = p0; p < p1; p++) {
        r.assign(a + a_off[p], (size_t)a_len[p]);
        q.assign(b + b_off[p], (size_t)b_len[p]);
        StripedSmithWaterman::Alignment aln;
        aln.Clear();
        int mask_len = std::max((int)(r.size() / 2), 15);
        bool ok = aligner.Align(q.c_str(), r.c_str(), (int)r.size(), filter, &aln, mask_len);
        ssw_ref_result &o = out[p];
        if (!ok) {
          o.score = -1;
          o.ref_begin = o.ref_end = o.query_begin = o.query_end = -2;
          mismatches[p] = -1;
          nops[p] = -1;
          continue;
        }
        o.score = aln.sw_score;
        o.ref_begin = aln.ref_begin;
        o.ref_end = aln.ref_end;
        o.query_begin = aln.query_begin;
        o.query_end = aln.query_end;
        mismatches[p] = aln.mismatches;
        nops[p] = (int32_t)aln.cigar.size();
        for (size_t k = 0; k < aln.cigar.size() && (int)k < maxops; k++) ops[(size_t)p * maxops + k] = aln.cigar[k];
      }
    }
  };
  if (nthreads == 1) {
    work();
  } else {
    std::vector<std::thread> th;
    for (int t = 0; t < nthreads; t++) th.emplace_back(work);
    for (auto &t : th) t.join();
  }
  return 0;
}

int ssw_ref_align_one(const char *query, const char *ref, int32_t ref_len, const ssw_ref_scoring *sc, int32_t *out) {
  StripedSmithWaterman::Aligner aligner((uint8_t)sc->match, (uint8_t)sc->mismatch, (uint8_t)sc->gap_open,
                                        (uint8_t)sc->gap_ext, (uint8_t)sc->ambiguity);
  StripedSmithWaterman::Filter filter(true, false, 0, 32767);
  StripedSmithWaterman::Alignment aln;
  aln.Clear();
  if (!aligner.Align(query, ref, ref_len, filter, &aln, std::max((int)(ref_len / 2), 15))) return -1;
  out[0] = aln.sw_score;
  out[1] = aln.sw_score_next_best;
  out[2] = aln.ref_begin;
  out[3] = aln.ref_end;
  out[4] = aln.query_begin;
  out[5] = aln.query_end;
  out[6] = aln.mismatches;
  return 0;
}

int ssw_ref_align_batch(int64_t npairs, const char *a, const int64_t *a_off, const int32_t *a_len, const char *b,
                        const int64_t *b_off, const int32_t *b_len, const ssw_ref_scoring *sc, ssw_ref_result *out,
                        int nthreads) {
  return ssw_ref_align_batch2(npairs, a, a_off, a_len, b, b_off, b_len, sc, out, nullptr, nthreads);
}

int ssw_ref_align_batch2(int64_t npairs, const char *a, const int64_t *a_off, const int32_t *a_len, const char *b,
                         const int64_t *b_off, const int32_t *b_len, const ssw_ref_scoring *sc, ssw_ref_result *out,
                         int32_t *score2, int nthreads) {
  if (nthreads <= 0) nthreads = (int)std::max(1u, std::thread::hardware_concurrency());
  std::atomic<int64_t> next(0);
  auto work = [&]() {
    StripedSmithWaterman::Aligner aligner((uint8_t)sc->match, (uint8_t)sc->mismatch, (uint8_t)sc->gap_open,
                                          (uint8_t)sc->gap_ext, (uint8_t)sc->ambiguity);
    StripedSmithWaterman::Filter filter(true, false, 0, 32767);
    std::string q, r;
    for (;;) {
      int64_t p0 = next.fetch_add(16);
      if (p0 >= npairs) break;
      int64_t p1 = std::min<int64_t>(p0 + 16, npairs);
      for (int64_t p = p0; p < p1; p++) {
        r.assign(a + a_off[p], (size_t)a_len[p]);
        q.assign(b + b_off[p], (size_t)b_len[p]);
        StripedSmithWaterman::Alignment aln;
        aln.Clear();
        int mask_len = std::max((int)(r.size() / 2), 15);
        bool ok = aligner.Align(q.c_str(), r.c_str(), (int)r.size(), filter, &aln, mask_len);
        ssw_ref_result &o = out[p];
        if (!ok) {
          o.score = -1;
          o.ref_begin = o.ref_end = o.query_begin = o.query_end = -2;
          if (score2) score2[p] = -1;
          continue;
        }
        if (score2) score2[p] = aln.sw_score_next_best;
        o.score = aln.sw_score;
        o.ref_begin = aln.ref_begin;
        o.ref_end = aln.ref_end;
        o.query_begin = aln.query_begin;
        o.query_end = aln.query_end;
      }
    }
  };
  if (nthreads == 1) {
    work();
  } else {
    std::vector<std::thread> th;
    for (int t = 0; t < nthreads; t++) th.emplace_back(work);
    for (auto &t : th) t.join();
  }
  return 0;
}
