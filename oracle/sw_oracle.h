/*
 * ORACLE -- TEST INFRASTRUCTURE ONLY.
 *
 * Scalar, int32, O(m*n) restatement of the result semantics of the reference
 * CPU aligner (StripedSmithWaterman / SSW, src/ssw/ssw_core.cpp and
 * src/ssw/ssw.cpp of mhm2).  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker
 * (or, for the CPU baseline, as the thing timed on host cores).  The product
 * path (mhm2_staging_amd) never links or calls it.
 *
 * Pinning (what checks this restatement independently of its own formula):
 *   - golden vectors produced by the reference SSW itself, compiled from
 *     /root/reference/src/ssw by oracle/Makefile into oracle/_ref/ and recorded
 *     under tests/golden/ by tests/golden/make_golden.py;
 *   - the 18 known-answer cases of test/ssw-test.cpp:151-238;
 *   - brute-force path enumeration on tiny inputs (tests/test_oracle_bruteforce.py),
 *     which defines the optimum and both tie-breaks from first principles;
 *   - for the CIGAR: CIGARs and mismatch counts recorded from the reference
 *     (tests/golden/make_golden_cigar.py), and invariants no restatement
 *     detail can satisfy by accident: the CIGAR spans exactly the query and
 *     the aligned rectangle, rescoring it gives the score, and = / X agree
 *     with the bases (tests/test_oracle_cigar.py).
 */
#ifndef MHMSW_ORACLE_H
#define MHMSW_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  int32_t match;      /* > 0 */
  int32_t mismatch;   /* cost, >= 0, used as -mismatch */
  int32_t gap_open;   /* cost of a gap of length 1 */
  int32_t gap_ext;    /* cost of each further gap position */
  int32_t ambiguity;  /* cost of any comparison involving code 4 (N) */
} oracle_scoring;

/* Result in SSW naming: "ref" is SSW's outer sequence, "query" its striped
 * inner sequence (ssw_core.hpp:46-56, s_align). */
typedef struct {
  int32_t score;
  int32_t ref_begin, ref_end;
  int32_t query_begin, query_end;
} oracle_result;

/* ASCII -> code 0..4, as kBaseTranslation (ssw.cpp:13-26): ACGT/acgt -> 0..3,
 * U/u -> 0, every other byte -> 4. */
int oracle_translate_base(int c);
void oracle_translate(const char *s, int n, int8_t *out);

/* One pair on translated codes.  ref (length m) is SSW-ref, query (length n)
 * is SSW-query. */
void oracle_align_codes(const int8_t *ref, int m, const int8_t *query, int n,
                        const oracle_scoring *sc, oracle_result *out);

/* Batch on ASCII: pair p aligns SSW-ref = a[a_off[p] .. a_off[p]+a_len[p])
 * against SSW-query = b[b_off[p] .. +b_len[p]).  nthreads <= 0 means all
 * hardware threads.  Returns 0. */
int oracle_align_batch(int64_t npairs,
                       const char *a, const int64_t *a_off, const int32_t *a_len,
                       const char *b, const int64_t *b_off, const int32_t *b_len,
                       const oracle_scoring *sc, oracle_result *out, int nthreads);

/* As oracle_align_batch, plus SSW's second-best score (sw_score_next_best for
 * klign's maskLen = max(|SSW-ref|/2, 15)) of each pair into score2[p]. */
int oracle_align_batch2(int64_t npairs,
                        const char *a, const int64_t *a_off, const int32_t *a_len,
                        const char *b, const int64_t *b_off, const int32_t *b_len,
                        const oracle_scoring *sc, oracle_result *out, int32_t *score2, int nthreads);

/* The masking rule alone: rowmax[i] = max_j H(i, j) over the m SSW-ref
 * positions (inner sequence padded by oracle_score2_pad positions scoring 0),
 * score and end_i from the forward pass. */
int oracle_score2_pad(int n, int32_t score, const oracle_scoring *sc);
int32_t oracle_score2_rule(const int32_t *rowmax, int m, int32_t score, int32_t end_i, const oracle_scoring *sc);

/* SSW's CIGAR and mismatch count for one pair whose result r came from
 * oracle_align_codes (klign's call with report_cigar set).  Ops use BAM's
 * encoding, length << 4 | op with op 0 M, 1 I, 2 D, 4 S, 7 '=', 8 X; at most
 * cap are written.  Returns the number of ops (0 when SSW gives no CIGAR), or
 * -1 when SSW's traceback would leave the band or fail. */
int oracle_cigar_codes(const int8_t *ref, int m, const int8_t *query, int n, const oracle_scoring *sc,
                       const oracle_result *r, uint32_t *ops, int cap, int32_t *mismatches);

/* Batch on ASCII (as oracle_align_batch): results, mismatches[p], nops[p] and
 * the ops of pair p at ops[p * maxops ..] (only the first maxops if more). */
int oracle_cigar_batch(int64_t npairs, const char *a, const int64_t *a_off, const int32_t *a_len, const char *b,
                       const int64_t *b_off, const int32_t *b_len, const oracle_scoring *sc, oracle_result *out,
                       int32_t *mismatches, int32_t *nops, uint32_t *ops, int maxops, int nthreads);

#ifdef __cplusplus
}
#endif
#endif
