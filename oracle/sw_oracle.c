/*
 * ORACLE -- TEST INFRASTRUCTURE ONLY (see sw_oracle.h for who may use it).
 *
 * Restates, as plain scalar int32 code, what the reference SSW computes for
 * klign's call (begin positions on, CIGAR off):
 *
 *  - translation and substitution matrix: ssw.cpp:13-26 (kBaseTranslation),
 *    ssw.cpp:28-55 (BuildSwScoreMatrix: match on the ACGT diagonal, -mismatch
 *    elsewhere among ACGT, -ambiguity in the whole row and column of code 4,
 *    including N against N);
 *  - recurrences: ssw_core.cpp:321-379 (byte) / :661-700 (word): local Gotoh
 *    with E/F = max(prev - gapE, H - gapO), H = max(0, diag + s, E, F).  The
 *    lazy-F rule of ssw_core.cpp:381 (no E update from a lazy-F-corrected H)
 *    only removes paths with a gap of length a directly followed by an
 *    opposite gap of length b >= a.  Such a path costs 2*gapO + (a+b-2)*gapE;
 *    replacing it by a diagonal steps plus one gap of b-a costs at most
 *    a*c + gapO + (b-a-1)*gapE with c = max(mismatch, ambiguity).  Whenever
 *    gapO + gapE >= c and 2*gapE >= c (true for every scoring mhm2 uses:
 *    1/1/1/1/1, 1/3/5/2/2, 2/2/3/1/2, 2/4/4/2/1) the removed paths never beat
 *    the kept ones, so H is textbook local Gotoh; the golden vectors check it;
 *  - forward end selection: ssw_core.cpp:464-486 (end_ref = first outer index
 *    where the global max strictly increases) and :496-506 (end_read = least
 *    inner index holding the max in that column); score 0 leaves end_ref = -1
 *    and end_read = 0 (:240-241 plus the scan of an all-zero column);
 *  - begin selection: ssw_core.cpp:1162-1178 -- the same DP on the outer
 *    prefix [0..ref_end] walked backwards and the reversed inner prefix
 *    [0..query_end], stopping at the first column whose max equals the score;
 *    ref_begin = that outer index, query_begin = query_end - least inner index.
 *
 * Width: SSW switches from 8-bit to 16-bit lanes at score+bias >= 255
 * (ssw_core.cpp:476-477, :1129-1134) without changing results; its 16-bit pass
 * saturates at 32767.  This restatement is exact int32 throughout, so it agrees
 * with SSW for every score below 32767.
 */
#include "sw_oracle.h"

#include <pthread.h>
#include <stdatomic.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

int oracle_translate_base(int c) {
  switch (c) {
    case 'A': case 'a': case 'U': case 'u': return 0;
    case 'C': case 'c': return 1;
    case 'G': case 'g': return 2;
    case 'T': case 't': return 3;
    default: return 4;
  }
}

void oracle_translate(const char *s, int n, int8_t *out) {
  for (int i = 0; i < n; i++) out[i] = (int8_t)oracle_translate_base((unsigned char)s[i]);
}

static inline int32_t subst(int8_t x, int8_t y, const oracle_scoring *sc) {
  if (x == 4 || y == 4) return -sc->ambiguity;
  return x == y ? sc->match : -sc->mismatch;
}

static inline int32_t max2(int32_t a, int32_t b) { return a > b ? a : b; }

/* Forward pass.  Returns the max score; *end_i / *end_j follow the SSW rules
 * (first outer index reaching the max, least inner index in that row).
 * npad > 0 appends npad inner positions that score 0 against everything:
 * SSW's striped profiles pad the inner sequence to a whole number of vectors
 * with the bias (8-bit, ssw_core.cpp:173-197) or 0 (16-bit, :543-564), i.e.
 * with score 0.  No real cell depends on them and they never exceed an
 * earlier maximum, so end points ignore them, but SSW's per-position maxima
 * (maxColumn, which score2 reads) include them; rowmax_out[i] does too. */
static int32_t forward(const int8_t *ref, int m, const int8_t *query, int n, int npad, const oracle_scoring *sc,
                       int32_t *end_i, int32_t *end_j, int32_t *hrow, int32_t *erow, int32_t *rowmax_out) {
  const int32_t gO = sc->gap_open, gE = sc->gap_ext;
  int32_t best = 0, bi = -1, bj = 0;
  /* hrow[j] holds H(i-1, j) while row i is computed; erow[j] holds E(i, j). */
  for (int j = 0; j < n + npad; j++) { hrow[j] = 0; erow[j] = -gO; }
  for (int i = 0; i < m; i++) {
    int32_t diag = 0;      /* H(i-1, j-1) */
    int32_t f = -gO;       /* F(i, j) */
    int32_t hleft = 0;     /* H(i, j-1) */
    int32_t rowmax = 0, rowarg = 0, rowall = 0;
    const int8_t r = ref[i];
    for (int j = 0; j < n + npad; j++) {
      int32_t up = hrow[j];
      int32_t e = max2(erow[j] - gE, up - gO);
      f = max2(f - gE, hleft - gO);
      int32_t h = diag + (j < n ? subst(r, query[j], sc) : 0);
      h = max2(h, e);
      h = max2(h, f);
      h = max2(h, 0);
      erow[j] = e;
      diag = up;
      hrow[j] = h;
      hleft = h;
      if (j < n && h > rowmax) { rowmax = h; rowarg = j; }
      if (h > rowall) rowall = h;
    }
    if (rowmax_out) rowmax_out[i] = rowall;
    if (rowmax > best) { best = rowmax; bi = i; bj = rowarg; }
  }
  *end_i = bi;
  *end_j = bj;
  return best;
}

/* Second-best score as SSW reports it (sw_score_next_best): the largest
 * per-position maximum (maxColumn[i] = max_j H(i, j) over the padded inner
 * sequence, ssw_core.cpp:489/800) over the
 * outer positions outside a window around the end (ssw_core.cpp:512-536 for
 * the 8-bit pass, :822-842 for the 16-bit pass, which starts its upper range
 * one position earlier).  SSW uses its 16-bit pass when score + bias >= 255
 * (ssw_core.cpp:476-477, 1129-1134), bias = max(0, mismatch, ambiguity)
 * (ssw_init, ssw_core.cpp:1076-1083).  klign's maskLen is max(m/2, 15)
 * (klign.cpp:257). */
static int32_t ssw_bias(const oracle_scoring *sc) {
  int32_t bias = 0;
  if (-sc->mismatch < bias) bias = -sc->mismatch;
  if (-sc->ambiguity < bias) bias = -sc->ambiguity;
  return -bias;
}

/* inner positions SSW appends: to 16 (8-bit pass) or 8 (16-bit pass) lanes */
int oracle_score2_pad(int n, int32_t score, const oracle_scoring *sc) {
  const int v = score + ssw_bias(sc) >= 255 ? 8 : 16;
  return (v - n % v) % v;
}

int32_t oracle_score2_rule(const int32_t *rowmax, int m, int32_t score, int32_t end_i, const oracle_scoring *sc) {
  const int word = score + ssw_bias(sc) >= 255;
  const int32_t mask = m / 2 > 15 ? m / 2 : 15;
  int32_t best = 0;
  int32_t edge = end_i - mask > 0 ? end_i - mask : 0;
  for (int32_t i = 0; i < edge; i++)
    if (rowmax[i] > best) best = rowmax[i];
  edge = end_i + mask > m ? m : end_i + mask;
  for (int32_t i = word ? edge : edge + 1; i < m; i++)
    if (rowmax[i] > best) best = rowmax[i];
  return best;
}

static void align_codes_ws(const int8_t *ref, int m, const int8_t *query, int n, const oracle_scoring *sc,
                           oracle_result *out, int32_t *ws, int8_t *rev, int32_t *rowmax, int32_t *score2) {
  out->score = 0;
  out->ref_begin = -1;
  out->ref_end = -1;
  out->query_begin = 0;
  out->query_end = 0;
  if (score2) *score2 = 0;
  if (m <= 0 || n <= 0) return;
  int32_t ei, ej;
  int32_t score = forward(ref, m, query, n, 0, sc, &ei, &ej, ws, ws + n + 16, NULL);
  out->score = score;
  if (score2) {
    int32_t e2, f2;
    forward(ref, m, query, n, oracle_score2_pad(n, score, sc), sc, &e2, &f2, ws, ws + n + 16, rowmax);
    *score2 = oracle_score2_rule(rowmax, m, score, ei, sc);
  }
  if (score <= 0) return; /* sentinels above: ssw_core.cpp:240-241 and the empty reverse pass */
  out->ref_end = ei;
  out->query_end = ej;
  /* reversed prefixes, ssw_core.cpp:1055-1066 and :1167-1173 */
  int8_t *rr = rev, *rq = rev + (ei + 1);
  for (int k = 0; k <= ei; k++) rr[k] = ref[ei - k];
  for (int l = 0; l <= ej; l++) rq[l] = query[ej - l];
  int32_t ki, lj;
  int32_t s2 = forward(rr, ei + 1, rq, ej + 1, 0, sc, &ki, &lj, ws, ws + n + 16, NULL);
  (void)s2; /* equals score: every reversed-prefix alignment is an alignment of the pair */
  out->ref_begin = ei - ki;
  out->query_begin = ej - lj;
}

void oracle_align_codes(const int8_t *ref, int m, const int8_t *query, int n, const oracle_scoring *sc,
                        oracle_result *out) {
  int32_t *ws = (int32_t *)malloc(sizeof(int32_t) * 2 * (size_t)((n > 0 ? n : 1) + 16));
  int8_t *rev = (int8_t *)malloc((size_t)(m > 0 ? m : 1) + (size_t)(n > 0 ? n : 1));
  align_codes_ws(ref, m, query, n, sc, out, ws, rev, NULL, NULL);
  free(rev);
  free(ws);
}

typedef struct {
  int64_t npairs;
  const char *a, *b;
  const int64_t *a_off, *b_off;
  const int32_t *a_len, *b_len;
  const oracle_scoring *sc;
  oracle_result *out;
  int32_t *score2;  /* optional */
  atomic_llong next;
} batch_job;

static void *batch_worker(void *arg) {
  batch_job *job = (batch_job *)arg;
  size_t cap_a = 0, cap_b = 0;
  int8_t *ca = NULL, *cb = NULL, *rev = NULL;
  int32_t *ws = NULL, *rowmax = NULL;
  for (;;) {
    int64_t p = atomic_fetch_add(&job->next, 64);
    if (p >= job->npairs) break;
    int64_t pe = p + 64 < job->npairs ? p + 64 : job->npairs;
    for (; p < pe; p++) {
      int m = job->a_len[p], n = job->b_len[p];
      if ((size_t)m + 1 > cap_a || (size_t)n + 1 > cap_b) {
        cap_a = cap_a > (size_t)m + 1 ? cap_a : (size_t)m + 1;
        cap_b = cap_b > (size_t)n + 1 ? cap_b : (size_t)n + 1;
        free(ca); free(cb); free(rev); free(ws); free(rowmax);
        ca = (int8_t *)malloc(cap_a);
        cb = (int8_t *)malloc(cap_b);
        rev = (int8_t *)malloc(cap_a + cap_b);
        ws = (int32_t *)malloc(sizeof(int32_t) * 2 * (cap_b + 16));
        rowmax = (int32_t *)malloc(sizeof(int32_t) * cap_a);
      }
      oracle_translate(job->a + job->a_off[p], m, ca);
      oracle_translate(job->b + job->b_off[p], n, cb);
      align_codes_ws(ca, m, cb, n, job->sc, &job->out[p], ws, rev, rowmax, job->score2 ? &job->score2[p] : NULL);
    }
  }
  free(ca); free(cb); free(rev); free(ws); free(rowmax);
  return NULL;
}

int oracle_align_batch(int64_t npairs, const char *a, const int64_t *a_off, const int32_t *a_len, const char *b,
                       const int64_t *b_off, const int32_t *b_len, const oracle_scoring *sc, oracle_result *out,
                       int nthreads) {
  return oracle_align_batch2(npairs, a, a_off, a_len, b, b_off, b_len, sc, out, NULL, nthreads);
}

int oracle_align_batch2(int64_t npairs, const char *a, const int64_t *a_off, const int32_t *a_len, const char *b,
                        const int64_t *b_off, const int32_t *b_len, const oracle_scoring *sc, oracle_result *out,
                        int32_t *score2, int nthreads) {
  batch_job job;
  job.npairs = npairs;
  job.a = a; job.b = b;
  job.a_off = a_off; job.b_off = b_off;
  job.a_len = a_len; job.b_len = b_len;
  job.sc = sc;
  job.out = out;
  job.score2 = score2;
  atomic_init(&job.next, 0);
  if (nthreads <= 0) nthreads = (int)sysconf(_SC_NPROCESSORS_ONLN);
  if (nthreads < 1) nthreads = 1;
  if (nthreads == 1) {
    batch_worker(&job);
    return 0;
  }
  pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * (size_t)nthreads);
  for (int t = 0; t < nthreads; t++) pthread_create(&th[t], NULL, batch_worker, &job);
  for (int t = 0; t < nthreads; t++) pthread_join(th[t], NULL);
  free(th);
  return 0;
}
