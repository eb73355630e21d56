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

/* ------------------------------------------------------------------------
 * CIGAR and mismatch count, as klign gets them when it reports CIGARs
 * (Filter report_begin_position = report_cigar = true, score_filter 0,
 * distance_filter 32767: klign.cpp:435 with ssw.hpp:59-64).
 *
 *  - ssw_align's CIGAR step (ssw_core.cpp:1180-1192): the rectangle from the
 *    begin to the end point, rows over the SSW-query segment and columns over
 *    the SSW-ref segment, goes through banded_sw with band half-width
 *    |columns - rows| + 1; no CIGAR when either span exceeds the distance
 *    filter (:1180);
 *  - banded_sw (ssw_core.cpp:854-1053): local affine DP restricted to
 *    |j - i| <= w, every cell outside the band (and above row 0 or left of
 *    column 0) counting as H = E = F = 0, E and F not clamped, H =
 *    max(max(E,0), max(F,0), H(i-1,j-1) + s).  Choices: E opened from H when
 *    H - gO > E - gE strictly (else extended), F likewise; H takes the diagonal
 *    when max(E,F clamped) <= diagonal, else E when E > F strictly, else F.
 *    The band doubles until its maximum reaches the score.  The traceback
 *    starts at the last cell in state H and follows the choices while the
 *    row index is above 0 (diagonal = M, E = I, F = D); the cell where it
 *    stops counts as one more M;
 *  - ConvertAlignment (ssw.cpp:57-93): soft clips of the SSW-query before the
 *    begin and after the end;
 *  - CalculateNumberMismatch (ssw.cpp:129-254): each M position compares the
 *    translated codes under the CIGAR's own walk from (ref_begin,
 *    query_begin) -- '=' when equal (N against N too), 'X' otherwise --
 *    and mismatches = X + I + D positions.
 *
 * A score-0 pair has ref_begin = -1, so SSW compares SSW-query[0] with the
 * byte before its translated SSW-ref buffer: the top byte of glibc's chunk
 * size word, 0 (code A) for every allocation below 2^56 bytes.  That byte is
 * taken as 0 here.
 * ---------------------------------------------------------------------- */
enum { CG_M = 0, CG_I = 1, CG_D = 2, CG_S = 4, CG_EQ = 7, CG_X = 8 };
#define CG_WRITTEN 0x80u
#define CG_E_OPEN 0x08u
#define CG_F_OPEN 0x10u

/* One band pass; dir[i * rl + j] gets the choices of every band cell: bits
 * 0-2 the value SSW's traceback reads in state H (1 diagonal, 2/3 E
 * extended/opened, 4/5 F extended/opened), CG_E_OPEN, CG_F_OPEN.  Returns the
 * band's maximum H. */
static int32_t cigar_band(const int8_t *R, int rl, const int8_t *Q, int ql, int bw, const oracle_scoring *sc,
                          uint8_t *dir, int32_t *hp, int32_t *ep, int32_t *hc, int32_t *ec) {
  const int32_t gO = sc->gap_open, gE = sc->gap_ext;
  int32_t best = 0;
  memset(dir, 0, (size_t)ql * (size_t)rl);
  for (int j = 0; j < rl; j++) { hp[j] = 0; ep[j] = 0; }
  for (int i = 0; i < ql; i++) {
    const int beg = i - bw > 0 ? i - bw : 0;
    const int end = i + bw < rl - 1 ? i + bw : rl - 1;
    for (int j = 0; j < rl; j++) { hc[j] = 0; ec[j] = 0; }
    int32_t f = 0, hleft = 0;
    for (int j = beg; j <= end; j++) {
      /* hp/ep hold row i-1 with zeros outside its band (all zeros for i = 0) */
      const int32_t hup = hp[j], eup = ep[j];
      const int32_t hdiag = j > 0 ? hp[j - 1] : 0;
      const int32_t eo = hup - gO, ee = eup - gE;
      const int32_t e = eo > ee ? eo : ee;
      const int32_t fo = hleft - gO, fe = f - gE;
      f = fo > fe ? fo : fe;
      const int32_t e1 = e > 0 ? e : 0, f1 = f > 0 ? f : 0;
      const int32_t t1 = e1 > f1 ? e1 : f1;
      const int32_t t2 = hdiag + subst(R[j], Q[i], sc);
      const int32_t h = t1 > t2 ? t1 : t2;
      uint8_t v;
      if (t1 <= t2) v = 1;
      else if (e1 > f1) v = eo > ee ? 3 : 2;
      else v = fo > fe ? 5 : 4;
      dir[(size_t)i * rl + j] = (uint8_t)(v | CG_WRITTEN | (eo > ee ? CG_E_OPEN : 0) | (fo > fe ? CG_F_OPEN : 0));
      hc[j] = h;
      ec[j] = e;
      hleft = h;
      if (h > best) best = h;
    }
    int32_t *t = hp; hp = hc; hc = t;
    t = ep; ep = ec; ec = t;
  }
  return best;
}

typedef struct {
  uint32_t *ops;
  int cap, n;
  int cur_op;
  uint32_t cur_len;
  int32_t mism;
} cigar_out;

static void cg_flush(cigar_out *o) {
  if (o->cur_len == 0) return;
  if (o->n < o->cap) o->ops[o->n] = (o->cur_len << 4) | (uint32_t)o->cur_op;
  o->n++;
  if (o->cur_op == CG_X || o->cur_op == CG_I || o->cur_op == CG_D) o->mism += (int32_t)o->cur_len;
  o->cur_len = 0;
}

static void cg_push(cigar_out *o, int op, uint32_t len) {
  if (len == 0) return;
  if (o->cur_len && o->cur_op != op) cg_flush(o);
  o->cur_op = op;
  o->cur_len += len;
}

int oracle_cigar_codes(const int8_t *ref, int m, const int8_t *query, int n, const oracle_scoring *sc,
                       const oracle_result *r, uint32_t *ops, int cap, int32_t *mismatches) {
  cigar_out o = {ops, cap, 0, 0, 0, 0};
  *mismatches = 0;
  if (n <= 0) return 0; /* SSW refuses an empty query (ssw.cpp:450) */
  const int rb = r->ref_begin, re = r->ref_end, qb = r->query_begin, qe = r->query_end;
  if (re - rb > 32767 || qe - qb > 32767) return 0; /* distance filter: no CIGAR, 0 mismatches */
  const int rl = re - rb + 1, ql = qe - qb + 1;
  /* trace steps in trace order: op and cell */
  int nst = 0;
  int32_t *st_op = NULL, *st_i = NULL, *st_j = NULL;
  int jfin = 0;
  if (r->score > 0) {
    uint8_t *dir = (uint8_t *)malloc((size_t)ql * (size_t)rl);
    int32_t *rows = (int32_t *)malloc(sizeof(int32_t) * 4 * (size_t)rl);
    int bw = (rl > ql ? rl - ql : ql - rl) + 1;
    int32_t mx = 0;
    for (;;) {
      int32_t b = cigar_band(ref + rb, rl, query + qb, ql, bw, sc, dir, rows, rows + rl, rows + 2 * rl, rows + 3 * rl);
      if (b > mx) mx = b;
      if (mx >= r->score) break;
      if (bw > 2 * (rl + ql) + 2) { free(dir); free(rows); return -1; } /* score unreachable: inconsistent input */
      bw *= 2;
    }
    st_op = (int32_t *)malloc(sizeof(int32_t) * 3 * (size_t)(rl + ql + 2));
    st_i = st_op + (rl + ql + 2);
    st_j = st_i + (rl + ql + 2);
    int i = ql - 1, j = rl - 1, state = 2; /* 0 E, 1 F, 2 H as in ssw_core.cpp:963-1009 */
    while (i > 0) {
      if (j < 0 || j - i > bw || i - j > bw) { free(dir); free(rows); free(st_op); return -1; }
      const uint8_t d = dir[(size_t)i * rl + j];
      if (!(d & CG_WRITTEN)) { free(dir); free(rows); free(st_op); return -1; }
      int v = state == 2 ? (d & 7) : state == 0 ? ((d & CG_E_OPEN) ? 3 : 2) : ((d & CG_F_OPEN) ? 5 : 4);
      st_i[nst] = i;
      st_j[nst] = j;
      switch (v) {
        case 1: st_op[nst] = CG_M; i--; j--; state = 2; break;
        case 2: st_op[nst] = CG_I; i--; state = 0; break;
        case 3: st_op[nst] = CG_I; i--; state = 2; break;
        case 4: st_op[nst] = CG_D; j--; state = 1; break;
        case 5: st_op[nst] = CG_D; j--; state = 2; break;
        default: free(dir); free(rows); free(st_op); return -1;
      }
      nst++;
    }
    jfin = j;
    free(dir);
    free(rows);
  }
  /* forward order: head clip, the cell the trace stopped at (one M), then the
   * trace steps backwards; M positions compared along the CIGAR's own walk */
  if (qb > 0) cg_push(&o, CG_S, (uint32_t)qb);
  for (int k = nst; k >= 0; k--) {
    const int op = k == nst ? CG_M : st_op[k];
    const int i = k == nst ? 0 : st_i[k], j = k == nst ? jfin : st_j[k];
    if (op == CG_M) {
      const int rj = rb + j - jfin;
      const int8_t rc = rj >= 0 ? ref[rj] : 0;
      cg_push(&o, rc == query[qb + i] ? CG_EQ : CG_X, 1);
    } else {
      cg_push(&o, op, 1);
    }
  }
  cg_flush(&o);
  if (n - qe - 1 > 0) cg_push(&o, CG_S, (uint32_t)(n - qe - 1));
  cg_flush(&o);
  free(st_op);
  *mismatches = o.mism;
  return o.n;
}

typedef struct {
  int64_t npairs;
  const char *a, *b;
  const int64_t *a_off, *b_off;
  const int32_t *a_len, *b_len;
  const oracle_scoring *sc;
  oracle_result *out;
  int32_t *mism, *nops;
  uint32_t *ops;
  int maxops;
  atomic_llong next;
} cigar_job;

static void *cigar_worker(void *arg) {
  cigar_job *job = (cigar_job *)arg;
  for (;;) {
    int64_t p = atomic_fetch_add(&job->next, 16);
    if (p >= job->npairs) break;
    int64_t pe = p + 16 < job->npairs ? p + 16 : job->npairs;
    for (; p < pe; p++) {
      int m = job->a_len[p], n = job->b_len[p];
      int8_t *ca = (int8_t *)malloc((size_t)m + 1), *cb = (int8_t *)malloc((size_t)n + 1);
      oracle_translate(job->a + job->a_off[p], m, ca);
      oracle_translate(job->b + job->b_off[p], n, cb);
      oracle_align_codes(ca, m, cb, n, job->sc, &job->out[p]);
      job->nops[p] = oracle_cigar_codes(ca, m, cb, n, job->sc, &job->out[p], job->ops + (size_t)p * job->maxops,
                                        job->maxops, &job->mism[p]);
      free(ca);
      free(cb);
    }
  }
  return NULL;
}

int oracle_cigar_batch(int64_t npairs, const char *a, const int64_t *a_off, const int32_t *a_len, const char *b,
                       const int64_t *b_off, const int32_t *b_len, const oracle_scoring *sc, oracle_result *out,
                       int32_t *mismatches, int32_t *nops, uint32_t *ops, int maxops, int nthreads) {
  cigar_job job;
  job.npairs = npairs;
  job.a = a; job.b = b;
  job.a_off = a_off; job.b_off = b_off;
  job.a_len = a_len; job.b_len = b_len;
  job.sc = sc;
  job.out = out;
  job.mism = mismatches;
  job.nops = nops;
  job.ops = ops;
  job.maxops = maxops;
  atomic_init(&job.next, 0);
  if (nthreads <= 0) nthreads = (int)sysconf(_SC_NPROCESSORS_ONLN);
  if (nthreads < 1) nthreads = 1;
  pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * (size_t)nthreads);
  for (int t = 0; t < nthreads; t++) pthread_create(&th[t], NULL, cigar_worker, &job);
  for (int t = 0; t < nthreads; t++) pthread_join(th[t], NULL);
  free(th);
  return 0;
}
