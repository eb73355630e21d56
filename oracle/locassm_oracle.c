/*
 * ORACLE -- TEST INFRASTRUCTURE ONLY.  See locassm_oracle.h for what this
 * restates and how it is pinned.  Scalar C: a byte-keyed open-addressing
 * table per mer length, one thread per contig.
 */
#define _GNU_SOURCE
#include "locassm_oracle.h"

#include <pthread.h>
#include <stdatomic.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

void oracle_la_default_params(oracle_la_params *p) {
  p->kmer_len = 33;
  p->min_kmer_len = 21;
  p->max_kmer_len = 121;
  p->shift = 8;
  p->walk_len_limit = 300;
  p->qual_offset = 33;
  p->min_qual = 10;
  p->min_hi_qual = 20;
  p->rating_thres = 0;
  p->max_reads = 3000;
  p->min_viable_depth = 0.2;
  p->min_expected_depth = 0.3;
}

/* get_base_rating, kernel.hpp:71-82: doubles compared with integer votes. */
int oracle_la_rating(const oracle_la_params *p, uint32_t nvotes, uint32_t hi, int depth) {
  double min_viable = p->min_viable_depth * depth;
  if (min_viable < 2.0) min_viable = 2.0;
  double min_expected = p->min_expected_depth * depth;
  if (min_expected < 2.0) min_expected = 2.0;
  const double n = (double)nvotes, h = (double)hi;
  if (nvotes == 0) return 0;
  if (nvotes == 1) return 1;
  if (n < min_viable) return 2;
  if (min_expected > n && n >= min_viable && h < min_viable) return 3;
  if (min_expected > n && n >= min_viable && h >= min_viable) return 4;
  if (n >= min_expected && h < min_viable) return 5;
  if (n >= min_expected && min_viable < h && h < min_expected) return 6;
  return 7;
}

/* comp_merbase, kernel.hpp:89-98: true when a ranks before b, and also when
 * all three keys are equal. */
static int la_before(const uint32_t *a, const uint32_t *b) {
  if (a[0] != b[0]) return a[0] > b[0];
  if (a[1] != b[1]) return a[1] > b[1];
  if (a[2] != b[2]) return a[2] > b[2];
  return 1;
}

/* sort_merbase, kernel.hpp:100-110: every ordered pair (i, j) of the four
 * entries in turn, exchanged when entry i ranks before entry j.  Entries
 * carry a fourth word (the base) in la_sort4. */
static void la_sort4(uint32_t e[4][4]) {
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++)
      if (la_before(e[i], e[j])) {
        uint32_t t[4];
        memcpy(t, e[i], sizeof t);
        memcpy(e[i], e[j], sizeof t);
        memcpy(e[j], t, sizeof t);
      }
}

void oracle_la_sort_keys(uint32_t *keys) {
  uint32_t e[4][4];
  for (int i = 0; i < 4; i++) {
    e[i][0] = keys[3 * i];
    e[i][1] = keys[3 * i + 1];
    e[i][2] = keys[3 * i + 2];
    e[i][3] = (uint32_t)i;
  }
  la_sort4(e);
  for (int i = 0; i < 4; i++) {
    keys[3 * i] = e[i][0];
    keys[3 * i + 1] = e[i][1];
    keys[3 * i + 2] = e[i][2];
  }
}

/* set_ext, kernel.hpp:112-161. */
int oracle_la_ext(const oracle_la_params *p, const uint32_t low[4], const uint32_t hi[4], int depth,
                  uint32_t *count) {
  static const char bases[4] = {'A', 'C', 'G', 'T'};
  uint32_t e[4][4];
  for (int i = 0; i < 4; i++) {
    e[i][0] = (uint32_t)oracle_la_rating(p, low[i], hi[i], depth);
    e[i][1] = hi[i];
    e[i][2] = low[i];
    e[i][3] = (uint32_t)bases[i];
  }
  la_sort4(e);
  const int top = (int)e[0][0], runner_up = (int)e[1][0];
  int ext = 'X';
  if (top > p->rating_thres) {
    if (top <= 3) {
      if (runner_up == 0) ext = (int)e[0][3];
    } else if (top < 6) {
      if (runner_up < 3) ext = (int)e[0][3];
    } else if (top == 6) {
      if (runner_up < 4) ext = (int)e[0][3];
    } else {
      if (runner_up < 7) {
        ext = (int)e[0][3];
      } else if (e[2][0] == 7 || e[0][2] == e[1][2]) {
        ext = 'F';
      } else if (e[0][2] > e[1][2]) {
        ext = (int)e[0][3];
      } else if (e[1][2] > e[0][2]) {
        ext = (int)e[1][3];
      }
    }
  }
  uint32_t c = 0;
  for (int i = 0; i < 4; i++)
    if ((int)e[i][3] == ext) {
      c = e[i][2];
      break;
    }
  if (count) *count = c;
  return ext;
}

/* ---- byte-keyed mer table ---------------------------------------------- */

typedef struct {
  const char *key; /* NULL: empty */
  uint32_t low[4], hi[4];
} la_entry;

typedef struct {
  la_entry *e;
  uint64_t cap; /* power of two */
  int k;
} la_table;

static uint64_t la_hash(const char *s, int k) {
  uint64_t h = 1469598103934665603ull; /* FNV-1a */
  for (int i = 0; i < k; i++) {
    h ^= (unsigned char)s[i];
    h *= 1099511628211ull;
  }
  return h;
}

static la_entry *la_find(la_table *t, const char *key, int insert) {
  uint64_t s = la_hash(key, t->k) & (t->cap - 1);
  for (;;) {
    la_entry *e = &t->e[s];
    if (!e->key) {
      if (!insert) return NULL;
      e->key = key;
      return e;
    }
    if (memcmp(e->key, key, (size_t)t->k) == 0) return e;
    s = (s + 1) & (t->cap - 1);
  }
}

static int la_base_index(char c) {
  switch (c) {
    case 'A': return 0;
    case 'C': return 1;
    case 'G': return 2;
    case 'T': return 3;
    default: return -1;
  }
}

/* count_mers, kernel.cpp:310-383, for one side and one mer length. */
static void la_count(const oracle_la_params *p, la_table *t, int k, int nreads, const char *read_seq,
                     const char *read_qual, const int64_t *read_off, const int32_t *read_len,
                     const int64_t *read_idx) {
  const int nr = nreads < p->max_reads ? nreads : p->max_reads;
  uint64_t occ = 0;
  for (int i = 0; i < nr; i++) {
    const int L = read_len[read_idx[i]];
    if (L > k) occ += (uint64_t)(L - k);
  }
  uint64_t cap = 16;
  while (cap < 2 * occ + 16) cap <<= 1;
  t->e = (la_entry *)calloc(cap, sizeof(la_entry));
  t->cap = cap;
  t->k = k;
  for (int i = 0; i < nr; i++) {
    const int64_t r = read_idx[i];
    const int L = read_len[r];
    if (k > L) continue;
    const char *s = read_seq + read_off[r];
    const char *q = read_qual + read_off[r];
    for (int start = 0; start < L - k; start++) {
      la_entry *e = la_find(t, s + start, 1);
      const char ext = s[start + k];
      if (ext == 'N') continue;
      const int qd = (int)(signed char)q[start + k] - p->qual_offset;
      const int b = la_base_index(ext);
      if (b < 0) continue;
      if (qd >= p->min_qual) e->low[b]++;
      if (qd >= p->min_hi_qual) e->hi[b]++;
    }
  }
}

/* walk_mers, kernel.cpp:243-307: buf holds the start mer in buf[0..k) and
 * receives the walk after it. */
static int la_walk(const oracle_la_params *p, la_table *t, char *buf, int k, int depth, int *walk_len) {
  const int limit = p->walk_len_limit;
  uint64_t vcap = 16;
  while (vcap < 2 * (uint64_t)limit + 16) vcap <<= 1;
  const char **visited = (const char **)calloc(vcap, sizeof(char *));
  int res = 'X', n = 0;
  for (int step = 0; step < limit; step++) {
    const char *mer = buf + step;
    uint64_t s = la_hash(mer, k) & (vcap - 1);
    int seen = 0;
    while (visited[s]) {
      if (memcmp(visited[s], mer, (size_t)k) == 0) {
        seen = 1;
        break;
      }
      s = (s + 1) & (vcap - 1);
    }
    if (seen) {
      res = 'R';
      break;
    }
    visited[s] = mer;
    la_entry *e = la_find(t, mer, 0);
    if (!e) {
      res = 'X';
      break;
    }
    const int ext = oracle_la_ext(p, e->low, e->hi, depth, NULL);
    if (ext == 'F' || ext == 'X') {
      res = ext;
      break;
    }
    buf[k + step] = (char)ext;
    n++;
  }
  free(visited);
  *walk_len = n;
  return res;
}

int oracle_la_walk_side(const oracle_la_params *p, const char *seq, int len, double depth, int nreads,
                        const char *read_seq, const char *read_qual, const int64_t *read_off,
                        const int32_t *read_len, const int64_t *read_idx, char *walk, int32_t *term_counts,
                        char *last_term) {
  const int idepth = (int)depth; /* set_ext(int) takes the double depth, kernel.cpp:377 */
  int max_len = p->max_kmer_len < len ? p->max_kmer_len : len;
  int longest = 0, shift = 0, tc[3] = {0, 0, 0};
  char term = 0;
  char *buf = (char *)malloc((size_t)(p->max_kmer_len + p->walk_len_limit + 8));
  for (int k = p->kmer_len; k >= p->min_kmer_len && k <= max_len; k += shift) {
    la_table t;
    la_count(p, &t, k, nreads, read_seq, read_qual, read_off, read_len, read_idx);
    memcpy(buf, seq + (len - k), (size_t)k);
    int wl = 0;
    const int res = la_walk(p, &t, buf, k, idepth, &wl);
    free(t.e);
    term = (char)res;
    tc[res == 'X' ? 0 : res == 'F' ? 1 : 2]++;
    if (wl > longest) {
      memcpy(walk, buf + k, (size_t)wl);
      longest = wl;
    }
    if (res == 'X') {
      if (shift == p->shift) break;
      shift = -p->shift;
    } else {
      if (shift == -p->shift) break;
      if (k > len) break;
      shift = p->shift;
    }
  }
  free(buf);
  if (last_term) *last_term = term;
  if (term_counts)
    for (int i = 0; i < 3; i++) term_counts[i] = tc[i];
  return longest;
}

/* locassm_driver::revcomp, driver.hpp:19-36. */
static int la_comp(char c) {
  switch (c) {
    case 'A': return 'T';
    case 'C': return 'G';
    case 'G': return 'C';
    case 'T': return 'A';
    case 'N': return 'N';
    case 'U': case 'R': case 'Y': case 'K': case 'M': case 'S': case 'W': case 'B': case 'D': case 'H':
    case 'V': return 'N';
    default: return -1;
  }
}

typedef struct {
  const oracle_la_params *p;
  int64_t nctgs;
  const char *ctg_seq;
  const int64_t *ctg_off;
  const int32_t *ctg_len;
  const double *depth;
  const int64_t *left_first, *right_first, *read_idx;
  const char *read_seq, *read_qual;
  const int64_t *read_off;
  const int32_t *read_len;
  int32_t *left_len, *right_len, *terms;
  char *left_ext, *right_ext;
  atomic_long next;
  atomic_int bad;
} la_job;

static void *la_worker(void *arg) {
  la_job *j = (la_job *)arg;
  const oracle_la_params *p = j->p;
  const int W = p->walk_len_limit;
  char *rc = NULL, *tmp = (char *)malloc((size_t)W + 1);
  size_t rc_cap = 0;
  for (;;) {
    const int64_t c = atomic_fetch_add(&j->next, 1);
    if (c >= j->nctgs) break;
    const char *s = j->ctg_seq + j->ctg_off[c];
    const int L = j->ctg_len[c];
    char *rext = j->right_ext + c * (int64_t)W, *lext = j->left_ext + c * (int64_t)W;
    j->right_len[c] = oracle_la_walk_side(p, s, L, j->depth[c], (int)(j->right_first[c + 1] - j->right_first[c]),
                                          j->read_seq, j->read_qual, j->read_off, j->read_len,
                                          j->read_idx + j->right_first[c], rext,
                                          j->terms ? j->terms + 6 * c + 3 : NULL, NULL);
    if ((size_t)L + 1 > rc_cap) {
      rc_cap = (size_t)L + 1;
      rc = (char *)realloc(rc, rc_cap);
    }
    int ok = 1;
    for (int i = 0; i < L; i++) {
      const int v = la_comp(s[L - 1 - i]);
      if (v < 0) ok = 0;
      rc[i] = (char)v;
    }
    if (!ok) {
      atomic_store(&j->bad, 1);
      j->left_len[c] = 0;
      continue;
    }
    const int n = oracle_la_walk_side(p, rc, L, j->depth[c], (int)(j->left_first[c + 1] - j->left_first[c]),
                                      j->read_seq, j->read_qual, j->read_off, j->read_len,
                                      j->read_idx + j->left_first[c], tmp, j->terms ? j->terms + 6 * c : NULL,
                                      NULL);
    for (int i = 0; i < n; i++) lext[i] = (char)la_comp(tmp[n - 1 - i]);
    j->left_len[c] = n;
  }
  free(rc);
  free(tmp);
  return NULL;
}

int oracle_la_extend_batch(const oracle_la_params *p, int64_t nctgs, const char *ctg_seq,
                           const int64_t *ctg_off, const int32_t *ctg_len, const double *depth,
                           const int64_t *left_first, const int64_t *right_first, const int64_t *read_idx,
                           const char *read_seq, const char *read_qual, const int64_t *read_off,
                           const int32_t *read_len, int32_t *left_len, char *left_ext, int32_t *right_len,
                           char *right_ext, int32_t *terms, int nthreads) {
  la_job j;
  memset(&j, 0, sizeof j);
  j.p = p;
  j.nctgs = nctgs;
  j.ctg_seq = ctg_seq;
  j.ctg_off = ctg_off;
  j.ctg_len = ctg_len;
  j.depth = depth;
  j.left_first = left_first;
  j.right_first = right_first;
  j.read_idx = read_idx;
  j.read_seq = read_seq;
  j.read_qual = read_qual;
  j.read_off = read_off;
  j.read_len = read_len;
  j.left_len = left_len;
  j.right_len = right_len;
  j.left_ext = left_ext;
  j.right_ext = right_ext;
  j.terms = terms;
  atomic_init(&j.next, 0);
  atomic_init(&j.bad, 0);
  if (nthreads <= 0) nthreads = (int)sysconf(_SC_NPROCESSORS_ONLN);
  if (nthreads < 1) nthreads = 1;
  if (nthreads == 1 || nctgs < 2) {
    la_worker(&j);
  } else {
    pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * (size_t)nthreads);
    for (int i = 0; i < nthreads; i++) pthread_create(&th[i], NULL, la_worker, &j);
    for (int i = 0; i < nthreads; i++) pthread_join(th[i], NULL);
    free(th);
  }
  return atomic_load(&j.bad) ? -1 : 0;
}
