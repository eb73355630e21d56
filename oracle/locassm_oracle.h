/*
 * ORACLE -- TEST INFRASTRUCTURE ONLY.
 *
 * Scalar C restatement of MHM2's GPU local assembly (src/gpu_loc_assem:
 * locassm_driver::local_assem_driver, driver.cpp:35-319, and
 * iterative_walks_kernel with count_mers / walk_mers / MerFreqs::set_ext,
 * kernel.cpp:243-545 and kernel.hpp:58-161).  Only tests/,
 * __graft_entry__.smoke() and the cpu_baseline leg of bench.py may load it,
 * as the checker (or, for the CPU baseline, as the thing timed on host
 * cores).  The product path never links or calls it.
 *
 * What it restates, per contig side (right: the contig as given; left: its
 * reverse complement, driver.cpp:250-266):
 *   - for mer lengths kmer_len, kmer_len -/+ shift, ... inside
 *     [min_kmer_len, min(max_kmer_len, contig length)] (kernel.cpp:459-537),
 *     count every mer of the first max_reads reads (a mer is any run of
 *     mer_len bytes, N included, kernel.cpp:343-360) with the votes of the
 *     base that follows it, split by quality (kernel.cpp:363-372);
 *   - pick each mer's extension from its votes (kernel.hpp:80-161);
 *   - walk from the contig's last mer until a repeat, a fork, a dead end or
 *     walk_len_limit steps (kernel.cpp:243-307), keep the longest walk
 *     (first of equal lengths), and shift the mer length as the kernel does;
 *   - the left walk is reverse-complemented for prepending (driver.cpp:297-302).
 *
 * Pinning (checks that do not retype this file's formulas):
 *   - hand-derived known answers: the rating thresholds and every branch of
 *     the extension rule on chosen vote counts, derived from the rules as
 *     stated in the reference, with the reasoning written beside each case
 *     (tests/test_oracle_locassm.py);
 *   - closed-form walks: error-free reads of a repeat-free sequence extend a
 *     contig exactly as far as read coverage of (mer_len + 1)-base windows
 *     reaches, for every mer length the shift schedule visits, a figure
 *     computed from the read intervals alone;
 *   - the exchange sort of kernel.hpp:100-110 ends in descending key order
 *     for every vote combination in a range (exhaustive), so the extension
 *     depends only on the keys;
 *   - invariants: read order does not change any result; each walked base
 *     is the extension the table gives the mer before it.
 * The reference's local assembly cannot be built here (CUDA sources, and
 * the CPU path links UPC++), and the reference holds no test or fixture for
 * it.
 */
#ifndef MHMLA_ORACLE_H
#define MHMLA_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  int32_t kmer_len;           /* first mer length (local_assem_driver mer_len) */
  int32_t min_kmer_len;       /* LASSM_MIN_KMER_LEN, 21 */
  int32_t max_kmer_len;       /* LASSM_MAX_KMER_LEN, 121 */
  int32_t shift;              /* LASSM_SHIFT_SIZE, 8 */
  int32_t walk_len_limit;     /* insert_avg + 2 * insert_stddev */
  int32_t qual_offset;        /* 33 or 64 */
  int32_t min_qual;           /* LASSM_MIN_QUAL, 10 */
  int32_t min_hi_qual;        /* LASSM_MIN_HI_QUAL, 20 */
  int32_t rating_thres;       /* LASSM_RATING_THRES, 0 */
  int32_t max_reads;          /* reads counted per side, 3000 (kernel.cpp:327) */
  double min_viable_depth;    /* LASSM_MIN_VIABLE_DEPTH, 0.2 */
  double min_expected_depth;  /* LASSM_MIN_EXPECTED_DEPTH, 0.3 */
} oracle_la_params;

void oracle_la_default_params(oracle_la_params *p);

/* MerBase::get_base_rating (kernel.hpp:71-82) for an integer depth. */
int oracle_la_rating(const oracle_la_params *p, uint32_t nvotes, uint32_t nvotes_hi_q, int depth);

/* MerFreqs::set_ext (kernel.hpp:112-161): low[b], hi[b] are the votes of
 * A, C, G, T; returns 'A', 'C', 'G', 'T', 'F' (fork) or 'X' (none) and, in
 * *count, the low-quality votes of the chosen base (0 for F and X). */
int oracle_la_ext(const oracle_la_params *p, const uint32_t low[4], const uint32_t hi[4], int depth,
                  uint32_t *count);

/* The exchange sort of kernel.hpp:100-110 on 4 (rating, hi, low) keys, in
 * place; keys[3*i + 0..2] = rating, hi, low of entry i. */
void oracle_la_sort_keys(uint32_t *keys);

/* One side: the walk from the end of seq[0..len) with the given reads.
 * Writes the longest walk (walk_len_limit bytes at most) into walk and
 * returns its length; term_counts (may be NULL) gets how many of the side's
 * walks ended in a dead end ('X'), a fork ('F') and a repeat ('R'), as the
 * CPU path counts them (localassm.cpp:780-799), and *last_term the end of
 * the last walk. */
int oracle_la_walk_side(const oracle_la_params *p, const char *seq, int len, double depth, int nreads,
                        const char *read_seq, const char *read_qual, const int64_t *read_off,
                        const int32_t *read_len, const int64_t *read_idx, char *walk, int32_t *term_counts,
                        char *last_term);

/* A batch of contigs.  Contig c is ctg_seq[ctg_off[c] .. +ctg_len[c]); its
 * left reads are read_idx[left_first[c] .. left_first[c+1]) and its right
 * reads read_idx[right_first[c] .. right_first[c+1]), indexes into
 * read_off / read_len, whose bases and quality bytes are read_seq / read_qual
 * at the same offsets.  Outputs per contig: the left extension, already
 * reverse-complemented (prepend it), and the right extension (append it),
 * each at ext + c * walk_len_limit with its length in *_len[c], and (terms
 * may be NULL) the walk terminations X F R of the left then the right side at
 * terms + 6 * c.  Returns 0, or -1 if a contig holds a byte the reverse complement does not
 * map (driver.hpp:19-36 accepts ACGTN and the IUPAC codes U R Y K M S W B D
 * H V, which become N).  nthreads <= 0: all hardware threads. */
int oracle_la_extend_batch(const oracle_la_params *p, int64_t nctgs, const char *ctg_seq,
                           const int64_t *ctg_off, const int32_t *ctg_len, const double *depth,
                           const int64_t *left_first, const int64_t *right_first, const int64_t *read_idx,
                           const char *read_seq, const char *read_qual, const int64_t *read_off,
                           const int32_t *read_len, int32_t *left_len, char *left_ext, int32_t *right_len,
                           char *right_ext, int32_t *terms, int nthreads);

#ifdef __cplusplus
}
#endif
#endif
