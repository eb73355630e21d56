/*
 * Seeded synthetic read / contig-window batches (host side, C ABI).
 *
 * No counterpart in the reference: mhm2 obtains its pairs from k-mer seeding
 * (src/klign.cpp:670-739).  These generators build the benchmark and test
 * workloads with the same shape: a read (std::string) and a contig window
 * (std::string) per pair, ASCII over {A,C,G,T,N}.
 */
#ifndef MHMSW_SYNTH_H
#define MHMSW_SYNTH_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  int64_t npairs;
  char *reads;        /* concatenated reads, NUL after the last byte */
  int64_t *read_off;  /* [npairs] byte offset of each read */
  int32_t *read_len;  /* [npairs] */
  char *ctgs;         /* concatenated contig windows */
  int64_t *ctg_off;   /* [npairs] */
  int32_t *ctg_len;   /* [npairs] */
  int64_t reads_bytes, ctgs_bytes;
} mhmsw_synth_batch;

/* config 1..5 (plain, uneven, length-skewed, repeats, serving pairs).
 * npairs <= 0 and seed == 0 select the defaults (10k/1M/4M/500k/100k pairs,
 * seed = config).  threads <= 0 uses every hardware thread.
 * Returns 0, or < 0 on bad arguments / allocation failure. */
int mhmsw_synth_generate(int config, int64_t npairs, uint64_t seed, int threads, mhmsw_synth_batch *out);
void mhmsw_synth_free(mhmsw_synth_batch *batch);

/* Serving workload: a set of contigs (i.i.d. ACGT, lengths log-uniform in
 * [min_len, max_len]) kept resident on the device, and requests whose pairs
 * are 150 bp reads at 1% substitutions / 0.1% indels against windows of
 * U[300, 1024] bases (clipped to the contig) that contain the read's source. */
typedef struct {
  int64_t nctgs;
  char *seqs;
  int64_t *off;
  int32_t *len;
  int64_t bytes;
} mhmsw_synth_contig_set;

typedef struct {
  int64_t npairs;
  char *reads;
  int64_t *read_off;
  int32_t *read_len;
  int32_t *ctg_id;
  int32_t *win_off;
  int32_t *win_len;
  int64_t reads_bytes;
} mhmsw_synth_window_batch;

int mhmsw_synth_contigs(uint64_t seed, int64_t nctgs, int32_t min_len, int32_t max_len, int threads,
                        mhmsw_synth_contig_set *out);
void mhmsw_synth_contigs_free(mhmsw_synth_contig_set *set);
int mhmsw_synth_window_pairs(const mhmsw_synth_contig_set *ctgs, uint64_t seed, int64_t npairs, int threads,
                             mhmsw_synth_window_batch *out);
void mhmsw_synth_window_pairs_free(mhmsw_synth_window_batch *batch);

#ifdef __cplusplus
}
#endif
#endif
