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

#ifdef __cplusplus
}
#endif
#endif
