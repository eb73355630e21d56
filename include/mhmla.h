/*
 * mhmla -- mhm2's local assembly (contig end extension by iterative k-mer
 * walks) on AMD Instinct MI355X (gfx950).  C ABI: plain pointers and sizes.
 *
 * Replaces the CUDA local assembly of mhm2 (src/gpu_loc_assem):
 * locassm_driver::local_assem_driver (driver.hpp:60, driver.cpp:35-319) and
 * its iterative_walks_kernel (kernel.cpp:391-545).  Results are the
 * reference kernel's, byte for byte: the same mers counted (any run of
 * mer_len bytes of the first max_reads reads of a side, N included), the same
 * extension rule (kernel.hpp:71-161), the same walks and mer-length schedule
 * (kernel.cpp:243-307, 459-537), the left side walked on the reverse
 * complement of the contig (driver.cpp:250-266) and returned
 * reverse-complemented for prepending (driver.cpp:297-302).
 *
 * Every entry point returns MHMLA_OK (0) or a negative MHMLA_E* code; the
 * message of the last failure is mhmla_last_error(ctx).  A context is not
 * thread-safe: one caller thread at a time.
 */
#ifndef MHMLA_H
#define MHMLA_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

#define MHMLA_OK 0
#define MHMLA_E_ARG (-1)       /* bad argument or input (message says which) */
#define MHMLA_E_HIP (-2)       /* HIP runtime error */
#define MHMLA_E_STATE (-3)     /* call out of sequence */
#define MHMLA_E_NODEVICE (-4)  /* no usable GPU */
#define MHMLA_E_NOMEM (-6)     /* allocation failed */

/* The reference's compile-time constants (CMakeDefinitions.txt:125-139) and
 * the driver's arguments, as run-time parameters. */
typedef struct {
  int32_t kmer_len;           /* first mer length: local_assem_driver's mer_len */
  int32_t min_kmer_len;       /* LASSM_MIN_KMER_LEN (21) */
  int32_t max_kmer_len;       /* LASSM_MAX_KMER_LEN (121) */
  int32_t shift;              /* LASSM_SHIFT_SIZE (8) */
  int32_t walk_len_limit;     /* walk_len_limit: insert_avg + 2 * insert_stddev */
  int32_t qual_offset;        /* qual_offset (33 or 64) */
  int32_t min_qual;           /* LASSM_MIN_QUAL (10) */
  int32_t min_hi_qual;        /* LASSM_MIN_HI_QUAL (20) */
  int32_t rating_thres;       /* LASSM_RATING_THRES (0) */
  int32_t max_reads;          /* reads counted per side (3000, kernel.cpp:327); at most 65535 */
  double min_viable_depth;    /* LASSM_MIN_VIABLE_DEPTH (0.2) */
  double min_expected_depth;  /* LASSM_MIN_EXPECTED_DEPTH (0.3) */
} mhmla_params;

typedef struct mhmla_ctx mhmla_ctx;

typedef struct {
  double job_ms;        /* device time of the last run (all launches) */
  int64_t contigs;      /* contigs in the staged batch */
  int64_t sides;        /* contig sides walked on the device (sides with reads) */
  int64_t launches;     /* kernel launches of the last run */
  int64_t mers;         /* mer occurrences counted (all sides, all mer lengths) */
  int64_t walk_steps;   /* walk steps taken (all sides, all mer lengths) */
  int64_t walks;        /* walks (one per side and mer length tried) */
  int64_t scratch_bytes;/* device scratch used by the largest launch */
  int64_t phase_cycles[4]; /* shader clock cycles per phase, summed over sides: table clear,
                              mer counting, extension and successor choice, walk */
  int64_t table_resizes; /* mer lengths whose count was redone in a larger table */
} mhmla_stats;

/* Defaults: the reference's constants, kmer_len 33, walk_len_limit 300,
 * qual_offset 33. */
void mhmla_default_params(mhmla_params *p);

/* Replaces locassm_driver::get_gpu_per_node (driver.cpp:28-33). */
int mhmla_device_count(void);

int mhmla_create(int device, mhmla_ctx **out);
void mhmla_destroy(mhmla_ctx *ctx);
const char *mhmla_last_error(const mhmla_ctx *ctx);

/* Copy a batch to the device (the flattening and copies of
 * driver.cpp:156-239).  Contig c is ctg_seq[ctg_off[c] .. +ctg_len[c]) with
 * depth[c]; its left reads are read_idx[left_first[c] .. left_first[c+1]) and
 * its right reads read_idx[right_first[c] .. right_first[c+1]), numbers into
 * read_off / read_len (nreads entries), whose bases and quality bytes are
 * read_seq / read_qual at the same offsets.  Contig bytes must be ACGTN or
 * the IUPAC codes U R Y K M S W B D H V (what driver.hpp:19-36 reverse-
 * complements); reads at most 65535 bases.  The host arrays are not kept. */
int mhmla_stage(mhmla_ctx *ctx, int64_t nctgs, const char *ctg_seq, const int64_t *ctg_off,
                const int32_t *ctg_len, const double *depth, const int64_t *left_first,
                const int64_t *right_first, const int64_t *read_idx, int64_t nreads, const char *read_seq,
                const char *read_qual, const int64_t *read_off, const int32_t *read_len);

/* Extend every staged contig on the device (both iterative_walks_kernel
 * launches of driver.cpp:246 and :273).  Asynchronous. */
int mhmla_run(mhmla_ctx *ctx, const mhmla_params *p);
int mhmla_wait(mhmla_ctx *ctx);

/* Results of the last run.  left_ext / right_ext: nctgs rows of
 * walk_len_limit bytes; row c holds left_len[c] (right_len[c]) bytes: the
 * left extension already reverse-complemented (prepend it) and the right
 * extension (append it).  terms (may be NULL): nctgs x 6 counts of walks
 * ending in a dead end, a fork and a repeat, left side then right side (the
 * walk terminations the CPU path counts, localassm.cpp:780-799). */
int mhmla_results(mhmla_ctx *ctx, int32_t *left_len, char *left_ext, int32_t *right_len, char *right_ext,
                  int32_t *terms);

/* stage + run + wait + results. */
int mhmla_extend(mhmla_ctx *ctx, const mhmla_params *p, int64_t nctgs, const char *ctg_seq,
                 const int64_t *ctg_off, const int32_t *ctg_len, const double *depth, const int64_t *left_first,
                 const int64_t *right_first, const int64_t *read_idx, int64_t nreads, const char *read_seq,
                 const char *read_qual, const int64_t *read_off, const int32_t *read_len, int32_t *left_len,
                 char *left_ext, int32_t *right_len, char *right_ext, int32_t *terms);

int mhmla_get_stats(const mhmla_ctx *ctx, mhmla_stats *out);

#ifdef __cplusplus
}
#endif
#endif
