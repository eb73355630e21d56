/*
 * mhmsw -- batched Smith-Waterman for mhm2's klign read-to-contig stage on
 * AMD Instinct MI355X (gfx950).  C ABI: plain pointers and sizes only.
 *
 * Results follow the reference CPU aligner's semantics exactly (SSW,
 * src/ssw/ssw_core.cpp:1104-1200 as klign calls it at src/klign.cpp:251-273):
 * local alignment with affine gaps (a gap of length L costs
 * gap_open + (L-1)*gap_ext), SSW's N handling and SSW's tie-breaks.
 *
 * Every entry point returns MHMSW_OK (0) or a negative MHMSW_E* code; the
 * message of the last failure on a context is mhmsw_last_error(ctx).
 * A context is not thread-safe: one caller thread at a time, at most one
 * batch in flight, as with adept_sw::GPUDriver (src/klign.cpp:586-591).
 */
#ifndef MHMSW_H
#define MHMSW_H
#include <stddef.h>
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

#define MHMSW_OK 0
#define MHMSW_E_ARG (-1)       /* bad argument */
#define MHMSW_E_HIP (-2)       /* HIP runtime error (message has the HIP error string) */
#define MHMSW_E_STATE (-3)     /* call out of sequence (e.g. results before a batch) */
#define MHMSW_E_NODEVICE (-4)  /* no usable GPU */
#define MHMSW_E_SCORING (-5)   /* scoring outside what SSW semantics support */
#define MHMSW_E_NOMEM (-6)     /* allocation failed */
#define MHMSW_E_INTERNAL (-7)  /* an internal consistency check failed */

/* Execution forms.  FAST: packed 16-bit persistent length-bucketed kernels with
 * an exact 32-bit rerun for any pair whose score could overflow.  PLAIN: one
 * 32-bit launch per batch, one wavefront per pair (the straightforward form). */
#define MHMSW_MODE_FAST 0
#define MHMSW_MODE_PLAIN 1

typedef struct mhmsw_ctx mhmsw_ctx;

/* Costs are positive magnitudes, as in mhm2's AlnScoring (include/klign.hpp:53-62). */
typedef struct {
  int32_t match;      /* 1..127 */
  int32_t mismatch;   /* 0..127 */
  int32_t gap_open;   /* 1..127, >= gap_ext */
  int32_t gap_ext;    /* 1..127 */
  int32_t ambiguity;  /* 0..127, cost of any comparison with N */
} mhmsw_scoring;

/* One alignment, 0-based inclusive coordinates.  A pair with score 0 reports
 * read/ctg begin and end as SSW does for its SSW-ref (-1, -1) and SSW-query
 * (0, 0) roles. */
typedef struct {
  int32_t score;
  int32_t read_begin, read_end;
  int32_t ctg_begin, ctg_end;
} mhmsw_result;

/* Work classes of the forward pass: 0..7 packed 16-bit, rows <= 32*P for
 * P = 1..8 (class P-1; class 4 ends at 152 rows and class 5 holds
 * 153..192), 8..15 packed 16-bit, rows <= 128*P, 16 the 32-bit form.  A
 * packed class runs 16 (classes 0..7) or 64 (8..15) lanes per pair of P
 * register pairs each, or, when it holds enough pairs, fewer lanes of more
 * registers: mhmsw_stats.class_lanes / class_regs give the shape that ran. */
#define MHMSW_NUM_CLASSES 17

/* Per-batch counters of the last completed batch. */
typedef struct {
  int64_t npairs;
  int64_t cells;            /* forward cells, sum of |read| * |ctg| */
  int64_t reverse_cells;    /* cells of the reverse (begin) pass */
  int64_t fast_pairs;       /* pairs whose forward pass ran packed 16-bit */
  int64_t plain_pairs;      /* pairs that ran 32-bit from the start */
  int64_t rerun_pairs;      /* pairs rerun in 32-bit after the 16-bit pass */
  double job_ms;            /* device time of the whole job (events on the context stream) */
  /* The per-phase and per-class times below are measured for jobs launched
   * eagerly (mhmsw_set_graphs(ctx, 0)); a replayed graph reports job_ms only. */
  double fwd_fast_ms;       /* device time of the forward packed kernels */
  double fwd_plain_ms;      /* device time of the forward 32-bit kernels (incl. reruns) */
  double rev_ms;            /* device time of the reverse pass */
  double prep_ms;           /* unpack + scheduling kernels */
  /* forward pass per work class: pairs, cells, device time of its launch */
  int64_t class_pairs[MHMSW_NUM_CLASSES];
  int64_t class_cells[MHMSW_NUM_CLASSES];
  double class_ms[MHMSW_NUM_CLASSES];
  double cigar_ms;          /* device time of the CIGAR pass (eager launches) */
  int64_t cigar_reruns;     /* CIGAR passes rerun with larger pools */
  double window_ms;         /* device time of the window extraction of a windowed batch (eager launches) */
  int64_t window_bases;     /* bases of the extracted windows */
  int32_t packed_scoring;   /* 1: the scoring fits the packed 16-bit form; 0: every pair runs 32-bit,
                               about 10x slower (16*(match+gap_open) > 255 or gap_open below mismatch
                               or ambiguity) */
  int64_t device_bytes;     /* device memory the context holds (its buffers only grow) */
  /* host time of the batch's staging and launch (steady clock): packing to
   * 2-bit codes, laying out the pinned staging and enqueueing its copies, and
   * enqueueing the job */
  double host_pack_ms, host_upload_ms, host_launch_ms;
  /* the packed kernel that ran each forward class: lanes per pair (2, 4, 8,
   * 16 or 64) and packed registers (2 rows each) per lane; 0 for the 32-bit
   * class.  A class of many pairs runs narrower groups with more rows per
   * lane, and one whose rows are all short enough one register fewer. */
  int32_t class_lanes[MHMSW_NUM_CLASSES];
  int32_t class_regs[MHMSW_NUM_CLASSES];
  /* the forward kernel's scale of H (H << class_tag_bits, the low bits
   * tagging the step of a row's maximum): 4, or 5 for the narrow shapes when
   * the scoring and the class's rows allow it; 0 for the 32-bit class */
  int32_t class_tag_bits[MHMSW_NUM_CLASSES];
} mhmsw_stats;

/* Scheduling parameters of the packed form (no counterpart in the reference,
 * whose launch shape is fixed: one block per pair, driver.cpp:347-368).
 *   blocks_per_cu  persistent 256-thread blocks launched per CU (wavefronts
 *                  per CU = 4 x blocks_per_cu), 1..16, for the forward and
 *                  rectangle kernels (the reverse band kernels run 2)
 *   g64_min_rows   pairs whose shorter side has at least this many bases (and
 *                  at most 1024) take the classes 8..15, shorter ones the
 *                  classes 0..7; 33..257 (257: classes 0..7 up to 256 rows)
 *   bin_shift      column-length bucket width is 1 << bin_shift bases, 3..7 */
typedef struct {
  int32_t blocks_per_cu;
  int32_t g64_min_rows;
  int32_t bin_shift;
} mhmsw_tuning;

/* --- devices (replace adept_sw free functions, src/adept-sw/driver.cpp:28-77) --- */
int mhmsw_device_count(int *count);
int mhmsw_device_total_mem(int device_id, size_t *bytes);
int mhmsw_device_name(int device_id, char *buf, size_t buflen);

/* Sustained packed 16-bit integer VALU rate of a device: a chain of
 * v_pk_add_u16 / v_pk_max_i16 on 32 wavefronts per CU for about 100 ms (the
 * roofline of the alignment kernels; no counterpart in the reference). */
typedef struct {
  double lane_ops_per_s; /* int16 lane-ops per second (2 per lane per packed instruction) */
  double clock_mhz;      /* shader clock sustained during the run (cycle counter / wall clock) */
  double max_clock_mhz;  /* the device's nominal maximum clock */
  double kernel_ms;
  int32_t cus;
} mhmsw_valu_peak;
int mhmsw_measure_valu_peak(int device_id, mhmsw_valu_peak *out);

/* --- context (replaces GPUDriver::init / ~GPUDriver, driver.cpp:198-268) --- */
int mhmsw_create(int device_id, mhmsw_ctx **out);
int mhmsw_destroy(mhmsw_ctx *ctx);
const char *mhmsw_last_error(const mhmsw_ctx *ctx);
/* Scoring incl. ambiguity (GPUDriver::init takes no ambiguity, driver.hpp:45-46). */
int mhmsw_set_scoring(mhmsw_ctx *ctx, const mhmsw_scoring *scoring);
/* 1 (default, klign's CPU call klign.cpp:257): the read is SSW's outer
 * sequence ("ref") and the contig window its striped sequence ("query").
 * 0: the opposite roles.  Tie-breaks depend on this choice. */
int mhmsw_set_orientation(mhmsw_ctx *ctx, int read_is_ssw_ref);
int mhmsw_set_mode(mhmsw_ctx *ctx, int mode);
/* 1: jobs also compute SSW's second-best score (Alignment::sw_score_next_best,
 * ssw_core.cpp:512-536 / :822-842 with klign's maskLen = max(|SSW-ref|/2, 15),
 * klign.cpp:257), which klign's GPU path leaves at 0 (klign.cpp:378-379).
 * Needs every SSW-query sequence to be at most 8177 bases.  0 (default): off. */
int mhmsw_set_score2(mhmsw_ctx *ctx, int enable);
/* 1: jobs also compute SSW's CIGAR and mismatch count of every pair
 * (Alignment::cigar and ::mismatches of klign's call with report_cigar set,
 * klign.cpp:251-273 and :435: ssw_core.cpp:1180-1192 banded_sw :854-1053,
 * ssw.cpp:57-93 ConvertAlignment, :129-254 CalculateNumberMismatch), which
 * ADEPT's GPU path cannot give (klign.cpp:589-599 runs such batches on the
 * CPU; its GPU mapping leaves them at 0 / "*", :378-385).  0 (default): off. */
int mhmsw_set_cigar(mhmsw_ctx *ctx, int enable);
/* 1 (default): the job of a staged batch (unpack, forward, reruns, reverse,
 * finalize: ~40 launches) runs eagerly the first time its geometry is seen,
 * is captured as a hipGraph when the next job repeats that geometry, and is
 * replayed while it stays the same (batches that change every call never pay
 * for a capture).  0: launches are always issued one by one. */
int mhmsw_set_graphs(mhmsw_ctx *ctx, int enable);
/* No counterpart in the reference (its driver times nothing).  1 (default):
 * an eagerly launched job records events around each phase and class, which
 * mhmsw_get_stats reports.  0: only the job's own two events, as for a
 * replayed graph (job_ms only), which saves their host time in every call;
 * the drop-in adept_sw::GPUDriver turns it off. */
int mhmsw_set_phase_timing(mhmsw_ctx *ctx, int enable);
int mhmsw_set_tuning(mhmsw_ctx *ctx, const mhmsw_tuning *tuning);
int mhmsw_get_tuning(const mhmsw_ctx *ctx, mhmsw_tuning *tuning);
/* Times candidate tunings on the staged batch (coordinate descent over the
 * three parameters, best of `repeats` jobs each), keeps the fastest and
 * returns it.  Results do not depend on the tuning. */
int mhmsw_autotune(mhmsw_ctx *ctx, int repeats, mhmsw_tuning *best);

/* --- batches (replace run_kernel_forwards / run_kernel_backwards,
 *     kernel_is_done / kernel_block and get_aln_results, driver.cpp:270-417) ---
 * Pair p aligns reads[read_off[p] .. +read_len[p]) with
 * ctgs[ctg_off[p] .. +ctg_len[p]) (ASCII; A/C/G/T/U any case, anything else is N).
 * mhmsw_stage packs and copies the batch to the device and returns when the
 * host buffers may be reused; mhmsw_align_staged starts the job on the staged
 * batch (it may be called again on the same staged batch);
 * mhmsw_submit = stage + align_staged, and its job also copies the results to
 * pinned host memory as its last step, so mhmsw_results after it is a host
 * copy.  Results stay valid until the next stage/submit. */
int mhmsw_stage(mhmsw_ctx *ctx, int64_t npairs, const char *reads, const int64_t *read_off, const int32_t *read_len,
                const char *ctgs, const int64_t *ctg_off, const int32_t *ctg_len);
int mhmsw_align_staged(mhmsw_ctx *ctx);
int mhmsw_submit(mhmsw_ctx *ctx, int64_t npairs, const char *reads, const int64_t *read_off, const int32_t *read_len,
                 const char *ctgs, const int64_t *ctg_off, const int32_t *ctg_len);
/* mhmsw_stage / mhmsw_submit with every sequence given by its own pointer:
 * read p is read_len[p] bytes at reads[p], window p ctg_len[p] bytes at
 * ctgs[p] (a pointer may be NULL when its length is 0).  This is the form a
 * caller holding std::string objects has without copying them; it replaces
 * the flattening of GPUDriver::run_kernel_forwards (src/adept-sw/driver.cpp:
 * 300-339, two vector copies and a byte memcpy per batch).  The strings are
 * read before the call returns. */
int mhmsw_stage_seqs(mhmsw_ctx *ctx, int64_t npairs, const char *const *reads, const int32_t *read_len,
                     const char *const *ctgs, const int32_t *ctg_len);
int mhmsw_submit_seqs(mhmsw_ctx *ctx, int64_t npairs, const char *const *reads, const int32_t *read_len,
                      const char *const *ctgs, const int32_t *ctg_len);
/* 1 when the job has finished, 0 while it runs, < 0 on error. */
int mhmsw_query(mhmsw_ctx *ctx);
int mhmsw_wait(mhmsw_ctx *ctx);
/* Copies npairs results (npairs must equal the batch size) to host memory. */
int mhmsw_results(mhmsw_ctx *ctx, mhmsw_result *out, int64_t npairs);
/* score2 of each pair of the last batch (which must have run with score2 on). */
int mhmsw_results_score2(mhmsw_ctx *ctx, int32_t *out, int64_t npairs);
/* CIGARs of the last batch (which must have run with the CIGAR pass on):
 * mismatches[p] (may be NULL); op_off[0..npairs] prefix offsets of each
 * pair's ops in ops (may be NULL); ops (may be NULL: only *total_ops and the
 * other outputs are filled) receives every pair's ops in pair order, BAM
 * encoding: length << 4 | op, op 0 M, 1 I, 2 D, 4 S, 7 '=', 8 X (SSW's
 * cigar after its =/X split: soft clips of the SSW-query, '=', 'X', I, D);
 * ops_cap = its capacity in ops. */
int mhmsw_results_cigar(mhmsw_ctx *ctx, int64_t npairs, int32_t *mismatches, int64_t *op_off, uint32_t *ops,
                        int64_t ops_cap, int64_t *total_ops);
int mhmsw_get_stats(mhmsw_ctx *ctx, mhmsw_stats *out);

/* No-allocation variant of mhmsw_submit + mhmsw_results, for callers that cannot
 * allocate device memory: code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc.  No counterpart in the reference.
 * Every pointer is DEVICE memory on the context's device: pair p aligns
 * reads[read_off[p] .. +read_len[p]) with ctgs[ctg_off[p] .. +ctg_len[p])
 * (ASCII, read as mhmsw_submit reads it), results[p] receives its result,
 * status[0] the batch's status and status[1] the first invalid pair (-1: none).
 * The work is enqueued on `stream` (a hipStream_t; NULL: the context's stream)
 * and the call returns without waiting; the outputs are complete when the
 * stream reaches that point.  It calls no allocator: the job's working memory
 * is a fixed array in device memory, one per GPU, held until the process ends.
 * All calls on one GPU share that array, so calls that overlap on one GPU race.
 * Limits: npairs <= MHMSW_NOALLOC_MAX_PAIRS (else MHMSW_E_ARG at once); the
 * sequences' lengths rounded up to 32 bases sum to at most
 * MHMSW_NOALLOC_MAX_BASES, and each pair passes mhmsw_stage's length checks,
 * else status[0] = MHMSW_E_ARG and every result is the score-0 sentinel.
 * Scoring, orientation, mode and tuning come from the context; its score2 and
 * CIGAR settings do not apply. */
#define MHMSW_NOALLOC_MAX_PAIRS 131072
#define MHMSW_NOALLOC_MAX_BASES 33554432
int mhmsw_submit_noalloc(mhmsw_ctx *ctx, int64_t npairs, const char *reads, const int64_t *read_off,
                         const int32_t *read_len, const char *ctgs, const int64_t *ctg_off, const int32_t *ctg_len,
                         mhmsw_result *results, int32_t *status, void *stream);

/* --- device-resident contigs (serving; no counterpart in the reference,
 *     whose windows are copied per batch as strings, klign.cpp:691-706) ---
 * Contigs are packed once to 2-bit codes plus an N bitmap and kept in HBM;
 * batches then name windows as (contig id, offset, length) and the window
 * codes are extracted on the device. */
typedef struct mhmsw_contigs mhmsw_contigs;
int mhmsw_contigs_load(int device_id, int64_t nctgs, const char *seqs, const int64_t *off, const int32_t *len,
                       mhmsw_contigs **out);
int mhmsw_contigs_free(mhmsw_contigs *contigs);
int64_t mhmsw_contigs_bytes(const mhmsw_contigs *contigs); /* device bytes held */

/* Stage a batch whose pair p aligns reads[read_off[p] .. +read_len[p]) with
 * contig ctg_id[p] at [win_off[p], win_off[p] + win_len[p]).  The contigs
 * must live on the context's device. */
int mhmsw_stage_windows(mhmsw_ctx *ctx, const mhmsw_contigs *contigs, int64_t npairs, const char *reads,
                        const int64_t *read_off, const int32_t *read_len, const int32_t *ctg_id, const int32_t *win_off,
                        const int32_t *win_len);

/* --- serving loop: continuous batching of alignment requests ---
 * A server owns `slots` contexts (one HIP stream each) on one device and a
 * dispatcher thread.  Requests are admitted in arrival order as soon as a
 * slot is free and their device-memory estimate fits what running jobs leave
 * of hbm_budget (bytes, 0 = 80% of the device less the resident contigs; a
 * request whose estimate alone exceeds it fails at submit with MHMSW_E_ARG);
 * requests queued together when a slot frees are
 * merged into one job (up to 2^20 pairs), and jobs on different slots run
 * concurrently on the device.  Results are identical to aligning the request
 * alone. */
typedef struct mhmsw_server mhmsw_server;
typedef struct {
  double submit_s;  /* host steady clock, seconds */
  double admit_s;
  double done_s;
  double device_ms; /* device time of the request's job */
  double launched_s; /* the job's copies and kernels were enqueued */
  double finished_s; /* the job's slot saw it finish (results not yet copied out) */
  double job_est_bytes; /* device memory the request's job may take: the admission estimate, counted against
                           hbm_budget while it runs */
} mhmsw_req_times;
int mhmsw_server_create(int device_id, const mhmsw_scoring *scoring, const mhmsw_contigs *contigs, int slots,
                        int64_t hbm_budget, mhmsw_server **out);
/* Copies the request's host data; returns at once with *req_id. */
int mhmsw_server_submit(mhmsw_server *srv, int64_t npairs, const char *reads, const int64_t *read_off,
                        const int32_t *read_len, const int32_t *ctg_id, const int32_t *win_off, const int32_t *win_len,
                        uint64_t *req_id);
/* 0 queued, 1 running, 2 done, < 0 failed. */
int mhmsw_server_status(mhmsw_server *srv, uint64_t req_id);
int mhmsw_server_wait(mhmsw_server *srv, uint64_t req_id);
/* Waits, copies npairs results and the request's times, and forgets the request. */
int mhmsw_server_results(mhmsw_server *srv, uint64_t req_id, mhmsw_result *out, int64_t npairs, mhmsw_req_times *times);
/* Second-best scores and CIGARs for the jobs admitted after the call (both
 * off by default), computed as mhmsw_set_score2 / mhmsw_set_cigar do. */
int mhmsw_server_set_outputs(mhmsw_server *srv, int score2, int cigar);
/* After mhmsw_server_results of a request that ran with them: its second-best
 * scores (score2, may be NULL) and CIGARs in the layout of
 * mhmsw_results_cigar (mismatches, op_off[0..npairs], ops; call with ops ==
 * NULL first to learn *total_ops).  The server keeps them until a call passes
 * ops (or, for a request without CIGARs, until the first call). */
int mhmsw_server_results_extra(mhmsw_server *srv, uint64_t req_id, int64_t npairs, int32_t *score2,
                               int32_t *mismatches, int64_t *op_off, uint32_t *ops, int64_t ops_cap,
                               int64_t *total_ops);
double mhmsw_server_clock(void); /* the clock of mhmsw_req_times */
const char *mhmsw_server_last_error(const mhmsw_server *srv);
int mhmsw_server_destroy(mhmsw_server *srv);

#ifdef __cplusplus
}
#endif
#endif
