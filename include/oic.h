/*
 * oic.h -- C ABI of the MI355X overlapInCore engine (liboic_hip.so).
 *
 * Plain C types only.  One host thread per context; a context owns its HIP stream and all device
 * memory; calls are not re-entrant.  Every entry point returns 0 on success and a negative
 * OIC_E* code on failure (never exits); oic_last_error() describes the last failure.
 *
 * Read IDs are 1-based, as in canu's sqStore.  Reads are passed as one byte array
 * (ACGTN, any case) with per-read start offsets and lengths; index 0 of start/len is unused.
 *
 * Reference interfaces replaced (paths relative to canu's src/):
 *   oic_create / oic_destroy   oicParameters + OverlapDriver's setup and teardown
 *                              (overlapInCore/overlapInCore.H:367-473, overlapInCore.C:161-277, :283-527)
 *   oic_load_reads             sqStore / sqCache reads of the hash and ref ranges
 *                              (overlapInCore-Build_Hash_Index.C:504-545, overlapInCore-Process_Overlaps.C:48-67,
 *                               stores/sqCache.C:262-318)
 *   oic_index_build            Build_Hash_Index (overlapInCore-Build_Hash_Index.C:414-631, overlapInCore.H:513-514)
 *   oic_search                 Process_Overlaps -> Find_Overlaps -> Process_String_Olaps -> Process_Matches
 *                              -> Output_Overlap over a ref range (overlapInCore-Process_Overlaps.C:24-123,
 *                              overlapInCore-Find_Overlaps.C:234-336, overlapInCore-Process_String_Overlaps.C:353-690,
 *                              overlapInCore-Output.C:26-264)
 *   oic_run                    OverlapDriver's hash-block loop (overlapInCore.C:204-263)
 *   oic_extend_batch           prefixEditDistance::Extend_Alignment (liboverlap/prefixEditDistance.H:93-98,
 *                              prefixEditDistance-extend.C:36-183)
 *   oic_prefix_dp_batch        prefixEditDistance::forward / ::reverse (prefixEditDistance.H:75-90,
 *                              prefixEditDistance-forward.C:94-314, prefixEditDistance-reverse.C:114-330)
 *   oic_tables                 prefixEditDistance constructor tables (prefixEditDistance.C:23-88,
 *                              Binomial_Bound.C:104-195)
 */
#ifndef OIC_H
#define OIC_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OIC_OK            0
#define OIC_EINVAL       -1   /* bad argument */
#define OIC_EHIP         -2   /* HIP runtime error */
#define OIC_ENOMEM       -3   /* device allocation failed */
#define OIC_ESTATE       -4   /* call out of order (e.g. search before index) */
#define OIC_ECAPACITY    -5   /* caller buffer too small; *_needed reports the size */
#define OIC_EUNSUPPORTED -6   /* option or input size the engine does not implement (e.g. a DP band
                                 wider than its widest tile) */

typedef struct oic_ctx oic_ctx;

/* overlapInCore command-line surface (overlapInCore.C:293-411).  Rates are given as the decimal
 * the user typed; the engine rounds them to float as strtof() does in the reference. */
typedef struct oic_params {
  uint32_t kmer;             /* -k */
  double   max_erate;        /* --maxerate */
  double   align_noise;      /* --alignnoise (1.0 default) */
  int32_t  min_olap;         /* --minlength */
  int32_t  partial;          /* -partial */
  int32_t  unique;           /* -u (1, default) / -m (0) */
  int32_t  hopeless;         /* 1 unless -z (forced 0 when max_erate > 0.06) */
  int32_t  minkmers;         /* --minkmers */
  uint64_t frag_olap_limit;  /* -l (UINT64_MAX = unlimited; the only supported value) */
  uint32_t hash_bits;        /* --hashbits (22) */
  double   hash_load;        /* --hashload (0.6) */
  uint64_t hash_data_len;    /* --hashdatalen (100000000) */
  uint32_t n_threads;        /* -t: only its effect on which ref reads are searched is reproduced */
  int32_t  device;           /* HIP device ordinal */
  int32_t  schedule;         /* OIC_SCHED_* */
} oic_params;

#define OIC_SCHED_PLAIN    0  /* one candidate pair per wavefront, pairs in discovery order */
#define OIC_SCHED_BATCHED  1  /* binned, load-balanced persistent launches (static length bins) */
#define OIC_SCHED_LEARNED  2  /* as BATCHED, pairs binned and ordered by the learned cost model */

void oic_params_default(oic_params *p);

typedef struct oic_reads {
  const char     *bases;   /* all reads, back to back */
  const uint64_t *start;   /* [nreads + 1], start[i] = offset of read i */
  const uint32_t *len;     /* [nreads + 1] */
  uint32_t        nreads;
} oic_reads;

/* ovOverlap as ovFile stores it (ovStoreFile.C:253-283, ovOverlap.H:49-78): the two 64-bit
 * ovOverlapDAT words hold ahg5:21 ahg3:21 evalue:16 flipped forOBT forDUP forUTG spare:2 and
 * bhg5:21 bhg3:21 span:21 spare:1, least significant bit first. */
typedef struct oic_overlap {
  uint32_t a_iid, b_iid;
  uint64_t dat[2];
} oic_overlap;

/* Caller-owned output sink.  records[count ..] receive the call's records while they fit.  If
 * they do not all fit, the search still runs to its end: stats and counters are complete, count
 * is the number of records written (at most capacity), records_needed the total, the context
 * holds the rest and the call returns OIC_ECAPACITY; oic_fetch_pending then appends them to a
 * sink with room, without searching again.  records may be NULL (capacity 0) to learn the size. */
typedef struct oic_sink {
  oic_overlap *records;
  uint64_t     capacity;
  uint64_t     count;           /* out */
  uint64_t     records_needed;  /* out */
  /* .stats counters, overlapInCore.C:550-558 order: kmer hits without olaps, with olaps,
   * multiple overlaps/pair, total overlaps, contained, dovetail, short window, long window */
  uint64_t     stats[8];        /* out, accumulated */
} oic_sink;

/* work counters of the last oic_search / oic_run (accumulated over blocks) */
typedef struct oic_counters {
  uint64_t kmer_probes;        /* S windows looked up */
  uint64_t hits;               /* seed hits with S_ID < T_ID */
  uint64_t pairs;              /* candidate (S, orientation, T) pairs */
  uint64_t nodes;              /* match nodes */
  uint64_t extensions;         /* Extend_Alignment calls */
  uint64_t dp_calls;           /* forward / reverse DP runs executed */
  uint64_t dp_cells;           /* (e, d) cell updates executed, row-0 slide counted as one */
  uint64_t dp_lane_slots;      /* lane-cells issued (cells plus idle lanes of partial chunks) */
  uint64_t dp_bases;           /* bases compared in slides, plus one terminating compare per cell */
  uint64_t overflow_retries;   /* DP tasks re-run with a wider band tile */
  uint64_t spec_tasks;         /* forward DPs run ahead of Process_Matches (batched schedules) */
  uint64_t spec_cells;         /* cells of those runs, used or not */
  uint64_t spec_rows;          /* DP rows (e >= 1) of those runs (with tuning.fwd_counters) */
  uint64_t spec_lane_slots;    /* lane-cells issued by those runs (with tuning.fwd_counters) */
  uint64_t spec_wave_rows;     /* row iterations wavefronts executed for them (a wavefront waits for
                                  its slowest packed DP) (with tuning.fwd_counters) */
  uint64_t spec_bases;         /* bases compared by those runs */
  uint64_t spec_slide_iters;   /* 16-base window compares of their slides, per lane (with tuning.fwd_counters) */
  uint64_t spec_wave_slide_iters; /* slide iterations wavefronts executed (with tuning.fwd_counters) */
  uint64_t fwd_launches;       /* launches of the packed narrow-band forward kernel */
  uint64_t fwd_cells;          /* cells and bases of those launches */
  uint64_t fwd_bases;
  uint64_t graphs_launched;    /* forward-pass hipGraphs launched */
  uint64_t list_wave_pairs;    /* pairs whose match list one wavefront built */
  uint64_t list_handed_back;   /* of those, pairs with an offset of more hits than the wavefront holds */
  uint64_t list_second_level;  /* of those, pairs taken on by the 512-hit wavefront level (an offset
                                  with more than 128 hits) */
  double   ms_index, ms_gather, ms_lists, ms_extend, ms_total;
  double   ms_dp_kernel;       /* device time of all DP kernels (HIP events) */
  double   ms_fwd_kernel;      /* of which the packed forward-DP kernels */
  double   ms_fwd_narrow;      /* of which the packed narrow-band kernel: HIP events around each launch,
                                  or, inside a forward-pass graph, the span from its first wavefront's
                                  start to its last one's end (device wall clock) */
  double   ms_fwd_graph;       /* forward-pass graphs, HIP events around their launches */
  double   ms_pp_kernel;       /* the Process_Matches kernel (k_process_pairs), HIP events around its launches */
  uint64_t pp_cells;           /* DP cells that kernel computes itself (full extensions with row storage) */
  uint64_t pp_bases;           /* bases it compares in their slides, plus one per cell */
  uint64_t learned_pairs;      /* pairs the learned cost model ordered (OIC_SCHED_LEARNED with a model) */
  uint64_t pp_launches;        /* launches of the Process_Matches kernel */
  uint64_t spec_wave_chunks;   /* band chunk iterations wavefronts of the row-stepped kernel executed (with tuning.fwd_counters) */
  uint64_t allocs;             /* device work-buffer allocations the context made (first use or growth) */
  double   ms_alloc;           /* host time those allocations took, with the free of the smaller buffer */
} oic_counters;

/* Launch-shape knobs.  No setting changes a result; the defaults suit MI355X.  No reference
 * counterpart: overlapInCore has no device work to shape. */
typedef struct oic_tuning {
  uint32_t list_wave_min_hits;  /* pairs with at least this many seed hits build their match list
                                   with one wavefront, smaller ones with one thread (default 2048) */
  uint32_t fwd_lanes;           /* lanes per packed forward DP: 1 (64 DPs per wavefront, each lane
                                   walking its own band cell by cell), or 4, 8, 16 or 32 (64/fwd_lanes
                                   DPs per wavefront, rows spread over the group; default 4) */
  uint32_t fwd_tile;            /* band cells a packed forward DP holds in LDS: 16, 32 or 64 with one
                                   lane per DP; 32 or 64 (8 or 16 lanes), 128, 256 or 512 otherwise;
                                   wider bands re-run one per wavefront, with a 512-cell tile and then
                                   an 8192-cell one (default 64) */
  uint32_t fwd_order;          /* speculative DP tasks in pair order (0) or by seed length,
                                   longest first (1); the autotuner picks */
  uint32_t fwd_graph;           /* 1 (default): each search batch's probe and speculation passes run
                                   as one captured hipGraph; 0: the same launches, issued directly */
  uint32_t fwd_counters;        /* 1: count rows, lane slots and wavefront rows of the packed
                                   forward DPs (spec_rows, spec_lane_slots, spec_wave_rows); costs
                                   time, off by default */
  uint32_t fwd_loop;            /* packed forward-DP loop: 0 advances every DP of a wavefront one
                                   row per iteration and starts a finished DP's next task at once;
                                   1 runs each DP to its end, so finished DPs wait for the slowest
                                   of their wavefront; 2 (default; flat rows, with 4 or 8 lanes per
                                   DP and fwd_tile 32, 64 or 128) is 0 with each row's cells of all
                                   the wavefront's DPs spread over its 64 lanes */
} oic_tuning;

int oic_get_tuning(const oic_ctx *ctx, oic_tuning *out);
int oic_set_tuning(oic_ctx *ctx, const oic_tuning *t);

typedef struct oic_tune_trial {
  oic_tuning tuning;
  int32_t    stage;             /* what was timed: 0 the packed forward-DP kernel, 1 the match-list
                                   stage, 2 the extension stage (forward passes through output) */
  double     ms;                /* time of that stage over the sample */
} oic_tune_trial;

/* Autotuner: searches ref reads [bgn_ref, end_ref] against the current index under each candidate
 * launch shape (forward DP: one DP per lane with a 16..64-cell ring, or 4/8/16/32 lanes per DP x
 * 32..512-cell tile; then loop form, then task order; list builder threshold 256..16384 hits).  Each
 * knob starts from the current setting, which a candidate replaces only when it is at least 3 %
 * faster.  Reports the trials (up to cap) when trials is not null.  Counters are left as they were.
 * No reference counterpart. */
int oic_autotune(oic_ctx *ctx, uint32_t bgn_ref, uint32_t end_ref, oic_tuning *best, oic_tune_trial *trials,
                 uint32_t cap, uint32_t *ntrials);

int         oic_create(const oic_params *p, oic_ctx **out);
void        oic_destroy(oic_ctx *ctx);
const char *oic_last_error(const oic_ctx *ctx);

/* -k <file>: the text of a skip-k-mer file, read as Mark_Skip_Kmers does
 * (overlapInCore-Build_Hash_Index.C:185-257).  Each k-mer and its reverse complement yield no seed
 * hits, and mark read ends within HOPELESS_MATCH of them as screened for the hopeless check.
 * Applies from the next index build.  Malformed line: OIC_EINVAL with the reference's message. */
int oic_set_skip_kmers(oic_ctx *ctx, const char *text, uint64_t len);

/* Upload and 2-bit pack reads 1..nreads (both strands) into device memory. */
int oic_load_reads(oic_ctx *ctx, const oic_reads *reads);

/* Build_Hash_Index over hash reads [bgn_id, end_id]: stops at the reference's data-length and
 * hash-load limits; *last_loaded receives the last read ID the block covers. */
int oic_index_build(oic_ctx *ctx, uint32_t bgn_id, uint32_t end_id, uint32_t *last_loaded);

/* Search ref reads [bgn_ref, end_ref] (both orientations) against the current index and append
 * the overlaps to sink.  Applies the reference's -t chunking rule for which ref reads are run. */
int oic_search(oic_ctx *ctx, uint32_t bgn_ref, uint32_t end_ref, oic_sink *sink);

/* The whole overlapInCore job: hash-block loop over [bgn_hash, end_hash] x [bgn_ref, end_ref]. */
int oic_run(oic_ctx *ctx, uint32_t bgn_hash, uint32_t end_hash, uint32_t bgn_ref, uint32_t end_ref,
            oic_sink *sink);

/* After OIC_ECAPACITY from oic_search / oic_run: append the records the context holds to
 * records[count ..] once capacity - count >= records_needed - count; OIC_ECAPACITY (with
 * records_needed) while it is smaller.  The next oic_search / oic_run drops held records.  No
 * reference counterpart: overlapInCore writes each thread's records to its ovFile as it goes
 * (overlapInCore-Process_Overlaps.C:96-112). */
int oic_fetch_pending(oic_ctx *ctx, oic_sink *sink);

int oic_get_counters(const oic_ctx *ctx, oic_counters *out);

/* ---- DP units ---- */

/* One Extend_Alignment call: seed Match_Node (Start on S, Offset on T, Len) between read s_id in
 * orientation s_dir (0 forward, 1 reverse complement) and read t_id (forward), both loaded. */
typedef struct oic_ext_task {
  uint32_t s_id, t_id;
  int32_t  s_dir;
  int32_t  start, offset, len;
} oic_ext_task;

typedef struct oic_ext_result {
  int32_t kind;                 /* Overlap_t: 0 NONE, 1 LEFT_BRANCH_PT, 2 RIGHT_BRANCH_PT, 3 DOVETAIL */
  int32_t s_lo, s_hi, t_lo, t_hi, errors;
  int32_t delta_len;            /* merged Left_Delta length */
  int32_t status;               /* 0 ok, else OIC_E* */
} oic_ext_result;

/* delta (optional): delta_stride ints per task receive the merged Left_Delta. */
int oic_extend_batch(oic_ctx *ctx, const oic_ext_task *tasks, uint32_t n, oic_ext_result *results,
                     int32_t *delta, uint32_t delta_stride);

/* One extension result offered to Add_Overlap: the aligned boxes on S and T and the quality
 * (errors / length) Process_Matches computes for it. */
typedef struct oic_olap_in {
  int32_t s_lo, s_hi, t_lo, t_hi;
  double  quality;
} oic_olap_in;

/* One distinct overlap of a pair (Olap_Info_t, overlapInCore.H:207-216, without the delta). */
typedef struct oic_olap_out {
  int32_t s_lo, s_hi, t_lo, t_hi;
  int32_t min_diag, max_diag;
  int32_t s_left, s_right, t_left, t_right;   /* s/t_left_boundary, s/t_right_boundary */
  int32_t deleted;                            /* removed by the selection: not output */
  int32_t pad;
  double  quality;
} oic_olap_out;

/* Add_Overlap over each pair's extension results in order, then the pair's selection
 * (Process_String_Overlaps.C:175-244, 42-162, 289-311, chosen as at :501-513) under the context's
 * partial and unique parameters: the step of Process_Matches that turns extension results into
 * output overlaps.  Pair g owns in[first[g] .. first[g+1]) (first has n_pairs + 1 entries, first[0]
 * is 0, non-decreasing); its distinct overlaps, at most 3, go to out[3g ..] and their number to
 * count[g]. */
int oic_distinct_olaps_batch(oic_ctx *ctx, const uint32_t *first, uint32_t n_pairs, const oic_olap_in *in,
                             oic_olap_out *out, int32_t *count);

/* One prefixEditDistance::forward (dir 0) or ::reverse (dir 1) call on explicit windows.
 * Forward compares a[a_pos ..] with b[b_pos ..] rightwards; reverse compares a[a_pos], a[a_pos-1], ...
 * with b[b_pos], b[b_pos-1], ... as the reference's negative indexing does.  Windows are
 * (read id, strand 0/1, position, length) of loaded reads; m <= n is required. */
typedef struct oic_dp_task {
  uint32_t a_id, b_id;
  int32_t  a_strand, b_strand;
  int32_t  a_pos, b_pos;
  int32_t  m, n;
  int32_t  error_limit;
  int32_t  dir;
} oic_dp_task;

typedef struct oic_dp_result {
  int32_t errors, a_end, t_end, match_to_end, leftover, delta_len;
  int32_t status;
  int32_t pad;
  uint64_t cells;
} oic_dp_result;

int oic_prefix_dp_batch(oic_ctx *ctx, const oic_dp_task *tasks, uint32_t n, oic_dp_result *results,
                        int32_t *delta, uint32_t delta_stride);

/* oic_prefix_dp_batch without any device allocation.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Same tasks, results and deltas as oic_prefix_dp_batch, on one
 * device array the library holds on each GPU it runs on (about 41 MB, from library load to process
 * end).  Limits per call: at most 16384 tasks, tasks x delta_stride at most 2^20, error limits at
 * most 1021 (OIC_EUNSUPPORTED beyond, nothing run); a DP whose traceback needs more than 131072
 * stored band cells gets status OIC_ECAPACITY, as one beyond oic_prefix_dp_batch's larger store
 * does.  delta entries past each result's delta_len are zero.  All calls on one GPU share its array:
 * calls that overlap on one GPU race. */
int oic_prefix_dp_batch_noalloc(oic_ctx *ctx, const oic_dp_task *tasks, uint32_t n, oic_dp_result *results,
                                int32_t *delta, uint32_t delta_stride);

/* Host copies of the DP tables the context uses. error_bound has 2^21 entries,
 * match_limit has max_errors entries. */
int oic_tables(const oic_ctx *ctx, const int32_t **error_bound, const int32_t **match_limit,
               uint32_t *max_errors);

/* ---- learned cost model (no counterpart in the reference: it schedules work, never changes a result) ----
 *
 * A 12 -> 64 -> 64 -> 2 ReLU MLP with FP32 weights predicts, from a candidate pair's seed
 * features, log(1 + DP cells) and the exit diagonal of its first extension.  OIC_SCHED_LEARNED
 * orders pairs and packed DP tasks by the prediction; OIC_SCHED_BATCHED uses a length-based
 * estimate instead.  Training runs on the device with the library's own forward, backward and
 * Adam kernels. */
#define OIC_NFEAT   12
#define OIC_NOUT    2
#define OIC_NPARAM  (64 * OIC_NFEAT + 64 + 64 * 64 + 64 + OIC_NOUT * 64 + OIC_NOUT)

/* keep per-pair traces (features, targets) of the following searches */
int oic_set_tracing(oic_ctx *ctx, int on);
/* features: count x 12, targets: count x 2 (log1p cells, exit diagonal / 1000) */
int oic_get_traces(oic_ctx *ctx, float *features, float *targets, uint64_t capacity, uint64_t *count);

typedef struct oic_train_params {
  uint32_t epochs, batch;
  float    lr, beta1, beta2, eps;
  float    w_exit;              /* loss weight of the exit-diagonal output */
  uint64_t seed;
} oic_train_params;

/* Train the context's model on (features, targets); returns the final mean loss. */
int oic_model_train(oic_ctx *ctx, const float *features, const float *targets, uint64_t n,
                    const oic_train_params *tp, float *final_loss);
/* weights: OIC_NPARAM floats (layout W1[64][12] b1 W2[64][64] b2 W3[2][64] b3) plus the
 * input standardisation mu[12], sd[12] */
int oic_model_get(const oic_ctx *ctx, float *weights, float *mu, float *sd);
int oic_model_set(oic_ctx *ctx, const float *weights, const float *mu, const float *sd);
/* device inference: out = n x 2 */
int oic_model_predict(oic_ctx *ctx, const float *features, uint64_t n, float *out);
/* loss and gradient of the (standardised-input) MLP on one batch, for checking the kernels */
int oic_model_grad(oic_ctx *ctx, const float *weights, const float *features, const float *targets, uint64_t n,
                   float w_exit, float *loss, float *grad);

/* The same tables without a context or device: error_bound needs 2^21 entries, match_limit
 * *max_errors entries (call once with match_limit = NULL to learn the size). */
int oic_dp_tables(double max_erate, double align_noise, int32_t *error_bound, int32_t *match_limit,
                  uint32_t *max_errors);

#ifdef __cplusplus
}
#endif
#endif
