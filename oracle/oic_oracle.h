/*
 * oic_oracle.h -- CPU restatement of canu's overlapInCore hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the tests' oracle: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The
 * product path (canu_amd/, liboic_hip.so) never links or calls it.
 *
 * Every function follows the reference statement by statement; each carries
 * the src/... file:line it restates.  Reads are 1-based IDs as in sqStore.
 */
#ifndef OIC_ORACLE_H
#define OIC_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OO_MAX_READLEN_BITS 21
#define OO_MAX_READLEN      ((1u << OO_MAX_READLEN_BITS) - 1u)   /* sqStore.H:57-58 */

/* ---- tables: prefixEditDistance.C:23-88, Binomial_Bound.C:36-195 ---- */
typedef struct oo_tables {
  double    maxErate;            /* (double)(float)  -- overlapInCore.C:380 parses with strtof */
  double    alignNoise;
  int       partial;
  uint32_t  MAX_ERRORS;          /* 1 + ceil(maxErate * AS_MAX_READLEN) */
  uint32_t  MIN_BRANCH_END_DIST; /* 20 (uint32 in the reference: comparisons are unsigned) */
  double    MIN_BRANCH_TAIL_SLOPE;
  double    Branch_Match_Value;
  int32_t  *Error_Bound;         /* [AS_MAX_READLEN + 1] */
  int32_t  *Edit_Match_Limit;    /* [MAX_ERRORS + 1], [MAX_ERRORS] unused (never written by the reference) */
} oo_tables;

int       oo_binomial_bound(int e, double p, int Start);
void      oo_init_match_limit(int32_t *ml, double maxErate, int32_t maxErrors);
oo_tables *oo_tables_new(double maxErate, double alignNoise, int partial);
void      oo_tables_free(oo_tables *tb);
/* float rounding of CLI parameters, as strtof() does at overlapInCore.C:380,382 */
double    oo_cli_rate(double v);

/* ---- DP workspace: prefixEditDistance state ---- */
typedef struct oo_dp oo_dp;

oo_dp    *oo_dp_new(const oo_tables *tb);
void      oo_dp_free(oo_dp *W);
void      oo_dp_set_poison(oo_dp *W, int on);    /* debug: poison rows, flag reads of unwritten cells */
int       oo_dp_poison_hits(const oo_dp *W);

/* prefixEditDistance::forward (prefixEditDistance-forward.C:94-314) */
int32_t   oo_forward(oo_dp *W, const char *A, int32_t m, const char *T, int32_t n,
                     int32_t Error_Limit, int32_t *A_End, int32_t *T_End, int32_t *Match_To_End);
/* prefixEditDistance::reverse (prefixEditDistance-reverse.C:114-330).  A and T point at the
 * rightmost base of the region; the DP walks leftwards with negative indexing. */
int32_t   oo_reverse(oo_dp *W, const char *A, int32_t m, const char *T, int32_t n,
                     int32_t Error_Limit, int32_t *A_End, int32_t *T_End, int32_t *Leftover,
                     int32_t *Match_To_End);

/* Match_Node_t (prefixEditDistance.H:57-62) */
typedef struct oo_match { int32_t Offset, Len, Start, Next; } oo_match;

/* prefixEditDistance::Extend_Alignment (prefixEditDistance-extend.C:36-183).
 * Returns Overlap_t: 0 NONE, 1 LEFT_BRANCH_PT, 2 RIGHT_BRANCH_PT, 3 DOVETAIL. */
int32_t   oo_extend(oo_dp *W, const oo_match *Match,
                    const char *S, uint32_t S_ID, int32_t S_Len,
                    const char *T, uint32_t T_ID, int32_t T_Len,
                    int32_t *S_Lo, int32_t *S_Hi, int32_t *T_Lo, int32_t *T_Hi, int32_t *Errors);

/* delta / counters accessors */
int32_t   oo_dp_left_delta(const oo_dp *W, int32_t *out, int32_t cap);
int32_t   oo_dp_right_delta(const oo_dp *W, int32_t *out, int32_t cap);
void      oo_dp_counters(const oo_dp *W, uint64_t out[4]);   /* cells, bases, calls, rows */
void      oo_dp_reset_counters(oo_dp *W);

/* ---- whole overlapInCore run ---- */
typedef struct oo_params {
  uint32_t  kmer;              /* -k */
  double    maxErate;          /* --maxerate (pass the CLI decimal; it is float-rounded inside) */
  double    alignNoise;        /* --alignnoise */
  int32_t   minOlap;           /* --minlength */
  int32_t   partial;           /* -partial */
  int32_t   unique;            /* -u (1, default) / -m (0) */
  int32_t   hopeless;          /* 1 unless -z; forced 0 when maxErate > 0.06 */
  int32_t   minkmers;          /* --minkmers */
  uint64_t  fragOlapLimit;     /* -l (UINT64_MAX default) */
  uint32_t  hashBits;          /* --hashbits */
  double    hashLoad;          /* --hashload */
  uint64_t  hashDataLen;       /* --hashdatalen (before the +AS_MAX_READLEN at overlapInCore.C:454) */
  uint32_t  bgnHash, endHash;  /* -h */
  uint32_t  bgnRef, endRef;    /* -r */
  uint32_t  nThreads;          /* -t: also defines the chunking quirk of OverlapDriver */
  uint32_t  ompThreads;        /* actual OpenMP threads used (0 = nThreads) */
  const char *skipKmers;       /* -k <file>: nSkip k-mers of kmer ACGT letters each, back to back, in
                                  file order (parsed as Mark_Skip_Kmers does); NULL = no file */
  uint64_t  nSkip;
  int32_t   restrictIndex;     /* oracle shortcut, no reference counterpart: index only the k-mers the
                                  searched ref reads and the skip k-mers hold (same search; the hash range
                                  must fit one block, else oo_run returns -2) */
  const uint32_t *refList;     /* oracle shortcut: when nRefList > 0, search exactly these ref IDs (in
                                  place of bgnRef..endRef and the -t chunk rule) */
  uint64_t  nRefList;
} oo_params;

void oo_params_default(oo_params *P);

/* One overlap record as ovFile writes it (ovStoreFile.C:253-283): IDs plus the two 64-bit
 * ovOverlapDAT words (ovOverlap.H:49-78, LSB-first bitfields). */
typedef struct oo_record { uint32_t a_iid, b_iid; uint64_t w0, w1; } oo_record;

/* per candidate pair (S, dir, T) trace: features and the exact work it took */
typedef struct oo_pair_trace {
  uint32_t s_id, t_id;
  int32_t  dir, s_len, t_len;
  int32_t  n_hits, n_nodes, diag_min, diag_max, consistent, longest;
  int32_t  n_ext, n_olaps;
  uint64_t cells;
  int32_t  max_mult;           /* max k-mer multiplicity over the pair's hits */
  int32_t  exit_d;             /* diagonal on which the first extension's forward DP left */
} oo_pair_trace;

typedef struct oo_result {
  oo_record     *rec;      uint64_t nrec;
  uint64_t       stats[8]; /* overlapInCore.C:550-558 order */
  uint64_t       cells, bases, dp_calls, extensions, hits, pairs;
  uint64_t       cells_used;     /* cells minus reverse() runs after a forward() that did not reach the end */
  oo_pair_trace *trace;    uint64_t ntrace;   /* only when collect_trace */
  uint32_t       hashEnd[64]; uint32_t nblocks;   /* hash block boundaries actually used */
  double         t_index, t_search;
} oo_result;

/* reads: bases[start[i] .. start[i]+len[i]) for IDs 1..nreads (index 0 unused), any case, ACGTN */
int  oo_run(const oo_params *P, const char *bases, const uint64_t *start, const uint32_t *len,
            uint32_t nreads, int collect_trace, oo_result *out);
void oo_result_free(oo_result *r);

/* ovOverlap helpers */
uint32_t oo_encode_evalue(double q);    /* ovOverlap.H:31-35 */

/* Test hook: Add_Match (overlapInCore-Find_Overlaps.C:26-96) over one pair's seed hits in arrival
 * order (offset[i] in S, toff[i] in T), from an empty list.  Writes the final list front first as
 * (Offset, Len, Start) triples and the consistent flag; returns the node count, or -1 when more
 * than cap nodes. */
int32_t oo_match_list(uint32_t kmer, const int32_t *offset, const int32_t *toff, int32_t n, int32_t *out,
                      int32_t cap, int32_t *consistent);
/* Add_Overlap over a sequence of extension results, then Choose_Best_Partial /
 * Combine_Into_One_Olap / Merge_Intersecting_Olaps (Process_String_Overlaps.C:175-311, 501-513);
 * rows of 11 ints per distinct overlap (box, min/max diag, boundaries, deleted). */
int32_t oo_distinct_olaps(int partial, int unique, int32_t n, const int32_t *box, const double *qual,
                          int32_t *out, double *outq);
/* Lies_On_Alignment (Process_String_Overlaps.C:260-281) against an explicit Left_Delta. */
int32_t oo_lies_on_alignment(int32_t start, int32_t offset, int32_t s_lo, int32_t t_lo, const int32_t *delta, int32_t n);

#ifdef __cplusplus
}
#endif
#endif
