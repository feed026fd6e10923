/*
 * readsim.c -- deterministic synthetic genome and read simulator for the overlapInCore workloads.
 *
 * Genome: uniform random ACGT, optionally with interspersed repeat families (copies diverged by
 * substitutions and indels) and tandem satellite arrays of a short unit.  Reads: lognormal lengths,
 * uniform start, strand Bernoulli(1/2), i.i.d. substitution / insertion / deletion errors, or
 * homopolymer-length indels; optional homopolymer compression of the reads.
 *
 * Output is 1-based like a seqStore: read i occupies bases[start[i] .. start[i]+len[i]) followed by a
 * NUL; index 0 of start/len/pos/strand is unused.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

typedef struct rs_config {
  uint64_t genome_len;
  uint64_t seed;
  double   coverage;
  double   ln_mu, ln_sigma;
  uint32_t min_len, max_len;
  double   p_sub, p_ins, p_del;
  int32_t  hp_errors;          /* 1: indels change homopolymer run lengths */
  /* interspersed repeats */
  double   repeat_frac;        /* fraction of the genome made of repeats (families + satellites) */
  uint32_t n_families, fam_min, fam_max;
  double   div_min, div_max;
  /* tandem satellites */
  uint32_t n_sat, sat_unit, sat_min, sat_max;
  double   sat_div;
  int32_t  hpc;                /* homopolymer-compress reads */
} rs_config;

typedef struct rs_reads {
  char     *bases;
  uint64_t *start;
  uint32_t *len;
  uint64_t *pos;               /* genome start of the template */
  uint32_t *tlen;              /* template length on the genome */
  uint8_t  *strand;            /* 0 forward, 1 reverse complement */
  uint32_t  nreads;
  uint64_t  total;             /* total read bases */
  char     *genome;
  uint64_t  genome_len;
} rs_reads;

/* xoshiro256** seeded by splitmix64 */
typedef struct { uint64_t s[4]; } rng_t;
static uint64_t splitmix64(uint64_t *x) {
  uint64_t z = (*x += 0x9e3779b97f4a7c15ULL);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}
static void rng_seed(rng_t *r, uint64_t seed) { uint64_t x = seed; for (int i = 0; i < 4; i++) r->s[i] = splitmix64(&x); }
static inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
static inline uint64_t rng_next(rng_t *r) {
  uint64_t *s = r->s;
  uint64_t res = rotl(s[1] * 5, 7) * 9, t = s[1] << 17;
  s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3]; s[2] ^= t; s[3] = rotl(s[3], 45);
  return res;
}
static inline double rng_unif(rng_t *r) { return (rng_next(r) >> 11) * (1.0 / 9007199254740992.0); }
static inline uint64_t rng_below(rng_t *r, uint64_t n) { return n ? rng_next(r) % n : 0; }
static double rng_normal(rng_t *r) {
  double u1 = rng_unif(r), u2 = rng_unif(r);
  if (u1 < 1e-300) u1 = 1e-300;
  return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

static const char ACGT[4] = {'A', 'C', 'G', 'T'};
static inline char comp(char c) {
  switch (c) { case 'A': return 'T'; case 'C': return 'G'; case 'G': return 'C'; case 'T': return 'A'; default: return c; }
}
static inline char other_base(rng_t *r, char c) {
  char b;
  do { b = ACGT[rng_next(r) & 3]; } while (b == c);
  return b;
}

typedef struct { char *s; uint64_t n, cap; } sbuf;
static void sb_push(sbuf *b, char c) {
  if (b->n == b->cap) { b->cap = b->cap ? 2 * b->cap : 1 << 16; b->s = (char *)realloc(b->s, b->cap); }
  b->s[b->n++] = c;
}

/* copy src with i.i.d. substitution / insertion / deletion at total rate div (equal thirds) */
static void mutate_copy(rng_t *r, const char *src, uint64_t n, double div, sbuf *out) {
  for (uint64_t i = 0; i < n; i++) {
    double u = rng_unif(r);
    if (u < div / 3)            sb_push(out, other_base(r, src[i]));
    else if (u < 2 * div / 3)   continue;
    else if (u < div)         { sb_push(out, ACGT[rng_next(r) & 3]); sb_push(out, src[i]); }
    else                        sb_push(out, src[i]);
  }
}

typedef struct { uint64_t len; int kind; uint32_t id; } element;   /* kind 0 family copy, 1 satellite */

static int cmp_u64(const void *a, const void *b) {
  uint64_t x = *(const uint64_t *)a, y = *(const uint64_t *)b;
  return x < y ? -1 : x > y;
}

static char *make_genome(const rs_config *C, rng_t *r, uint64_t *glen_out) {
  uint64_t G = C->genome_len;
  if (C->repeat_frac <= 0 || (C->n_families == 0 && C->n_sat == 0)) {
    char *g = (char *)malloc(G + 1);
    for (uint64_t i = 0; i < G; i++) g[i] = ACGT[rng_next(r) & 3];
    g[G] = 0;
    *glen_out = G;
    return g;
  }
  /* repeat content */
  sbuf rep = {0};
  uint64_t target = (uint64_t)(C->repeat_frac * G);
  uint64_t nel = 0, capel = 1024;
  element *els = (element *)malloc(sizeof(element) * capel);
  uint64_t *el_off = (uint64_t *)malloc(sizeof(uint64_t) * capel);

  /* satellites: one random unit per array, each unit copy diverged by substitutions */
  for (uint32_t a = 0; a < C->n_sat; a++) {
    char unit[4096];
    uint32_t U = C->sat_unit < 4096 ? C->sat_unit : 4096;
    for (uint32_t i = 0; i < U; i++) unit[i] = ACGT[rng_next(r) & 3];
    uint64_t alen = C->sat_min + rng_below(r, (uint64_t)(C->sat_max - C->sat_min + 1));
    uint64_t b0 = rep.n;
    while (rep.n - b0 < alen)
      for (uint32_t i = 0; i < U && rep.n - b0 < alen; i++)
        sb_push(&rep, rng_unif(r) < C->sat_div ? other_base(r, unit[i]) : unit[i]);
    if (nel == capel) { capel *= 2; els = realloc(els, sizeof(element) * capel); el_off = realloc(el_off, sizeof(uint64_t) * capel); }
    els[nel].len = rep.n - b0; els[nel].kind = 1; els[nel].id = a; el_off[nel] = b0; nel++;
  }
  /* families: consensus per family, copies round-robin until the repeat target is met */
  char **cons = (char **)malloc(sizeof(char *) * (C->n_families ? C->n_families : 1));
  uint64_t *clen = (uint64_t *)malloc(sizeof(uint64_t) * (C->n_families ? C->n_families : 1));
  for (uint32_t f = 0; f < C->n_families; f++) {
    clen[f] = C->fam_min + rng_below(r, (uint64_t)(C->fam_max - C->fam_min + 1));
    cons[f] = (char *)malloc(clen[f]);
    for (uint64_t i = 0; i < clen[f]; i++) cons[f][i] = ACGT[rng_next(r) & 3];
  }
  for (uint32_t f = 0; C->n_families && rep.n < target; f = (f + 1) % C->n_families) {
    double div = C->div_min + (C->div_max - C->div_min) * rng_unif(r);
    uint64_t b0 = rep.n;
    mutate_copy(r, cons[f], clen[f], div, &rep);
    if (nel == capel) { capel *= 2; els = realloc(els, sizeof(element) * capel); el_off = realloc(el_off, sizeof(uint64_t) * capel); }
    els[nel].len = rep.n - b0; els[nel].kind = 0; els[nel].id = f; el_off[nel] = b0; nel++;
  }
  for (uint32_t f = 0; f < C->n_families; f++) free(cons[f]);
  free(cons); free(clen);

  /* shuffle element order (Fisher-Yates on an index array) */
  uint64_t *ord = (uint64_t *)malloc(sizeof(uint64_t) * (nel ? nel : 1));
  for (uint64_t i = 0; i < nel; i++) ord[i] = i;
  for (uint64_t i = nel; i > 1; i--) { uint64_t j = rng_below(r, i); uint64_t t = ord[i - 1]; ord[i - 1] = ord[j]; ord[j] = t; }

  uint64_t unique = G > rep.n ? G - rep.n : 0;
  /* nel+1 random gaps summing to unique: sorted uniform cut points */
  uint64_t *cut = (uint64_t *)malloc(sizeof(uint64_t) * (nel + 2));
  for (uint64_t i = 0; i < nel; i++) cut[i] = rng_below(r, unique + 1);
  qsort(cut, nel, sizeof(uint64_t), cmp_u64);
  uint64_t glen = unique + rep.n;
  char *g = (char *)malloc(glen + 1);
  uint64_t gp = 0, prev = 0;
  for (uint64_t e = 0; e <= nel; e++) {
    uint64_t c = (e < nel) ? cut[e] : unique;
    for (uint64_t i = prev; i < c; i++) g[gp++] = ACGT[rng_next(r) & 3];
    prev = c;
    if (e < nel) { memcpy(g + gp, rep.s + el_off[ord[e]], els[ord[e]].len); gp += els[ord[e]].len; }
  }
  g[gp] = 0;
  *glen_out = gp;
  free(cut); free(ord); free(els); free(el_off); free(rep.s);
  return g;
}

/* homopolymer compression: keep the first base of each maximal run */
static uint64_t hpc_inplace(char *s, uint64_t n) {
  if (n == 0) return 0;
  uint64_t o = 1;
  for (uint64_t i = 1; i < n; i++)
    if (s[i] != s[o - 1]) s[o++] = s[i];
  return o;
}

int rs_generate(const rs_config *C, rs_reads *out) {
  memset(out, 0, sizeof(*out));
  rng_t r;
  rng_seed(&r, C->seed);
  uint64_t G;
  char *g = make_genome(C, &r, &G);
  out->genome = g;
  out->genome_len = G;

  uint64_t want = (uint64_t)(C->coverage * (double)G);
  uint64_t capr = 1024, capb = want + want / 8 + 1024;
  out->bases  = (char *)malloc(capb);
  out->start  = (uint64_t *)malloc(sizeof(uint64_t) * (capr + 1));
  out->len    = (uint32_t *)malloc(sizeof(uint32_t) * (capr + 1));
  out->pos    = (uint64_t *)malloc(sizeof(uint64_t) * (capr + 1));
  out->tlen   = (uint32_t *)malloc(sizeof(uint32_t) * (capr + 1));
  out->strand = (uint8_t *)malloc(capr + 1);
  out->start[0] = 0; out->len[0] = 0; out->pos[0] = 0; out->tlen[0] = 0; out->strand[0] = 0;

  sbuf rd = {0};
  uint64_t tot_template = 0, bp = 0;
  uint32_t n = 0;
  while (tot_template < want) {
    double L = exp(C->ln_mu + C->ln_sigma * rng_normal(&r));
    uint64_t len = (uint64_t)(L + 0.5);
    if (len < C->min_len || (C->max_len && len > C->max_len) || len > G) continue;
    uint64_t s0 = rng_below(&r, G - len + 1);
    int rc = (int)(rng_next(&r) & 1);
    rd.n = 0;
    const char *t = g + s0;
    for (uint64_t i = 0; i < len; i++) {
      double u = rng_unif(&r);
      if (!C->hp_errors) {
        if (u < C->p_sub)                              sb_push(&rd, other_base(&r, t[i]));
        else if (u < C->p_sub + C->p_del)              continue;
        else if (u < C->p_sub + C->p_del + C->p_ins) { sb_push(&rd, ACGT[rng_next(&r) & 3]); sb_push(&rd, t[i]); }
        else                                           sb_push(&rd, t[i]);
      } else {
        /* homopolymer-length errors: at the first base of a run, grow or shrink the run by one */
        int run_start = (i == 0) || (t[i - 1] != t[i]);
        if (u < C->p_sub)                                          sb_push(&rd, other_base(&r, t[i]));
        else if (run_start && u < C->p_sub + C->p_del)             continue;
        else if (run_start && u < C->p_sub + C->p_del + C->p_ins) { sb_push(&rd, t[i]); sb_push(&rd, t[i]); }
        else                                                       sb_push(&rd, t[i]);
      }
    }
    if (rc) {
      for (uint64_t i = 0, j = rd.n ? rd.n - 1 : 0; rd.n && i <= j; i++, j--) {
        char a = comp(rd.s[i]), b = comp(rd.s[j]);
        rd.s[i] = b; rd.s[j] = a;
        if (j == 0) break;
      }
    }
    uint64_t rl = rd.n;
    if (C->hpc) rl = hpc_inplace(rd.s, rd.n);
    if (rl >= (1u << 21)) continue;
    n++;
    if (n > capr) {
      capr *= 2;
      out->start  = (uint64_t *)realloc(out->start, sizeof(uint64_t) * (capr + 1));
      out->len    = (uint32_t *)realloc(out->len, sizeof(uint32_t) * (capr + 1));
      out->pos    = (uint64_t *)realloc(out->pos, sizeof(uint64_t) * (capr + 1));
      out->tlen   = (uint32_t *)realloc(out->tlen, sizeof(uint32_t) * (capr + 1));
      out->strand = (uint8_t *)realloc(out->strand, capr + 1);
    }
    if (bp + rl + 1 > capb) { capb = 2 * capb + rl + 1; out->bases = (char *)realloc(out->bases, capb); }
    out->start[n] = bp; out->len[n] = (uint32_t)rl; out->pos[n] = s0; out->tlen[n] = (uint32_t)len; out->strand[n] = (uint8_t)rc;
    memcpy(out->bases + bp, rd.s, rl);
    out->bases[bp + rl] = 0;
    bp += rl + 1;
    tot_template += len;
  }
  free(rd.s);
  out->nreads = n;
  out->total = bp - n;
  return 0;
}

void rs_free(rs_reads *R) {
  free(R->bases); free(R->start); free(R->len); free(R->pos); free(R->tlen); free(R->strand); free(R->genome);
  memset(R, 0, sizeof(*R));
}
