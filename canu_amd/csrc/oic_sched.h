// oic_sched.h -- processing order of candidate pairs.
//
// PLAIN    : discovery order (pair key order), one pair per wavefront.
// BATCHED  : longest-predicted-first by a static length-based cost proxy, so the heavy tail of
//            repeat pairs starts early and the persistent waves drain evenly.
// LEARNED  : longest-predicted-first by the cost model's predicted DP cells.
// The order never changes a result: every pair is processed completely by one wavefront.
#pragma once

#include <algorithm>
#include <cmath>
#include <numeric>
#include <vector>

#define OIC_NFEAT 12

struct PairFeat {
  float x[OIC_NFEAT];
  double cells;   // target: DP cells the pair took
  int32_t exit_d; // target: exit diagonal of the first extension
  int32_t n_ext;
};

// learned cost model: 12 -> 64 -> 64 -> 2 MLP, ReLU, inputs standardised
struct oic_sched_model {
  float mu[OIC_NFEAT], sd[OIC_NFEAT];
  float w1[64 * OIC_NFEAT], b1[64];
  float w2[64 * 64], b2[64];
  float w3[2 * 64], b3[2];
};

static inline float flog1p(double v) { return (float)std::log1p(v > 0 ? v : 0); }

// features of a candidate pair (host copy of its PairRec)
static inline void pair_features(const PairRec &p, const int32_t *len, PairFeat *f) {
  const double sl = len[p.s_id], tl = len[p.t_id];
  const double hits = (double)(p.he - p.hb);
  f->x[0] = flog1p(sl);
  f->x[1] = flog1p(tl);
  f->x[2] = flog1p(hits);
  f->x[3] = flog1p(p.nnodes);
  f->x[4] = (float)(p.dmin / 1000.0);
  f->x[5] = (float)(p.dmax / 1000.0);
  f->x[6] = flog1p((double)p.dmax - (double)p.dmin);
  f->x[7] = flog1p(p.max_mult);
  f->x[8] = (float)p.consistent;
  f->x[9] = (float)p.dir;
  f->x[10] = flog1p(std::min(sl, tl));
  f->x[11] = flog1p(p.longest);
  f->cells = 0;
  f->exit_d = 0;
  f->n_ext = 0;
}

// static proxy: one extension per match node, each about the length of the shorter read
static inline double static_cost(const PairRec &p, const int32_t *len) {
  const double m = std::min(len[p.s_id], len[p.t_id]);
  return (p.consistent ? 1.0 : (double)p.nnodes) * m;
}

static inline double model_predict(const oic_sched_model &M, const float *x) {
  float h1[64], h2[64], z[OIC_NFEAT];
  for (int i = 0; i < OIC_NFEAT; i++) z[i] = (x[i] - M.mu[i]) / M.sd[i];
  for (int j = 0; j < 64; j++) {
    float s = M.b1[j];
    for (int i = 0; i < OIC_NFEAT; i++) s += M.w1[j * OIC_NFEAT + i] * z[i];
    h1[j] = s > 0 ? s : 0;
  }
  for (int j = 0; j < 64; j++) {
    float s = M.b2[j];
    for (int i = 0; i < 64; i++) s += M.w2[j * 64 + i] * h1[i];
    h2[j] = s > 0 ? s : 0;
  }
  float o = M.b3[0];
  for (int i = 0; i < 64; i++) o += M.w3[i] * h2[i];
  return (double)o;
}

static inline int make_order(int schedule, const PairRec *d_pairs, uint64_t np, const int32_t *len, const oic_sched_model *model,
                             hipStream_t st, std::vector<uint32_t> &order) {
  order.resize(np);
  std::iota(order.begin(), order.end(), 0u);
  if (schedule == 0 || np < 2) return 0;
  std::vector<PairRec> pr(np);
  if (hipMemcpyAsync(pr.data(), d_pairs, np * sizeof(PairRec), hipMemcpyDeviceToHost, st) != hipSuccess) return -1;
  if (hipStreamSynchronize(st) != hipSuccess) return -1;
  std::vector<double> cost(np);
  if (schedule == 2 && model) {
    PairFeat f;
    for (uint64_t p = 0; p < np; p++) {
      pair_features(pr[p], len, &f);
      cost[p] = model_predict(*model, f.x);
    }
  } else {
    for (uint64_t p = 0; p < np; p++) cost[p] = static_cost(pr[p], len);
  }
  std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return cost[a] > cost[b]; });
  return 0;
}
