// oic_sched.h -- processing order of candidate pairs.
//
// PLAIN    : discovery order (pair key order), one pair per wavefront.
// BATCHED  : longest-predicted-first by a static length-based cost proxy, so the heavy tail of
//            repeat pairs starts early and the persistent waves drain evenly.
// LEARNED  : longest-predicted-first by the cost model's predicted DP cells (oic_mlp.h).
// The order never changes a result: every pair is processed completely by one wavefront.
#pragma once

#include <algorithm>
#include <cmath>
#include <numeric>
#include <vector>

// static proxy: one extension per match node, each about the length of the shorter read
static inline double static_cost(const PairRec &p, const int32_t *len) {
  const double m = std::min(len[p.s_id], len[p.t_id]);
  return (p.consistent ? 1.0 : (double)p.nnodes) * m;
}

static inline int make_order(int schedule, const PairRec *d_pairs, uint64_t np, const int32_t *len, const float *pred,
                             hipStream_t st, std::vector<uint32_t> &order) {
  order.resize(np);
  std::iota(order.begin(), order.end(), 0u);
  if (schedule == 0 || np < 2) return 0;
  std::vector<PairRec> pr(np);
  if (hipMemcpyAsync(pr.data(), d_pairs, np * sizeof(PairRec), hipMemcpyDeviceToHost, st) != hipSuccess) return -1;
  if (hipStreamSynchronize(st) != hipSuccess) return -1;
  std::vector<double> cost(np);
  // a prediction that is not a finite number (a diverged or corrupt model) would make the
  // comparison below an invalid ordering: the whole batch then falls back to the static estimate
  bool use_pred = schedule == 2 && pred;
  for (uint64_t p = 0; use_pred && p < np; p++) use_pred = std::isfinite(pred[p]);
  if (use_pred) {
    for (uint64_t p = 0; p < np; p++) cost[p] = pred[p];
  } else {
    for (uint64_t p = 0; p < np; p++) cost[p] = static_cost(pr[p], len);
  }
  std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return cost[a] > cost[b]; });
  return 0;
}
