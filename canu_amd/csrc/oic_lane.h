// oic_lane.h -- packed forward-extension DPs, one DP per lane: 64 DPs per wavefront, each advanced
// one cell per loop iteration.
//
// The same tasks and outcomes as k_fwd_sm (oic_batch.h): a node's forward DP on the right tails
// (prefixEditDistance-extend.C:81-99), then, when it reaches the end, the reverse DP on the left
// tails (:112-132), cached as 1 (not a dovetail), 2 (dovetail) or 0 (band outgrew the tile).
//
// Why one lane per DP.  Bands are narrow (about 15 cells on repeat-rich reads at 4.5 %), so a DP
// spread over a group of lanes spends most instructions on per-row work shared by the group:
// sentinels, ballots, group reductions for pruning and the first maximum, and a chunk loop that runs
// as long as the widest band of the wavefront.  Here a lane walks its own band cell by cell and keeps
// the row's bookkeeping in its own registers: the pruning range and the first maximal cell are
// running values, no lane waits on another.  Lanes are out of step: in one iteration one lane computes
// a cell while another continues a slide (one 16- or 64-base window per iteration), closes a row or
// starts its next DP.
//
// Row storage: one ring of CAP ints per lane in LDS, laid out [slot][lane] so that every access of a
// wavefront hits 64 distinct banks.  A row is updated in place: cell d reads prev[d + 1] from the ring
// and keeps prev[d - 1], prev[d] in registers, then overwrites prev[d] (forward.C:155-167 reads row
// e - 1 only at d - 1, d, d + 1, and the four cells just outside the old band are the -2 sentinels
// of forward.C:147-153).  A band wider than CAP goes to the retry levels as in k_fwd_sm.
//
// Bases are compared from HBM (the packed reads, through L2), 16 at a time with one alignbit per
// strand, then 64 at a time while a slide goes on.
#pragma once

#include "oic_batch.h"

// 16 bases from position q (in bases, q >= 0) of the words at w, against 16 of w2 from q2: the
// number of leading equal bases (16 if all are)
__device__ __forceinline__ int32_t lane_cmp16(const uint32_t *__restrict__ wa, int32_t qa, const uint32_t *__restrict__ wb, int32_t qb) {
  const uint32_t *pa = wa + (qa >> 4), *pb = wb + (qb >> 4);
  const uint32_t xa = __builtin_amdgcn_alignbit(pa[1], pa[0], (uint32_t)(qa & 15) * 2u);
  const uint32_t xb = __builtin_amdgcn_alignbit(pb[1], pb[0], (uint32_t)(qb & 15) * 2u);
  uint32_t x = xa ^ xb;
  x = (x | (x >> 1)) & 0x55555555u;
  return x ? (int32_t)(__builtin_ctz(x) >> 1) : 16;
}

// the same with 'n' masks (reads holding N): na / nb are the N-mask bit positions of base 0 of each
// view (or -1 when that read has no N)
__device__ __forceinline__ int32_t lane_cmp16_n(const uint32_t *__restrict__ wa, int32_t qa, const uint32_t *__restrict__ wb, int32_t qb,
                                               const uint32_t *__restrict__ nmask, int64_t na, int64_t nb) {
  const uint32_t *pa = wa + (qa >> 4), *pb = wb + (qb >> 4);
  const uint32_t xa = __builtin_amdgcn_alignbit(pa[1], pa[0], (uint32_t)(qa & 15) * 2u);
  const uint32_t xb = __builtin_amdgcn_alignbit(pb[1], pb[0], (uint32_t)(qb & 15) * 2u);
  uint32_t x = xa ^ xb;
  x = (x | (x >> 1)) & 0x55555555u;
  if (na >= 0) x &= ~nwin16(nmask, na + qa);
  if (nb >= 0) x &= ~nwin16(nmask, nb + qb);
  return x ? (int32_t)(__builtin_ctz(x) >> 1) : 16;
}

// 64 bases: five words of each strand loaded together
__device__ __forceinline__ int32_t lane_cmp64(const uint32_t *__restrict__ wa, int32_t qa, const uint32_t *__restrict__ wb, int32_t qb,
                                             const uint32_t *__restrict__ nmask, int64_t na, int64_t nb) {
  const uint32_t *pa = wa + (qa >> 4), *pb = wb + (qb >> 4);
  uint32_t a[5], b[5];
#pragma unroll
  for (int k = 0; k < 5; k++) { a[k] = pa[k]; b[k] = pb[k]; }
  const uint32_t sa = (uint32_t)(qa & 15) * 2u, sb = (uint32_t)(qb & 15) * 2u;
#pragma unroll
  for (int k = 0; k < 4; k++) {
    uint32_t x = __builtin_amdgcn_alignbit(a[k + 1], a[k], sa) ^ __builtin_amdgcn_alignbit(b[k + 1], b[k], sb);
    x = (x | (x >> 1)) & 0x55555555u;
    if (na >= 0) x &= ~nwin16(nmask, na + qa + 16 * k);
    if (nb >= 0) x &= ~nwin16(nmask, nb + qb + 16 * k);
    if (x) return 16 * k + (int32_t)(__builtin_ctz(x) >> 1);
  }
  return 64;
}

// Wavefronts per SIMD the ring allows (160 KB of LDS per CU over 4 SIMDs), at most 8.
template <int CAP>
constexpr int lane_waves_per_eu() {
  constexpr int lds = CAP * OIC_WAVE * 4;
  constexpr int w = 163840 / lds / 4;
  return w < 1 ? 1 : (w > 8 ? 8 : w);
}

template <int CAP, bool STATS>
__global__ void __launch_bounds__(OIC_WAVE) __attribute__((amdgpu_waves_per_eu(lane_waves_per_eu<CAP>(), 8)))
k_fwd_lane(FwdArgs A) {
  static_assert((CAP & (CAP - 1)) == 0, "ring size is a power of two");
  constexpr int32_t MASK = CAP - 1;
  constexpr int32_t NONE = 0x7fffffff;
  __shared__ int32_t ring[CAP * OIC_WAVE];
  const int lane = lane_id();
  int32_t *const my = ring + lane;              // slot s of this lane: my[s * 64]
  const uint64_t ntasks = A.ntasks_dev ? (uint64_t)*A.ntasks_dev : A.ntasks;
  if (A.tstamp && lane == 0) atomicMin(&A.tstamp[0], (unsigned long long)wall_clock64());
  const uint32_t *__restrict__ words = A.R.words;
  const DPTables &tb = A.tb;

  // task
  int state = 0;                 // 0 needs a task, 1 forward DP, 2 reverse DP, 3 no tasks left
  uint64_t ti = 0;
  int32_t limit = 0, right_e = 0, exit_d = 0;
  bool s_short = false, rmte = true;
  uint32_t cells = 0, bases = 0, rows = 0, windows = 0;
  // the DP in progress
  const uint32_t *wa = words, *wb = words;      // word holding base 0 of each view
  int32_t ra = 0, rb = 0;                       // base 0's position within that word
  int64_t na = -1, nb = -1;                     // N-mask position of base 0 (-1: no N in the read)
  int32_t m = 0, n = 0, lim = 0;
  int32_t e = 0, L = 0, R = 0, oR = 0, d = 0, eml = 0;
  int32_t pv = -2, cv = -2, nv = -2;            // prev[d - 1], prev[d], prev[d + 1]
  int32_t Row = 0, rem = 0, r0 = 0;
  bool sliding = false, cont = false, row0 = false;
  int32_t nL = NONE, nR = 0, cbv = -1, cbd = 0, pbv = -1, pbd = 0;
  int32_t Longest = 0, MSL = 0, MSd = 0, MSe = 0;
  double MaxScore = 0.0;
  // outcome of the DP that just ended
  bool dmte = false, dovf = false;
  int32_t de = 0, dd = 0;
  uint64_t cells_acc = 0, bases_acc = 0, rows_acc = 0, win_acc = 0, iters = 0;

  // a DP on views (word, base within word, N position, length) starts with its row-0 slide
  auto begin_dp = [&](int64_t a_b0, bool aN, int32_t am, int64_t b_b0, bool bN, int32_t bn, int32_t l) {
    wa = words + (a_b0 >> 4); ra = (int32_t)(a_b0 & 15);
    wb = words + (b_b0 >> 4); rb = (int32_t)(b_b0 & 15);
    na = aN ? a_b0 - ra : -1;
    nb = bN ? b_b0 - rb : -1;
    m = am; n = bn; lim = l;
    row0 = true; sliding = true; cont = false;
    d = 0; Row = 0; r0 = 0; rem = m;            // m <= n
    Longest = 0; MSL = 0; MSd = 0; MSe = 0; MaxScore = 0.0;
  };
  // the reverse DP on the left tails when the forward one reached the end; false: none to run
  auto begin_reverse = [&]() -> bool {
    const FwdTask tk = A.tasks[ti];
    const PairRec pr = A.pairs[tk.pair];
    const int4 nd = A.nodes[pr.hb + tk.node];
    const int32_t S_Left_Begin = nd.z - 1, T_Left_Begin = nd.x - 1;
    if (S_Left_Begin < 0 || T_Left_Begin < 0) return false;
    const int32_t S_Len = A.R.len[pr.s_id], T_Len = A.R.len[pr.t_id];
    const int32_t S_Right_Begin = nd.z + nd.y, T_Right_Begin = nd.x + nd.y;
    const int64_t s0 = (pr.dir ? A.R.b0f[pr.s_id] : A.R.b0r[pr.s_id]) + (S_Len - 1 - S_Left_Begin);
    const int64_t t0 = A.R.b0r[pr.t_id] + (T_Len - 1 - T_Left_Begin);
    const bool sN = A.R.hasN[pr.s_id], tN = A.R.hasN[pr.t_id];
    const bool s_first = S_Right_Begin <= T_Right_Begin;
    if (s_first) begin_dp(s0, sN, S_Left_Begin + 1, t0, tN, T_Left_Begin + 1, limit - right_e);
    else         begin_dp(t0, tN, T_Left_Begin + 1, s0, sN, S_Left_Begin + 1, limit - right_e);
    return true;
  };

  while (true) {
    // ---- lanes without a DP take the next tasks, one each, in queue order
    const uint64_t want = __ballot(state == 0);
    if (want) {
      uint32_t base = 0;
      if (lane == (int)__builtin_ctzll(want)) base = atomicAdd(A.queue, (uint32_t)__builtin_popcountll(want));
      base = __shfl(base, (int)__builtin_ctzll(want));
      if (state == 0) {
        ti = (uint64_t)base + (uint64_t)__builtin_popcountll(want & ((1ull << lane) - 1));
        if (ti >= ntasks) {
          state = 3;
        } else {
          cells = 0; bases = 0; rows = 0; windows = 0; rmte = true; dovf = false;
          const FwdTask tk = A.tasks[ti];
          const PairRec pr = A.pairs[tk.pair];
          const int4 nd = A.nodes[pr.hb + tk.node];
          const int32_t S_Len = A.R.len[pr.s_id], T_Len = A.R.len[pr.t_id];
          const int32_t Start = nd.z, Offset = nd.x, Len = nd.y;
          const int32_t S_Right_Begin = Start + Len, S_Right_Len = S_Len - S_Right_Begin;
          const int32_t T_Right_Begin = Offset + Len, T_Right_Len = T_Len - T_Right_Begin;
          const int32_t Total_Olap = (Start < Offset ? Start : Offset) + Len + (S_Right_Len < T_Right_Len ? S_Right_Len : T_Right_Len);
          limit = tb.error_bound[Total_Olap];
          exit_d = T_Right_Begin - S_Right_Begin;
          right_e = 0;
          if (S_Right_Len == 0 || T_Right_Len == 0) {
            // no forward DP: the right end is reached with 0 errors
            state = begin_reverse() ? 2 : 0;
            if (state == 0) {
              const uint64_t gnode = pr.hb + tk.node;
              A.cache[gnode] = 2;
              A.ncells[gnode] = 0;
              A.nexit[gnode] = exit_d;
            }
          } else {
            s_short = S_Right_Len <= T_Right_Len;   // the shorter tail is A (prefixEditDistance-extend.C:81-99)
            const int64_t s0 = (pr.dir ? A.R.b0r[pr.s_id] : A.R.b0f[pr.s_id]) + S_Right_Begin, t0 = A.R.b0f[pr.t_id] + T_Right_Begin;
            const bool sN = A.R.hasN[pr.s_id], tN = A.R.hasN[pr.t_id];
            if (s_short) begin_dp(s0, sN, S_Right_Len, t0, tN, T_Right_Len, limit);
            else         begin_dp(t0, tN, T_Right_Len, s0, sN, S_Right_Len, limit);
            state = 1;
          }
        }
      }
    }
    const bool run = state == 1 || state == 2;
    if (__ballot(run) == 0) {
      if (__ballot(state == 0) == 0) break;   // every lane has run out of tasks
      continue;
    }
    if (STATS) iters++;
    bool dp_end = false, row_end = false;
    if (run) {
      // ---- one step of the current cell: its start (the furthest-reaching maximum), then one
      // window of its slide (prefixEditDistance-forward.C:155-167)
      if (!sliding) {
        nv = (d + 1 <= oR) ? my[((d + 1) & MASK) * OIC_WAVE] : -2;
        Row = cv + 1;
        Row = pv > Row ? pv : Row;
        Row = nv + 1 > Row ? nv + 1 : Row;
        r0 = Row;
        const int32_t r1 = m - Row, r2 = n - Row - d;
        rem = r1 < r2 ? r1 : r2;
        sliding = true;
        cont = false;
      }
      int32_t k = 0;
      if (rem > 0) {
        if (STATS) windows++;
        if (!cont) k = (na & nb) < 0 ? lane_cmp16(wa, ra + Row, wb, rb + Row + d)
                                     : lane_cmp16_n(wa, ra + Row, wb, rb + Row + d, A.R.nmask, na, nb);
        else       k = lane_cmp64(wa, ra + Row, wb, rb + Row + d, A.R.nmask, na, nb);
      }
      const int32_t wlen = cont ? 64 : 16;
      if (k >= rem) { Row += rem; rem = 0; }
      else          { Row += k; rem = (k == wlen) ? rem - wlen : 0; }
      if (rem > 0) {
        cont = true;                             // the slide goes on in the next iteration
      } else {
        sliding = false;
        cells++;
        bases += (uint32_t)(Row - r0) + 1u;
        if (row0) {
          // the exact-prefix slide (forward.C:113-132)
          row0 = false;
          if (Row == m) { dmte = true; de = 0; dd = 0; dp_end = true; }
          else {
            my[0] = Row;
            e = 0; L = 0; R = 0;
            if (lim < 1) { dmte = false; de = 0; dd = 0; dp_end = true; }
            else row_end = true;                 // opens row 1 below
          }
        } else {
          my[(d & MASK) * OIC_WAVE] = Row;
          if (Row == m || Row + d == n) {
            // the end of A or T: branch-point test (forward.C:172-212)
            const double Score = (double)Row * tb.bmv - (double)e;
            const int32_t Tail_Len = Row - MSL;
            const double slope = (double)(MaxScore - Score) / (double)Tail_Len;
            const bool abort_ = ((uint32_t)e > tb.min_branch_end_dist / 2u) && ((uint32_t)Tail_Len >= tb.min_branch_end_dist) &&
                                (slope >= tb.slope_min);
            dmte = !abort_;
            de = abort_ ? MSe : e;
            dd = abort_ ? MSd : d;
            dp_end = true;
          } else {
            // pruning range and first maximal cell of the row, as running values (forward.C:245-299):
            // "committed" is the best cell up to the latest kept cell, "pending" the best after it
            const bool keep = Row + (d > 0 ? d : 0) >= eml;
            if (keep) {
              if (nL == NONE) nL = d;
              nR = d;
              if (pbv > cbv) { cbv = pbv; cbd = pbd; }
              pbv = -1;
              if (Row > cbv) { cbv = Row; cbd = d; }
            } else if (nL != NONE && Row > pbv) {
              pbv = Row;
              pbd = d;
            }
            pv = cv;
            cv = nv;
            d++;
            row_end = d > R;
          }
        }
      }
    }
    if (row_end) {
      // ---- the row is done: prune, take its first maximum, open the next row
      bool go = true;
      if (e > 0) {
        if (nL == NONE) { dmte = false; de = MSe; dd = MSd; dp_end = true; go = false; }   // band emptied
        else {
          L = nL;
          R = nR;
          if (cbv > Longest) {
            Longest = cbv;
            const double Score = (double)cbv * tb.bmv - (double)e;
            if (Score > MaxScore) { MaxScore = Score; MSL = cbv; MSd = cbd; MSe = e; }
          }
          if (e >= lim) { dmte = false; de = MSe; dd = MSd; dp_end = true; go = false; }
        }
      }
      if (go) {
        e++;
        if (STATS) rows++;
        oR = R;
        L = L - 1;                               // max(Left - 1, -e): Left >= -(e - 1)
        R = R + 1;
        if (R - L + 1 > CAP) { dovf = true; dp_end = true; }
        else {
          eml = tb.match_limit[e];
          d = L;
          pv = -2;
          cv = -2;
          nL = NONE; nR = 0; cbv = -1; cbd = 0; pbv = -1; pbd = 0;
        }
      }
    }
    if (dp_end) {
      if (state == 1 && !dovf) {
        rmte = dmte;
        right_e = de;
        exit_d += s_short ? dd : -dd;
        if (!(rmte && begin_reverse())) state = 0;
        else state = 2;
      } else {
        state = 0;
      }
      if (state == 0) {
        // record the task's outcome: 1 not a dovetail (Process_Matches only removes the node),
        // 2 dovetail, 0 not computed (band outgrew the tile)
        const FwdTask tk = A.tasks[ti];
        const uint64_t gnode = A.pairs[tk.pair].hb + tk.node;
        A.cache[gnode] = dovf ? 0 : ((rmte && dmte) ? 2 : 1);
        if (dovf && A.ovf) A.ovf[atomicAdd(A.novf, 1u)] = tk;
        A.ncells[gnode] = cells;
        A.nexit[gnode] = exit_d;
        if (!dovf) cells_acc += cells;
        bases_acc += bases;
        rows_acc += rows;
        win_acc += windows;
      }
    }
  }
  for (int o = 32; o > 0; o >>= 1) {
    cells_acc += __shfl_xor(cells_acc, o);
    bases_acc += __shfl_xor(bases_acc, o);
    rows_acc += __shfl_xor(rows_acc, o);
    win_acc += __shfl_xor(win_acc, o);
  }
  if (A.tstamp && lane == 0) atomicMax(&A.tstamp[1], (unsigned long long)wall_clock64());
  if (lane == 0 && A.stats) {
    atomicAdd(&A.stats[0], (unsigned long long)cells_acc);
    atomicAdd(&A.stats[4], (unsigned long long)bases_acc);
    if (STATS) {
      atomicAdd(&A.stats[1], (unsigned long long)rows_acc);
      atomicAdd(&A.stats[2], (unsigned long long)(iters * OIC_WAVE));   // lane steps issued
      atomicAdd(&A.stats[3], (unsigned long long)iters);
      atomicAdd(&A.stats[5], (unsigned long long)win_acc);
      atomicAdd(&A.stats[6], (unsigned long long)iters);
    }
  }
}
