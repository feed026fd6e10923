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

// words of each view a lane keeps staged in LDS (128 bases)
#define LN_SW 8

// Wavefronts per SIMD the ring and staged words allow (160 KB of LDS per CU over 4 SIMDs), at most 8.
template <int CAP>
constexpr int lane_waves_per_eu() {
  constexpr int lds = (CAP + 2 * LN_SW) * OIC_WAVE * 4;
  constexpr int w = 163840 / lds / 4;
  return w < 1 ? 1 : (w > 8 ? 8 : w);
}

// lane states
#define LN_CELL   0   // computing cells of its current row (the inner loop)
#define LN_ROWEND 1   // its row is done: prune, first maximum, open the next row
#define LN_DPEND  2   // its DP is over (dp_* hold the outcome, or the reaching cell when dp_reach)
#define LN_FETCH  3   // needs its next task
#define LN_DONE   4   // no tasks left

template <int CAP, bool STATS>
__global__ void __launch_bounds__(OIC_WAVE) __attribute__((amdgpu_waves_per_eu(lane_waves_per_eu<CAP>(), 8)))
k_fwd_lane(FwdArgs A) {
  static_assert((CAP & (CAP - 1)) == 0, "ring size is a power of two");
  constexpr int32_t MASK = CAP - 1;
  constexpr int32_t NONE = 0x7fffffff;
  __shared__ int32_t ring[CAP * OIC_WAVE];
  __shared__ uint32_t stage[2 * LN_SW * OIC_WAVE];
  const int lane = lane_id();
  int32_t *const my = ring + lane;              // slot s of this lane: my[s * 64]
  uint32_t *const sa = stage + lane, *const sb = stage + LN_SW * OIC_WAVE + lane;   // word k: s?[k * 64]
  const uint64_t ntasks = A.ntasks_dev ? (uint64_t)*A.ntasks_dev : A.ntasks;
  if (A.tstamp && lane == 0) atomicMin(&A.tstamp[0], (unsigned long long)wall_clock64());
  const uint32_t *__restrict__ words = A.R.words;
  const DPTables &tb = A.tb;

  int state = LN_FETCH;
  // task
  uint64_t ti = 0;
  int32_t limit = 0, right_e = 0, exit_d = 0;
  bool s_short = false, fwd = true, rmte = true;
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
  int32_t sta = 0, stb = 0;                     // first staged word of each view (relative to wa / wb)
  int32_t mina = 0, minb = 0;                   // lowest A and T position the row's cells started from
  bool fell = false;                            // a cell of the row read past the staged words
  int32_t nL = NONE, nR = 0, cbv = -1, cbd = 0, pbv = -1, pbd = 0;
  int32_t Longest = 0, MSL = 0, MSd = 0, MSe = 0;
  double MaxScore = 0.0;
  // outcome of the DP that just ended
  bool dp_mte = false, dp_ovf = false, dp_reach = false;
  int32_t dp_e = 0, dp_d = 0;
  uint64_t cells_acc = 0, bases_acc = 0, rows_acc = 0, win_acc = 0, iters = 0;

  // copy LN_SW words of each view from words ka / kb into the lane's staging slots
  auto restage = [&](int32_t ka, int32_t kb) {
    sta = ka;
    stb = kb;
#pragma unroll
    for (int k = 0; k < LN_SW; k++) { sa[k * OIC_WAVE] = wa[ka + k]; sb[k * OIC_WAVE] = wb[kb + k]; }
  };
  // a DP on two views starts with its row-0 slide (prefixEditDistance-forward.C:113-132)
  auto begin_dp = [&](int64_t a_b0, bool aN, int32_t am, int64_t b_b0, bool bN, int32_t bn, int32_t l) {
    wa = words + (a_b0 >> 4); ra = (int32_t)(a_b0 & 15);
    wb = words + (b_b0 >> 4); rb = (int32_t)(b_b0 & 15);
    na = aN ? a_b0 - ra : -1;
    nb = bN ? b_b0 - rb : -1;
    m = am; n = bn; lim = l;
    row0 = true; sliding = true; cont = false;
    d = 0; Row = 0; r0 = 0; rem = m;            // m <= n
    Longest = 0; MSL = 0; MSd = 0; MSe = 0; MaxScore = 0.0;
    restage(0, 0);
    state = LN_CELL;
  };
  // the reverse DP on the left tails when the forward one reached the end (extend.C:112-132)
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
    fwd = false;
    return true;
  };
  // the task's outcome: 1 not a dovetail (Process_Matches only removes the node), 2 dovetail, 0 not
  // computed (band outgrew the ring)
  auto record = [&](bool ovf, bool dove) {
    const FwdTask tk = A.tasks[ti];
    const uint64_t gnode = A.pairs[tk.pair].hb + tk.node;
    A.outc[gnode] = pack_outcome(ovf ? 0 : (dove ? 2 : 1), exit_d, cells);
    if (ovf && A.ovf) A.ovf[atomicAdd(A.novf, 1u)] = tk;
    if (!ovf) cells_acc += cells;
    bases_acc += bases;
    rows_acc += rows;
    win_acc += windows;
    state = LN_FETCH;
  };

  while (true) {
    // ---------------------------------------------------------------- outer: rows, DPs, tasks
    // Lanes leave the inner loop with a finished row, a finished DP or no task; each pass below
    // moves them on, until every lane computes cells again or has run out of tasks.
    while (__ballot(state == LN_ROWEND || state == LN_DPEND || state == LN_FETCH)) {
      if (state == LN_ROWEND) {
        // prune to the first and last kept cell, take the row's first maximum (forward.C:245-299),
        // open the next row with -2 sentinels just outside the old band (forward.C:147-153)
        bool go = true;
        if (e > 0) {
          if (nL == NONE) { dp_mte = false; dp_e = MSe; dp_d = MSd; state = LN_DPEND; go = false; }   // band emptied
          else {
            L = nL;
            R = nR;
            if (cbv > Longest) {
              Longest = cbv;
              const double Score = (double)cbv * tb.bmv - (double)e;
              if (Score > MaxScore) { MaxScore = Score; MSL = cbv; MSd = cbd; MSe = e; }
            }
            if (e >= lim) { dp_mte = false; dp_e = MSe; dp_d = MSd; state = LN_DPEND; go = false; }
          }
        }
        if (go) {
          e++;
          if (STATS) rows++;
          oR = R;
          L = L - 1;                             // max(Left - 1, -e): Left >= -(e - 1)
          R = R + 1;
          if (R - L + 1 > CAP) { dp_ovf = true; state = LN_DPEND; }
          else {
            eml = tb.match_limit[e];
            d = L;
            pv = -2;
            cv = -2;
            nL = NONE; nR = 0; cbv = -1; cbd = 0; pbv = -1; pbd = 0;
            // The next row's cells start at or after the lowest A and T positions this row's cells
            // started from (Edit_Array rows only grow, and pruning drops the lowest).  Move the
            // staged words there when they lag by LN_SW / 4 words or a cell read past them.
            const int32_t ka = (ra + mina) >> 4, kb = (rb + minb) >> 4;
            if (fell || ka - sta >= LN_SW / 4 || kb - stb >= LN_SW / 4) restage(ka, kb);
            fell = false;
            mina = 0x7fffffff;
            minb = 0x7fffffff;
            state = LN_CELL;
          }
        }
      }
      if (state == LN_DPEND) {
        if (dp_reach) {
          // the end of A or T was reached at (Row, d): branch-point test (forward.C:172-212)
          dp_reach = false;
          const double Score = (double)Row * tb.bmv - (double)e;
          const int32_t Tail_Len = Row - MSL;
          const double slope = (double)(MaxScore - Score) / (double)Tail_Len;
          const bool abort_ = ((uint32_t)e > tb.min_branch_end_dist / 2u) && ((uint32_t)Tail_Len >= tb.min_branch_end_dist) &&
                              (slope >= tb.slope_min);
          dp_mte = !abort_;
          dp_e = abort_ ? MSe : e;
          dp_d = abort_ ? MSd : d;
        }
        if (dp_ovf) record(true, false);
        else if (fwd) {
          rmte = dp_mte;
          right_e = dp_e;
          exit_d += s_short ? dp_d : -dp_d;
          if (!(rmte && begin_reverse())) record(false, rmte);
        } else {
          record(false, rmte && dp_mte);
        }
      }
      const uint64_t want = __ballot(state == LN_FETCH);
      if (want) {
        uint32_t base = 0;
        if (lane == (int)__builtin_ctzll(want)) base = atomicAdd(A.queue, (uint32_t)__builtin_popcountll(want));
        base = __shfl(base, (int)__builtin_ctzll(want));
        if (state == LN_FETCH) {
          ti = (uint64_t)base + (uint64_t)__builtin_popcountll(want & ((1ull << lane) - 1));
          if (ti >= ntasks) {
            state = LN_DONE;
          } else {
            cells = 0; bases = 0; rows = 0; windows = 0;
            rmte = true; fwd = true; dp_ovf = false; dp_reach = false;
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
              // no forward DP: the right end is reached with 0 errors (prefixEditDistance-extend.C:81-99)
              if (!begin_reverse()) record(false, true);
            } else {
              s_short = S_Right_Len <= T_Right_Len;   // the shorter tail is A
              const int64_t s0 = (pr.dir ? A.R.b0r[pr.s_id] : A.R.b0f[pr.s_id]) + S_Right_Begin, t0 = A.R.b0f[pr.t_id] + T_Right_Begin;
              const bool sN = A.R.hasN[pr.s_id], tN = A.R.hasN[pr.t_id];
              if (s_short) begin_dp(s0, sN, S_Right_Len, t0, tN, T_Right_Len, limit);
              else         begin_dp(t0, tN, T_Right_Len, s0, sN, S_Right_Len, limit);
            }
          }
        }
      }
    }
    const uint64_t live = __ballot(state != LN_DONE);
    if (live == 0) break;
    // ---------------------------------------------------------------- inner: one cell step per lane
    // A step starts a cell (the furthest-reaching maximum of forward.C:155-160) or continues its
    // slide by one window (:161-167); a finished cell is stored in place and updates the row's
    // pruning range and first maximum.  The loop runs while three quarters of the live lanes are
    // still in their row's cells, so the outer work above is done for many lanes at once.
    const int nlive = __builtin_popcountll(live);
    do {
      if (STATS) iters++;
      if (state == LN_CELL) {
        if (!sliding) {
          nv = (d + 1 <= oR) ? my[((d + 1) & MASK) * OIC_WAVE] : -2;
          Row = cv + 1;
          Row = pv > Row ? pv : Row;
          Row = nv + 1 > Row ? nv + 1 : Row;
          r0 = Row;
          const int32_t r1 = m - Row, r2 = n - Row - d;
          rem = r1 < r2 ? r1 : r2;
          cont = false;
          mina = Row < mina ? Row : mina;
          minb = Row + d < minb ? Row + d : minb;
        }
        int32_t k = 0;
        bool staged = false;
        if (rem > 0) {
          if (STATS) windows++;
          const int32_t qa = ra + Row, qb = rb + Row + d;
          const int32_t ia = (qa >> 4) - sta, ib = (qb >> 4) - stb;
          staged = (uint32_t)ia < (uint32_t)(LN_SW - 1) && (uint32_t)ib < (uint32_t)(LN_SW - 1) && (na & nb) < 0;
          if (staged) {
            // 16 bases of each view from the staged words (two LDS reads each)
            const uint32_t xa = __builtin_amdgcn_alignbit(sa[(ia + 1) * OIC_WAVE], sa[ia * OIC_WAVE], (uint32_t)(qa & 15) * 2u);
            const uint32_t xb = __builtin_amdgcn_alignbit(sb[(ib + 1) * OIC_WAVE], sb[ib * OIC_WAVE], (uint32_t)(qb & 15) * 2u);
            uint32_t x = xa ^ xb;
            x = (x | (x >> 1)) & 0x55555555u;
            k = x ? (int32_t)(__builtin_ctz(x) >> 1) : 16;
          } else {
            // past the staged words: from HBM, 64 bases at a time once a slide goes on
            fell = true;
            if (!cont) k = (na & nb) < 0 ? lane_cmp16(wa, qa, wb, qb) : lane_cmp16_n(wa, qa, wb, qb, A.R.nmask, na, nb);
            else       k = lane_cmp64(wa, qa, wb, qb, A.R.nmask, na, nb);
          }
        }
        const int32_t wlen = (cont && !staged) ? 64 : 16;
        const bool last = k >= rem || k < wlen;
        Row += k < rem ? k : rem;
        rem -= wlen;
        sliding = !last;
        cont = !last;
        if (last) {
          cells++;
          bases += (uint32_t)(Row - r0) + 1u;
          if (row0) {
            row0 = false;
            if (Row == m) { dp_mte = true; dp_e = 0; dp_d = 0; state = LN_DPEND; }
            else {
              my[0] = Row;
              e = 0; L = 0; R = 0;
              mina = 0; minb = 0; fell = false;
              if (lim < 1) { dp_mte = false; dp_e = 0; dp_d = 0; state = LN_DPEND; }
              else state = LN_ROWEND;            // opens row 1
            }
          } else {
            my[(d & MASK) * OIC_WAVE] = Row;
            if (Row == m || Row + d == n) {
              dp_reach = true;
              state = LN_DPEND;
            } else {
              // "committed" is the best cell up to the latest kept cell, "pending" the best after it
              const bool keep = Row + (d > 0 ? d : 0) >= eml;
              if (keep) {
                nL = nL == NONE ? d : nL;
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
              if (d > R) state = LN_ROWEND;
            }
          }
        }
      }
    } while (4 * __builtin_popcountll(__ballot(state == LN_CELL)) >= 3 * nlive);
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
