// oic_batch.h -- packed forward-extension DPs: G lanes per DP, 64/G DPs per wavefront.
//
// Extend_Alignment always runs forward() on the right tails first (prefixEditDistance-extend.C:81-99).
// When it does not reach the end of a read (and -partial is off), Process_Matches can only remove
// that match node (overlapInCore-Process_String_Overlaps.C:457-490): the reverse DP, the deltas and
// the coordinates are never used.  A node's forward DP depends on nothing but the node, so the
// forward DPs of many nodes run here ahead of Process_Matches, as independent tasks, and their
// outcome is cached per node, one 64-bit word (pack_outcome) holding its kind,
//   1  not a dovetail extension (Process_Matches skips the DP, removes the node)
//   2  dovetail (Process_Matches runs the whole extension for the deltas)
//   0  not computed (band outgrew the tile),
// with the run's DP cells and right-end diagonal, so that a finished task makes a single store.
// Tasks run 64/G to a wavefront (G lanes each, a T-cell band tile in LDS; the autotuner picks G and
// T, 8 x 64 on repeat-rich reads); those whose band outgrows the tile are queued and re-run one to a
// wavefront with a 512-cell tile, then an 8192-cell one, which holds any band the error limit
// allows (2 x 2700 + 3 cells at 60 kbp and 4.5 %).
// The DP is wave_prefix_dp's row loop without row storage or traceback, on a G-lane group, with the
// first 1024 bases of both tails staged in LDS.  When the forward DP reaches the end, the reverse DP
// on the left tails runs too, so the cached outcome is the extension's kind.  Two loop forms:
// k_fwd_tasks runs each DP to its end (group_prefix_dp); k_fwd_sm advances every DP of the
// wavefront by one row per iteration and starts a finished DP's successor at once (sm_start, sm_row).
#pragma once

#include "oic_search.h"

#define OIC_SEQ_WIN 64      // words (16 bases each) of each tail staged in LDS
#define FWD_CHUNK 8         // consecutive tasks a group takes from the queue at once

struct FwdTask {
  uint32_t pair;
  uint32_t node;            // index within the pair's node list
};

struct FwdArgs {
  ReadsDev R;
  DPTables tb;
  const PairRec *pairs;
  const int4 *nodes;
  const FwdTask *tasks;
  uint64_t ntasks;
  uint32_t *queue;
  uint64_t *outc;           // per node (global node index = pair.hb + node): the cached outcome (pack_outcome)
  unsigned long long *stats;
  FwdTask *ovf;             // tasks whose band outgrew the tile are appended here (if not null)
  uint32_t *novf;
  const uint32_t *ntasks_dev;  // if not null, the task count is read here when the kernel starts
  unsigned long long *tstamp;  // if not null: [0] earliest wavefront start, [1] latest end (wall clock)
};

__global__ void k_fill_u32(uint32_t *__restrict__ p, uint64_t n, uint32_t v) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// the total of an exclusive scan, for launches that read their size from device memory
__global__ void k_scan_total(const uint32_t *__restrict__ in, const uint64_t *__restrict__ out, uint64_t n, uint32_t *__restrict__ total) {
  if (blockIdx.x == 0 && threadIdx.x == 0) *total = n ? (uint32_t)(out[n - 1] + in[n - 1]) : 0u;
}

__device__ __forceinline__ void wave_fence() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

// A tail staged in LDS: its first OIC_SEQ_WIN words, from the word holding base 0 of the view.
// Positions are 32-bit offsets from that word's first base; HBM is read only past the copy.
// The copy is held through an LDS-typed pointer: a pointer that may address either LDS or HBM
// (one chosen by a select) compiles to flat loads, which wait on both memory counters.
typedef const __attribute__((address_space(3))) uint32_t *lds_words;

struct Staged {
  int64_t w0;               // first staged word
  int32_t r0;               // base 0 of the view within word w0 (0..15)
  lds_words lds;
};

__device__ __forceinline__ Staged make_staged(int64_t b0, const uint32_t *lds) {
  Staged t;
  t.w0 = b0 >> 4;
  t.r0 = (int32_t)(b0 & 15);
  t.lds = (lds_words)lds;
  return t;
}

// 16 bases from view position p (p >= 0); words = all packed reads; SW words of the tail are staged
template <int SW = OIC_SEQ_WIN>
__device__ __forceinline__ uint32_t win16_staged(const uint32_t *__restrict__ words, const Staged &t, int32_t p) {
  const int32_t q = t.r0 + p;
  const int32_t wi = q >> 4;
  const uint32_t sh = (uint32_t)q * 2u;   // alignbit shifts by this modulo 32: 2 (q mod 16)
  if (wi + 1 < SW) return __builtin_amdgcn_alignbit(t.lds[wi + 1], t.lds[wi], sh);
  return __builtin_amdgcn_alignbit(words[t.w0 + wi + 1], words[t.w0 + wi], sh);
}

// maximum over the G lanes of a group, by DPP moves within rows of 16
template <int G>
__device__ __forceinline__ uint64_t group_max_u64(uint64_t v) {
  auto step = [&](uint32_t lo, uint32_t hi) {
    const uint64_t o = ((uint64_t)hi << 32) | lo;
    v = o > v ? o : v;
  };
#define OIC_DPP64(ctl) step(__builtin_amdgcn_mov_dpp((uint32_t)v, ctl, 0xF, 0xF, false), \
                            __builtin_amdgcn_mov_dpp((uint32_t)(v >> 32), ctl, 0xF, 0xF, false))
  if (G >= 2) OIC_DPP64(0xB1);    // quad_perm [1,0,3,2]
  if (G >= 4) OIC_DPP64(0x4E);    // quad_perm [2,3,0,1]
  if (G >= 8) OIC_DPP64(0x141);   // row_half_mirror
  if (G >= 16) OIC_DPP64(0x140);  // row_mirror
#undef OIC_DPP64
  if (G >= 32) step(__shfl_xor((uint32_t)v, 16), __shfl_xor((uint32_t)(v >> 32), 16));
  if (G >= 64) step(__shfl_xor((uint32_t)v, 32), __shfl_xor((uint32_t)(v >> 32), 32));
  return v;
}

// minimum of lo and maximum of hi over the G lanes of a group, by DPP moves within rows of 16
template <int G>
__device__ __forceinline__ void group_min_max(int32_t &lo, int32_t &hi) {
#define OIC_DPPMM(ctl)                                                               \
  {                                                                                  \
    const int32_t ol = __builtin_amdgcn_mov_dpp(lo, ctl, 0xF, 0xF, false);           \
    const int32_t oh = __builtin_amdgcn_mov_dpp(hi, ctl, 0xF, 0xF, false);           \
    lo = ol < lo ? ol : lo;                                                          \
    hi = oh > hi ? oh : hi;                                                          \
  }
  if (G >= 2) OIC_DPPMM(0xB1);    // quad_perm [1,0,3,2]
  if (G >= 4) OIC_DPPMM(0x4E);    // quad_perm [2,3,0,1]
  if (G >= 8) OIC_DPPMM(0x141);   // row_half_mirror
  if (G >= 16) OIC_DPPMM(0x140);  // row_mirror
#undef OIC_DPPMM
  if (G >= 32) { const int32_t ol = __shfl_xor(lo, 16), oh = __shfl_xor(hi, 16); lo = ol < lo ? ol : lo; hi = oh > hi ? oh : hi; }
  if (G >= 64) { const int32_t ol = __shfl_xor(lo, 32), oh = __shfl_xor(hi, 32); lo = ol < lo ? ol : lo; hi = oh > hi ? oh : hi; }
}

// lexicographic (max v, then min d) over the G lanes of a group, as one 64-bit key per lane
// (v >= -1, |d| < 2^30; a lane's d matters only where its v is the group's maximum and >= 0)
template <int G>
__device__ __forceinline__ void group_argmax(int32_t &v, int32_t &d) {
  uint64_t k = ((uint64_t)(uint32_t)(v + 1) << 32) | (uint32_t)(0x40000000 - d);
  k = group_max_u64<G>(k);
  v = (int32_t)(uint32_t)(k >> 32) - 1;
  d = 0x40000000 - (int32_t)(uint32_t)k;
}

template <int G>
__device__ __forceinline__ uint64_t group_ballot(bool p, int g) {
  const uint64_t b = __ballot(p);
  return (G == 64) ? b : ((b >> (g * G)) & ((1ull << G) - 1));
}

struct GroupDP {
  bool mte;
  bool overflow;
  int32_t e;          // errors at the end (mte) or of the best branch point
  int32_t exit_d;     // diagonal of the end cell / best branch point
  uint32_t cells;
  uint32_t rows;      // rows e >= 1 entered
  uint32_t slots;     // lane-cells issued (G per band chunk)
  uint32_t wrows;     // row iterations the whole wavefront executed (counted on its first active lane)
  uint32_t bases;     // this lane's bases compared: slide lengths plus one terminating compare per cell
  uint32_t siters;    // this lane's 16-base window compares in slides
  uint32_t wsiters;   // slide-loop iterations the whole wavefront executed (counted on its first active lane)
};

// One prefix DP (forward() row loop, no row storage) on a G-lane group; sa/st are the group's LDS
// copies of the first OIC_SEQ_WIN words of a and b.
// a / b: start (in bases) of the two tails in the packed reads, and whether their reads hold N;
// words / nmask are the packed reads and N mask of the whole read set.
template <int G, int CAP, bool STATS>
__device__ GroupDP group_prefix_dp(const uint32_t *__restrict__ words, const uint32_t *__restrict__ nmask, const int64_t a_b0,
                                   const bool nA, const int32_t m, const int64_t b_b0, const bool nB, const int32_t n,
                                   const int32_t limit, const DPTables &tb, int32_t *pbuf, int32_t *cbuf, uint32_t *sa, uint32_t *sb,
                                   int g, int gl) {
  constexpr int32_t MASK = CAP - 1;
  GroupDP r;
  r.mte = false; r.overflow = false; r.e = 0; r.exit_d = 0; r.cells = 1; r.rows = 0; r.slots = 0; r.wrows = 0; r.bases = 0; r.siters = 0; r.wsiters = 0;
  const Staged ta = make_staged(a_b0, sa), tb_ = make_staged(b_b0, sb);
  for (int i = gl; i < OIC_SEQ_WIN; i += G) {
    sa[i] = words[ta.w0 + i];
    sb[i] = words[tb_.w0 + i];
  }
  wave_fence();
  int32_t row0 = 0;
  for (int32_t base = 0;; base += 16 * G) {
    const int32_t off = base + 16 * gl;
    const int32_t rem = m - off;
    int32_t k = 0;
    bool stop = true;
    if (rem > 0) {
      uint32_t x = win16_staged(words, ta, off) ^ win16_staged(words, tb_, off);
      x = (x | (x >> 1)) & 0x55555555u;
      if (nA) x &= ~nwin16(nmask, a_b0 + off);
      if (nB) x &= ~nwin16(nmask, b_b0 + off);
      const int32_t kk = x ? (int32_t)(__builtin_ctz(x) >> 1) : 16;
      if (kk >= rem) { k = rem; stop = true; } else { k = kk; stop = kk < 16; }
    }
    const uint64_t bl = group_ballot<G>(stop, g);
    if (bl) {
      const int f = __builtin_ctzll(bl);
      row0 = base + 16 * f + __shfl(k, g * G + f);
      break;
    }
  }
  if (row0 == m) { r.mte = true; return r; }
  if (gl == 0) pbuf[0] = row0;
  int32_t Left = 0, Right = 0, Longest = 0, Best_d = 0, Best_e = 0, MSL = 0, MSd = 0, MSe = 0;
  double MaxScore = 0.0;
  // slide along diagonal d from Row while the bases match (prefixEditDistance-forward.C:155-167)
  auto slide = [&](int32_t Row, int32_t d) -> int32_t {
    const int32_t r1 = m - Row, r2 = n - Row - d;
    int32_t rem = r1 < r2 ? r1 : r2;
    int32_t adv = 0;
    while (rem > 0) {
      if (STATS) {
        r.siters++;
        if (lane_id() == __builtin_ctzll(__ballot(1))) r.wsiters++;
      }
      uint32_t x = win16_staged(words, ta, Row + adv) ^ win16_staged(words, tb_, Row + d + adv);
      x = (x | (x >> 1)) & 0x55555555u;
      if (nA) x &= ~nwin16(nmask, a_b0 + Row + adv);
      if (nB) x &= ~nwin16(nmask, b_b0 + Row + d + adv);
      const int32_t kk = x ? (int32_t)(__builtin_ctz(x) >> 1) : 16;
      if (kk >= rem) { adv += rem; break; }
      adv += kk;
      if (kk < 16) break;
      rem -= 16;
    }
    r.bases += (uint32_t)adv + 1u;
    return Row + adv;
  };
  // the end of A or T was reached on diagonal rd at row rrow: branch-point test (:172-212)
  auto reached = [&](int32_t e, int32_t rrow, int32_t rd) {
    const double Score = (double)rrow * tb.bmv - (double)e;
    const int32_t Tail_Len = rrow - MSL;
    const double slope = (double)(MaxScore - Score) / (double)Tail_Len;
    const bool abort_ = ((uint32_t)e > tb.min_branch_end_dist / 2u) && ((uint32_t)Tail_Len >= tb.min_branch_end_dist) &&
                        (slope >= tb.slope_min);
    if (abort_) { r.mte = false; r.e = MSe; r.exit_d = MSd; }
    else        { r.mte = true;  r.e = e;   r.exit_d = rd; }
  };
  for (int32_t e = 1; e <= limit; e++) {
    const int32_t eml = tb.match_limit[e];
    Left = (Left - 1 > -e) ? Left - 1 : -e;
    Right = (Right + 1 < e) ? Right + 1 : e;
    if (STATS) {
      r.rows++;
      if (lane_id() == __builtin_ctzll(__ballot(1))) r.wrows++;
    }
    if (Right - Left + 3 > CAP) { r.overflow = true; return r; }
    wave_fence();
    if (gl < 4) {
      const int32_t sd = (gl == 0) ? Left : (gl == 1) ? Left - 1 : (gl == 2) ? Right : Right + 1;
      pbuf[sd & MASK] = -2;
    }
    wave_fence();
    // One pass over the band computes the row and, on the fly, the pruned range [nL, nR] (first and
    // last cell with Row + max(d, 0) >= Edit_Match_Limit[e], forward.C:245-280) and the first-max
    // cell inside it: per lane, "committed" holds the best cell at or before the latest kept cell,
    // "pending" the best cell after it, promoted when a later kept cell appears.
    int32_t nL = 0x7fffffff, nR = 0;
    int32_t cbv = -1, cbd = 0x7fffffff, pbv = -1, pbd = 0x7fffffff;
    for (int32_t c0 = Left; c0 <= Right; c0 += G) {
      const int32_t d = c0 + gl;
      const bool act = d <= Right;
      if (STATS) r.slots += G;
      int32_t Row = 0;
      if (act) {
        const int32_t x1 = 1 + pbuf[d & MASK], x2 = pbuf[(d - 1) & MASK], x3 = 1 + pbuf[(d + 1) & MASK];
        Row = x1 > x2 ? x1 : x2;
        Row = Row > x3 ? Row : x3;
        Row = slide(Row, d);
        cbuf[d & MASK] = Row;
      }
      const bool reach = act && (Row == m || Row + d == n);
      const uint64_t bl = group_ballot<G>(reach, g);
      if (bl) {
        const int f = __builtin_ctzll(bl);
        r.cells += (uint32_t)(f + 1);
        reached(e, __shfl(Row, g * G + f), c0 + f);
        return r;
      }
      r.cells += (uint32_t)((Right - c0 + 1) < G ? (Right - c0 + 1) : G);
      const bool keep = act && (Row + (d > 0 ? d : 0) >= eml);
      const uint64_t kb = group_ballot<G>(keep, g);
      if (kb) {
        const int32_t kl = c0 + (63 - __builtin_clzll(kb));
        if (nL == 0x7fffffff) nL = c0 + __builtin_ctzll(kb);
        nR = kl;
        if (pbv > cbv || (pbv == cbv && pbd < cbd)) { cbv = pbv; cbd = pbd; }
        pbv = -1;
        pbd = 0x7fffffff;
        if (act && d >= nL) {
          if (d <= kl) {
            if (Row > cbv || (Row == cbv && d < cbd)) { cbv = Row; cbd = d; }
          } else {
            pbv = Row;
            pbd = d;
          }
        }
      } else if (act && nL != 0x7fffffff) {
        if (Row > pbv || (Row == pbv && d < pbd)) { pbv = Row; pbd = d; }
      }
    }
    if (nL == 0x7fffffff) break;
    Left = nL;
    Right = nR;
    group_argmax<G>(cbv, cbd);
    // Score = Longest * bmv - e only rises when Longest does: with Longest unchanged since row e',
    // this row's score is at most row e''s, which MaxScore already holds (rounding is monotonic).
    if (cbv > Longest) {
      Longest = cbv; Best_d = cbd; Best_e = e;
      const double Score = (double)Longest * tb.bmv - (double)e;
      if (Score > MaxScore) { MaxScore = Score; MSL = Longest; MSd = Best_d; MSe = Best_e; }
    }
    int32_t *t = pbuf; pbuf = cbuf; cbuf = t;
  }
  r.mte = false; r.e = MSe; r.exit_d = MSd;
  return r;
}

// Extend_Alignment's outcome for one node without deltas: forward on the right tails, then (only
// if it reached the end) reverse on the left tails with the remaining error budget.
template <int G, int CAP, bool STATS>
__global__ void __launch_bounds__(OIC_WAVE) k_fwd_tasks(FwdArgs A) {
  constexpr int NG = OIC_WAVE / G;
  __shared__ int32_t tile[NG][2][CAP];
  __shared__ uint32_t seqA[NG][OIC_SEQ_WIN], seqT[NG][OIC_SEQ_WIN];
  const int lane = lane_id();
  const int g = lane / G, gl = lane % G;
  // per-launch totals per lane stay far below 2^32 (about 10^6 per group at the c3 workload)
  uint64_t cells_acc = 0;
  uint32_t rows_acc = 0, slots_acc = 0, wrows_acc = 0, bases_acc = 0, siters_acc = 0, wsiters_acc = 0;
  uint64_t ti = 0, tend = 0;
  const uint64_t ntasks = A.ntasks_dev ? (uint64_t)*A.ntasks_dev : A.ntasks;
  if (A.tstamp && lane == 0) atomicMin(&A.tstamp[0], (unsigned long long)wall_clock64());

  while (true) {
    if (ti == tend) {                        // group-uniform: take the next FWD_CHUNK tasks
      uint32_t f = 0;
      if (gl == 0) f = atomicAdd(A.queue, (uint32_t)FWD_CHUNK);
      ti = (uint32_t)__shfl((int)f, g * G);  // from the group's first lane
      tend = ti + FWD_CHUNK < ntasks ? ti + FWD_CHUNK : ntasks;
    }
    if (ti >= ntasks) break;                 // group-uniform
    uint32_t cells = 0, rows = 0, slots = 0, wrows = 0, bases = 0, siters = 0, wsiters = 0;
    bool rmte = true, overflow = false, lmte = true;
    int32_t right_e = 0, exit_d = 0, limit = 0;
    {
      const FwdTask tk = A.tasks[ti];
      const PairRec pr = A.pairs[tk.pair];
      const int4 nd = A.nodes[pr.hb + tk.node];
      const int32_t S_Len = A.R.len[pr.s_id], T_Len = A.R.len[pr.t_id];
      const int32_t Start = nd.z, Offset = nd.x, Len = nd.y;
      const int32_t S_Right_Begin = Start + Len, S_Right_Len = S_Len - S_Right_Begin;
      const int32_t T_Right_Begin = Offset + Len, T_Right_Len = T_Len - T_Right_Begin;
      const int32_t Total_Olap = (Start < Offset ? Start : Offset) + Len + (S_Right_Len < T_Right_Len ? S_Right_Len : T_Right_Len);
      limit = A.tb.error_bound[Total_Olap];
      exit_d = T_Right_Begin - S_Right_Begin;
      if (!(S_Right_Len == 0 || T_Right_Len == 0)) {
        const int64_t s0 = (pr.dir ? A.R.b0r[pr.s_id] : A.R.b0f[pr.s_id]) + S_Right_Begin, t0 = A.R.b0f[pr.t_id] + T_Right_Begin;
        const bool sN = A.R.hasN[pr.s_id], tN = A.R.hasN[pr.t_id];
        const bool s_short = S_Right_Len <= T_Right_Len;   // the shorter tail is A (prefixEditDistance-extend.C:81-99)
        const GroupDP f = group_prefix_dp<G, CAP, STATS>(A.R.words, A.R.nmask, s_short ? s0 : t0, s_short ? sN : tN,
                                                  s_short ? S_Right_Len : T_Right_Len, s_short ? t0 : s0, s_short ? tN : sN,
                                                  s_short ? T_Right_Len : S_Right_Len, limit, A.tb, tile[g][0], tile[g][1], seqA[g],
                                                  seqT[g], g, gl);
        cells += f.cells;
        rows += f.rows;
        slots += f.slots;
        wrows += f.wrows;
        bases += f.bases;
        siters += f.siters;
        wsiters += f.wsiters;
        overflow = f.overflow;
        rmte = f.mte;
        right_e = f.e;
        exit_d += s_short ? f.exit_d : -f.exit_d;
      }
    }
    // The task's fields are read again below rather than kept in registers across the DP.
    __asm__ volatile("" ::: "memory");
    const FwdTask tk = A.tasks[ti];
    const PairRec pr = A.pairs[tk.pair];
    const int4 nd = A.nodes[pr.hb + tk.node];
    const uint64_t gnode = pr.hb + tk.node;
    const int32_t S_Left_Begin = nd.z - 1, T_Left_Begin = nd.x - 1;
    if (!overflow && rmte && !(S_Left_Begin < 0 || T_Left_Begin < 0)) {
      const int32_t S_Len = A.R.len[pr.s_id], T_Len = A.R.len[pr.t_id];
      const int32_t S_Right_Begin = nd.z + nd.y, T_Right_Begin = nd.x + nd.y;
      // left tails, read rightwards on the reverse-complement strands
      const int64_t s0 = (pr.dir ? A.R.b0f[pr.s_id] : A.R.b0r[pr.s_id]) + (S_Len - 1 - S_Left_Begin);
      const int64_t t0 = A.R.b0r[pr.t_id] + (T_Len - 1 - T_Left_Begin);
      const bool sN = A.R.hasN[pr.s_id], tN = A.R.hasN[pr.t_id];
      const bool s_first = S_Right_Begin <= T_Right_Begin;   // prefixEditDistance-extend.C:112-132
      const GroupDP l = group_prefix_dp<G, CAP, STATS>(A.R.words, A.R.nmask, s_first ? s0 : t0, s_first ? sN : tN,
                                                s_first ? S_Left_Begin + 1 : T_Left_Begin + 1, s_first ? t0 : s0, s_first ? tN : sN,
                                                s_first ? T_Left_Begin + 1 : S_Left_Begin + 1, limit - right_e, A.tb, tile[g][0],
                                                tile[g][1], seqA[g], seqT[g], g, gl);
      cells += l.cells;
      rows += l.rows;
      slots += l.slots;
      wrows += l.wrows;
      bases += l.bases;
      siters += l.siters;
      wsiters += l.wsiters;
      overflow = l.overflow;
      lmte = l.mte;
    }
    if (gl == 0) {
      // 1: not a dovetail (Process_Matches only removes the node), 2: dovetail, 0: not computed
      A.outc[gnode] = pack_outcome(overflow ? 0 : ((rmte && lmte) ? 2 : 1), exit_d, cells);
      if (overflow && A.ovf) A.ovf[atomicAdd(A.novf, 1u)] = tk;
      if (!overflow) cells_acc += cells;
      rows_acc += rows;
      slots_acc += slots;
    }
    wrows_acc += wrows;
    bases_acc += bases;
    siters_acc += siters;
    wsiters_acc += wsiters;
    ti++;
  }
  uint64_t rows64 = rows_acc, slots64 = slots_acc, wrows64 = wrows_acc, bases64 = bases_acc, si64 = siters_acc,
           wsi64 = wsiters_acc;
  for (int o = 32; o > 0; o >>= 1) {
    cells_acc += __shfl_xor(cells_acc, o);
    rows64 += __shfl_xor(rows64, o);
    slots64 += __shfl_xor(slots64, o);
    wrows64 += __shfl_xor(wrows64, o);
    bases64 += __shfl_xor(bases64, o);
    si64 += __shfl_xor(si64, o);
    wsi64 += __shfl_xor(wsi64, o);
  }
  if (A.tstamp && lane == 0) atomicMax(&A.tstamp[1], (unsigned long long)wall_clock64());
  if (lane == 0 && A.stats) {
    atomicAdd(&A.stats[3], (unsigned long long)wrows64);
    atomicAdd(&A.stats[4], (unsigned long long)bases64);
    atomicAdd(&A.stats[5], (unsigned long long)si64);
    atomicAdd(&A.stats[6], (unsigned long long)wsi64);
    atomicAdd(&A.stats[0], (unsigned long long)cells_acc);
    atomicAdd(&A.stats[1], (unsigned long long)rows64);
    atomicAdd(&A.stats[2], (unsigned long long)slots64);
  }
}

// ---------------------------------------------------------------- state-machine form
//
// The same tasks and results as k_fwd_tasks, with the wavefront's loop turned inside out: each
// iteration advances every group by one DP row, and a group whose DP ends takes its next DP (the
// reverse one, or the next task) before the following row.  In k_fwd_tasks a group that finishes
// waits at the reconvergence point for the slowest DP of its wavefront.

// one group's DP in progress (every lane of the group holds the same values)
struct SMDP {
  int64_t a_b0, b_b0;
  Staged ta, tb;
  int32_t m, n, lim, e, Left, Right, Longest, MSL, MSd, MSe;
  double MaxScore;
  bool nA, nB;
};

// a DP's outcome, and the work it took (accumulated over a task's forward and reverse DPs)
struct SMRes {
  bool mte, overflow;
  int32_t e, d;
  uint32_t cells, rows, slots, wrows, bases, siters, wsiters, wchunks;
};

// N bases match any base (prefixEditDistance-forward.C:155-167 compares bytes, and an 'n' there is
// treated as equal): the two XOR bits of every N base in a 16-base window, cleared by the caller
__device__ __forceinline__ uint32_t nbits16(const uint32_t *__restrict__ nmask, int64_t p) {
  const uint32_t n = nwin16(nmask, p);
  return n | (n << 1);
}

// 16 bases of both tails from positions pa (in a) and pb (in b), compared: the number of leading
// equal bases (16 if all are)
template <int SW = OIC_SEQ_WIN>
__device__ __forceinline__ int32_t sm_match16(const SMDP &s, const uint32_t *__restrict__ words,
                                              const uint32_t *__restrict__ nmask, int32_t pa, int32_t pb) {
  // the XOR's lowest set bit lies in the first differing base (two bits a base)
  uint32_t x = win16_staged<SW>(words, s.ta, pa) ^ win16_staged<SW>(words, s.tb, pb);
  if (s.nA) x &= ~nbits16(nmask, s.a_b0 + pa);
  if (s.nB) x &= ~nbits16(nmask, s.b_b0 + pb);
  return x ? (int32_t)(__builtin_ctz(x) >> 1) : 16;
}

// the five words from word wi of a staged tail: from its LDS copy when they lie inside it, else
// from HBM (the packed reads carry 64 guard bases per strand and slack after the last one, so the
// words past a view's end exist)
template <int SW = OIC_SEQ_WIN>
__device__ __forceinline__ void words5_staged(const uint32_t *__restrict__ words, const Staged &t, int32_t wi, uint32_t (&w)[5]) {
  if (wi + 4 < SW) {
#pragma unroll
    for (int k = 0; k < 5; k++) w[k] = t.lds[wi + k];
  } else {
    const uint32_t *g = words + (t.w0 + wi);
#pragma unroll
    for (int k = 0; k < 5; k++) w[k] = g[k];
  }
}

// the three words from word wi of a staged tail: read from its LDS copy (at a clamped index, so
// always in bounds), and reloaded from HBM by the lanes whose words lie past the copy
template <int SW = OIC_SEQ_WIN>
__device__ __forceinline__ void words3_staged(const uint32_t *__restrict__ words, const Staged &t, int32_t wi, uint32_t &w0,
                                              uint32_t &w1, uint32_t &w2) {
  const int32_t wl = wi < SW - 3 ? wi : SW - 3;
  w0 = t.lds[wl]; w1 = t.lds[wl + 1]; w2 = t.lds[wl + 2];
  if (wi > SW - 3) {
    const uint32_t *g = words + (t.w0 + wi);
    w0 = g[0]; w1 = g[1]; w2 = g[2];
  }
}

// 32 bases of both tails from positions pa and pb, compared: the number of leading equal bases (32
// if all are).  A slide's first window: at 8 % divergence about 2 % of cells slide past 16 bases,
// and one such lane makes the whole wavefront run the 64-base continuation; past 32 bases far fewer.
template <int SW = OIC_SEQ_WIN>
__device__ __forceinline__ int32_t sm_match32(const SMDP &s, const uint32_t *__restrict__ words,
                                              const uint32_t *__restrict__ nmask, int32_t pa, int32_t pb) {
  const int32_t qa = s.ta.r0 + pa, qb = s.tb.r0 + pb;
  uint32_t a0, a1, a2, b0, b1, b2;
  words3_staged<SW>(words, s.ta, qa >> 4, a0, a1, a2);
  words3_staged<SW>(words, s.tb, qb >> 4, b0, b1, b2);
  // alignbit shifts by its operand modulo 32: 2q is 2 (q mod 16) there
  const uint32_t sa = (uint32_t)qa * 2u, sb = (uint32_t)qb * 2u;
  uint32_t x0 = __builtin_amdgcn_alignbit(a1, a0, sa) ^ __builtin_amdgcn_alignbit(b1, b0, sb);
  uint32_t x1 = __builtin_amdgcn_alignbit(a2, a1, sa) ^ __builtin_amdgcn_alignbit(b2, b1, sb);
  if (s.nA) { x0 &= ~nbits16(nmask, s.a_b0 + pa); x1 &= ~nbits16(nmask, s.a_b0 + pa + 16); }
  if (s.nB) { x0 &= ~nbits16(nmask, s.b_b0 + pb); x1 &= ~nbits16(nmask, s.b_b0 + pb + 16); }
  // the lowest set bit of x1:x0 lies in the first differing base (two bits a base); 64 when none
  return (int32_t)(__builtin_ctzg(((uint64_t)x1 << 32) | x0, 64) >> 1);
}

// 64 bases of both tails from positions pa and pb, compared: the number of leading equal bases (64
// if all are).  The five words of each tail are loaded together.
template <int SW = OIC_SEQ_WIN>
__device__ __forceinline__ int32_t sm_match64(const SMDP &s, const uint32_t *__restrict__ words,
                                              const uint32_t *__restrict__ nmask, int32_t pa, int32_t pb) {
  const int32_t qa = s.ta.r0 + pa, qb = s.tb.r0 + pb;
  uint32_t a[5], b[5];
  words5_staged<SW>(words, s.ta, qa >> 4, a);
  words5_staged<SW>(words, s.tb, qb >> 4, b);
  const uint32_t sa = (uint32_t)qa * 2u, sb = (uint32_t)qb * 2u;   // alignbit takes them modulo 32
#pragma unroll
  for (int k = 0; k < 4; k++) {
    uint32_t x = __builtin_amdgcn_alignbit(a[k + 1], a[k], sa) ^ __builtin_amdgcn_alignbit(b[k + 1], b[k], sb);
    if (s.nA) x &= ~nbits16(nmask, s.a_b0 + pa + 16 * k);
    if (s.nB) x &= ~nbits16(nmask, s.b_b0 + pb + 16 * k);
    if (x) return 16 * k + (int32_t)(__builtin_ctz(x) >> 1);
  }
  return 64;
}

// Stage the tails and run row 0 (the exact-prefix slide, prefixEditDistance-forward.C:113-132).
// Returns true when the DP is already over (res holds its outcome).
template <int G, int SW = OIC_SEQ_WIN>
__device__ __forceinline__ bool sm_start(SMDP &s, SMRes &res, const uint32_t *__restrict__ words,
                                         const uint32_t *__restrict__ nmask, int32_t *pbuf, uint32_t *sa, uint32_t *sb, int g,
                                         int gl) {
  res.mte = false; res.overflow = false; res.e = 0; res.d = 0;
  s.ta = make_staged(s.a_b0, sa);
  s.tb = make_staged(s.b_b0, sb);
  for (int i = gl; i < SW; i += G) {
    sa[i] = words[s.ta.w0 + i];
    sb[i] = words[s.tb.w0 + i];
  }
  wave_fence();
  int32_t row0 = 0;
  for (int32_t base = 0;; base += 16 * G) {
    const int32_t off = base + 16 * gl;
    const int32_t rem = s.m - off;
    int32_t k = 0;
    bool stop = true;
    if (rem > 0) {
      const int32_t kk = sm_match16<SW>(s, words, nmask, off, off);
      if (kk >= rem) { k = rem; stop = true; } else { k = kk; stop = kk < 16; }
    }
    const uint64_t bl = group_ballot<G>(stop, g);
    if (bl) {
      const int f = __builtin_ctzll(bl);
      row0 = base + 16 * f + __shfl(k, g * G + f);
      break;
    }
  }
  res.cells += 1;
  if (row0 == s.m) { res.mte = true; return true; }
  if (gl == 0) pbuf[0] = row0;
  s.e = 0; s.Left = 0; s.Right = 0; s.Longest = 0; s.MSL = 0; s.MSd = 0; s.MSe = 0;
  s.MaxScore = 0.0;
  return s.lim < 1;
}

// first maximum (largest Row, then smallest d) of a stored row over diagonals [lo, hi], on a group;
// v = -1 when the range is empty
template <int G, int CAP>
__device__ __forceinline__ void sm_row_first_max(const int32_t *buf, int32_t lo, int32_t hi, int gl, int32_t &v, int32_t &dv) {
  v = -1;
  dv = 0;
  for (int32_t d = lo + gl; d <= hi; d += G) {
    const int32_t r = buf[d & (CAP - 1)];
    dv = r > v ? d : dv;
    v = r > v ? r : v;
  }
  group_argmax<G>(v, dv);
}

// Longest / Max_Score bookkeeping of row er, from that row's first maximum (v at diagonal dv)
// (prefixEditDistance-forward.C:292-306)
__device__ __forceinline__ void sm_longest(SMDP &s, const DPTables &tb, int32_t v, int32_t dv, int32_t er) {
  if (v > s.Longest) {
    s.Longest = v;
    const double Score = (double)v * tb.bmv - (double)er;
    if (Score > s.MaxScore) { s.MaxScore = Score; s.MSL = v; s.MSd = dv; s.MSe = er; }
  }
}

// Row s.e + 1 of the DP: the body of group_prefix_dp's row loop.  Returns true when the DP is over.
//
// Per cell, each lane keeps only its own first and last kept diagonal and its first maximum of the
// previous row; the group reduces them once per row.  The previous row's first maximum (its
// Longest update, prefixEditDistance-forward.C:292-306) is read from the same three band loads the
// recurrence makes: this row's band [Left, Right] is the previous row's kept range widened by one
// on each side, whose two new ends hold the -2 sentinels, so every previous-row cell is read
// exactly once as 1 + EAL[e-1][d].  The update therefore lands one row late, which nothing can
// see: the branch-point test reads Max_Score only when a cell reaches an end, and there (a rare
// path) the previous row is scanned first; the last row (e = Error_Limit) is scanned at its end.
template <int G, int CAP, bool STATS>
__device__ __forceinline__ bool sm_row(SMDP &s, SMRes &res, const uint32_t *__restrict__ words,
                                       const uint32_t *__restrict__ nmask, const DPTables &tb, int32_t *&pbuf, int32_t *&cbuf,
                                       int g, int gl) {
  constexpr int32_t MASK = CAP - 1;
  const int32_t e = ++s.e;
  const int32_t eml = tb.match_limit[e];
  s.Left = (s.Left - 1 > -e) ? s.Left - 1 : -e;
  s.Right = (s.Right + 1 < e) ? s.Right + 1 : e;
  if (STATS) {
    res.rows++;
    if (lane_id() == __builtin_ctzll(__ballot(1))) res.wrows++;
  }
  if (s.Right - s.Left + 3 > CAP) { res.overflow = true; return true; }
  wave_fence();
  {
    // sentinels at Left - 1, Left, Right, Right + 1 (prefixEditDistance-forward.C:147-153), from
    // lanes 0-3 (positions from gl itself: no per-lane constant is kept across rows)
    const int32_t sd = gl + ((gl < 2) ? s.Left - 1 : s.Right - 2);
    if (gl < 4) pbuf[sd & MASK] = -2;
  }
  wave_fence();
  const int32_t Left = s.Left, Right = s.Right;
  int32_t kmin = 0x7fffffff, kmax = -0x7fffffff;   // this lane's first and last kept diagonal
  int32_t pv = -1, pd = 0;                         // this lane's first maximum of the previous row
  for (int32_t c0 = Left; c0 <= Right; c0 += G) {
    const int32_t d = c0 + gl;
    const bool act = d <= Right;
    if (STATS) res.slots += G;
    if (STATS && lane_id() == __builtin_ctzll(__ballot(1))) res.wchunks++;
    bool reach = false;
    int32_t Row = 0;
    if (act) {
      const int32_t p1 = pbuf[d & MASK], x2 = pbuf[(d - 1) & MASK], x3 = 1 + pbuf[(d + 1) & MASK];
      pd = p1 > pv ? d : pd;
      pv = p1 > pv ? p1 : pv;
      Row = 1 + p1;
      Row = Row > x2 ? Row : x2;
      Row = Row > x3 ? Row : x3;
      // slide along diagonal d while the bases match (prefixEditDistance-forward.C:155-167): a
      // 32-base window first (nearly all slides end in it), then 64-base windows (long matches)
      const int32_t r1 = s.m - Row, r2 = s.n - Row - d;
      const int32_t rem0 = r1 < r2 ? r1 : r2;     // >= 0: no cell of the previous row reached an end
      const int32_t kk = sm_match32(s, words, nmask, Row, Row + d);
      int32_t adv = kk < rem0 ? kk : rem0;
      if (STATS) {
        res.siters++;
        if (lane_id() == __builtin_ctzll(__ballot(1))) res.wsiters++;
      }
      int32_t rem = (kk == 32 && rem0 > 32) ? rem0 - 32 : 0;
      while (rem > 0) {
        if (STATS) {
          res.siters++;
          if (lane_id() == __builtin_ctzll(__ballot(1))) res.wsiters++;
        }
        const int32_t k64 = sm_match64(s, words, nmask, Row + adv, Row + d + adv);
        if (k64 >= rem) { adv += rem; break; }
        adv += k64;
        if (k64 < 64) break;
        rem -= 64;
      }
      res.bases += (uint32_t)adv + 1u;
      reach = adv == rem0;                        // Row + adv is m, or Row + adv + d is n
      Row += adv;
      cbuf[d & MASK] = Row;
      const bool keep = Row + (d > 0 ? d : 0) >= eml;   // not pruned (prefixEditDistance-forward.C:245-285)
      kmin = (keep & (kmin == 0x7fffffff)) ? d : kmin;
      kmax = keep ? d : kmax;
    }
    if (__ballot(reach)) {
      const uint64_t bl = group_ballot<G>(reach, g);
      if (bl) {
        // the end of A or T was reached: branch-point test (prefixEditDistance-forward.C:172-212),
        // after the previous row's Longest update
        const int f = __builtin_ctzll(bl);
        res.cells += (uint32_t)(c0 + f - Left + 1);
        const int32_t rrow = __shfl(Row, g * G + f);
        if (e >= 2) {
          int32_t v, dv;
          sm_row_first_max<G, CAP>(pbuf, Left + 1, Right - 1, gl, v, dv);
          sm_longest(s, tb, v, dv, e - 1);
        }
        const double Score = (double)rrow * tb.bmv - (double)e;
        const int32_t Tail_Len = rrow - s.MSL;
        const double slope = (double)(s.MaxScore - Score) / (double)Tail_Len;
        const bool abort_ = ((uint32_t)e > tb.min_branch_end_dist / 2u) && ((uint32_t)Tail_Len >= tb.min_branch_end_dist) &&
                            (slope >= tb.slope_min);
        if (abort_) { res.mte = false; res.e = s.MSe; res.d = s.MSd; }
        else        { res.mte = true;  res.e = e;     res.d = c0 + f; }
        return true;
      }
    }
  }
  res.cells += (uint32_t)(Right - Left + 1);
  if (e >= 2) {
    group_argmax<G>(pv, pd);
    sm_longest(s, tb, pv, pd, e - 1);
  }
  // the kept range (prefixEditDistance-forward.C:245-285): first and last kept diagonal
  int32_t nL = kmin, nR = kmax;
  group_min_max<G>(nL, nR);
  if (nL == 0x7fffffff) { res.mte = false; res.e = s.MSe; res.d = s.MSd; return true; }
  s.Left = nL;
  s.Right = nR;
  if (e >= s.lim) {
    int32_t v, dv;
    sm_row_first_max<G, CAP>(cbuf, nL, nR, gl, v, dv);
    sm_longest(s, tb, v, dv, e);
    res.mte = false; res.e = s.MSe; res.d = s.MSd;
    return true;
  }
  int32_t *t = pbuf; pbuf = cbuf; cbuf = t;
  return false;
}

// Wavefronts per SIMD the kernel's LDS allows (160 KB per CU over 4 SIMDs, rounded up), at most 5:
// registers are held to that occupancy (5: at most 102 VGPRs), no further.  (8 x 64: 8 KiB per
// wavefront, 20 per CU.)
template <int G, int CAP>
constexpr int sm_waves_per_eu() {
  constexpr int lds = (OIC_WAVE / G) * (2 * CAP * 4 + 2 * OIC_SEQ_WIN * 4);
  constexpr int w = (163840 / lds + 3) / 4;
  return w < 1 ? 1 : (w > 5 ? 5 : w);
}

template <int G, int CAP, bool STATS>
__global__ void __launch_bounds__(OIC_WAVE) __attribute__((amdgpu_waves_per_eu(sm_waves_per_eu<G, CAP>(), 8)))
k_fwd_sm(FwdArgs A) {
  constexpr int NG = OIC_WAVE / G;
  __shared__ int32_t tile[NG][2][CAP];
  __shared__ uint32_t seqA[NG][OIC_SEQ_WIN], seqT[NG][OIC_SEQ_WIN];
  const int lane = lane_id();
  const int g = lane / G, gl = lane % G;
  int32_t *pbuf = tile[g][0], *cbuf = tile[g][1];
  uint64_t cells_acc = 0;
  uint32_t rows_acc = 0, slots_acc = 0, wrows_acc = 0, bases_acc = 0, siters_acc = 0, wsiters_acc = 0, wchunks_acc = 0;
  uint64_t ti = 0, tend = 0;
  const uint64_t ntasks = A.ntasks_dev ? (uint64_t)*A.ntasks_dev : A.ntasks;
  if (A.tstamp && lane == 0) atomicMin(&A.tstamp[0], (unsigned long long)wall_clock64());

  SMDP s;
  SMRes res;
  int state = 0;   // 0: take the next task, 1: forward rows, 2: reverse rows, 3: no tasks left
  int32_t limit = 0, right_e = 0, exit_d = 0;
  bool s_short = false, rmte = true, lmte = true;

  // the reverse DP on the left tails, when the forward one reached the end (prefixEditDistance-extend.C:112-132);
  // returns true when it has rows to run
  auto start_reverse = [&]() -> bool {
    const FwdTask tk = A.tasks[ti];
    const PairRec pr = A.pairs[tk.pair];
    const int4 nd = A.nodes[pr.hb + tk.node];
    const int32_t S_Left_Begin = nd.z - 1, T_Left_Begin = nd.x - 1;
    if (res.overflow || !rmte || S_Left_Begin < 0 || T_Left_Begin < 0) return false;
    const int32_t S_Len = A.R.len[pr.s_id], T_Len = A.R.len[pr.t_id];
    const int32_t S_Right_Begin = nd.z + nd.y, T_Right_Begin = nd.x + nd.y;
    const int64_t s0 = (pr.dir ? A.R.b0f[pr.s_id] : A.R.b0r[pr.s_id]) + (S_Len - 1 - S_Left_Begin);
    const int64_t t0 = A.R.b0r[pr.t_id] + (T_Len - 1 - T_Left_Begin);
    const bool sN = A.R.hasN[pr.s_id], tN = A.R.hasN[pr.t_id];
    const bool s_first = S_Right_Begin <= T_Right_Begin;
    s.a_b0 = s_first ? s0 : t0;
    s.nA = s_first ? sN : tN;
    s.m = s_first ? S_Left_Begin + 1 : T_Left_Begin + 1;
    s.b_b0 = s_first ? t0 : s0;
    s.nB = s_first ? tN : sN;
    s.n = s_first ? T_Left_Begin + 1 : S_Left_Begin + 1;
    s.lim = limit - right_e;
    if (!sm_start<G>(s, res, A.R.words, A.R.nmask, pbuf, seqA[g], seqT[g], g, gl)) return true;
    lmte = res.mte;                          // over at row 0
    return false;
  };
  // record the task's outcome and move on
  auto finish = [&]() {
    const FwdTask tk = A.tasks[ti];
    const PairRec pr = A.pairs[tk.pair];
    const uint64_t gnode = pr.hb + tk.node;
    const bool overflow = res.overflow;
    if (gl == 0) {
      // 1: not a dovetail (Process_Matches only removes the node), 2: dovetail, 0: not computed
      A.outc[gnode] = pack_outcome(overflow ? 0 : ((rmte && lmte) ? 2 : 1), exit_d, res.cells);
      if (overflow && A.ovf) A.ovf[atomicAdd(A.novf, 1u)] = tk;
      if (!overflow) cells_acc += res.cells;
      rows_acc += res.rows;
      slots_acc += res.slots;
    }
    wrows_acc += res.wrows;
    bases_acc += res.bases;
    siters_acc += res.siters;
    wsiters_acc += res.wsiters;
    wchunks_acc += res.wchunks;
    ti++;
    state = 0;
  };
  // the forward DP is over: take its outcome, then run the reverse DP or finish the task
  auto forward_done = [&]() {
    rmte = res.mte;
    right_e = res.e;
    exit_d += s_short ? res.d : -res.d;
    if (start_reverse()) state = 2;
    else finish();
  };

  while (true) {
    while (state == 0) {
      if (ti == tend) {                      // group-uniform: take the next FWD_CHUNK tasks
        uint32_t f = 0;
        if (gl == 0) f = atomicAdd(A.queue, (uint32_t)FWD_CHUNK);
        ti = (uint32_t)__shfl((int)f, g * G);  // from the group's first lane
        tend = ti + FWD_CHUNK < ntasks ? ti + FWD_CHUNK : ntasks;
      }
      if (ti >= ntasks) { state = 3; break; }
      res.cells = 0; res.rows = 0; res.slots = 0; res.wrows = 0; res.bases = 0; res.siters = 0; res.wsiters = 0; res.wchunks = 0;
      res.mte = true; res.overflow = false; res.e = 0; res.d = 0;
      lmte = true;
      const FwdTask tk = A.tasks[ti];
      const PairRec pr = A.pairs[tk.pair];
      const int4 nd = A.nodes[pr.hb + tk.node];
      const int32_t S_Len = A.R.len[pr.s_id], T_Len = A.R.len[pr.t_id];
      const int32_t Start = nd.z, Offset = nd.x, Len = nd.y;
      const int32_t S_Right_Begin = Start + Len, S_Right_Len = S_Len - S_Right_Begin;
      const int32_t T_Right_Begin = Offset + Len, T_Right_Len = T_Len - T_Right_Begin;
      const int32_t Total_Olap = (Start < Offset ? Start : Offset) + Len + (S_Right_Len < T_Right_Len ? S_Right_Len : T_Right_Len);
      limit = A.tb.error_bound[Total_Olap];
      exit_d = T_Right_Begin - S_Right_Begin;
      s_short = S_Right_Len <= T_Right_Len;   // the shorter tail is A (prefixEditDistance-extend.C:81-99)
      if (S_Right_Len == 0 || T_Right_Len == 0) {
        res.d = 0;
        forward_done();                      // no forward DP: the right end is reached with 0 errors
        continue;
      }
      const int64_t s0 = (pr.dir ? A.R.b0r[pr.s_id] : A.R.b0f[pr.s_id]) + S_Right_Begin, t0 = A.R.b0f[pr.t_id] + T_Right_Begin;
      const bool sN = A.R.hasN[pr.s_id], tN = A.R.hasN[pr.t_id];
      s.a_b0 = s_short ? s0 : t0;
      s.nA = s_short ? sN : tN;
      s.m = s_short ? S_Right_Len : T_Right_Len;
      s.b_b0 = s_short ? t0 : s0;
      s.nB = s_short ? tN : sN;
      s.n = s_short ? T_Right_Len : S_Right_Len;
      s.lim = limit;
      if (!sm_start<G>(s, res, A.R.words, A.R.nmask, pbuf, seqA[g], seqT[g], g, gl)) { state = 1; break; }
      forward_done();
    }
    if (__ballot(state != 3) == 0) break;    // every group has run out of tasks
    if (state == 1 || state == 2) {
      if (sm_row<G, CAP, STATS>(s, res, A.R.words, A.R.nmask, A.tb, pbuf, cbuf, g, gl)) {
        if (state == 1) {
          forward_done();
        } else {
          lmte = res.mte;
          finish();
        }
      }
    }
  }
  uint64_t rows64 = rows_acc, slots64 = slots_acc, wrows64 = wrows_acc, bases64 = bases_acc, si64 = siters_acc,
           wsi64 = wsiters_acc, wch64 = wchunks_acc;
  for (int o = 32; o > 0; o >>= 1) {
    cells_acc += __shfl_xor(cells_acc, o);
    rows64 += __shfl_xor(rows64, o);
    slots64 += __shfl_xor(slots64, o);
    wrows64 += __shfl_xor(wrows64, o);
    bases64 += __shfl_xor(bases64, o);
    si64 += __shfl_xor(si64, o);
    wsi64 += __shfl_xor(wsi64, o);
    wch64 += __shfl_xor(wch64, o);
  }
  if (A.tstamp && lane == 0) atomicMax(&A.tstamp[1], (unsigned long long)wall_clock64());
  if (lane == 0 && A.stats) {
    atomicAdd(&A.stats[3], (unsigned long long)wrows64);
    atomicAdd(&A.stats[4], (unsigned long long)bases64);
    atomicAdd(&A.stats[5], (unsigned long long)si64);
    atomicAdd(&A.stats[6], (unsigned long long)wsi64);
    atomicAdd(&A.stats[7], (unsigned long long)wch64);
    atomicAdd(&A.stats[0], (unsigned long long)cells_acc);
    atomicAdd(&A.stats[1], (unsigned long long)rows64);
    atomicAdd(&A.stats[2], (unsigned long long)slots64);
  }
}

// probe tasks: each pair's first-longest node (Process_String_Overlaps.C:424-431)
__global__ void k_probe_tasks(const PairRec *__restrict__ pairs, uint64_t np, FwdTask *__restrict__ tasks) {
  const uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  tasks[p].pair = (uint32_t)p;
  tasks[p].node = (uint32_t)pairs[p].first_longest;   // found by the list builders
}

// Pairs whose other nodes are all worth a forward DP ahead of Process_Matches: those whose probe
// was not a dovetail (each later node is reached unless it is itself the next one removed), and
// some of those with a dovetail probe, whose nodes on the probe's alignment Process_Matches will
// remove without extending them (overlapInCore-Process_String_Overlaps.C:476-490) while repeat
// copies on shifted diagonals survive.  For a dovetail probe the static rule speculates when the
// pair has at least SPEC_DOVE_MIN nodes; with a cost model (pred: its log(1 + DP cells) of the
// pair's Process_Matches, per pair) it speculates when the model expects more than SPEC_DOVE_GAIN
// times the probe's own cells, i.e. further extensions beyond the probe.  Counts are written in
// processing order.  Which nodes run ahead never changes a result.
#define SPEC_DOVE_MIN 16
#define SPEC_DOVE_GAIN 2.0f
__global__ void k_spec_count(const PairRec *__restrict__ pairs, uint64_t np, const FwdTask *__restrict__ probe,
                             const uint64_t *__restrict__ outc, const float *__restrict__ pred,
                             const uint32_t *__restrict__ order, uint32_t *__restrict__ cnt) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  const uint32_t p = order[i];
  const PairRec pr = pairs[p];
  const uint64_t po = outc[pr.hb + probe[p].node];
  const int pc = outcome_kind(po);
  bool dove_spec = pr.nnodes >= SPEC_DOVE_MIN;
  if (pred) dove_spec = expm1f(pred[2 * (uint64_t)p]) > SPEC_DOVE_GAIN * (float)outcome_cells(po);
  const bool spec = !pr.consistent && pr.nnodes > 1 && (pc == 1 || (pc == 2 && dove_spec));
  cnt[i] = spec ? (uint32_t)(pr.nnodes - 1) : 0u;
}

__global__ void k_spec_write(const PairRec *__restrict__ pairs, uint64_t np, const FwdTask *__restrict__ probe,
                             const uint32_t *__restrict__ order, const uint32_t *__restrict__ cnt, const uint64_t *__restrict__ off,
                             FwdTask *__restrict__ tasks) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np || cnt[i] == 0) return;
  const uint32_t p = order[i];
  const int32_t nn = pairs[p].nnodes;
  const uint32_t skip = probe[p].node;
  uint64_t w = off[i];
  for (int32_t i = 0; i < nn; i++) {
    if ((uint32_t)i == skip) continue;
    tasks[w].pair = (uint32_t)p;
    tasks[w].node = (uint32_t)i;
    w++;
  }
}

// sort keys of packed DP tasks by the seed length of their node, longest first, so that the
// DPs sharing a wavefront (which waits for its slowest) tend to be alike; values are the tasks
__global__ void k_task_keys(const FwdTask *__restrict__ t, uint64_t n, const PairRec *__restrict__ pairs, const int4 *__restrict__ nodes,
                            uint64_t *__restrict__ keys, uint64_t *__restrict__ vals) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const FwdTask x = t[i];
  const int32_t len = nodes[pairs[x.pair].hb + x.node].y;
  keys[i] = 0xFFFFull - (uint64_t)(len < 0xFFFF ? len : 0xFFFF);
  vals[i] = (uint64_t)x.pair | ((uint64_t)x.node << 32);
}
