// oic_batch.h -- packed forward-extension DPs: G lanes per DP, 64/G DPs per wavefront.
//
// Extend_Alignment always runs forward() on the right tails first (prefixEditDistance-extend.C:81-99).
// When it does not reach the end of a read (and -partial is off), Process_Matches can only remove
// that match node (overlapInCore-Process_String_Overlaps.C:457-490): the reverse DP, the deltas and
// the coordinates are never used.  A node's forward DP depends on nothing but the node, so the
// forward DPs of many nodes run here ahead of Process_Matches, as independent tasks, and their
// outcome is cached per node:
//   cache[node] = 1  forward ended at a branch point / error limit (Process_Matches skips the DP)
//                 2  forward reached the end (Process_Matches runs the whole extension)
//                 0  not computed (band outgrew the tile)
// The DP below is wave_prefix_dp's row loop without row storage or traceback, on a G-lane group,
// with the first 1024 bases of both tails staged in LDS.
#pragma once

#include "oic_search.h"

#define OIC_SEQ_WIN 64      // words (16 bases each) of each tail staged in LDS

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
  uint8_t *cache;           // per node (global node index = pair.hb + node)
  uint32_t *ncells;         // per node: DP cells of the cached forward run
  int32_t *nexit;           // per node: right-end diagonal (T_Hi - S_Hi) of the cached run
  unsigned long long *stats;
};

__device__ __forceinline__ void wave_fence() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

// window of 16 bases at view position p: from the LDS copy when it covers p, else from HBM
__device__ __forceinline__ uint32_t win16_staged(const SeqV &v, const uint32_t *__restrict__ lw, int64_t w0, int64_t p) {
  const int64_t q = v.b0 + p;
  const int64_t wi = (q >> 4) - w0;
  const uint32_t s = (uint32_t)(q & 15) * 2u;
  if (wi >= 0 && wi + 1 < OIC_SEQ_WIN) return __builtin_amdgcn_alignbit(lw[wi + 1], lw[wi], s);
  return __builtin_amdgcn_alignbit(v.w[(q >> 4) + 1], v.w[q >> 4], s);
}

template <int G>
__device__ __forceinline__ uint64_t group_ballot(bool p, int g) {
  const uint64_t b = __ballot(p);
  return (G == 64) ? b : ((b >> (g * G)) & ((1ull << G) - 1));
}

template <int G, int CAP>
__global__ void __launch_bounds__(OIC_WAVE) k_fwd_tasks(FwdArgs A) {
  constexpr int NG = OIC_WAVE / G;
  constexpr int32_t MASK = CAP - 1;
  __shared__ int32_t tile[NG][2][CAP];
  __shared__ uint32_t seqA[NG][OIC_SEQ_WIN], seqT[NG][OIC_SEQ_WIN];
  __shared__ uint32_t fetched[NG];
  const int lane = lane_id();
  const int g = lane / G, gl = lane % G;
  uint64_t cells_acc = 0;

  while (true) {
    // ---- fetch a task for this group
    if (gl == 0) fetched[g] = atomicAdd(A.queue, 1u);
    wave_fence();
    const uint32_t ti = fetched[g];
    wave_fence();
    if ((uint64_t)ti >= A.ntasks) break;   // group-uniform
    const FwdTask tk = A.tasks[ti];
    const PairRec pr = A.pairs[tk.pair];
    const int4 nd = A.nodes[pr.hb + tk.node];
    const int32_t S_Len = A.R.len[pr.s_id], T_Len = A.R.len[pr.t_id];
    const int32_t Start = nd.z, Offset = nd.x, Len = nd.y;
    const int32_t S_Right_Begin = Start + Len, S_Right_Len = S_Len - S_Right_Begin;
    const int32_t T_Right_Begin = Offset + Len, T_Right_Len = T_Len - T_Right_Begin;
    const uint64_t gnode = pr.hb + tk.node;
    if (S_Right_Len == 0 || T_Right_Len == 0) {   // nothing to the right: matches to the end
      if (gl == 0) { A.cache[gnode] = 2; A.ncells[gnode] = 0; A.nexit[gnode] = 0; }
      continue;
    }
    const int32_t Total_Olap = (Start < Offset ? Start : Offset) + Len + (S_Right_Len < T_Right_Len ? S_Right_Len : T_Right_Len);
    const int32_t limit = A.tb.error_bound[Total_Olap];
    SeqV S = read_view(A.R, pr.s_id, pr.dir), T = read_view(A.R, pr.t_id, 0);
    SeqV a, b;
    int32_t m, n;
    const bool s_short = S_Right_Len <= T_Right_Len;
    if (s_short) { a = S; a.b0 += S_Right_Begin; m = S_Right_Len; b = T; b.b0 += T_Right_Begin; n = T_Right_Len; }
    else         { a = T; a.b0 += T_Right_Begin; m = T_Right_Len; b = S; b.b0 += S_Right_Begin; n = S_Right_Len; }
    const bool nA = a.nm != nullptr, nB = b.nm != nullptr;
    // stage the first OIC_SEQ_WIN words of each tail
    const int64_t wa0 = a.b0 >> 4, wb0 = b.b0 >> 4;
    for (int i = gl; i < OIC_SEQ_WIN; i += G) {
      seqA[g][i] = a.w[wa0 + i];
      seqT[g][i] = b.w[wb0 + i];
    }
    wave_fence();
    int32_t *pbuf = tile[g][0], *cbuf = tile[g][1];

    // row 0: group-cooperative slide of 16 * G bases per step
    int32_t row0 = 0;
    for (int32_t base = 0;; base += 16 * G) {
      const int32_t off = base + 16 * gl;
      const int32_t rem = m - off;
      int32_t k = 0;
      bool stop = true;
      if (rem > 0) {
        uint32_t x = win16_staged(a, seqA[g], wa0, off) ^ win16_staged(b, seqT[g], wb0, off);
        x = (x | (x >> 1)) & 0x55555555u;
        if (nA) x &= ~nwin16(a.nm, a.b0 + off);
        if (nB) x &= ~nwin16(b.nm, b.b0 + off);
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
    uint64_t cells = 1;
    bool mte = false;
    int32_t exit_d = 0;
    if (row0 == m) {
      mte = true;
    } else {
      if (gl == 0) pbuf[0] = row0;
      int32_t Left = 0, Right = 0, Longest = 0, Best_d = 0, MSL = 0, MSd = 0;
      double MaxScore = 0.0;
      bool ended = false, overflow = false;
      for (int32_t e = 1; e <= limit; e++) {
        Left = (Left - 1 > -e) ? Left - 1 : -e;
        Right = (Right + 1 < e) ? Right + 1 : e;
        if (Right - Left + 3 > CAP) { overflow = true; break; }
        wave_fence();
        if (gl < 4) {
          const int32_t sd = (gl == 0) ? Left : (gl == 1) ? Left - 1 : (gl == 2) ? Right : Right + 1;
          pbuf[sd & MASK] = -2;
        }
        wave_fence();
        int32_t rd = 0, rrow = 0;
        bool reached = false;
        for (int32_t c0 = Left; c0 <= Right; c0 += G) {
          const int32_t d = c0 + gl;
          const bool act = d <= Right;
          int32_t Row = 0;
          if (act) {
            const int32_t x1 = 1 + pbuf[d & MASK], x2 = pbuf[(d - 1) & MASK], x3 = 1 + pbuf[(d + 1) & MASK];
            Row = x1 > x2 ? x1 : x2;
            Row = Row > x3 ? Row : x3;
            const int32_t r1 = m - Row, r2 = n - Row - d;
            int32_t rem = r1 < r2 ? r1 : r2;
            int32_t adv = 0;
            while (rem > 0) {
              uint32_t x = win16_staged(a, seqA[g], wa0, Row + adv) ^ win16_staged(b, seqT[g], wb0, (int64_t)Row + d + adv);
              x = (x | (x >> 1)) & 0x55555555u;
              if (nA) x &= ~nwin16(a.nm, a.b0 + Row + adv);
              if (nB) x &= ~nwin16(b.nm, b.b0 + Row + d + adv);
              const int32_t kk = x ? (int32_t)(__builtin_ctz(x) >> 1) : 16;
              if (kk >= rem) { adv += rem; break; }
              adv += kk;
              if (kk < 16) break;
              rem -= 16;
            }
            Row += adv;
            cbuf[d & MASK] = Row;
          }
          const bool reach = act && (Row == m || Row + d == n);
          const uint64_t bl = group_ballot<G>(reach, g);
          if (bl) {
            const int f = __builtin_ctzll(bl);
            cells += (uint64_t)(f + 1);
            rd = c0 + f;
            rrow = __shfl(Row, g * G + f);
            reached = true;
            break;
          }
          cells += (uint64_t)((Right - c0 + 1) < G ? (Right - c0 + 1) : G);
        }
        if (reached) {
          const double Score = (double)rrow * A.tb.bmv - (double)e;
          const int32_t Tail_Len = rrow - MSL;
          const double slope = (double)(MaxScore - Score) / (double)Tail_Len;
          const bool abort_ = ((uint32_t)e > A.tb.min_branch_end_dist / 2u) && ((uint32_t)Tail_Len >= A.tb.min_branch_end_dist) &&
                              (slope >= A.tb.slope_min);
          if (abort_) { mte = false; exit_d = MSd; }
          else        { mte = true; exit_d = rd; }
          ended = true;
          break;
        }
        wave_fence();
        const int32_t eml = A.tb.match_limit[e];
        int32_t nL = Right + 1;
        for (int32_t c0 = Left; c0 <= Right; c0 += G) {
          const int32_t d = c0 + gl;
          const bool keep = (d <= Right) && (cbuf[d & MASK] + (d > 0 ? d : 0) >= eml);
          const uint64_t bl = group_ballot<G>(keep, g);
          if (bl) { nL = c0 + __builtin_ctzll(bl); break; }
        }
        if (nL > Right) break;
        int32_t nR = nL;
        for (int32_t c1 = Right; c1 >= nL; c1 -= G) {
          const int32_t d = c1 - gl;
          const bool keep = (d >= nL) && (cbuf[d & MASK] + (d > 0 ? d : 0) >= eml);
          const uint64_t bl = group_ballot<G>(keep, g);
          if (bl) { nR = c1 - __builtin_ctzll(bl); break; }
        }
        Left = nL;
        Right = nR;
        int32_t bv = -1, bd = 0x7fffffff;
        for (int32_t c0 = Left; c0 <= Right; c0 += G) {
          const int32_t d = c0 + gl;
          if (d <= Right) {
            const int32_t v = cbuf[d & MASK];
            if (v > bv) { bv = v; bd = d; }
          }
        }
#pragma unroll
        for (int o = 1; o < G; o <<= 1) {
          const int32_t ov = __shfl_xor(bv, o), od = __shfl_xor(bd, o);
          if (ov > bv || (ov == bv && od < bd)) { bv = ov; bd = od; }
        }
        if (bv > Longest) { Longest = bv; Best_d = bd; }
        const double Score = (double)Longest * A.tb.bmv - (double)e;
        if (Score > MaxScore) { MaxScore = Score; MSL = Longest; MSd = Best_d; }
        int32_t *t = pbuf; pbuf = cbuf; cbuf = t;
      }
      if (!ended && !overflow) { mte = false; exit_d = MSd; }
      if (overflow) {
        if (gl == 0) A.cache[gnode] = 0;
        continue;
      }
    }
    if (gl == 0) {
      A.cache[gnode] = mte ? 2 : 1;
      A.ncells[gnode] = (uint32_t)(cells > 0xffffffffull ? 0xffffffffull : cells);
      // right-end diagonal T_Hi - S_Hi of the run, in S/T orientation
      A.nexit[gnode] = (T_Right_Begin - S_Right_Begin) + (s_short ? exit_d : -exit_d);
      cells_acc += cells;
    }
  }
  for (int o = 32; o > 0; o >>= 1) cells_acc += __shfl_xor(cells_acc, o);
  if (lane == 0 && A.stats) atomicAdd(&A.stats[0], (unsigned long long)cells_acc);
}

// probe tasks: each pair's first-longest node (Process_String_Overlaps.C:424-431)
__global__ void k_probe_tasks(const PairRec *__restrict__ pairs, uint64_t np, const int4 *__restrict__ nodes, FwdTask *__restrict__ tasks) {
  const uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  const PairRec pr = pairs[p];
  int32_t best = 0, bl = -1;
  for (int32_t i = 0; i < pr.nnodes; i++) {
    const int32_t l = nodes[pr.hb + i].y;
    if (l > bl) { bl = l; best = i; }
  }
  tasks[p].pair = (uint32_t)p;
  tasks[p].node = (uint32_t)best;
}

// pairs whose probe ended at a branch point: count their remaining nodes
__global__ void k_spec_count(const PairRec *__restrict__ pairs, uint64_t np, const FwdTask *__restrict__ probe,
                             const uint8_t *__restrict__ cache, uint32_t *__restrict__ cnt) {
  const uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  const PairRec pr = pairs[p];
  const bool spec = !pr.consistent && pr.nnodes > 1 && cache[pr.hb + probe[p].node] == 1;
  cnt[p] = spec ? (uint32_t)(pr.nnodes - 1) : 0u;
}

__global__ void k_spec_write(const PairRec *__restrict__ pairs, uint64_t np, const FwdTask *__restrict__ probe,
                             const uint32_t *__restrict__ cnt, const uint64_t *__restrict__ off, FwdTask *__restrict__ tasks) {
  const uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np || cnt[p] == 0) return;
  const int32_t nn = pairs[p].nnodes;
  const uint32_t skip = probe[p].node;
  uint64_t w = off[p];
  for (int32_t i = 0; i < nn; i++) {
    if ((uint32_t)i == skip) continue;
    tasks[w].pair = (uint32_t)p;
    tasks[w].node = (uint32_t)i;
    w++;
  }
}
