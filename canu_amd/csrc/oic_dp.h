// oic_dp.h -- wavefront-cooperative banded prefix edit distance for gfx950.
//
// One 64-lane wavefront (one 64-thread workgroup) runs one DP.  Diagonals of a row are spread
// across lanes (lane l handles d = Left + l + 64c in chunk c).  The previous and current rows live
// in two LDS tiles of CAP ints indexed by d mod CAP; every post-prune row is also written to a
// per-wave global scratch so the traceback can walk it.  Bases are read as 2-bit packed words: a
// cell's slide compares 16 bases per step with XOR + count-trailing-zeros.
//
// Semantics follow prefixEditDistance::forward / ::reverse
// (src/overlapInCore/liboverlap/prefixEditDistance-forward.C:93-314, -reverse.C:113-330):
//   * row e widens the band by one on each side, reads row e-1 at d-1, d, d+1 with -2 sentinels,
//   * the first diagonal (lowest d) that reaches the end of A or T ends the call,
//     with the branch-point test in double precision (forward.C:172-212),
//   * after each row, diagonals are pruned from both ends while EAL[e][d] + max(d,0) <
//     Edit_Match_Limit[e] (prefixEditDistance.H:249-257), and the first-max Longest is tracked.
// A reverse() call is run as a rightward DP over the reverse-complement strands: comparing
// X[p-r] with Y[q-r] is comparing X_rc[|X|-1-p+r] with Y_rc[|Y|-1-q+r].
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define OIC_WAVE 64

struct DPTables {
  const int32_t *error_bound;    // [2^21]
  const int32_t *match_limit;    // [max_errors]
  double   bmv;                  // Branch_Match_Value = e/(1+e)
  double   slope_min;            // MIN_BRANCH_TAIL_SLOPE
  uint32_t min_branch_end_dist;  // MIN_BRANCH_END_DIST (uint32 in the reference)
  int32_t  partial;
};

// rightward sequence view: view base i is packed base (b0 + i)
struct SeqV {
  const uint32_t *w;     // 2-bit packed, base p at word p>>4, bits 2*(p&15)
  const uint32_t *nm;    // N mask with the same base indexing, or nullptr
  int64_t b0;
};

// per-wave global scratch for row storage and deltas
struct DPScratch {
  int32_t *rowL, *rowR;  // [max_rows]
  int64_t *rowOff;       // [max_rows]
  int32_t *rowVal;       // [val_cap]
  int64_t  val_cap;
  int32_t  max_rows;
  int32_t *dbuf;         // [max_rows + 2] raw traceback stack
  int32_t *ldelta;       // [max_rows + 2] Left_Delta
  int32_t *rdelta;       // [max_rows + 2] Right_Delta
};

struct DPOut {
  int32_t e, a_end, t_end, mte, leftover, ndelta;
  int32_t overflow;
};

__device__ __forceinline__ int lane_id() { return threadIdx.x & (OIC_WAVE - 1); }

__device__ __forceinline__ uint32_t win16(const uint32_t *__restrict__ w, int64_t p) {
  const int64_t wi = p >> 4;
  const uint32_t s = (uint32_t)(p & 15) * 2u;
  return __builtin_amdgcn_alignbit(w[wi + 1], w[wi], s);
}

// 16 N-mask bits starting at base p, spread to the even bit of each 2-bit base slot
__device__ __forceinline__ uint32_t nwin16(const uint32_t *__restrict__ nm, int64_t p) {
  const int64_t wi = p >> 5;
  const uint32_t s = (uint32_t)(p & 31);
  uint32_t v = __builtin_amdgcn_alignbit(nm[wi + 1], nm[wi], s) & 0xffffu;
  v = (v | (v << 8)) & 0x00ff00ffu;
  v = (v | (v << 4)) & 0x0f0f0f0fu;
  v = (v | (v << 2)) & 0x33333333u;
  v = (v | (v << 1)) & 0x55555555u;
  return v;
}

// mismatch bits (one per base, even positions) of the 16 bases at A[ia..], T[it..]
__device__ __forceinline__ uint32_t mism16(const SeqV &A, int64_t ia, const SeqV &T, int64_t it) {
  uint32_t x = win16(A.w, A.b0 + ia) ^ win16(T.w, T.b0 + it);
  x = (x | (x >> 1)) & 0x55555555u;
  if (A.nm) x &= ~nwin16(A.nm, A.b0 + ia);
  if (T.nm) x &= ~nwin16(T.nm, T.b0 + it);
  return x;
}

// equal bases (with 'n' matching anything) among the 64 from A[ia..], T[it..]: five words of each
// loaded together (each strand carries 4 guard words past its last base, so they exist)
__device__ __forceinline__ int32_t match64(const SeqV &A, int64_t ia, const SeqV &T, int64_t it) {
  const int64_t pa = A.b0 + ia, pt = T.b0 + it;
  const uint32_t *wa = A.w + (pa >> 4), *wt = T.w + (pt >> 4);
  const uint32_t sa = (uint32_t)(pa & 15) * 2u, st = (uint32_t)(pt & 15) * 2u;
  uint32_t a[5], b[5];
#pragma unroll
  for (int k = 0; k < 5; k++) {
    a[k] = wa[k];
    b[k] = wt[k];
  }
#pragma unroll
  for (int k = 0; k < 4; k++) {
    uint32_t x = __builtin_amdgcn_alignbit(a[k + 1], a[k], sa) ^ __builtin_amdgcn_alignbit(b[k + 1], b[k], st);
    x = (x | (x >> 1)) & 0x55555555u;
    if (A.nm) x &= ~nwin16(A.nm, pa + 16 * k);
    if (T.nm) x &= ~nwin16(T.nm, pt + 16 * k);
    if (x) return 16 * k + (int32_t)(__builtin_ctz(x) >> 1);
  }
  return 64;
}

// number of equal bases (with 'n' matching anything) from A[ia..], T[it..], at most rem: a 16-base
// window first (most slides end in it), then 64-base windows (long matches: one memory round trip
// per 64 bases)
__device__ __forceinline__ int32_t lane_slide(const SeqV &A, int64_t ia, const SeqV &T, int64_t it, int32_t rem) {
  if (rem <= 0) return 0;
  const uint32_t x = mism16(A, ia, T, it);
  const int32_t k = x ? (int32_t)(__builtin_ctz(x) >> 1) : 16;
  if (k >= rem) return rem;
  if (k < 16) return k;
  int32_t adv = 16;
  rem -= 16;
  while (rem > 0) {
    const int32_t k64 = match64(A, ia + adv, T, it + adv);
    if (k64 >= rem) return adv + rem;
    adv += k64;
    if (k64 < 64) return adv;
    rem -= 64;
  }
  return adv;
}

// wave-cooperative version for potentially long exact runs (row 0): 1024 bases per step
__device__ __forceinline__ int32_t wave_slide(const SeqV &A, int64_t ia, const SeqV &T, int64_t it, int32_t maxlen) {
  const int lane = lane_id();
  for (int32_t base = 0;; base += 16 * OIC_WAVE) {
    const int32_t off = base + 16 * lane;
    const int32_t rem = maxlen - off;
    int32_t k = 0;
    bool stop = true;
    if (rem > 0) {
      const uint32_t x = mism16(A, ia + off, T, it + off);
      const int32_t kk = x ? (int32_t)(__builtin_ctz(x) >> 1) : 16;
      if (kk >= rem) { k = rem; stop = true; }
      else           { k = kk;  stop = (kk < 16); }
    }
    const uint64_t b = __ballot(stop);
    if (b) {
      const int f = __builtin_ctzll(b);
      return base + 16 * f + __shfl(k, f);
    }
  }
}

__device__ __forceinline__ int32_t stored_cell(const DPScratch &sc, int32_t k, int32_t d) {
  const int32_t L = sc.rowL[k], R = sc.rowR[k];
  return (d >= L && d <= R) ? sc.rowVal[sc.rowOff[k] + (d - L)] : -2;
}

// Set_Right_Delta (forward.C:31-71) / Set_Left_Delta (reverse.C:35-93); lane 0 only.
// Returns the delta length; delta goes to sc.rdelta (forward) or sc.ldelta (reverse).
__device__ int32_t traceback(const DPScratch &sc, int32_t e, int32_t d, int32_t last, bool is_forward,
                             int32_t *leftover, int32_t *t_end, int32_t t_len) {
  int32_t len = 0;
  int32_t *buf = is_forward ? sc.dbuf : sc.ldelta;
  for (int32_t k = e; k > 0; k--) {
    int32_t from = d;
    int32_t mx = 1 + stored_cell(sc, k - 1, d);
    int32_t j;
    if ((j = stored_cell(sc, k - 1, d - 1)) > mx) { from = d - 1; mx = j; }
    if ((j = 1 + stored_cell(sc, k - 1, d + 1)) > mx) { from = d + 1; mx = j; }
    if (from == d - 1) {
      buf[len++] = mx - last - 1;
      d--;
      last = stored_cell(sc, k - 1, from);
    } else if (from == d + 1) {
      buf[len++] = last - (mx - 1);
      d++;
      last = stored_cell(sc, k - 1, from);
    }
  }
  if (is_forward) {
    buf[len++] = last + 1;
    int32_t o = 0;
    for (int32_t i = len - 1; i > 0; i--) {
      const int32_t v = buf[i], w = buf[i - 1];
      const int32_t sg = (w > 0) - (w < 0);
      sc.rdelta[o++] = (v < 0 ? -v : v) * sg;
    }
    return len - 1;
  }
  *leftover = last;
  if (len > 1 && buf[0] == 1 && *t_end + t_len > 0) {
    buf[0] = (buf[1] > 0) ? buf[1] + 1 : buf[1] - 1;
    for (int32_t i = 2; i < len; i++) buf[i - 1] = buf[i];
    len--;
    (*t_end)--;
    if (len == 0) (*leftover)++;
  }
  return len;
}

// Wave-uniform (value, first-d) maximum.
__device__ __forceinline__ void wave_argmax_first(int32_t &v, int32_t &d) {
#pragma unroll
  for (int o = 1; o < OIC_WAVE; o <<= 1) {
    const int32_t ov = __shfl_xor(v, o), od = __shfl_xor(d, o);
    if (ov > v || (ov == v && od < d)) { v = ov; d = od; }
  }
}

// Run one prefix DP.  All 64 lanes of the wave must call it with identical arguments.
//   is_forward  : forward() semantics (force-last-error fix-up, Right_Delta) vs reverse()
//   want_delta  : run Set_*_Delta when the call ends (rows are stored either way when set)
//   Returns e, A_End, T_End in the reference's convention (negated for reverse).
template <int CAP>
__device__ DPOut wave_prefix_dp(const SeqV A, const int32_t m, const SeqV T, const int32_t n, const int32_t limit,
                                const bool is_forward, const bool want_delta, const DPTables &tb,
                                int32_t *__restrict__ bufA, int32_t *__restrict__ bufB, const DPScratch &sc,
                                uint64_t &cells, uint64_t &slots, uint64_t &lane_bases) {
  constexpr int32_t MASK = CAP - 1;
  const int lane = lane_id();
  DPOut out;
  out.overflow = 0;
  out.ndelta = 0;
  out.leftover = 0;

  int32_t *pbuf = bufA, *cbuf = bufB;

  const int32_t row0 = wave_slide(A, 0, T, 0, m);
  cells += 1;
  slots += 1;
  if (lane == 0) lane_bases += (uint64_t)row0 + 1;
  __syncthreads();
  if (lane == 0) pbuf[0] = row0;
  if (want_delta && lane == 0) {
    sc.rowL[0] = 0; sc.rowR[0] = 0; sc.rowOff[0] = 0; sc.rowVal[0] = row0;
  }
  if (row0 == m) {
    out.e = 0;
    out.a_end = is_forward ? m : -m;
    out.t_end = is_forward ? m : -m;
    out.mte = 1;
    out.leftover = m;
    return out;
  }

  int32_t Left = 0, Right = 0;
  int32_t Longest = 0, Best_d = 0, Best_e = 0;
  int32_t MSL = 0, MSd = 0, MSe = 0;
  double MaxScore = 0.0;
  int64_t valOff = 1;

  int32_t tb_e = 0, tb_d = 0, tb_last = 0;   // traceback start
  int32_t ret_e = 0;
  bool ended = false;

  for (int32_t e = 1; e <= limit; e++) {
    Left = (Left - 1 > -e) ? Left - 1 : -e;
    Right = (Right + 1 < e) ? Right + 1 : e;
    if (Right - Left + 3 > CAP || (want_delta && e >= sc.max_rows)) {
      out.overflow = 1;
      return out;
    }
    __syncthreads();
    if (lane < 4) {
      const int32_t sd = (lane == 0) ? Left : (lane == 1) ? Left - 1 : (lane == 2) ? Right : Right + 1;
      pbuf[sd & MASK] = -2;
    }
    __syncthreads();

    int32_t rd = 0, rrow = 0;
    bool reached = false;
    for (int32_t c0 = Left; c0 <= Right; c0 += OIC_WAVE) {
      const int32_t d = c0 + lane;
      const bool act = d <= Right;
      int32_t Row = 0;
      if (act) {
        const int32_t a = 1 + pbuf[d & MASK];
        const int32_t b = pbuf[(d - 1) & MASK];
        const int32_t c = 1 + pbuf[(d + 1) & MASK];
        Row = a > b ? a : b;
        Row = Row > c ? Row : c;
        const int32_t r1 = m - Row, r2 = n - Row - d;
        const int32_t adv = lane_slide(A, Row, T, (int64_t)Row + d, r1 < r2 ? r1 : r2);
        Row += adv;
        lane_bases += (uint64_t)adv + 1;
        cbuf[d & MASK] = Row;
      }
      const bool reach = act && (Row == m || Row + d == n);
      const uint64_t bal = __ballot(reach);
      const int32_t nact = (Right - c0 + 1) < OIC_WAVE ? (Right - c0 + 1) : OIC_WAVE;
      slots += OIC_WAVE;
      if (bal) {
        const int f = __builtin_ctzll(bal);
        cells += (uint64_t)(f + 1);
        rd = c0 + f;
        rrow = __shfl(Row, f);
        reached = true;
        break;
      }
      cells += (uint64_t)nact;
    }

    if (reached) {
      // branch-point test (forward.C:172-212); -ffp-contract=off keeps it two roundings
      const double Score = (double)rrow * tb.bmv - (double)e;
      const int32_t Tail_Len = rrow - MSL;
      const double slope = (double)(MaxScore - Score) / (double)Tail_Len;
      bool abort_ = false;
      if (tb.partial && (Score < MaxScore)) abort_ = true;
      if (((uint32_t)e > tb.min_branch_end_dist / 2u) && ((uint32_t)Tail_Len >= tb.min_branch_end_dist) &&
          (slope >= tb.slope_min))
        abort_ = true;
      if (abort_) {
        ret_e = MSe;
        out.mte = 0;
        out.a_end = is_forward ? MSL : -MSL;
        out.t_end = is_forward ? MSL + MSd : -MSL - MSd;
        tb_e = MSe; tb_d = MSd;
        tb_last = -1;  // read from storage
      } else {
        int32_t d = rd;
        if (is_forward && rrow == m && (1 + pbuf[(d + 1) & MASK] == rrow) && d < Right) d++;
        ret_e = e;
        out.mte = 1;
        out.a_end = is_forward ? rrow : -rrow;
        out.t_end = is_forward ? rrow + d : -rrow - d;
        tb_e = e; tb_d = d; tb_last = rrow;
      }
      ended = true;
      break;
    }

    __syncthreads();
    // prune both ends (forward.C:244-282)
    const int32_t eml = tb.match_limit[e];
    int32_t nL = Right + 1;
    for (int32_t c0 = Left; c0 <= Right; c0 += OIC_WAVE) {
      const int32_t d = c0 + lane;
      const bool keep = (d <= Right) && (cbuf[d & MASK] + (d > 0 ? d : 0) >= eml);
      const uint64_t bal = __ballot(keep);
      if (bal) { nL = c0 + __builtin_ctzll(bal); break; }
    }
    if (nL > Right) break;
    int32_t nR = nL;
    for (int32_t c1 = Right; c1 >= nL; c1 -= OIC_WAVE) {
      const int32_t d = c1 - lane;
      const bool keep = (d >= nL) && (cbuf[d & MASK] + (d > 0 ? d : 0) >= eml);
      const uint64_t bal = __ballot(keep);
      if (bal) { nR = c1 - __builtin_ctzll(bal); break; }
    }
    Left = nL;
    Right = nR;

    // first-max Longest over the surviving band (forward.C:284-299)
    int32_t bv = -1, bd = 0x7fffffff;
    for (int32_t c0 = Left; c0 <= Right; c0 += OIC_WAVE) {
      const int32_t d = c0 + lane;
      if (d <= Right) {
        const int32_t v = cbuf[d & MASK];
        if (v > bv) { bv = v; bd = d; }
      }
    }
    wave_argmax_first(bv, bd);
    if (bv > Longest) { Longest = bv; Best_d = bd; Best_e = e; }
    const double Score = (double)Longest * tb.bmv - (double)e;
    if (Score > MaxScore) { MaxScore = Score; MSL = Longest; MSd = Best_d; MSe = Best_e; }

    if (want_delta) {
      const int32_t W = Right - Left + 1;
      if (valOff + W > sc.val_cap) { out.overflow = 1; return out; }
      if (lane == 0) { sc.rowL[e] = Left; sc.rowR[e] = Right; sc.rowOff[e] = valOff; }
      for (int32_t i = lane; i < W; i += OIC_WAVE) sc.rowVal[valOff + i] = cbuf[(Left + i) & MASK];
      valOff += W;
    }

    int32_t *t = pbuf; pbuf = cbuf; cbuf = t;
  }

  if (!ended) {   // error limit exhausted or band emptied (forward.C:306-313)
    ret_e = MSe;
    out.mte = 0;
    out.a_end = is_forward ? MSL : -MSL;
    out.t_end = is_forward ? MSL + MSd : -MSL - MSd;
    tb_e = MSe; tb_d = MSd; tb_last = -1;
  }
  out.e = ret_e;

  if (want_delta) {
    __threadfence_block();
    __syncthreads();
    int32_t nd = 0, lo = 0, te = out.t_end;
    if (lane == 0) {
      const int32_t last = (tb_last >= 0) ? tb_last : stored_cell(sc, tb_e, tb_d);
      nd = traceback(sc, tb_e, tb_d, last, is_forward, &lo, &te, n);
    }
    out.ndelta = __shfl(nd, 0);
    out.leftover = __shfl(lo, 0);
    out.t_end = __shfl(te, 0);
    __threadfence_block();
    __syncthreads();
  }
  return out;
}
