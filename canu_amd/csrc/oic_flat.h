// Row-stepped packed forward DPs with flat rows (tuning.fwd_loop = 2).
//
// The same tasks, results and DP semantics as k_fwd_sm (oic_batch.h), which gives each of a
// wavefront's 64/G DPs a fixed group of G lanes: a row there costs as many band chunks as the
// widest band among the DPs needs, and on C3 half of the lane slots of those chunks are idle.  Here
// the G lanes of a group keep only the DP's own state (task, row bookkeeping, staging, row 0), and
// each row's cells of all the wavefront's DPs are spread over all 64 lanes:
//   * every active DP's band width w is summed over the wavefront (W), and the row takes the
//     fewest chunk iterations C such that giving each DP ceil(w / C) lanes needs at most 64;
//   * each DP gets a run of consecutive lanes; a lane takes diagonals Left + j, Left + j + n, ...
//     of its DP (j its place in the run, n the run's length), so its diagonals rise;
//   * per lane, the first and last kept diagonal, the reaching diagonal and the previous row's
//     first maximum (as in sm_row) are reduced per DP by LDS atomics (min, max, 64-bit max).
// A row therefore costs ceil(W / 64) chunk iterations, or one more, however unequal the bands.
// LDS decides how many wavefronts a CU holds, so each DP keeps one band buffer, updated in place:
// in a chunk every lane reads its three cells of the previous row before any lane writes, and the
// one cell a run's first lane needs from the chunk before (its left neighbour, already
// overwritten) comes from the lane that read it there, by a shuffle.  With 4 lanes per DP the
// first 128 bases of each tail are staged (SW = 8 words), else 1024: longer windows measured slower
// on both the repeat-rich and the low-error reads, their copy at every DP start costing more than
// the cached reads past them.

#pragma once

// one DP's current row, for the lanes that run its cells
struct FlatDesc {
  int4 row;             // this row: Left, Right, Edit_Match_Limit[e], lanes its cells are spread over
  int64_t w0a, w0b;     // first staged word of each tail in the packed reads (base 0 = 16 w0 + r0)
  int32_t m, n, r0a, r0b;
  int32_t nab;          // bit 0: A's read has N, bit 1: B's read has N
};

template <int G>
constexpr int flat_sw() { return G == 4 ? 8 : OIC_SEQ_WIN; }

// Each DP's band buffer holds its CAP ring cells, then copies of cells 0 and 1 (so the three cells
// d - 1, d, d + 1 are three consecutive words from (d - 1) mod CAP), then one spare word: strides of
// CAP + 3 and SW + 1 words, both odd, put the same cell of different DPs in different LDS banks.
template <int G, int CAP>
struct FlatLDS {
  static constexpr int NG = OIC_WAVE / G;
  static constexpr int SW = flat_sw<G>();
  // a DP's two staged tails side by side (A, then T; SW + 1 words each): one base register
  // addresses both within the 8-bit dword offsets of paired LDS reads
  uint32_t seq[NG][2][SW + 1];
  int32_t tile[NG][CAP + 3];
  FlatDesc desc[NG];
  unsigned long long pkey[NG];   // the previous row's first maximum: (Row + 1) << 32 | (2^30 - d)
  int4 red[NG];                  // first kept, last kept and first reaching diagonal, -
  int32_t start[OIC_WAVE];
};

// inclusive prefix sum over the wavefront's lanes: within each row of 16 lanes by row shifts
// (lanes whose source lies before the row start add 0), then each row's total carried into the
// rows after it by the row broadcasts of lane 15 (into rows 1 and 3) and lane 31 (into rows 2, 3)
__device__ __forceinline__ int32_t wave_incl_scan(int32_t v) {
  v += __builtin_amdgcn_update_dpp(0, v, 0x111, 0xF, 0xF, false);   // row_shr:1
  v += __builtin_amdgcn_update_dpp(0, v, 0x112, 0xF, 0xF, false);   // row_shr:2
  v += __builtin_amdgcn_update_dpp(0, v, 0x114, 0xF, 0xF, false);   // row_shr:4
  v += __builtin_amdgcn_update_dpp(0, v, 0x118, 0xF, 0xF, false);   // row_shr:8
  v += __builtin_amdgcn_update_dpp(0, v, 0x142, 0xA, 0xF, false);   // row_bcast:15
  v += __builtin_amdgcn_update_dpp(0, v, 0x143, 0xC, 0xF, false);   // row_bcast:31
  return v;
}

// Wavefronts per SIMD the flat kernel's LDS allows, at most 4: registers are held to that
// occupancy (128 VGPRs), which the cell loop's per-lane DP copy needs without spilling.
// (4 x 64: 10 KB of LDS per wavefront, 16 per CU.)
template <int G, int CAP>
constexpr int flat_waves_per_eu() {
  constexpr int w = (163840 / (int)sizeof(FlatLDS<G, CAP>) + 3) / 4;
  return w < 1 ? 1 : (w > 4 ? 4 : w);
}

// the descriptor fields that stay fixed for a DP's rows, written when the DP starts
template <int G, int CAP>
__device__ __forceinline__ void flat_desc_start(FlatLDS<G, CAP> &L, const SMDP &s, int g, int gl) {
  if (gl == 0) {
    FlatDesc &D = L.desc[g];
    D.w0a = s.ta.w0; D.w0b = s.tb.w0;
    D.m = s.m; D.n = s.n; D.r0a = s.ta.r0; D.r0b = s.tb.r0;
    D.nab = (s.nA ? 1 : 0) | (s.nB ? 2 : 0);
    L.tile[g][CAP] = L.tile[g][0];   // the copy of cell 0 (row 0)
  }
}

// One row of every active DP of the wavefront; called by all lanes.  rowg: this lane's group has
// a DP in its row phase.  Returns true (group-uniform) when the group's DP is over (res holds its
// outcome).  The group's DP state s is updated as sm_row updates its own.
template <int G, int CAP, bool STATS>
__device__ __forceinline__ bool flat_row(FlatLDS<G, CAP> &L, bool rowg, SMDP &s, SMRes &res,
                                         const uint32_t *__restrict__ words, const uint32_t *__restrict__ nmask,
                                         const DPTables &tb, int g, int gl, uint64_t (&lstat)[6]) {
  constexpr int NG = OIC_WAVE / G;
  constexpr int SW = flat_sw<G>();
  constexpr int32_t MASK = CAP - 1;
  constexpr int32_t INF = 0x7fffffff;
  const int lane = lane_id();
  bool over = false;
  int32_t e = 0, eml_row = 0;
  // group phase: the row's band, sentinels (prefixEditDistance-forward.C:140-153)
  if (rowg) {
    e = ++s.e;
    eml_row = tb.match_limit[e];
    s.Left = (s.Left - 1 > -e) ? s.Left - 1 : -e;
    s.Right = (s.Right + 1 < e) ? s.Right + 1 : e;
    if (STATS) res.rows++;
    if (s.Right - s.Left + 3 > CAP) {
      res.overflow = true;
      over = true;
    } else {
      int32_t *pbuf = L.tile[g];
      const int32_t sd = (gl + ((gl < 2) ? s.Left - 1 : s.Right - 2)) & MASK;
      if (gl < 4) {
        pbuf[sd] = -2;
        if (sd < 2) pbuf[sd + CAP] = -2;
      }
    }
  }
  if (lane < NG) {
    L.red[lane] = make_int4(INF, -INF, INF, 0);
    L.pkey[lane] = 0ull;
  }
  L.start[lane] = -1;
  // lanes per DP: few chunk iterations C with sum over DPs of ceil(w / C) <= 64.  With A active
  // DPs that sum is below W / C + A, so C = ceil(W / (64 - A)) always fits: try ceil(W / 64) and
  // the next value first (the fewest in about 95 % of rows of up to 16 random bands up to 125 wide).
  const int32_t w = (rowg && !over && gl == 0) ? s.Right - s.Left + 1 : 0;
  const int32_t W = __builtin_amdgcn_readlane(wave_incl_scan(w), OIC_WAVE - 1);
  const int32_t A = __builtin_popcountll(__ballot(w > 0));
  int32_t C = (W + OIC_WAVE - 1) / OIC_WAVE;
  int32_t nv = 0, incl = 0;
  if (W > 0) {
    const int32_t Csafe = (W + (OIC_WAVE - A) - 1) / (OIC_WAVE - A);
    for (int it = 0; it < 3; it++) {
      // ceil(w / C) for w, C <= 2^12: the float quotient's floor is floor(w / C) or one less
      // (exact quotients), so one correction step gives the ceiling
      const float rc = __builtin_amdgcn_rcpf((float)C);
      const int32_t q = (int32_t)((float)w * rc);
      nv = q + ((int32_t)__umul24((uint32_t)q, (uint32_t)C) < w ? 1 : 0);
      incl = wave_incl_scan(nv);
      if (__builtin_amdgcn_readlane(incl, OIC_WAVE - 1) <= OIC_WAVE) break;
      C = it == 0 ? C + 1 : Csafe;
    }
    if (gl == 0 && nv > 0) L.desc[g].row = make_int4(s.Left, s.Right, eml_row, nv);
  }
  const int32_t S = __builtin_amdgcn_readlane(incl, OIC_WAVE - 1);
  wave_fence();
  if (nv > 0 && incl <= OIC_WAVE) L.start[incl - nv] = g;
  wave_fence();
  if (STATS && lane == 0) lstat[5]++;    // wavefront rows
  if (W > 0) {
    // this lane's DP (the run that starts at or below it) and its place in the run
    const int32_t sv = L.start[lane];
    const uint64_t M = __ballot(sv >= 0);
    const bool used = lane < S && S <= OIC_WAVE;
    const int s0 = 63 - __builtin_clzll((M & (~0ull >> (63 - lane))) | 1ull);
    const int cg = __shfl(sv, s0);
    const int32_t j = lane - s0;
    const int dg = used ? cg : 0;
    const FlatDesc &D = L.desc[dg];
    SMDP c;
    c.ta.w0 = D.w0a;
    c.ta.r0 = D.r0a;
    c.ta.lds = (lds_words)L.seq[dg][0];
    c.tb.w0 = D.w0b;
    c.tb.r0 = D.r0b;
    c.tb.lds = (lds_words)L.seq[dg][1];
    c.a_b0 = D.w0a * 16 + D.r0a;
    c.b_b0 = D.w0b * 16 + D.r0b;
    c.nA = (D.nab & 1) != 0;
    c.nB = (D.nab & 2) != 0;
    c.m = D.m;
    c.n = D.n;
    const int4 rw = D.row;
    const int32_t Left = rw.x, Right = rw.y, eml = rw.z, nl = rw.w;
    int32_t *buf = L.tile[dg];
    const int last = s0 + nl - 1;   // the run's last lane
    int32_t kmin = INF, kmax = -INF, rmin = INF, pv = -1, pd = 0, p1 = -2;
    uint32_t lb = 0;   // bases compared: slide lengths plus one terminating compare per cell
    for (int32_t k = 0; k < C; k++) {
      const int32_t d = Left + j + nl * k;
      const bool act = used && d <= Right;
      if (STATS) {
        lstat[0] += used ? 1 : 0;                       // lane slots
        if (lane == 0) lstat[4]++;                      // wavefront chunk iterations
      }
      // the previous row's cell d - 1 as the run's last lane read it in the chunk before
      const int32_t left = __shfl(p1, used ? last : lane);
      if (act) {
        const int32_t *b3 = buf + ((d - 1) & MASK);
        const int32_t l0 = b3[0], l1 = b3[1], l2 = b3[2];   // all three loaded: no branch on the lane
        const int32_t x2 = (j == 0 && k > 0) ? left : l0;
        p1 = l1;
        const int32_t x3 = 1 + l2;
        pd = p1 > pv ? d : pd;
        pv = p1 > pv ? p1 : pv;
        int32_t Row = 1 + p1;
        Row = Row > x2 ? Row : x2;
        Row = Row > x3 ? Row : x3;
        // slide along diagonal d (prefixEditDistance-forward.C:155-167), in 32-base windows
        const int32_t r1 = c.m - Row, r2 = c.n - Row - d;
        const int32_t rem0 = r1 < r2 ? r1 : r2;
        const int32_t kk = sm_match32<SW>(c, words, nmask, Row, Row + d);
        int32_t adv = kk < rem0 ? kk : rem0;
        if (STATS) {
          lstat[2]++;
          if (lane == __builtin_ctzll(__ballot(1))) lstat[3]++;
        }
        int32_t rem = (kk == 32 && rem0 > 32) ? rem0 - 32 : 0;
        while (rem > 0) {
          if (STATS) {
            lstat[2]++;
            if (lane == __builtin_ctzll(__ballot(1))) lstat[3]++;
          }
          const int32_t k32 = sm_match32<SW>(c, words, nmask, Row + adv, Row + d + adv);
          if (k32 >= rem) { adv += rem; break; }
          adv += k32;
          if (k32 < 32) break;
          rem -= 32;
        }
        lb += (uint32_t)adv + 1u;
        const bool reach = adv == rem0;
        Row += adv;
        const int32_t pos = d & MASK;
        buf[pos] = Row;
        if (pos < 2) buf[pos + CAP] = Row;
        rmin = (reach & (rmin == INF)) ? d : rmin;
        const bool keep = Row + (d > 0 ? d : 0) >= eml;   // prefixEditDistance-forward.C:245-285
        kmin = (keep & (kmin == INF)) ? d : kmin;
        kmax = keep ? d : kmax;
      }
      // the next chunk's reads of the band follow this chunk's writes (one wavefront's LDS
      // operations complete in order; this keeps the compiler from reordering them)
      __atomic_signal_fence(__ATOMIC_SEQ_CST);
    }
    lstat[1] += lb;
    if (used) {
      if (kmin != INF) {
        atomicMin(&L.red[dg].x, kmin);
        atomicMax(&L.red[dg].y, kmax);
      }
      if (rmin != INF) atomicMin(&L.red[dg].z, rmin);
      if (pv >= 0)
        atomicMax(&L.pkey[dg], ((unsigned long long)(uint32_t)(pv + 1) << 32) | (uint32_t)(0x40000000 - pd));
    }
  }
  wave_fence();
  // group phase: the row's outcome (prefixEditDistance-forward.C:172-306)
  if (rowg && !over) {
    const int32_t Left = s.Left, Right = s.Right;
    const int4 kr = L.red[g];
    const int32_t km = kr.x, kx = kr.y, rm = kr.z;
    const unsigned long long pk = L.pkey[g];
    const int32_t pvv = (int32_t)(uint32_t)(pk >> 32) - 1, pdd = 0x40000000 - (int32_t)(uint32_t)pk;
    const int32_t *cbuf = L.tile[g];
    over = true;
    if (rm != INF) {
      // the end of A or T was reached at diagonal rm: branch-point test after the previous row's
      // Longest update
      res.cells += (uint32_t)(rm - Left + 1);
      const int32_t rrow = cbuf[rm & MASK];
      if (e >= 2) sm_longest(s, tb, pvv, pdd, e - 1);
      const double Score = (double)rrow * tb.bmv - (double)e;
      const int32_t Tail_Len = rrow - s.MSL;
      const double slope = (double)(s.MaxScore - Score) / (double)Tail_Len;
      const bool abort_ = ((uint32_t)e > tb.min_branch_end_dist / 2u) && ((uint32_t)Tail_Len >= tb.min_branch_end_dist) &&
                          (slope >= tb.slope_min);
      if (abort_) { res.mte = false; res.e = s.MSe; res.d = s.MSd; }
      else        { res.mte = true;  res.e = e;     res.d = rm; }
    } else {
      res.cells += (uint32_t)(Right - Left + 1);
      if (e >= 2) sm_longest(s, tb, pvv, pdd, e - 1);
      if (km == INF) {
        res.mte = false; res.e = s.MSe; res.d = s.MSd;
      } else {
        s.Left = km;
        s.Right = kx;
        if (e >= s.lim) {
          int32_t v, dv;
          sm_row_first_max<G, CAP>(cbuf, km, kx, gl, v, dv);
          sm_longest(s, tb, v, dv, e);
          res.mte = false; res.e = s.MSe; res.d = s.MSd;
        } else {
          over = false;
        }
      }
    }
  }
  return over;
}

template <int G, int CAP, bool STATS>
__global__ void __launch_bounds__(OIC_WAVE) __attribute__((amdgpu_waves_per_eu(flat_waves_per_eu<G, CAP>(), 8)))
k_fwd_flat(FwdArgs A) {
  __shared__ FlatLDS<G, CAP> L;
  const int lane = lane_id();
  const int g = lane / G, gl = lane % G;
  constexpr int SW = flat_sw<G>();
  uint64_t cells_acc = 0;
  uint32_t rows_acc = 0;
  uint64_t lstat[6] = {0, 0, 0, 0, 0, 0};   // lane slots, bases, slide windows, wave slide windows, wave chunks, wave rows
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
    if (!sm_start<G, SW>(s, res, A.R.words, A.R.nmask, L.tile[g], L.seq[g][0], L.seq[g][1], g, gl)) {
      flat_desc_start<G, CAP>(L, s, g, gl);
      return true;
    }
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
    }
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
      if (!sm_start<G, SW>(s, res, A.R.words, A.R.nmask, L.tile[g], L.seq[g][0], L.seq[g][1], g, gl)) {
        flat_desc_start<G, CAP>(L, s, g, gl);
        state = 1;
        break;
      }
      forward_done();
    }
    if (__ballot(state != 3) == 0) break;    // every group has run out of tasks
    const bool rowg = state == 1 || state == 2;
    if (flat_row<G, CAP, STATS>(L, rowg, s, res, A.R.words, A.R.nmask, A.tb, g, gl, lstat) && rowg) {
      if (state == 1) {
        forward_done();
      } else {
        lmte = res.mte;
        finish();
      }
    }
  }
  uint64_t rows64 = rows_acc, slots64 = lstat[0], bases64 = lstat[1], si64 = lstat[2], wsi64 = lstat[3],
           wrows64 = lstat[5], wch64 = lstat[4];
  for (int o = 32; o > 0; o >>= 1) {
    cells_acc += __shfl_xor(cells_acc, o);
    rows64 += __shfl_xor(rows64, o);
    slots64 += __shfl_xor(slots64, o);
    bases64 += __shfl_xor(bases64, o);
    si64 += __shfl_xor(si64, o);
    wsi64 += __shfl_xor(wsi64, o);
    wrows64 += __shfl_xor(wrows64, o);
    wch64 += __shfl_xor(wch64, o);
  }
  if (A.tstamp && lane == 0) atomicMax(&A.tstamp[1], (unsigned long long)wall_clock64());
  if (lane == 0 && A.stats) {
    atomicAdd(&A.stats[0], (unsigned long long)cells_acc);
    atomicAdd(&A.stats[1], (unsigned long long)rows64);
    atomicAdd(&A.stats[2], (unsigned long long)slots64);
    atomicAdd(&A.stats[3], (unsigned long long)wrows64);
    atomicAdd(&A.stats[4], (unsigned long long)bases64);
    atomicAdd(&A.stats[5], (unsigned long long)si64);
    atomicAdd(&A.stats[6], (unsigned long long)wsi64);
    atomicAdd(&A.stats[7], (unsigned long long)wch64);
  }
}
