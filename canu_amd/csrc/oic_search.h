// oic_search.h -- seed table, seed-hit gathering, match lists and per-pair Process_Matches on gfx950.
//
// Layout in HBM
//   reads     : 2-bit packed words; read r's forward strand starts at base b0f[r], its reverse
//               complement at b0r[r] (both multiples of 32, two guard words after each strand);
//               an N mask (one bit per base, same base index) flags non-ACGT bases.
//   k-mer table (one hash block): every valid k-mer position of the block's reads as
//               (key, t_local << 32 | offset), stably radix-sorted by key, so each k-mer's
//               occurrences are in ascending (read, offset) order; unique keys plus a directory
//               over the key's top bits locate a k-mer's run with one read and a short search.
//   hits      : for a batch of ref reads, one (pair key, hit) per seed hit with S_ID < T_ID, emitted
//               in (S offset asc, T desc, T offset desc) order and stably sorted by pair key =
//               (S, orientation, T): each candidate pair's hits come out in the order Add_Ref sees
//               them in Find_Overlaps (overlapInCore-Find_Overlaps.C:251-332).
//   nodes     : per pair, Match_Node_t list in list order (Offset, Len, Start, removed).
#pragma once

#include "oic_dp.h"

#define OIC_MAX_DISTINCT 3       // MAX_DISTINCT_OLAPS (overlapInCore.H:107)
#define OIC_HOPELESS     90      // HOPELESS_MATCH (overlapInCore.H:78)
#define OIC_MIN_INTER    10      // MIN_INTERSECTION (overlapInCore.H:120)
#define OIC_SHIFT_SLACK  1       // SHIFT_SLACK (overlapInCore.H:146)

struct ReadsDev {
  const uint32_t *words;
  const uint32_t *nmask;
  const int64_t  *b0f, *b0r;
  const int32_t  *len;
  const uint8_t  *hasN;
};

__device__ __forceinline__ SeqV read_view(const ReadsDev &R, uint32_t r, int strand) {
  SeqV v;
  v.w = R.words;
  v.nm = R.hasN[r] ? R.nmask : nullptr;
  v.b0 = strand ? R.b0r[r] : R.b0f[r];
  return v;
}

__device__ __forceinline__ uint32_t base_code(uint8_t c, bool *isN) {
  switch (c) {
    case 'A': case 'a': *isN = false; return 0;
    case 'C': case 'c': *isN = false; return 1;
    case 'G': case 'g': *isN = false; return 2;
    case 'T': case 't': *isN = false; return 3;
    default: *isN = true; return 0;
  }
}

// one block per read: pack forward and reverse-complement strands, N masks
__global__ void __launch_bounds__(256) k_pack(const uint8_t *__restrict__ raw, const uint64_t *__restrict__ start,
                                             const int32_t *__restrict__ len, const int64_t *__restrict__ b0f,
                                             const int64_t *__restrict__ b0r, uint32_t *__restrict__ words,
                                             uint32_t *__restrict__ nmask, uint8_t *__restrict__ hasN, uint32_t nreads) {
  const uint32_t r = blockIdx.x + 1;
  if (r > nreads) return;
  const int32_t L = len[r];
  const uint8_t *s = raw + start[r];
  __shared__ int anyN;
  if (threadIdx.x == 0) anyN = 0;
  __syncthreads();
  const int32_t nw = (L + 31) / 32 + 1;   // 32-base groups incl. one guard group
  for (int32_t g = threadIdx.x; g < nw; g += blockDim.x) {
    uint32_t wf0 = 0, wf1 = 0, wr0 = 0, wr1 = 0, mf = 0, mr = 0;
    for (int j = 0; j < 32; j++) {
      const int32_t i = g * 32 + j;
      if (i < L) {
        bool n1;
        const uint32_t c = base_code(s[i], &n1);
        if (j < 16) wf0 |= c << (2 * j); else wf1 |= c << (2 * (j - 16));
        if (n1) mf |= 1u << j;
        bool n2;
        const uint32_t cr = base_code(s[L - 1 - i], &n2);
        const uint32_t cc = n2 ? 0u : (3u - cr);
        if (j < 16) wr0 |= cc << (2 * j); else wr1 |= cc << (2 * (j - 16));
        if (n2) mr |= 1u << j;
      }
    }
    const int64_t pf = b0f[r] + (int64_t)g * 32, pr = b0r[r] + (int64_t)g * 32;
    words[pf >> 4] = wf0; words[(pf >> 4) + 1] = wf1;
    words[pr >> 4] = wr0; words[(pr >> 4) + 1] = wr1;
    nmask[pf >> 5] = mf; nmask[pr >> 5] = mr;
    if (mf) atomicOr(&anyN, 1);
  }
  // guard word after the last group
  if (threadIdx.x == 0) {
    const int64_t pf = b0f[r] + (int64_t)nw * 32, pr = b0r[r] + (int64_t)nw * 32;
    words[pf >> 4] = 0; words[(pf >> 4) + 1] = 0; words[pr >> 4] = 0; words[(pr >> 4) + 1] = 0;
    nmask[pf >> 5] = 0; nmask[pr >> 5] = 0;
  }
  __syncthreads();
  if (threadIdx.x == 0) hasN[r] = (uint8_t)anyN;
}

// k-mer key of view bases [p, p+k): base j at bits 2j (k <= 31)
__device__ __forceinline__ uint64_t kmer_at(const uint32_t *__restrict__ w, int64_t p, int k) {
  const int64_t wi = p >> 4;
  const uint32_t s = (uint32_t)(p & 15) * 2u;
  const uint64_t lo = ((uint64_t)w[wi + 1] << 32) | w[wi];
  const uint64_t hi = w[wi + 2];
  uint64_t v = s ? ((lo >> s) | (hi << (64 - s))) : lo;
  return (k >= 32) ? v : (v & ((1ull << (2 * k)) - 1));
}

__device__ __forceinline__ bool kmer_has_n(const uint32_t *__restrict__ nm, int64_t p, int k) {
  const int64_t wi = p >> 5;
  const uint32_t s = (uint32_t)(p & 31);
  const uint64_t v = (((uint64_t)nm[wi + 1] << 32) | nm[wi]) >> s;   // >= 33 valid bits, k <= 31
  return (v & ((1ull << k) - 1)) != 0;
}

// emit (key, t_local << 32 | offset) for every window of every loaded hash read; invalid
// windows (a non-ACGT base) get a key above every valid one.
__global__ void __launch_bounds__(256) k_kmer_emit(ReadsDev R, uint32_t bgn, uint32_t nblock, const uint64_t *__restrict__ posbase,
                                                  const uint8_t *__restrict__ loaded, int k, uint64_t invalid_key,
                                                  uint64_t *__restrict__ keys, uint64_t *__restrict__ vals) {
  const uint32_t t = blockIdx.x;
  if (t >= nblock || !loaded[t]) return;
  const uint32_t r = bgn + t;
  const int32_t L = R.len[r];
  const int32_t npos = L - k + 1;
  const int64_t b0 = R.b0f[r];
  const bool hn = R.hasN[r];
  const uint64_t pb = posbase[t];
  for (int32_t i = threadIdx.x; i < npos; i += blockDim.x) {
    uint64_t key = kmer_at(R.words, b0 + i, k);
    if (hn && kmer_has_n(R.nmask, b0 + i, k)) key = invalid_key;
    keys[pb + i] = key;
    vals[pb + i] = ((uint64_t)t << 32) | (uint32_t)i;
  }
}

__global__ void k_unique_flags(const uint64_t *__restrict__ keys, uint64_t n, uint64_t invalid_key, uint32_t *__restrict__ flag) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint64_t k = keys[i];
  flag[i] = (k != invalid_key && (i == 0 || keys[i - 1] != k)) ? 1u : 0u;
}

__global__ void k_unique_write(const uint64_t *__restrict__ keys, const uint64_t *__restrict__ vals, uint64_t n,
                               const uint32_t *__restrict__ flag, const uint64_t *__restrict__ uidx, uint64_t *__restrict__ ukey,
                               uint64_t *__restrict__ ustart, uint32_t *__restrict__ first_t) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || !flag[i]) return;
  const uint64_t u = uidx[i];
  ukey[u] = keys[i];
  ustart[u] = i;
  if (first_t) atomicAdd(&first_t[(uint32_t)(vals[i] >> 32)], 1u);
}

// dir[b] = first unique index whose key >> shift >= b, for b in [0, 2^bits]
__global__ void k_build_dir(const uint64_t *__restrict__ ukey, uint64_t nu, int shift, uint64_t nb, uint64_t *__restrict__ dir) {
  const uint64_t u = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (u > nu) return;
  const int64_t bu = (u < nu) ? (int64_t)(ukey[u] >> shift) : (int64_t)nb;
  const int64_t bp = (u == 0) ? -1 : (int64_t)(ukey[u - 1] >> shift);
  for (int64_t b = bp + 1; b <= bu; b++) dir[b] = u;
}

struct KmerTable {
  const uint64_t *ukey;     // [nu] sorted unique keys
  const uint64_t *ustart;   // [nu + 1] run starts in occ (ustart[nu] = number of valid positions)
  const uint64_t *occ;      // [nvalid] t_local << 32 | offset, ascending within a key
  const uint64_t *dir;      // [2^dbits + 1]
  uint64_t nu;
  int shift;
};

// locate key; returns false if absent
__device__ __forceinline__ bool table_find(const KmerTable &X, uint64_t key, uint64_t *lo, uint64_t *hi) {
  const uint64_t b = key >> X.shift;
  uint64_t a = X.dir[b], z = X.dir[b + 1];
  while (a < z) {
    const uint64_t mid = (a + z) >> 1;
    if (X.ukey[mid] < key) a = mid + 1; else z = mid;
  }
  if (a >= X.nu || X.ukey[a] != key) return false;
  *lo = X.ustart[a];
  *hi = X.ustart[a + 1];
  return true;
}

// first index in occ[lo, hi) whose read (t_local) is > tl
__device__ __forceinline__ uint64_t occ_upper(const uint64_t *__restrict__ occ, uint64_t lo, uint64_t hi, uint32_t tl) {
  while (lo < hi) {
    const uint64_t mid = (lo + hi) >> 1;
    if ((uint32_t)(occ[mid] >> 32) <= tl) lo = mid + 1; else hi = mid;
  }
  return lo;
}

struct GatherArgs {
  ReadsDev R;
  KmerTable X;
  uint32_t s0, ns;          // ref batch: IDs s0 .. s0+ns-1
  uint32_t hbgn;            // Hash_String_Num_Offset
  int32_t  k;
  int32_t  min_olap;
  const uint8_t *searched;  // [ns] 1 if this ref read is searched
  int      tbits;           // bits for t_local in the pair key
  const uint64_t *skip;     // sorted keys of the -k skip k-mers, both strands (nskip = 0: no file)
  uint32_t nskip;
  uint32_t *sscreen;        // [2 ns] per (S, orientation): 1 left end screened, 2 right end screened
};

__device__ __forceinline__ bool skip_has(const uint64_t *__restrict__ skip, uint32_t n, uint64_t key) {
  uint32_t lo = 0, hi = n;
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (skip[mid] < key) lo = mid + 1; else hi = mid;
  }
  return lo < n && skip[lo] == key;
}

// Mark_Skip_Kmers' screened ends of the hash reads (overlapInCore-Build_Hash_Index.C:98-121,
// :132-177): every occurrence of a skip k-mer within HOPELESS_MATCH of a read's end flags that end.
__global__ void k_mark_screens(const uint64_t *__restrict__ skip, uint32_t nskip, KmerTable X, const int32_t *__restrict__ len,
                               uint32_t hbgn, int32_t k, uint32_t *__restrict__ tscreen) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nskip) return;
  uint64_t lo, hi;
  if (!table_find(X, skip[i], &lo, &hi)) return;
  for (uint64_t q = lo; q < hi; q++) {
    const uint64_t oc = X.occ[q];
    const uint32_t tl = (uint32_t)(oc >> 32);
    const int32_t off = (int32_t)(uint32_t)oc;
    uint32_t f = 0;
    if (off < OIC_HOPELESS) f |= 1u;
    if (len[hbgn + tl] - off - k + 1 < OIC_HOPELESS) f |= 2u;
    if (f) atomicOr(&tscreen[tl], f);
  }
}

// one block per (S, orientation): count hits with S_ID < T_ID for every window; sdsum gets the
// block's total
__global__ void __launch_bounds__(256) k_hit_count(GatherArgs A, uint32_t *__restrict__ cnt, const uint64_t *__restrict__ pbase,
                                                   unsigned long long *__restrict__ sdsum, uint64_t *__restrict__ phi,
                                                   uint64_t *__restrict__ pval) {
  __shared__ unsigned long long bsum;
  __shared__ uint32_t screen;
  if (threadIdx.x == 0) { bsum = 0; screen = 0; }
  __syncthreads();
  const uint32_t sd = blockIdx.x;
  const uint32_t si = sd >> 1, dir = sd & 1;
  if (si >= A.ns || !A.searched[si]) {
    if (threadIdx.x == 0) { sdsum[sd] = 0; A.sscreen[sd] = 0; }
    return;
  }
  const uint32_t S = A.s0 + si;
  const int32_t L = A.R.len[S];
  const int32_t npos = L - A.k + 1;
  const int64_t b0 = dir ? A.R.b0r[S] : A.R.b0f[S];
  const bool hn = A.R.hasN[S];
  const uint64_t pb = pbase[sd];
  const uint32_t tl_max = (S >= A.hbgn) ? (S - A.hbgn) : 0xffffffffu;   // T_local must exceed this
  unsigned long long mine = 0;
  for (int32_t o = threadIdx.x; o < npos; o += blockDim.x) {
    uint32_t c = 0;
    if (!(hn && kmer_has_n(A.R.nmask, b0 + o, A.k))) {
      const uint64_t key = kmer_at(A.R.words, b0 + o, A.k);
      uint64_t lo, hi;
      if (A.nskip && skip_has(A.skip, A.nskip, key)) {
        // an Empty hash entry: no hits, and Hash_Find's hi_hits screens the S ends (Find_Overlaps.C:266-310)
        uint32_t f = 0;
        if (o == 0) f = 1u;
        else {
          if (o < OIC_HOPELESS) f |= 1u;
          if (L - o - A.k + 1 < OIC_HOPELESS) f |= 2u;
        }
        if (f) atomicOr(&screen, f);
      } else if (table_find(A.X, key, &lo, &hi)) {
        const uint64_t f = (tl_max == 0xffffffffu) ? lo : occ_upper(A.X.occ, lo, hi, tl_max);
        c = (uint32_t)(hi - f);
        if (c) {
          const uint64_t mult = hi - lo;
          const uint64_t mcap = mult > 0x3fffffull ? 0x3fffffull : mult;
          phi[pb + o] = hi;
          pval[pb + o] = (uint64_t)o | (mcap << 21) | ((uint64_t)sd << 43);
        }
      }
    }
    cnt[pb + o] = c;
    mine += c;
  }
  atomicAdd(&bsum, mine);
  __syncthreads();
  if (threadIdx.x == 0) {
    sdsum[sd] = bsum;
    A.sscreen[sd] = screen;
  }
}

// Write the hits in (o asc, T desc, T offset desc) order at the scanned offsets, load-balanced:
// workgroup t writes output slots [t*EMIT_TILE, (t+1)*EMIT_TILE), so the 10^4 occurrences of a
// satellite k-mer spread over many workgroups and stores are coalesced.  The S windows feeding a tile are found by binary search over the exclusive
// scan of per-window counts; each slot's window is the last one whose first slot is <= the slot
// (a window with no hits shares its first slot with the next window, so the search lands on the
// window that holds the slot).  k_hit_count saved each window's occurrence bound and the packed
// (o, multiplicity, S/orientation), so no table lookup is repeated.
#define EMIT_TILE 4096
#define EMIT_PCAP 4096

__device__ __forceinline__ uint64_t last_le(const uint64_t *__restrict__ a, uint64_t lo, uint64_t hi, uint64_t s) {
  // largest i in [lo, hi] with a[i] <= s; a[lo] <= s is given
  while (lo < hi) {
    const uint64_t mid = lo + (hi - lo + 1) / 2;
    if (a[mid] <= s) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

__global__ void __launch_bounds__(256) k_hit_emit_tiles(const uint64_t *__restrict__ off, const uint64_t *__restrict__ phi,
                                                        const uint64_t *__restrict__ pval, uint64_t npos, uint64_t nhits,
                                                        const uint64_t *__restrict__ occ, int tbits, uint64_t *__restrict__ hkey,
                                                        uint64_t *__restrict__ hval) {
  __shared__ uint64_t s_off[EMIT_PCAP];
  __shared__ uint64_t s_p0, s_np;
  const uint64_t s0 = (uint64_t)blockIdx.x * EMIT_TILE;
  if (s0 >= nhits) return;
  const uint64_t s1 = (s0 + EMIT_TILE < nhits) ? s0 + EMIT_TILE : nhits;
  if (threadIdx.x == 0) {
    const uint64_t p0 = last_le(off, 0, npos - 1, s0);
    const uint64_t p1 = last_le(off, p0, npos - 1, s1 - 1);
    s_p0 = p0;
    s_np = p1 - p0 + 1;
  }
  __syncthreads();
  const uint64_t p0 = s_p0, np = s_np;
  const bool in_lds = np <= EMIT_PCAP;
  if (in_lds)
    for (uint64_t i = threadIdx.x; i < np; i += blockDim.x) s_off[i] = off[p0 + i];
  __syncthreads();
  for (uint64_t s = s0 + threadIdx.x; s < s1; s += blockDim.x) {
    uint64_t p, first;
    if (in_lds) {
      uint64_t lo = 0, hi = np - 1;
      while (lo < hi) {
        const uint64_t mid = lo + (hi - lo + 1) / 2;
        if (s_off[mid] <= s) lo = mid;
        else hi = mid - 1;
      }
      p = p0 + lo;
      first = s_off[lo];
    } else {
      p = last_le(off, p0, p0 + np - 1, s);
      first = off[p];
    }
    const uint64_t q = phi[p] - 1 - (s - first);
    const uint64_t oc = occ[q];
    const uint64_t v = pval[p];
    hkey[s] = ((v >> 43) << tbits) | (oc >> 32);
    hval[s] = (v & 0x1fffffull) | ((oc & 0xffffffffull) << 21) | (((v >> 21) & 0x3fffffull) << 42);
  }
}

__global__ void k_pair_flags(const uint64_t *__restrict__ hkey, uint64_t n, uint32_t *__restrict__ flag) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  flag[i] = (i == 0 || hkey[i] != hkey[i - 1]) ? 1u : 0u;
}

struct PairRec {
  uint64_t hb, he;          // hit / node range
  uint32_t s_id, t_id;
  int32_t  dir;
  int32_t  nnodes;
  int32_t  consistent;
  int32_t  diag_bgn, diag_end;   // min / max S offset of the pair's hits
  int32_t  max_mult;
  int32_t  longest;              // longest node Len
  int32_t  dmin, dmax;           // min / max node diagonal
  int32_t  first_longest;        // list position of the first node of Len == longest
};

__global__ void k_pair_write(const uint64_t *__restrict__ hkey, uint64_t n, const uint32_t *__restrict__ flag,
                             const uint64_t *__restrict__ pidx, PairRec *__restrict__ pairs, uint32_t s0, uint32_t hbgn, int tbits) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || !flag[i]) return;
  const uint64_t p = pidx[i];
  const uint64_t k = hkey[i];
  PairRec pr = {};
  pr.hb = i;
  pr.t_id = (uint32_t)(k & ((1ull << tbits) - 1)) + hbgn;
  const uint64_t sd = k >> tbits;
  pr.s_id = s0 + (uint32_t)(sd >> 1);
  pr.dir = (int32_t)(sd & 1);
  pairs[p] = pr;
}

__global__ void k_pair_ends(PairRec *__restrict__ pairs, uint64_t np, uint64_t nhits) {
  const uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  pairs[p].he = (p + 1 < np) ? pairs[p + 1].hb : nhits;
}

// Add_Ref + Add_Match (overlapInCore-Find_Overlaps.C:24-96, :104-163), one thread per candidate
// pair, in O(1) per hit.
//
// Add_Match walks the match list from its head.  With hits in (S offset asc, T offset desc) order,
// the list at S offset o is [nodes touched at o] ++ [nodes touched at o-1, in list order] ++ [dead
// nodes]: a node touched at o-1 expects offset o, anything older ends the walk.  So a hit at o
// extends the node of its diagonal among those touched at o-1 (at most one), which moves to the
// front iff an untouched node of o-1 precedes it; otherwise a new node goes to the front, and
// the pair stops being consistent if any live node was passed, or (none live) if the list's head
// is more than 3 diagonals away or does not end near o.  After offset o the list is
//   reverse(created/moved at o) ++ (o-1 nodes extended in place) ++ (o-1 nodes not hit) ++ older dead,
// and the hits of o, taken in order, are also the new live nodes sorted by diagonal (descending),
// which lets the next offset find diagonals by a merge instead of a walk.
//
// Scratch, all indexed from the pair's first hit: tmp (node records), and the int arrays
// sA/sB (live nodes by diagonal), oA/oB (live nodes in list order), fr (created / moved at o),
// dq (dead nodes, blocks in order of death), db (block starts), pos (list position of a live node),
// fl (per live position: 1 touched, 2 moved).
struct ListScratch {
  int32_t *sA, *sB, *oA, *oB, *fr, *dq, *db, *pos;
  uint8_t *fl;
};

// Pairs with done[p] == want: 0 for pairs never given to a wavefront, 2 for pairs a wavefront handed back.
__global__ void k_build_lists(PairRec *__restrict__ pairs, uint64_t np, const uint64_t *__restrict__ hval, int4 *__restrict__ tmp,
                              int4 *__restrict__ nodes, int32_t K, ListScratch X, const uint8_t *__restrict__ done, uint8_t want,
                              unsigned long long *__restrict__ nnodes_total) {
  const uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np || done[p] != want) return;
  PairRec pr = pairs[p];
  const uint64_t hb = pr.hb, he = pr.he;
  int4 *N = tmp + hb;                     // node records: x Offset, y Len, z Start
  int32_t *sCur = X.sA + hb, *sNew = X.sB + hb, *oCur = X.oA + hb, *oNew = X.oB + hb;
  int32_t *fr = X.fr + hb, *dq = X.dq + hb, *db = X.db + hb, *pos = X.pos + hb;
  uint8_t *fl = X.fl + hb;
  int32_t nA = 0, nnode = 0, ndq = 0, ndb = 0;
  int32_t consistent = 1, maxm = 0;
  int32_t prev_o = -2;
  uint64_t h = hb;
  while (h < he) {
    const int32_t o = (int32_t)(hval[h] & 0x1fffff);
    if (o != prev_o + 1 && nA > 0) {       // a gap: every live node is dead now
      db[ndb++] = ndq;
      for (int32_t i = 0; i < nA; i++) dq[ndq++] = oCur[i];
      nA = 0;
    }
    for (int32_t i = 0; i < nA; i++) fl[i] = 0;
    int32_t f = 0, nfr = 0, nmoved = 0, nnew = 0, q = 0;
    for (; h < he && (int32_t)(hval[h] & 0x1fffff) == o; h++) {
      const uint64_t v = hval[h];
      const int32_t toff = (int32_t)((v >> 21) & 0x1fffff);
      const int32_t mult = (int32_t)(v >> 42);
      if (mult > maxm) maxm = mult;
      const int32_t g = toff - o;
      while (q < nA && (N[sCur[q]].x - N[sCur[q]].z) > g) q++;
      if (q < nA && (N[sCur[q]].x - N[sCur[q]].z) == g) {
        const int32_t x = sCur[q], i = pos[x];
        N[x].y += 1;
        fl[i] = 1;
        if (f < i) { fl[i] = 2; fr[nfr++] = x; nmoved++; }
        while (f < nA && fl[f]) f++;
        sNew[nnew++] = x;
        q++;
      } else {
        if (nnode > 0 && consistent) {
          const int32_t live = nfr + nA - nmoved;
          if (live > 0) consistent = 0;
          else {
            const int4 hd = N[dq[db[ndb - 1]]];          // head: first node of the latest dead block
            const int32_t es = hd.z + hd.y - K + 1, dg = hd.x - hd.z;
            if (abs(dg - g) > 3 || o < es + K - 2) consistent = 0;
          }
        }
        N[nnode] = make_int4(toff, K, o, 0);
        fr[nfr++] = nnode;
        sNew[nnew++] = nnode;
        nnode++;
      }
    }
    // new list order of live nodes; the o-1 nodes not hit are dead (block prepended to the dead list)
    int32_t n2 = 0;
    for (int32_t i = nfr - 1; i >= 0; i--) { oNew[n2] = fr[i]; pos[fr[i]] = n2; n2++; }
    bool dead_block = false;
    for (int32_t i = 0; i < nA; i++) {
      if (fl[i] == 1) { oNew[n2] = oCur[i]; pos[oCur[i]] = n2; n2++; }
      else if (fl[i] == 0) {
        if (!dead_block) { db[ndb++] = ndq; dead_block = true; }
        dq[ndq++] = oCur[i];
      }
    }
    nA = n2;
    int32_t *t = sCur; sCur = sNew; sNew = t;
    t = oCur; oCur = oNew; oNew = t;
    prev_o = o;
  }
  // list: live nodes, then dead blocks newest first
  int32_t n = 0, longest = 0, first_longest = 0, dmin = 0x7fffffff, dmax = -0x7fffffff;
  auto emit = [&](int32_t x) {
    const int4 nd = N[x];
    nodes[hb + n] = make_int4(nd.x, nd.y, nd.z, 0);
    if (nd.y > longest) { longest = nd.y; first_longest = n; }
    n++;
    const int32_t dg = nd.x - nd.z;
    if (dg < dmin) dmin = dg;
    if (dg > dmax) dmax = dg;
  };
  for (int32_t i = 0; i < nA; i++) emit(oCur[i]);
  for (int32_t b = ndb - 1; b >= 0; b--) {
    const int32_t e = (b + 1 < ndb) ? db[b + 1] : ndq;
    for (int32_t i = db[b]; i < e; i++) emit(dq[i]);
  }
  pr.nnodes = n;
  atomicAdd(nnodes_total, (unsigned long long)n);
  pr.consistent = consistent;
  pr.diag_bgn = (int32_t)(hval[hb] & 0x1fffff);
  pr.diag_end = (int32_t)(hval[he - 1] & 0x1fffff);
  pr.max_mult = maxm;
  pr.longest = longest;
  pr.first_longest = first_longest;
  pr.dmin = dmin;
  pr.dmax = dmax;
  pairs[p] = pr;
}

// The same list, built by one wavefront per pair for pairs with many hits (satellite and repeat
// pairs reach millions).  Offsets are still taken in order, but the hits of one offset are handled
// by all lanes at once, with the live nodes in LDS:
//   - a hit's node is found by binary search over the live diagonals (descending);
//   - with t(k) = index of the hit touching the node at list position k (INF if none), the node
//     of hit j moves to the front iff max(t(0..k-1)) > j, i.e. some node before it is not yet
//     touched when hit j is taken;
//   - the new list is reverse(created/moved hits) ++ (nodes extended in place) ++ (untouched
//     nodes, now dead), positions found by ballot ranks;
//   - a created node makes the pair inconsistent unless it is the offset's first hit with no
//     live node, in which case the head-of-list test of Add_Match applies to the first node of the
//     latest dead block.
// Dead nodes go to tmp in blocks as they die, each record tagged so that its final position
// (live nodes, then blocks newest first, each in list order) is nlive + ndead - tag.  LDS holds up
// to CAP hits of one offset: a pair with a more crowded offset goes on to the next level's list
// (the 512-hit kernel), or, from the last level, back to the one-thread builder (done[p] = 2).
#define LW_CAP 512        // hits per offset of the last wavefront level
#define LW_CAP_FIRST 128  // of the first: a quarter of the LDS, four times the wavefronts per CU

__device__ __forceinline__ int32_t wave_incl_max_i32(int32_t v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int32_t t = __shfl_up(v, o);
    if (lane >= o) v = max(v, t);
  }
  return v;
}

__device__ __forceinline__ int32_t wave_max_i32(int32_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o));
  return v;
}

__device__ __forceinline__ int32_t wave_min_i32(int32_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o));
  return v;
}

struct BigListArgs {
  PairRec *pairs;
  const uint32_t *big;      // indices of pairs with many hits
  const uint32_t *nbig;
  const uint32_t *bcnt;     // if not null: big holds LW_BUCKETS lists of bstride entries, bcnt[b]
  uint64_t bstride;         //   of them used, taken in bucket order (largest pairs first)
  uint32_t *next;           // work counter, starts at 0
  uint32_t *handed_back;    // pairs left to the one-thread builder
  uint32_t *up;             // if not null: pairs too crowded for this level go here, for the next one
  uint32_t *nup;
  uint8_t *done;            // per pair: 0 = one-thread builder, 1 = a wavefront builds it, 2 = handed back
  unsigned long long *nnodes_total;
  const uint64_t *hval;
  int4 *tmp;
  int4 *nodes;
  int32_t K;
};

__global__ void k_fill_u8(uint8_t *__restrict__ p, uint64_t n) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

// Pairs with an S offset of more than LW_CAP_FIRST hits (crowded[p] = 1).  Within a pair, hits are in
// S-offset order, so an offset holds more than LW_CAP_FIRST of them exactly when some hit i and hit
// i + LW_CAP_FIRST have the same pair key and the same offset.  keys are the pair-grouped hit keys,
// flag marks each pair's first hit and idx is the exclusive scan of flag.
__global__ void k_crowded_pairs(const uint64_t *__restrict__ keys, const uint64_t *__restrict__ hval,
                                const uint32_t *__restrict__ flag, const uint64_t *__restrict__ idx, uint64_t n,
                                uint8_t *__restrict__ crowded) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i + LW_CAP_FIRST >= n) return;
  const uint64_t v = hval[i];
  if ((v >> 42) <= LW_CAP_FIRST) return;   // a k-mer with no more occurrences cannot crowd an offset
  const uint64_t j = i + LW_CAP_FIRST;
  if (keys[j] == keys[i] && (hval[j] & 0x1fffff) == (v & 0x1fffff)) crowded[idx[i] + flag[i] - 1] = 1;
}

// Pairs for the wavefront builder, bucketed by hit count (>= 2^17, 2^14, 2^11, the rest) so that
// the largest, whose offsets run one after another, start first and do not end the launch late.
// Crowded pairs go to the 512-hit level's list (up) directly.
#define LW_BUCKETS 4
__global__ void k_big_select(const PairRec *__restrict__ pairs, uint64_t np, uint64_t min_hits, uint8_t *__restrict__ done,
                             uint32_t *__restrict__ big, uint32_t *__restrict__ nbig, uint32_t *__restrict__ bcnt,
                             const uint8_t *__restrict__ crowded, uint32_t *__restrict__ up, uint32_t *__restrict__ nup) {
  const uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  const uint64_t nh = pairs[p].he - pairs[p].hb;
  const bool b = nh >= min_hits;
  done[p] = b ? 1 : 0;
  if (!b) return;
  if (crowded[p]) {
    up[atomicAdd(nup, 1u)] = (uint32_t)p;
    return;
  }
  const int k = nh >= (1u << 17) ? 0 : nh >= (1u << 14) ? 1 : nh >= (1u << 11) ? 2 : 3;
  big[(uint64_t)k * np + atomicAdd(&bcnt[k], 1u)] = (uint32_t)p;
  atomicAdd(nbig, 1u);
}

template <int CAP>
__global__ void __launch_bounds__(64) k_build_lists_wave(BigListArgs A) {
  __shared__ int32_t sLd[2][CAP];   // live diagonals in hit order of the last offset (descending)
  __shared__ int4 sR[2][CAP];       // live nodes: Offset, Len, Start, list position
  __shared__ int32_t sLq[2][CAP];   // list position -> live index
  __shared__ int32_t sT[CAP];       // per list position: the hit touching it, or INF
  __shared__ int32_t sM[CAP];       // exclusive prefix max of sT
  __shared__ int32_t sQ[CAP > LW_CAP_FIRST ? CAP : 1];    // per hit: live index of its node, or -1 (512-hit level)
  __shared__ int32_t sRk[CAP > LW_CAP_FIRST ? CAP : 1];   // per hit: 2 x rank among created / moved hits, +1 if it is one
  __shared__ int32_t sIp[CAP];      // per list position: rank among nodes extended in place
  __shared__ uint64_t sW[128];      // hit values [wb, wb + 128) of the pair, refilled every 64 hits
  __shared__ uint32_t sWork;
  const int lane = threadIdx.x;
  const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  const int32_t INF = 0x7fffffff;
  const int32_t K = A.K;
  const uint32_t nbig = *A.nbig;
  for (;;) {
    if (lane == 0) sWork = atomicAdd(A.next, 1u);
    __syncthreads();
    const uint32_t w = sWork;
    __syncthreads();
    if (w >= nbig) break;
    uint32_t p;
    if (A.bcnt) {
      uint32_t r = w;
      int k = 0;
      while (k < LW_BUCKETS - 1 && r >= A.bcnt[k]) r -= A.bcnt[k++];
      p = A.big[(uint64_t)k * A.bstride + r];
    } else {
      p = A.big[w];
    }
    PairRec pr = A.pairs[p];
    const uint64_t hb = pr.hb, he = pr.he;
    int4 *dq = A.tmp + hb;
    int cur = 0;
    int32_t nA = 0, nnode = 0, ndq = 0, maxm = 0, prev_o = -2;
    bool consistent = true, over = false;
    int4 head = make_int4(0, 0, 0, 0);
    uint64_t h = hb, wb = hb;
    auto fill = [&]() {   // the window from wb (hits past the pair's end read as 0, never used)
      for (int i = lane; i < 128; i += 64) sW[i] = (wb + (uint64_t)i < he) ? A.hval[wb + i] : 0ull;
      __syncthreads();
    };
    // hit value hh of the pair: from the window when it holds it
    auto hit = [&](uint64_t hh) -> uint64_t { return (hh - wb < 128) ? sW[hh - wb] : A.hval[hh]; };
    fill();
    while (h < he) {
      if (h - wb >= 64) {   // keep [h, h + 64) in the window
        __syncthreads();
        wb = h;
        fill();
      }
      if (nA == 1) {
        // The next offsets, each holding one hit on the live node's diagonal, only extend that node
        // (the one-hit step below, with nothing dying and nothing created): take up to 64 at once.
        // Hit h + lane and h + lane + 1 lie in the window ([h, h + 64) is held, the window is 128).
        const uint64_t hh = h + (uint64_t)lane;
        bool ext = false;
        uint64_t v = 0;
        if (hh < he) {
          v = sW[hh - wb];
          const int32_t oi = (int32_t)(v & 0x1fffff);
          const int32_t gi = (int32_t)((v >> 21) & 0x1fffff) - oi;
          const bool single = hh + 1 >= he || (int32_t)(sW[hh + 1 - wb] & 0x1fffff) != oi;
          ext = oi == prev_o + 1 + lane && gi == sLd[cur][0] && single;
        }
        const uint64_t b = __ballot(ext);
        const int32_t r = (b == ~0ull) ? 64 : __builtin_ctzll(~b);
        if (r > 0) {
          if (lane < r) maxm = max(maxm, (int32_t)(v >> 42));
          if (lane == 0) sR[cur][0].y += r;
          __syncthreads();
          h += (uint64_t)r;
          prev_o += r;
          continue;
        }
      }
      const int32_t o = (int32_t)(sW[h - wb] & 0x1fffff);
      int32_t c = 0;
      for (;;) {
        const uint64_t hh = h + (uint64_t)c + (uint64_t)lane;
        const bool same = hh < he && (int32_t)(hit(hh) & 0x1fffff) == o;
        const uint64_t b = __ballot(same);
        const int32_t run = (b == ~0ull) ? 64 : __builtin_ctzll(~b);
        c += run;
        if (run < 64 || c > CAP) break;
      }
      if (c > CAP) { over = true; break; }
      int4 *R = sR[cur], *R2 = sR[cur ^ 1];
      int32_t *Ld = sLd[cur], *Ld2 = sLd[cur ^ 1], *Lq = sLq[cur], *Lq2 = sLq[cur ^ 1];
      if (o != prev_o + 1 && nA > 0) {               // a gap: every live node dies, in list order
        for (int32_t k = lane; k < nA; k += 64) {
          int4 r = R[Lq[k]];
          r.w = ndq + nA - k;
          dq[ndq + k] = r;
        }
        head = R[Lq[0]];
        ndq += nA;
        nA = 0;
      }
      if (c == 1) {
        // One hit (most offsets of plain overlaps): its node -- the live one on its diagonal, extended,
        // or a new one -- is the only node left live, at list position 0 (it moves to the front when
        // untouched nodes precede it), and every other live node dies, in list order.
        const uint64_t v = hit(h);
        const int32_t g = (int32_t)((v >> 21) & 0x1fffff) - o;
        maxm = max(maxm, (int32_t)(v >> 42));
        int32_t q = -1;
        for (int32_t k0 = 0; k0 < nA; k0 += 64) {
          const uint64_t bm = __ballot(k0 + lane < nA && Ld[k0 + lane] == g);
          if (bm) { q = k0 + __builtin_ctzll(bm); break; }
        }
        const int32_t kp = (q >= 0) ? R[q].w : -1;   // the hit node's list position
        const int32_t sz = (q >= 0) ? nA - 1 : nA;   // nodes that die
        for (int32_t k = lane; k < nA; k += 64) {
          if (k == kp) continue;
          const int32_t r = (kp >= 0 && k > kp) ? k - 1 : k;
          int4 rec = R[Lq[k]];
          rec.w = ndq + sz - r;
          dq[ndq + r] = rec;
        }
        int4 nrec;
        if (q >= 0) {
          nrec = R[q];
          nrec.y += 1;
        } else {
          nrec = make_int4(g + o, K, o, 0);
          bool bad = nA > 0;   // a created node makes the pair inconsistent unless no node is live
          if (nA == 0 && nnode > 0) {
            const int32_t es = head.z + head.y - K + 1, dg = head.x - head.z;
            if (abs(dg - g) > 3 || o < es + K - 2) bad = true;
          }
          if (bad) consistent = false;
          nnode += 1;
        }
        nrec.w = 0;
        ndq += sz;
        if (lane == 0) {
          R2[0] = nrec;
          Ld2[0] = g;
          Lq2[0] = 0;
        }
        __syncthreads();
        nA = 1;
        prev_o = o;
        h += 1;
        cur ^= 1;
        continue;
      }
      int32_t ncre = 0;
      if constexpr (CAP <= LW_CAP_FIRST) {
        // hits -> live nodes.  Each lane keeps, for the hits it holds (j = lane + 64 t), the live
        // index of the hit's node (-1: none), that node's list position, the hit's diagonal and
        // its rank among created / moved hits, in registers.
        constexpr int NJ = CAP / 64;
        int32_t qr[NJ], wr[NJ], gr[NJ], rkr[NJ];
  #pragma unroll
        for (int t = 0; t < NJ; t++) {
          const int32_t j = lane + 64 * t;
          qr[t] = -1; wr[t] = -1; gr[t] = 0; rkr[t] = 0;
          if (64 * t < c && j < c) {   // the first test is uniform: rounds past c are skipped whole
            const uint64_t v = hit(h + j);
            const int32_t g = (int32_t)((v >> 21) & 0x1fffff) - o;
            maxm = max(maxm, (int32_t)(v >> 42));
            int32_t lo = 0, hi = nA;
            while (lo < hi) {
              const int32_t mid = (lo + hi) >> 1;
              if (Ld[mid] > g) lo = mid + 1;
              else hi = mid;
            }
            qr[t] = (lo < nA && Ld[lo] == g) ? lo : -1;
            gr[t] = g;
            Ld2[j] = g;
          }
        }
        for (int32_t k = lane; k < nA; k += 64) sT[k] = INF;
        __syncthreads();
  #pragma unroll
        for (int t = 0; t < NJ; t++) {
          const int32_t j = lane + 64 * t;
          if (64 * t < c && j < c && qr[t] >= 0) {
            wr[t] = R[qr[t]].w;
            sT[wr[t]] = j;
          }
        }
        __syncthreads();
        {
          int32_t carry = -1;
          for (int32_t k0 = 0; k0 < nA; k0 += 64) {
            const int32_t k = k0 + lane;
            const int32_t t = (k < nA) ? sT[k] : -1;
            const int32_t inc = wave_incl_max_i32(t);
            const int32_t up = __shfl_up(inc, 1);
            if (k < nA) sM[k] = (lane == 0) ? carry : max(up, carry);
            carry = max(carry, __shfl(inc, 63));
          }
        }
        __syncthreads();
        // created / moved hits
        int32_t nfr = 0;
        bool bad = false;
  #pragma unroll
        for (int t = 0; t < NJ; t++) {
          if (64 * t < c) {   // uniform
            const int32_t j = lane + 64 * t;
            bool cre = false, cf = false;
            if (j < c) {
              cre = qr[t] < 0;
              cf = cre || sM[wr[t]] > j;
            }
            const uint64_t bcf = __ballot(cf), bcre = __ballot(cre);
            rkr[t] = (nfr + (int32_t)__popcll(bcf & lt)) * 2 + (cf ? 1 : 0);
            if (cre && (j > 0 || nA > 0)) bad = true;
            nfr += (int32_t)__popcll(bcf);
            ncre += (int32_t)__popcll(bcre);
          }
        }
        if (nA == 0 && nnode > 0) {                    // first hit makes a node with no live node
          const int32_t es = head.z + head.y - K + 1, dg = head.x - head.z;
          if (abs(dg - Ld2[0]) > 3 || o < es + K - 2) bad = true;
        }
        if (__ballot(bad)) consistent = false;
        // list positions: extended in place, dead
        const int32_t sz = nA - (c - ncre);
        {
          int32_t ni = 0, nd = 0;
          for (int32_t k0 = 0; k0 < nA; k0 += 64) {
            const int32_t k = k0 + lane;
            bool inpl = false, dead = false;
            if (k < nA) {
              const int32_t t = sT[k];
              dead = t == INF;
              inpl = !dead && sM[k] < t;
            }
            const uint64_t bi = __ballot(inpl), bd = __ballot(dead);
            if (k < nA) sIp[k] = ni + (int32_t)__popcll(bi & lt);
            if (dead) {
              const int32_t r = nd + (int32_t)__popcll(bd & lt);
              int4 rec = R[Lq[k]];
              rec.w = ndq + sz - r;
              dq[ndq + r] = rec;
            }
            ni += (int32_t)__popcll(bi);
            nd += (int32_t)__popcll(bd);
          }
          ndq += sz;
        }
        __syncthreads();
        // the new live set, in hit order, with list positions
  #pragma unroll
        for (int t = 0; t < NJ; t++) {
          const int32_t j = lane + 64 * t;
          if (64 * t < c && j < c) {
            const int32_t q = qr[t], rk = rkr[t];
            int4 rec;
            int32_t P;
            if (q < 0) {
              rec = make_int4(gr[t] + o, K, o, 0);
              P = nfr - 1 - (rk >> 1);
            } else {
              rec = R[q];
              rec.y += 1;
              P = (rk & 1) ? nfr - 1 - (rk >> 1) : nfr + sIp[wr[t]];
            }
            rec.w = P;
            R2[j] = rec;
            Lq2[P] = j;
          }
        }
      } else {
        // the 512-hit level: per-hit values in LDS (eight rounds of registers measured slower there)
        // hits -> live nodes
        for (int32_t j = lane; j < c; j += 64) {
          const uint64_t v = hit(h + j);
          const int32_t g = (int32_t)((v >> 21) & 0x1fffff) - o;
          maxm = max(maxm, (int32_t)(v >> 42));
          int32_t lo = 0, hi = nA;
          while (lo < hi) {
            const int32_t mid = (lo + hi) >> 1;
            if (Ld[mid] > g) lo = mid + 1;
            else hi = mid;
          }
          sQ[j] = (lo < nA && Ld[lo] == g) ? lo : -1;
          Ld2[j] = g;
        }
        for (int32_t k = lane; k < nA; k += 64) sT[k] = INF;
        __syncthreads();
        for (int32_t j = lane; j < c; j += 64) {
          const int32_t q = sQ[j];
          if (q >= 0) sT[R[q].w] = j;
        }
        __syncthreads();
        {
          int32_t carry = -1;
          for (int32_t k0 = 0; k0 < nA; k0 += 64) {
            const int32_t k = k0 + lane;
            const int32_t t = (k < nA) ? sT[k] : -1;
            const int32_t inc = wave_incl_max_i32(t);
            const int32_t up = __shfl_up(inc, 1);
            if (k < nA) sM[k] = (lane == 0) ? carry : max(up, carry);
            carry = max(carry, __shfl(inc, 63));
          }
        }
        __syncthreads();
        // created / moved hits
        int32_t nfr = 0;
        bool bad = false;
        for (int32_t j0 = 0; j0 < c; j0 += 64) {
          const int32_t j = j0 + lane;
          bool cre = false, cf = false;
          if (j < c) {
            const int32_t q = sQ[j];
            cre = q < 0;
            cf = cre || sM[R[q].w] > j;
          }
          const uint64_t bcf = __ballot(cf), bcre = __ballot(cre);
          if (j < c) sRk[j] = (nfr + (int32_t)__popcll(bcf & lt)) * 2 + (cf ? 1 : 0);
          if (cre && (j > 0 || nA > 0)) bad = true;
          nfr += (int32_t)__popcll(bcf);
          ncre += (int32_t)__popcll(bcre);
        }
        if (nA == 0 && nnode > 0) {                    // first hit makes a node with no live node
          const int32_t es = head.z + head.y - K + 1, dg = head.x - head.z;
          if (abs(dg - Ld2[0]) > 3 || o < es + K - 2) bad = true;
        }
        if (__ballot(bad)) consistent = false;
        // list positions: extended in place, dead
        const int32_t sz = nA - (c - ncre);
        {
          int32_t ni = 0, nd = 0;
          for (int32_t k0 = 0; k0 < nA; k0 += 64) {
            const int32_t k = k0 + lane;
            bool inpl = false, dead = false;
            if (k < nA) {
              const int32_t t = sT[k];
              dead = t == INF;
              inpl = !dead && sM[k] < t;
            }
            const uint64_t bi = __ballot(inpl), bd = __ballot(dead);
            if (k < nA) sIp[k] = ni + (int32_t)__popcll(bi & lt);
            if (dead) {
              const int32_t r = nd + (int32_t)__popcll(bd & lt);
              int4 rec = R[Lq[k]];
              rec.w = ndq + sz - r;
              dq[ndq + r] = rec;
            }
            ni += (int32_t)__popcll(bi);
            nd += (int32_t)__popcll(bd);
          }
          ndq += sz;
        }
        __syncthreads();
        // the new live set, in hit order, with list positions
        for (int32_t j = lane; j < c; j += 64) {
          const int32_t q = sQ[j], rk = sRk[j];
          int4 rec;
          int32_t P;
          if (q < 0) {
            rec = make_int4(Ld2[j] + o, K, o, 0);
            P = nfr - 1 - (rk >> 1);
          } else {
            rec = R[q];
            rec.y += 1;
            P = (rk & 1) ? nfr - 1 - (rk >> 1) : nfr + sIp[rec.w];
          }
          rec.w = P;
          R2[j] = rec;
          Lq2[P] = j;
        }
      }
      __syncthreads();
      nnode += ncre;
      nA = c;
      prev_o = o;
      h += (uint64_t)c;
      cur ^= 1;
    }
    if (over) {
      if (lane == 0) {
        if (A.up) {
          A.up[atomicAdd(A.nup, 1u)] = p;
        } else {
          A.done[p] = 2;
          atomicAdd(A.handed_back, 1u);
        }
      }
      continue;
    }
    __threadfence();
    __syncthreads();
    int4 *out = A.nodes + hb;
    const int4 *R = sR[cur];
    const int32_t *Lq = sLq[cur];
    int32_t longest = 0, dmin = 0x7fffffff, dmax = -0x7fffffff;
    uint64_t lkey = 0;   // Len << 32 | ~position: the maximum is the first node of the longest Len
    for (int32_t k = lane; k < nA; k += 64) {
      int4 r = R[Lq[k]];
      r.w = 0;
      out[k] = r;
      lkey = max(lkey, ((uint64_t)(uint32_t)r.y << 32) | (uint32_t)~(uint32_t)k);
      longest = max(longest, r.y);
      dmin = min(dmin, r.x - r.z);
      dmax = max(dmax, r.x - r.z);
    }
    for (int32_t i = lane; i < ndq; i += 64) {
      int4 r = dq[i];
      const int32_t d = nA + ndq - r.w;
      r.w = 0;
      out[d] = r;
      lkey = max(lkey, ((uint64_t)(uint32_t)r.y << 32) | (uint32_t)~(uint32_t)d);
      longest = max(longest, r.y);
      dmin = min(dmin, r.x - r.z);
      dmax = max(dmax, r.x - r.z);
    }
    longest = wave_max_i32(longest);
    for (int o = 32; o > 0; o >>= 1) lkey = max(lkey, (uint64_t)__shfl_xor(lkey, o));
    dmin = wave_min_i32(dmin);
    dmax = wave_max_i32(dmax);
    maxm = wave_max_i32(maxm);
    if (lane == 0) {
      pr.nnodes = nA + ndq;
      atomicAdd(A.nnodes_total, (unsigned long long)(nA + ndq));
      pr.consistent = consistent ? 1 : 0;
      pr.diag_bgn = (int32_t)(A.hval[hb] & 0x1fffff);
      pr.diag_end = (int32_t)(A.hval[he - 1] & 0x1fffff);
      pr.max_mult = maxm;
      pr.longest = longest;
      pr.first_longest = (int32_t)~(uint32_t)lkey;
      pr.dmin = dmin;
      pr.dmax = dmax;
      A.pairs[p] = pr;
    }
  }
}

// ------------------------------------------------------------------ per-pair Process_Matches

struct OlapDev {
  int32_t s_lo, s_hi, t_lo, t_hi;
  double  quality;
  int32_t delta_ct;
  int32_t slb, srb, tlb, trb;
  int32_t min_diag, max_diag;
};

struct ExtOut {
  int32_t kind, s_lo, s_hi, t_lo, t_hi, errors, ndelta, overflow;
  int32_t fwd_e;       // errors of the forward DP (exit row)
};

struct SearchConst {
  DPTables tb;
  int32_t  min_olap;
  int32_t  unique;
  int32_t  partial;
  int32_t  hopeless;
  int32_t  K;
  const uint32_t *min_kmers; // --minkmers threshold by diagonal span, or nullptr
  int32_t  elide;          // skip reverse DP / tracebacks whose results Process_Matches cannot use
  const uint32_t *sscreen;  // screened ends per (S - s0, orientation) and per hash read (T - hbgn),
  const uint32_t *tscreen;  // bit 1 left, bit 2 right; nullptr without a skip file
  uint32_t s0, hbgn;
};

// Extend_Alignment (prefixEditDistance-extend.C:35-183) for one wave.  Deltas land in sc.ldelta
// (merged Left_Delta) when needed.
template <int CAP>
__device__ ExtOut wave_extend(int32_t Start, int32_t Offset, int32_t Len, const SeqV S, const SeqV Src, int32_t S_Len,
                              const SeqV T, const SeqV Trc, int32_t T_Len, const SearchConst &C, int32_t *bufA,
                              int32_t *bufB, const DPScratch &sc, uint64_t &cells, uint64_t &slots, uint32_t &dpcalls,
                              uint64_t &lane_bases, bool known_dovetail = false) {
  const int lane = lane_id();
  ExtOut r = {};
  const int32_t S_Left_Begin = Start - 1, S_Right_Begin = Start + Len, S_Right_Len = S_Len - S_Right_Begin;
  const int32_t T_Left_Begin = Offset - 1, T_Right_Begin = Offset + Len, T_Right_Len = T_Len - T_Right_Begin;
  const int32_t Total_Olap = (Start < Offset ? Start : Offset) + Len + (S_Right_Len < T_Right_Len ? S_Right_Len : T_Right_Len);
  const int32_t Error_Limit = C.tb.error_bound[Total_Olap];
  // With elision, deltas only after both sides end, by a second run with row storage; an extension
  // known to be a dovetail (its forward DPs already ran ahead) keeps rows in its only run.
  const bool need_delta = !C.elide || C.partial || known_dovetail;

  int32_t Right_Errors = 0, Left_Errors = 0, Leftover = 0, nr = 0, nl = 0;
  bool rmte = true, lmte = true;
  int32_t S_Hi, T_Hi, S_Lo, T_Lo;
  bool right_negate = false;

  if (S_Right_Len == 0 || T_Right_Len == 0) {
    S_Hi = 0; T_Hi = 0; rmte = true;
  } else {
    SeqV a = S, b = T;
    int32_t am, bn;
    if (S_Right_Len <= T_Right_Len) {
      a.b0 += S_Right_Begin; am = S_Right_Len; b.b0 += T_Right_Begin; bn = T_Right_Len; right_negate = true;
    } else {
      a = T; b = S; a.b0 += T_Right_Begin; am = T_Right_Len; b.b0 += S_Right_Begin; bn = S_Right_Len;
    }
    DPOut o = wave_prefix_dp<CAP>(a, am, b, bn, Error_Limit, true, need_delta, C.tb, bufA, bufB, sc, cells, slots, lane_bases);
    dpcalls++;
    if (o.overflow) { r.overflow = 1; return r; }
    Right_Errors = o.e; rmte = o.mte; nr = o.ndelta;
    if (right_negate) { S_Hi = o.a_end; T_Hi = o.t_end; } else { T_Hi = o.a_end; S_Hi = o.t_end; }
    if (!need_delta && !rmte) {
      // Process_Matches only distinguishes DOVETAIL from the rest when not doing partials
      r.kind = 0; r.fwd_e = Right_Errors;
      r.s_hi = S_Hi + S_Right_Begin - 1; r.t_hi = T_Hi + T_Right_Begin - 1;
      r.errors = Right_Errors;
      return r;
    }
  }
  r.fwd_e = Right_Errors;
  S_Hi += S_Right_Begin - 1;
  T_Hi += T_Right_Begin - 1;

  // with elision the right side matched to the end here; its delta is needed if the left side
  // also does: recompute it lazily below (rows of the right DP were not kept)
  if (S_Left_Begin < 0 || T_Left_Begin < 0) {
    S_Lo = 0; T_Lo = 0; lmte = true;
  } else {
    SeqV a, b;
    int32_t am, bn;
    bool swapped;
    if (S_Right_Begin <= T_Right_Begin) {
      a = Src; a.b0 += S_Len - 1 - S_Left_Begin; am = S_Left_Begin + 1;
      b = Trc; b.b0 += T_Len - 1 - T_Left_Begin; bn = T_Left_Begin + 1;
      swapped = false;
    } else {
      a = Trc; a.b0 += T_Len - 1 - T_Left_Begin; am = T_Left_Begin + 1;
      b = Src; b.b0 += S_Len - 1 - S_Left_Begin; bn = S_Left_Begin + 1;
      swapped = true;
    }
    DPOut o = wave_prefix_dp<CAP>(a, am, b, bn, Error_Limit - Right_Errors, false, need_delta, C.tb, bufA, bufB, sc, cells, slots, lane_bases);
    dpcalls++;
    if (o.overflow) { r.overflow = 1; return r; }
    Left_Errors = o.e; lmte = o.mte; nl = o.ndelta; Leftover = o.leftover;
    if (!swapped) { S_Lo = o.a_end; T_Lo = o.t_end; } else { T_Lo = o.a_end; S_Lo = o.t_end; }
    if (need_delta && swapped && lane == 0)
      for (int32_t i = 0; i < nl; i++) sc.ldelta[i] = -sc.ldelta[i];
  }
  S_Lo += S_Left_Begin + 1;
  T_Lo += T_Left_Begin + 1;

  r.errors = Left_Errors + Right_Errors;
  r.kind = (!rmte) ? (!lmte ? 0 : 2) : (!lmte ? 1 : 3);
  r.s_lo = S_Lo; r.s_hi = S_Hi; r.t_lo = T_Lo; r.t_hi = T_Hi;

  if (!need_delta && r.kind == 3) {
    // both sides matched to the end: rerun the two DPs with row storage for the deltas (their
    // cells repeat work already counted)
    uint64_t rc_cells = 0, rc_slots = 0;
    int32_t re_e = 0, rr = 0;
    {
      SeqV a = S, b = T;
      int32_t am, bn;
      if (S_Right_Len == 0 || T_Right_Len == 0) { am = 0; bn = 0; }
      else if (S_Right_Len <= T_Right_Len) { a.b0 += S_Right_Begin; am = S_Right_Len; b.b0 += T_Right_Begin; bn = T_Right_Len; }
      else { a = T; b = S; a.b0 += T_Right_Begin; am = T_Right_Len; b.b0 += S_Right_Begin; bn = S_Right_Len; }
      if (am > 0) {
        uint64_t lb = 0;
        DPOut o = wave_prefix_dp<CAP>(a, am, b, bn, Error_Limit, true, true, C.tb, bufA, bufB, sc, rc_cells, rc_slots, lb);
        if (o.overflow) { r.overflow = 1; return r; }
        rr = o.ndelta; re_e = o.e;
      }
      // keep Right_Delta aside in dbuf while the left DP reuses the row storage
      __syncthreads();
      if (lane == 0) for (int32_t i = 0; i < rr; i++) sc.dbuf[i] = sc.rdelta[i];
      __syncthreads();
    }
    (void)re_e;
    if (!(S_Left_Begin < 0 || T_Left_Begin < 0)) {
      SeqV a, b;
      int32_t am, bn;
      bool swapped;
      if (S_Right_Begin <= T_Right_Begin) {
        a = Src; a.b0 += S_Len - 1 - S_Left_Begin; am = S_Left_Begin + 1;
        b = Trc; b.b0 += T_Len - 1 - T_Left_Begin; bn = T_Left_Begin + 1;
        swapped = false;
      } else {
        a = Trc; a.b0 += T_Len - 1 - T_Left_Begin; am = T_Left_Begin + 1;
        b = Src; b.b0 += S_Len - 1 - S_Left_Begin; bn = S_Left_Begin + 1;
        swapped = true;
      }
      uint64_t lb = 0;
      DPOut o = wave_prefix_dp<CAP>(a, am, b, bn, Error_Limit - Right_Errors, false, true, C.tb, bufA, bufB, sc, rc_cells, rc_slots, lb);
      if (o.overflow) { r.overflow = 1; return r; }
      nl = o.ndelta; Leftover = o.leftover;
      // Set_Left_Delta may move T_End by one (reverse.C:76-92): take the coordinates from this run
      if (!swapped) { r.s_lo = o.a_end + S_Left_Begin + 1; r.t_lo = o.t_end + T_Left_Begin + 1; }
      else          { r.t_lo = o.a_end + T_Left_Begin + 1; r.s_lo = o.t_end + S_Left_Begin + 1; }
      if (swapped && lane == 0)
        for (int32_t i = 0; i < nl; i++) sc.ldelta[i] = -sc.ldelta[i];
    } else {
      nl = 0; Leftover = 0;
    }
    __syncthreads();
    if (lane == 0) for (int32_t i = 0; i < rr; i++) sc.rdelta[i] = sc.dbuf[i];
    __syncthreads();
    nr = rr;
  }

  // merge deltas (extend.C:164-175); Right_Delta was negated when S was the shorter right tail
  if (need_delta || r.kind == 3) {
    __syncthreads();
    if (lane == 0) {
      int32_t n = nl;
      if (nr > 0) {
        const int32_t r0 = right_negate ? -sc.rdelta[0] : sc.rdelta[0];
        sc.ldelta[n++] = (r0 > 0) ? -(r0 + Leftover + Len) : -(r0 - Leftover - Len);
      }
      for (int32_t i = 1; i < nr; i++) {
        const int32_t ri = right_negate ? -sc.rdelta[i] : sc.rdelta[i];
        sc.ldelta[n++] = -ri;
      }
      nl = n;
    }
    r.ndelta = __shfl(nl, 0);
    __syncthreads();
  }
  return r;
}

// Lies_On_Alignment (overlapInCore-Process_String_Overlaps.C:260-281)
__device__ __forceinline__ bool lies_on_alignment(int32_t start, int32_t offset, int32_t s_lo, int32_t t_lo,
                                                  const int32_t *__restrict__ delta, int32_t nd) {
  int32_t diag = t_lo - s_lo;
  const int32_t new_diag = offset - start;
  for (int32_t i = 0; i < nd; i++) {
    const int32_t dl = delta[i];
    s_lo += dl < 0 ? -dl : dl;
    if (start < s_lo) return abs(new_diag - diag) <= OIC_SHIFT_SLACK;
    if (dl < 0) diag++;
    else { s_lo++; diag--; }
  }
  return abs(new_diag - diag) <= OIC_SHIFT_SLACK;
}

// Add_Overlap (Process_String_Overlaps.C:175-244), executed by lane 0 on the LDS copy
__device__ void add_overlap(const SearchConst &C, int32_t s_lo, int32_t s_hi, int32_t t_lo, int32_t t_hi, double qual,
                            int32_t delta_ct, OlapDev *olap, int32_t *ct) {
  if (!C.partial) {
    const int32_t new_diag = t_lo - s_lo;
    for (int32_t i = 0; i < *ct; i++) {
      const int32_t old_diag = olap[i].t_lo - olap[i].s_lo;
      if ((new_diag > 0 && old_diag > 0 && olap[i].trb - new_diag - olap[i].slb >= OIC_MIN_INTER) ||
          (new_diag <= 0 && old_diag <= 0 && olap[i].srb + new_diag - olap[i].tlb >= OIC_MIN_INTER)) {
        if (new_diag < olap[i].min_diag) olap[i].min_diag = new_diag;
        if (new_diag > olap[i].max_diag) olap[i].max_diag = new_diag;
        if (s_lo < olap[i].slb) olap[i].slb = s_lo;
        if (s_hi > olap[i].srb) olap[i].srb = s_hi;
        if (t_lo < olap[i].tlb) olap[i].tlb = t_lo;
        if (t_hi > olap[i].trb) olap[i].trb = t_hi;
        if (qual < olap[i].quality) {
          olap[i].s_lo = s_lo; olap[i].s_hi = s_hi; olap[i].t_lo = t_lo; olap[i].t_hi = t_hi;
          olap[i].quality = qual;
          olap[i].delta_ct = delta_ct;
        }
        return;
      }
    }
  }
  if (*ct >= OIC_MAX_DISTINCT) return;
  OlapDev &o = olap[*ct];
  o.s_lo = o.slb = s_lo; o.s_hi = o.srb = s_hi; o.t_lo = o.tlb = t_lo; o.t_hi = o.trb = t_hi;
  o.quality = qual;
  o.delta_ct = delta_ct;
  o.min_diag = o.max_diag = t_lo - s_lo;
  (*ct)++;
}

// The per-pair selection of Process_String_Overlaps.C:501-513 over the distinct overlaps Add_Overlap
// left: Choose_Best_Partial under partial and unique, Combine_Into_One_Olap under unique,
// Merge_Intersecting_Olaps otherwise.  Updates the kept entries' bounds; deleted[] starts at zero.
__device__ void select_distinct(const SearchConst &C, OlapDev *olap, const int32_t ct, int32_t *deleted) {
  if (ct > 0) {
    if (C.partial) {
      if (C.unique) {   // Choose_Best_Partial (:289-311)
        int32_t best = 0;
        double mb0 = (1.0 - olap[0].quality) * (2 + olap[0].s_hi - olap[0].s_lo + olap[0].t_hi - olap[0].t_lo);
        for (int32_t i = 1; i < ct; i++) {
          const double mb = (1.0 - olap[i].quality) * (2 + olap[i].s_hi - olap[i].s_lo + olap[i].t_hi - olap[i].t_lo);
          if (mb0 < mb || (mb0 == mb && olap[i].quality < olap[best].quality)) best = i;
        }
        for (int32_t i = 0; i < ct; i++) deleted[i] = (i != best);
      }
    } else if (C.unique) {   // Combine_Into_One_Olap (:42-96)
      int32_t best = 0;
      int32_t mnd = olap[0].min_diag, mxd = olap[0].max_diag, slb = olap[0].slb, srb = olap[0].srb, tlb = olap[0].tlb, trb = olap[0].trb;
      for (int32_t i = 1; i < ct; i++) {
        const int32_t si = olap[i].s_hi - olap[i].s_lo, ti = olap[i].t_hi - olap[i].t_lo;
        const int32_t sb = olap[best].s_hi - olap[best].s_lo, tb2 = olap[best].t_hi - olap[best].t_lo;
        const int32_t leni = 1 + (si < ti ? si : ti), lenb = 1 + (sb < tb2 ? sb : tb2);
        if (olap[i].quality < olap[best].quality || (olap[i].quality == olap[best].quality && leni > lenb)) best = i;
        if (olap[i].min_diag < mnd) mnd = olap[i].min_diag;
        if (olap[i].max_diag > mxd) mxd = olap[i].max_diag;
        if (olap[i].slb < slb) slb = olap[i].slb;
        if (olap[i].srb > srb) srb = olap[i].srb;
        if (olap[i].tlb < tlb) tlb = olap[i].tlb;
        if (olap[i].trb > trb) trb = olap[i].trb;
      }
      olap[best].min_diag = mnd; olap[best].max_diag = mxd; olap[best].slb = slb; olap[best].srb = srb;
      olap[best].tlb = tlb; olap[best].trb = trb;
      for (int32_t i = 0; i < ct; i++) deleted[i] = (i != best);
    } else {   // Merge_Intersecting_Olaps (:106-162)
      for (int32_t i = 0; i < ct - 1; i++)
        for (int32_t j = i + 1; j < ct; j++) {
          if (deleted[i] || deleted[j]) continue;
          const int32_t lo = olap[i].min_diag, hi = olap[i].max_diag;
          if ((lo <= 0 && olap[j].min_diag > 0) || (lo > 0 && olap[j].min_diag <= 0)) continue;
          if ((lo >= 0 && olap[j].trb - lo - olap[j].slb >= OIC_MIN_INTER) || (lo <= 0 && olap[j].srb + lo - olap[j].tlb >= OIC_MIN_INTER) ||
              (hi >= 0 && olap[j].trb - hi - olap[j].slb >= OIC_MIN_INTER) || (hi <= 0 && olap[j].srb + hi - olap[j].tlb >= OIC_MIN_INTER)) {
            OlapDev *keep, *disc;
            if (olap[i].quality < olap[j].quality) { keep = &olap[i]; disc = &olap[j]; deleted[j] = 1; }
            else                                   { keep = &olap[j]; disc = &olap[i]; deleted[i] = 1; }
            if (disc->min_diag < keep->min_diag) keep->min_diag = disc->min_diag;
            if (disc->max_diag > keep->max_diag) keep->max_diag = disc->max_diag;
            if (disc->slb < keep->slb) keep->slb = disc->slb;
            if (disc->srb > keep->srb) keep->srb = disc->srb;
            if (disc->tlb < keep->tlb) keep->tlb = disc->tlb;
            if (disc->trb > keep->trb) keep->trb = disc->trb;
          }
        }
    }
  }
}

__device__ __forceinline__ uint32_t encode_evalue(double q) {
  return (q < (65535 / 100000.0)) ? (uint32_t)(100000.0 * q + 0.5) : 65535u;
}

struct OutRec { uint32_t a_iid, b_iid; uint64_t w0, w1; };

__device__ __forceinline__ uint64_t pack_w0(uint32_t ahg5, uint32_t ahg3, uint32_t ev, uint32_t fl, uint32_t obt, uint32_t dup, uint32_t utg) {
  const uint64_t F = 0x1fffffull;
  return ((uint64_t)ahg5 & F) | (((uint64_t)ahg3 & F) << 21) | (((uint64_t)ev & 0xffff) << 42) | ((uint64_t)(fl & 1) << 58) |
         ((uint64_t)(obt & 1) << 59) | ((uint64_t)(dup & 1) << 60) | ((uint64_t)(utg & 1) << 61);
}
__device__ __forceinline__ uint64_t pack_w1(uint32_t bhg5, uint32_t bhg3, uint32_t span) {
  const uint64_t F = 0x1fffffull;
  return ((uint64_t)bhg5 & F) | (((uint64_t)bhg3 & F) << 21) | (((uint64_t)span & F) << 42);
}

// Output_Overlap (overlapInCore-Output.C:26-191); returns 1 if counted as contained
__device__ int output_overlap(uint32_t S_ID, int32_t S_Len, int32_t S_Dir, uint32_t T_ID, int32_t T_Len, const OlapDev &o, OutRec *rec) {
  const uint32_t span = (uint32_t)((o.s_hi - o.s_lo) + (o.t_hi - o.t_lo) + o.delta_ct) / 2u;
  const int32_t S_Right_Hang = S_Len - o.s_hi - 1;
  const int32_t T_Right_Hang = T_Len - o.t_hi - 1;
  const bool Sleft = (o.s_lo > o.t_lo) || ((o.s_lo == o.t_lo) && (S_Right_Hang > T_Right_Hang));
  rec->a_iid = Sleft ? S_ID : T_ID;
  rec->b_iid = Sleft ? T_ID : S_ID;
  char orient;
  int32_t ahg, bhg;
  if (Sleft) { orient = (S_Dir == 0) ? 'N' : 'O'; ahg = o.s_lo; bhg = T_Right_Hang - S_Right_Hang; }
  else       { orient = (S_Dir == 0) ? 'N' : 'I'; ahg = o.t_lo; bhg = S_Right_Hang - T_Right_Hang; }
  if (orient == 'O' && S_Right_Hang >= T_Right_Hang) {
    orient = 'I';
    ahg = -(T_Right_Hang - S_Right_Hang);
    bhg = -(o.s_lo);
  }
  int32_t ah, bh;
  uint32_t fl;
  if (orient == 'N')      { ah = ahg;  bh = bhg;  fl = 0; }
  else if (orient == 'I') { ah = ahg;  bh = bhg;  fl = 1; }
  else                    { ah = -bhg; bh = -ahg; fl = 1; }
  const uint32_t ahg5 = ah < 0 ? 0u : (uint32_t)ah, bhg5 = ah < 0 ? (uint32_t)(-ah) : 0u;
  const uint32_t bhg3 = bh < 0 ? 0u : (uint32_t)bh, ahg3 = bh < 0 ? (uint32_t)(-bh) : 0u;
  rec->w0 = pack_w0(ahg5, ahg3, encode_evalue(o.quality), fl, 0, 0, 1);
  rec->w1 = pack_w1(bhg5, bhg3, span);
  return bhg <= 0;
}

// Output_Partial_Overlap (overlapInCore-Output.C:195-264)
__device__ void output_partial(uint32_t s_id, uint32_t t_id, int32_t dir, const OlapDev &o, int32_t s_len, int32_t t_len, OutRec *rec) {
  const uint32_t span = (uint32_t)((o.s_hi - o.s_lo) + (o.t_hi - o.t_lo) + o.delta_ct) / 2u;
  rec->a_iid = s_id;
  rec->b_iid = t_id;
  uint32_t ahg5, ahg3, bhg5, bhg3, fl;
  if (dir == 0) { ahg5 = o.s_lo; ahg3 = s_len - (o.s_hi + 1); bhg5 = o.t_lo; bhg3 = t_len - (o.t_hi + 1); fl = 0; }
  else          { ahg5 = s_len - (o.s_hi + 1); ahg3 = o.s_lo; bhg5 = t_len - (o.t_hi + 1); bhg3 = o.t_lo; fl = 1; }
  rec->w0 = pack_w0(ahg5, ahg3, encode_evalue(o.quality), fl, 1, 1, 0);
  rec->w1 = pack_w1(bhg5, bhg3, span);
}

// per-pair outcome, also the trace the cost model learns from
struct PairOut {
  int32_t  nout;          // records written (<= 3)
  int32_t  n_ext;         // Extend_Alignment calls
  uint64_t cells;         // DP cells executed for this pair
  int32_t  status;        // 0 ok, 1 overflow (retry with a wider tile), 2 hopeless, 3 skipped by --minkmers
  int32_t  exit_d;        // first extension: diagonal at the right end
  uint32_t contained, dovetail;
};

#define STAT_WITHOUT 0
#define STAT_WITH    1
#define STAT_MULTI   2
#define STAT_TOTAL   3
#define STAT_CONT    4
#define STAT_DOVE    5
#define STAT_CELLS   8
#define STAT_SLOTS   9
#define STAT_EXT     10
#define STAT_DPCALLS 11
#define STAT_BASES   12
#define STAT_PPCELLS 13   // cells of the DPs this kernel runs itself (full extensions)
#define STAT_NSTATS  16

// A node's cached outcome in one 64-bit word, so a finished task makes one store: the kind in bits
// 62-63 (0 not computed, 1 not a dovetail, 2 dovetail), the right-end diagonal (T_Hi - S_Hi) of
// the run in bits 32-61 (signed), its DP cells in bits 0-31.  A zeroed word reads "not computed".
__host__ __device__ __forceinline__ uint64_t pack_outcome(int kind, int32_t exit_d, uint32_t cells) {
  return ((uint64_t)(uint32_t)kind << 62) | (((uint64_t)(uint32_t)exit_d & 0x3FFFFFFFull) << 32) | (uint64_t)cells;
}
__host__ __device__ __forceinline__ int outcome_kind(uint64_t o) { return (int)(o >> 62); }
__host__ __device__ __forceinline__ uint32_t outcome_cells(uint64_t o) { return (uint32_t)o; }
__host__ __device__ __forceinline__ int32_t outcome_exit(uint64_t o) { return (int32_t)((int64_t)(o << 2) >> 34); }

struct PairKernelArgs {
  ReadsDev R;
  SearchConst C;
  PairRec *pairs;
  const uint32_t *order;     // processing order (pair indices)
  uint64_t norder;
  int4 *nodes;               // per pair node list (x Offset, y Len, z Start, w removed)
  OutRec *recs;              // [np * 3]
  PairOut *pout;             // [np]
  uint32_t *queue;           // work counter
  unsigned long long *stats; // [STAT_NSTATS]
  const uint64_t *outc;      // per node cached forward-extension outcome (pack_outcome), or nullptr
  DPScratch sc0;             // scratch of wave 0; wave w adds w * stride
  int64_t row_stride, val_stride, delta_stride;
};

template <int CAP>
__device__ __forceinline__ DPScratch wave_scratch(const PairKernelArgs &A, uint32_t w) {
  DPScratch s = A.sc0;
  s.rowL = A.sc0.rowL + (int64_t)w * A.row_stride;
  s.rowR = A.sc0.rowR + (int64_t)w * A.row_stride;
  s.rowOff = A.sc0.rowOff + (int64_t)w * A.row_stride;
  s.rowVal = A.sc0.rowVal + (int64_t)w * A.val_stride;
  s.dbuf = A.sc0.dbuf + (int64_t)w * A.delta_stride;
  s.ldelta = A.sc0.ldelta + (int64_t)w * A.delta_stride;
  s.rdelta = A.sc0.rdelta + (int64_t)w * A.delta_stride;
  return s;
}

// Process_Matches (overlapInCore-Process_String_Overlaps.C:353-547) for one candidate pair per
// wavefront.  Persistent: each wave pulls the next pair index from a global counter.
template <int CAP>
__global__ void __launch_bounds__(OIC_WAVE) k_process_pairs(PairKernelArgs A) {
  __shared__ int32_t bufA[CAP], bufB[CAP];
  __shared__ OlapDev olap[OIC_MAX_DISTINCT];
  __shared__ int32_t olap_ct;
  __shared__ uint32_t next_item;
  const int lane = lane_id();
  const uint32_t w = blockIdx.x;
  const DPScratch sc = wave_scratch<CAP>(A, w);
  const SearchConst &C = A.C;

  uint64_t cells = 0, slots = 0, lane_bases = 0, own_cells = 0;
  uint32_t dpcalls = 0, exts = 0;
  while (true) {
    if (lane == 0) next_item = atomicAdd(A.queue, 1u);
    __syncthreads();
    const uint32_t qi = next_item;
    __syncthreads();
    if (qi >= A.norder) break;
    const uint32_t p = A.order[qi];
    const PairRec pr = A.pairs[p];
    int4 *nodes = A.nodes + pr.hb;
    const int32_t nn = pr.nnodes;
    const uint32_t S_ID = pr.s_id, T_ID = pr.t_id;
    const int32_t S_Len = A.R.len[S_ID], T_Len = A.R.len[T_ID];
    const SeqV S = read_view(A.R, S_ID, pr.dir), Src = read_view(A.R, S_ID, 1 - pr.dir);
    const SeqV T = read_view(A.R, T_ID, 0), Trc = read_view(A.R, T_ID, 1);
    PairOut po = {};
    const uint64_t cells0 = cells;

    // --minkmers filter (Process_String_Overlaps.C:618-621); the threshold per diagonal span is a
    // host-computed table (computeMinimumKmers, :22-33)
    if (C.min_kmers) {
      const uint64_t mk = C.min_kmers[pr.diag_end - pr.diag_bgn];
      if (mk > (uint64_t)(int64_t)(pr.he - pr.hb)) {
        if (lane == 0) { po.status = 3; A.pout[p] = po; }
        continue;
      }
    }

    // hopeless check for a singleton match (Process_String_Overlaps.C:384-415)
    if (C.hopeless && nn == 1 && !C.partial) {
      const int4 nd = nodes[0];
      const int32_t s_head = nd.z, t_head = nd.x;
      const uint32_t ss = C.sscreen ? C.sscreen[2 * (S_ID - C.s0) + pr.dir] : 0u;
      const uint32_t ts = C.tscreen ? C.tscreen[T_ID - C.hbgn] : 0u;
      bool hopeless = false;
      if (s_head <= t_head) { if (s_head > OIC_HOPELESS && !(ss & 1u)) hopeless = true; }
      else                  { if (t_head > OIC_HOPELESS && !(ts & 1u)) hopeless = true; }
      const int32_t s_tail = S_Len - s_head - nd.y + 1, t_tail = T_Len - t_head - nd.y + 1;
      if (s_tail <= t_tail) { if (s_tail > OIC_HOPELESS && !(ss & 2u)) hopeless = true; }
      else                  { if (t_tail > OIC_HOPELESS && !(ts & 2u)) hopeless = true; }
      if (hopeless) {
        if (lane == 0) {
          po.status = 2;
          A.pout[p] = po;
          atomicAdd(&A.stats[STAT_WITHOUT], 1ull);
        }
        continue;
      }
    }

    if (lane == 0) olap_ct = 0;
    __syncthreads();
    bool overflow = false;
    bool first = true;
    const uint64_t *outc = A.outc ? A.outc + pr.hb : nullptr;   // cached forward-extension outcomes (pack_outcome)
    while (true) {
      // first node with the maximal Len (Process_String_Overlaps.C:424-431)
      int32_t bl = -1, bi = 0x7fffffff;
      for (int32_t i = lane; i < nn; i += OIC_WAVE) {
        const int4 nd = nodes[i];
        if (!nd.w && nd.y > bl) { bl = nd.y; bi = i; }
      }
      wave_argmax_first(bl, bi);
      if (bl < 0) break;
      if (outc && outcome_kind(outc[bi]) == 1) {
        // A node whose extension is known not to be a dovetail is only removed (:473-490); so are
        // all such nodes that would be picked before the first node without that knowledge.
        int32_t ul = -1, ui = 0x7fffffff;
        for (int32_t i = lane; i < nn; i += OIC_WAVE) {
          const int4 nd = nodes[i];
          if (!nd.w && outcome_kind(outc[i]) != 1 && nd.y > ul) { ul = nd.y; ui = i; }
        }
        wave_argmax_first(ul, ui);
        uint64_t cnt = 0, cs = 0;
        if (pr.consistent) {
          if (lane == 0) { cnt = 1; cs = outcome_cells(outc[bi]); nodes[bi].w = 1; }
        } else {
          for (int32_t i = lane; i < nn; i += OIC_WAVE) {
            const int4 nd = nodes[i];
            const uint64_t oi = outc[i];
            if (!nd.w && outcome_kind(oi) == 1 && (nd.y > ul || (nd.y == ul && i < ui))) {
              nodes[i].w = 1;
              cnt++;
              cs += outcome_cells(oi);
            }
          }
        }
        for (int o = 32; o > 0; o >>= 1) { cnt += __shfl_xor(cnt, o); cs += __shfl_xor(cs, o); }
        exts += (uint32_t)cnt;
        po.n_ext += (int32_t)cnt;
        cells += cs;
        if (first) { po.exit_d = outcome_exit(outc[bi]); first = false; }
        __threadfence_block();
        __syncthreads();
        if (pr.consistent) break;
        continue;
      }
      const int4 lm = nodes[bi];
      const uint64_t cx = cells;
      ExtOut x = wave_extend<CAP>(lm.z, lm.x, lm.y, S, Src, S_Len, T, Trc, T_Len, C, bufA, bufB, sc, cells, slots, dpcalls, lane_bases,
                                  outc && outcome_kind(outc[bi]) == 2);
      own_cells += cells - cx;
      exts++;
      po.n_ext++;
      if (x.overflow) { overflow = true; break; }
      if (first) { po.exit_d = x.t_hi - x.s_hi; first = false; }
      if (x.kind == 3 || C.partial) {
        if (1 + x.s_hi - x.s_lo >= C.min_olap && 1 + x.t_hi - x.t_lo >= C.min_olap) {
          const int32_t a = x.s_hi - x.s_lo, b = x.t_hi - x.t_lo;
          const int32_t olen = 1 + (a < b ? a : b);
          const double q = (double)x.errors / (double)olen;
          if (x.errors <= C.tb.error_bound[olen]) {
            if (lane == 0) add_overlap(C, x.s_lo, x.s_hi, x.t_lo, x.t_hi, q, x.ndelta, olap, &olap_ct);
          }
        }
      }
      __syncthreads();
      if (pr.consistent) break;
      // remove the longest match and, after a dovetail, every match on the alignment (:476-490)
      if (lane == 0) nodes[bi].w = 1;
      if (x.kind == 3 || C.partial) {
        for (int32_t i = lane; i < nn; i += OIC_WAVE) {
          const int4 nd = nodes[i];
          if (nd.w || i == bi) continue;
          if (x.s_lo - OIC_SHIFT_SLACK <= nd.z && nd.z + nd.y <= (x.s_hi + 1) + OIC_SHIFT_SLACK - 1 &&
              lies_on_alignment(nd.z, nd.x, x.s_lo, x.t_lo, sc.ldelta, x.ndelta))
            nodes[i].w = 1;
        }
      }
      __threadfence_block();
      __syncthreads();
    }

    if (overflow) {
      // leave the node list untouched for the retry: clear removal marks
      for (int32_t i = lane; i < nn; i += OIC_WAVE) nodes[i].w = 0;
      if (lane == 0) { po.status = 1; A.pout[p] = po; }
      __syncthreads();
      continue;
    }

    if (lane == 0) {
      const int32_t ct = olap_ct;
      int32_t deleted[OIC_MAX_DISTINCT] = {0, 0, 0};
      select_distinct(C, olap, ct, deleted);
      int32_t nout = 0;
      uint32_t cont = 0, dove = 0;
      for (int32_t i = 0; i < ct; i++) {
        if (deleted[i]) continue;
        OutRec rr;
        if (C.partial) output_partial(S_ID, T_ID, pr.dir, olap[i], S_Len, T_Len, &rr);
        else {
          if (output_overlap(S_ID, S_Len, pr.dir, T_ID, T_Len, olap[i], &rr)) cont++; else dove++;
        }
        A.recs[(uint64_t)p * OIC_MAX_DISTINCT + nout] = rr;
        nout++;
      }
      po.nout = nout;
      po.contained = cont;
      po.dovetail = dove;
      po.cells = cells - cells0;
      A.pout[p] = po;
      if (nout == 0) atomicAdd(&A.stats[STAT_WITHOUT], 1ull);
      else {
        atomicAdd(&A.stats[STAT_WITH], 1ull);
        if (nout > 1) atomicAdd(&A.stats[STAT_MULTI], 1ull);
      }
      atomicAdd(&A.stats[STAT_TOTAL], (unsigned long long)nout);
      if (!C.partial) {
        atomicAdd(&A.stats[STAT_CONT], (unsigned long long)cont);
        atomicAdd(&A.stats[STAT_DOVE], (unsigned long long)dove);
      }
    }
    __syncthreads();
  }
  for (int o = 32; o > 0; o >>= 1) lane_bases += __shfl_xor(lane_bases, o);
  if (lane == 0) {
    atomicAdd(&A.stats[STAT_BASES], (unsigned long long)lane_bases);
    atomicAdd(&A.stats[STAT_CELLS], (unsigned long long)cells);
    atomicAdd(&A.stats[STAT_PPCELLS], (unsigned long long)own_cells);
    atomicAdd(&A.stats[STAT_SLOTS], (unsigned long long)slots);
    atomicAdd(&A.stats[STAT_EXT], (unsigned long long)exts);
    atomicAdd(&A.stats[STAT_DPCALLS], (unsigned long long)dpcalls);
  }
}

// record compaction: pairs' records into a dense array at scanned offsets
__global__ void k_compact_recs(const OutRec *__restrict__ recs, const PairOut *__restrict__ pout, const uint64_t *__restrict__ off,
                               uint64_t np, OutRec *__restrict__ out) {
  const uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  const int32_t n = pout[p].nout;
  for (int32_t i = 0; i < n; i++) out[off[p] + i] = recs[p * OIC_MAX_DISTINCT + i];
}

__global__ void k_nout(const PairOut *__restrict__ pout, uint64_t np, uint32_t *__restrict__ n) {
  const uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p < np) n[p] = (uint32_t)pout[p].nout;
}
