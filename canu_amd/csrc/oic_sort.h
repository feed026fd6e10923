// oic_sort.h -- device-wide exclusive scan and stable LSD radix sort (8-bit digits) for gfx950.
//
// Radix sort: per pass, (1) each 256-thread block histograms its tile of 256 x ITEMS keys,
// (2) the digit-major histogram is exclusive-scanned, (3) each block ranks its tile stably --
// ITEMS rounds of 256 keys in order; within a 64-lane wave, lanes holding the same digit are found
// with 8 ballots -- and scatters keys and values.  Stability is what the callers rely on: the
// k-mer table keeps positions in (read, offset) order within a k-mer, and the hit grouping keeps
// hits in (S offset, T offset) generation order within a read pair.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define SORT_THREADS 256
#define SORT_ITEMS   32
#define SORT_TILE    (SORT_THREADS * SORT_ITEMS)
#define SCAN_THREADS 256
#define SCAN_ITEMS   8
#define SCAN_TILE    (SCAN_THREADS * SCAN_ITEMS)

// ---------------------------------------------------------------- exclusive scan (uint64)

__device__ __forceinline__ uint64_t wave_incl_scan_u64(uint64_t v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint64_t t = __shfl_up(v, o);
    if (lane >= o) v += t;
  }
  return v;
}

// block-wide exclusive scan of one value per thread; returns exclusive prefix, total in *tot
__device__ __forceinline__ uint64_t block_excl_scan_u64(uint64_t v, uint64_t *tot) {
  __shared__ uint64_t wsum[SCAN_THREADS / 64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const uint64_t inc = wave_incl_scan_u64(v);
  if (lane == 63) wsum[w] = inc;
  __syncthreads();
  uint64_t off = 0, all = 0;
  for (int i = 0; i < SCAN_THREADS / 64; i++) {
    if (i < w) off += wsum[i];
    all += wsum[i];
  }
  __syncthreads();
  *tot = all;
  return off + inc - v;
}

template <typename T>
__global__ void __launch_bounds__(SCAN_THREADS) k_scan_tiles(const T *__restrict__ in, uint64_t *__restrict__ out,
                                                             uint64_t *__restrict__ tile_sums, uint64_t n) {
  const uint64_t base = (uint64_t)blockIdx.x * SCAN_TILE + (uint64_t)threadIdx.x * SCAN_ITEMS;
  uint64_t v[SCAN_ITEMS];
  uint64_t s = 0;
#pragma unroll
  for (int i = 0; i < SCAN_ITEMS; i++) {
    v[i] = (base + i < n) ? (uint64_t)in[base + i] : 0;
    s += v[i];
  }
  uint64_t tot;
  uint64_t off = block_excl_scan_u64(s, &tot);
#pragma unroll
  for (int i = 0; i < SCAN_ITEMS; i++) {
    if (base + i < n) out[base + i] = off;
    off += v[i];
  }
  if (threadIdx.x == 0 && tile_sums) tile_sums[blockIdx.x] = tot;
}

__global__ void __launch_bounds__(SCAN_THREADS) k_scan_add(uint64_t *__restrict__ out, const uint64_t *__restrict__ tile_off,
                                                           uint64_t n) {
  const uint64_t add = tile_off[blockIdx.x];
  const uint64_t base = (uint64_t)blockIdx.x * SCAN_TILE;
  for (int i = threadIdx.x; i < SCAN_TILE; i += SCAN_THREADS)
    if (base + i < n) out[base + i] += add;
}

// scratch needed by scan_excl for n items (uint64 words)
static inline uint64_t scan_scratch_words(uint64_t n) {
  uint64_t words = 0;
  while (n > SCAN_TILE) {
    n = (n + SCAN_TILE - 1) / SCAN_TILE;
    words += 2 * n + 2;
  }
  return words + 2;
}

// out[i] = sum(in[0..i)); returns total via *total (device->host copy); scratch from scan_scratch_words
template <typename T>
static hipError_t scan_excl(const T *in, uint64_t *out, uint64_t n, uint64_t *scratch, hipStream_t st, uint64_t *total_host) {
  if (n == 0) {
    if (total_host) *total_host = 0;
    return hipSuccess;
  }
  const uint64_t tiles = (n + SCAN_TILE - 1) / SCAN_TILE;
  uint64_t *sums = scratch;
  uint64_t *sums_off = scratch + tiles + 1;
  hipLaunchKernelGGL(k_scan_tiles<T>, dim3((unsigned)tiles), dim3(SCAN_THREADS), 0, st, in, out, sums, n);
  if (tiles > 1) {
    hipError_t e = scan_excl<uint64_t>(sums, sums_off, tiles, scratch + 2 * tiles + 2, st, nullptr);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_scan_add, dim3((unsigned)tiles), dim3(SCAN_THREADS), 0, st, out, sums_off, n);
  }
  if (total_host) {
    // total = out[n-1] + in[n-1]
    uint64_t last_out = 0;
    T last_in = 0;
    hipMemcpyAsync(&last_out, out + n - 1, sizeof(uint64_t), hipMemcpyDeviceToHost, st);
    hipMemcpyAsync(&last_in, in + n - 1, sizeof(T), hipMemcpyDeviceToHost, st);
    hipError_t e = hipStreamSynchronize(st);
    if (e != hipSuccess) return e;
    *total_host = last_out + (uint64_t)last_in;
  }
  return hipGetLastError();
}

// ---------------------------------------------------------------- radix sort (u64 key, u64 value)

__global__ void __launch_bounds__(SORT_THREADS) k_radix_hist(const uint64_t *__restrict__ keys, uint64_t n, int shift,
                                                             uint32_t *__restrict__ hist, uint32_t nblocks) {
  __shared__ uint32_t h[256];
  h[threadIdx.x] = 0;
  __syncthreads();
  const uint64_t base = (uint64_t)blockIdx.x * SORT_TILE;
#pragma unroll
  for (int i = 0; i < SORT_ITEMS; i++) {
    const uint64_t idx = base + (uint64_t)i * SORT_THREADS + threadIdx.x;
    if (idx < n) atomicAdd(&h[(keys[idx] >> shift) & 0xff], 1u);
  }
  __syncthreads();
  hist[(uint64_t)threadIdx.x * nblocks + blockIdx.x] = h[threadIdx.x];
}

__global__ void __launch_bounds__(SORT_THREADS) k_radix_scatter(const uint64_t *__restrict__ keys, const uint64_t *__restrict__ vals,
                                                                uint64_t *__restrict__ okeys, uint64_t *__restrict__ ovals,
                                                                uint64_t n, int shift, const uint64_t *__restrict__ offs,
                                                                uint32_t nblocks) {
  __shared__ uint32_t run[256];                       // running count per digit in this tile
  __shared__ uint32_t wcnt[SORT_THREADS / 64][256];   // per-wave digit counts of the current round
  __shared__ uint64_t gbase[256];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  run[tid] = 0;
  gbase[tid] = offs[(uint64_t)tid * nblocks + blockIdx.x];
  const uint64_t base = (uint64_t)blockIdx.x * SORT_TILE;
  const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));

  for (int r = 0; r < SORT_ITEMS; r++) {
    const uint64_t idx = base + (uint64_t)r * SORT_THREADS + tid;
    const bool valid = idx < n;
    uint64_t k = 0, v = 0;
    uint32_t dg = 0;
    if (valid) { k = keys[idx]; v = vals[idx]; dg = (uint32_t)(k >> shift) & 0xffu; }
    // lanes of this wave with the same digit (invalid lanes get their own class)
    uint64_t peers = __ballot(valid);
    if (!valid) peers = ~peers;
#pragma unroll
    for (int b = 0; b < 8; b++) {
      const uint64_t bb = __ballot((dg >> b) & 1u);
      peers &= ((dg >> b) & 1u) ? bb : ~bb;
    }
    const uint32_t rank_w = (uint32_t)__popcll(peers & lt_mask);
    const bool leader = (peers & lt_mask) == 0;
    for (int i = tid; i < (SORT_THREADS / 64) * 256; i += SORT_THREADS) (&wcnt[0][0])[i] = 0;
    __syncthreads();
    if (valid && leader) wcnt[w][dg] = (uint32_t)__popcll(peers);
    __syncthreads();
    if (valid) {
      uint32_t off = run[dg] + rank_w;
      for (int i = 0; i < w; i++) off += wcnt[i][dg];
      const uint64_t dst = gbase[dg] + off;
      okeys[dst] = k;
      ovals[dst] = v;
    }
    __syncthreads();
    {
      uint32_t s = 0;
      for (int i = 0; i < SORT_THREADS / 64; i++) s += wcnt[i][tid];
      run[tid] += s;
    }
    __syncthreads();
  }
}

struct RadixTemp {
  uint64_t *k2, *v2;        // n each
  uint32_t *hist;           // 256 * nblocks
  uint64_t *offs;           // 256 * nblocks
  uint64_t *scan_scratch;   // scan_scratch_words(256 * nblocks)
};

static inline uint64_t radix_blocks(uint64_t n) { return (n + SORT_TILE - 1) / SORT_TILE; }

// Sorts (keys, vals) stably by key bits [0, nbits).  Result ends in (*keys, *vals), which may be
// swapped with (tmp.k2, tmp.v2).
static hipError_t radix_sort(uint64_t **keys, uint64_t **vals, uint64_t n, int nbits, RadixTemp &tmp, hipStream_t st) {
  if (n <= 1) return hipSuccess;
  const uint32_t nb = (uint32_t)radix_blocks(n);
  uint64_t *ka = *keys, *va = *vals, *kb = tmp.k2, *vb = tmp.v2;
  for (int shift = 0; shift < nbits; shift += 8) {
    hipLaunchKernelGGL(k_radix_hist, dim3(nb), dim3(SORT_THREADS), 0, st, ka, n, shift, tmp.hist, nb);
    hipError_t e = scan_excl<uint32_t>(tmp.hist, tmp.offs, (uint64_t)256 * nb, tmp.scan_scratch, st, nullptr);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_radix_scatter, dim3(nb), dim3(SORT_THREADS), 0, st, ka, va, kb, vb, n, shift, tmp.offs, nb);
    uint64_t *t;
    t = ka; ka = kb; kb = t;
    t = va; va = vb; vb = t;
  }
  *keys = ka; *vals = va;
  tmp.k2 = kb; tmp.v2 = vb;
  return hipGetLastError();
}
