// oic_sort.h -- device-wide exclusive scan (tiles of 256 threads x 8 items, tile sums scanned
// recursively) and a stable LSD radix sort of (uint64 key, uint64 value) pairs for gfx950.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

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

// ---------------------------------------------------------------- stable LSD radix sort
//
// One pass per BITS-bit digit (8 to 11 bits, chosen per sort so that few passes cover the key
// bits).  A pass: k_rs_hist counts each tile's digits (4096 keys per tile, LDS histogram); the
// counts, laid out digit-major (count[digit][tile]), are exclusively scanned, which gives every
// (digit, tile) its first output slot; k_rs_scatter then ranks the tile's keys among equal digits in
// input order -- a round of 256 keys at a time: within a wavefront by matching digits with one
// ballot per digit bit, across the 4 wavefronts through LDS counts, across rounds through a running
// count per digit -- and writes each pair to its slot.  Traffic per key and pass: the key read
// twice, the value once, key and value written (40 B).

#define RS_THREADS 256
#define RS_ITEMS   16
#define RS_TILE    (RS_THREADS * RS_ITEMS)
#define RS_WAVES   (RS_THREADS / 64)

template <int BITS>
__global__ void __launch_bounds__(RS_THREADS) k_rs_hist(const uint64_t *__restrict__ keys, uint64_t n, int shift,
                                                        uint32_t *__restrict__ counts, uint32_t ntiles) {
  constexpr int NB = 1 << BITS;
  __shared__ uint32_t h[NB];
  for (int i = threadIdx.x; i < NB; i += RS_THREADS) h[i] = 0;
  __syncthreads();
  const uint64_t base = (uint64_t)blockIdx.x * RS_TILE;
#pragma unroll 4
  for (int r = 0; r < RS_ITEMS; r++) {
    const uint64_t i = base + (uint64_t)r * RS_THREADS + threadIdx.x;
    if (i < n) atomicAdd(&h[(uint32_t)(keys[i] >> shift) & (NB - 1)], 1u);
  }
  __syncthreads();
  for (int b = threadIdx.x; b < NB; b += RS_THREADS) counts[(uint64_t)b * ntiles + blockIdx.x] = h[b];
}

template <int BITS>
__global__ void __launch_bounds__(RS_THREADS) k_rs_scatter(const uint64_t *__restrict__ kin, const uint64_t *__restrict__ vin,
                                                           uint64_t *__restrict__ kout, uint64_t *__restrict__ vout, uint64_t n,
                                                           int shift, const uint64_t *__restrict__ offs, uint32_t ntiles) {
  constexpr int NB = 1 << BITS;
  __shared__ uint64_t next[NB];              // this tile's next output slot per digit
  __shared__ uint32_t wcnt[RS_WAVES][NB];    // keys of each digit in each wavefront, this round
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  for (int b = tid; b < NB; b += RS_THREADS) {
    next[b] = offs[(uint64_t)b * ntiles + blockIdx.x];
#pragma unroll
    for (int x = 0; x < RS_WAVES; x++) wcnt[x][b] = 0;
  }
  __syncthreads();
  const uint64_t below = (1ull << lane) - 1ull;
  const uint64_t base = (uint64_t)blockIdx.x * RS_TILE;
  for (int r = 0; r < RS_ITEMS; r++) {
    const uint64_t i = base + (uint64_t)r * RS_THREADS + tid;
    const bool valid = i < n;
    uint64_t k = 0, v = 0;
    if (valid) { k = kin[i]; v = vin[i]; }
    const uint32_t dg = (uint32_t)(k >> shift) & (NB - 1);
    // lanes of this wavefront holding the same digit
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < BITS; b++) {
      const uint64_t bl = __ballot((dg >> b) & 1u);
      peers &= ((dg >> b) & 1u) ? bl : ~bl;
    }
    const uint32_t rank = (uint32_t)__builtin_popcountll(peers & below);
    const uint32_t cnt = (uint32_t)__builtin_popcountll(peers);
    const bool leader = valid && rank == 0;
    if (leader) wcnt[w][dg] = cnt;
    __syncthreads();
    if (valid) {
      uint64_t pos = next[dg] + rank;
      for (int x = 0; x < w; x++) pos += wcnt[x][dg];
      kout[pos] = k;
      vout[pos] = v;
    }
    __syncthreads();
    if (leader) {
      atomicAdd((unsigned long long *)&next[dg], (unsigned long long)cnt);
      wcnt[w][dg] = 0;
    }
    __syncthreads();
  }
}
