// oic_mlp.h -- the learned cost model: a 12 -> 64 -> 64 -> 2 ReLU MLP in FP32 on gfx950.
//
// Parameter vector (OIC_NPARAM floats): W1[64][12] b1[64] W2[64][64] b2[64] W3[2][64] b3[2].
// Inputs are standardised features z = (x - mu) / sd; outputs are log(1 + DP cells) and the
// exit diagonal / 1000.  Loss per sample: 0.5 (y0 - t0)^2 + 0.5 w_exit (y1 - t1)^2, averaged.
//
// Kernels: the forward pass runs one thread per sample with the weights in LDS.  The gradient
// kernel runs persistent waves over 64-sample chunks with the chunk's activations in LDS; each
// lane accumulates the gradient rows of one hidden unit in registers, and every wave adds its
// accumulators into the global gradient once at the end.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define MLP_IN  12
#define MLP_H   64
#define MLP_OUT 2
#define MLP_OFF_W1 0
#define MLP_OFF_B1 (MLP_OFF_W1 + MLP_H * MLP_IN)
#define MLP_OFF_W2 (MLP_OFF_B1 + MLP_H)
#define MLP_OFF_B2 (MLP_OFF_W2 + MLP_H * MLP_H)
#define MLP_OFF_W3 (MLP_OFF_B2 + MLP_H)
#define MLP_OFF_B3 (MLP_OFF_W3 + MLP_OUT * MLP_H)
#define MLP_NPARAM (MLP_OFF_B3 + MLP_OUT)

__device__ __forceinline__ float wave_sum_f(float v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

// y[i] = MLP(x[i]) with x standardised by (mu, sd) when mu != nullptr
__global__ void __launch_bounds__(256) k_mlp_forward(const float *__restrict__ X, uint64_t n, const float *__restrict__ P,
                                                     const float *__restrict__ mu, const float *__restrict__ sd,
                                                     float *__restrict__ Y) {
  __shared__ float w[MLP_NPARAM];
  for (int i = threadIdx.x; i < MLP_NPARAM; i += blockDim.x) w[i] = P[i];
  __syncthreads();
  const uint64_t s = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  float z[MLP_IN];
#pragma unroll
  for (int i = 0; i < MLP_IN; i++) {
    const float x = X[s * MLP_IN + i];
    z[i] = mu ? (x - mu[i]) / sd[i] : x;
  }
  float h1[MLP_H];
#pragma unroll
  for (int j = 0; j < MLP_H; j++) {
    float a = w[MLP_OFF_B1 + j];
#pragma unroll
    for (int i = 0; i < MLP_IN; i++) a += w[MLP_OFF_W1 + j * MLP_IN + i] * z[i];
    h1[j] = a > 0.f ? a : 0.f;
  }
  float y0 = w[MLP_OFF_B3], y1 = w[MLP_OFF_B3 + 1];
  for (int k = 0; k < MLP_H; k++) {
    float a = w[MLP_OFF_B2 + k];
#pragma unroll
    for (int j = 0; j < MLP_H; j++) a += w[MLP_OFF_W2 + k * MLP_H + j] * h1[j];
    const float h2 = a > 0.f ? a : 0.f;
    y0 += w[MLP_OFF_W3 + k] * h2;
    y1 += w[MLP_OFF_W3 + MLP_H + k] * h2;
  }
  Y[s * MLP_OUT] = y0;
  Y[s * MLP_OUT + 1] = y1;
}

// Mean loss and gradient over samples idx[0..n) (or 0..n when idx is null); Z is standardised.
// grad must be zeroed before the launch; loss_sum receives the sum of per-sample losses.
// One wave per block; per 64-sample chunk the activations live in LDS columns (one per sample):
//   forward (lane = sample)      H1[j][s], H2[k][s]
//   dW3 (lane = k), then G2 = W3^T g3 masked by H2 > 0, written over H2
//   dW2 (lane = k), then G1 = W2^T G2 masked by H1 > 0, written over H1
//   dW1 (lane = j)
#define MLP_LD (MLP_H + 1)   // padded LDS row: lanes reading row k of a column block avoid bank conflicts
__global__ void __launch_bounds__(64) k_mlp_grad(const float *__restrict__ Z, const float *__restrict__ Tg,
                                                 const uint32_t *__restrict__ idx, uint64_t n, const float *__restrict__ P,
                                                 float w_exit, float inv_n, float *__restrict__ grad, float *__restrict__ loss_sum) {
  __shared__ float w[MLP_NPARAM];
  __shared__ float H1[MLP_H * MLP_LD];    // [unit][sample]
  __shared__ float H2[MLP_H * MLP_LD];
  __shared__ float Zs[MLP_IN * MLP_LD];
  __shared__ float G3[MLP_OUT * MLP_LD];
  const int lane = threadIdx.x;
  for (int i = lane; i < MLP_NPARAM; i += 64) w[i] = P[i];
  __syncthreads();
  // per-lane accumulators: lane owns unit `lane` of each layer
  float aW1[MLP_IN], aW2[MLP_H], aW3[MLP_OUT];
  float ab1 = 0.f, ab2 = 0.f, ab3 = 0.f;
#pragma unroll
  for (int i = 0; i < MLP_IN; i++) aW1[i] = 0.f;
#pragma unroll
  for (int j = 0; j < MLP_H; j++) aW2[j] = 0.f;
  aW3[0] = aW3[1] = 0.f;
  float lsum = 0.f;

  for (uint64_t c0 = (uint64_t)blockIdx.x * 64; c0 < n; c0 += (uint64_t)gridDim.x * 64) {
    const uint64_t si = c0 + lane;
    const bool valid = si < n;
    const uint64_t s = valid ? (idx ? idx[si] : si) : 0;
    // ---- forward, lane = sample
    for (int i = 0; i < MLP_IN; i++) Zs[i * MLP_LD + lane] = valid ? Z[s * MLP_IN + i] : 0.f;
    for (int j = 0; j < MLP_H; j++) {
      float a = w[MLP_OFF_B1 + j];
#pragma unroll
      for (int i = 0; i < MLP_IN; i++) a += w[MLP_OFF_W1 + j * MLP_IN + i] * Zs[i * MLP_LD + lane];
      H1[j * MLP_LD + lane] = a > 0.f ? a : 0.f;
    }
    float y0 = w[MLP_OFF_B3], y1 = w[MLP_OFF_B3 + 1];
    for (int k = 0; k < MLP_H; k++) {
      float a = w[MLP_OFF_B2 + k];
#pragma unroll 16
      for (int j = 0; j < MLP_H; j++) a += w[MLP_OFF_W2 + k * MLP_H + j] * H1[j * MLP_LD + lane];
      const float h = a > 0.f ? a : 0.f;
      H2[k * MLP_LD + lane] = h;
      y0 += w[MLP_OFF_W3 + k] * h;
      y1 += w[MLP_OFF_W3 + MLP_H + k] * h;
    }
    const float t0 = valid ? Tg[s * MLP_OUT] : 0.f, t1 = valid ? Tg[s * MLP_OUT + 1] : 0.f;
    const float e0 = y0 - t0, e1 = y1 - t1;
    if (valid) lsum += 0.5f * e0 * e0 + 0.5f * w_exit * e1 * e1;
    G3[lane] = valid ? e0 * inv_n : 0.f;
    G3[MLP_LD + lane] = valid ? w_exit * e1 * inv_n : 0.f;
    __syncthreads();
    // ---- dW3[o][k], db3 (lane = k)
    for (int q = 0; q < 64; q++) {
      const float g0 = G3[q], g1 = G3[MLP_LD + q], h = H2[lane * MLP_LD + q];
      aW3[0] += g0 * h;
      aW3[1] += g1 * h;
      if (lane < MLP_OUT) ab3 += (lane == 0) ? g0 : g1;
    }
    __syncthreads();
    // ---- G2 over H2 (lane = sample)
    {
      const float g0 = G3[lane], g1 = G3[MLP_LD + lane];
      for (int k = 0; k < MLP_H; k++) {
        const float h = H2[k * MLP_LD + lane];
        H2[k * MLP_LD + lane] = h > 0.f ? (w[MLP_OFF_W3 + k] * g0 + w[MLP_OFF_W3 + MLP_H + k] * g1) : 0.f;
      }
    }
    __syncthreads();
    // ---- dW2[k][j], db2 (lane = k)
    for (int q = 0; q < 64; q++) {
      const float g = H2[lane * MLP_LD + q];
      ab2 += g;
#pragma unroll
      for (int j = 0; j < MLP_H; j++) aW2[j] += g * H1[j * MLP_LD + q];
    }
    __syncthreads();
    // ---- G1 = W2^T G2 masked by H1 > 0, over H1 (lane = sample)
    for (int j = 0; j < MLP_H; j++) {
      float a = 0.f;
#pragma unroll 16
      for (int k = 0; k < MLP_H; k++) a += w[MLP_OFF_W2 + k * MLP_H + j] * H2[k * MLP_LD + lane];
      const float h = H1[j * MLP_LD + lane];
      H1[j * MLP_LD + lane] = h > 0.f ? a : 0.f;
    }
    __syncthreads();
    // ---- dW1[j][i], db1 (lane = j)
    for (int q = 0; q < 64; q++) {
      const float g = H1[lane * MLP_LD + q];
      ab1 += g;
#pragma unroll
      for (int i = 0; i < MLP_IN; i++) aW1[i] += g * Zs[i * MLP_LD + q];
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < MLP_IN; i++) atomicAdd(&grad[MLP_OFF_W1 + lane * MLP_IN + i], aW1[i]);
  atomicAdd(&grad[MLP_OFF_B1 + lane], ab1);
#pragma unroll
  for (int j = 0; j < MLP_H; j++) atomicAdd(&grad[MLP_OFF_W2 + lane * MLP_H + j], aW2[j]);
  atomicAdd(&grad[MLP_OFF_B2 + lane], ab2);
  atomicAdd(&grad[MLP_OFF_W3 + lane], aW3[0]);
  atomicAdd(&grad[MLP_OFF_W3 + MLP_H + lane], aW3[1]);
  if (lane < MLP_OUT) atomicAdd(&grad[MLP_OFF_B3 + lane], ab3);
  lsum = wave_sum_f(lsum);
  if (lane == 0 && loss_sum) atomicAdd(loss_sum, lsum);
}

__global__ void k_fill_f(float *__restrict__ p, uint64_t n, float v) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// Adam (Kingma & Ba 2015) with bias correction
__global__ void k_adam(float *__restrict__ P, const float *__restrict__ g, float *__restrict__ m, float *__restrict__ v, int n,
                       float lr, float b1, float b2, float eps, float c1, float c2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float mh = mi / c1, vh = vi / c2;
  P[i] -= lr * mh / (sqrtf(vh) + eps);
}

// the 12 features of a candidate pair: log1p of S length, T length, seed hits, match nodes;
// min and max node diagonal / 1000; log1p diagonal spread; log1p max k-mer multiplicity;
// consistent flag; orientation; log1p shorter read length; log1p longest node length
__global__ void k_pair_features(const PairRec *__restrict__ pairs, uint64_t np, const int32_t *__restrict__ len,
                                float *__restrict__ X) {
  const uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  const PairRec pr = pairs[p];
  const double sl = len[pr.s_id], tl = len[pr.t_id];
  float *x = X + p * MLP_IN;
  x[0] = (float)log1p(sl);
  x[1] = (float)log1p(tl);
  x[2] = (float)log1p((double)(pr.he - pr.hb));
  x[3] = (float)log1p((double)pr.nnodes);
  x[4] = (float)(pr.dmin / 1000.0);
  x[5] = (float)(pr.dmax / 1000.0);
  x[6] = (float)log1p((double)pr.dmax - (double)pr.dmin);
  x[7] = (float)log1p((double)pr.max_mult);
  x[8] = (float)pr.consistent;
  x[9] = (float)pr.dir;
  x[10] = (float)log1p(sl < tl ? sl : tl);
  x[11] = (float)log1p((double)pr.longest);
}
