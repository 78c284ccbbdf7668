// gs_cull.hip - training and evaluation kernels of the learned cull.
//
// The training set is gathered from instrumented plain frames: projection
// writes each visible Gaussian's 16 features, the raster raises a
// per-Gaussian record to every alpha * T it blends, and k_cull_collect
// appends a hashed sample of (features, log max alpha * T) pairs. Training
// minimises a weighted squared error of the log prediction with Adam: an
// under-prediction of a Gaussian above the cull threshold weighs
// 1 + lambda * (its max alpha * T), since culling it costs image quality in
// proportion to that contribution. The forward and backward products run
// on the matrix cores, one wave per 64 samples.
#include "gs_cull.h"
#include "gs_internal.h"

namespace gs {

namespace {

constexpr int kTrainThreads = 256;
constexpr int kTrainWaves = kTrainThreads / 64;

__device__ __forceinline__ uint32_t hash32(uint32_t x) {  // lowbias32
  x ^= x >> 16;
  x *= 0x7feb352du;
  x ^= x >> 15;
  x *= 0x846ca68bu;
  x ^= x >> 16;
  return x;
}

__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
  return v;
}

// Log target of a recorded max alpha * T, clamped to [kLogFloor, 0].
constexpr float kLogFloor = -16.0f;
__device__ __forceinline__ float log_target(uint32_t bits) {
  float c = gs_b2f(bits);
  float t = gs_logf(c > 1e-7f ? c : 1e-7f);
  return t < kLogFloor ? kLogFloor : (t > 0.0f ? 0.0f : t);
}

// Append the visible Gaussians whose hash falls below `keep` (of 2^32).
__global__ __launch_bounds__(256) void k_cull_collect(
    const float4* __restrict__ pmean, const float* __restrict__ feats,
    const uint32_t* __restrict__ record, uint32_t active, uint32_t keep, uint32_t seed,
    float* __restrict__ ds_feats, float* __restrict__ ds_targets, uint32_t* __restrict__ count,
    uint32_t capacity) {
  uint32_t gid = blockIdx.x * 256 + threadIdx.x;
  bool take = gid < active && pmean[gid].w > 0.0f && hash32(gid ^ seed) < keep;
  uint64_t mask = __ballot(take);
  if (mask == 0ull) return;
  uint32_t lane = __lane_id();
  uint32_t base = 0;
  int leader = __ffsll((long long)mask) - 1;
  if ((int)lane == leader) base = atomicAdd(count, (uint32_t)__popcll(mask));
  base = __shfl(base, leader, 64);
  if (!take) return;
  uint32_t slot = base + (uint32_t)__popcll(mask & ((1ull << lane) - 1ull));
  if (slot >= capacity) return;
#pragma unroll
  for (int k = 0; k < kCullIn; k++)
    ds_feats[(size_t)slot * kCullIn + k] = feats[(size_t)gid * kCullIn + k];
  ds_targets[slot] = log_target(record[gid]);
}

// Sums and sums of squares of every feature over n samples (binary64).
__global__ __launch_bounds__(256) void k_cull_moments(const float* __restrict__ feats, uint32_t n,
                                                      double* __restrict__ out) {
  __shared__ double part[2 * kCullIn];
  if (threadIdx.x < 2 * kCullIn) part[threadIdx.x] = 0.0;
  __syncthreads();
  double s[kCullIn], q[kCullIn];
#pragma unroll
  for (int k = 0; k < kCullIn; k++) s[k] = q[k] = 0.0;
  for (uint32_t j = blockIdx.x * 256 + threadIdx.x; j < n; j += gridDim.x * 256)
#pragma unroll
    for (int k = 0; k < kCullIn; k++) {
      double v = feats[(size_t)j * kCullIn + k];
      s[k] += v;
      q[k] += v * v;
    }
#pragma unroll
  for (int k = 0; k < kCullIn; k++) {
    atomicAdd(&part[k], s[k]);
    atomicAdd(&part[kCullIn + k], q[k]);
  }
  __syncthreads();
  if (threadIdx.x < 2 * kCullIn) atomicAdd(&out[threadIdx.x], part[threadIdx.x]);
}

struct alignas(16) TrainWave {
  uint16_t x[64 * kCullIn];
  uint16_t a1[64 * kRow];
  uint16_t dz[64 * kRow];  // dZ2, later dZ1 (bf16)
  float y[64];
  float dy[64];
};

struct alignas(16) TrainLds {
  CullLdsNet net;
  uint16_t w2t[kCullHidden * kRow];  // W2 transposed: [in][out]
  union {
    TrainWave wave[kTrainWaves];
    float grad[kCullParams];  // block gradient, after the sample loop
  };
};

// One minibatch gradient: samples perm(start + k), k < batch, with
// perm(j) = (a j + b) mod n (a coprime to n). Adds the block's gradient of
// the loss (1/batch) sum w (y - t)^2 to grad and its loss sum to loss.
__global__ __launch_bounds__(kTrainThreads) void k_cull_train(
    const float* __restrict__ feats, const float* __restrict__ targets, uint32_t n,
    uint32_t perm_a, uint32_t perm_b, uint32_t start, uint32_t batch,
    const CullNet* __restrict__ net, float under_weight, float* __restrict__ grad,
    float* __restrict__ loss) {
  __shared__ TrainLds L;
  cull_stage_net(net, L.net);
  for (int t = threadIdx.x; t < kCullHidden * kCullHidden; t += kTrainThreads) {
    int o = t / kCullHidden, in = t % kCullHidden;
    L.w2t[in * kRow + o] = net->w2[t];
  }
  __syncthreads();
  const int wv = threadIdx.x >> 6;
  const int l = __lane_id();
  const int i = l & 15, g = l >> 4;
  TrainWave& W = L.wave[wv];
  const float inv_batch = 1.0f / (float)batch;
  const float log_thr = L.net.log_threshold;
  gs_f32x4 acc_w2[4][4], acc_w1[4];
#pragma unroll
  for (int mt = 0; mt < 4; mt++) {
    acc_w1[mt] = gs_f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int nt = 0; nt < 4; nt++) acc_w2[mt][nt] = gs_f32x4{0.f, 0.f, 0.f, 0.f};
  }
  float gb1[4] = {0.f, 0.f, 0.f, 0.f}, gb2[4] = {0.f, 0.f, 0.f, 0.f};
  float gw3[4] = {0.f, 0.f, 0.f, 0.f};
  float gb3 = 0.0f, loss_acc = 0.0f;
  uint32_t chunks = (batch + 63) / 64;
  for (uint32_t chunk = blockIdx.x * kTrainWaves + wv; chunk < chunks;
       chunk += gridDim.x * kTrainWaves) {
    uint32_t k = chunk * 64 + (uint32_t)l;
    bool valid = k < batch;
    uint32_t idx = 0;
    if (valid) idx = (uint32_t)(((uint64_t)perm_a * (uint64_t)(start + k) + perm_b) % n);
    float f[kCullIn];
#pragma unroll
    for (int q = 0; q < kCullIn; q++) f[q] = valid ? feats[(size_t)idx * kCullIn + q] : 0.0f;
    float t = valid ? targets[idx] : 0.0f;
    gs_f32x4 z2[4][4];
    float y = cull_forward_wave<true>(L.net, W.x, W.a1, W.y, f, z2);
    float d = valid ? y - t : 0.0f;
    float wgt = (t > log_thr && y < t) ? 1.0f + under_weight * gs_expf(t) : 1.0f;
    loss_acc += wgt * d * d;
    float dy = 2.0f * wgt * d * inv_batch;
    gb3 += dy;
    W.dy[l] = dy;
    wave_sync();
    // dZ2 = dy w3 [Z2 > 0]; dw3 += dy A2; db2 += dZ2.
#pragma unroll
    for (int mt = 0; mt < 4; mt++)
#pragma unroll
      for (int nt = 0; nt < 4; nt++) {
        float w3 = L.net.w3[16 * nt + i];
#pragma unroll
        for (int r = 0; r < 4; r++) {
          int s = 16 * mt + 4 * g + r;
          float z = z2[mt][nt][r];
          float dys = W.dy[s];
          float dz = z > 0.0f ? dys * w3 : 0.0f;
          gw3[nt] += z > 0.0f ? dys * z : 0.0f;
          gb2[nt] += dz;
          W.dz[s * kRow + 16 * nt + i] = gs_f2bf(dz);
        }
      }
    wave_sync();
    // dW2 += dZ2^T A1 (M = out, N = in, K = samples).
#pragma unroll
    for (int kt = 0; kt < 4; kt++)
#pragma unroll
      for (int mt = 0; mt < 4; mt++) {
        gs_bf16x4 a = lds4_col(&W.dz[(16 * kt + 4 * g) * kRow + 16 * mt + i], kRow);
#pragma unroll
        for (int nt = 0; nt < 4; nt++) {
          gs_bf16x4 b = lds4_col(&W.a1[(16 * kt + 4 * g) * kRow + 16 * nt + i], kRow);
          acc_w2[mt][nt] = mfma16(a, b, acc_w2[mt][nt]);
        }
      }
    // dA1 = dZ2 W2 (M = samples, N = in, K = out).
    gs_f32x4 da1[4][4];
#pragma unroll
    for (int mt = 0; mt < 4; mt++)
#pragma unroll
      for (int nt = 0; nt < 4; nt++) {
        gs_f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kt = 0; kt < 4; kt++) {
          gs_bf16x4 a = lds4(&W.dz[(16 * mt + i) * kRow + 16 * kt + 4 * g]);
          gs_bf16x4 b = lds4(&L.w2t[(16 * nt + i) * kRow + 16 * kt + 4 * g]);
          acc = mfma16(a, b, acc);
        }
        da1[mt][nt] = acc;
      }
    wave_sync();  // dZ2 is overwritten by dZ1
    // dZ1 = dA1 [A1 > 0]; db1 += dZ1.
#pragma unroll
    for (int mt = 0; mt < 4; mt++)
#pragma unroll
      for (int nt = 0; nt < 4; nt++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
          int s = 16 * mt + 4 * g + r;
          float a = gs_bf2f(W.a1[s * kRow + 16 * nt + i]);
          float dz = a > 0.0f ? da1[mt][nt][r] : 0.0f;
          gb1[nt] += dz;
          W.dz[s * kRow + 16 * nt + i] = gs_f2bf(dz);
        }
    wave_sync();
    // dW1 += dZ1^T X (M = hidden, N = inputs, K = samples).
#pragma unroll
    for (int kt = 0; kt < 4; kt++)
#pragma unroll
      for (int mt = 0; mt < 4; mt++) {
        gs_bf16x4 a = lds4_col(&W.dz[(16 * kt + 4 * g) * kRow + 16 * mt + i], kRow);
        gs_bf16x4 b = lds4_col(&W.x[(16 * kt + 4 * g) * kCullIn + i], kCullIn);
        acc_w1[mt] = mfma16(a, b, acc_w1[mt]);
      }
    wave_sync();  // the next chunk overwrites this wave's rows
  }
  __syncthreads();
  // Block reduction in LDS (over the waves' scratch), then one global add
  // per parameter.
  for (int p = threadIdx.x; p < kCullParams; p += kTrainThreads) L.grad[p] = 0.0f;
  __syncthreads();
#pragma unroll
  for (int mt = 0; mt < 4; mt++) {
#pragma unroll
    for (int nt = 0; nt < 4; nt++)
#pragma unroll
      for (int r = 0; r < 4; r++)
        atomicAdd(&L.grad[kOffW2 + (16 * mt + 4 * g + r) * kCullHidden + 16 * nt + i],
                  acc_w2[mt][nt][r]);
#pragma unroll
    for (int r = 0; r < 4; r++)
      atomicAdd(&L.grad[kOffW1 + (16 * mt + 4 * g + r) * kCullIn + i], acc_w1[mt][r]);
  }
#pragma unroll
  for (int nt = 0; nt < 4; nt++) {
    float b1 = gb1[nt], b2 = gb2[nt], w3 = gw3[nt];
    b1 += __shfl_xor(b1, 16, 64);
    b1 += __shfl_xor(b1, 32, 64);
    b2 += __shfl_xor(b2, 16, 64);
    b2 += __shfl_xor(b2, 32, 64);
    w3 += __shfl_xor(w3, 16, 64);
    w3 += __shfl_xor(w3, 32, 64);
    if (g == 0) {
      atomicAdd(&L.grad[kOffB1 + 16 * nt + i], b1);
      atomicAdd(&L.grad[kOffB2 + 16 * nt + i], b2);
      atomicAdd(&L.grad[kOffW3 + 16 * nt + i], w3);
    }
  }
  gb3 = wave_sum(gb3);
  loss_acc = wave_sum(loss_acc);
  if (l == 0) {
    atomicAdd(&L.grad[kOffB3], gb3);
    atomicAdd(loss, loss_acc);
  }
  __syncthreads();
  for (int p = threadIdx.x; p < kCullParams; p += kTrainThreads) atomicAdd(&grad[p], L.grad[p]);
}

// Adam step on the fp32 parameters; writes the kernels' copy of the network
// (bf16 weights, fp32 biases) and clears the gradient.
__global__ __launch_bounds__(256) void k_cull_adam(float* __restrict__ params,
                                                   float* __restrict__ m, float* __restrict__ v,
                                                   float* __restrict__ grad, float lr,
                                                   float beta1, float beta2, float eps,
                                                   float corr1, float corr2,
                                                   CullNet* __restrict__ net) {
  int p = blockIdx.x * 256 + threadIdx.x;
  if (p >= kCullParams) return;
  float gr = grad[p];
  grad[p] = 0.0f;
  float mm = beta1 * m[p] + (1.0f - beta1) * gr;
  float vv = beta2 * v[p] + (1.0f - beta2) * gr * gr;
  m[p] = mm;
  v[p] = vv;
  float w = params[p] - lr * (mm / corr1) / (sqrtf(vv / corr2) + eps);
  params[p] = w;
  if (p < kOffB1) net->w1[p - kOffW1] = gs_f2bf(w);
  else if (p < kOffW2) net->b1[p - kOffB1] = w;
  else if (p < kOffB2) net->w2[p - kOffW2] = gs_f2bf(w);
  else if (p < kOffW3) net->b2[p - kOffB2] = w;
  else if (p < kOffB3) net->w3[p - kOffW3] = gs_f2bf(w);
  else net->b3 = w;
}

// Predictions for n feature rows.
__global__ __launch_bounds__(256) void k_cull_eval(const float* __restrict__ feats, uint32_t n,
                                                   const CullNet* __restrict__ net,
                                                   float* __restrict__ out) {
  __shared__ CullLdsNet lnet;
  __shared__ __align__(16) uint16_t lx[4][64 * kCullIn];
  cull_stage_net(net, lnet);
  __syncthreads();
  uint32_t j = blockIdx.x * 256 + threadIdx.x;
  float f[kCullIn];
#pragma unroll
  for (int k = 0; k < kCullIn; k++) f[k] = j < n ? feats[(size_t)j * kCullIn + k] : 0.0f;
  float y = cull_infer_wave(lnet, lx[threadIdx.x >> 6], f);  // the projection's forward pass
  if (j < n) out[j] = y;
}

}  // namespace

hipError_t cull_collect(const float4* pmean, const float* feats, const uint32_t* record,
                        uint32_t active, uint32_t keep, uint32_t seed, float* ds_feats,
                        float* ds_targets, uint32_t* count, uint32_t capacity, hipStream_t s) {
  uint32_t blocks = (active + 255) / 256;
  if (blocks == 0) return hipSuccess;
  k_cull_collect<<<blocks, 256, 0, s>>>(pmean, feats, record, active, keep, seed, ds_feats,
                                         ds_targets, count, capacity);
  return hipGetLastError();
}

hipError_t cull_moments(const float* feats, uint32_t n, double* out, hipStream_t s) {
  uint32_t blocks = (n + 255) / 256;
  if (blocks > 1024) blocks = 1024;
  if (blocks == 0) return hipSuccess;
  k_cull_moments<<<blocks, 256, 0, s>>>(feats, n, out);
  return hipGetLastError();
}

hipError_t cull_train_step(const float* feats, const float* targets, uint32_t n, uint32_t perm_a,
                           uint32_t perm_b, uint32_t start, uint32_t batch, const CullNet* net,
                           float under_weight, float* grad, float* loss, hipStream_t s) {
  uint32_t chunks = (batch + 63) / 64;
  uint32_t blocks = (chunks + kTrainWaves - 1) / kTrainWaves;
  if (blocks > 512) blocks = 512;
  if (blocks == 0) return hipSuccess;
  k_cull_train<<<blocks, kTrainThreads, 0, s>>>(feats, targets, n, perm_a, perm_b, start, batch,
                                                net, under_weight, grad, loss);
  return hipGetLastError();
}

hipError_t cull_adam(float* params, float* m, float* v, float* grad, float lr, float beta1,
                     float beta2, float eps, float corr1, float corr2, CullNet* net,
                     hipStream_t s) {
  k_cull_adam<<<(kCullParams + 255) / 256, 256, 0, s>>>(params, m, v, grad, lr, beta1, beta2,
                                                         eps, corr1, corr2, net);
  return hipGetLastError();
}

hipError_t cull_eval(const float* feats, uint32_t n, const CullNet* net, float* out,
                     hipStream_t s) {
  uint32_t blocks = (n + 255) / 256;
  if (blocks == 0) return hipSuccess;
  k_cull_eval<<<blocks, 256, 0, s>>>(feats, n, net, out);
  return hipGetLastError();
}

}  // namespace gs
