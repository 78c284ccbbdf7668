// gs_cull.h - the learned cull: a 16-64-64-1 MLP that predicts, from a
// projected Gaussian's features, the log of the largest alpha * T it
// contributes to any pixel; Gaussians predicted below a threshold are
// dropped before duplication and sort. No counterpart in the reference (its
// nearest mechanisms are the opacity and subpixel culls of
// ProjectionStage.ts:533, 572-587).
//
// Weights are bf16, products run on the matrix cores
// (v_mfma_f32_16x16x16bf16_1k) with fp32 accumulation, one wave per 64
// Gaussians. Operand layout of that instruction, lane l = 16 g + i
// (g = l / 16, i = l % 16): A (16x16, M x K) lane holds A[i][4g .. 4g+3];
// B (16x16, K x N) lane holds B[4g .. 4g+3][i]; D (16x16) lane holds
// D[4g + r][i], r = 0..3.
#ifndef GS_CULL_H
#define GS_CULL_H

#include <stddef.h>
#include <stdint.h>

#include "gs_math.h"

namespace gs {

constexpr int kCullIn = 16;
constexpr int kCullHidden = 64;
// fp32 parameter vector: W1 [64][16], b1 [64], W2 [64][64], b2 [64], w3 [64], b3.
constexpr int kOffW1 = 0;
constexpr int kOffB1 = kOffW1 + kCullHidden * kCullIn;
constexpr int kOffW2 = kOffB1 + kCullHidden;
constexpr int kOffB2 = kOffW2 + kCullHidden * kCullHidden;
constexpr int kOffW3 = kOffB2 + kCullHidden;
constexpr int kOffB3 = kOffW3 + kCullHidden;
constexpr int kCullParams = kOffB3 + 1;  // 5313
// Padded row length (bf16) of the 64-wide LDS matrices.
constexpr int kRow = 72;

// The network as the kernels read it.
struct CullNet {
  uint16_t w1[kCullHidden * kCullIn];      // bf16 [hidden][input]
  uint16_t w2[kCullHidden * kCullHidden];  // bf16 [out][in]
  uint16_t w3[kCullHidden];                // bf16
  float b1[kCullHidden], b2[kCullHidden];
  float b3;
  float log_threshold;  // cull when the prediction is below this
  float mu[kCullIn], inv_sigma[kCullIn];  // feature standardisation
};
static_assert(offsetof(CullNet, w2) % 16 == 0 && offsetof(CullNet, w1) % 16 == 0,
              "the weight matrices are staged 16 bytes at a time");

// binary32 -> bfloat16, round to nearest even (NaN -> quiet NaN).
GS_HD inline uint16_t gs_f2bf(float f) {
  uint32_t u = gs_f2b(f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return 0x7fc0;
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
GS_HD inline float gs_bf2f(uint16_t h) { return gs_b2f((uint32_t)h << 16); }

// The 16 features of a projected Gaussian: footprint (log pixel radius,
// log conic diagonal, conic correlation), opacity, the position along the
// viewing ray (object-space position . view direction, a proxy for how
// deep in the scene the Gaussian sits), log view depth, the object-space
// view direction, the object-space position, the log2 scale of the
// strongest higher-band SH coefficient (its rgb8e8 exponent minus the bias)
// and the normalised screen position (inv_vw, inv_vh: 1 / viewport size).
// Every log argument is positive and normal.
GS_HD inline void cull_features(float f[kCullIn], float rx, float ry, const float conic[3],
                                float opacity, float depth, const float dir[3],
                                const float pos[3], float sh_top, float cx, float cy,
                                float inv_vw, float inv_vh) {
  f[0] = gs_logf_pos(rx > ry ? rx : ry);
  f[1] = gs_logf_pos(conic[0]);
  f[2] = conic[1] / gs_sqrtf(conic[0] * conic[2]);
  f[3] = gs_logf_pos(conic[2]);
  f[4] = opacity;
  f[5] = pos[0] * dir[0] + pos[1] * dir[1] + pos[2] * dir[2];
  f[6] = gs_logf_pos(depth);
  f[7] = dir[0];
  f[8] = dir[1];
  f[9] = dir[2];
  f[10] = pos[0];
  f[11] = pos[1];
  f[12] = pos[2];
  f[13] = sh_top;
  f[14] = cx * inv_vw;
  f[15] = cy * inv_vh;
}

#ifdef __HIPCC__

typedef short gs_bf16x4 __attribute__((ext_vector_type(4)));
typedef float gs_f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ gs_f32x4 mfma16(gs_bf16x4 a, gs_bf16x4 b, gs_f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}
typedef __bf16 gs_bf16x8 __attribute__((ext_vector_type(8)));
typedef short gs_i16x8 __attribute__((ext_vector_type(8)));
// v_mfma_f32_16x16x32_bf16 (gfx950) on two 16x16x16 operand pairs: K slots
// 0..3 of lane group g are the first pair's, 4..7 the second's. A and B
// put the same K index in the same slot of a lane group, so any assignment
// of K to slots that A and B share gives the same contraction:
// D = C + a0 b0 + a1 b1 over the 32 K indices of both pairs.
__device__ __forceinline__ gs_f32x4 mfma32(gs_bf16x4 a0, gs_bf16x4 a1, gs_bf16x4 b0,
                                           gs_bf16x4 b1, gs_f32x4 c) {
  const gs_i16x8 a = {a0.x, a0.y, a0.z, a0.w, a1.x, a1.y, a1.z, a1.w};
  const gs_i16x8 b = {b0.x, b0.y, b0.z, b0.w, b1.x, b1.y, b1.z, b1.w};
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(gs_bf16x8, a),
                                                 __builtin_bit_cast(gs_bf16x8, b), c, 0, 0, 0);
}

// Wave-scope ordering of LDS writes before reads by other lanes of the
// same wave.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Four consecutive bf16 of an LDS row.
__device__ __forceinline__ gs_bf16x4 lds4(const uint16_t* p) {
  return *reinterpret_cast<const gs_bf16x4*>(p);
}
// Four bf16 a row apart (a column slice).
__device__ __forceinline__ gs_bf16x4 lds4_col(const uint16_t* p, int stride) {
  gs_bf16x4 v;
  v.x = (short)p[0];
  v.y = (short)p[stride];
  v.z = (short)p[2 * stride];
  v.w = (short)p[3 * stride];
  return v;
}

// The network's weights staged in LDS for a block (16-byte aligned: the
// matrix fragments are read 8 bytes at a time).
struct alignas(16) CullLdsNet {
  uint16_t w1[kCullHidden * kCullIn];
  uint16_t w2[kCullHidden * kRow];  // [out][in], rows padded
  float w3[kCullHidden];
  float b1[kCullHidden], b2[kCullHidden];
  float b3, log_threshold;
  float mu[kCullIn], inv_sigma[kCullIn];
};

// Cooperative copy of the network into LDS (all threads of the block), the
// weight matrices 16 bytes at a time (the CullNet comes from a 256-byte
// aligned allocation; w1 and w2 sit at 16-byte multiples).
__device__ inline void cull_stage_net(const CullNet* __restrict__ net, CullLdsNet& L) {
  const uint4* w1 = reinterpret_cast<const uint4*>(net->w1);
  const uint4* w2 = reinterpret_cast<const uint4*>(net->w2);
  constexpr int kW1Chunks = kCullHidden * kCullIn / 8, kW2Chunks = kCullHidden * kCullHidden / 8;
  for (int t = threadIdx.x; t < kW1Chunks; t += blockDim.x)
    reinterpret_cast<uint4*>(L.w1)[t] = w1[t];
  for (int t = threadIdx.x; t < kW2Chunks; t += blockDim.x)  // 8 chunks per row
    *reinterpret_cast<uint4*>(&L.w2[(t >> 3) * kRow + (t & 7) * 8]) = w2[t];
  for (int t = threadIdx.x; t < kCullHidden; t += blockDim.x) {
    L.w3[t] = gs_bf2f(net->w3[t]);
    L.b1[t] = net->b1[t];
    L.b2[t] = net->b2[t];
  }
  for (int t = threadIdx.x; t < kCullIn; t += blockDim.x) {
    L.mu[t] = net->mu[t];
    L.inv_sigma[t] = net->inv_sigma[t];
  }
  if (threadIdx.x == 0) {
    L.b3 = net->b3;
    L.log_threshold = net->log_threshold;
  }
}

// Forward pass of one wave: lane l holds the raw features of sample l
// (zeros for idle lanes); returns that sample's prediction. x (64 x 16) and
// a1 (64 x kRow) are the wave's LDS rows for the bf16 inputs and hidden
// activations (they may overlap: x is dead once layer 1 has run); y holds
// 64 outputs. With kKeepZ2 the layer-2 pre-activations (bias added) are
// returned in z2, in the D layout above as [mt][nt] tiles.
template <bool kKeepZ2>
__device__ inline float cull_forward_wave(const CullLdsNet& N, uint16_t* x, uint16_t* a1,
                                          float* y, const float f[kCullIn],
                                          gs_f32x4 (*z2)[4]) {
  const int l = __lane_id();
  const int i = l & 15, g = l >> 4;
  // Standardised inputs, bf16, row l.
#pragma unroll
  for (int k = 0; k < kCullIn; k++) x[l * kCullIn + k] = gs_f2bf((f[k] - N.mu[k]) * N.inv_sigma[k]);
  wave_sync();
  // Layer 1: Z1 = X W1^T (64 x 64), K = 16. The A fragments of all four
  // row tiles are taken into registers first (the x rows may share bytes
  // with a1), then each row tile's products go straight to a1.
  gs_bf16x4 xa[4];
#pragma unroll
  for (int mt = 0; mt < 4; mt++) xa[mt] = lds4(&x[(16 * mt + i) * kCullIn + 4 * g]);
  wave_sync();
#pragma unroll
  for (int mt = 0; mt < 4; mt++)
#pragma unroll
    for (int nt = 0; nt < 4; nt++) {
      gs_bf16x4 b = lds4(&N.w1[(16 * nt + i) * kCullIn + 4 * g]);
      gs_f32x4 z = mfma16(xa[mt], b, gs_f32x4{0.f, 0.f, 0.f, 0.f});
      float bias = N.b1[16 * nt + i];
#pragma unroll
      for (int r = 0; r < 4; r++) {
        float v = z[r] + bias;
        a1[(16 * mt + 4 * g + r) * kRow + 16 * nt + i] = gs_f2bf(v > 0.0f ? v : 0.0f);
      }
    }
  wave_sync();
  // Layer 2: Z2 = A1 W2^T (64 x 64), K = 64; layer 3 folded in.
  float part[4][4];
#pragma unroll
  for (int mt = 0; mt < 4; mt++)
#pragma unroll
    for (int r = 0; r < 4; r++) part[mt][r] = 0.0f;
#pragma unroll
  for (int mt = 0; mt < 4; mt++)
#pragma unroll
    for (int nt = 0; nt < 4; nt++) {
      gs_f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int kt = 0; kt < 4; kt++) {
        gs_bf16x4 a = lds4(&a1[(16 * mt + i) * kRow + 16 * kt + 4 * g]);
        gs_bf16x4 b = lds4(&N.w2[(16 * nt + i) * kRow + 16 * kt + 4 * g]);
        acc = mfma16(a, b, acc);
      }
      float bias = N.b2[16 * nt + i], w3 = N.w3[16 * nt + i];
#pragma unroll
      for (int r = 0; r < 4; r++) {
        float v = acc[r] + bias;
        if (kKeepZ2) z2[mt][nt][r] = v;
        part[mt][r] += (v > 0.0f ? v : 0.0f) * w3;
      }
    }
  // Sum over the 16 hidden units each group of 16 lanes holds.
#pragma unroll
  for (int mt = 0; mt < 4; mt++)
#pragma unroll
    for (int r = 0; r < 4; r++) {
      float v = part[mt][r];
      v += __shfl_xor(v, 1, 64);
      v += __shfl_xor(v, 2, 64);
      v += __shfl_xor(v, 4, 64);
      v += __shfl_xor(v, 8, 64);
      if (i == 0) y[16 * mt + 4 * g + r] = v + N.b3;
    }
  wave_sync();
  return y[l];
}

typedef __bf16 gs_bf16x2 __attribute__((ext_vector_type(2)));
typedef float gs_f32x2 __attribute__((ext_vector_type(2)));
typedef uint32_t gs_u32x2 __attribute__((ext_vector_type(2)));

// Two binary32 -> bfloat16 (round to nearest even), `lo` in the low half:
// one v_cvt_pk_bf16_f32 on gfx950. Equal to gs_f2bf for every non-NaN input.
__device__ __forceinline__ uint32_t bf16_pack2(float lo, float hi) {
  return __builtin_bit_cast(uint32_t, __builtin_convertvector(gs_f32x2{lo, hi}, gs_bf16x2));
}
typedef short gs_i16x2 __attribute__((ext_vector_type(2)));

// ReLU on bit patterns: a negative binary32 or bfloat16 has its sign bit
// set, so as a signed integer it is below zero (one integer max instead of
// a NaN-canonicalising float max pair); -0 becomes +0 as with x > 0 ? x : 0.
__device__ __forceinline__ float relu_bits(float v) {
  int b = __builtin_bit_cast(int, v);
  return __builtin_bit_cast(float, b > 0 ? b : 0);
}
__device__ __forceinline__ uint32_t relu_bf16x2(uint32_t v) {
  return __builtin_bit_cast(uint32_t, __builtin_elementwise_max(__builtin_bit_cast(gs_i16x2, v),
                                                                gs_i16x2{0, 0}));
}
// bf16(relu(z)) for the four rows of a D fragment (rounding first gives the
// same bits: rounding preserves the sign).
__device__ __forceinline__ gs_bf16x4 bf16x4_relu(gs_f32x4 z) {
  return __builtin_bit_cast(gs_bf16x4, gs_u32x2{relu_bf16x2(bf16_pack2(z[0], z[1])),
                                                relu_bf16x2(bf16_pack2(z[2], z[3]))});
}

// Inference forward pass of one wave, computed transposed so that no
// activation goes through LDS: Z1^T = W1 X^T (M = hidden, N = samples, K =
// inputs) leaves each lane holding hidden units 16mt + 4g + r of sample
// 16nt + i, which is exactly the B operand of Z2^T = W2 A1^T for K tile
// kt = mt. The biases enter as the accumulators' initial values (a lane's
// four rows are four consecutive hidden units), and layer 3 sums over the
// hidden units in registers and across the four lane groups. Lane l holds
// the raw features of sample l (zeros for idle lanes) and gets that
// sample's prediction; xs is the wave's 64 x 16 bf16 LDS rows for the
// standardised inputs. Same bf16 operands as cull_forward_wave, fp32
// accumulation (the bias is added inside the accumulation rather than
// after it); layer 2 takes two K = 32 products (v_mfma_f32_16x16x32_bf16)
// per output tile instead of four K = 16 ones.
__device__ inline float cull_infer_wave(const CullLdsNet& N, uint16_t* xs,
                                        const float f[kCullIn]) {
  const int l = __lane_id();
  const int i = l & 15, g = l >> 4;
  uint32_t packed[kCullIn / 2];
#pragma unroll
  for (int k = 0; k < kCullIn / 2; k++)
    packed[k] = bf16_pack2((f[2 * k] - N.mu[2 * k]) * N.inv_sigma[2 * k],
                           (f[2 * k + 1] - N.mu[2 * k + 1]) * N.inv_sigma[2 * k + 1]);
  uint4* row = reinterpret_cast<uint4*>(&xs[l * kCullIn]);
  row[0] = make_uint4(packed[0], packed[1], packed[2], packed[3]);
  row[1] = make_uint4(packed[4], packed[5], packed[6], packed[7]);
  wave_sync();
  gs_bf16x4 xb[4];  // B operand: X[16 nt + i][4g .. 4g + 3]
#pragma unroll
  for (int nt = 0; nt < 4; nt++) xb[nt] = lds4(&xs[(16 * nt + i) * kCullIn + 4 * g]);
  wave_sync();  // the rows may be rewritten by the next call
  // Layer 1 (transposed), ReLU, bf16: the B operands of layer 2.
  gs_bf16x4 a1t[4][4];  // [kt][nt]
#pragma unroll
  for (int mt = 0; mt < 4; mt++) {
    gs_bf16x4 a = lds4(&N.w1[(16 * mt + i) * kCullIn + 4 * g]);
    float4 b1 = *reinterpret_cast<const float4*>(&N.b1[16 * mt + 4 * g]);
#pragma unroll
    for (int nt = 0; nt < 4; nt++)
      a1t[mt][nt] = bf16x4_relu(mfma16(a, xb[nt], gs_f32x4{b1.x, b1.y, b1.z, b1.w}));
  }
  // Layer 2 (transposed) with layer 3 folded in.
  float yp[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int mt = 0; mt < 4; mt++) {
    gs_bf16x4 w[4];
#pragma unroll
    for (int kt = 0; kt < 4; kt++) w[kt] = lds4(&N.w2[(16 * mt + i) * kRow + 16 * kt + 4 * g]);
    float4 b2 = *reinterpret_cast<const float4*>(&N.b2[16 * mt + 4 * g]);
    float4 w3 = *reinterpret_cast<const float4*>(&N.w3[16 * mt + 4 * g]);
#pragma unroll
    for (int nt = 0; nt < 4; nt++) {
      gs_f32x4 acc = {b2.x, b2.y, b2.z, b2.w};
      acc = mfma32(w[0], w[1], a1t[0][nt], a1t[1][nt], acc);
      acc = mfma32(w[2], w[3], a1t[2][nt], a1t[3][nt], acc);
      yp[nt] = __builtin_fmaf(relu_bits(acc[0]), w3.x, yp[nt]);
      yp[nt] = __builtin_fmaf(relu_bits(acc[1]), w3.y, yp[nt]);
      yp[nt] = __builtin_fmaf(relu_bits(acc[2]), w3.z, yp[nt]);
      yp[nt] = __builtin_fmaf(relu_bits(acc[3]), w3.w, yp[nt]);
    }
  }
#pragma unroll
  for (int nt = 0; nt < 4; nt++) {
    yp[nt] += __shfl_xor(yp[nt], 16, 64);
    yp[nt] += __shfl_xor(yp[nt], 32, 64);
  }
  // Sample l = 16 g + i is column i of N tile nt = g.
  float y = g == 0 ? yp[0] : (g == 1 ? yp[1] : (g == 2 ? yp[2] : yp[3]));
  return y + N.b3;
}

#endif  // __HIPCC__

}  // namespace gs

#endif
