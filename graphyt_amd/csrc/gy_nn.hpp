// gy_nn.hpp — forward pass of the return-map network for one wavefront (gfx950 MFMA).
//
// Network: 16 inputs (GY_MLP_IN used) -> 64 -> 64 -> 64 -> 16 outputs (GY_MLP_OUT used),
// SiLU on the hidden layers; weights W[k][n] (K x N per layer) and FP32 biases in the
// parameter layout of gy_mlp.hip.
//
// One wavefront evaluates the samples of its 64 lanes as four blocks of 16. Each layer
// computes H^T = W^T X^T on the matrix cores (v_mfma_f32_16x16x16bf16_1k, BF16 operands,
// FP32 accumulation) with the samples along N. The result layout of that instruction
// (lane l holds rows 4(l/16)+j, j = 0..3, of column l%16) is the layout of its B operand
// (lane l supplies k = 4(l/16)+j of column l%16), so the SiLU of a layer's result, packed
// to BF16 in the lane that computed it, is the next layer's B operand without any data
// movement. W^T is the A operand, read from LDS in fragment order: one 8-byte read per
// lane per 16x16 tile, tile (mt, kt) of a layer holding W[16kt + 4(l/16) + j][16mt + l%16]
// in lane l. The inputs enter each block by a cross-lane transpose (ds_bpermute) and the
// output leaves by one. The biases initialise the accumulators.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/graphyt_amd.h"

namespace gy {

constexpr int NN_KIN = 16, NN_HID = 64, NN_KOUT = 16;
// A-fragment tiles: layer 1 (4 x 1), layers 2 and 3 (4 x 4 each), layer 4 (1 x 4)
constexpr int NN_T1 = 0, NN_T2 = 4, NN_T3 = 20, NN_T4 = 36, NN_TILES = 40;
constexpr int NN_NBIAS = 3 * NN_HID + NN_KOUT;
// normalisation block: input mean, 1/std, trained range [lo, hi] of the normalised
// inputs (16 each), output scale and offset (network output 0 -> scale * y + offset),
// trained range [ylo, yhi] of the scaled output
constexpr int NN_NORM_MEAN = 0, NN_NORM_ISTD = 16, NN_NORM_LO = 32, NN_NORM_HI = 48, NN_NORM_OUT = 64;
constexpr int NN_NORM_YLO = 66, NN_NORM_YHI = 67, NN_NORM_N = 68;
static_assert(GY_MLP_IN <= NN_KIN && GY_MLP_OUT <= 4, "padded widths of the network");

// Network as the step kernels read it (device pointers; all null when none is set).
struct NetDev {
  const uint2 *frag;   // NN_TILES * 64 A fragments (BF16)
  const float *bias;   // NN_NBIAS
  const float *norm;   // NN_NORM_N
};

namespace nn {
typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));
typedef __bf16 bf2 __attribute__((ext_vector_type(2)));
typedef float f2 __attribute__((ext_vector_type(2)));

// two floats to a BF16 pair (round to nearest even; v_cvt_pk_bf16_f32), a in the low half
__device__ __forceinline__ uint32_t pack2(float a, float b) {
  const f2 v = {a, b};
  const bf2 h = __builtin_convertvector(v, bf2);
  uint32_t u;
  __builtin_memcpy(&u, &h, 4);
  return u;
}
__device__ __forceinline__ v4s as_v4s(uint32_t lo, uint32_t hi) {
  const uint32_t t[2] = {lo, hi};
  v4s r;
  __builtin_memcpy(&r, t, 8);
  return r;
}
// z * sigmoid(z) with the hardware exp2 and reciprocal (v_exp_f32, v_rcp_f32): exact to
// a few FP32 ulps, far below the BF16 rounding of the result
__device__ __forceinline__ float silu(float z) {
  return z * __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(-1.44269504088896341f * z));
}
__device__ __forceinline__ v4f mfma(uint2 a, uint32_t b0, uint32_t b1, v4f c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(as_v4s(a.x, a.y), as_v4s(b0, b1), c, 0, 0, 0);
}
__device__ __forceinline__ v4f bias4(const float *b) {
  const float4 q = *reinterpret_cast<const float4 *>(b);
  return v4f{q.x, q.y, q.z, q.w};
}
}  // namespace nn

namespace nn {
// Blocks t0 .. t0 + NB - 1 of 16 samples through the network; y receives output 0 of this
// lane's sample when its block is among them. NB > 1 interleaves independent blocks (more
// instruction-level parallelism, more registers).
template <int NB>
__device__ __forceinline__ void blocks_forward(const uint2 *sf, const float *sb, const uint32_t x[8], int t0,
                                               float &y) {
  const int lane = threadIdx.x & 63, g = lane >> 4, r = lane & 15;
  uint32_t act[NB][8];  // B operand of the next layer: act[b][2kt], act[b][2kt+1]
#pragma unroll
  for (int b = 0; b < NB; b++) {
    // B operand of layer 1: inputs 4g .. 4g+3 of sample 16t + r (lane 16t + r)
    uint32_t w[8];
#pragma unroll
    for (int k = 0; k < 8; k++) w[k] = __shfl(x[k], 16 * (t0 + b) + r, 64);
    act[b][0] = g == 0 ? w[0] : g == 1 ? w[2] : g == 2 ? w[4] : w[6];
    act[b][1] = g == 0 ? w[1] : g == 1 ? w[3] : g == 2 ? w[5] : w[7];
  }
  {
    v4f c[NB][4];
#pragma unroll
    for (int mt = 0; mt < 4; mt++) {
      const v4f bi = bias4(sb + 16 * mt + 4 * g);
      const uint2 a = sf[(NN_T1 + mt) * 64 + lane];
#pragma unroll
      for (int b = 0; b < NB; b++) c[b][mt] = mfma(a, act[b][0], act[b][1], bi);
    }
#pragma unroll
    for (int b = 0; b < NB; b++)
#pragma unroll
      for (int mt = 0; mt < 4; mt++) {
        act[b][2 * mt] = pack2(silu(c[b][mt][0]), silu(c[b][mt][1]));
        act[b][2 * mt + 1] = pack2(silu(c[b][mt][2]), silu(c[b][mt][3]));
      }
  }
#pragma unroll
  for (int L = 0; L < 2; L++) {
    const int T0 = L == 0 ? NN_T2 : NN_T3;
    const float *bb = sb + NN_HID * (L + 1);
    v4f c[NB][4];
#pragma unroll
    for (int mt = 0; mt < 4; mt++) {
      const v4f bi = bias4(bb + 16 * mt + 4 * g);
#pragma unroll
      for (int b = 0; b < NB; b++) c[b][mt] = bi;
#pragma unroll
      for (int kt = 0; kt < 4; kt++) {
        const uint2 a = sf[(T0 + 4 * mt + kt) * 64 + lane];
#pragma unroll
        for (int b = 0; b < NB; b++) c[b][mt] = mfma(a, act[b][2 * kt], act[b][2 * kt + 1], c[b][mt]);
      }
    }
#pragma unroll
    for (int b = 0; b < NB; b++)
#pragma unroll
      for (int mt = 0; mt < 4; mt++) {
        act[b][2 * mt] = pack2(silu(c[b][mt][0]), silu(c[b][mt][1]));
        act[b][2 * mt + 1] = pack2(silu(c[b][mt][2]), silu(c[b][mt][3]));
      }
  }
  // output layer: unit 4g + j of sample 16t + r in component j
  const v4f bi = bias4(sb + 3 * NN_HID + 4 * g);
  v4f c[NB];
#pragma unroll
  for (int b = 0; b < NB; b++) c[b] = bi;
#pragma unroll
  for (int kt = 0; kt < 4; kt++) {
    const uint2 a = sf[(NN_T4 + kt) * 64 + lane];
#pragma unroll
    for (int b = 0; b < NB; b++) c[b] = mfma(a, act[b][2 * kt], act[b][2 * kt + 1], c[b]);
  }
  // unit 0 of sample 16t + r sits in lane r; the sample's own lane is 16t + r
#pragma unroll
  for (int b = 0; b < NB; b++) {
    const float v = __shfl(c[b][0], r, 64);
    if (g == t0 + b) y = v;
  }
}
}  // namespace nn

// Output 0 of the network for this lane's sample. x: the lane's 16 normalised inputs as
// BF16 pairs (x[k] = inputs 2k, 2k+1). sf: the NN_TILES * 64 fragments, sb: the biases
// (16-byte aligned), both in LDS. want: the lanes whose output is needed (a ballot);
// blocks of 16 lanes without one are skipped (their lanes get 0). NB: blocks evaluated
// together (1, 2 or 4). Every lane of the wavefront must take part.
template <int NB = 1>
__device__ __forceinline__ float nn_wave_forward(const uint2 *sf, const float *sb, const uint32_t x[8],
                                                 unsigned long long want) {
  static_assert(NB == 1 || NB == 2 || NB == 4, "blocks per pass");
  constexpr unsigned long long MASK = NB == 4 ? ~0ull : ((1ull << (16 * (NB & 3))) - 1);
  float y = 0.0f;
#pragma unroll 1
  for (int t0 = 0; t0 < 4; t0 += NB) {
    if (((want >> (16 * t0)) & MASK) == 0) continue;  // wave-uniform
    nn::blocks_forward<NB>(sf, sb, x, t0, y);
  }
  return y;
}

// Stage the fragments, biases and normalisation block of a network into LDS (all threads
// of the block). The gate reads the normalisation block per sample: from LDS those are
// broadcast reads, from global memory they were per-lane loads (64 per sample).
__device__ __forceinline__ void nn_stage(const NetDev &N, uint2 *sf, float *sb, float *sn) {
  for (int i = threadIdx.x; i < NN_TILES * 64; i += blockDim.x) sf[i] = N.frag[i];
  for (int i = threadIdx.x; i < NN_NBIAS; i += blockDim.x) sb[i] = N.bias[i];
  for (int i = threadIdx.x; i < NN_NORM_N; i += blockDim.x) sn[i] = N.norm[i];
}

// Normalised network inputs of a sample (raw values xr[0 .. GY_MLP_IN)) as BF16 pairs;
// returns whether every input lies in the trained range.
__device__ __forceinline__ bool nn_inputs(const float *norm, const float xr[GY_MLP_IN], uint32_t x[8]) {
  float xn[NN_KIN];
  bool in = true;
#pragma unroll
  for (int k = 0; k < NN_KIN; k++) {
    if (k < GY_MLP_IN) {
      xn[k] = (xr[k] - norm[NN_NORM_MEAN + k]) * norm[NN_NORM_ISTD + k];
      in = in && xn[k] >= norm[NN_NORM_LO + k] && xn[k] <= norm[NN_NORM_HI + k];
    } else {
      xn[k] = 0.0f;
    }
  }
#pragma unroll
  for (int k = 0; k < 8; k++) x[k] = nn::pack2(xn[2 * k], xn[2 * k + 1]);
  return in;
}

}  // namespace gy
