// gy_mlp.hip — learned Drucker-Prager return map (gfx950).
//
// A 14-64-64-64-1 multilayer perceptron with SiLU activations gives the multiplier of
// the cone return of the Drucker-Prager map (druckerPrager_Damage,
// ConstitutiveModels.cpp:238-428) on yielding samples; the new stress and accumulated
// plastic strain are rebuilt from it along the exact return direction (dp_return in
// gy_models.hpp). Inputs: strain increment (6), prior stress (6), accumulated plastic
// strain (1), friction angle (1), normalised; output: the multiplier relative to the
// linear estimate F / (G + K eta eta_bar), scaled and offset by host-chosen constants.
//
// Every matrix product runs on the matrix cores as v_mfma_f32_16x16x16bf16_1k: BF16
// operands (weights, activations, back-propagated errors), FP32 accumulation; biases,
// pre-activations, losses, gradients and the Adam state stay FP32.
//
// Kernels:
//   k_mlp_pass<BACKWARD>  training: forward (outputs) or forward + backward of an MSE
//                         loss, one 64-sample tile per wavefront with operands staged in
//                         LDS (one FP32 gradient partial per tile)
//   k_mlp_adam            sums the partials and applies one Adam step
//   k_dp_exact<LOOP>      the return map per sample, closed form or the reference's loop
//   k_dp_surrogate        the gated surrogate: elastic trial exactly; yielding samples in
//                         the trained range through the network (nn_wave_forward, the
//                         evaluation the step kernel runs: four wavefronts per workgroup
//                         share the LDS copy of the weight fragments); yielding samples
//                         outside it in closed form
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "../../include/graphyt_amd.h"
#include "gy_internal.hpp"
#include "gy_models.hpp"

using namespace gy;

namespace {

constexpr int NIN = GY_MLP_IN, NOUT = GY_MLP_OUT;
constexpr int KIN = 16, HID = 64, KOUT = 16;  // padded widths
constexpr int OW1 = 0;                        // W1 [KIN][HID], row-major [k][n]
constexpr int OW2 = OW1 + KIN * HID;          // W2 [HID][HID]
constexpr int OW3 = OW2 + HID * HID;          // W3 [HID][HID]
constexpr int OW4 = OW3 + HID * HID;          // W4 [HID][KOUT]
constexpr int NW = OW4 + HID * KOUT;
constexpr int OB1 = NW, OB2 = OB1 + HID, OB3 = OB2 + HID, OB4 = OB3 + HID;
constexpr int NPARAM = OB4 + KOUT;
static_assert(NPARAM == GY_MLP_PARAMS, "parameter count of the header");
constexpr int ZW = 3 * HID;  // saved pre-activations per sample

typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

__host__ __device__ inline uint16_t f2bf(float x) {  // round to nearest even
  uint32_t u;
  __builtin_memcpy(&u, &x, 4);
  if ((u & 0x7fffffffu) > 0x7f800000u) return 0x7fc0;
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ inline float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

__device__ inline float sigm(float z) { return 1.0f / (1.0f + expf(-z)); }
__device__ inline float silu(float z) { return z * sigm(z); }
__device__ inline float silu_d(float z) {
  const float s = sigm(z);
  return s * (1.0f + z * (1.0f - s));
}

// D(16x16) += A(16xK) B(Kx16) in steps of 16 along K, with A(m, k) = a(m, k) and
// B(k, n) = b(k, n) given as BF16 bits. Operand layout of the 16x16x16 BF16 MFMA: lane l
// supplies row l%16 of A and column l%16 of B at k = 4*(l/16) + j (j = 0..3); result
// element (m, n) = (4*(l/16) + j, l%16) is component j of lane l.
template <class FA, class FB>
__device__ __forceinline__ v4f mfma_tile(FA a, FB b, int K, v4f acc) {
  const int l = threadIdx.x & 63, r = l & 15, kq = (l >> 4) * 4;
  for (int k0 = 0; k0 < K; k0 += 16) {
    v4s av, bv;
#pragma unroll
    for (int j = 0; j < 4; j++) {
      av[j] = (short)a(r, k0 + kq + j);
      bv[j] = (short)b(k0 + kq + j, r);
    }
    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bv, acc, 0, 0, 0);
  }
  return acc;
}

// One dense layer over 64 rows: Z = Hin W + b (Hin: 64 x K BF16, row stride ldi), then
// Hout = silu(Z) in BF16 (ACT) or Y = Z in FP32; Z saved to zs (row stride ZW) if given.
template <int K, int N, bool ACT>
__device__ __forceinline__ void dense(const uint16_t *hin, int ldi, const uint16_t *sW, int ow, const float *bias,
                                      uint16_t *hout, float *yout, float *zs) {
  const int l = threadIdx.x & 63, r = l & 15, kq = (l >> 4) * 4;
  for (int mt = 0; mt < 4; mt++)
    for (int nt = 0; nt < N / 16; nt++) {
      v4f acc = {0.f, 0.f, 0.f, 0.f};
      acc = mfma_tile([&](int m, int k) { return hin[(mt * 16 + m) * ldi + k]; },
                      [&](int k, int n) { return sW[ow + k * N + nt * 16 + n]; }, K, acc);
#pragma unroll
      for (int j = 0; j < 4; j++) {
        const int row = mt * 16 + kq + j, col = nt * 16 + r;
        const float z = acc[j] + bias[col];
        if (zs) zs[row * ZW + col] = z;
        if (ACT) hout[row * N + col] = f2bf(silu(z));
        else yout[row * KOUT + col] = z;
      }
    }
}

struct MlpLds {
  uint16_t W[NW];          // BF16 weights
  float B[NPARAM - NW];    // FP32 biases
  uint16_t X[64 * KIN];    // inputs (H0)
  uint16_t H1[64 * HID], H2[64 * HID], H3[64 * HID];  // activations; reused for dZ in backward
  float Y[64 * KOUT];      // outputs, then dY
  uint16_t DY[64 * KOUT];  // dY in BF16
};

__device__ void load_weights(MlpLds &S, const uint16_t *w16, const float *p32) {
  for (int i = threadIdx.x; i < NW; i += blockDim.x) S.W[i] = w16[i];
  for (int i = threadIdx.x; i < NPARAM - NW; i += blockDim.x) S.B[i] = p32[NW + i];
}

// Forward of the 64 rows staged in S.X; outputs in S.Y; pre-activations to zs if given.
__device__ void forward64(MlpLds &S, float *zs) {
  dense<KIN, HID, true>(S.X, KIN, S.W, OW1, S.B + (OB1 - NW), S.H1, nullptr, zs ? zs + 0 : nullptr);
  dense<HID, HID, true>(S.H1, HID, S.W, OW2, S.B + (OB2 - NW), S.H2, nullptr, zs ? zs + HID : nullptr);
  dense<HID, HID, true>(S.H2, HID, S.W, OW3, S.B + (OB3 - NW), S.H3, nullptr, zs ? zs + 2 * HID : nullptr);
  dense<HID, KOUT, false>(S.H3, HID, S.W, OW4, S.B + (OB4 - NW), nullptr, S.Y, nullptr);
}

// dW[k][n] = sum_rows Hin[row][k] dZ[row][n] (K x N), written to g[ow + k*N + n];
// db[n] = sum_rows dZ[row][n] to g[ob + n].
template <int K, int N>
__device__ __forceinline__ void weight_grad(const uint16_t *hin, int ldi, const uint16_t *dz, int ldz, float *g, int ow,
                                            int ob) {
  const int l = threadIdx.x & 63, r = l & 15, kq = (l >> 4) * 4;
  for (int mt = 0; mt < K / 16; mt++)
    for (int nt = 0; nt < N / 16; nt++) {
      v4f acc = {0.f, 0.f, 0.f, 0.f};
      acc = mfma_tile([&](int m, int kk) { return hin[kk * ldi + mt * 16 + m]; },
                      [&](int kk, int n) { return dz[kk * ldz + nt * 16 + n]; }, 64, acc);
#pragma unroll
      for (int j = 0; j < 4; j++) g[ow + (mt * 16 + kq + j) * N + nt * 16 + r] = acc[j];
    }
  for (int n = l; n < N; n += 64) {
    float s = 0.f;
    for (int row = 0; row < 64; row++) s += bf2f(dz[row * ldz + n]);
    g[ob + n] = s;
  }
}

// dZin[row][k] = (sum_n dZ[row][n] W[k][n]) * silu'(Zin[row][k]) for the K-wide layer
// below, written in BF16 to out (row stride K).
template <int K, int N>
__device__ __forceinline__ void back_prop(const uint16_t *dz, int ldz, const uint16_t *sW, int ow, const float *zin,
                                          uint16_t *out) {
  const int l = threadIdx.x & 63, r = l & 15, kq = (l >> 4) * 4;
  for (int mt = 0; mt < 4; mt++)
    for (int nt = 0; nt < K / 16; nt++) {
      v4f acc = {0.f, 0.f, 0.f, 0.f};
      acc = mfma_tile([&](int m, int kk) { return dz[(mt * 16 + m) * ldz + kk]; },
                      [&](int kk, int n) { return sW[ow + (nt * 16 + n) * N + kk]; }, N, acc);
#pragma unroll
      for (int j = 0; j < 4; j++) {
        const int row = mt * 16 + kq + j, col = nt * 16 + r;
        out[row * K + col] = f2bf(acc[j] * silu_d(zin[row * ZW + col]));
      }
    }
}

// Forward (BACKWARD = false: Y[n][NOUT] written) or forward + backward of
// L = sum (Y - T)^2 * loss_scale (BACKWARD = true: one gradient partial of NPARAM floats
// per block in gpart, and the block's loss in gpart_loss). One wave = 64 rows per block.
template <bool BACKWARD>
__global__ __launch_bounds__(64) void k_mlp_pass(const float *X, const float *T, long long n, const uint16_t *w16,
                                                 const float *p32, float *Yout, float *zscratch, float *gpart,
                                                 float *gloss, float loss_scale) {
  __shared__ MlpLds S;
  const int l = threadIdx.x;
  const long long r0 = (long long)blockIdx.x * 64;
  load_weights(S, w16, p32);
  {
    const long long row = r0 + l;
    for (int k = 0; k < KIN; k++) S.X[l * KIN + k] = (row < n && k < NIN) ? f2bf(X[row * NIN + k]) : (uint16_t)0;
  }
  __syncthreads();
  float *zs = BACKWARD ? zscratch + (long long)blockIdx.x * 64 * ZW : nullptr;
  forward64(S, zs);
  __syncthreads();
  if (!BACKWARD) {
    const long long row = r0 + l;
    if (row < n)
      for (int o = 0; o < NOUT; o++) Yout[row * NOUT + o] = S.Y[l * KOUT + o];
    return;
  }
  // dY = 2 (Y - T) * loss_scale on the NOUT used columns of valid rows
  {
    const long long row = r0 + l;
    float loss = 0.f;
    for (int o = 0; o < KOUT; o++) {
      float dy = 0.f;
      if (row < n && o < NOUT) {
        const float e = S.Y[l * KOUT + o] - T[row * NOUT + o];
        loss += e * e;
        dy = 2.0f * e * loss_scale;
      }
      S.Y[l * KOUT + o] = dy;
      S.DY[l * KOUT + o] = f2bf(dy);
    }
    for (int w = 32; w > 0; w >>= 1) loss += __shfl_xor(loss, w, 64);
    if (l == 0) gloss[blockIdx.x] = loss * loss_scale;
  }
  __syncthreads();
  float *g = gpart + (long long)blockIdx.x * NPARAM;
  // layer 4: dW4 = H3^T dY (BF16 dY), db4 = sum dY (FP32)
  {
    const int r = l & 15, kq = (l >> 4) * 4;
    for (int mt = 0; mt < HID / 16; mt++) {
      v4f acc = {0.f, 0.f, 0.f, 0.f};
      acc = mfma_tile([&](int m, int kk) { return S.H3[kk * HID + mt * 16 + m]; },
                      [&](int kk, int nn) { return S.DY[kk * KOUT + nn]; }, 64, acc);
#pragma unroll
      for (int j = 0; j < 4; j++) g[OW4 + (mt * 16 + kq + j) * KOUT + r] = acc[j];
    }
    if (l < KOUT) {
      float s = 0.f;
      for (int row = 0; row < 64; row++) s += S.Y[row * KOUT + l];
      g[OB4 + l] = s;
    }
  }
  __syncthreads();
  back_prop<HID, KOUT>(S.DY, KOUT, S.W, OW4, zs + 2 * HID, S.H3);  // dZ3 -> H3 buffer
  __syncthreads();
  weight_grad<HID, HID>(S.H2, HID, S.H3, HID, g, OW3, OB3);
  __syncthreads();
  back_prop<HID, HID>(S.H3, HID, S.W, OW3, zs + HID, S.H2);        // dZ2 -> H2 buffer
  __syncthreads();
  weight_grad<HID, HID>(S.H1, HID, S.H2, HID, g, OW2, OB2);
  __syncthreads();
  back_prop<HID, HID>(S.H2, HID, S.W, OW2, zs + 0, S.H1);          // dZ1 -> H1 buffer
  __syncthreads();
  weight_grad<KIN, HID>(S.X, KIN, S.H1, HID, g, OW1, OB1);
}

// Sum of the per-block gradient partials, then one Adam step (Kingma & Ba 2015) on the
// FP32 master parameters and their BF16 copies. grad (optional) receives the summed
// gradient.
__global__ __launch_bounds__(256) void k_mlp_adam(const float *gpart, int nblocks, float *p32, uint16_t *w16,
                                                  float *m1, float *m2, float *grad, float lr, float b1, float b2,
                                                  float eps, float c1, float c2, int apply) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= NPARAM) return;
  float g = 0.f;
  for (int b = 0; b < nblocks; b++) g += gpart[(long long)b * NPARAM + p];
  if (grad) grad[p] = g;
  if (!apply) return;
  const float a = b1 * m1[p] + (1.0f - b1) * g;
  const float v = b2 * m2[p] + (1.0f - b2) * g * g;
  m1[p] = a;
  m2[p] = v;
  const float w = p32[p] - lr * (a * c1) / (sqrtf(v * c2) + eps);
  p32[p] = w;
  if (p < NW) w16[p] = f2bf(w);
}

// ---- the return map over a batch of samples ------------------------------------------
// sample i: in[i*GY_DP_IN + 0..5] strain increment (xx yy zz xy yz xz), 6..11 prior
// stress, 12 accumulated plastic strain; out[i*GY_DP_OUT + 0..5] new stress, 6 new
// accumulated plastic strain; mult[i] (optional) the cone multiplier used (0 if elastic).
__device__ __forceinline__ void load_sample(const float *in, long long i, bool valid, float de[6], float s[6],
                                            float &epb) {
#pragma unroll
  for (int k = 0; k < 6; k++) {
    de[k] = valid ? in[i * GY_DP_IN + k] : 0.0f;
    s[k] = valid ? in[i * GY_DP_IN + 6 + k] : 0.0f;
  }
  epb = valid ? in[i * GY_DP_IN + 12] : 0.0f;
}
__device__ __forceinline__ void store_sample(float *out, float *mult, long long i, const float s[6], float epb,
                                             float dL) {
#pragma unroll
  for (int k = 0; k < 6; k++) out[i * GY_DP_OUT + k] = s[k];
  out[i * GY_DP_OUT + 6] = epb;
  if (mult) mult[i] = dL;
}

template <bool LOOP>
__global__ __launch_bounds__(256) void k_dp_exact(MatDev M, const float *in, float *out, float *mult, long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float de[6], s[6], p[6] = {0, 0, 0, 0, 0, 0}, epb;
  load_sample(in, i, true, de, s, epb);
  DpTrial T;
  dp_trial(M, de, s, 1.0f, T);
  float dL = 0.0f;
  if (T.yield) {
    dL = dp_cone_multiplier<LOOP>(M, T, epb);
    dp_return<LOOP>(M, T, dL, s, p, epb);
  } else {
#pragma unroll
    for (int k = 0; k < 6; k++) s[k] = T.tr[k];
  }
  store_sample(out, mult, i, s, epb, dL);
}

// Gated surrogate, persistent over 64-sample tiles per wavefront. used[0] counts samples
// that took the network's multiplier, used[1] yielding samples outside the trained range
// or near the return's branch switch (closed form).
template <int NB>  // 16-sample blocks per network pass
__global__ __launch_bounds__(256) void k_dp_surrogate(MatDev M, NetDev N, const float *in, float *out, float *mult,
                                                      long long n, unsigned long long *used) {
  __shared__ uint2 sf[NN_TILES * 64];
  __shared__ __align__(16) float sb[NN_NBIAS];
  __shared__ __align__(16) float sn[NN_NORM_N];
  nn_stage(N, sf, sb, sn);
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const long long wave = (long long)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  const long long nwaves = (long long)gridDim.x * (blockDim.x >> 6);
  const float oscale = sn[NN_NORM_OUT], ooff = sn[NN_NORM_OUT + 1];
  const float ylo = sn[NN_NORM_YLO], yhi = sn[NN_NORM_YHI];
  unsigned n_net = 0, n_fb = 0;
  for (long long r0 = wave * 64; r0 < n; r0 += nwaves * 64) {
    const long long i = r0 + lane;
    const bool valid = i < n;
    float de[6], s[6], p[6] = {0, 0, 0, 0, 0, 0}, epb;
    load_sample(in, i, valid, de, s, epb);
    DpTrial T;
    dp_trial(M, de, s, 1.0f, T);
    const bool yield = valid && T.yield;
    float xr[GY_MLP_IN];
#pragma unroll
    for (int k = 0; k < 6; k++) {
      xr[k] = de[k];
      xr[6 + k] = s[k];
    }
    xr[12] = epb;
    xr[13] = M.phi_deg;
    uint32_t xb[8];
    const bool net = yield && nn_inputs(sn, xr, xb);
    float y = 0.0f;
    const unsigned long long want = __ballot(net);
    if (want) y = nn_wave_forward<NB>(sf, sb, xb, want);  // wave-uniform
    float dL = net ? dp_learned_multiplier(M, T, oscale * y + ooff, ylo, yhi) : 0.0f;
    if (net && dL >= 0.0f) {
      dp_return<false>(M, T, dL, s, p, epb);
      n_net++;
    } else if (yield) {
      dL = dp_cone_multiplier<false>(M, T, epb);
      dp_return<false>(M, T, dL, s, p, epb);
      n_fb++;
    } else {
#pragma unroll
      for (int k = 0; k < 6; k++) s[k] = T.tr[k];
    }
    if (valid) store_sample(out, mult, i, s, epb, dL);
  }
  for (int w = 32; w > 0; w >>= 1) {
    n_net += __shfl_xor(n_net, w, 64);
    n_fb += __shfl_xor(n_fb, w, 64);
  }
  if (lane == 0 && (n_net | n_fb)) {
    atomicAdd(&used[0], (unsigned long long)n_net);
    atomicAdd(&used[1], (unsigned long long)n_fb);
  }
}

}  // namespace

// ---------------------------------------------------------------------------------------
struct gy_mlp {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  float *p32 = nullptr, *m1 = nullptr, *m2 = nullptr, *grad = nullptr;
  uint16_t *w16 = nullptr;
  // training set and scratch
  float *X = nullptr, *T = nullptr;
  long long ndata = 0;
  float *zs = nullptr, *gpart = nullptr, *gloss = nullptr;
  int cap_blocks = 0;
  int adam_t = 0;
  float norm[NN_NORM_N] = {};  // layout of gy_nn.hpp (NN_NORM_*)
  void *nnbuf = nullptr;       // device copy of the network in the step's layout (gy_dp_map)
};

namespace {

int mfail(gy_mlp *m, int code, const char *msg) {
  if (m) m->err = msg;
  return code;
}

#define MCHECK(m, call)                                                            \
  do {                                                                             \
    hipError_t e_ = (call);                                                        \
    if (e_ != hipSuccess) {                                                        \
      if (m) m->err = std::string(#call) + " failed: " + hipGetErrorString(e_);    \
      return GY_ERR_HIP;                                                           \
    }                                                                              \
  } while (0)

int ensure_scratch(gy_mlp *m, int blocks) {
  if (blocks <= m->cap_blocks) return GY_OK;
  if (m->zs) (void)hipFree(m->zs);
  if (m->gpart) (void)hipFree(m->gpart);
  if (m->gloss) (void)hipFree(m->gloss);
  m->zs = m->gpart = m->gloss = nullptr;
  MCHECK(m, hipMalloc((void **)&m->zs, (size_t)blocks * 64 * ZW * sizeof(float)));
  MCHECK(m, hipMalloc((void **)&m->gpart, (size_t)blocks * NPARAM * sizeof(float)));
  MCHECK(m, hipMalloc((void **)&m->gloss, (size_t)blocks * sizeof(float)));
  m->cap_blocks = blocks;
  return GY_OK;
}

int upload_params(gy_mlp *m, const float *p) {
  std::vector<uint16_t> w(NW);
  for (int i = 0; i < NW; i++) w[i] = f2bf(p[i]);
  MCHECK(m, hipMemcpy(m->p32, p, NPARAM * sizeof(float), hipMemcpyHostToDevice));
  MCHECK(m, hipMemcpy(m->w16, w.data(), NW * sizeof(uint16_t), hipMemcpyHostToDevice));
  MCHECK(m, hipMemset(m->m1, 0, NPARAM * sizeof(float)));
  MCHECK(m, hipMemset(m->m2, 0, NPARAM * sizeof(float)));
  m->adam_t = 0;
  return GY_OK;
}

// one batch [row0, row0 + rows) of the loaded training set: gradients and (apply) an
// Adam step; returns the batch loss
int batch_step(gy_mlp *m, const float *X, const float *T, long long rows, float lr, int apply, float *grad_out,
               double *loss_out) {
  const int blocks = (int)((rows + 63) / 64);
  int rc = ensure_scratch(m, blocks);
  if (rc) return rc;
  const float scale = 1.0f / (float)(rows * NOUT);
  hipLaunchKernelGGL(k_mlp_pass<true>, dim3(blocks), dim3(64), 0, m->stream, X, T, rows, m->w16, m->p32, nullptr,
                     m->zs, m->gpart, m->gloss, scale);
  float c1 = 1.0f, c2 = 1.0f;
  const float b1 = 0.9f, b2 = 0.999f;
  if (apply) {
    m->adam_t++;
    c1 = 1.0f / (1.0f - std::pow(b1, (float)m->adam_t));
    c2 = 1.0f / (1.0f - std::pow(b2, (float)m->adam_t));
  }
  hipLaunchKernelGGL(k_mlp_adam, dim3((NPARAM + 255) / 256), dim3(256), 0, m->stream, m->gpart, blocks, m->p32,
                     m->w16, m->m1, m->m2, grad_out, lr, b1, b2, 1e-8f, c1, c2, apply);
  MCHECK(m, hipGetLastError());
  if (loss_out) {
    std::vector<float> bl(blocks);
    MCHECK(m, hipMemcpyAsync(bl.data(), m->gloss, blocks * sizeof(float), hipMemcpyDeviceToHost, m->stream));
    MCHECK(m, hipStreamSynchronize(m->stream));
    double s = 0;
    for (float x : bl) s += x;
    *loss_out = s;
  }
  return GY_OK;
}

}  // namespace

namespace gy {
int mlp_export(gy_mlp *m, uint2 *frag, float *bias, float *norm) {
  if (!m) return GY_ERR_INVALID;
  std::vector<float> p(NPARAM);
  MCHECK(m, hipStreamSynchronize(m->stream));
  MCHECK(m, hipMemcpy(p.data(), m->p32, NPARAM * sizeof(float), hipMemcpyDeviceToHost));
  // tile (mt, kt) of a K x N layer at ow: lane l holds W[16kt + 4(l/16) + j][16mt + l%16]
  auto tiles = [&](int t0, int ow, int K, int N) {
    for (int mt = 0; mt < N / 16; mt++)
      for (int kt = 0; kt < K / 16; kt++)
        for (int l = 0; l < 64; l++) {
          const int col = 16 * mt + (l & 15), k0 = 16 * kt + 4 * (l >> 4);
          auto w = [&](int k) { return (uint32_t)f2bf(p[ow + k * N + col]); };
          frag[(t0 + mt * (K / 16) + kt) * 64 + l] = make_uint2(w(k0) | (w(k0 + 1) << 16), w(k0 + 2) | (w(k0 + 3) << 16));
        }
  };
  tiles(NN_T1, OW1, KIN, HID);
  tiles(NN_T2, OW2, HID, HID);
  tiles(NN_T3, OW3, HID, HID);
  tiles(NN_T4, OW4, HID, KOUT);
  for (int i = 0; i < NN_NBIAS; i++) bias[i] = p[OB1 + i];
  for (int i = 0; i < NN_NORM_N; i++) norm[i] = m->norm[i];
  return GY_OK;
}
}  // namespace gy

extern "C" {

int gy_mlp_create(gy_mlp **out, uint64_t seed) {
  if (!out) return GY_ERR_INVALID;
  *out = nullptr;
  gy_mlp *m = new gy_mlp();
  *out = m;
  MCHECK(m, hipGetDevice(&m->device));
  MCHECK(m, hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking));
  MCHECK(m, hipMalloc((void **)&m->p32, NPARAM * sizeof(float)));
  MCHECK(m, hipMalloc((void **)&m->m1, NPARAM * sizeof(float)));
  MCHECK(m, hipMalloc((void **)&m->m2, NPARAM * sizeof(float)));
  MCHECK(m, hipMalloc((void **)&m->grad, NPARAM * sizeof(float)));
  MCHECK(m, hipMalloc((void **)&m->w16, NW * sizeof(uint16_t)));
  // He-uniform weights on the used inputs, zero biases and zero padding
  std::vector<float> p(NPARAM, 0.f);
  std::mt19937_64 rng(seed);
  auto fill = [&](int ow, int K, int N, int kused, int nused) {
    const float a = std::sqrt(6.0f / (float)kused);
    std::uniform_real_distribution<float> U(-a, a);
    for (int k = 0; k < K; k++)
      for (int n = 0; n < N; n++) p[ow + k * N + n] = (k < kused && n < nused) ? U(rng) : 0.f;
  };
  fill(OW1, KIN, HID, NIN, HID);
  fill(OW2, HID, HID, HID, HID);
  fill(OW3, HID, HID, HID, HID);
  fill(OW4, HID, KOUT, HID, NOUT);
  for (int i = 0; i < NN_KIN; i++) {
    m->norm[NN_NORM_ISTD + i] = i < NIN ? 1.f : 0.f;
    m->norm[NN_NORM_LO + i] = -1e30f;
    m->norm[NN_NORM_HI + i] = 1e30f;
  }
  m->norm[NN_NORM_OUT] = 1.f;
  m->norm[NN_NORM_YLO] = -1e30f;
  m->norm[NN_NORM_YHI] = 1e30f;
  return upload_params(m, p.data());
}

int gy_mlp_param_count(void) { return NPARAM; }

int gy_mlp_set_params(gy_mlp *m, const float *params) {
  if (!m || !params) return GY_ERR_INVALID;
  return upload_params(m, params);
}

int gy_mlp_get_params(gy_mlp *m, float *params) {
  if (!m || !params) return GY_ERR_INVALID;
  MCHECK(m, hipStreamSynchronize(m->stream));
  MCHECK(m, hipMemcpy(params, m->p32, NPARAM * sizeof(float), hipMemcpyDeviceToHost));
  return GY_OK;
}

int gy_mlp_set_norm(gy_mlp *m, const float *in_mean, const float *in_std, const float *in_lo, const float *in_hi,
                    float out_scale, float out_offset, float out_lo, float out_hi) {
  if (!m || !in_mean || !in_std || !in_lo || !in_hi) return GY_ERR_INVALID;
  for (int i = 0; i < NIN; i++) {
    if (!(in_std[i] > 0.f)) return mfail(m, GY_ERR_INVALID, "input std must be positive");
    m->norm[NN_NORM_MEAN + i] = in_mean[i];
    m->norm[NN_NORM_ISTD + i] = 1.0f / in_std[i];
    m->norm[NN_NORM_LO + i] = in_lo[i];
    m->norm[NN_NORM_HI + i] = in_hi[i];
  }
  if (!std::isfinite(out_scale) || !std::isfinite(out_offset)) return mfail(m, GY_ERR_INVALID, "output scale and offset must be finite");
  if (!(out_lo <= out_hi)) return mfail(m, GY_ERR_INVALID, "output range needs out_lo <= out_hi");
  m->norm[NN_NORM_OUT] = out_scale;
  m->norm[NN_NORM_OUT + 1] = out_offset;
  m->norm[NN_NORM_YLO] = out_lo;
  m->norm[NN_NORM_YHI] = out_hi;
  return GY_OK;
}

int gy_mlp_forward(gy_mlp *m, const float *x, int64_t n, float *y) {
  if (!m || !x || !y || n <= 0) return GY_ERR_INVALID;
  float *dx = nullptr, *dy = nullptr;
  MCHECK(m, hipMalloc((void **)&dx, (size_t)n * NIN * 4));
  MCHECK(m, hipMalloc((void **)&dy, (size_t)n * NOUT * 4));
  int rc = GY_OK;
  if (hipMemcpy(dx, x, (size_t)n * NIN * 4, hipMemcpyHostToDevice) != hipSuccess) rc = GY_ERR_HIP;
  if (!rc) {
    hipLaunchKernelGGL(k_mlp_pass<false>, dim3((unsigned)((n + 63) / 64)), dim3(64), 0, m->stream, dx, nullptr, n,
                       m->w16, m->p32, dy, nullptr, nullptr, nullptr, 0.f);
    if (hipStreamSynchronize(m->stream) != hipSuccess ||
        hipMemcpy(y, dy, (size_t)n * NOUT * 4, hipMemcpyDeviceToHost) != hipSuccess)
      rc = mfail(m, GY_ERR_HIP, "forward failed");
  }
  (void)hipFree(dx);
  (void)hipFree(dy);
  return rc;
}

int gy_mlp_gradient(gy_mlp *m, const float *x, const float *t, int64_t n, float *grad, double *loss) {
  if (!m || !x || !t || !grad || n <= 0) return GY_ERR_INVALID;
  float *dx = nullptr, *dt = nullptr;
  MCHECK(m, hipMalloc((void **)&dx, (size_t)n * NIN * 4));
  MCHECK(m, hipMalloc((void **)&dt, (size_t)n * NOUT * 4));
  int rc = GY_OK;
  if (hipMemcpy(dx, x, (size_t)n * NIN * 4, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(dt, t, (size_t)n * NOUT * 4, hipMemcpyHostToDevice) != hipSuccess)
    rc = mfail(m, GY_ERR_HIP, "copy failed");
  if (!rc) rc = batch_step(m, dx, dt, n, 0.f, 0, m->grad, loss);
  if (!rc && hipMemcpy(grad, m->grad, NPARAM * 4, hipMemcpyDeviceToHost) != hipSuccess)
    rc = mfail(m, GY_ERR_HIP, "copy failed");
  (void)hipFree(dx);
  (void)hipFree(dt);
  return rc;
}

int gy_mlp_load_data(gy_mlp *m, const float *x, const float *t, int64_t n) {
  if (!m || !x || !t || n <= 0) return GY_ERR_INVALID;
  if (m->X) (void)hipFree(m->X);
  if (m->T) (void)hipFree(m->T);
  m->X = m->T = nullptr;
  MCHECK(m, hipMalloc((void **)&m->X, (size_t)n * NIN * 4));
  MCHECK(m, hipMalloc((void **)&m->T, (size_t)n * NOUT * 4));
  MCHECK(m, hipMemcpy(m->X, x, (size_t)n * NIN * 4, hipMemcpyHostToDevice));
  MCHECK(m, hipMemcpy(m->T, t, (size_t)n * NOUT * 4, hipMemcpyHostToDevice));
  m->ndata = n;
  return GY_OK;
}

int gy_mlp_train(gy_mlp *m, int32_t epochs, int32_t batch, double lr, double lr_decay, double *last_loss) {
  if (!m || !m->X || epochs <= 0 || batch < 64 || batch > (1 << 20)) return GY_ERR_INVALID;
  double loss_sum = 0.0;
  long long rows_sum = 0;
  float rate = (float)lr;
  for (int e = 0; e < epochs; e++) {
    loss_sum = 0.0;
    rows_sum = 0;
    for (long long r0 = 0; r0 < m->ndata; r0 += batch) {
      const long long rows = std::min<long long>(batch, m->ndata - r0);
      double bl = 0.0;
      const bool last = (e == epochs - 1);
      int rc = batch_step(m, m->X + r0 * NIN, m->T + r0 * NOUT, rows, rate, 1, nullptr, last ? &bl : nullptr);
      if (rc) return rc;
      loss_sum += bl * (double)rows;
      rows_sum += rows;
    }
    rate *= (float)lr_decay;
  }
  MCHECK(m, hipStreamSynchronize(m->stream));
  if (last_loss) *last_loss = rows_sum ? loss_sum / (double)rows_sum : 0.0;
  return GY_OK;
}

int gy_dp_map(const gy_material *mat, gy_mlp *m, int32_t mode, const float *in, float *out, float *multiplier,
              int64_t n, int64_t *n_network, int64_t *n_fallback, double *kernel_ms) {
  if (!mat || !in || !out || n <= 0) return GY_ERR_INVALID;
  if (mode < GY_DP_RETURN_CLOSED || mode > GY_DP_RETURN_LEARNED) return mfail(m, GY_ERR_INVALID, "unknown return-map mode");
  if (mode == GY_DP_RETURN_LEARNED && !m) return GY_ERR_INVALID;
  if (mat->model != 4) return GY_ERR_UNSUPPORTED;
  const MatDev M = material_constants(*mat);
  hipStream_t st = m ? m->stream : nullptr;
  if (mode == GY_DP_RETURN_LEARNED) {
    std::vector<uint2> frag(NN_TILES * 64);
    std::vector<float> bias(NN_NBIAS), norm(NN_NORM_N);
    int rc = mlp_export(m, frag.data(), bias.data(), norm.data());
    if (rc) return rc;
    const size_t fb = frag.size() * sizeof(uint2), bb = bias.size() * sizeof(float), nb = norm.size() * sizeof(float);
    if (!m->nnbuf) MCHECK(m, hipMalloc(&m->nnbuf, fb + bb + nb));
    char *base = (char *)m->nnbuf;
    MCHECK(m, hipMemcpy(base, frag.data(), fb, hipMemcpyHostToDevice));
    MCHECK(m, hipMemcpy(base + fb, bias.data(), bb, hipMemcpyHostToDevice));
    MCHECK(m, hipMemcpy(base + fb + bb, norm.data(), nb, hipMemcpyHostToDevice));
  }
  float *din = nullptr, *dout = nullptr, *dmult = nullptr;
  unsigned long long *used = nullptr;
  if (hipMalloc((void **)&din, (size_t)n * GY_DP_IN * 4) != hipSuccess ||
      hipMalloc((void **)&dout, (size_t)n * GY_DP_OUT * 4) != hipSuccess ||
      (multiplier && hipMalloc((void **)&dmult, (size_t)n * 4) != hipSuccess) ||
      hipMalloc((void **)&used, 2 * sizeof(unsigned long long)) != hipSuccess) {
    (void)hipFree(din); (void)hipFree(dout); (void)hipFree(dmult); (void)hipFree(used);
    return mfail(m, GY_ERR_HIP, "allocation failed");
  }
  int rc = GY_OK;
  hipEvent_t a = nullptr, b = nullptr;
  (void)hipEventCreate(&a);
  (void)hipEventCreate(&b);
  if (hipMemcpy(din, in, (size_t)n * GY_DP_IN * 4, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemset(used, 0, 2 * sizeof(unsigned long long)) != hipSuccess)
    rc = mfail(m, GY_ERR_HIP, "copy failed");
  if (!rc) {
    int cus = 256;
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, m ? m->device : 0);
    (void)hipEventRecord(a, st);
    if (mode == GY_DP_RETURN_LEARNED) {
      // four wavefronts per workgroup share the weights; as many workgroups as fit
      const char *base = (const char *)m->nnbuf;
      const NetDev N{(const uint2 *)base, (const float *)(base + NN_TILES * 64 * sizeof(uint2)),
                     (const float *)(base + NN_TILES * 64 * sizeof(uint2) + NN_NBIAS * sizeof(float))};
      // four 16-sample blocks per pass (2 waves per SIMD): 0.135 ms per 1M samples, against
      // 0.155 with two (4 waves per SIMD) and 0.174 with one (5 waves per SIMD)
      int per_cu = 1;
      (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_dp_surrogate<4>, 256, 0);
      const long long tiles = (n + 63) / 64;
      const int grid = (int)std::max<long long>(
          1, std::min<long long>((tiles + 3) / 4, (long long)cus * std::max(1, per_cu)));
      hipLaunchKernelGGL(k_dp_surrogate<4>, dim3(grid), dim3(256), 0, st, M, N, din, dout, dmult, n, used);
    } else if (mode == GY_DP_RETURN_REFERENCE) {
      hipLaunchKernelGGL(k_dp_exact<true>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, M, din, dout, dmult, n);
    } else {
      hipLaunchKernelGGL(k_dp_exact<false>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, M, din, dout, dmult, n);
    }
    (void)hipEventRecord(b, st);
    unsigned long long u[2] = {0, 0};
    if (hipStreamSynchronize(st) != hipSuccess || hipGetLastError() != hipSuccess ||
        hipMemcpy(out, dout, (size_t)n * GY_DP_OUT * 4, hipMemcpyDeviceToHost) != hipSuccess ||
        (multiplier && hipMemcpy(multiplier, dmult, (size_t)n * 4, hipMemcpyDeviceToHost) != hipSuccess) ||
        hipMemcpy(u, used, sizeof u, hipMemcpyDeviceToHost) != hipSuccess) {
      rc = mfail(m, GY_ERR_HIP, "return-map kernel failed");
    } else {
      float ms = 0.f;
      (void)hipEventElapsedTime(&ms, a, b);
      if (kernel_ms) *kernel_ms = ms;
      if (n_network) *n_network = (int64_t)u[0];
      if (n_fallback) *n_fallback = (int64_t)u[1];
    }
  }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  (void)hipFree(din);
  (void)hipFree(dout);
  (void)hipFree(dmult);
  (void)hipFree(used);
  return rc;
}

const char *gy_mlp_last_error(const gy_mlp *m) { return m ? m->err.c_str() : "null network"; }

void gy_mlp_destroy(gy_mlp *m) {
  if (!m) return;
  if (m->stream) (void)hipStreamSynchronize(m->stream);
  for (void *p : {(void *)m->p32, (void *)m->m1, (void *)m->m2, (void *)m->grad, (void *)m->w16, (void *)m->X,
                  (void *)m->T, (void *)m->zs, (void *)m->gpart, (void *)m->gloss, m->nnbuf})
    if (p) (void)hipFree(p);
  if (m->stream) (void)hipStreamDestroy(m->stream);
  delete m;
}

}  // extern "C"
