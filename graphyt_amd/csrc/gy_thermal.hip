// gy_thermal.hip — heat transfer of the MPM step, for contexts with a nonzero temperature.
//
// The reference carries per particle a temperature T, its gradient gT and its rate, and per
// node a heat content and a conduction rate, neither paired across the periodic faces
// (Extrapolation.cpp:281-283, 359-362; Grid.cpp:797-860; TimeIntegration.cpp:50, 160):
//   P2G   node heat += N m cp T,   node rate -= (m/rho) k (gradN . gT)   (m = 1 for rigid)
//   G2P   at nodes with mass:  gT = sum gradN * node heat,  Tdot = sum N * node rate
//   update  T += Tdot * dt
// Here the two transfers are kernels of their own, outside the fused particle kernel (which
// only carries T and gT along when it moves a particle's state): k_therm_g2p runs between
// the grid update and the particle kernel (same positions as this step's P2G, node mass of
// this step's solve), k_therm_p2g after the re-binning (positions, temperature and gradient
// of the new state). FP32 global atomics for the node sums.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "gy_device.hpp"
#include "gy_internal.hpp"

namespace gy {
namespace {

// The 3 x 3 x 3 stencil of a particle (3 x 3 x 1 in a plane, where shape_z = 1 and
// gradShape_z = 0: Extrapolation.cpp:184-188, 304-308).
struct Stencil {
  int b[3];
  Axis W[3];
  int nk;
};
__device__ __forceinline__ Stencil stencil_of(const uint32_t *A, long long cap, int s, const InstDesc &I) {
  Stencil S;
  bool bad = false;
  const int na = I.plane ? 2 : 3;
  for (int a = 0; a < na; a++) {
    const int q = (int)A[(F_QX + a) * cap + s];
    S.b[a] = stencil_base(q, I.lpq, I.fbits, I.n[a] - 3, bad);
    S.W[a] = gimp3((float)(q - (S.b[a] << I.fbits)) * I.inv_2F, I.lp, I.inv_lp, I.inv_h);
  }
  if (I.plane) {
    S.b[2] = 0;
    S.W[2].w[0] = 1.0f; S.W[2].w[1] = 0.0f; S.W[2].w[2] = 0.0f;
    S.W[2].g[0] = 0.0f; S.W[2].g[1] = 0.0f; S.W[2].g[2] = 0.0f;
  }
  S.nk = I.plane ? 1 : 3;
  return S;
}

__global__ __launch_bounds__(WG) void k_therm_clear(DevPtrs d) {
  float4 *t = reinterpret_cast<float4 *>(d.therm);
  const long long n4 = d.nnodes / 2;  // two nodes per float4
  for (long long i = (long long)blockIdx.x * WG + threadIdx.x; i < n4; i += (long long)gridDim.x * WG)
    t[i] = make_float4(0.f, 0.f, 0.f, 0.f);
  if (blockIdx.x == 0 && threadIdx.x == 0 && (d.nnodes & 1)) {
    d.therm[2 * (d.nnodes - 1)] = 0.f;
    d.therm[2 * (d.nnodes - 1) + 1] = 0.f;
  }
}

// Extrapolation.cpp:281-283 for every node of the particle's stencil.
__global__ __launch_bounds__(WG) void k_therm_p2g(DevPtrs d) {
  const int s = blockIdx.x * WG + threadIdx.x;
  if (s >= d.np) return;
  const uint32_t *A = (d.ctrl->parity ? d.pbuf[1] : d.pbuf[0]);
  const long long cap = d.cap;
  const InstDesc &I = d.inst[inst_of_slot(d, s)];
  const uint32_t tag = A[F_TAG * cap + s];
  const MatDev &M = d.mats[I.mat_base + tag_mat(tag)];
  const Stencil S = stencil_of(A, cap, s, I);
  const float mp = M.rigid ? 1.0f : f32(A[F_MASS * cap + s]);
  const float heat = f32(A[F_T * cap + s]) * mp * M.cp;
  const float cond = -(mp / f32(A[F_RHO * cap + s])) * M.kt;
  const float tg[3] = {f32(A[F_TG0 * cap + s]), f32(A[F_TG1 * cap + s]), f32(A[F_TG2 * cap + s])};
  for (int k = 0; k < S.nk; k++)
    for (int j = 0; j < 3; j++)
#pragma unroll
      for (int i = 0; i < 3; i++) {
        const float wyz = S.W[1].w[j] * S.W[2].w[k];
        const float w = S.W[0].w[i] * wyz;
        const float gx = S.W[0].g[i] * wyz, gy = S.W[0].w[i] * S.W[1].g[j] * S.W[2].w[k],
                    gz = S.W[0].w[i] * S.W[1].w[j] * S.W[2].g[k];
        if (w == 0.0f && gx == 0.0f && gy == 0.0f && gz == 0.0f) continue;
        const long long ni =
            I.node_base + (S.b[0] + i) + (long long)I.n[0] * ((S.b[1] + j) + (long long)I.n[1] * (S.b[2] + k));
        unsafeAtomicAdd(&d.therm[2 * ni], w * heat);
        unsafeAtomicAdd(&d.therm[2 * ni + 1], cond * (gx * tg[0] + gy * tg[1] + gz * tg[2]));
      }
}

// Extrapolation.cpp:127-132, 359-362 and TimeIntegration.cpp:50, in place on the state
// buffer the particle kernel reads next.
__global__ __launch_bounds__(WG) void k_therm_g2p(DevPtrs d) {
  const int s = blockIdx.x * WG + threadIdx.x;
  if (s >= d.np) return;
  uint32_t *A = (d.ctrl->parity ? d.pbuf[1] : d.pbuf[0]);
  const long long cap = d.cap;
  const InstDesc &I = d.inst[inst_of_slot(d, s)];
  const Stencil S = stencil_of(A, cap, s, I);
  float tg[3] = {0.f, 0.f, 0.f}, rate = 0.f;
  for (int k = 0; k < S.nk; k++)
    for (int j = 0; j < 3; j++)
#pragma unroll
      for (int i = 0; i < 3; i++) {
        const long long ni =
            I.node_base + (S.b[0] + i) + (long long)I.n[0] * ((S.b[1] + j) + (long long)I.n[1] * (S.b[2] + k));
        if (d.vel[ni * NODE_W + 6] == 0.0f) continue;  // node mass (Extrapolation.cpp:291)
        const float2 t = reinterpret_cast<const float2 *>(d.therm)[ni];
        const float wyz = S.W[1].w[j] * S.W[2].w[k];
        tg[0] += S.W[0].g[i] * wyz * t.x;
        tg[1] += S.W[0].w[i] * S.W[1].g[j] * S.W[2].w[k] * t.x;
        tg[2] += S.W[0].w[i] * S.W[1].w[j] * S.W[2].g[k] * t.x;
        rate += S.W[0].w[i] * wyz * t.y;
      }
  A[F_T * cap + s] = u32(f32(A[F_T * cap + s]) + rate * I.dt);
  A[F_TG0 * cap + s] = u32(tg[0]);
  A[F_TG1 * cap + s] = u32(tg[1]);
  A[F_TG2 * cap + s] = u32(tg[2]);
}

inline hipStream_t S(void *s) { return reinterpret_cast<hipStream_t>(s); }
inline int blocks(long long n, int per) { return (int)((n + per - 1) / per); }

}  // namespace

void launch_therm_p2g(const DevPtrs &d, void *st) {
  const int g = std::max(1, std::min(4096, blocks(d.nnodes / 2 + 1, WG)));
  hipLaunchKernelGGL(k_therm_clear, dim3(g), dim3(WG), 0, S(st), d);
  if (d.np > 0) hipLaunchKernelGGL(k_therm_p2g, dim3(blocks(d.np, WG)), dim3(WG), 0, S(st), d);
}
void launch_therm_g2p(const DevPtrs &d, void *st) {
  if (d.np > 0) hipLaunchKernelGGL(k_therm_g2p, dim3(blocks(d.np, WG)), dim3(WG), 0, S(st), d);
}

}  // namespace gy
