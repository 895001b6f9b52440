// gy_device.hpp — device helpers shared by the step kernels of the 3D path
// (gy_kernels.hip) and the plane path (gy_plane.hip): launch constants, fixed-point
// stencil bases, the GIMP weights of one axis, particle BC handlers, the P2G terms of one
// node, the histogram rank of a wave's keys and the active-tile work queue.
#pragma once
#include <hip/hip_runtime.h>

#include "gy_internal.hpp"

namespace gy {
namespace {

constexpr int WG = 256;
constexpr int SCAN_ITEMS = 16;
constexpr int SCAN_BLOCK = WG * SCAN_ITEMS;
constexpr int MAXMAT = 16;
constexpr int CHUNK_ENTRIES = WG;  // list entries per chunk of the fused particle kernel

__device__ __forceinline__ float f32(uint32_t u) { return __uint_as_float(u); }
__device__ __forceinline__ uint32_t u32(float f) { return __float_as_uint(f); }
// The particle state is read once per step: streaming (nontemporal) loads leave L2 to the
// node velocities, tile partials and masks that several workgroups read. Used by the 3D
// particle kernels (measured on the 1M-particle column: fused kernel -2 %, grid kernel
// -12 %, step -3 %); the plane kernel reads plain (there streaming measured 6 % slower).
__device__ __forceinline__ uint32_t ld_stream(const uint32_t *p) { return __builtin_nontemporal_load(p); }
// A load through the constant address space, for tables no kernel of the step writes while
// the reading kernel runs (instance and tile tables, offsets, the active tile list): at a
// workgroup-uniform address it becomes a scalar load, which keeps it out of the in-order
// vector-memory counter that the vector loads a kernel issues ahead of use are waited on.
template <class T>
__device__ __forceinline__ T cld(const T *p) {
#if defined(__HIP_DEVICE_COMPILE__)
  return *reinterpret_cast<const __attribute__((address_space(4))) T *>(reinterpret_cast<uintptr_t>(p));
#else
  return *p;  // the host pass parses device functions but never runs them
#endif
}

// Extended position and velocity (the F_XV words). A position is the 48-bit fixed-point
// number (q << 16) + lo in units of 2^-16 quantum; a step advances it by d quanta with one
// rounding at 2^-16 quantum. A velocity is v_hi + r with r held as a bfloat16 (8 bits below
// v_hi's last bit); vel_add adds an increment as an error-free sum of v_hi and the
// increment plus the carried residual (Knuth's two-sum, then a renormalisation), so the
// rounding of each step's update stays at ~2^-32 of v instead of 2^-24.
__device__ __forceinline__ uint32_t xv_lo(uint32_t w) { return w & 0xffffu; }
__device__ __forceinline__ float xv_res(uint32_t w) { return __uint_as_float(w & 0xffff0000u); }
__device__ __forceinline__ uint32_t xv_pack(uint32_t lo, float res) {
  uint32_t u = __float_as_uint(res);
  u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even at bfloat16
  return (u & 0xffff0000u) | (lo & 0xffffu);
}
__device__ __forceinline__ void advance_q(int &q, uint32_t &lo, float dquanta) {
  const long long Q = ((long long)q << 16) + (long long)lo + __float2ll_rn(dquanta * 65536.0f);
  q = (int)(Q >> 16);
  lo = (uint32_t)(Q & 0xffff);
}
__device__ __forceinline__ void vel_add(float &vh, float &vr, float inc) {
#pragma clang fp contract(off)
  const float s = vh + inc;
  const float bb = s - vh;
  const float err = (vh - (s - bb)) + (inc - bb);
  const float lo = vr + err;
  const float nh = s + lo;
  vr = lo - (nh - s);
  vh = nh;
}

// Stencil base cell of a fixed-point coordinate, clamped to the grid (getGIMPNodes,
// Extrapolation.cpp:658-660: int((x - psep/2 - x0)/h)); `bad` is set when clamped.
__device__ __forceinline__ int stencil_base(int q, int lpq, int F, int maxb, bool &bad) {
  int b = (q - lpq) >> F;
  if (b < 0) { b = 0; bad = true; }
  if (b > maxb) { b = maxb; bad = true; }
  return b;
}

__device__ __forceinline__ int key_of(const InstDesc &I, const int b[3]) {
  const int t = I.tile_base + (b[0] / TILE) + I.nt[0] * ((b[1] / TILE) + I.nt[1] * (b[2] / TILE));
  return t * TILE3 + (b[0] % TILE) + TILE * ((b[1] % TILE) + TILE * (b[2] % TILE));
}

// GIMP weights of the three stencil nodes along one axis (Extrapolation.cpp:556-647 with
// h = 1), for a particle at offset u in [lp, 1+lp) from the base node and lp <= 1/2.
// S[i] = S(d_i), G[i] = dS/dd at d_i = i - u; the physical gradient dN/dx_p is -G/h.
struct Axis {
  float w[3], g[3];
};
// one of three values by a loop index k in {0,1,2}, by value so no array is spilled
__device__ __forceinline__ float sel3(int k, float a0, float a1, float a2) { return k == 0 ? a0 : (k == 1 ? a1 : a2); }
__device__ __forceinline__ Axis gimp3(float u, float lp, float inv_lp, float inv_h) {
  Axis A;
  const float d0 = -u, d1 = 1.0f - u, d2 = 2.0f - u;
  const float e0 = 1.0f + lp + d0;  // distance past -(1+lp)
  // node 0: d0 in (-1-lp, -lp]
  if (d0 <= -1.0f + lp) { A.w[0] = e0 * e0 * (0.25f * inv_lp); } else { A.w[0] = 1.0f + d0; }
  const float G0 = (d0 < -1.0f + lp) ? e0 * (0.5f * inv_lp) : 1.0f;
  // node 1: d1 in (-lp, 1-lp]
  if (d1 <= lp) { A.w[1] = 1.0f - (d1 * d1 + lp * lp) * (0.5f * inv_lp); } else { A.w[1] = 1.0f - d1; }
  const float G1 = (d1 < lp) ? -d1 * inv_lp : -1.0f;
  // node 2: d2 in (1-lp, 2-lp]
  const float e2 = 1.0f + lp - d2;
  float G2;
  if (d2 <= 1.0f + lp) { A.w[2] = e2 * e2 * (0.25f * inv_lp); G2 = -e2 * (0.5f * inv_lp); }
  else { A.w[2] = 0.0f; G2 = 0.0f; }
  A.g[0] = -G0 * inv_h;
  A.g[1] = -G1 * inv_h;
  A.g[2] = -G2 * inv_h;
  return A;
}

// Particle BC handlers, TimeIntegration.cpp:177-296 (phase is ignored by the reference).
__device__ __forceinline__ void particle_bc(int type, const float *bv, float t, float v[3], float a[3],
                                            float vg[3], float s[6], float mass) {
  switch (type) {
    case 33: v[0] = bv[0]; break;
    case 34: v[0] = 0.0f; v[1] = bv[0]; v[2] = 0.0f; break;
    case 35: v[0] = 0.0f; v[1] = 0.0f; v[2] = bv[0]; break;
    case 40: v[0] = bv[0]; v[1] = bv[1]; v[2] = 0.0f; break;
    case 65: v[0] = bv[0] * sinf(2.0f * 3.14159f * bv[1] * t); break;
    case 66: v[1] = bv[0] * sinf(2.0f * 3.14159f * bv[1] * t); vg[1] = 0.0f; a[1] = 0.0f; break;
    case 67:
      s[1] = bv[0] * sinf(2.0f * 3.14159f * bv[1] * t);
      v[0] = v[1] = v[2] = 0.0f; vg[0] = vg[1] = vg[2] = 0.0f; break;
    case 37: a[0] += bv[0] / mass; break;
    case 38: a[0] = 0.0f; a[1] += bv[0] / mass; break;
    case 39: a[2] += bv[0] / mass; break;
    case 43: a[0] += bv[0] / mass; a[1] += bv[1] / mass; break;
    case 63: s[0] = bv[0]; s[1] = bv[1]; break;
    default: break;
  }
}

__device__ __forceinline__ int inst_of_slot(const DevPtrs &d, int s) {
  int lo = 0, hi = d.ninst - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (d.inst[mid].pstart <= s) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// Scatter one particle's P2G terms of one node: m, m*v, and the internal force
// -V * sigma . gradN (Extrapolation.cpp:265-273).
struct P2GTerms {
  float m, mv[3], vs[6];  // vs = -V*sigma
};

__device__ __forceinline__ void node_terms(const P2GTerms &T, float w, float gx, float gy, float gz,
                                           float out[7]) {
  out[0] = w * T.m;
  out[1] = w * T.mv[0];
  out[2] = w * T.mv[1];
  out[3] = w * T.mv[2];
  out[4] = T.vs[0] * gx + T.vs[3] * gy + T.vs[5] * gz;
  out[5] = T.vs[3] * gx + T.vs[1] * gy + T.vs[4] * gz;
  out[6] = T.vs[5] * gx + T.vs[4] * gy + T.vs[2] * gz;
}

// Model-set bits: the fused kernel is instantiated per set of constitutive models a
// context uses, so a single-material run carries only its own model's registers.
constexpr int KM_ELASTIC = 1, KM_NEWTON = 2, KM_PERFECT = 4, KM_DP = 8, KM_GK = 16, KM_PBC = 32,
              KM_TURB = 64;
constexpr int KM_ALL = 127;
// Drucker-Prager multipliers other than the closed form, outside KM_ALL: the reference's
// loop control (a material with dp_return = GY_DP_RETURN_REFERENCE) and the network
// (GY_DP_RETURN_LEARNED); single-body 3D kernels only
constexpr int KM_DPLOOP = 128, KM_DPNN = 256;

// Rank of a lane's key among the equal keys of its wave, with one global atomic per run
// of equal keys (lanes hold consecutive entries of a cell-grouped list, so equal keys
// are adjacent). key < 0: the lane takes no part (only trailing lanes may do so).
__device__ __forceinline__ int wave_hist_rank(int *count, int key) {
  const int lane = threadIdx.x & 63;
  const int prev = __shfl_up(key, 1, 64);
  const bool head = key >= 0 && (lane == 0 || prev != key);
  const unsigned long long heads = __ballot(head);
  const unsigned long long valid = __ballot(key >= 0);
  const unsigned long long upto = (2ull << lane) - 1ull;  // lanes 0..lane
  int base = 0;
  if (head) {
    const unsigned long long later = heads & ~upto;
    const int vend = 64 - __clzll((long long)valid);
    const int end = later ? __ffsll((long long)later) - 1 : vend;
    base = atomicAdd(&count[key], end - lane);
  }
  const unsigned long long mine = heads & upto;
  const int hl = mine ? 63 - __clzll((long long)mine) : lane;
  base = __shfl(base, hl, 64);
  return base + lane - hl;
}

// The same rank in two halves, so the atomic's round trip can overlap other work:
// hist_issue adds the run lengths (head lanes get the old counts in `base`), hist_rank
// then spreads each run's base over its lanes.
__device__ __forceinline__ int hist_issue(int *count, int key, int &hl) {
  const int lane = threadIdx.x & 63;
  const int prev = __shfl_up(key, 1, 64);
  const bool head = key >= 0 && (lane == 0 || prev != key);
  const unsigned long long heads = __ballot(head);
  const unsigned long long valid = __ballot(key >= 0);
  const unsigned long long upto = (2ull << lane) - 1ull;
  int base = 0;
  if (head) {
    const unsigned long long later = heads & ~upto;
    const int vend = 64 - __clzll((long long)valid);
    const int end = later ? __ffsll((long long)later) - 1 : vend;
    base = atomicAdd(&count[key], end - lane);
  }
  const unsigned long long mine = heads & upto;
  hl = mine ? 63 - __clzll((long long)mine) : lane;
  return base;
}
__device__ __forceinline__ int hist_rank(int base, int hl) {
  return __shfl(base, hl, 64) + (int)(threadIdx.x & 63) - hl;
}

// The particle kernels' work queue over the active tiles, longest first (entry-count
// classes), so the claims that end a launch are the short ones; one
// counter for the whole device (balance across the XCDs matters more here than keeping
// neighbouring tiles in one L2: measured). Bounds read once per launch.
struct TileQueue {
  int nh, n;
  __device__ __forceinline__ explicit TileQueue(const DevPtrs &d) : nh(d.ctrl->heavy_tiles), n(d.ctrl->active_tiles) {}
  // tile of claim k, -1 past the end
  __device__ __forceinline__ int tile(const DevPtrs &d, int k) const {
    if (k < nh) return d.active[k];
    if (k < n) return d.active[n - 1 - (k - nh)];
    return -1;
  }
};

// The box (tile coordinates: box[0..2] lo, box[3..5] hi) of the tiles whose partials a
// grid kernel gathers this step, written by the binning that fed the particle pass (Ctrl
// tbox); the whole grid in a context of several instances (each has its own tile
// coordinates).
__device__ __forceinline__ void tile_box(const DevPtrs &d, int epoch, int box[6]) {
#pragma unroll
  for (int a = 0; a < 3; a++) {
    box[a] = 0;
    box[3 + a] = TBOX_BIG;
  }
  if (d.ninst != 1) return;
  const int *bx = d.ctrl->tbox[epoch & 1];
#pragma unroll
  for (int a = 0; a < 3; a++) {
    box[a] = bx[a];
    box[3 + a] = TBOX_BIG - bx[3 + a];
  }
}

// Whether actual node a (one axis) can hold partials of a tile in [lo, hi] on that axis,
// for tiles of E cells per edge: its candidate tiles are (a + 1) / E - 1 and (a + 1) / E.
// A periodic-class member outside the box has no partials, and its gather (a chain of two
// dependent loads) is skipped: the mirrors of the nodes next to a periodic face usually lie
// at the far side of the grid, away from the particles.
template <int E>
__device__ __forceinline__ bool in_tile_box(int a, int lo, int hi) {
  const int t = (a + 1) / E;
  return t >= lo && t - 1 <= hi;
}

// calculateColoumbFrictionContactNormals (Grid.cpp:247-308) for one body: the contact
// normal n + min(mu, mu0) * (omega x ...) with the reference's component formulas.
__device__ __forceinline__ void contact_normal(const float rel[3], const float n[3], float mu, float cn[3]) {
  float om[3] = {rel[1] * n[2] - rel[2] * n[1], rel[0] * n[2] - rel[2] * n[0], rel[0] * n[1] - rel[1] * n[0]};
  const float omag = sqrtf(om[0] * om[0] + om[1] * om[1] + om[2] * om[2]) + 0.001f;
  for (int a = 0; a < 3; a++) om[a] /= omag;
  const float mu0 = omag / (rel[0] * n[0] + rel[1] * n[1] + rel[2] * n[2]) + 0.001f;
  const float on[3] = {om[2] * n[1] - om[1] * n[2], om[2] * n[0] - om[0] * n[2], om[1] * n[0] - om[0] * n[1]};
  const float m = fminf(mu, mu0);
  for (int a = 0; a < 3; a++) cn[a] = n[a] + m * on[a];
}

// The direct-scatter centre-of-mass sums of canonical node ci (added to a7) and its body
// sums (into ab), read and cleared for the next step with 16-byte accesses.
__device__ __forceinline__ void take_node_sums(const DevPtrs &d, long long ci, float a7[7], float ab[MB_ACC2_W]) {
  float4 *ap = reinterpret_cast<float4 *>(d.acc + ci * NODE_W);
  const float4 x0 = ap[0], x1 = ap[1];
  a7[0] += x0.x; a7[1] += x0.y; a7[2] += x0.z; a7[3] += x0.w;
  a7[4] += x1.x; a7[5] += x1.y; a7[6] += x1.z;
  ap[0] = make_float4(0.f, 0.f, 0.f, 0.f);
  ap[1] = make_float4(0.f, 0.f, 0.f, 0.f);
  float4 *bp = reinterpret_cast<float4 *>(d.acc2 + ci * MB_ACC2_W);
#pragma unroll
  for (int q = 0; q < MB_ACC2_W / 4; q++) {
    const float4 b = bp[q];
    ab[4 * q] = b.x; ab[4 * q + 1] = b.y; ab[4 * q + 2] = b.z; ab[4 * q + 3] = b.w;
    bp[q] = make_float4(0.f, 0.f, 0.f, 0.f);
  }
}

// Contact solve of one actual node (Grid.cpp:9-245) and its BCs on the centre of mass and
// both bodies (Grid.cpp:967-982, 1012-1031): m, a7 the centre-of-mass sums of its periodic
// class (m, p, f_int), ab the body sums [m, p, f_int, rigid mass] x 2, vc and ac the
// centre-of-mass velocity p/m and acceleration f/m + g. Writes VEL (centre of mass) and
// VELB (body velocities and accelerations) of the node and clears its normal sums.
__device__ __forceinline__ void mb_solve_node(const DevPtrs &d, const InstDesc &I, long long ni, float m,
                                              const float a7[7], const float ab[MB_ACC2_W], const float vc[3],
                                              const float ac[3], float dt, float kick) {
  // body sums: [m, p, f_int, rigid mass]; gravity only on the non-rigid mass
  float mb[2], pb[2][3], fb[2][3];
  auto body_sums = [&]() {
    for (int s2 = 0; s2 < 2; s2++) {
      mb[s2] = ab[8 * s2];
      for (int k = 0; k < 3; k++) {
        pb[s2][k] = ab[8 * s2 + 1 + k];
        fb[s2][k] = ab[8 * s2 + 4 + k] + (ab[8 * s2] - ab[8 * s2 + 7]) * I.gravity[k];
      }
    }
  };
  const int4 cl = d.claim[ni];
  body_sums();
  float nr[2][3];
  for (int k = 0; k < 3; k++) {
    nr[0][k] = d.nrm[ni * 6 + k];
    nr[1][k] = d.nrm[ni * 6 + 3 + k];
    d.nrm[ni * 6 + k] = 0.f;
    d.nrm[ni * 6 + 3 + k] = 0.f;
  }
  float vb[2][3] = {{0.f, 0.f, 0.f}, {0.f, 0.f, 0.f}}, acb[2][3] = {{0.f, 0.f, 0.f}, {0.f, 0.f, 0.f}};
  if (cl.y == -1) {  // one body: its fields are the centre-of-mass sums
    mb[0] = m;
    for (int k = 0; k < 3; k++) {
      pb[0][k] = a7[1 + k];
      fb[0][k] = a7[4 + k] + (m - ab[7]) * I.gravity[k];
    }
  }
  if (m != 0.0f) {
    for (int s2 = 0; s2 < 2; s2++) {  // normalise (Grid.cpp:19-43)
      const float mag = sqrtf(nr[s2][0] * nr[s2][0] + nr[s2][1] * nr[s2][1] + nr[s2][2] * nr[s2][2]);
      for (int k = 0; k < 3; k++) nr[s2][k] = (mag > 1e-8f) ? nr[s2][k] / mag : 0.f;
    }
    for (int k = 0; k < 3; k++) nr[1][k] = -nr[0][k];  // Grid.cpp:45-47
    auto free_body = [&](int s2) {
      if (mb[s2] == 0.0f) return;
      for (int k = 0; k < 3; k++) {
        acb[s2][k] = fb[s2][k] / mb[s2];
        vb[s2][k] = pb[s2][k] / mb[s2] + acb[s2][k] * kick;
      }
    };
    if (cl.y != -1) {
      float rel[2][3] = {{0.f, 0.f, 0.f}, {0.f, 0.f, 0.f}}, ap[2] = {0.f, 0.f};
      for (int s2 = 0; s2 < 2; s2++)
        if (mb[s2] != 0.0f) {
          for (int k = 0; k < 3; k++) rel[s2][k] = pb[s2][k] / mb[s2] - vc[k];
          ap[s2] = rel[s2][0] * nr[s2][0] + rel[s2][1] * nr[s2][1] + rel[s2][2] * nr[s2][2];
        }
      if (ap[0] > 0.01f || ap[1] > 0.01f) {
        if (cl.z || cl.w) {  // a rigid body: the other takes its velocity (Grid.cpp:78-105)
          const int r = cl.z ? 0 : 1, o2 = 1 - r;
          if (mb[r] != 0.0f) {
            for (int k = 0; k < 3; k++) vb[r][k] = pb[r][k] / mb[r];
            if (mb[o2] != 0.0f)
              for (int k = 0; k < 3; k++)
                vb[o2][k] = pb[o2][k] / mb[o2] - (pb[o2][k] - mb[o2] * vb[r][k]) / mb[o2];
          }
        } else {  // two deformable bodies (Grid.cpp:106-170)
          for (int s2 = 0; s2 < 2; s2++) {
            if (mb[s2] == 0.0f) continue;
            float cn[3];
            contact_normal(rel[s2], nr[s2], I.mu, cn);
            for (int k = 0; k < 3; k++) {
              vb[s2][k] = pb[s2][k] / mb[s2] - ap[s2] * cn[k];
              acb[s2][k] = fb[s2][k] / mb[s2] - (ap[s2] * cn[k]) / dt;
              vb[s2][k] += acb[s2][k] * kick;
            }
          }
        }
      } else {
        free_body(0);
        free_body(1);
      }
    } else {
      free_body(0);
    }
  }
  float vv[3] = {vc[0], vc[1], vc[2]}, aa[3] = {ac[0], ac[1], ac[2]};
  const int t = d.bc_type[ni];
  if (t == 1 || t == 2 || t == 11) {  // Grid.cpp:967-982, 1012-1031
    const float b0 = d.bc_val[2 * ni], b1v = d.bc_val[2 * ni + 1];
    const int ax0 = (t == 2) ? 1 : 0;
    const float val0 = b0;
    vv[ax0] = val0; aa[ax0] = 0.f;
    for (int s2 = 0; s2 < 2; s2++) { vb[s2][ax0] = val0; acb[s2][ax0] = 0.f; }
    if (t == 11) {
      vv[1] = b1v; aa[1] = 0.f;
      for (int s2 = 0; s2 < 2; s2++) { vb[s2][1] = b1v; acb[s2][1] = 0.f; }
    }
  }
  float4 *vp = reinterpret_cast<float4 *>(d.vel + ni * NODE_W);
  vp[0] = make_float4(vv[0], vv[1], vv[2], aa[0]);
  vp[1] = make_float4(aa[1], aa[2], m, 0.f);
  float *vq = d.velb + ni * MB_VELB_W;
  for (int k = 0; k < 3; k++) {
    vq[k] = vb[0][k];
    vq[3 + k] = acb[0][k];
    vq[6 + k] = vb[1][k];
    vq[9 + k] = acb[1][k];
  }
}

}  // namespace
}  // namespace gy
