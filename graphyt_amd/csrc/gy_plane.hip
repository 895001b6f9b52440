// gy_plane.hip — HIP kernels (gfx950) of the plane MPM step (Parameters::is3D false).
//
// The reference's plane path works per unit depth on one node layer: shape_z = 1 and
// gradShape_z = 0 in both transfers (Extrapolation.cpp:184-188, 304-308), the 9-node
// stencil of getGIMPNodes' plane branch (Extrapolation.cpp:683-705), the 4x4 cell window of
// the node gather (Extrapolation.cpp:89-121), periodic pairs along x and y only
// (Grid.cpp:604-800), no z advance of velocity or position and no z wrap
// (TimeIntegration.cpp:30-40, 98-108), and the plane branches of the Newtonian, turbulent
// and perfect-plastic models and of Grady-Kipp damage. The nodal solve and the BCs are
// those of the 3D path.
//
// Same step structure and buffers as the 3D kernels (gy_kernels.hip), on plane tiles of
// PT x PT cells (TILE3 keys, so the shared scan/place kernels and the active-tile queue
// serve both) and plane bricks of PB x PB canonical nodes (one wave each):
//   k_plane_grid    persistent over the flagged bricks: gather the tile partials of every
//                   node of the periodic class (at most 2 x 2 tiles per node), solve, BCs,
//                   fan-out to the mirror nodes
//   k_plane_g2p2g   persistent over the active tiles: G2P from the LDS-staged node
//                   velocities, update, models, wrap, new key + histogram, and the 9-node
//                   P2G of the updated particle into an FP64 LDS region [o-1, o+PT+3)^2,
//                   written out as FP32 tile partials (core and halo shell) with node masks
#include <hip/hip_runtime.h>

#include <algorithm>

#include "gy_device.hpp"
#include "gy_internal.hpp"
#include "gy_models.hpp"
#include "gy_scan.hpp"

namespace gy {
namespace {

constexpr int PCHUNK = CHUNK_ENTRIES;
constexpr int PG_N = PG_R * PG_R;

__device__ __forceinline__ int plane_key(const InstDesc &I, int b0, int b1) {
  const int t = I.tile_base + b0 / PT + I.nt[0] * (b1 / PT);
  return t * TILE3 + (b0 % PT) + PT * (b1 % PT);
}

// Canonical (periodic-class representative) node of actual node (g0, g1): index i folded
// to i - (N-3) when i >= N-3 (Grid.cpp:604-800).
__device__ __forceinline__ long long plane_canon(const InstDesc &I, int g0, int g1, int &c0, int &c1) {
  c0 = (g0 >= I.cper[0]) ? g0 - I.cper[0] : g0;
  c1 = (g1 >= I.cper[1]) ? g1 - I.cper[1] : g1;
  return I.node_base + c0 + (long long)I.n[0] * c1;
}

__device__ __forceinline__ void plane_flag(uint8_t *brick_flag, const InstDesc &I, int c0, int c1) {
  brick_flag[I.brick_base + c0 / PB + I.nb[0] * (c1 / PB)] = 1;
}

// Keys + histogram of the current buffer (first step after an upload).
__global__ __launch_bounds__(WG) void k_plane_keys(DevPtrs d) {
  const int s = blockIdx.x * WG + threadIdx.x;
  if (s >= d.np) return;
  const int par = d.ctrl->parity;
  const uint32_t *A = (par ? d.pbuf[1] : d.pbuf[0]);
  const InstDesc &I = d.inst[inst_of_slot(d, s)];
  bool bad = false;
  const int b0 = stencil_base((int)A[F_QX * d.cap + s], I.lpq, I.fbits, I.n[0] - 3, bad);
  const int b1 = stencil_base((int)A[F_QY * d.cap + s], I.lpq, I.fbits, I.n[1] - 3, bad);
  const int k = plane_key(I, b0, b1);
  (par ? d.key[1] : d.key[0])[s] = k;
  d.rank[s] = atomicAdd(&d.count[k], 1);
  if (bad) { atomicAdd(&d.ctrl->oor, 1); atomicOr(&d.ctrl->status, 1); }
}

// ---------------------------------------------------------------------------------------
// Grid

// Sum of the tile partials covering actual node (a0, a1): the node lies in the P2G region
// [o-1, o+PT+3) of the tiles with o = PT*(tt-1) and PT*tt, tt = (a+1)/PT, at region
// coordinate l = a - o + 1 (< PR). Coordinates 1..PC are the tile's core partials, the
// rest its halo shell. A tile contributes when it wrote partials this step and its node
// mask has the node's bit; fixed order (y, x offset): deterministic.
__device__ __forceinline__ void plane_gather(const DevPtrs &d, const InstDesc &I, int epoch, int a0, int a1,
                                             float acc7[7]) {
  const int t0 = (a0 + 1) / PT, t1 = (a1 + 1) / PT;
  bool use[4];
  const float4 *pp[4];
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const int tx = t0 - 1 + (k & 1), ty = t1 - 1 + (k >> 1);
    const int lx = a0 - PT * tx + 1, ly = a1 - PT * ty + 1;  // in [0, 2*PT)
    const bool in = tx >= 0 && tx < I.nt[0] && ty >= 0 && ty < I.nt[1] && lx < PR && ly < PR;
    const int t = in ? I.tile_base + tx + I.nt[0] * ty : I.tile_base;
    const int l = in ? lx + PR * ly : 0;
    const uint64_t word = d.nodemask[(long long)t * (P2G_NODES / 64) + (l >> 6)];
    use[k] = in && d.tile_epoch[t] == epoch && ((word >> (l & 63)) & 1ull);
    const bool core = lx >= 1 && lx <= PC && ly >= 1 && ly <= PC;
    const int c = (lx - 1) + PC * (ly - 1);
    pp[k] = reinterpret_cast<const float4 *>(core ? d.pcore + (long long)t * CORE_W + 8 * c
                                                  : d.partial + (long long)t * PARTIAL_W + 8 * l);
  }
#pragma unroll
  for (int k = 0; k < 4; k++) {
    if (!use[k]) continue;
    const float4 u = pp[k][0], w = pp[k][1];
    acc7[0] += u.x; acc7[1] += u.y; acc7[2] += u.z; acc7[3] += u.w;
    acc7[4] += w.x; acc7[5] += w.y; acc7[6] += w.z;
  }
}

// Nodal solve of one flagged plane brick (one wave, lane = node of the 8 x 8 brick): the
// P2G sums of the periodic class, v = p/m, a = f/m + g, the kick (half at step 1), BCs per
// actual node (Grid.cpp:9-245, 862-1031), fan-out to every node of the class.
__device__ __forceinline__ void plane_brick(const DevPtrs &d, int gb, int lane, int step, int epoch,
                                            const int box[6]) {
  const InstDesc &I = d.inst[d.brick_inst[gb]];
  const int lb = gb - I.brick_base;
  const int bx = lb % I.nb[0], by = lb / I.nb[0];
  const int c0 = bx * PB + (lane & 7), c1 = by * PB + (lane >> 3);
  if (c0 < I.cper[0] && c1 < I.cper[1]) {
    float a7[7] = {0, 0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int my = 0; my < 2; my++) {
      if (my && c1 >= 3) break;
      const int a1 = c1 + my * I.cper[1];
      if (!in_tile_box<PT>(a1, box[1], box[4])) continue;
#pragma unroll
      for (int mx = 0; mx < 2; mx++) {
        if (mx && c0 >= 3) break;
        const int a0 = c0 + mx * I.cper[0];
        if (!in_tile_box<PT>(a0, box[0], box[3])) continue;
        plane_gather(d, I, epoch, a0, a1, a7);
      }
    }
    // particles scattered straight to the grid (outside their tile's region)
    const long long ci = I.node_base + c0 + (long long)I.n[0] * c1;
    float4 *ap = reinterpret_cast<float4 *>(d.acc + ci * NODE_W);
    const float4 x0 = ap[0], x1 = ap[1];
    if (x0.x != 0.f || x0.y != 0.f || x0.z != 0.f || x0.w != 0.f || x1.x != 0.f || x1.y != 0.f || x1.z != 0.f) {
      a7[0] += x0.x; a7[1] += x0.y; a7[2] += x0.z; a7[3] += x0.w;
      a7[4] += x1.x; a7[5] += x1.y; a7[6] += x1.z;
      ap[0] = make_float4(0.f, 0.f, 0.f, 0.f);
      ap[1] = make_float4(0.f, 0.f, 0.f, 0.f);
    }
    const float m = a7[0];
    float v[3] = {0.f, 0.f, 0.f}, acc[3] = {0.f, 0.f, 0.f};
    if (m != 0.0f) {
      const float kick = (step == 1) ? 0.5f * I.dt : I.dt;
#pragma unroll
      for (int k = 0; k < 3; k++) {
        acc[k] = (a7[4 + k] + m * I.gravity[k]) / m;
        v[k] = a7[1 + k] / m + acc[k] * kick;
      }
    }
#pragma unroll
    for (int my = 0; my < 2; my++) {
      if (my && c1 >= 3) break;
      const int n1 = c1 + my * I.cper[1];
#pragma unroll
      for (int mx = 0; mx < 2; mx++) {
        if (mx && c0 >= 3) break;
        const int n0 = c0 + mx * I.cper[0];
        const long long ni = I.node_base + n0 + (long long)I.n[0] * n1;
        float vv[3] = {v[0], v[1], v[2]}, aa[3] = {acc[0], acc[1], acc[2]};
        const int t = d.bc_type[ni];
        if (t == 1 || t == 2 || t == 11) {
          const float b0 = d.bc_val[2 * ni], b1 = d.bc_val[2 * ni + 1];
          if (t == 1) { vv[0] = b0; aa[0] = 0.f; }
          else if (t == 2) { vv[1] = b0; aa[1] = 0.f; }
          else { vv[0] = b0; vv[1] = b1; aa[0] = 0.f; aa[1] = 0.f; }
        }
        float4 *vp = reinterpret_cast<float4 *>(d.vel + ni * NODE_W);
        vp[0] = make_float4(vv[0], vv[1], vv[2], aa[0]);
        vp[1] = make_float4(aa[1], aa[2], m, 0.f);
      }
    }
  }
  if (lane == 0) d.brick_flag[gb] = 0;
}

// The plane grid update as block bid of nblk: one wave per listed brick.
__device__ __forceinline__ void plane_grid_body(const DevPtrs &d, int bid, int nblk) {
  // the particle kernel that follows walks the active tile list from the start
  if (bid == 0 && threadIdx.x == 0) d.ctrl->next = 0;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int n = d.ctrl->active_bricks, step = d.ctrl->step, epoch = d.ctrl->epoch;
  int box[6];
  tile_box(d, epoch, box);  // plane tiles: x, y tile coordinates, one layer
  for (int k = bid * (WG / 64) + wave; k < n; k += nblk * (WG / 64))
    plane_brick(d, d.brick_list[k], lane, step, epoch, box);
}

__global__ __launch_bounds__(WG) void k_plane_grid(DevPtrs d) { plane_grid_body(d, blockIdx.x, gridDim.x); }

// The second scan kernel of step s beside the plane grid update of step s + 1, as the 3D
// k_scan_final_grid: blocks [0, ngrid) update the grid, the others scan.
__global__ __launch_bounds__(WG) void k_plane_scan_final_grid(DevPtrs d, int ngrid) {
  if ((int)blockIdx.x < ngrid) plane_grid_body(d, blockIdx.x, ngrid);
  else scan_final_body(d, blockIdx.x - ngrid, gridDim.x - ngrid, false);
}

// ---------------------------------------------------------------------------------------
// Particles

// P2G record of one particle handed from phase A to phase B in LDS: three 16-byte words
// (q0 q1 m mv0 | mv1 mv2 vs0 vs1 | vs3 vs4 vs5 -), vs = -V*sigma (sigma_zz does not reach
// the plane's nodes: its gradient factor is gradShape_z = 0). One word of padding after
// every 16 columns, so the 16 columns of a phase-A pass land on different banks.
struct PlaneRec {
  float4 w[3 * PCHUNK + PCHUNK / 16];
  __device__ __forceinline__ static int wi(int col) { return 3 * col + (col >> 4); }
};

// FP64 LDS accumulator of the P2G region, x fastest.
constexpr int PACC_N = PR_N;

// P2G of one particle straight into the global accumulators (its stencil left the tile's
// LDS region, e.g. after a periodic wrap). Out of line, with plain pointer arguments: a
// reference to the kernel's DevPtrs argument would make every lane keep a private copy.
__device__ __noinline__ void plane_p2g_global(float *acc, uint8_t *brick_flag, const InstDesc *Ip, int b0, int b1,
                                              Axis W0, Axis W1,
                                              float m, float mv0, float mv1, float mv2, float vs0, float vs1,
                                              float vs3, float vs4, float vs5) {
  const InstDesc &I = *Ip;
#pragma unroll
  for (int j = 0; j < 3; j++)
#pragma unroll
    for (int i = 0; i < 3; i++) {
      const float w = W0.w[i] * W1.w[j];
      if (w == 0.0f) continue;
      const float gx = W0.g[i] * W1.w[j], gy = W0.w[i] * W1.g[j];
      const float t[7] = {w * m, w * mv0, w * mv1, w * mv2, vs0 * gx + vs3 * gy, vs3 * gx + vs1 * gy,
                          vs5 * gx + vs4 * gy};
      int c0, c1;
      const long long ci = plane_canon(I, b0 + i, b1 + j, c0, c1);
#pragma unroll
      for (int f = 0; f < 7; f++) unsafeAtomicAdd(&acc[ci * NODE_W + f], t[f]);
      plane_flag(brick_flag, I, c0, c1);
    }
}

// Phase A for list entry i (slot `slot` of the source buffer): G2P + update + models +
// store at sorted position i (UPDATE), and the P2G record of the (new) state into column
// `col` of s_rec. Returns the new key (UPDATE) or -1.
// MB (contact between objects): G2P reads the velocity field of the body slot the
// particle's object holds at each node (s_vel: body 1, s_vel2: body 2, s_obj: the slots'
// objects; Extrapolation.cpp:288-367). The record is written when s_rec is given.
template <bool UPDATE, int MASK, bool MB = false>
__device__ __forceinline__ int plane_update(const DevPtrs &d, const InstDesc &I, int o0, int o1, double tl, int i,
                                            int slot, const float4 (*s_vel)[2], float3 vref, const MatDev *s_mat, int par,
                                            PlaneRec *s_rec, int col, const float4 (*s_vel2)[2] = nullptr,
                                            const int (*s_obj)[PG_N] = nullptr) {
  const uint32_t *A = (par ? d.pbuf[1] : d.pbuf[0]);
  uint32_t *B = (par ? d.pbuf[0] : d.pbuf[1]);
  const long long cap = d.cap;
  int q0 = (int)A[F_QX * cap + slot], q1 = (int)A[F_QY * cap + slot];
  const uint32_t tag = A[F_TAG * cap + slot];
  const MatDev &M = s_mat[tag_mat(tag)];
  float v[3], s[6];
#pragma unroll
  for (int a = 0; a < 3; a++) v[a] = f32(A[(F_VX + a) * cap + slot]);
#pragma unroll
  for (int a = 0; a < 6; a++) s[a] = f32(A[(F_S0 + a) * cap + slot]);
  float rho = f32(A[F_RHO * cap + slot]);
  const float mass = f32(A[F_MASS * cap + slot]);
  int knew = -1;

  if (UPDATE) {
    const int out = i;
    uint32_t xv[3];  // position fraction | velocity residual (gy_internal.hpp)
#pragma unroll
    for (int a = 0; a < 3; a++) xv[a] = A[(F_XV0 + a) * cap + slot];
    bool bad = false;
    // ---- G2P (Extrapolation.cpp:127-163, 288-367 with shape_z = 1, gradShape_z = 0) ----
    const int b0 = stencil_base(q0, I.lpq, I.fbits, I.n[0] - 3, bad);
    const int b1 = stencil_base(q1, I.lpq, I.fbits, I.n[1] - 3, bad);
    const Axis W0 = gimp3((float)(q0 - (b0 << I.fbits)) * I.inv_2F, I.lp, I.inv_lp, I.inv_h);
    const Axis W1 = gimp3((float)(q1 - (b1 << I.fbits)) * I.inv_2F, I.lp, I.inv_lp, I.inv_h);
    int r0 = b0 - o0, r1 = b1 - o1;
    if ((unsigned)r0 >= (unsigned)PT || (unsigned)r1 >= (unsigned)PT) {
      // a particle listed under a tile it is not in: never expected; keep LDS reads in range
      atomicOr(&d.ctrl->status, 2);
      r0 = min(max(r0, 0), PT - 1);
      r1 = min(max(r1, 0), PT - 1);
    }
    float vg[3] = {0, 0, 0}, ap[3] = {0, 0, 0}, Lx[3] = {0, 0, 0}, Ly[3] = {0, 0, 0};
#pragma unroll 1
    for (int jj = 0; jj < 3; jj++) {
      float S0[3] = {0, 0, 0}, S1[3] = {0, 0, 0}, Sa[3] = {0, 0, 0};
      const int row = r0 + PG_R * (r1 + jj);
#pragma unroll
      for (int ii = 0; ii < 3; ii++) {
        float4 a0 = s_vel[row + ii][0], a1 = s_vel[row + ii][1];
        if (MB) {
          const int n = row + ii, ob = tag_obj(tag);
          const int sl = (s_obj[0][n] == ob) ? 0 : (s_obj[1][n] == ob) ? 1 : 2;
          const float4 z4 = make_float4(0.f, 0.f, 0.f, 0.f);
          if (sl != 0) {
            a0 = (sl == 1) ? s_vel2[n][0] : z4;
            a1 = (sl == 1) ? s_vel2[n][1] : z4;
          }
        }
        const float vn[3] = {a0.x, a0.y, a0.z}, an[3] = {a0.w, a1.x, a1.y};
        const float wx = W0.w[ii], gx = W0.g[ii];
#pragma unroll
        for (int c = 0; c < 3; c++) {
          S0[c] += wx * vn[c];
          S1[c] += gx * vn[c];
          Sa[c] += wx * an[c];
        }
      }
      const float wy = sel3(jj, W1.w[0], W1.w[1], W1.w[2]), gy = sel3(jj, W1.g[0], W1.g[1], W1.g[2]);
#pragma unroll
      for (int c = 0; c < 3; c++) {
        vg[c] += wy * S0[c];
        ap[c] += wy * Sa[c];
        Lx[c] += wy * S1[c];
        Ly[c] += gy * S0[c];
      }
    }
    vg[0] += vref.x;
    vg[1] += vref.y;
    vg[2] += vref.z;
    float er[6];
    er[0] = Lx[0];
    er[1] = Ly[1];
    er[2] = 0.0f;
    er[3] = 0.5f * (Lx[1] + Ly[0]);
    er[4] = 0.5f * Ly[2];
    er[5] = 0.5f * Lx[2];
    const float rho_dot = -rho * (Lx[0] + Ly[1]);

    const int bct = (MASK & KM_PBC) ? tag_bct(tag) : 0;
    float bv[4] = {0, 0, 0, 0};
    const int pid = (int)A[F_PID * cap + slot];
    if (MASK & KM_PBC) {
      if (bct) {
#pragma unroll
        for (int k = 0; k < 4; k++) bv[k] = d.pbc_val[4 * (long long)pid + k];
        const float v0[3] = {v[0], v[1], v[2]};
        particle_bc(bct, bv, (float)tl, v, ap, vg, s, mass);  // phase 4 (Solver.cpp:219-223)
#pragma unroll
        for (int a = 0; a < 3; a++)
          if (v[a] != v0[a]) xv[a] &= 0xffffu;  // a prescribed velocity has no residual
      }
    }
    // ---- particle update (TimeIntegration.cpp:11-117): x and y only ----
    const float dt = I.dt;
    rho += rho_dot * dt;
    if (!tag_hold(tag)) {
      int qq[2] = {q0, q1};
#pragma unroll
      for (int a = 0; a < 2; a++) {
        float vr = xv_res(xv[a]);
        uint32_t lo = xv_lo(xv[a]);
        vel_add(v[a], vr, (ap[a] - I.damping * v[a]) * dt);
        advance_q(qq[a], lo, vg[a] * I.qscale);
        xv[a] = xv_pack(lo, vr);
      }
      q0 = qq[0];
      q1 = qq[1];
    }
    float e[6];
#pragma unroll
    for (int a = 0; a < 6; a++) e[a] = f32(A[(F_E0 + a) * cap + slot]) + er[a] * dt;
#pragma unroll
    for (int a = 0; a < 6; a++) B[(F_E0 + a) * cap + out] = u32(e[a]);
    if (!M.rigid) {
      if ((MASK & KM_ELASTIC) && M.model == 1) cm_elastic(M, er, s, dt);
      if ((MASK & KM_NEWTON) && M.model == 2) cm_newtonian<true>(M, er, e, s, dt);
      if ((MASK & (KM_PERFECT | KM_DP)) && M.carry_plastic) {
        float p[6];
#pragma unroll
        for (int a = 0; a < 6; a++) p[a] = f32(A[(F_P0 + a) * cap + slot]);
        float epb = f32(A[F_EPB * cap + slot]);
        if ((MASK & KM_PERFECT) && M.model == 3) cm_perfect<true>(M, er, s, p, epb, dt);
        if ((MASK & KM_DP) && M.model == 4) cm_drucker_prager(M, er, s, p, epb, dt);
#pragma unroll
        for (int a = 0; a < 6; a++) B[(F_P0 + a) * cap + out] = u32(p[a]);
        B[F_EPB * cap + out] = u32(epb);
      }
      if (MASK & (KM_GK | KM_TURB)) {
        if (M.carry_damage) {
          float dmg = f32(A[F_DMG * cap + slot]), smax = f32(A[F_SMAX * cap + slot]);
          if ((MASK & KM_TURB) && M.model == 5) cm_turbulent<true>(M, er, e, v, mass, dmg, s, dt);
          if ((MASK & KM_GK) && M.gk) gk_damage_plane(M, s, dmg, smax, mass, rho, dt);
          B[F_DMG * cap + out] = u32(dmg);
          B[F_SMAX * cap + out] = u32(smax);
        }
      }
    } else {
      // rigid particles keep their stress, plastic and damage fields
      if (M.carry_plastic) {
#pragma unroll
        for (int a = 0; a < 6; a++) B[(F_P0 + a) * cap + out] = A[(F_P0 + a) * cap + slot];
        B[F_EPB * cap + out] = A[F_EPB * cap + slot];
      }
      if (M.carry_damage) {
        B[F_DMG * cap + out] = A[F_DMG * cap + slot];
        B[F_SMAX * cap + out] = A[F_SMAX * cap + slot];
      }
    }
    if (!(MASK & (KM_PERFECT | KM_DP)) && M.carry_plastic && !M.rigid) {
#pragma unroll
      for (int a = 0; a < 6; a++) B[(F_P0 + a) * cap + out] = A[(F_P0 + a) * cap + slot];
      B[F_EPB * cap + out] = A[F_EPB * cap + slot];
    }
    if (!(MASK & (KM_GK | KM_TURB)) && M.carry_damage && !M.rigid) {
      B[F_DMG * cap + out] = A[F_DMG * cap + slot];
      B[F_SMAX * cap + out] = A[F_SMAX * cap + slot];
    }
    // periodic wrap with period L-2h along x and y (TimeIntegration.cpp:80-97)
    if (q0 > I.wrapq[0]) q0 = I.oneq + q0 % I.wrapq[0];
    if (q0 < I.oneq) q0 += I.wrapq[0] - I.oneq;
    if (q1 > I.wrapq[1]) q1 = I.oneq + q1 % I.wrapq[1];
    if (q1 < I.oneq) q1 += I.wrapq[1] - I.oneq;
    if (MASK & KM_PBC) {
      if (bct) {
        const float v0[3] = {v[0], v[1], v[2]};
        particle_bc(bct, bv, (float)tl, v, ap, vg, s, mass);  // phase 5 (Solver.cpp:249-253)
#pragma unroll
        for (int a = 0; a < 3; a++)
          if (v[a] != v0[a]) xv[a] &= 0xffffu;
      }
    }

    // ---- store the new state at its sorted position ----
    B[F_QX * cap + out] = (uint32_t)q0;
    B[F_QY * cap + out] = (uint32_t)q1;
    B[F_QZ * cap + out] = A[F_QZ * cap + slot];
#pragma unroll
    for (int a = 0; a < 3; a++) B[(F_XV0 + a) * cap + out] = xv[a];
#pragma unroll
    for (int a = 0; a < 3; a++) B[(F_VX + a) * cap + out] = u32(v[a]);
#pragma unroll
    for (int a = 0; a < 6; a++) B[(F_S0 + a) * cap + out] = u32(s[a]);
    B[F_RHO * cap + out] = u32(rho);
    B[F_MASS * cap + out] = u32(mass);
    B[F_TAG * cap + out] = tag;
    B[F_PID * cap + out] = (uint32_t)pid;
    if (d.thermal) {  // temperature and gradient travel with the particle (gy_thermal.hip)
#pragma unroll
      for (int a = 0; a < 4; a++) B[(F_T + a) * cap + out] = A[(F_T + a) * cap + slot];
    }

    // ---- new binning key ----
    bad = false;
    const int n0 = stencil_base(q0, I.lpq, I.fbits, I.n[0] - 3, bad);
    const int n1 = stencil_base(q1, I.lpq, I.fbits, I.n[1] - 3, bad);
    knew = plane_key(I, n0, n1);
    (par ? d.key[0] : d.key[1])[out] = knew;
    if (bad) { atomicAdd(&d.ctrl->oor, 1); atomicOr(&d.ctrl->status, 1); }
  }

  if (s_rec == nullptr) return knew;
  // ---- P2G record of the (new) state (Extrapolation.cpp:166-286) ----
  const float mp = M.rigid ? 1.0f : mass;  // rigid particles scatter unit mass (Extrapolation.cpp:183)
  const float nV = -(mp / rho);
  float4 *w = &s_rec->w[PlaneRec::wi(col)];
  w[0] = make_float4(__int_as_float(q0), __int_as_float(q1), mp, mp * v[0]);
  w[1] = make_float4(mp * v[1], mp * v[2], nV * s[0], nV * s[1]);
  w[2] = make_float4(nV * s[3], nV * s[4], nV * s[5], 0.0f);
  return knew;
}

// Contact G2P staging of node n of a plane tile's G2P region: both body slots' velocity
// and acceleration and the objects that hold them; zero at nodes without mass, which G2P
// skips (Extrapolation.cpp:290).
__device__ __forceinline__ void plane_stage_body_vel(const DevPtrs &d, const InstDesc &I, int o0, int o1, int n,
                                                     float4 (*s_vel)[2], float4 (*s_vel2)[2], int (*s_obj)[PG_N]) {
  const int g0 = o0 + n % PG_R, g1 = o1 + n / PG_R;
  float vq[MB_VELB_W];
  for (int f = 0; f < MB_VELB_W; f++) vq[f] = 0.f;
  int4 cl = make_int4(-1, -1, 0, 0);
  if (g0 < I.n[0] && g1 < I.n[1]) {
    const long long ni = I.node_base + g0 + (long long)I.n[0] * g1;
    if (d.vel[ni * NODE_W + 6] != 0.0f) {
      for (int f = 0; f < MB_VELB_W; f++) vq[f] = d.velb[ni * MB_VELB_W + f];
      cl = d.claim[ni];
    }
  }
  s_vel[n][0] = make_float4(vq[0], vq[1], vq[2], vq[3]);
  s_vel[n][1] = make_float4(vq[4], vq[5], 0.f, 0.f);
  s_vel2[n][0] = make_float4(vq[6], vq[7], vq[8], vq[9]);
  s_vel2[n][1] = make_float4(vq[10], vq[11], 0.f, 0.f);
  s_obj[0][n] = cl.x;
  s_obj[1][n] = cl.y;
}

// Phase B: 9-node scatter of the record in column `col` into the LDS region, or by global
// atomics when its stencil is outside the region.
__device__ __forceinline__ void plane_scatter(const DevPtrs &d, const InstDesc &I, int o0, int o1, int inst,
                                              const PlaneRec *s_rec, int col, double (*s_acc)[PACC_N]) {
  const float4 *w = &s_rec->w[PlaneRec::wi(col)];
  const float4 w0 = w[0], w1 = w[1], w2 = w[2];
  const int q0 = __float_as_int(w0.x), q1 = __float_as_int(w0.y);
  bool bad = false;
  const int b0 = stencil_base(q0, I.lpq, I.fbits, I.n[0] - 3, bad);
  const int b1 = stencil_base(q1, I.lpq, I.fbits, I.n[1] - 3, bad);
  const Axis W0 = gimp3((float)(q0 - (b0 << I.fbits)) * I.inv_2F, I.lp, I.inv_lp, I.inv_h);
  const Axis W1 = gimp3((float)(q1 - (b1 << I.fbits)) * I.inv_2F, I.lp, I.inv_lp, I.inv_h);
  const float m = w0.z, mv0 = w0.w, mv1 = w1.x, mv2 = w1.y;
  const float vs0 = w1.z, vs1 = w1.w, vs3 = w2.x, vs4 = w2.y, vs5 = w2.z;
  const int r0 = b0 - o0 + 1, r1 = b1 - o1 + 1;
  if (!(r0 >= 0 && r1 >= 0 && r0 + 2 < PR && r1 + 2 < PR)) {
    plane_p2g_global(d.acc, d.brick_flag, &d.inst[inst], b0, b1, W0, W1, m, mv0, mv1, mv2, vs0, vs1, vs3, vs4, vs5);
    return;
  }
#pragma unroll
  for (int jj = 0; jj < 3; jj++) {
    const float wy = W1.w[jj], gy = W1.g[jj];
#pragma unroll
    for (int ii = 0; ii < 3; ii++) {
      const int a = (r0 + ii) + PR * (r1 + jj);
      const float wgt = W0.w[ii] * wy, gx = W0.g[ii] * wy, gyy = W0.w[ii] * gy;
      const float t[7] = {wgt * m, wgt * mv0, wgt * mv1, wgt * mv2, vs0 * gx + vs3 * gyy, vs3 * gx + vs1 * gyy,
                          vs5 * gx + vs4 * gyy};
#pragma unroll
      for (int f = 0; f < 7; f++) atomicAdd(&s_acc[f][a], (double)t[f]);
    }
  }
}

// The plane fused particle kernel, persistent over the active tile list (UPDATE = false:
// P2G of the current state only). Per chunk of a tile's entries: phase A with lane =
// consecutive entry, the record of entry j written to column 16*(j mod P) + j/P
// (P = ceil(n/16) passes), so the entries of one cell, consecutive in the list, are
// scattered by different passes of phase B (lane = column); then the region is flushed as
// FP32 partials (core, and the halo shell when reached), node masks and brick flags.
// MB (contact between objects): phase A is the contact G2P on the staged body fields,
// phase B the centre-of-mass P2G of the new state; the body fields follow the re-binning
// and the claims (k_plane_mb_p2g).
template <bool UPDATE, int MASK, bool MB = false>
__global__ __launch_bounds__(WG) __attribute__((amdgpu_waves_per_eu(3))) void k_plane_g2p2g(DevPtrs d) {
  static_assert(!MB || UPDATE, "the contact variant moves particles");
  constexpr int NB2 = MB ? PG_N : 1;
  __shared__ float4 s_vel[PG_N][2];
  __shared__ float4 s_vel2[NB2][2];
  __shared__ int s_obj[2][NB2];
  __shared__ __align__(16) double s_acc[7][PACC_N];
  __shared__ PlaneRec s_rec;
  __shared__ int s_off[TILE3 + 1];
  __shared__ MatDev s_mat[MAXMAT];
  __shared__ int s_tile;
  __shared__ int s_refkey[WG / 64];
  __shared__ float4 s_refv[WG / 64];
  const int tid = threadIdx.x;
  const int par = d.ctrl->parity;
  const int epoch_w = d.ctrl->epoch + (UPDATE ? 1 : 0);
  const TileQueue queue(d);
  // the grid update before this kernel has read the brick list: the bricks this kernel
  // flags are listed from an empty list (by either scan kernel)
  if (blockIdx.x == 0 && tid == 0) d.ctrl->active_bricks = 0;

  for (;;) {
    if (tid == 0) s_tile = queue.tile(d, atomicAdd(&d.ctrl->next, 1));
    __syncthreads();
    const int t = __builtin_amdgcn_readfirstlane(s_tile);
    if (t < 0) break;
    const int k0 = t * TILE3;
    const int ii = d.tile_inst[t];
    const InstDesc I = d.inst[ii];
    const int lt = t - I.tile_base;
    const int o0 = (lt % I.nt[0]) * PT, o1 = (lt / I.nt[0]) * PT;
    float3 vref = make_float3(0.f, 0.f, 0.f);  // staged node velocities are relative to vref

    for (int n = tid; n < 7 * PACC_N; n += WG) (&s_acc[0][0])[n] = 0.0;
    if (tid <= TILE3) s_off[tid] = d.offsets[k0 + tid];
    if (tid < I.nmat && tid < MAXMAT) s_mat[tid] = d.mats[I.mat_base + tid];
    if constexpr (MB) {
      for (int n = tid; n < PG_N; n += WG) plane_stage_body_vel(d, I, o0, o1, n, s_vel, s_vel2, s_obj);
    } else if (UPDATE) {
      // Node velocities staged relative to the tile's reference velocity (the massive node
      // nearest the region centre), nodes without mass as zero: the 9 weights of a plane
      // stencil sum to 1 and their gradients to 0, so the FP32 G2P sums round velocity
      // differences, not full velocities (as in the 3D kernel).
      static_assert(PG_N <= WG, "one staged node per thread");
      const int n = tid;
      float4 v0 = make_float4(0.f, 0.f, 0.f, 0.f), v1 = v0;
      int key = 0x7fffffff;
      if (n < PG_N) {
        const int x = n % PG_R, y = n / PG_R;
        const int g0 = o0 + x, g1 = o1 + y;
        if (g0 < I.n[0] && g1 < I.n[1]) {
          const float4 *vp = reinterpret_cast<const float4 *>(d.vel + (I.node_base + g0 + (long long)I.n[0] * g1) * NODE_W);
          v0 = vp[0];
          v1 = vp[1];
        }
        if (v1.z == 0.0f) {
          v0 = make_float4(0.f, 0.f, 0.f, 0.f);
          v1 = v0;
        } else {
          key = ((abs(2 * x - (PG_R - 1)) + abs(2 * y - (PG_R - 1))) << 8) | n;
        }
      }
      int kmin = key;
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) kmin = min(kmin, __shfl_xor(kmin, off, 64));
      if (key == kmin && key != 0x7fffffff) s_refv[tid >> 6] = v0;
      if ((tid & 63) == 0) s_refkey[tid >> 6] = kmin;
      __syncthreads();
      int best = 0;
#pragma unroll
      for (int w = 1; w < WG / 64; w++) best = (s_refkey[w] < s_refkey[best]) ? w : best;
      const float4 r = (s_refkey[best] != 0x7fffffff) ? s_refv[best] : make_float4(0.f, 0.f, 0.f, 0.f);
      vref = make_float3(r.x, r.y, r.z);
      if (n < PG_N) {
        s_vel[n][0] = make_float4(v0.x - r.x, v0.y - r.y, v0.z - r.z, v0.w);
        s_vel[n][1] = v1;
      }
    }
    __syncthreads();

    const double tl = UPDATE ? d.t_last[ii] : 0.0;
    const int p0 = __builtin_amdgcn_readfirstlane(s_off[0]), p1 = __builtin_amdgcn_readfirstlane(s_off[TILE3]);
    for (int c0 = p0; c0 < p1; c0 += PCHUNK) {
      const int c1 = min(c0 + PCHUNK, p1), nc = c1 - c0;
      const int i = c0 + tid;
      const int P = (nc + 15) >> 4;
      const int j = tid;
      const int col = 16 * (j % P) + j / P;
      int knew = -1;
      if (i < c1)
        knew = plane_update<UPDATE, MASK, MB>(d, I, o0, o1, tl, i, d.tmp[i], s_vel, vref, s_mat, par, &s_rec, col,
                                              MB ? s_vel2 : nullptr,
                                              MB ? reinterpret_cast<const int (*)[PG_N]>(s_obj) : nullptr);
      int hl = 0, hbase = 0;
      if (UPDATE) hbase = hist_issue(d.count, knew, hl);
      __syncthreads();
      // column tid holds entry (tid % 16) * P + tid / 16
      if ((tid >> 4) < P && (tid & 15) * P + (tid >> 4) < nc) plane_scatter(d, I, o0, o1, ii, &s_rec, tid, s_acc);
      if (UPDATE) {
        const int r = hist_rank(hbase, hl);
        if (knew >= 0) d.rank[i] = r;
      }
      __syncthreads();
    }

    // ---- flush ----
    float *pc = d.pcore + (long long)t * CORE_W;
    bool halo = false;
    constexpr int NR = (PR_N + 63) / 64 * 64;  // whole waves per mask word
    for (int n = tid; n < NR; n += WG) {
      const bool live = n < PR_N;
      const int x = n % PR, y = n / PR;
      const bool shell = x == 0 || y == 0 || x > PC || y > PC;
      const bool has = live && s_acc[0][live ? n : 0] != 0.0;
      if (has && !shell) {  // nodes without mass are never read
        const int c = (x - 1) + PC * (y - 1);
        float4 *q = reinterpret_cast<float4 *>(pc + 8 * c);
        q[0] = make_float4((float)s_acc[0][n], (float)s_acc[1][n], (float)s_acc[2][n], (float)s_acc[3][n]);
        q[1] = make_float4((float)s_acc[4][n], (float)s_acc[5][n], (float)s_acc[6][n], 0.0f);
      }
      const unsigned long long bits = __ballot(has);
      if ((tid & 63) == 0) d.nodemask[(long long)t * (P2G_NODES / 64) + (n >> 6)] = bits;
      if (!has) continue;
      halo = halo || shell;
      const int g0 = o0 - 1 + x, g1 = o1 - 1 + y;
      if (g0 < 0 || g1 < 0 || g0 >= I.n[0] || g1 >= I.n[1]) continue;
      int cc0, cc1;
      (void)plane_canon(I, g0, g1, cc0, cc1);
      plane_flag(d.brick_flag, I, cc0, cc1);
    }
    if (__syncthreads_or(halo)) {
      float *pt = d.partial + (long long)t * PARTIAL_W;
      for (int n = tid; n < PR_N; n += WG) {
        const int x = n % PR, y = n / PR;
        if (!(x == 0 || y == 0 || x > PC || y > PC)) continue;
        if (s_acc[0][n] == 0.0) continue;
        float4 *q = reinterpret_cast<float4 *>(pt + 8 * n);
        q[0] = make_float4((float)s_acc[0][n], (float)s_acc[1][n], (float)s_acc[2][n], (float)s_acc[3][n]);
        q[1] = make_float4((float)s_acc[4][n], (float)s_acc[5][n], (float)s_acc[6][n], 0.0f);
      }
      if (tid == 0) d.halo_epoch[t] = epoch_w;
    }
    if (tid == 0) d.tile_epoch[t] = epoch_w;
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------------------
// Contact between objects in a plane (the plane counterparts of the 3D k_mb_* kernels).

// Body claims of every node a particle of an active tile can reach (the tile's 10 x 10
// core): the node's window is the 4 x 4 cells nidx-2..nidx+1 with linear cell ids and only
// a range check (Extrapolation.cpp:89-121), met in ascending (j, i); per cell its objects
// (k_mb_cells) in ascending object id.
__global__ __launch_bounds__(WG) void k_plane_mb_claim(DevPtrs d) {
  const long long nact = d.ctrl->active_tiles;
  for (long long it = (long long)blockIdx.x * WG + threadIdx.x; it < nact * PC_N;
       it += (long long)gridDim.x * WG) {
    const int t = d.active[it / PC_N], ln = (int)(it % PC_N);
    const InstDesc &I = d.inst[d.tile_inst[t]];
    const int lt = t - I.tile_base;
    const int g0 = (lt % I.nt[0]) * PT + ln % PC, g1 = (lt / I.nt[0]) * PT + ln / PC;
    if (g0 >= I.n[0] || g1 >= I.n[1]) continue;
    const int Cx = I.n[0] - 1, Cy = I.n[1] - 1;
    const long long ncell = (long long)Cx * Cy;
    // the 16 cells of the window, all loads issued together (an empty or absent cell
    // holds no object); then the two passes over them in the window's order
    int4 co[16];
#pragma unroll
    for (int q = 0; q < 16; q++) {
      const int ix = (q & 3) - 2, jy = (q >> 2) - 2;
      const long long lin = (g0 + ix) + (long long)(g1 + jy) * Cx;
      int k = -1;
      if (lin >= 0 && lin < ncell) {
        const int c0 = (int)(lin % Cx), c1 = (int)(lin / Cx);
        if (c0 <= I.n[0] - 3 && c1 <= I.n[1] - 3) k = plane_key(I, c0, c1);  // else no particle there
      }
      int4 v = make_int4(-1, -1, 0, 0);
      if (k >= 0) {
        const int lo = d.offsets[k], hi = d.offsets[k + 1];
        const int4 c = d.cellobj[k];
        if (hi != lo) v = c;
      }
      co[q] = v;
    }
    int b1 = -1, b2 = -1, r1 = 0, r2 = 0;
#pragma unroll
    for (int q = 0; q < 16; q++) {
      const int ob[2] = {co[q].x, co[q].y};
      for (int u = 0; u < 2; u++) {
        if (ob[u] < 0) continue;
        if (b1 < 0) b1 = ob[u];
        else if (b2 < 0 && ob[u] != b1) b2 = ob[u];
      }
    }
#pragma unroll
    for (int q = 0; q < 16; q++) {
      r1 |= (co[q].x == b1 && co[q].z) || (co[q].y == b1 && co[q].w);
      r2 |= b2 >= 0 && ((co[q].x == b2 && co[q].z) || (co[q].y == b2 && co[q].w));
    }
    d.claim[I.node_base + g0 + (long long)I.n[0] * g1] = make_int4(b1, b2, r1, r2);
    if (b2 >= 0) d.twotile[t] = 1;  // the tile's particles have a two-body node in reach
  }
}

// Body-field P2G of one particle at the nodes where two bodies are claimed (rigid
// particles: unit mass, their mass also added to the slot's rigid-mass sum at every node),
// folded onto the canonical node; normals per actual node with the particle's volume m/rho
// (Extrapolation.cpp:210-256). FP32 global atomics.
__global__ __launch_bounds__(WG) void k_plane_mb_p2g(DevPtrs d) {
  const int s = blockIdx.x * WG + threadIdx.x;
  const int lane = threadIdx.x & 63;
  const bool live = s < d.np;
  const int sc = live ? s : d.np - 1;  // dead lanes read a valid slot and contribute nothing
  const uint32_t *A = (d.ctrl->parity ? d.pbuf[1] : d.pbuf[0]);
  const long long cap = d.cap;
  const InstDesc &I = d.inst[inst_of_slot(d, sc)];
  const uint32_t tag = A[F_TAG * cap + sc];
  const MatDev &M = d.mats[I.mat_base + tag_mat(tag)];
  // a particle's 9 nodes lie in its tile's claim region: without a two-body node there
  // only rigid particles add (their mass to their slot)
  const int key = (d.ctrl->parity ? d.key[1] : d.key[0])[sc];
  const bool active = live && (M.rigid || d.twotile[key / TILE3]);
  if (__ballot(active) == 0ull) return;  // wave-uniform
  const int obj = tag_obj(tag);
  bool bad = false;
  const int q0 = (int)A[F_QX * cap + sc], q1 = (int)A[F_QY * cap + sc];
  const int b0 = stencil_base(q0, I.lpq, I.fbits, I.n[0] - 3, bad);
  const int b1 = stencil_base(q1, I.lpq, I.fbits, I.n[1] - 3, bad);
  const Axis W0 = gimp3((float)(q0 - (b0 << I.fbits)) * I.inv_2F, I.lp, I.inv_lp, I.inv_h);
  const Axis W1 = gimp3((float)(q1 - (b1 << I.fbits)) * I.inv_2F, I.lp, I.inv_lp, I.inv_h);
  const float mass = f32(A[F_MASS * cap + sc]), rho = f32(A[F_RHO * cap + sc]);
  const float mp = M.rigid ? 1.0f : mass;
  const float nV = -(mp / rho), Vn = mass / rho;
  float mv[3], vs[6];
  for (int a = 0; a < 3; a++) mv[a] = mp * f32(A[(F_VX + a) * cap + sc]);
  for (int a = 0; a < 6; a++) vs[a] = nV * f32(A[(F_S0 + a) * cap + sc]);
  // Adjacent lanes with the same stencil base node, object and rigid flag share all 9
  // target nodes and body slots: their terms are summed over the run (segmented shuffle
  // reduction) and the run's first lane issues the atomics. Runs are numbered by the heads
  // at or below a lane; every shuffle runs with the whole wave active.
  const int base = (int)(I.node_base + b0 + (long long)I.n[0] * b1);
  const int seg_base = active ? base : -1 - lane;
  const int seg_obj = obj * 2 + (M.rigid ? 1 : 0);
  const int up_base = __shfl_up(seg_base, 1, 64), up_obj = __shfl_up(seg_obj, 1, 64);
  const bool head = lane == 0 || up_base != seg_base || up_obj != seg_obj;
  const unsigned long long heads = __ballot(head);
  const int run = __popcll(heads & (lane == 63 ? ~0ull : ((2ull << lane) - 1ull)));
  bool same[6];
#pragma unroll
  for (int r = 0, off = 1; r < 6; r++, off <<= 1) {
    const int down_run = __shfl_down(run, off, 64);
    same[r] = lane + off < 64 && down_run == run;
  }
  // the claimed objects of the 9 stencil nodes, loaded together
  const int2 *claim2 = reinterpret_cast<const int2 *>(d.claim);
  int2 cls[9];
#pragma unroll
  for (int j = 0; j < 3; j++)
#pragma unroll
    for (int i = 0; i < 3; i++) cls[3 * j + i] = claim2[2 * (base + i + (long long)I.n[0] * j)];
#pragma unroll
  for (int j = 0; j < 3; j++)
#pragma unroll
    for (int i = 0; i < 3; i++) {
      const float w = W0.w[i] * W1.w[j], gx = W0.g[i] * W1.w[j], gy = W0.w[i] * W1.g[j];
      const int2 cl = cls[3 * j + i];
      const bool two = cl.y != -1;
      const int sl = (cl.x == obj) ? 0 : (cl.y == obj) ? 1 : -1;
      const bool use = active && sl >= 0 && (two || M.rigid);  // uniform over a run
      if (__ballot(use) == 0ull) continue;  // wave-uniform: nothing to add at this node
      float v[10] = {w * mp, w * mv[0], w * mv[1], w * mv[2], vs[0] * gx + vs[3] * gy, vs[3] * gx + vs[1] * gy,
                     vs[5] * gx + vs[4] * gy, gx * Vn, gy * Vn, w * mp};
#pragma unroll
      for (int f = 0; f < 10; f++) v[f] = use ? v[f] : 0.0f;
#pragma unroll
      for (int r = 0, off = 1; r < 6; r++, off <<= 1)
#pragma unroll
        for (int f = 0; f < 10; f++) {
          const float o = __shfl_down(v[f], off, 64);
          v[f] += same[r] ? o : 0.0f;
        }
      if (!(use && head)) continue;
      int c0, c1;
      const long long ci = plane_canon(I, b0 + i, b1 + j, c0, c1);
      float *ab = d.acc2 + ci * MB_ACC2_W + sl * 8;
      if (M.rigid) unsafeAtomicAdd(&ab[7], v[9]);
      if (!two) continue;
      for (int f = 0; f < 7; f++) unsafeAtomicAdd(&ab[f], v[f]);
      const long long gi = I.node_base + (b0 + i) + (long long)I.n[0] * (b1 + j);
      float *nb = d.nrm + gi * 6 + sl * 3;
      unsafeAtomicAdd(&nb[0], v[7]);
      unsafeAtomicAdd(&nb[1], v[8]);
    }
}

// Contact nodal solve on the flagged plane bricks (one wave per 8 x 8 brick): centre-of-mass
// sums of the periodic class (tile partials and direct scatter), body sums, then per actual
// node its own claims, normals, solve and BCs (mb_solve_node).
__device__ __forceinline__ void plane_mb_grid_brick(const DevPtrs &d, int gb, int lane, int step, int epoch,
                                                    const int box[6]) {
  const InstDesc &I = d.inst[d.brick_inst[gb]];
  const int lb = gb - I.brick_base;
  const int c0 = (lb % I.nb[0]) * PB + (lane & 7), c1 = (lb / I.nb[0]) * PB + (lane >> 3);
  if (c0 < I.cper[0] && c1 < I.cper[1]) {
    const long long ci = I.node_base + c0 + (long long)I.n[0] * c1;
    float a7[7] = {0, 0, 0, 0, 0, 0, 0}, ab[MB_ACC2_W];
    for (int my = 0; my < 2; my++) {
      if (my && c1 >= 3) break;
      const int a1 = c1 + my * I.cper[1];
      if (!in_tile_box<PT>(a1, box[1], box[4])) continue;
      for (int mx = 0; mx < 2; mx++) {
        if (mx && c0 >= 3) break;
        const int a0 = c0 + mx * I.cper[0];
        if (!in_tile_box<PT>(a0, box[0], box[3])) continue;
        plane_gather(d, I, epoch, a0, a1, a7);
      }
    }
    take_node_sums(d, ci, a7, ab);
    const float m = a7[0];
    const float dt = I.dt, kick = (step == 1) ? 0.5f * dt : dt;
    float vc[3] = {0.f, 0.f, 0.f}, ac[3] = {0.f, 0.f, 0.f};
    if (m != 0.0f)
      for (int k = 0; k < 3; k++) {
        vc[k] = a7[1 + k] / m;
        ac[k] = (a7[4 + k] + m * I.gravity[k]) / m;
      }
    for (int my = 0; my < 2; my++) {
      if (my && c1 >= 3) break;
      for (int mx = 0; mx < 2; mx++) {
        if (mx && c0 >= 3) break;
        const long long ni = I.node_base + (c0 + mx * I.cper[0]) + (long long)I.n[0] * (c1 + my * I.cper[1]);
        mb_solve_node(d, I, ni, m, a7, ab, vc, ac, dt, kick);
      }
    }
  }
  if (lane == 0) d.brick_flag[gb] = 0;
}

// Plane contact grid update, persistent over the list of flagged bricks (one wave per brick).
__global__ __launch_bounds__(WG) void k_plane_mb_grid(DevPtrs d) {
  // the particle kernel that follows walks the active tile list from the start
  if (blockIdx.x == 0 && threadIdx.x == 0) d.ctrl->next = 0;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int n = d.ctrl->active_bricks, step = d.ctrl->step, epoch = d.ctrl->epoch;
  int box[6];
  tile_box(d, epoch, box);
  for (int k = blockIdx.x * (WG / 64) + wave; k < n; k += gridDim.x * (WG / 64))
    plane_mb_grid_brick(d, d.brick_list[k], lane, step, epoch, box);
}

inline hipStream_t S(void *s) { return reinterpret_cast<hipStream_t>(s); }
inline int blocks(long long n, int per) { return (int)((n + per - 1) / per); }

}  // namespace

void launch_plane_keys(const DevPtrs &d, void *st) {
  if (d.np > 0) hipLaunchKernelGGL(k_plane_keys, dim3(blocks(d.np, WG)), dim3(WG), 0, S(st), d);
}
void launch_plane_p2g_only(const DevPtrs &d, void *st) {
  hipLaunchKernelGGL((k_plane_g2p2g<false, KM_ALL>), dim3(d.persist_wgs), dim3(WG), 0, S(st), d);
}
// Smallest built model-set specialization covering the context's mask.
void launch_plane_g2p2g(const DevPtrs &d, void *st) {
  const int m = d.kmask;
  const dim3 g(d.persist_wgs), b(WG);
  if ((m & ~KM_ELASTIC) == 0) hipLaunchKernelGGL((k_plane_g2p2g<true, KM_ELASTIC>), g, b, 0, S(st), d);
  else if ((m & ~KM_NEWTON) == 0) hipLaunchKernelGGL((k_plane_g2p2g<true, KM_NEWTON>), g, b, 0, S(st), d);
  else if ((m & ~KM_DP) == 0) hipLaunchKernelGGL((k_plane_g2p2g<true, KM_DP>), g, b, 0, S(st), d);
  else if ((m & ~(KM_ELASTIC | KM_GK)) == 0) hipLaunchKernelGGL((k_plane_g2p2g<true, KM_ELASTIC | KM_GK>), g, b, 0, S(st), d);
  else hipLaunchKernelGGL((k_plane_g2p2g<true, KM_ALL>), g, b, 0, S(st), d);
}
void launch_plane_mb_claim(const DevPtrs &d, void *st) {
  const int g = std::max(8, std::min(4096, blocks((long long)d.ntiles * PC_N, WG)));
  hipLaunchKernelGGL(k_plane_mb_claim, dim3(g), dim3(WG), 0, S(st), d);
}
void launch_plane_mb_p2g(const DevPtrs &d, void *st) {
  if (d.np > 0) hipLaunchKernelGGL(k_plane_mb_p2g, dim3(blocks(d.np, WG)), dim3(WG), 0, S(st), d);
}
void launch_plane_mb_grid(const DevPtrs &d, void *st) {
  const int nb = std::max(1, std::min(blocks(d.nbricks, WG / 64), std::max(1, d.grid_factor) * d.persist_wgs));
  hipLaunchKernelGGL(k_plane_mb_grid, dim3(nb), dim3(WG), 0, S(st), d);
}
void launch_plane_mb_g2p2g(const DevPtrs &d, void *st) {
  const int m = d.kmask;
  const dim3 g(d.persist_wgs), b(WG);
  // rigid platens count as elastic: a damaging Drucker-Prager body between them is 1 + 8 + 16
  if ((m & ~(KM_ELASTIC | KM_DP | KM_GK)) == 0)
    hipLaunchKernelGGL((k_plane_g2p2g<true, KM_ELASTIC | KM_DP | KM_GK, true>), g, b, 0, S(st), d);
  else if ((m & ~(KM_ELASTIC | KM_GK)) == 0)
    hipLaunchKernelGGL((k_plane_g2p2g<true, KM_ELASTIC | KM_GK, true>), g, b, 0, S(st), d);
  else hipLaunchKernelGGL((k_plane_g2p2g<true, KM_ALL, true>), g, b, 0, S(st), d);
}
static int plane_grid_blocks(const DevPtrs &d) {
  return std::max(1, std::min(blocks(d.nbricks, WG / 64), std::max(1, d.grid_factor) * d.persist_wgs));
}
void launch_plane_grid(const DevPtrs &d, void *st) {
  hipLaunchKernelGGL(k_plane_grid, dim3(plane_grid_blocks(d)), dim3(WG), 0, S(st), d);
}
void launch_plane_scan_final_grid(const DevPtrs &d, void *st) {
  const int ng = plane_grid_blocks(d);
  hipLaunchKernelGGL(k_plane_scan_final_grid, dim3(ng + d.nblocks_scan), dim3(WG), 0, S(st), d, ng);
}

}  // namespace gy
