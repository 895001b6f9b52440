// gy_kernels.hip — HIP kernels (gfx950) of the MPM time step.
//
// One step is five launches on one stream:
//   k_grid      nodal solve, persistent over the list of flagged canonical bricks: gather
//               the P2G tile partials of every node of the periodic class, v = p/m,
//               a = f/m + g, body-1 kick (half at step 1), BCs, fan-out to every mirror
//               node (Grid.cpp:9-245, 310-603, 862-1094)
//   k_g2p2g     persistent over the active cell tiles (one work queue, tiles with the most
//               particles first): G2P from LDS-staged node
//               velocities, particle update, constitutive/damage models, periodic wrap,
//               new cell key + histogram, and P2G of the updated particle into an FP64 LDS
//               region with a one-node halo, written out as per-tile FP32 partial sums
//               with plain stores (Extrapolation.cpp:31-367, TimeIntegration.cpp:11-296,
//               ConstitutiveModels.cpp, DamageModels.cpp)
//   k_scan_*    exclusive scan of the per-cell histogram (counting sort offsets), the
//               active tile list (longest first, by entry-count class) and
//               the flagged-brick list; the first also does the step/parity/time
//               bookkeeping
//   k_place     cell-grouped slot list (within a cell in histogram order; the stable
//               order, ascending slot, is computed only for gy_debug_binning and
//               gy_counting_sort)
// Binning replaces sortParticlesByCellID_CPU (Extrapolation.cpp:370-503) and runs every step.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "gy_device.hpp"
#include "gy_internal.hpp"
#include "gy_models.hpp"
#include "gy_scan.hpp"

namespace gy {

namespace {

constexpr int NODES_G2P = G2P_R * G2P_R * G2P_R;
constexpr int NODES_P2G = P2G_NODES;


// ---------------------------------------------------------------------------------------
__global__ void k_set_ctrl(DevPtrs d, int step, double t, int reset) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < d.ninst) d.t_cur[i] = t;
  if (i == 0) {
    d.ctrl->step = step;
    if (reset) {
      d.ctrl->parity = 0; d.ctrl->status = 0; d.ctrl->oor = 0;
      d.ctrl->nn_count[0] = 0; d.ctrl->nn_count[1] = 0;
    }
    if (d.ctrl->epoch == 0) d.ctrl->epoch = 1;  // tile_epoch starts at 0: never current
  }
}


// Keys + histogram of the current buffer (first step after an upload).
__global__ __launch_bounds__(WG) void k_keys(DevPtrs d) {
  const int s = blockIdx.x * WG + threadIdx.x;
  if (s >= d.np) return;
  const int par = d.ctrl->parity;
  const uint32_t *A = (par ? d.pbuf[1] : d.pbuf[0]);
  const InstDesc &I = d.inst[inst_of_slot(d, s)];
  bool bad = false;
  int b[3];
  for (int a = 0; a < 3; a++)
    b[a] = stencil_base((int)A[(F_QX + a) * d.cap + s], I.lpq, I.fbits, I.n[a] - 3, bad);
  const int k = key_of(I, b);
  (par ? d.key[1] : d.key[0])[s] = k;
  d.rank[s] = atomicAdd(&d.count[k], 1);
  if (bad) { atomicAdd(&d.ctrl->oor, 1); atomicOr(&d.ctrl->status, 1); }
}

// Histogram of caller-supplied keys (the standalone counting sorts): keys outside
// [0, nkeys) are skipped here, in k_place and in k_stable_order.
__global__ __launch_bounds__(WG) void k_hist(DevPtrs d) {
  const int s = blockIdx.x * WG + threadIdx.x;
  if (s >= d.np) return;
  const int k = d.key[0][s];
  if (k < 0 || k >= d.nkeys) return;
  d.rank[s] = atomicAdd(&d.count[k], 1);
}

// With `advance`, block 0 also performs the step bookkeeping: times, step, parity and
// epoch (no block of this kernel reads them; the kernels after it see the new values).
// bricks: this kernel also lists the flagged bricks (the grid update then runs beside the
// second scan kernel); the list's length was reset by the particle kernel before it.
__global__ __launch_bounds__(WG) void k_scan_partial(DevPtrs d, int advance, int bricks) {
  if (advance && blockIdx.x == 0) {
    for (int i = threadIdx.x; i < d.ninst; i += WG) {
      d.t_last[i] = d.t_cur[i];
      d.t_cur[i] += (double)d.inst[i].dt;
    }
    if (threadIdx.x == 0) {
      d.ctrl->step += 1;
      d.ctrl->parity ^= 1;
      d.ctrl->epoch += 1;
    }
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    // the tile box entry of the particle pass this binning feeds: it stamps epoch + 1
    // after a step's binning, the epoch itself after an upload's
    d.ctrl->tbox_w = (d.ctrl->epoch + advance) & 1;
  }
  // the brick list is rebuilt by the scan's second kernel (or below)
  if (!bricks && blockIdx.x == 0 && threadIdx.x == 0) d.ctrl->active_bricks = 0;
  const long long base = (long long)blockIdx.x * SCAN_BLOCK + threadIdx.x * SCAN_ITEMS;
  int s = 0;
  if (base < d.nkeys) {
    int c[SCAN_ITEMS];
    load16(d.count + base, c);
#pragma unroll
    for (int j = 0; j < SCAN_ITEMS; j++) s += c[j];
  }
  int ts = s + __shfl_xor(s, 1, 64);
  ts += __shfl_xor(ts, 2, 64);
  const int bk = ((threadIdx.x & 3) == 0) ? tile_bucket(ts) : -1;
  int v5[5] = {s, bk == 0, bk == 1, bk == 2, bk == 3}, t5[5];
  block_scan_n(v5, t5);
  if (threadIdx.x == 0) {
    d.blocksum[blockIdx.x] = t5[0];
    // tiles per bucket, at most SCAN_BLOCK / TILE3 = 64 each: one byte each
    d.blocktiles[blockIdx.x] = t5[1] | (t5[2] << 8) | (t5[3] << 16) | (t5[4] << 24);
  }
  if (d.ninst == 1 && d.blockbox) {  // uniform
    // the box of this block's occupied tiles (tile coordinates: lo, TBOX_BIG - hi)
    const InstDesc &I = d.inst[0];
    const int tl = (int)(base / TILE3);
    const int tc[3] = {tl % I.nt[0], (tl / I.nt[0]) % I.nt[1], tl / (I.nt[0] * I.nt[1])};
    int v[6];
#pragma unroll
    for (int a = 0; a < 3; a++) {
      v[a] = bk >= 0 ? tc[a] : TBOX_BIG;
      v[3 + a] = bk >= 0 ? TBOX_BIG - tc[a] : TBOX_BIG;
    }
    block_min6(v);
    if (threadIdx.x == 0)
#pragma unroll
      for (int a = 0; a < 6; a++) d.blockbox[6 * blockIdx.x + a] = v[a];
  }
  if (bricks) append_bricks(d, blockIdx.x * WG + threadIdx.x, gridDim.x * WG);
}

__global__ __launch_bounds__(WG) void k_scan_final(DevPtrs d) { scan_final_body(d, blockIdx.x, gridDim.x, true); }

// Brick list outside the step sequence (after the first P2G of an upload).
__global__ void k_brick_list_reset(DevPtrs d) { d.ctrl->active_bricks = 0; }
__global__ __launch_bounds__(WG) void k_brick_list(DevPtrs d) {
  append_bricks(d, blockIdx.x * WG + threadIdx.x, gridDim.x * WG);
}

__global__ __launch_bounds__(WG) void k_place(DevPtrs d) {
  const int s = blockIdx.x * WG + threadIdx.x;
  if (s >= d.np) return;
  const int k = (d.ctrl->parity ? d.key[1] : d.key[0])[s];
  if (k < 0 || k >= d.nkeys) return;
  d.tmp[d.offsets[k] + d.rank[s]] = s;
}

// Stable position of the slot listed at sorted index i: the key-grouped list holds each
// cell's slots in arbitrary order; ascending slot order inside the cell is the stable
// counting-sort order.
__device__ __forceinline__ int stable_pos(const DevPtrs &d, const int *keys, int i, int slot) {
  const int k = keys[slot];
  const int lo = d.offsets[k], hi = d.offsets[k + 1];
  int r = 0;
  for (int j = lo; j < hi; j++) r += (d.tmp[j] < slot) ? 1 : 0;
  return lo + r;
}

__global__ __launch_bounds__(WG) void k_stable_order(DevPtrs d, int *order) {
  const int i = blockIdx.x * WG + threadIdx.x;
  if (i >= d.np) return;
  if (i >= d.offsets[d.nkeys]) return;  // entries of skipped (out-of-range) keys
  const int slot = d.tmp[i];
  order[stable_pos(d, (d.ctrl->parity ? d.key[1] : d.key[0]), i, slot)] = slot;
}

// ---------------------------------------------------------------------------------------
// Sum of the tile partials covering actual node (a0, a1, a2): a node n lies in the P2G
// region [o-1, o+TILE+3) of the tiles with o = TILE*(tt-1) and TILE*tt, tt = (n+1)/TILE,
// at region coordinate l = n - o + 1. Coordinates 1..TILE+2 on every axis are the tile's
// core partials, the rest its halo shell (written only when a particle that changed cell
// reached it). A tile contributes when it wrote partials this step (tile_epoch == epoch)
// and its node mask has the node's bit (the node has mass in that tile's region), so
// zero partials are never read. Fixed order (z, y, x offset): deterministic.
__device__ __forceinline__ void gather_partials(const DevPtrs &d, const InstDesc &I, int epoch, int a0, int a1,
                                                int a2, float acc7[7]) {
  const int t0 = (a0 + 1) / TILE, t1 = (a1 + 1) / TILE, t2 = (a2 + 1) / TILE;
  bool use[8];
  const float4 *pp[8];
#pragma unroll
  for (int k = 0; k < 8; k++) {
    const int tx = t0 - 1 + (k & 1), ty = t1 - 1 + ((k >> 1) & 1), tz = t2 - 1 + (k >> 2);
    const bool in = tx >= 0 && tx < I.nt[0] && ty >= 0 && ty < I.nt[1] && tz >= 0 && tz < I.nt[2];
    const int lx = a0 - TILE * tx + 1, ly = a1 - TILE * ty + 1, lz = a2 - TILE * tz + 1;  // in [0, P2G_R)
    const int t = in ? I.tile_base + tx + I.nt[0] * (ty + I.nt[1] * tz) : I.tile_base;
    const int l = lx + P2G_R * (ly + P2G_R * lz);
    const uint64_t word = d.nodemask[(long long)t * (P2G_NODES / 64) + (l >> 6)];
    use[k] = in && d.tile_epoch[t] == epoch && ((word >> (l & 63)) & 1ull);
    const bool core = lx >= 1 && lx <= CORE_R && ly >= 1 && ly <= CORE_R && lz >= 1 && lz <= CORE_R;
    const int c = (lx - 1) + CORE_R * ((ly - 1) + CORE_R * (lz - 1));
    pp[k] = reinterpret_cast<const float4 *>(core ? d.pcore + (long long)t * CORE_W + 8 * c
                                                  : d.partial + (long long)t * PARTIAL_W + 8 * l);
  }
  // a node's partial is (m, mv0, mv1, mv2 | f0, f1, f2, -): two 16-byte loads
  float4 u[8], w[8];
#pragma unroll
  for (int k = 0; k < 8; k++) {
    const float4 z = make_float4(0.f, 0.f, 0.f, 0.f);
    u[k] = use[k] ? pp[k][0] : z;
    w[k] = use[k] ? pp[k][1] : z;
  }
#pragma unroll
  for (int k = 0; k < 8; k++) {
    const float v[7] = {u[k].x, u[k].y, u[k].z, u[k].w, w[k].x, w[k].y, w[k].z};
#pragma unroll
    for (int f = 0; f < 7; f++) acc7[f] = use[k] ? acc7[f] + v[f] : acc7[f];
  }
}

// Grid update of one flagged canonical brick (one wave per 4x4x4 brick): gather the P2G
// partials of every actual node of the periodic class inside the tile box (tile_box),
// then solve.
__device__ __forceinline__ void grid_brick(const DevPtrs &d, int gb, int lane, int step, int epoch,
                                           const int box[6]) {
  // one instance: no brick -> instance lookup in the load chain
  const InstDesc &I = d.inst[d.ninst == 1 ? 0 : d.brick_inst[gb]];
  const int lb = gb - I.brick_base;
  const int bx = lb % I.nb[0], by = (lb / I.nb[0]) % I.nb[1], bz = lb / (I.nb[0] * I.nb[1]);
  const int c0 = bx * BRICK + (lane & 3), c1 = by * BRICK + ((lane >> 2) & 3), c2 = bz * BRICK + (lane >> 4);
  if (c0 < I.cper[0] && c1 < I.cper[1] && c2 < I.cper[2]) {
    // the canonical node's direct-scatter sums and BC, loaded before the gather's chain
    const long long ci = I.node_base + c0 + (long long)I.n[0] * (c1 + (long long)I.n[1] * c2);
    float4 *ap = reinterpret_cast<float4 *>(d.acc + ci * NODE_W);
    const float4 x0 = ap[0], x1 = ap[1];
    const int bt0 = d.bc_type[ci];
    const float bv0 = d.bc_val[2 * ci], bv1 = d.bc_val[2 * ci + 1];
    // P2G sums over the periodic class (Grid.cpp:320-602: node i and node N-3+i share sums)
    float a7[7] = {0, 0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int mz = 0; mz < 2; mz++) {
      if (mz && c2 >= 3) break;
      const int a2 = c2 + mz * I.cper[2];
      if (!in_tile_box<TILE>(a2, box[2], box[5])) continue;
#pragma unroll
      for (int my = 0; my < 2; my++) {
        if (my && c1 >= 3) break;
        const int a1 = c1 + my * I.cper[1];
        if (!in_tile_box<TILE>(a1, box[1], box[4])) continue;
#pragma unroll
        for (int mx = 0; mx < 2; mx++) {
          if (mx && c0 >= 3) break;
          const int a0 = c0 + mx * I.cper[0];
          if (!in_tile_box<TILE>(a0, box[0], box[3])) continue;
          gather_partials(d, I, epoch, a0, a1, a2, a7);
        }
      }
    }
    // particles scattered straight to the grid (outside their tile's region)
    if (x0.x != 0.f || x0.y != 0.f || x0.z != 0.f || x0.w != 0.f || x1.x != 0.f || x1.y != 0.f || x1.z != 0.f) {
      a7[0] += x0.x; a7[1] += x0.y; a7[2] += x0.z; a7[3] += x0.w;
      a7[4] += x1.x; a7[5] += x1.y; a7[6] += x1.z;
      ap[0] = make_float4(0.f, 0.f, 0.f, 0.f);
      ap[1] = make_float4(0.f, 0.f, 0.f, 0.f);
    }
    const float m = a7[0];
    float v[3] = {0.f, 0.f, 0.f}, acc[3] = {0.f, 0.f, 0.f};
    if (m != 0.0f) {
      // Grid.cpp:14-16, 220-243: v = p/m; a = f/m (gravity enters as m*g, the sum of the
      // per-pair N*m_p*g terms of Extrapolation.cpp:226-228); v += a*dt (0.5*dt at step 1)
      const float kick = (step == 1) ? 0.5f * I.dt : I.dt;
#pragma unroll
      for (int k = 0; k < 3; k++) {
        acc[k] = (a7[4 + k] + m * I.gravity[k]) / m;
        v[k] = a7[1 + k] / m + acc[k] * kick;
      }
    }
    // fan out to the actual nodes of the periodic class with each node's own BC
    // (Grid.cpp:862-1031)
#pragma unroll
    for (int mz = 0; mz < 2; mz++) {
      if (mz && c2 >= 3) break;
      const int n2 = c2 + mz * I.cper[2];
#pragma unroll
      for (int my = 0; my < 2; my++) {
        if (my && c1 >= 3) break;
        const int n1 = c1 + my * I.cper[1];
#pragma unroll
        for (int mx = 0; mx < 2; mx++) {
          if (mx && c0 >= 3) break;
          const int n0 = c0 + mx * I.cper[0];
          const long long ni = I.node_base + n0 + (long long)I.n[0] * (n1 + (long long)I.n[1] * n2);
          float vv[3] = {v[0], v[1], v[2]}, aa[3] = {acc[0], acc[1], acc[2]};
          const bool canon = (ni == ci);
          const int t = canon ? bt0 : d.bc_type[ni];
          if (t == 1 || t == 2 || t == 11) {
            const float b0 = canon ? bv0 : d.bc_val[2 * ni], b1 = canon ? bv1 : d.bc_val[2 * ni + 1];
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
  }
  if (lane == 0) d.brick_flag[gb] = 0;
}

// Grid update, persistent over the list of flagged bricks.
// The grid update as block bid of nblk: one wave per listed brick.
__device__ __forceinline__ void grid_body(const DevPtrs &d, int bid, int nblk) {
  // the particle kernel that follows walks the active tile list from the start
  if (bid == 0 && threadIdx.x == 0) d.ctrl->next = 0;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int n = d.ctrl->active_bricks, step = d.ctrl->step, epoch = d.ctrl->epoch;
  int box[6];
  tile_box(d, epoch, box);
  for (int k = bid * (WG / 64) + wave; k < n; k += nblk * (WG / 64))
    grid_brick(d, d.brick_list[k], lane, step, epoch, box);
}

__global__ __launch_bounds__(WG) void k_grid(DevPtrs d) { grid_body(d, blockIdx.x, gridDim.x); }

// The second scan kernel of step s and the grid update of step s + 1 in one launch: blocks
// [0, ngrid) update the grid, the others scan. The two share no data: the grid update
// reads the tile partials of step s, the brick list (listed by the first scan kernel) and
// the tile box of the pass that wrote those partials; the scan writes the offsets, the
// active tile list and the other tile box entry. Both are chains of dependent round trips
// with little work, so the shorter one hides beside the other.
__global__ __launch_bounds__(WG) void k_scan_final_grid(DevPtrs d, int ngrid) {
  if ((int)blockIdx.x < ngrid) grid_body(d, blockIdx.x, ngrid);
  else scan_final_body(d, blockIdx.x - ngrid, gridDim.x - ngrid, false);
}


__device__ __forceinline__ void flag_brick(const DevPtrs &d, const InstDesc &I, const int c[3]) {
  const int gb = I.brick_base + c[0] / BRICK + I.nb[0] * (c[1] / BRICK + I.nb[1] * (c[2] / BRICK));
  d.brick_flag[gb] = 1;
}

__device__ __forceinline__ long long canon_index(const InstDesc &I, const int g[3], int c[3]) {
#pragma unroll
  for (int a = 0; a < 3; a++) c[a] = (g[a] >= I.cper[a]) ? g[a] - I.cper[a] : g[a];
  return I.node_base + c[0] + (long long)I.n[0] * (c[1] + (long long)I.n[1] * c[2]);
}

// P2G of one particle straight into the global accumulators (used when its stencil lies
// outside the tile's LDS region, e.g. after a periodic wrap). Kept out of line and fully
// unrolled so the hot path never materialises the weight arrays in scratch.
__device__ __noinline__ void p2g_global(float *acc, uint8_t *brick_flag, const InstDesc *Ip, int b0, int b1, int b2,
                                        Axis W0, Axis W1, Axis W2, P2GTerms T) {
  const InstDesc &I = *Ip;
  const int b[3] = {b0, b1, b2};
#pragma unroll
  for (int k = 0; k < 3; k++)
#pragma unroll
    for (int j = 0; j < 3; j++)
#pragma unroll
      for (int i = 0; i < 3; i++) {
        const float A = W1.w[j] * W2.w[k];
        const float w = W0.w[i] * A;
        if (w == 0.0f) continue;
        float t[7];
        node_terms(T, w, W0.g[i] * A, W0.w[i] * W1.g[j] * W2.w[k], W0.w[i] * W1.w[j] * W2.g[k], t);
        const int g[3] = {b[0] + i, b[1] + j, b[2] + k};
        int c[3];
        const long long ci = canon_index(I, g, c);
#pragma unroll
        for (int f = 0; f < 7; f++) unsafeAtomicAdd(&acc[ci * NODE_W + f], t[f]);
        brick_flag[I.brick_base + c[0] / BRICK + I.nb[0] * (c[1] / BRICK + I.nb[1] * (c[2] / BRICK))] = 1;
      }
}


// The fused kernel works through a tile's entries in chunks of CHUNK list entries.
constexpr int CHUNK = CHUNK_ENTRIES;
// FP64 LDS accumulator region: P2G_R^3 nodes, x fastest, unpadded, with the x index of
// rows y = 2, 3 (mod 4) XOR-ed with 4. ds_add_f64 serves 16 lanes at a time over 32
// four-byte banks, so a double's bank pair is its index mod 16: (x ^ 4*bit1(y)) + 8*(y
// mod 2) for this layout. Over any four consecutive x and four consecutive y -- the 16
// cells of one z-layer of a tile, at any stencil offset -- these are 16 distinct pairs,
// so any 16 consecutive cells of a tile (the lanes of one pass in rank-major order) add
// without bank conflicts. Without the swizzle rows y and y+2 share banks (two-way
// conflicts on every pass, measured).
constexpr int ACC_YS = P2G_R;
constexpr int ACC_ZS = P2G_R * P2G_R;
constexpr int ACC_N = ACC_ZS * P2G_R;
__device__ __forceinline__ int acc_index(int x, int y, int z) {
  return (x ^ (((y >> 1) & 1) << 2)) + ACC_YS * y + ACC_ZS * z;
}
// P2G record of one particle, handed from the update phase to the scatter phase in LDS:
// fixed-point position (3), m, m*v (3), -V*sigma (6).
constexpr int NREC = 13;
// Records in LDS: three 16-byte words (q0 q1 q2 m | mv0 mv1 mv2 vs0 | vs1 vs2 vs3 vs4) and
// one more float (vs5) per record, with one word (one float) of padding after every 16
// columns. Phase A writes the particles of one cell to columns 16 apart: unpadded, those
// are 768 bytes apart, a multiple of the 256-byte LDS bank period (a 16-way conflict for
// a pass of one cell's lanes); padded, 784 bytes, which spreads them over all banks.
struct RecLds {
  float4 w[3 * CHUNK + CHUNK / 16];
  float t[CHUNK + CHUNK / 16];
  __device__ __forceinline__ static int wi(int col) { return 3 * col + (col >> 4); }
  __device__ __forceinline__ static int ti(int col) { return col + (col >> 4); }
};



// A yielding Drucker-Prager sample of a learned material, held between phase A and the
// workgroup's network evaluation (k_g2p2g with KM_DPNN).
struct DpPend {
  int state;        // 0 none, 1 waiting for the network, 2 outside the trained range (closed form taken)
  int mat;          // material index
  int slot;         // source slot (the stress before the step is read there again when the return keeps it)
  uint32_t x[8];    // normalised network inputs, BF16 pairs
  DpTrial T;        // elastic trial state
  float p[6], epb;  // plastic strain and accumulated plastic strain before the step
  float nV;         // -m/rho of the record
};

// Learned return map, phase A part: the trial state exactly; an elastic sample takes the
// trial stress; a yielding one inside the trained input range waits for the network
// (s becomes the trial stress until then); outside it, the closed form.
__device__ __forceinline__ void dp_learned_gate(const float *norm, const MatDev &M, int mat, int slot,
                                                const float er[6], float s[6], float p[6], float &epb, float dt,
                                                DpPend &pd) {
  DpTrial T;
  dp_trial(M, er, s, dt, T);
  if (!T.yield) {
#pragma unroll
    for (int k = 0; k < 6; k++) s[k] = T.tr[k];
    return;
  }
  float xr[GY_MLP_IN];
#pragma unroll
  for (int k = 0; k < 6; k++) {
    xr[k] = er[k] * dt;
    xr[6 + k] = s[k];
  }
  xr[12] = epb;
  xr[13] = M.phi_deg;
  if (nn_inputs(norm, xr, pd.x)) {
    pd.state = 1;
    pd.mat = mat;
    pd.slot = slot;
    pd.T = T;
#pragma unroll
    for (int k = 0; k < 6; k++) {
      s[k] = T.tr[k];
      pd.p[k] = p[k];
    }
    pd.epb = epb;
    return;
  }
  pd.state = 2;
  const float dL = dp_cone_multiplier<false>(M, T, epb);
  dp_return<false>(M, T, dL, s, p, epb);
}

// Tile of claim k over the active list (longest first), -1 past the end; scalar loads.
__device__ __forceinline__ int claimed_tile(const DevPtrs &d, int nh, int n, int k) {
  if (k < nh) return cld(&d.active[k]);
  if (k < n) return cld(&d.active[n - 1 - (k - nh)]);
  return -1;
}

// Instance and grid placement of tile t (scalar loads).
struct TileGeo {
  int ii;               // instance
  int o[3];             // first cell (= first node of the G2P region) per axis
  int n[3];             // nodes per axis
  long long node_base;  // first global node of the instance
};
__device__ __forceinline__ TileGeo tile_geo(const DevPtrs &d, int t) {
  TileGeo g;
  g.ii = d.ninst == 1 ? 0 : cld(&d.tile_inst[t]);
  const InstDesc *Ip = &d.inst[g.ii];
  const int lt = t - cld(&Ip->tile_base), nt0 = cld(&Ip->nt[0]), nt1 = cld(&Ip->nt[1]);
  g.o[0] = (lt % nt0) * TILE;
  g.o[1] = ((lt / nt0) % nt1) * TILE;
  g.o[2] = (lt / (nt0 * nt1)) * TILE;
#pragma unroll
  for (int a = 0; a < 3; a++) g.n[a] = cld(&Ip->n[a]);
  g.node_base = cld(&Ip->node_base);
  return g;
}

// The vector loads a tile starts with: thread tid's entry of its 65 offsets and (VEL) the
// velocity record of node tid of its G2P region, every lane loading a valid address (an
// out-of-grid node reads the instance's first node and gets keep = 0). Nothing here uses
// the values, so nothing waits for them here.
template <bool VEL>
__device__ __forceinline__ void tile_loads(const DevPtrs &d, int t, const TileGeo &g, int tid, int &off,
                                           float4 &v0, float4 &v1, uint32_t &keep) {
  off = d.offsets[t * TILE3 + min(tid, TILE3)];
  if constexpr (VEL) {
    const int n = min(tid, NODES_G2P - 1);
    const int x = n % G2P_R, y = (n / G2P_R) % G2P_R, z = n / (G2P_R * G2P_R);
    const int g0 = g.o[0] + x, g1 = g.o[1] + y, g2 = g.o[2] + z;
    const bool in = tid < NODES_G2P && g0 < g.n[0] && g1 < g.n[1] && g2 < g.n[2];
    const long long ni = g.node_base + (in ? g0 + (long long)g.n[0] * (g1 + (long long)g.n[1] * g2) : 0);
    const float4 *vp = reinterpret_cast<const float4 *>(d.vel + ni * NODE_W);
    v0 = vp[0];
    v1 = vp[1];
    keep = in ? 0xffffffffu : 0u;
  }
}

// ... into the places the tile's first chunk reads: s_off (tid <= TILE3) and (VEL) the raw
// node records, made relative to the tile's reference velocity when the tile starts.
// Out-of-grid records are zeroed by masking their bits (a select of the values lets the
// compiler select the address instead: a generic load of a stack zero).
template <bool VEL>
__device__ __forceinline__ void stage_ahead(int tid, int off, float4 a, float4 b, uint32_t m, int *s_off,
                                            float4 (*s_vel)[2]) {
  if (tid <= TILE3) s_off[tid] = off;
  if constexpr (VEL) {
    if (tid < NODES_G2P) {
      s_vel[tid][0] = make_float4(f32(u32(a.x) & m), f32(u32(a.y) & m), f32(u32(a.z) & m), f32(u32(a.w) & m));
      s_vel[tid][1] = make_float4(f32(u32(b.x) & m), f32(u32(b.y) & m), f32(u32(b.z) & m), f32(u32(b.w) & m));
    }
  }
}

// `publish` runs in an update after its G2P and before its first store (the fused kernel
// hands the next tile's claim over there).
struct NoPublish {
  __device__ void operator()() const {}
};

// Phase A for one list entry i (slot `slot` of the source buffer): G2P + update +
// constitutive + store at sorted position i (UPDATE), and the P2G record of the (new)
// state into column j of s_rec. Returns the new cell key (UPDATE) or -1.
// Node velocity and acceleration of the staged G2P region: s_vel[n] = {(vx, vy, vz, ax),
// (ay, az, -, -)}, two 16-byte LDS reads per node; the velocity relative to the tile's
// reference velocity vref (zero in the contact variant).
// MB (two-body contact): G2P reads the velocity field of the body slot the particle's
// object holds at each node (s_vel: body 1, s_vel2: body 2, s_obj: the slots' objects;
// Extrapolation.cpp:288-367). The record is written when s_rec is given.
template <bool UPDATE, int MASK, bool MB = false, class Publish = NoPublish>
__device__ __forceinline__ int particle_update(const DevPtrs &d, const InstDesc &I, const int o[3], double tl,
                                               int i, int slot, const float4 (*s_vel)[2], float3 vref,
                                               const MatDev *s_mat,
                                               int par, RecLds *s_rec, int j,
                                               const float4 (*s_vel2)[2] = nullptr,
                                               int (*s_obj)[NODES_G2P] = nullptr, DpPend *pd = nullptr,
                                               const float *s_nnn = nullptr,
                                               const Publish &publish = Publish()) {
  constexpr bool DPNN = (MASK & KM_DPNN) != 0, DPLOOP = (MASK & KM_DPLOOP) != 0;
  const uint32_t *A = (par ? d.pbuf[1] : d.pbuf[0]);
  uint32_t *B = (par ? d.pbuf[0] : d.pbuf[1]);
  const long long cap = d.cap;
  int q[3];
#pragma unroll
  for (int a = 0; a < 3; a++) q[a] = (int)ld_stream(&A[(F_QX + a) * cap + slot]);
  const uint32_t tag = ld_stream(&A[F_TAG * cap + slot]);
  const MatDev &M = s_mat[tag_mat(tag)];
  float v[3], s[6];
#pragma unroll
  for (int a = 0; a < 3; a++) v[a] = f32(ld_stream(&A[(F_VX + a) * cap + slot]));
#pragma unroll
  for (int a = 0; a < 6; a++) s[a] = f32(ld_stream(&A[(F_S0 + a) * cap + slot]));
  float rho = f32(ld_stream(&A[F_RHO * cap + slot]));
  const float mass = f32(ld_stream(&A[F_MASS * cap + slot]));
  int knew = -1;

  if (UPDATE) {
    const int out = i;
    uint32_t xv[3];  // position fraction | velocity residual (gy_internal.hpp)
#pragma unroll
    for (int a = 0; a < 3; a++) xv[a] = ld_stream(&A[(F_XV0 + a) * cap + slot]);
    bool bad = false;
    int b[3];
    Axis W[3];
    // ---- G2P (Extrapolation.cpp:127-163, 288-367) ----
#pragma unroll
    for (int a = 0; a < 3; a++) {
      b[a] = stencil_base(q[a], I.lpq, I.fbits, I.n[a] - 3, bad);
      W[a] = gimp3((float)(q[a] - (b[a] << I.fbits)) * I.inv_2F, I.lp, I.inv_lp, I.inv_h);
    }
    float vg[3] = {0, 0, 0}, ap[3] = {0, 0, 0};
    float Lx[3] = {0, 0, 0}, Ly[3] = {0, 0, 0}, Lz[3] = {0, 0, 0};  // d v_b / d x_a
    int r0 = b[0] - o[0], r1 = b[1] - o[1], r2 = b[2] - o[2];
    if ((unsigned)r0 >= (unsigned)TILE || (unsigned)r1 >= (unsigned)TILE || (unsigned)r2 >= (unsigned)TILE) {
      // a particle listed under a tile it is not in: never expected; keep LDS reads in range
      atomicOr(&d.ctrl->status, 2);
      r0 = min(max(r0, 0), TILE - 1);
      r1 = min(max(r1, 0), TILE - 1);
      r2 = min(max(r2, 0), TILE - 1);
    }
#pragma unroll 1
    for (int kk = 0; kk < 3; kk++) {
      const float wz = sel3(kk, W[2].w[0], W[2].w[1], W[2].w[2]), gz = sel3(kk, W[2].g[0], W[2].g[1], W[2].g[2]);
#pragma unroll
      for (int jj = 0; jj < 3; jj++) {
        const float Aw = W[1].w[jj] * wz;
        const float By = W[1].g[jj] * wz;
        const float Bz = W[1].w[jj] * gz;
        float S0[3] = {0, 0, 0}, S1[3] = {0, 0, 0}, Sa[3] = {0, 0, 0};
        const int row = r0 + G2P_R * ((r1 + jj) + G2P_R * (r2 + kk));
#pragma unroll
        for (int ii2 = 0; ii2 < 3; ii2++) {
          const int n = row + ii2;
          const float wx = W[0].w[ii2], gx = W[0].g[ii2];
          float4 q0 = s_vel[n][0], q1 = s_vel[n][1];
          if (MB) {
            const int ob = tag_obj(tag);
            const int sl = (s_obj[0][n] == ob) ? 0 : (s_obj[1][n] == ob) ? 1 : 2;
            const float4 z4 = make_float4(0.f, 0.f, 0.f, 0.f);
            if (sl != 0) {
              q0 = (sl == 1) ? s_vel2[n][0] : z4;
              q1 = (sl == 1) ? s_vel2[n][1] : z4;
            }
          }
          const float vn[3] = {q0.x, q0.y, q0.z}, an[3] = {q0.w, q1.x, q1.y};
#pragma unroll
          for (int c = 0; c < 3; c++) {
            S0[c] += wx * vn[c];
            S1[c] += gx * vn[c];
            Sa[c] += wx * an[c];
          }
        }
#pragma unroll
        for (int c = 0; c < 3; c++) {
          vg[c] += Aw * S0[c];
          ap[c] += Aw * Sa[c];
          Lx[c] += Aw * S1[c];
          Ly[c] += By * S0[c];
          Lz[c] += Bz * S0[c];
        }
      }
    }
    vg[0] += vref.x;
    vg[1] += vref.y;
    vg[2] += vref.z;
    float er[6];
    er[0] = Lx[0];
    er[1] = Ly[1];
    er[2] = Lz[2];
    er[3] = 0.5f * (Lx[1] + Ly[0]);
    er[4] = 0.5f * (Ly[2] + Lz[1]);
    er[5] = 0.5f * (Lx[2] + Lz[0]);
    const float rho_dot = -rho * (Lx[0] + Ly[1] + Lz[2]);
    publish();

    const int bct = (MASK & KM_PBC) ? tag_bct(tag) : 0;
    float bv[4] = {0, 0, 0, 0};
    const int pid = (int)ld_stream(&A[F_PID * cap + slot]);
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
    // ---- particle update (TimeIntegration.cpp:11-117) ----
    const float dt = I.dt;
    rho += rho_dot * dt;
    if (!tag_hold(tag)) {
#pragma unroll
      for (int a = 0; a < 3; a++) {
        float vr = xv_res(xv[a]);
        uint32_t lo = xv_lo(xv[a]);
        vel_add(v[a], vr, (ap[a] - I.damping * v[a]) * dt);
        advance_q(q[a], lo, vg[a] * I.qscale);
        xv[a] = xv_pack(lo, vr);
      }
    }
    float e[6];
#pragma unroll
    for (int a = 0; a < 6; a++) e[a] = f32(ld_stream(&A[(F_E0 + a) * cap + slot])) + er[a] * dt;
#pragma unroll
    for (int a = 0; a < 6; a++) B[(F_E0 + a) * cap + out] = u32(e[a]);
    if (!M.rigid) {
      if ((MASK & KM_ELASTIC) && M.model == 1) cm_elastic(M, er, s, dt);
      if ((MASK & KM_NEWTON) && M.model == 2) cm_newtonian(M, er, e, s, dt);
      if ((MASK & (KM_PERFECT | KM_DP)) && M.carry_plastic) {
        float p[6];
#pragma unroll
        for (int a = 0; a < 6; a++) p[a] = f32(ld_stream(&A[(F_P0 + a) * cap + slot]));
        float epb = f32(ld_stream(&A[F_EPB * cap + slot]));
        if ((MASK & KM_PERFECT) && M.model == 3) cm_perfect(M, er, s, p, epb, dt);
        if ((MASK & KM_DP) && M.model == 4) {
          if (DPNN && M.dp_mode == GY_DP_RETURN_LEARNED)
            dp_learned_gate(s_nnn, M, tag_mat(tag), slot, er, s, p, epb, dt, *pd);
          else
            cm_drucker_prager<DPLOOP>(M, er, s, p, epb, dt);
        }
        if (!DPNN || pd->state != 1) {  // a sample waiting for the network stores after it
#pragma unroll
          for (int a = 0; a < 6; a++) B[(F_P0 + a) * cap + out] = u32(p[a]);
          B[F_EPB * cap + out] = u32(epb);
        }
      }
      if (MASK & (KM_GK | KM_TURB)) {
        if (M.carry_damage) {
          float dmg = f32(ld_stream(&A[F_DMG * cap + slot])), smax = f32(ld_stream(&A[F_SMAX * cap + slot]));
          if ((MASK & KM_TURB) && M.model == 5) cm_turbulent(M, er, e, v, mass, dmg, s, dt);
          if ((MASK & KM_GK) && M.gk) gk_damage(M, s, dmg, smax, mass, rho, dt);
          B[F_DMG * cap + out] = u32(dmg);
          B[F_SMAX * cap + out] = u32(smax);
        }
      }
    } else {
      // rigid particles keep their stress, plastic and damage fields
      if (M.carry_plastic) {
#pragma unroll
        for (int a = 0; a < 6; a++) B[(F_P0 + a) * cap + out] = ld_stream(&A[(F_P0 + a) * cap + slot]);
        B[F_EPB * cap + out] = ld_stream(&A[F_EPB * cap + slot]);
      }
      if (M.carry_damage) {
        B[F_DMG * cap + out] = ld_stream(&A[F_DMG * cap + slot]);
        B[F_SMAX * cap + out] = ld_stream(&A[F_SMAX * cap + slot]);
      }
    }
    if (!(MASK & (KM_PERFECT | KM_DP)) && M.carry_plastic && !M.rigid) {
      // plastic fields uploaded for a model that never changes them
#pragma unroll
      for (int a = 0; a < 6; a++) B[(F_P0 + a) * cap + out] = ld_stream(&A[(F_P0 + a) * cap + slot]);
      B[F_EPB * cap + out] = ld_stream(&A[F_EPB * cap + slot]);
    }
    if (!(MASK & (KM_GK | KM_TURB)) && M.carry_damage && !M.rigid) {
      B[F_DMG * cap + out] = ld_stream(&A[F_DMG * cap + slot]);
      B[F_SMAX * cap + out] = ld_stream(&A[F_SMAX * cap + slot]);
    }
    // periodic wrap with period L-2h (TimeIntegration.cpp:80-108)
#pragma unroll
    for (int a = 0; a < 3; a++) {
      if (q[a] > I.wrapq[a]) q[a] = I.oneq + q[a] % I.wrapq[a];
      if (q[a] < I.oneq) q[a] += I.wrapq[a] - I.oneq;
    }
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
#pragma unroll
    for (int a = 0; a < 3; a++) B[(F_QX + a) * cap + out] = (uint32_t)q[a];
#pragma unroll
    for (int a = 0; a < 3; a++) B[(F_VX + a) * cap + out] = u32(v[a]);
#pragma unroll
    for (int a = 0; a < 3; a++) B[(F_XV0 + a) * cap + out] = xv[a];
    if (!DPNN || pd->state != 1) {
#pragma unroll
      for (int a = 0; a < 6; a++) B[(F_S0 + a) * cap + out] = u32(s[a]);
    }
    B[F_RHO * cap + out] = u32(rho);
    B[F_MASS * cap + out] = u32(mass);
    B[F_TAG * cap + out] = tag;
    B[F_PID * cap + out] = (uint32_t)pid;
    if (d.thermal) {  // temperature and gradient travel with the particle (gy_thermal.hip)
#pragma unroll
      for (int a = 0; a < 4; a++) B[(F_T + a) * cap + out] = ld_stream(&A[(F_T + a) * cap + slot]);
    }

    // ---- new binning key ----
    bad = false;
#pragma unroll
    for (int a = 0; a < 3; a++) b[a] = stencil_base(q[a], I.lpq, I.fbits, I.n[a] - 3, bad);
    knew = key_of(I, b);
    (par ? d.key[0] : d.key[1])[out] = knew;
    if (bad) { atomicAdd(&d.ctrl->oor, 1); atomicOr(&d.ctrl->status, 1); }
  }

  if (s_rec == nullptr) return knew;
  // ---- P2G record of the (new) state (Extrapolation.cpp:166-286) ----
  const float mp = M.rigid ? 1.0f : mass;  // rigid particles scatter unit mass (Extrapolation.cpp:183)
  const float nV = -(mp / rho);
  if (DPNN) pd->nV = nV;
  float4 *w = &s_rec->w[RecLds::wi(j)];
  w[0] = make_float4(__int_as_float(q[0]), __int_as_float(q[1]), __int_as_float(q[2]), mp);
  w[1] = make_float4(mp * v[0], mp * v[1], mp * v[2], nV * s[0]);
  w[2] = make_float4(nV * s[1], nV * s[2], nV * s[3], nV * s[4]);
  s_rec->t[RecLds::ti(j)] = nV * s[5];
  return knew;
}

// Contact G2P staging of node n of a tile's G2P region: both body slots' velocity and
// acceleration and the objects that hold them; zero at nodes without mass, which G2P
// skips (Extrapolation.cpp:290).
__device__ __forceinline__ void stage_body_vel(const DevPtrs &d, const InstDesc &I, const int o[3], int n,
                                               float4 (*s_vel)[2], float4 (*s_vel2)[2], int (*s_obj)[NODES_G2P]) {
  const int x = n % G2P_R, y = (n / G2P_R) % G2P_R, z = n / (G2P_R * G2P_R);
  const int g0 = o[0] + x, g1 = o[1] + y, g2 = o[2] + z;
  float vq[MB_VELB_W];
  for (int f = 0; f < MB_VELB_W; f++) vq[f] = 0.f;
  int4 cl = make_int4(-1, -1, 0, 0);
  if (g0 < I.n[0] && g1 < I.n[1] && g2 < I.n[2]) {
    const long long ni = I.node_base + g0 + (long long)I.n[0] * (g1 + (long long)I.n[1] * g2);
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

// Phase B: GIMP weights of one record's position (Extrapolation.cpp:556-709).
struct ScatterRec {
  int b[3];
  Axis W[3];
  P2GTerms T;
};
__device__ __forceinline__ void load_rec(const InstDesc &I, const RecLds *s_rec, int j, ScatterRec &R) {
  const float4 *w = &s_rec->w[RecLds::wi(j)];
  const float4 w0 = w[0], w1 = w[1], w2 = w[2];
  const float qf[3] = {w0.x, w0.y, w0.z};
  bool bad = false;
#pragma unroll
  for (int a = 0; a < 3; a++) {
    const int q = __float_as_int(qf[a]);
    R.b[a] = stencil_base(q, I.lpq, I.fbits, I.n[a] - 3, bad);
    R.W[a] = gimp3((float)(q - (R.b[a] << I.fbits)) * I.inv_2F, I.lp, I.inv_lp, I.inv_h);
  }
  R.T.m = w0.w;
  R.T.mv[0] = w1.x; R.T.mv[1] = w1.y; R.T.mv[2] = w1.z;
  R.T.vs[0] = w1.w; R.T.vs[1] = w2.x; R.T.vs[2] = w2.y; R.T.vs[3] = w2.z; R.T.vs[4] = w2.w;
  R.T.vs[5] = s_rec->t[RecLds::ti(j)];
}

__device__ __forceinline__ bool region_offset(const ScatterRec &R, const int o[3], int r[3]) {
#pragma unroll
  for (int a = 0; a < 3; a++) r[a] = R.b[a] - o[a] + 1;
  return r[0] >= 0 && r[1] >= 0 && r[2] >= 0 && r[0] + 2 < P2G_R && r[1] + 2 < P2G_R && r[2] + 2 < P2G_R;
}

// 27-node scatter of m, m*v and -V*sigma.gradN (Extrapolation.cpp:166-286) of one
// particle into the tile's FP64 LDS region.
__device__ __forceinline__ void scatter_lds(const ScatterRec &P, const int r[3], double (*s_acc)[ACC_N]) {
#pragma unroll 1
  for (int kk = 0; kk < 3; kk++) {
    const float wz = sel3(kk, P.W[2].w[0], P.W[2].w[1], P.W[2].w[2]);
    const float gz = sel3(kk, P.W[2].g[0], P.W[2].g[1], P.W[2].g[2]);
#pragma unroll
    for (int jj = 0; jj < 3; jj++) {
      const float Aw = P.W[1].w[jj] * wz, By = P.W[1].g[jj] * wz, Bz = P.W[1].w[jj] * gz;
#pragma unroll
      for (int ii2 = 0; ii2 < 3; ii2++) {
        const int a = acc_index(r[0] + ii2, r[1] + jj, r[2] + kk);
        float tt[7];
        node_terms(P.T, P.W[0].w[ii2] * Aw, P.W[0].g[ii2] * Aw, P.W[0].w[ii2] * By, P.W[0].w[ii2] * Bz, tt);
#pragma unroll
        for (int f = 0; f < 7; f++) atomicAdd(&s_acc[f][a], (double)tt[f]);
      }
    }
  }
}

// Phase B for the record in column j: into the LDS region, or (a particle whose new
// stencil left the region) by global atomics.
__device__ __forceinline__ void particle_scatter(const DevPtrs &d, const InstDesc &I, const int o[3], int inst,
                                                 const RecLds *s_rec, int j, double (*s_acc)[ACC_N]) {
  ScatterRec P;
  load_rec(I, s_rec, j, P);
  int r[3];
  if (region_offset(P, o, r)) scatter_lds(P, r, s_acc);
  else p2g_global(d.acc, d.brick_flag, &d.inst[inst], P.b[0], P.b[1], P.b[2], P.W[0], P.W[1], P.W[2], P.T);
}

// Learned return map, after the network: the cone multiplier from the network output y
// (the closed form near the return's branch switch), the return along the exact
// direction (apex return exact), the stores of the stress and plastic fields at sorted
// position out, and the stress terms of the record in column j. Returns whether the
// network's multiplier was used.
__device__ __forceinline__ bool learned_return(const DevPtrs &d, const MatDev *s_mat, const float *norm,
                                               const DpPend &pd, float y, int par, int out, RecLds *s_rec, int j) {
  const MatDev &M = s_mat[pd.mat];
  float dL = dp_learned_multiplier(M, pd.T, norm[NN_NORM_OUT] * y + norm[NN_NORM_OUT + 1], norm[NN_NORM_YLO],
                                   norm[NN_NORM_YHI]);
  const bool used = dL >= 0.0f;
  const uint32_t *A = (par ? d.pbuf[1] : d.pbuf[0]);
  uint32_t *B = (par ? d.pbuf[0] : d.pbuf[1]);
  const long long cap = d.cap;
  float s[6], p[6], epb = pd.epb;
#pragma unroll
  for (int k = 0; k < 6; k++) p[k] = pd.p[k];
  if (!used) dL = dp_cone_multiplier<false>(M, pd.T, epb);
  if (!dp_return<false>(M, pd.T, dL, s, p, epb)) {
    // an apex return with a negative multiplier keeps the stress before the step
#pragma unroll
    for (int k = 0; k < 6; k++) s[k] = f32(ld_stream(&A[(F_S0 + k) * cap + pd.slot]));
  }
#pragma unroll
  for (int a = 0; a < 6; a++) {
    B[(F_S0 + a) * cap + out] = u32(s[a]);
    B[(F_P0 + a) * cap + out] = u32(p[a]);
  }
  B[F_EPB * cap + out] = u32(epb);
  const float nV = pd.nV;
  float4 *w = &s_rec->w[RecLds::wi(j)];
  w[1].w = nV * s[0];
  w[2] = make_float4(nV * s[1], nV * s[2], nV * s[3], nV * s[4]);
  s_rec->t[RecLds::ti(j)] = nV * s[5];
  return used;
}

// The fused particle kernel, persistent over the active tile list. UPDATE = false: P2G of
// the current state only (first step after an upload). UPDATE = true: G2P + update +
// constitutive + P2G of the new state.
//
// A tile's list entries are taken in chunks of CHUNK, in two phases per chunk:
//  A. lane = consecutive entry (coalesced state reads and writes): G2P from LDS-staged
//     node velocities, update, constitutive/damage models, store at the entry's sorted
//     position, new key with one histogram atomic per run of equal keys, and the P2G
//     record into LDS;
//  B. lane = record, 27-node scatter into the FP64 LDS region [o-1, o+TILE+3)^3.
// ds_add_f64 is a fast LDS atomic on gfx950, unlike FP32 LDS atomic adds; it serves 16
// lanes per pass, and two lanes of a pass conflict when their nodes share a bank pair.
// With the region layout of acc_index, cells c and c' share bank pairs (at every stencil
// offset) iff c = c' mod 16, their "class". Phase A therefore writes each record to the
// column that a class-sorted, transposed order gives it: the chunk's entries sorted by
// class (whole cells, so the entries of one class are consecutive: position j), and
// j -> column 16*(j mod P) + j/P for P = ceil(n/16) passes. The entries of a class, the
// particles of one cell among them, land in different passes unless the class holds
// more than P entries of the chunk. After the last
// chunk the region is written as FP32 partial sums with plain stores and the tile's epoch
// is stamped; the grid kernel gathers the partials of the tiles covering each node in a
// fixed order.
//
// MB (two-body contact): phase A is the contact G2P (body velocity fields s_vel, s_vel2
// and the slots' objects staged per node) and phase B the centre-of-mass P2G; the body
// fields follow the re-binning and the claims (k_mb_p2g). Body 2's fields take LDS
// beyond three workgroups per CU, so the contact variant runs two.

// WAVES: resident waves per SIMD the build targets, 3 (168 VGPRs) or 2 (256 VGPRs, no
// spills, fewer waves to hide latency); the autotuner picks one per workload.
template <bool UPDATE, int MASK, bool MB = false, int WAVES = 3>
__global__ __launch_bounds__(WG) __attribute__((amdgpu_waves_per_eu(MB ? 2 : WAVES))) void k_g2p2g(DevPtrs d) {
  static_assert(WG == 4 * TILE3, "four waves per tile");
  static_assert(!MB || UPDATE, "the contact variant moves particles");
  constexpr bool DPNN = (MASK & KM_DPNN) != 0;
  static_assert(!DPNN || (UPDATE && !MB), "the network runs in the single-body stepping kernel");
  constexpr int NB2 = MB ? NODES_G2P : 1;
  __shared__ float4 s_vel[NODES_G2P][2];
  __shared__ float4 s_vel2[NB2][2];
  __shared__ int s_obj[2][NB2];
  __shared__ __align__(16) double s_acc[7][ACC_N];
  __shared__ RecLds s_rec;
  __shared__ int s_off[TILE3 + 1];
  __shared__ MatDev s_mat[MAXMAT];
  __shared__ int s_tile;
  __shared__ int s_refkey[WG / 64];
  __shared__ float4 s_refv[WG / 64];
  __shared__ uint2 s_nnw[DPNN ? NN_TILES * 64 : 1];               // network fragments
  __shared__ __align__(16) float s_nnb[DPNN ? NN_NBIAS : 4];       // network biases
  __shared__ __align__(16) float s_nnn[DPNN ? NN_NORM_N : 4];      // network normalisation
  // a chunk's waiting samples, compacted: network inputs 0-3, 4-5 and 6 (input 7 is the
  // zero padding), and the network's output in place of input 6; per wave the count
  __shared__ uint4 s_nnx0[DPNN ? WG : 1];
  __shared__ uint2 s_nnx1[DPNN ? WG : 1];
  __shared__ uint32_t s_nnx2[DPNN ? WG : 1];
  __shared__ int s_nncnt[DPNN ? WG / 64 : 1];
  const int tid = threadIdx.x;
  const int par = d.ctrl->parity;
  const int epoch_w = d.ctrl->epoch + (UPDATE ? 1 : 0);
  const int nh = cld(&d.ctrl->heavy_tiles), na = cld(&d.ctrl->active_tiles);
  unsigned nn_net = 0, nn_fb = 0;  // this wave's network and fallback samples
  if constexpr (DPNN) nn_stage(d.nn, s_nnw, s_nnb, s_nnn);  // visible after the loop's first barrier
  // the grid update before this kernel has read the brick list: the bricks this kernel
  // flags are listed from an empty list (by either scan kernel)
  if (blockIdx.x == 0 && tid == 0) d.ctrl->active_bricks = 0;

  // A tile's successor is claimed when its last chunk starts (one work counter, longest
  // tiles first); the successor's offsets and node velocities are loaded during that
  // chunk's scatter and put in place after it, so a tile starts with no load round trip.
  constexpr bool VEL = UPDATE && !MB;
  int off_r = 0;
  float4 va = make_float4(0.f, 0.f, 0.f, 0.f), vb = va;
  uint32_t vkeep = 0;
  if (tid == 0) s_tile = atomicAdd(&d.ctrl->next, 1);
  __syncthreads();
  int t = claimed_tile(d, nh, na, __builtin_amdgcn_readfirstlane(s_tile));
  if (t >= 0) {
    tile_loads<VEL>(d, t, tile_geo(d, t), tid, off_r, va, vb, vkeep);
    stage_ahead<VEL>(tid, off_r, va, vb, vkeep, s_off, s_vel);
  }
  int ii_mat = -1;  // instance whose materials s_mat holds
  for (;;) {
    __syncthreads();  // the staged offsets and node records are in place
    if (t < 0) {
      if (DPNN && (tid & 63) == 0 && (nn_net | nn_fb)) {
        atomicAdd(&d.ctrl->nn_count[0], (unsigned long long)nn_net);
        atomicAdd(&d.ctrl->nn_count[1], (unsigned long long)nn_fb);
      }
      break;
    }
    // workgroup-uniform values (scalar loads): keep them, the instance description
    // included, in scalar registers
    const TileGeo g = tile_geo(d, t);
    const int ii = g.ii;
    const InstDesc I = cld(&d.inst[ii]);
    const int o[3] = {g.o[0], g.o[1], g.o[2]};
    float3 vref = make_float3(0.f, 0.f, 0.f);  // staged node velocities are relative to vref

    {
      double2 *z = reinterpret_cast<double2 *>(&s_acc[0][0]);
      for (int n = tid; n < 7 * ACC_N / 2; n += WG) z[n] = make_double2(0.0, 0.0);
    }
    if (ii != ii_mat) {  // workgroup-uniform; a single-instance context loads them once
      if (tid < I.nmat && tid < MAXMAT) s_mat[tid] = d.mats[I.mat_base + tid];
      ii_mat = ii;
    }
    if constexpr (MB) {
      for (int n = tid; n < NODES_G2P; n += WG) stage_body_vel(d, I, o, n, s_vel, s_vel2, s_obj);
    } else if (UPDATE) {
      // Stage the region's node velocity and acceleration; nodes without mass stage zero
      // (G2P skips them, Extrapolation.cpp:290). The velocities are then staged relative to
      // a reference velocity of the tile, the massive node nearest the region centre: the
      // 27 GIMP weights of an interior stencil sum to 1 and their gradients to 0, so
      // sum(N v) = v_ref + sum(N (v - v_ref)) and sum(gradN v) = sum(gradN (v - v_ref)),
      // and the FP32 sums then round the small differences instead of the full velocities
      // (the velocity gradient of a slowly deforming, fast moving body is otherwise lost
      // to cancellation).
      static_assert(NODES_G2P <= WG, "one staged node per thread");
      const int n = tid;
      float4 v0 = make_float4(0.f, 0.f, 0.f, 0.f), v1 = v0;
      int key = 0x7fffffff;
      if (n < NODES_G2P) {
        const int x = n % G2P_R, y = (n / G2P_R) % G2P_R, z = n / (G2P_R * G2P_R);
        v0 = s_vel[n][0];  // the raw record this thread staged ahead of the tile
        v1 = s_vel[n][1];
        if (v1.z == 0.0f) {
          v0 = make_float4(0.f, 0.f, 0.f, 0.f);
          v1 = v0;
        } else {
          key = ((abs(2 * x - (G2P_R - 1)) + abs(2 * y - (G2P_R - 1)) + abs(2 * z - (G2P_R - 1))) << 8) | n;
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
      float4 r = s_refv[best];  // read unconditionally: a select of addresses is a generic load
      if (s_refkey[best] == 0x7fffffff) r = make_float4(0.f, 0.f, 0.f, 0.f);
      vref = make_float3(r.x, r.y, r.z);
      if (n < NODES_G2P) {
        s_vel[n][0] = make_float4(v0.x - r.x, v0.y - r.y, v0.z - r.z, v0.w);
        s_vel[n][1] = v1;
      }
    }
    __syncthreads();

    const double tl = UPDATE ? cld(&d.t_last[ii]) : 0.0;
    const int p0 = __builtin_amdgcn_readfirstlane(s_off[0]), p1 = __builtin_amdgcn_readfirstlane(s_off[TILE3]);
    int t_next = -1;
    if (p0 >= p1) {  // a tile without entries (never in the active list)
      __syncthreads();  // every thread has read p0 and p1
      if (tid == 0) s_tile = atomicAdd(&d.ctrl->next, 1);
      __syncthreads();
      t_next = claimed_tile(d, nh, na, __builtin_amdgcn_readfirstlane(s_tile));
      if (t_next >= 0) {
        tile_loads<VEL>(d, t_next, tile_geo(d, t_next), tid, off_r, va, vb, vkeep);
        stage_ahead<VEL>(tid, off_r, va, vb, vkeep, s_off, s_vel);
      }
    }
    for (int c0 = p0; c0 < p1; c0 += CHUNK) {
      const bool last = c0 + CHUNK >= p1;  // workgroup-uniform
      // the successor's claim, handed over in LDS after this update's G2P (before its stores,
      // so nothing waits for the counter's round trip behind them)
      int kclaim = 0;
      if (last && tid == 0) kclaim = atomicAdd(&d.ctrl->next, 1);
      const auto publish = [&] {
        if (last && tid == 0) s_tile = kclaim;
      };
      const int c1 = min(c0 + CHUNK, p1), nc = c1 - c0;
      // cell of this lane's entry: the last cell whose first entry is <= i (two rounds of
      // LDS reads over the tile's 65 offsets)
      const int i = c0 + tid;
      const int lane = tid & 63;
      int ce = 0;
      {
        int k = 0;
#pragma unroll
        for (int m = 1; m < 8; m++) k += (s_off[8 * m] <= i) ? 1 : 0;
        int c = 8 * k;
#pragma unroll
        for (int m = 1; m < 8; m++) c += (s_off[8 * k + m] <= i) ? 1 : 0;
        ce = c;
      }
      // class-sorted position j of the entry: the chunk's cells in class order (c mod 16,
      // then c / 16), an exclusive scan of their in-chunk counts over the wave (lane = cell)
      const int clo = max(s_off[lane], c0), chi = min(s_off[lane + 1], c1);
      const int ncell = max(chi - clo, 0);
      const int nL = __shfl(ncell, 16 * (lane & 3) + (lane >> 2), 64);  // lane L: cell in class order
      int incl = nL;
#pragma unroll
      for (int w = 1; w < 64; w <<= 1) {
        const int u = __shfl_up(incl, w, 64);
        if (lane >= w) incl += u;
      }
      const int cpos = __shfl(incl - nL, 4 * (ce & 15) + (ce >> 4), 64);
      const int cfirst = __shfl(clo, ce, 64);
      const int P = (nc + 15) >> 4;
      const float invP = 1.0f / (float)P;
      const int j = cpos + (i - cfirst);
      const int jq = (int)(((float)j + 0.5f) * invP);  // j / P (exact: j < 256, P <= 16)
      const int col = 16 * (j - jq * P) + jq;
      // phase A
      int knew = -1;
      DpPend pd;
      if constexpr (DPNN) {
        pd.state = 0;
#pragma unroll
        for (int k = 0; k < 8; k++) pd.x[k] = 0u;
      }
      if (i < c1) {
        knew = particle_update<UPDATE, MASK, MB>(d, I, o, tl, i, d.tmp[i], s_vel, vref, s_mat, par, &s_rec, col,
                                                 MB ? s_vel2 : nullptr,
                                                 MB ? reinterpret_cast<int (*)[NODES_G2P]>(s_obj) : nullptr,
                                                 DPNN ? &pd : nullptr, DPNN ? s_nnn : nullptr, publish);
        if constexpr (!UPDATE) publish();  // (no G2P in the P2G-only pass)
      }
      if constexpr (DPNN) {
        // the chunk's waiting samples through the network, compacted over the workgroup
        // (position = waiting samples of lower waves and lanes), then their return, stores
        // and record. Wave w evaluates 16-sample blocks w, w + 4, w + 8, w + 12 of the
        // compacted list, block w + 4m in its lanes 16m .. 16m + 15. On the 4.2M-particle
        // column about 40 % of the lanes wait; evaluating every wave with a waiting lane
        // cost 1055 us per launch against 1003 us compacted (the kept plastic fields
        // included).
        const int w = tid >> 6, lane = tid & 63;
        const unsigned long long wait = __ballot(pd.state == 1);
        bool closed = pd.state == 2;
        if (lane == 0) s_nncnt[w] = (int)__popcll(wait);
        __syncthreads();
        const int n0 = s_nncnt[0], n1 = s_nncnt[1], n2 = s_nncnt[2], n3 = s_nncnt[3];
        const int nw = n0 + n1 + n2 + n3;
        if (nw) {  // uniform over the workgroup
          const int q = (w > 0 ? n0 : 0) + (w > 1 ? n1 : 0) + (w > 2 ? n2 : 0) +
                        (int)__popcll(wait & ((1ull << lane) - 1ull));
          if (pd.state == 1) {
            s_nnx0[q] = make_uint4(pd.x[0], pd.x[1], pd.x[2], pd.x[3]);
            s_nnx1[q] = make_uint2(pd.x[4], pd.x[5]);
            s_nnx2[q] = pd.x[6];
          }
          __syncthreads();
          const int kw = (((nw + 15) >> 4) - w + 3) >> 2;  // blocks of this wave
          if (kw > 0) {  // uniform over the wave
            const int sq = 16 * w + 64 * (lane >> 4) + (lane & 15);
            const bool v = sq < nw;
            const int sr = v ? sq : 16 * w;  // lanes past the end read the wave's first row
            const uint4 a = s_nnx0[sr];
            const uint2 b = s_nnx1[sr];
            const uint32_t x[8] = {a.x, a.y, a.z, a.w, b.x, b.y, s_nnx2[sr], 0u};
            const unsigned long long want = __ballot(v);
            // blocks interleaved: more independent matrix-core and SiLU work in flight
            float y;
            if (kw >= 3) y = nn_wave_forward<4>(s_nnw, s_nnb, x, want);
            else if (kw == 2) y = nn_wave_forward<2>(s_nnw, s_nnb, x, want);
            else y = nn_wave_forward<1>(s_nnw, s_nnb, x, want);
            if (v) s_nnx2[sq] = __float_as_uint(y);  // this wave's own rows only
          }
          __syncthreads();
          if (pd.state == 1)
            closed = !learned_return(d, s_mat, s_nnn, pd, __uint_as_float(s_nnx2[q]), par, i, &s_rec, col);
        }
        const unsigned nfb = (unsigned)__popcll(__ballot(closed));
        nn_fb += nfb;
        nn_net += (unsigned)__popcll(wait) + (unsigned)__popcll(__ballot(pd.state == 2)) - nfb;
      }
      int hl = 0, hbase = 0;
      if (UPDATE) hbase = hist_issue(d.count, knew, hl);
      __syncthreads();
      // the successor's offsets and node records, in flight during the scatter
      if (last) {
        t_next = claimed_tile(d, nh, na, __builtin_amdgcn_readfirstlane(s_tile));
        if (t_next >= 0) tile_loads<VEL>(d, t_next, tile_geo(d, t_next), tid, off_r, va, vb, vkeep);
      }
      // phase B: column tid holds class-sorted entry (tid % 16) * P + tid / 16
      if ((tid >> 4) < P && (tid & 15) * P + (tid >> 4) < nc) particle_scatter(d, I, o, ii, &s_rec, tid, s_acc);
      int r = 0;
      if (UPDATE) r = hist_rank(hbase, hl);
      // this tile's last reads of s_off and s_vel were in the update above
      if (last && t_next >= 0) stage_ahead<VEL>(tid, off_r, va, vb, vkeep, s_off, s_vel);
      if (UPDATE && knew >= 0) d.rank[i] = r;
      __syncthreads();
    }

    // ---- flush: core sums as FP32 partials (plain stores), the halo shell only when a
    // particle reached it, brick flags, epochs ----
    float *pc = d.pcore + (long long)t * CORE_W;
    bool halo = false;
    static_assert(P2G_NODES % WG == 0, "whole waves per flush round");
    for (int n = tid; n < P2G_NODES; n += WG) {
      const int x = n % P2G_R, y = (n / P2G_R) % P2G_R, z = n / (P2G_R * P2G_R);
      const int a = acc_index(x, y, z);
      const bool shell = x == 0 || y == 0 || z == 0 || x > CORE_R || y > CORE_R || z > CORE_R;
      if (!shell && s_acc[0][a] != 0.0) {  // nodes without mass are never read
        const int c = (x - 1) + CORE_R * ((y - 1) + CORE_R * (z - 1));
        float4 *q = reinterpret_cast<float4 *>(pc + 8 * c);
        q[0] = make_float4((float)s_acc[0][a], (float)s_acc[1][a], (float)s_acc[2][a], (float)s_acc[3][a]);
        q[1] = make_float4((float)s_acc[4][a], (float)s_acc[5][a], (float)s_acc[6][a], 0.0f);
      }
      // node mask: the 64 nodes of this wave's round are region nodes n - lane .. n - lane + 63
      const bool has = s_acc[0][a] != 0.0;
      const unsigned long long bits = __ballot(has);
      if ((tid & 63) == 0) d.nodemask[(long long)t * (P2G_NODES / 64) + (n >> 6)] = bits;
      if (!has) continue;
      halo = halo || shell;
      const int g[3] = {o[0] - 1 + x, o[1] - 1 + y, o[2] - 1 + z};
      if (g[0] < 0 || g[1] < 0 || g[2] < 0 || g[0] >= I.n[0] || g[1] >= I.n[1] || g[2] >= I.n[2]) continue;
      int c[3];
      (void)canon_index(I, g, c);
      flag_brick(d, I, c);
    }
    if (__syncthreads_or(halo)) {
      float *pt = d.partial + (long long)t * PARTIAL_W;
      for (int n = tid; n < P2G_NODES; n += WG) {
        const int x = n % P2G_R, y = (n / P2G_R) % P2G_R, z = n / (P2G_R * P2G_R);
        if (!(x == 0 || y == 0 || z == 0 || x > CORE_R || y > CORE_R || z > CORE_R)) continue;
        const int a = acc_index(x, y, z);
        if (s_acc[0][a] == 0.0) continue;  // not in the node mask: never read
        float4 *q = reinterpret_cast<float4 *>(pt + 8 * n);
        q[0] = make_float4((float)s_acc[0][a], (float)s_acc[1][a], (float)s_acc[2][a], (float)s_acc[3][a]);
        q[1] = make_float4((float)s_acc[4][a], (float)s_acc[5][a], (float)s_acc[6][a], 0.0f);
      }
      if (tid == 0) d.halo_epoch[t] = epoch_w;
    }
    if (tid == 0) d.tile_epoch[t] = epoch_w;
    t = t_next;
  }
}

// ---------------------------------------------------------------------------------------
// Two-body contact (contexts holding several objects or rigid materials).
//
// Per node the reference keeps centre-of-mass fields and two body slots: the first object
// a node's gather meets claims slot 1, the first different object slot 2
// (Extrapolation.cpp:191-208); P2G adds every particle to the centre-of-mass fields and
// to its object's slot (Extrapolation.cpp:210-280); the nodal solve corrects approaching
// bodies (Grid.cpp:9-245, 247-308); G2P reads the slot of the particle's object
// (Extrapolation.cpp:288-367). Here the step is: mb_grid, the contact variant of the
// fused tile kernel (G2P from the body slots + update + store + new key, then the
// centre-of-mass P2G of the new state into the tile partials), advance, binning, mb_claim
// (cell objects, then node claims from the new binning) and mb_p2g for the body fields. A
// node with one claimed body has body fields equal to its centre-of-mass fields (every
// particle its window sees is of that object), so body fields and normals are scattered
// only to nodes with two claimed bodies; rigid particles also scatter their mass to their
// body's slot wherever gravity is nonzero (gravity acts on the body fields of non-rigid
// mass only).

// Objects present in each cell of the active tiles: the two smallest object ids and
// whether each has a rigid particle there. A cell's particles are met in ascending object
// id when a node claims slots (the reference's order inside a cell is its sort order).
__global__ __launch_bounds__(WG) void k_mb_cells(DevPtrs d) {
  const int nact = d.ctrl->active_tiles;
  const uint32_t *A = (d.ctrl->parity ? d.pbuf[1] : d.pbuf[0]);
  for (int ta = blockIdx.x * (WG / TILE3) + (threadIdx.x / TILE3); ta < nact; ta += gridDim.x * (WG / TILE3)) {
    const int t = d.active[ta];
    const InstDesc &I = d.inst[d.tile_inst[t]];
    const int k = t * TILE3 + (threadIdx.x % TILE3);
    if (threadIdx.x % TILE3 == 0) d.twotile[t] = 0;  // set again by the claims that follow
    int a = 0x7fffffff, b = 0x7fffffff, ra = 0, rb = 0;
    for (int j = d.offsets[k]; j < d.offsets[k + 1]; j++) {
      const uint32_t tag = A[F_TAG * d.cap + d.tmp[j]];
      const int ob = tag_obj(tag), rg = d.mats[I.mat_base + tag_mat(tag)].rigid ? 1 : 0;
      if (ob == a) ra |= rg;
      else if (ob < a) { b = a; rb = ra; a = ob; ra = rg; }
      else if (ob == b) rb |= rg;
      else if (ob < b) { b = ob; rb = rg; }
    }
    d.cellobj[k] = make_int4(a == 0x7fffffff ? -1 : a, b == 0x7fffffff ? -1 : b, ra, rb);
  }
}

// Body claims of every node a particle of an active tile can reach (the core region of
// the tile): the node's window is the 4x4x4 cells nidx-2..nidx+1 with linear cell ids
// and only a range check (Extrapolation.cpp:36-66, so ids past a row end continue on the
// next row), met in ascending (k, j, i).
__global__ __launch_bounds__(WG) void k_mb_claim(DevPtrs d) {
  const long long nact = d.ctrl->active_tiles;
  for (long long it = (long long)blockIdx.x * WG + threadIdx.x; it < nact * CORE_NODES;
       it += (long long)gridDim.x * WG) {
    const int t = d.active[it / CORE_NODES], ln = (int)(it % CORE_NODES);
    const InstDesc &I = d.inst[d.tile_inst[t]];
    const int lt = t - I.tile_base;
    const int g0 = (lt % I.nt[0]) * TILE + ln % CORE_R, g1 = ((lt / I.nt[0]) % I.nt[1]) * TILE + (ln / CORE_R) % CORE_R,
              g2 = (lt / (I.nt[0] * I.nt[1])) * TILE + ln / (CORE_R * CORE_R);
    if (g0 >= I.n[0] || g1 >= I.n[1] || g2 >= I.n[2]) continue;
    const int Cx = I.n[0] - 1, Cy = I.n[1] - 1, Cz = I.n[2] - 1;
    const long long ncell = (long long)Cx * Cy * Cz;
    int b1 = -1, b2 = -1, r1 = 0, r2 = 0;
    for (int pass = 0; pass < 2; pass++)
      for (int kz = -2; kz < 2; kz++)
        for (int jy = -2; jy < 2; jy++)
          for (int ix = -2; ix < 2; ix++) {
            const long long lin = (g0 + ix) + (long long)(g1 + jy) * Cx + (long long)(g2 + kz) * Cx * Cy;
            if (lin < 0 || lin >= ncell) continue;
            const int c[3] = {(int)(lin % Cx), (int)((lin / Cx) % Cy), (int)(lin / ((long long)Cx * Cy))};
            if (c[0] > I.n[0] - 3 || c[1] > I.n[1] - 3 || c[2] > I.n[2] - 3) continue;  // no particle there
            const int k = key_of(I, c);
            if (d.offsets[k + 1] == d.offsets[k]) continue;
            const int4 co = d.cellobj[k];
            if (pass == 0) {
              const int ob[2] = {co.x, co.y};
              for (int q = 0; q < 2; q++) {
                if (ob[q] < 0) continue;
                if (b1 < 0) b1 = ob[q];
                else if (b2 < 0 && ob[q] != b1) b2 = ob[q];
              }
            } else {
              r1 |= (co.x == b1 && co.z) || (co.y == b1 && co.w);
              r2 |= b2 >= 0 && ((co.x == b2 && co.z) || (co.y == b2 && co.w));
            }
          }
    d.claim[I.node_base + g0 + (long long)I.n[0] * (g1 + (long long)I.n[1] * g2)] = make_int4(b1, b2, r1, r2);
  }
}

// The same claims with one workgroup per active tile: the objects of the 9 x 9 x 9 cells
// the windows of the tile's 6 x 6 x 6 core nodes cover, [o-2, o+7), are staged in LDS (an
// empty cell holds no object), then each core node walks its 64 window cells from LDS in
// the reference's order. One pass: a slot is claimed at the first cell holding its object,
// so the rigid flag of a slot only needs the cells from there on. Nodes whose window
// leaves the grid (where linear cell ids wrap to the next row) take the global walk.
constexpr int CLAIM_R = CORE_R + 3;  // staged cell edge: windows g-2..g+1 of nodes o..o+CORE_R-1
__global__ __launch_bounds__(WG) void k_mb_claim_tiles(DevPtrs d) {
  __shared__ int4 s_co[CLAIM_R * CLAIM_R * CLAIM_R];
  const int nact = d.ctrl->active_tiles;
  for (int ta = blockIdx.x; ta < nact; ta += gridDim.x) {
    const int t = d.active[ta];
    const InstDesc &I = d.inst[d.tile_inst[t]];
    const int lt = t - I.tile_base;
    const int o0 = (lt % I.nt[0]) * TILE, o1 = ((lt / I.nt[0]) % I.nt[1]) * TILE, o2 = (lt / (I.nt[0] * I.nt[1])) * TILE;
    for (int n = threadIdx.x; n < CLAIM_R * CLAIM_R * CLAIM_R; n += WG) {
      const int c[3] = {o0 - 2 + n % CLAIM_R, o1 - 2 + (n / CLAIM_R) % CLAIM_R, o2 - 2 + n / (CLAIM_R * CLAIM_R)};
      int4 co = make_int4(-1, -1, 0, 0);
      if (c[0] >= 0 && c[1] >= 0 && c[2] >= 0 && c[0] <= I.n[0] - 3 && c[1] <= I.n[1] - 3 && c[2] <= I.n[2] - 3) {
        const int k = key_of(I, c);
        if (d.offsets[k + 1] != d.offsets[k]) co = d.cellobj[k];
      }
      s_co[n] = co;
    }
    __syncthreads();
    int two = 0;
    if (threadIdx.x < CORE_NODES) {
      const int ln = threadIdx.x;
      const int g0 = o0 + ln % CORE_R, g1 = o1 + (ln / CORE_R) % CORE_R, g2 = o2 + ln / (CORE_R * CORE_R);
      const int Cx = I.n[0] - 1, Cy = I.n[1] - 1, Cz = I.n[2] - 1;
      if (g0 < I.n[0] && g1 < I.n[1] && g2 < I.n[2]) {
        int b1 = -1, b2 = -1, r1 = 0, r2 = 0;
        if (g0 >= 2 && g1 >= 2 && g2 >= 2 && g0 + 1 < Cx && g1 + 1 < Cy && g2 + 1 < Cz) {
          // window inside the grid: cell (g + d) for d in [-2, 1]^3, in (k, j, i) order
          for (int kz = -2; kz < 2; kz++)
            for (int jy = -2; jy < 2; jy++)
              for (int ix = -2; ix < 2; ix++) {
                const int4 co = s_co[(g0 + ix - (o0 - 2)) + CLAIM_R * ((g1 + jy - (o1 - 2)) + CLAIM_R * (g2 + kz - (o2 - 2)))];
                const int ob[2] = {co.x, co.y}, rg[2] = {co.z, co.w};
                for (int q = 0; q < 2; q++) {
                  if (ob[q] < 0) continue;
                  if (b1 < 0) b1 = ob[q];
                  else if (b2 < 0 && ob[q] != b1) b2 = ob[q];
                  if (ob[q] == b1) r1 |= rg[q];
                  else if (ob[q] == b2) r2 |= rg[q];
                }
              }
        } else {
          const long long ncell = (long long)Cx * Cy * Cz;
          for (int pass = 0; pass < 2; pass++)
            for (int kz = -2; kz < 2; kz++)
              for (int jy = -2; jy < 2; jy++)
                for (int ix = -2; ix < 2; ix++) {
                  const long long lin = (g0 + ix) + (long long)(g1 + jy) * Cx + (long long)(g2 + kz) * Cx * Cy;
                  if (lin < 0 || lin >= ncell) continue;
                  const int c[3] = {(int)(lin % Cx), (int)((lin / Cx) % Cy), (int)(lin / ((long long)Cx * Cy))};
                  if (c[0] > I.n[0] - 3 || c[1] > I.n[1] - 3 || c[2] > I.n[2] - 3) continue;
                  const int k = key_of(I, c);
                  if (d.offsets[k + 1] == d.offsets[k]) continue;
                  const int4 co = d.cellobj[k];
                  if (pass == 0) {
                    const int ob[2] = {co.x, co.y};
                    for (int q = 0; q < 2; q++) {
                      if (ob[q] < 0) continue;
                      if (b1 < 0) b1 = ob[q];
                      else if (b2 < 0 && ob[q] != b1) b2 = ob[q];
                    }
                  } else {
                    r1 |= (co.x == b1 && co.z) || (co.y == b1 && co.w);
                    r2 |= b2 >= 0 && ((co.x == b2 && co.z) || (co.y == b2 && co.w));
                  }
                }
        }
        d.claim[I.node_base + g0 + (long long)I.n[0] * (g1 + (long long)I.n[1] * g2)] = make_int4(b1, b2, r1, r2);
        two = b2 >= 0;
      }
    }
    // the tile's particles reach only its core nodes: without a two-body node there, its
    // non-rigid particles have no body fields to scatter (k_mb_p2g skips them)
    if (__syncthreads_or(two) ? 1 : 0) {
      if (threadIdx.x == 0) d.twotile[t] = 1;
    } else if (threadIdx.x == 0) {
      d.twotile[t] = 0;
    }
    __syncthreads();
  }
}

// Body-field P2G of one particle (slot order) at the nodes where two bodies are claimed:
// body fields to the slot its object holds at the actual node, folded onto the canonical
// node (rigid particles: unit mass, no gravity in their body's force,
// Extrapolation.cpp:183, 224-229), normals sum gradN * V per actual node
// (Extrapolation.cpp:231-234); rigid particles add their mass to their slot's rigid-mass
// sum at every node. FP32 global atomics.
__global__ __launch_bounds__(WG) void k_mb_p2g(DevPtrs d) {
  const int s = blockIdx.x * WG + threadIdx.x;
  const int lane = threadIdx.x & 63;
  const bool live = s < d.np;
  const int sc = live ? s : d.np - 1;  // dead lanes read a valid slot and contribute nothing
  const uint32_t *A = (d.ctrl->parity ? d.pbuf[1] : d.pbuf[0]);
  const long long cap = d.cap;
  const int ii = inst_of_slot(d, sc);
  const InstDesc &I = d.inst[ii];
  const uint32_t tag = A[F_TAG * cap + sc];
  const MatDev &M = d.mats[I.mat_base + tag_mat(tag)];
  // the rigid-mass sums only remove the rigid share of m*g from the body forces
  const bool rigid_mass = M.rigid && (I.gravity[0] != 0.0f || I.gravity[1] != 0.0f || I.gravity[2] != 0.0f);
  const int key = (d.ctrl->parity ? d.key[1] : d.key[0])[sc];
  const bool active = live && (rigid_mass || d.twotile[key / TILE3]);
  if (__ballot(active) == 0ull) return;  // wave-uniform
  const int obj = tag_obj(tag);
  int q[3], b[3];
  bool bad = false;
  Axis W[3];
  for (int a = 0; a < 3; a++) {
    q[a] = (int)A[(F_QX + a) * cap + sc];
    b[a] = stencil_base(q[a], I.lpq, I.fbits, I.n[a] - 3, bad);
    W[a] = gimp3((float)(q[a] - (b[a] << I.fbits)) * I.inv_2F, I.lp, I.inv_lp, I.inv_h);
  }
  const float mass = f32(A[F_MASS * cap + sc]), rho = f32(A[F_RHO * cap + sc]);
  const float mp = M.rigid ? 1.0f : mass;
  P2GTerms T;
  T.m = mp;
  for (int a = 0; a < 3; a++) T.mv[a] = mp * f32(A[(F_VX + a) * cap + sc]);
  const float V = mp / rho;
  for (int a = 0; a < 6; a++) T.vs[a] = -V * f32(A[(F_S0 + a) * cap + sc]);
  // Adjacent lanes with the same stencil base node, object and rigid-mass flag -- consecutive
  // slots are mostly the particles of one cell -- share all 27 target nodes and body slots:
  // their terms are summed over the run (segmented shuffle reduction) and the run's first
  // lane issues the atomics, which cuts the contended global atomics by the particles per
  // cell. Runs are numbered by the heads at or below a lane, so two runs of equal base in one
  // wave stay apart.
  const int base = (int)(I.node_base + b[0] + (long long)I.n[0] * (b[1] + (long long)I.n[1] * b[2]));
  const int seg_base = active ? base : -1 - lane;
  const int seg_obj = obj * 2 + (rigid_mass ? 1 : 0);
  // Every shuffle runs with the whole wave active (a lane switched off reads as 0), so none
  // sits behind a short-circuit test.
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
  // the claimed objects (the first two ints of a node's claim) of one z-layer of the
  // stencil are loaded together, ahead of that layer's atomics: one load latency per layer
  const int2 *claim2 = reinterpret_cast<const int2 *>(d.claim);
  for (int k = 0; k < 3; k++) {
    int2 cls[9];
#pragma unroll
    for (int j = 0; j < 3; j++)
#pragma unroll
      for (int i = 0; i < 3; i++)
        cls[3 * j + i] = claim2[2 * (base + i + (long long)I.n[0] * (j + (long long)I.n[1] * k))];
#pragma unroll
    for (int j = 0; j < 3; j++)
#pragma unroll
      for (int i = 0; i < 3; i++) {
        const float A_ = W[1].w[j] * W[2].w[k];
        const float w = W[0].w[i] * A_;
        const float gx = W[0].g[i] * A_, gy = W[0].w[i] * W[1].g[j] * W[2].w[k], gz = W[0].w[i] * W[1].w[j] * W[2].g[k];
        const int g[3] = {b[0] + i, b[1] + j, b[2] + k};
        const long long gi = I.node_base + g[0] + (long long)I.n[0] * (g[1] + (long long)I.n[1] * g[2]);
        const int2 cl = cls[3 * j + i];
        const bool two = cl.y != -1;
        const int sl = (cl.x == obj) ? 0 : (cl.y == obj) ? 1 : -1;
        const bool use = active && sl >= 0 && (two || rigid_mass);  // uniform over a run
        float v[11];
        node_terms(T, w, gx, gy, gz, v);
        v[7] = gx * V;
        v[8] = gy * V;
        v[9] = gz * V;
        v[10] = w * T.m;
#pragma unroll
        for (int f = 0; f < 11; f++) v[f] = use ? v[f] : 0.0f;
#pragma unroll
        for (int r = 0, off = 1; r < 6; r++, off <<= 1)
#pragma unroll
          for (int f = 0; f < 11; f++) {
            const float o = __shfl_down(v[f], off, 64);
            v[f] += same[r] ? o : 0.0f;
          }
        if (!(use && head)) continue;
        int c[3];
        const long long ci = canon_index(I, g, c);
        float *ab = d.acc2 + ci * MB_ACC2_W + sl * 8;
        if (rigid_mass) unsafeAtomicAdd(&ab[7], v[10]);
        if (!two) continue;
        for (int f = 0; f < 7; f++) unsafeAtomicAdd(&ab[f], v[f]);
        float *nb = d.nrm + gi * 6 + sl * 3;
        unsafeAtomicAdd(&nb[0], v[7]);
        unsafeAtomicAdd(&nb[1], v[8]);
        unsafeAtomicAdd(&nb[2], v[9]);
      }
  }
}


// Two-body nodal solve on the active canonical bricks: centre-of-mass and body sums of the
// periodic class (P2G folded them onto the canonical node), then per actual node its own
// claims and normals (Grid.cpp:9-245), BCs on the centre of mass and both bodies
// (Grid.cpp:967-1031). VEL gets the centre-of-mass velocity p/m, its acceleration and m;
// VELB the body velocities and accelerations G2P reads.
__device__ __forceinline__ void mb_grid_brick(const DevPtrs &d, int gb, int lane, int step, int epoch,
                                              const int box[6]) {
  const InstDesc &I = d.inst[d.brick_inst[gb]];
  const int lb = gb - I.brick_base;
  const int bx = lb % I.nb[0], by = (lb / I.nb[0]) % I.nb[1], bz = lb / (I.nb[0] * I.nb[1]);
  const int c0 = bx * BRICK + (lane & 3), c1 = by * BRICK + ((lane >> 2) & 3), c2 = bz * BRICK + (lane >> 4);
  if (c0 < I.cper[0] && c1 < I.cper[1] && c2 < I.cper[2]) {
    const long long ci = I.node_base + c0 + (long long)I.n[0] * (c1 + (long long)I.n[1] * c2);
    // centre-of-mass sums: tile partials of the periodic class and the direct scatter
    float a7[7] = {0, 0, 0, 0, 0, 0, 0}, ab[MB_ACC2_W];
    for (int mz = 0; mz < 2; mz++) {
      if (mz && c2 >= 3) break;
      const int a2 = c2 + mz * I.cper[2];
      if (!in_tile_box<TILE>(a2, box[2], box[5])) continue;
      for (int my = 0; my < 2; my++) {
        if (my && c1 >= 3) break;
        const int a1 = c1 + my * I.cper[1];
        if (!in_tile_box<TILE>(a1, box[1], box[4])) continue;
        for (int mx = 0; mx < 2; mx++) {
          if (mx && c0 >= 3) break;
          const int a0 = c0 + mx * I.cper[0];
          if (!in_tile_box<TILE>(a0, box[0], box[3])) continue;
          gather_partials(d, I, epoch, a0, a1, a2, a7);
        }
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
    for (int mz = 0; mz < 2; mz++) {
      if (mz && c2 >= 3) break;
      const int n2 = c2 + mz * I.cper[2];
      for (int my = 0; my < 2; my++) {
        if (my && c1 >= 3) break;
        const int n1 = c1 + my * I.cper[1];
        for (int mx = 0; mx < 2; mx++) {
          if (mx && c0 >= 3) break;
          const int n0 = c0 + mx * I.cper[0];
          const long long ni = I.node_base + n0 + (long long)I.n[0] * (n1 + (long long)I.n[1] * n2);
          mb_solve_node(d, I, ni, m, a7, ab, vc, ac, dt, kick);
        }
      }
    }
  }
  if (lane == 0) d.brick_flag[gb] = 0;
}

// Contact grid update, persistent over the list of flagged bricks (one wave per brick).
__global__ __launch_bounds__(WG) void k_mb_grid(DevPtrs d) {
  // the particle kernel that follows walks the active tile list from the start
  if (blockIdx.x == 0 && threadIdx.x == 0) d.ctrl->next = 0;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int n = d.ctrl->active_bricks, step = d.ctrl->step, epoch = d.ctrl->epoch;
  int box[6];
  tile_box(d, epoch, box);
  for (int k = blockIdx.x * (WG / 64) + wave; k < n; k += gridDim.x * (WG / 64))
    mb_grid_brick(d, d.brick_list[k], lane, step, epoch, box);
}

// FP64 audit: per block sums of mass, momentum, kinetic energy, a stress-energy proxy and
// non-finite count.
constexpr int AUD = 8;
__global__ __launch_bounds__(WG) void k_audit(DevPtrs d) {
  __shared__ double sh[AUD][WG];
  const int s = blockIdx.x * WG + threadIdx.x;
  double r[AUD] = {0, 0, 0, 0, 0, 0, 0, 0};
  if (s < d.np) {
    const uint32_t *A = (d.ctrl->parity ? d.pbuf[1] : d.pbuf[0]);
    const long long cap = d.cap;
    const double m = f32(A[F_MASS * cap + s]);
    double v[3], ss = 0;
    bool fin = true;
    for (int a = 0; a < 3; a++) {
      const float x = f32(A[(F_VX + a) * cap + s]);
      fin = fin && isfinite(x);
      v[a] = (double)x + (double)xv_res(A[(F_XV0 + a) * cap + s]);
    }
    for (int a = 0; a < 6; a++) {
      const float x = f32(A[(F_S0 + a) * cap + s]);
      fin = fin && isfinite(x);
      ss += (a < 3 ? 1.0 : 2.0) * (double)x * (double)x;
    }
    const double rho = f32(A[F_RHO * cap + s]);
    r[0] = m;
    r[1] = m * v[0]; r[2] = m * v[1]; r[3] = m * v[2];
    r[4] = 0.5 * m * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    r[5] = 0.5 * (m / rho) * ss;
    r[6] = fin ? 0.0 : 1.0;
  }
  for (int k = 0; k < AUD; k++) sh[k][threadIdx.x] = r[k];
  __syncthreads();
  for (int w = WG / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w)
      for (int k = 0; k < AUD; k++) sh[k][threadIdx.x] += sh[k][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x < AUD) d.audit[(long long)blockIdx.x * AUD + threadIdx.x] = sh[threadIdx.x][0];
}

__global__ __launch_bounds__(WG) void k_audit_final(DevPtrs d, int nblocks) {
  __shared__ double sh[AUD][WG];
  double r[AUD] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int b = threadIdx.x; b < nblocks; b += WG)
    for (int k = 0; k < AUD; k++) r[k] += d.audit[(long long)b * AUD + k];
  for (int k = 0; k < AUD; k++) sh[k][threadIdx.x] = r[k];
  __syncthreads();
  for (int w = WG / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w)
      for (int k = 0; k < AUD; k++) sh[k][threadIdx.x] += sh[k][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x < AUD) d.audit[(long long)nblocks * AUD + threadIdx.x] = sh[threadIdx.x][0];
}

inline hipStream_t S(void *s) { return reinterpret_cast<hipStream_t>(s); }
inline int blocks(long long n, int per) { return (int)((n + per - 1) / per); }

}  // namespace

// ---------------------------------------------------------------------------------------
int scan_items_per_block() { return SCAN_BLOCK; }

void launch_set_ctrl(const DevPtrs &d, int step, double t, int reset, void *st) {
  hipLaunchKernelGGL(k_set_ctrl, dim3(blocks(d.ninst, WG)), dim3(WG), 0, S(st), d, step, t, reset);
}
void launch_keys(const DevPtrs &d, void *st) {
  if (d.np > 0) hipLaunchKernelGGL(k_keys, dim3(blocks(d.np, WG)), dim3(WG), 0, S(st), d);
}
void launch_hist(const DevPtrs &d, void *st) {
  if (d.np > 0) hipLaunchKernelGGL(k_hist, dim3(blocks(d.np, WG)), dim3(WG), 0, S(st), d);
}
void launch_scan_partial(const DevPtrs &d, void *st, int advance, int bricks) {
  hipLaunchKernelGGL(k_scan_partial, dim3(d.nblocks_scan), dim3(WG), 0, S(st), d, advance, bricks);
}
void launch_scan_final(const DevPtrs &d, void *st) {
  hipLaunchKernelGGL(k_scan_final, dim3(d.nblocks_scan), dim3(WG), 0, S(st), d);
}
void launch_place(const DevPtrs &d, void *st) {
  if (d.np > 0) hipLaunchKernelGGL(k_place, dim3(blocks(d.np, WG)), dim3(WG), 0, S(st), d);
}
// Smallest built model-set specialization covering the context's mask.
template <bool U, int W>
void launch_fused_w(const DevPtrs &d, hipStream_t st) {
  const int m = U ? d.kmask : (d.kmask & KM_ALL);  // the P2G of an upload runs no constitutive model
  const dim3 g(d.persist_wgs), b(WG);
  if constexpr (U) {
    if (m & (KM_DPLOOP | KM_DPNN)) {
      // other Drucker-Prager multipliers; the network's LDS leaves two workgroups per CU
      // (the default three per CU scaled to two)
      const dim3 gn(std::max(8, (d.persist_wgs * 2 / 3 + 7) / 8 * 8));
      if (m == (KM_DP | KM_DPLOOP)) hipLaunchKernelGGL((k_g2p2g<true, KM_DP | KM_DPLOOP, false, W>), g, b, 0, st, d);
      else if (m == (KM_DP | KM_DPNN)) hipLaunchKernelGGL((k_g2p2g<true, KM_DP | KM_DPNN, false, 2>), gn, b, 0, st, d);
      else hipLaunchKernelGGL((k_g2p2g<true, KM_ALL | KM_DPLOOP | KM_DPNN, false, 2>), gn, b, 0, st, d);
      return;
    }
  }
  if ((m & ~KM_DP) == 0) hipLaunchKernelGGL((k_g2p2g<U, KM_DP, false, W>), g, b, 0, st, d);
  else if ((m & ~KM_ELASTIC) == 0) hipLaunchKernelGGL((k_g2p2g<U, KM_ELASTIC, false, W>), g, b, 0, st, d);
  else if ((m & ~KM_PERFECT) == 0) hipLaunchKernelGGL((k_g2p2g<U, KM_PERFECT, false, W>), g, b, 0, st, d);
  else if ((m & ~(KM_ELASTIC | KM_GK)) == 0) hipLaunchKernelGGL((k_g2p2g<U, KM_ELASTIC | KM_GK, false, W>), g, b, 0, st, d);
  else hipLaunchKernelGGL((k_g2p2g<U, KM_ALL, false, W>), g, b, 0, st, d);
}
// the P2G of an upload runs once: only the stepping kernel has the 2-wave build
template <bool U>
void launch_fused(const DevPtrs &d, hipStream_t st) {
  if (U && d.fused_waves == 2) launch_fused_w<U, 2>(d, st);
  else launch_fused_w<U, 3>(d, st);
}

void launch_p2g_only(const DevPtrs &d, void *st) { launch_fused<false>(d, S(st)); }
// persistent: grid_factor workgroups per persistent particle workgroup at most
static int grid_blocks(const DevPtrs &d) {
  return std::max(1, std::min(blocks(d.nbricks, WG / 64), std::max(1, d.grid_factor) * d.persist_wgs));
}
void launch_grid(const DevPtrs &d, void *st) {
  hipLaunchKernelGGL(k_grid, dim3(grid_blocks(d)), dim3(WG), 0, S(st), d);
}
void launch_scan_final_grid(const DevPtrs &d, void *st) {
  const int ng = grid_blocks(d);
  hipLaunchKernelGGL(k_scan_final_grid, dim3(ng + d.nblocks_scan), dim3(WG), 0, S(st), d, ng);
}
void launch_brick_list(const DevPtrs &d, void *st) {
  hipLaunchKernelGGL(k_brick_list_reset, dim3(1), dim3(1), 0, S(st), d);
  hipLaunchKernelGGL(k_brick_list, dim3(blocks(d.nbricks, WG)), dim3(WG), 0, S(st), d);
}
void launch_g2p2g(const DevPtrs &d, void *st) { launch_fused<true>(d, S(st)); }
void launch_audit(const DevPtrs &d, void *st) {
  const int nb = blocks(d.np, WG);
  hipLaunchKernelGGL(k_audit, dim3(nb), dim3(WG), 0, S(st), d);
  hipLaunchKernelGGL(k_audit_final, dim3(1), dim3(WG), 0, S(st), d, nb);
}
void launch_mb_grid(const DevPtrs &d, void *st) {
  const int nb = std::max(1, std::min(blocks(d.nbricks, WG / 64), std::max(1, d.grid_factor) * d.persist_wgs));
  hipLaunchKernelGGL(k_mb_grid, dim3(nb), dim3(WG), 0, S(st), d);
}
void launch_mb_g2p2g(const DevPtrs &d, void *st) {
  // two resident workgroups per CU (LDS), in the proportion of the single-body grid
  const int m = d.kmask;
  const dim3 g(std::max(8, (d.persist_wgs * 2 / 3 + 7) / 8 * 8)), b(WG);
  const hipStream_t s = S(st);
  if ((m & ~KM_DP) == 0) hipLaunchKernelGGL((k_g2p2g<true, KM_DP, true>), g, b, 0, s, d);
  else if ((m & ~KM_ELASTIC) == 0) hipLaunchKernelGGL((k_g2p2g<true, KM_ELASTIC, true>), g, b, 0, s, d);
  else if ((m & ~(KM_ELASTIC | KM_GK)) == 0) hipLaunchKernelGGL((k_g2p2g<true, KM_ELASTIC | KM_GK, true>), g, b, 0, s, d);
  else hipLaunchKernelGGL((k_g2p2g<true, KM_ALL, true>), g, b, 0, s, d);
}
void launch_mb_cells(const DevPtrs &d, void *st) {
  hipLaunchKernelGGL(k_mb_cells, dim3(std::max(8, std::min(4096, blocks(d.ntiles, WG / TILE3)))), dim3(WG), 0, S(st), d);
}
void launch_mb_claim(const DevPtrs &d, void *st) {
  launch_mb_cells(d, st);
  hipLaunchKernelGGL(k_mb_claim_tiles, dim3(std::max(8, std::min(8192, d.ntiles))), dim3(WG), 0, S(st), d);
}
void launch_mb_p2g(const DevPtrs &d, void *st) {
  if (d.np > 0) hipLaunchKernelGGL(k_mb_p2g, dim3(blocks(d.np, WG)), dim3(WG), 0, S(st), d);
}
void launch_stable_order(const DevPtrs &d, int *order, void *st) {
  if (d.np > 0) hipLaunchKernelGGL(k_stable_order, dim3(blocks(d.np, WG)), dim3(WG), 0, S(st), d, order);
}

}  // namespace gy
