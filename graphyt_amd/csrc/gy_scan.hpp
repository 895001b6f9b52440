// Binning-scan helpers shared by the 3D and plane step kernels: block-wide minima and
// scans, the tile size classes of the particle kernels' claim order, the flagged-brick
// list, and the second scan kernel's body as a device function (so that a kernel can run
// it beside other work).
#pragma once

#include <hip/hip_runtime.h>

#include "gy_device.hpp"
#include "gy_internal.hpp"

namespace gy {
namespace {

// Block-wide minimum of six values per thread; the result in every thread. Every thread of
// the block calls it.
__device__ __forceinline__ void block_min6(int (&v)[6]) {
  __shared__ int wmin[6][WG / 64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1)
#pragma unroll
    for (int k = 0; k < 6; k++) v[k] = min(v[k], __shfl_xor(v[k], o, 64));
  if (lane == 0)
#pragma unroll
    for (int k = 0; k < 6; k++) wmin[k][w] = v[k];
  __syncthreads();
#pragma unroll
  for (int k = 0; k < 6; k++)
#pragma unroll
    for (int q = 0; q < WG / 64; q++) v[k] = min(v[k], wmin[k][q]);
}

// N independent block-wide exclusive scans in one pass (one shuffle chain per value, one
// pair of barriers for all): v[k] becomes its exclusive prefix, total[k] the block total.
template <int N>
__device__ __forceinline__ void block_scan_n(int (&v)[N], int (&total)[N]) {
  __shared__ int wsum[N][WG / 64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  int x[N];
#pragma unroll
  for (int k = 0; k < N; k++) x[k] = v[k];
#pragma unroll
  for (int o = 1; o < 64; o <<= 1)
#pragma unroll
    for (int k = 0; k < N; k++) {
      const int y = __shfl_up(x[k], o, 64);
      if (lane >= o) x[k] += y;
    }
  if (lane == 63)
#pragma unroll
    for (int k = 0; k < N; k++) wsum[k][w] = x[k];
  __syncthreads();
#pragma unroll
  for (int k = 0; k < N; k++) {
    int before = 0, all = 0;
#pragma unroll
    for (int q = 0; q < WG / 64; q++) {
      const int t = wsum[k][q];
      if (q < w) before += t;
      all += t;
    }
    total[k] = all;
    v[k] = before + x[k] - v[k];
  }
  __syncthreads();
}

// Scan blocks cover SCAN_BLOCK keys = SCAN_BLOCK / TILE3 whole tiles; four consecutive
// threads (16 keys each) hold one tile, so tile occupancy is a 4-lane shuffle sum.
static_assert(TILE3 == 4 * SCAN_ITEMS, "a tile is four scan threads");

// The key count is a multiple of TILE3, so a thread's SCAN_ITEMS keys are all present or
// all absent; they are read and written as four 16-byte vectors.
static_assert(SCAN_ITEMS == 16, "four int4 per scan thread");
__device__ __forceinline__ void load16(const int *p, int c[16]) {
  const int4 *v = reinterpret_cast<const int4 *>(p);
#pragma unroll
  for (int q = 0; q < 4; q++) {
    const int4 x = v[q];
    c[4 * q] = x.x; c[4 * q + 1] = x.y; c[4 * q + 2] = x.z; c[4 * q + 3] = x.w;
  }
}

// Size class of a tile holding n entries, for the particle kernels' claim order (longest
// first): more than two chunks, two, more than half of one, the rest; -1 when empty.
constexpr int NBUCKET = 4;
__device__ __forceinline__ int tile_bucket(int n) {
  if (n <= 0) return -1;
  if (n > 2 * CHUNK_ENTRIES) return 0;
  if (n > CHUNK_ENTRIES) return 1;
  if (n > CHUNK_ENTRIES / 2) return 2;
  return 3;
}

// Append the flagged bricks among gb, gb + stride, ... to the brick list: one atomic per
// wave and round. All lanes of the wave take part (uniform trip count). first >= 0: the
// raw flag byte of brick min(gb, nbricks - 1), loaded by the caller.
__device__ __forceinline__ void append_bricks(const DevPtrs &d, int gb, int stride, int first = -1) {
  const int lane = threadIdx.x & 63;
  const int wave0 = gb - lane;
  for (int w = wave0; w < d.nbricks; w += stride) {
    const int b = w + lane;
    const bool f = b < d.nbricks && ((w == wave0 && first >= 0) ? first != 0 : d.brick_flag[b] != 0);
    const unsigned long long m = __ballot(f);
    if (m == 0ull) continue;
    const int leader = __ffsll((long long)m) - 1;
    int base = 0;
    if (lane == leader) base = atomicAdd(&d.ctrl->active_bricks, __popcll(m));
    base = __shfl(base, leader, 64);
    if (f) d.brick_list[base + __popcll(m & ((1ull << lane) - 1ull))] = b;
  }
}

// The second scan kernel as block bid of nblk; bricks: it also lists the flagged bricks.
__device__ __forceinline__ void scan_final_body(const DevPtrs &d, int bid, int nblk, bool bricks) {
  // The loads that depend on nothing computed here go out first, so that their round trips
  // overlap: this thread's counts, its first brick flag and (block 0) the per-block tile
  // boxes; then the prefixes of the earlier blocks.
  const long long base = (long long)bid * SCAN_BLOCK + threadIdx.x * SCAN_ITEMS;
  const bool have = base < d.nkeys;
  int c[SCAN_ITEMS];
  if (have) load16(d.count + base, c);
  const int gb = bid * WG + threadIdx.x;
  int flag0 = 0;  // compared where used; contexts without a grid (the sorts) have no bricks
  if (bricks && d.nbricks > 0) flag0 = d.brick_flag[min(gb, d.nbricks - 1)];
  const bool boxes = bid == 0 && d.ninst == 1 && d.blockbox;  // uniform
  int box[6] = {TBOX_BIG, TBOX_BIG, TBOX_BIG, TBOX_BIG, TBOX_BIG, TBOX_BIG};
  if (boxes)
    for (int j = threadIdx.x; j < nblk; j += WG)
#pragma unroll
      for (int a = 0; a < 6; a++) box[a] = min(box[a], d.blockbox[6 * j + a]);
  // prefixes of all earlier blocks (particles and active tiles per bucket), and the totals
  int pre = 0, preb[NBUCKET] = {0, 0, 0, 0}, allb[NBUCKET] = {0, 0, 0, 0};
  for (int j = threadIdx.x; j < nblk; j += WG) {
    const int bt = d.blocktiles[j];
#pragma unroll
    for (int b = 0; b < NBUCKET; b++) {
      const int nb = (bt >> (8 * b)) & 0xff;
      if (j < bid) preb[b] += nb;
      allb[b] += nb;
    }
    if (j < bid) pre += d.blocksum[j];
  }
  {
    int v9[9] = {pre, preb[0], preb[1], preb[2], preb[3], allb[0], allb[1], allb[2], allb[3]}, t9[9];
    block_scan_n(v9, t9);
    pre = t9[0];
#pragma unroll
    for (int b = 0; b < NBUCKET; b++) { preb[b] = t9[1 + b]; allb[b] = t9[5 + b]; }
  }
  int s = 0;
#pragma unroll
  for (int j = 0; j < SCAN_ITEMS; j++) {
    if (!have) c[j] = 0;
    s += c[j];
  }
  // exclusive in-block prefixes of the particle count and of each bucket's tiles
  int ts = s + __shfl_xor(s, 1, 64);
  ts += __shfl_xor(ts, 2, 64);
  const int bk = ((threadIdx.x & 3) == 0) ? tile_bucket(ts) : -1;
  int v5[5] = {s, bk == 0, bk == 1, bk == 2, bk == 3}, t5[5];
  block_scan_n(v5, t5);
  const int tot = t5[0];
  int run = pre + v5[0];
  if (have) {
    int4 *ov = reinterpret_cast<int4 *>(d.offsets + base);
    int4 *cv = reinterpret_cast<int4 *>(d.count + base);
#pragma unroll
    for (int q = 0; q < 4; q++) {
      int4 x;
      x.x = run; run += c[4 * q];
      x.y = run; run += c[4 * q + 1];
      x.z = run; run += c[4 * q + 2];
      x.w = run; run += c[4 * q + 3];
      ov[q] = x;
      // counts are left zero for the next step: only cells that held particles need it
      if ((c[4 * q] | c[4 * q + 1] | c[4 * q + 2] | c[4 * q + 3]) != 0) cv[q] = make_int4(0, 0, 0, 0);
    }
  }
  // active tile list, longest first: the buckets in order (most entries first), each in
  // ascending tile order; claim k of the particle kernels takes entry k
  if (bk >= 0) {
    int start = 0, mine = 0;
#pragma unroll
    for (int b = 0; b < NBUCKET; b++) {
      if (b < bk) start += allb[b];
      if (b == bk) mine = preb[b] + v5[1 + b];
    }
    d.active[start + mine] = (int)(base / TILE3);
  }
  if (boxes) {
    // the box of the listed tiles, from the first kernel's per-block boxes
    block_min6(box);
    if (threadIdx.x == 0) {
      int *bx = d.ctrl->tbox[d.ctrl->tbox_w];
#pragma unroll
      for (int a = 0; a < 6; a++) bx[a] = box[a];
    }
  }
  const int allt = allb[0] + allb[1] + allb[2] + allb[3];
  if (bid == 0 && threadIdx.x == 0) d.ctrl->heavy_tiles = allt;  // the list is in claim order
  if (bid == nblk - 1 && threadIdx.x == 0) {
    d.offsets[d.nkeys] = pre + tot;
    d.ctrl->active_tiles = allt;
  }
  if (bid == 0 && threadIdx.x == 0) d.ctrl->next = 0;
  // list of the bricks the particle kernel flagged, for the next grid update
  if (bricks) append_bricks(d, gb, nblk * WG, flag0);
}

}  // namespace
}  // namespace gy
