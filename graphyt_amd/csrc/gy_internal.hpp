// gy_internal.hpp — device data layout shared by the host context and the HIP kernels.
//
// Particles: one 32-bit-word SoA block per state buffer, field f of slot s at
// base[f * cap + s] (fields listed in PField). Positions are per-axis fixed point:
// q = x / h * 2^F (int32), so a particle's cell and its offset inside the cell are exact
// integers/floats and single-precision round-off never depends on the distance from the
// origin. Position and velocity are the two fields that accumulate a small increment into
// a large value every step, so each carries 16 more bits per axis (F_XV0..2): the position
// fraction below the quantum h/2^F (units of 2^-16 quantum, low half) and the bfloat16 of
// the velocity residual v - v_hi (high half), v = v_hi + residual to about 32 bits. Two state buffers (A, B) alternate every step: the fused particle kernel reads
// slot order of one binning and writes the other in cell-sorted order.
//
// Grid: per node 8 floats (AoS, 32 B). ACC holds P2G sums (m, px, py, pz, fx, fy, fz, -)
// on the periodic-canonical nodes (index i folded to i-(N-3) when i >= N-3, the pair
// identification of Grid.cpp:320-602); VEL holds the solved body-1 velocity and
// acceleration (vx, vy, vz, ax, ay, az, m, -) on every actual node, BCs applied.
//
// Binning: cells are numbered tile-major (tiles of TILE^3 cells, global tile id across
// all instances), so a counting sort by cell also groups particles by tile.
//
// Plane instances (Parameters::is3D false, one node layer): tiles of PT x PT cells (the
// same TILE3 keys per tile), bricks of PB x PB canonical nodes (one wave), and the same
// per-tile partial buffers holding the smaller plane regions.
#pragma once
#include <cstdint>

#include "../../include/graphyt_amd.h"
#include "gy_nn.hpp"

struct gy_material;

namespace gy {

constexpr int TILE = 4;                    // cells per tile edge
constexpr int TILE3 = TILE * TILE * TILE;  // cells (keys) per tile
constexpr int BRICK = 4;                   // canonical nodes per grid-brick edge
constexpr int G2P_R = TILE + 2;            // staged VEL region edge (nodes)
constexpr int P2G_R = TILE + 4;            // LDS accumulator region edge (nodes), halo -1/+1
constexpr int NODE_W = 8;                  // floats per node record
constexpr int P2G_NODES = P2G_R * P2G_R * P2G_R;  // nodes of a tile's P2G region
constexpr int PARTIAL_W = 8 * P2G_NODES;   // floats of one tile's full-region partial sums
constexpr int CORE_R = TILE + 2;           // edge of the region a tile's own cells reach
constexpr int CORE_NODES = CORE_R * CORE_R * CORE_R;
constexpr int CORE_W = 8 * CORE_NODES;     // floats of one tile's core partial sums

// plane path
constexpr int PT = 8;                      // cells per plane tile edge
constexpr int PB = 8;                      // canonical nodes per plane brick edge
constexpr int PG_R = PT + 2;               // staged VEL region edge: nodes [o, o+PT+2)
constexpr int PR = PT + 4;                 // LDS P2G region edge: nodes [o-1, o+PT+3)
constexpr int PR_N = PR * PR;
constexpr int PC = PT + 2;                 // core edge: region coordinates 1..PT+2
constexpr int PC_N = PC * PC;
static_assert(PT * PT == TILE3, "a plane tile has the keys of a 3D tile");
static_assert(PB * PB == 64, "a plane brick is one wave");
static_assert(8 * PC_N <= CORE_W && 8 * PR_N <= PARTIAL_W && PR_N <= P2G_NODES, "plane regions fit the tile buffers");

enum PField : int {
  F_QX = 0, F_QY, F_QZ,                     // fixed-point position (int32 bits)
  F_VX, F_VY, F_VZ,                         // velocity
  F_S0, F_S1, F_S2, F_S3, F_S4, F_S5,       // stress xx yy zz xy yz xz
  F_E0, F_E1, F_E2, F_E3, F_E4, F_E5,       // total strain
  F_P0, F_P1, F_P2, F_P3, F_P4, F_P5,       // plastic strain
  F_EPB,                                    // accumulated plastic strain
  F_RHO,                                    // density
  F_MASS,                                   // mass
  F_DMG,                                    // damage
  F_SMAX,                                   // max principal stress
  F_TAG,                                    // packed: mat(8) | hold(1) | bc type(7) | obj(16)
  F_PID,                                    // caller's particle index
  F_T,                                      // temperature
  F_TG0, F_TG1, F_TG2,                      // temperature gradient of the last G2P
  F_XV0, F_XV1, F_XV2,                      // per axis: position fraction (low 16) | velocity residual bf16 (high 16)
  NF
};

__host__ __device__ inline uint32_t make_tag(int mat, int hold, int bct, int obj) {
  return (uint32_t)(mat & 0xff) | ((uint32_t)(hold & 1) << 8) | ((uint32_t)(bct & 0x7f) << 9) |
         ((uint32_t)(obj & 0xffff) << 16);
}
__host__ __device__ inline int tag_mat(uint32_t t) { return (int)(t & 0xff); }
__host__ __device__ inline int tag_hold(uint32_t t) { return (int)((t >> 8) & 1); }
__host__ __device__ inline int tag_bct(uint32_t t) { return (int)((t >> 9) & 0x7f); }
__host__ __device__ inline int tag_obj(uint32_t t) { return (int)(t >> 16); }

// Per-instance description (one MPM problem; a context holds one or many).
struct InstDesc {
  int n[3];          // nodes per axis (Nodes.cpp:71-73)
  int cper[3];       // canonical extent per axis = n - 3
  int nt[3];         // tiles per axis
  int nb[3];         // canonical bricks per axis
  int tile_base;     // first global tile
  int brick_base;    // first global brick
  long long node_base;  // first global node
  int fbits;         // fixed-point fraction bits F
  int lpq;           // lp (= psep/2) in fixed units
  int oneq;          // h in fixed units (2^F)
  int wrapq[3];      // particle wrap period L-2h in fixed units (TimeIntegration.cpp:81)
  int mat_base;      // first entry in the material table
  int nmat;          // materials of this instance
  int pstart;        // first particle slot of this instance (slots stay grouped)
  int pcount;        // particles of this instance
  float h, inv_h;
  float lp;          // lp / h
  float inv_lp;      // h / lp
  float dt;
  float qscale;      // dt * 2^F / h: fixed-point units per (m/s) of grid velocity
  float gravity[3];
  float damping;
  float inv_2F;      // 2^-F
  float mu;          // contact friction coefficient (Parameters::mu)
  int plane;         // Parameters::is3D false: n[2] = 1, plane tiles and bricks
};

// Per-material constants, derived on the host in FP64 (ConstitutiveModels.cpp,
// DamageModels.cpp) and stored in FP32 except for the Grady-Kipp block.
struct MatDev {
  int model;        // 1 elastic, 2 Newtonian, 3 perfect plastic, 4 Drucker-Prager, 5 turbulent liquid
  int rigid;
  int damage;       // isDamage (also switches the Drucker-Prager criterion)
  int gk;           // Grady-Kipp damage update runs (isDamage && damageModel == 1)
  int apex_guard;
  int carry_plastic;  // plastic strain fields are live for this material
  int carry_damage;   // damage / smax fields are live for this material (Grady-Kipp, model 5's flag)
  int dp_mode;        // Drucker-Prager multipliers: GY_DP_RETURN_CLOSED / _REFERENCE / _LEARNED
  float phi_deg;      // friction angle (degrees), the network's material input
  float K, G, visc, sy;
  float turb;         // turbStress of model 5
  float cp, kt;       // specificHeatCapacity, thermalConductivity
  float c, ecr, eta, xi, eta_bar;
  float c_over_ecr, c_floor;   // c/eps_cr, c/10
  float xi_over_eta, xi_over_etab;
  double gk_K43;      // K + 4G/3
  double gk_alpha13;  // alpha^(1/3)
  double gk_m, gk_k, gk_m3;  // m, k, (m+3)/3
};

// Step control read by every kernel of a step (updated on the device, so a captured
// graph can be replayed for many steps).
struct Ctrl {
  int step;          // reference step number (half kick at step == 1, Grid.cpp:126)
  int parity;        // which state buffer is the source this step
  int status;        // sticky error bits (1 = particle stencil left the grid)
  int oor;           // count of out-of-range particles (clamped)
  int epoch;         // tile partials written with tile_epoch == epoch are this step's P2G
  int active_tiles;  // tiles holding particles in the current binning
  int heavy_tiles;   // leading active-list entries claimed in list order (all: the list is longest first)
  // work counter of the persistent particle kernels, on a line of its own: every
  // workgroup adds to it, and loads of fields sharing its line stall behind those atomics
  alignas(256) int next;
  int active_bricks;  // length of the flagged-brick list (appended with atomics)
  // learned return map: yielding samples through the network, and outside its trained
  // range (closed form), since the upload or the last reset
  unsigned long long nn_count[2];
  // per epoch parity (epoch & 1 of the particle pass that writes the partials): the box of
  // the tiles in that pass's list, i.e. the tiles it flushes, in tile coordinates, as
  // lo[3] and TBOX_BIG - hi[3] (all zero is the whole grid: a fresh block bounds nothing).
  // The binning's first kernel picks the entry of the pass it feeds (tbox_w) and writes
  // per-block boxes, its second kernel writes the entry; the grid kernel reads the entry of
  // the pass whose partials it gathers. Single-instance contexts only.
  alignas(256) int tbox[2][6];
  int tbox_w;
};
constexpr int TBOX_BIG = 1 << 30;

// All device pointers of a context (passed by value to kernels).
struct DevPtrs {
  uint32_t *pbuf[2];     // particle state buffers, NF * cap words each
  int *key[2];           // cell key per slot, per buffer
  int *rank;             // provisional rank within the key (histogram atomic)
  int *tmp;              // key-grouped slot list of the current binning
  int *count;            // per-key histogram (nkeys)
  int *offsets;          // per-key exclusive offsets (nkeys + 1)
  int *blocksum;         // scan partials (particle counts per scan block)
  int *blocktiles;       // scan partials (active tiles per scan block)
  int *blockbox;         // scan partials: box of each scan block's occupied tiles (6 ints;
                         // single-instance contexts, null where no grid kernel reads it)
  int *active;           // active tiles, longest first: entry-count classes, each in ascending tile order
  int *tile_epoch;       // per tile: epoch of its last written core partials
  int *halo_epoch;       // per tile: epoch of its last written halo-shell partials
  uint64_t *nodemask;    // per tile: 8 words, bit l set when region node l has mass
  float *pcore;          // per tile: P2G sums of the core [o, o+TILE+2)^3, [CORE_R^3][8] (m, mv, f, 0)
  float *partial;        // per tile: P2G sums of the LDS region [o-1, o+TILE+3)^3,
                         // [P2G_R^3][8]; only the shell outside the core is written and
                         // read, the nodes with a bit in nodemask
  float *acc;            // NODE_W per global node: P2G of particles outside their tile's region
  float *vel;            // NODE_W per global node
  uint8_t *brick_flag;   // per global canonical brick
  int *brick_list;       // flagged bricks (any order), built after the particle kernel
  uint8_t *bc_type;      // per global node
  float *bc_val;         // 2 per global node
  float *pbc_val;        // 4 per particle (indexed by pid)
  int *tile_inst;        // global tile -> instance
  int *brick_inst;       // global brick -> instance
  InstDesc *inst;
  MatDev *mats;
  Ctrl *ctrl;
  double *t_last;        // per instance: time of the previous step (sine BCs)
  double *t_cur;         // per instance: time of the current step
  double *audit;         // audit partials
  long long cap;         // slots per buffer
  int np;                // live particles (all instances)
  int ninst;
  int nkeys;             // total keys = total tiles * TILE3
  int ntiles;
  int nbricks;
  int nblocks_scan;
  int kmask;             // model-set bits of the fused kernel specialization
  int persist_wgs;       // workgroups of the persistent particle kernel (multiple of 8)
  int fused_waves;       // build of the single-body fused kernel: 3 or 2 waves per SIMD
  int grid_factor;       // grid-kernel workgroups per persistent particle workgroup (1, 2, 4)
  // two-body contact (more than one object or rigid materials); see the MB kernels
  int multibody;
  float *acc2;           // MB_ACC2_W per canonical node: body 1 and 2 (m, p, f, rigid mass)
  float *nrm;            // 6 per node: body 1 and body 2 normal sums (sum gradN * V)
  int4 *claim;           // per node: object of body 1, object of body 2, rigid flags
  float *velb;           // MB_VELB_W per node: v1, a1, v2, a2 after the contact solve
  int4 *cellobj;         // per key: first and second object in the cell, rigid flags
  int *twotile;          // per tile: a node of its core region has two claimed bodies
  // heat transfer (any uploaded temperature nonzero); see gy_thermal.hip
  int thermal;
  float *therm;          // 2 per node: heat content sum(N m cp T), conduction rate
  long long nnodes;      // nodes of all instances
  NetDev nn;             // network of the learned return map (null pointers when none)
};

constexpr int MB_ACC2_W = 16;   // per canonical node: [m p3 f3 mr] for body 1, then body 2
constexpr int MB_VELB_W = 12;   // per node: v1 a1 v2 a2

// Kernel identifiers for per-kernel timing.
enum KernelId : int {
  K_KEYS = 0, K_SCAN_PARTIAL, K_SCAN_FINAL, K_PLACE, K_P2G, K_GRID, K_G2P2G, K_ADVANCE,
  K_AUDIT, K_ORDER, K_MB_CLAIM, K_MB_P2G, K_THERM_P2G, K_THERM_G2P, K_SCAN_GRID, K_COUNT_
};

// Device material constants of a gy_material (defined in gy_ctx.hip).
MatDev material_constants(const gy_material &m);
// A network's parameters in the step's layout (gy_nn.hpp): NN_TILES * 64 fragments,
// NN_NBIAS biases, NN_NORM_N normalisation values (defined in gy_mlp.hip).
int mlp_export(gy_mlp *m, uint2 *frag, float *bias, float *norm);

// Host-side launchers (defined in gy_kernels.hip). Every kernel reads the step number and
// the buffer parity from Ctrl, so a captured sequence replays correctly step after step.
// Order of one step: grid, g2p2g, scan_partial (with the advance bookkeeping), scan_final,
// place.
void launch_set_ctrl(const DevPtrs &d, int step, double t, int reset, void *stream);
void launch_keys(const DevPtrs &d, void *stream);
void launch_hist(const DevPtrs &d, void *stream);  // histogram of key[0] (standalone sort)
// advance: + step bookkeeping; bricks: + the flagged-brick list (for launch_scan_final_grid)
void launch_scan_partial(const DevPtrs &d, void *stream, int advance = 0, int bricks = 0);
void launch_scan_final(const DevPtrs &d, void *stream);
// the second scan kernel beside the next step's grid update (one launch)
void launch_scan_final_grid(const DevPtrs &d, void *stream);
void launch_place(const DevPtrs &d, void *stream);
void launch_p2g_only(const DevPtrs &d, void *stream);
void launch_grid(const DevPtrs &d, void *stream);
void launch_brick_list(const DevPtrs &d, void *stream);  // flagged-brick list outside the step
void launch_g2p2g(const DevPtrs &d, void *stream);
void launch_audit(const DevPtrs &d, void *stream);
void launch_stable_order(const DevPtrs &d, int *order, void *stream);
// two-body contact step (Grid.cpp:9-308, Extrapolation.cpp:166-367): order of one step is
// mb_grid, mb_update, scan_partial (+ advance), scan_final, place, mb_claim, p2g_only, mb_p2g
void launch_mb_grid(const DevPtrs &d, void *stream);
void launch_mb_g2p2g(const DevPtrs &d, void *stream);
void launch_mb_claim(const DevPtrs &d, void *stream);   // cell objects + node body claims
void launch_mb_p2g(const DevPtrs &d, void *stream);
int scan_items_per_block();
// plane path (gy_plane.hip): keys, P2G of the uploaded state, the fused particle kernel and
// the grid kernel of plane instances; scan, place and brick list are shared
void launch_plane_keys(const DevPtrs &d, void *stream);
void launch_plane_p2g_only(const DevPtrs &d, void *stream);
void launch_plane_g2p2g(const DevPtrs &d, void *stream);
void launch_plane_grid(const DevPtrs &d, void *stream);
void launch_plane_scan_final_grid(const DevPtrs &d, void *stream);  // as launch_scan_final_grid
// plane contact step: plane_mb_grid, plane_mb_update, scan_partial (+ advance), scan_final,
// place, mb_cells, plane_mb_claim, plane_p2g_only, plane_mb_p2g
void launch_mb_cells(const DevPtrs &d, void *stream);  // objects per cell (both paths)
void launch_plane_mb_claim(const DevPtrs &d, void *stream);
void launch_plane_mb_p2g(const DevPtrs &d, void *stream);
void launch_plane_mb_grid(const DevPtrs &d, void *stream);
void launch_plane_mb_g2p2g(const DevPtrs &d, void *stream);
// heat transfer (gy_thermal.hip): node sums of the current state (cleared first), and the
// particle gradient, rate and temperature update before the fused particle kernel
void launch_therm_p2g(const DevPtrs &d, void *stream);
void launch_therm_g2p(const DevPtrs &d, void *stream);

}  // namespace gy
