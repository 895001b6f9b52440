// gy_ctx.hip — host side of the C ABI (include/graphyt_amd.h): context, device memory,
// upload/download between the caller's FP64 arrays and the FP32 device state, the
// step sequence and its hipGraph replay.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/graphyt_amd.h"
#include "gy_internal.hpp"

using namespace gy;

namespace {

// bfloat16 of x, round to nearest even, in the high half of a word (the velocity residual
// of the F_XV words)
uint32_t bf16_bits_rn(float x) {
  uint32_t u;
  std::memcpy(&u, &x, 4);
  u += 0x7fffu + ((u >> 16) & 1u);
  return u & 0xffff0000u;
}

const char *kKernelNames[K_COUNT_] = {"keys",  "scan_partial", "scan_final", "place",   "p2g",
                                      "grid",  "g2p2g",        "advance",    "audit",   "order",
                                      "mb_claim", "mb_p2g", "therm_p2g", "therm_g2p", "scan_final_grid"};

struct Timed {
  int id;
  hipEvent_t a, b;
};

// Per-instance host information kept for upload/download conversions.
struct InstHost {
  gy_grid grid;
  gy_params params;
  std::vector<gy_material> mats;
  int n[3];
  double dt;
  int nparticles = 0;
  std::vector<double> pos0;  // caller's reference positions (displacement)
  std::vector<uint8_t> carry_plastic, carry_damage;
};

}  // namespace

struct gy_ctx {
  hipStream_t stream = nullptr;
  int device = 0;
  std::string err;
  std::vector<InstHost> ih;
  std::vector<InstDesc> inst;
  std::vector<MatDev> mats;
  DevPtrs d{};
  long long nnodes = 0;
  bool plane = false;  // Parameters::is3D false for every instance
  bool uploaded = false;
  bool acc_valid = false;
  int host_step = 0;
  // options
  bool use_graph = true;
  bool profile = false;
  hipGraphExec_t graph = nullptr;       // one step (STEP_FULL)
  hipGraphExec_t graph_ahead = nullptr; // one step with the next step's grid update (STEP_AHEAD)
  hipGraphExec_t graph_last = nullptr;  // one step after a STEP_AHEAD (STEP_LAST)
  bool graph_valid = false;
  std::vector<Timed> pending;
  int64_t launches[K_COUNT_] = {0};
  double total_ms[K_COUNT_] = {0};
  std::vector<hipEvent_t> event_pool;
  void *nnbuf = nullptr;  // network of the learned return map: fragments, biases, normalisation
};

namespace {

int fail(gy_ctx *c, int code, const char *fmt, ...) {
  if (c) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    c->err = buf;
  }
  return code;
}

#define HIPCHECK(ctx, call)                                                                 \
  do {                                                                                      \
    hipError_t e_ = (call);                                                                 \
    if (e_ != hipSuccess)                                                                   \
      return fail(ctx, GY_ERR_HIP, "%s failed: %s", #call, hipGetErrorString(e_));          \
  } while (0)

template <class T>
int dalloc(gy_ctx *c, T **p, size_t count) {
  if (*p) { (void)hipFree(*p); *p = nullptr; }
  if (count == 0) count = 1;
  HIPCHECK(c, hipMalloc((void **)p, count * sizeof(T)));
  HIPCHECK(c, hipMemsetAsync(*p, 0, count * sizeof(T), c->stream));
  return GY_OK;
}

void dfree(void *p) {
  if (p) (void)hipFree(p);
}

hipEvent_t get_event(gy_ctx *c) {
  if (!c->event_pool.empty()) {
    hipEvent_t e = c->event_pool.back();
    c->event_pool.pop_back();
    return e;
  }
  hipEvent_t e;
  (void)hipEventCreate(&e);
  return e;
}

// Launch with optional per-kernel event timing.
template <class F>
void timed(gy_ctx *c, int id, F &&launch) {
  if (!c->profile) {
    launch();
    return;
  }
  Timed t{id, get_event(c), get_event(c)};
  (void)hipEventRecord(t.a, c->stream);
  launch();
  (void)hipEventRecord(t.b, c->stream);
  c->pending.push_back(t);
}

void collect_timing(gy_ctx *c) {
  for (auto &t : c->pending) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, t.a, t.b) == hipSuccess) {
      c->launches[t.id] += 1;
      c->total_ms[t.id] += ms;
    }
    c->event_pool.push_back(t.a);
    c->event_pool.push_back(t.b);
  }
  c->pending.clear();
}

int fixed_bits(const int n[3]) {
  int nmax = std::max(n[0], std::max(n[1], n[2])) + 8;
  int b = 0;
  while ((1 << b) < nmax) b++;
  return std::min(23, 31 - b);
}

MatDev make_mat(const gy_material &m) {
  MatDev M{};
  M.model = m.model;
  M.rigid = m.is_rigid != 0;
  M.damage = m.is_damage != 0;
  M.gk = (m.is_damage != 0) && m.damage_model == 1;
  M.apex_guard = m.dp_apex_guard != 0;
  M.dp_mode = m.model == 4 ? m.dp_return : GY_DP_RETURN_CLOSED;
  M.phi_deg = (float)m.phi_deg;
  M.carry_plastic = (m.model == 3 || m.model == 4);
  M.carry_damage = M.gk || m.model == 5;  // model 5 keeps its turbulence flag in the damage field
  M.turb = (float)m.turb_stress;
  M.cp = (float)m.specific_heat;
  M.kt = (float)m.thermal_conductivity;
  M.K = (float)m.bulk_mod;
  M.G = (float)m.shear_mod;
  M.visc = (float)m.viscosity;
  M.sy = (float)m.yield_stress;
  // Drucker-Prager constants, ConstitutiveModels.cpp:246-295 (pi = 3.14159 as there)
  const double phi = m.phi_deg * 3.14159 / 180.0, psi = m.psi_deg * 3.14159 / 180.0;
  const double eta = 6.0 * sin(phi) / (sqrt(3.0) * (3.0 - sin(phi)));
  const double xi = 6.0 * cos(phi) / (sqrt(3.0) * (3.0 - sin(phi)));
  const double etab = (phi == psi) ? eta : 6.0 * sin(psi) / (sqrt(3.0) * (3.0 - sin(psi)));
  M.c = (float)m.cohesion;
  M.ecr = (float)m.critical_strain;
  M.eta = (float)eta;
  M.xi = (float)xi;
  M.eta_bar = (float)etab;
  M.c_over_ecr = (float)(m.cohesion / m.critical_strain);
  M.c_floor = (float)(m.cohesion / 10.0);
  M.xi_over_eta = (float)(xi / eta);
  M.xi_over_etab = (float)(xi / etab);
  // Grady-Kipp constants, DamageModels.cpp:96-106
  const double Cg = m.crack_speed, dm = m.damage_m, dk = m.damage_k;
  M.gk_K43 = m.bulk_mod + 4.0 * m.shear_mod / 3.0;
  const double alpha = (8.0 * 3.14159 * Cg * Cg * Cg * dk) / (((dm + 1.0) * (dm + 2.0) * (dm + 3.0)));
  M.gk_alpha13 = pow(alpha, 1.0 / 3.0);
  M.gk_m = dm;
  M.gk_k = dk;
  M.gk_m3 = (dm + 3.0) / 3.0;
  return M;
}

// Validate one instance and fill its descriptor (grid part).
int describe(gy_ctx *c, const gy_grid *g, const gy_material *mats, int nmat, const gy_params *p,
             InstHost &H, InstDesc &I) {
  if (!g || !mats || !p || nmat <= 0) return fail(c, GY_ERR_INVALID, "null grid/materials/params");
  if (nmat > 16) return fail(c, GY_ERR_UNSUPPORTED, "at most 16 materials per instance");
  const double h = g->cellsize;
  if (!(h > 0)) return fail(c, GY_ERR_INVALID, "cellsize must be positive");
  // plane path: one node layer, Nz = 1 + L[2]/h = 1 (Nodes.cpp:71-73 with L[2] = 0)
  const bool plane = !p->is3D;
  if (plane && g->L[2] != 0.0) return fail(c, GY_ERR_INVALID, "a plane (is3D = false) grid needs L[2] = 0");
  I.plane = plane ? 1 : 0;
  I.n[2] = 1;
  for (int a = 0; a < (plane ? 2 : 3); a++) {
    const double r = g->L[a] / h;
    if (std::fabs(r - std::round(r)) > 1e-9 * std::max(1.0, r))
      return fail(c, GY_ERR_INVALID, "L[%d]/cellsize = %.12g is not an integer (Nodes.cpp:12 vs 71)", a, r);
    I.n[a] = (int)(1 + (g->L[a] / h));
    if (I.n[a] != (int)std::llround(1 + r))
      return fail(c, GY_ERR_INVALID, "L[%d]/cellsize = %.17g truncates to %d nodes (Nodes.cpp:71-73); lengthen L by an ulp",
                  a, r, I.n[a]);
    if (I.n[a] < 6) return fail(c, GY_ERR_INVALID, "need at least 6 nodes per axis for the periodic pairs");
  }
  if (!(p->psep > 0) || p->psep > h) return fail(c, GY_ERR_UNSUPPORTED, "psep must be in (0, cellsize]");
  for (int i = 0; i < nmat; i++) {
    const int md = mats[i].model;
    if (md < 1 || md > 5) return fail(c, GY_ERR_UNSUPPORTED, "material %d: model %d not implemented", i, md);
    const int dr = mats[i].dp_return;
    if (dr < GY_DP_RETURN_CLOSED || dr > GY_DP_RETURN_LEARNED)
      return fail(c, GY_ERR_INVALID, "material %d: dp_return %d is not a return-map mode", i, dr);
    if (dr != GY_DP_RETURN_CLOSED && md != 4)
      return fail(c, GY_ERR_INVALID, "material %d: dp_return applies to the Drucker-Prager model (4) only", i);
    if (dr == GY_DP_RETURN_LEARNED && mats[i].is_damage)
      return fail(c, GY_ERR_UNSUPPORTED, "material %d: the learned return map is built for is_damage = 0", i);
  }
  H.grid = *g;
  H.params = *p;
  H.mats.assign(mats, mats + nmat);
  for (int a = 0; a < 3; a++) H.n[a] = I.n[a];
  // dt: Solver::Solver (Solver.cpp:15-32) unless given
  double maxK = 1, minrho = 1e15;
  for (int i = 0; i < nmat; i++) {
    if (mats[i].bulk_mod > maxK) maxK = mats[i].bulk_mod;
    if (mats[i].density < minrho) minrho = mats[i].density;
  }
  H.dt = p->dt > 0 ? p->dt : p->cfl * (h / sqrt(maxK / minrho));
  const int F = fixed_bits(I.n);
  I.fbits = F;
  I.oneq = 1 << F;
  I.lp = (float)(0.5 * p->psep / h);
  I.inv_lp = (float)(h / (0.5 * p->psep));
  I.lpq = (int)llround(0.5 * p->psep / h * (double)(1 << F));
  for (int a = 0; a < 3; a++) {
    I.gravity[a] = (float)p->gravity[a];
    if (plane && a == 2) {  // one node layer: no z tiles, bricks, pairs or wrap
      I.cper[a] = 1;
      I.nt[a] = 1;
      I.nb[a] = 1;
      I.wrapq[a] = 0;
      continue;
    }
    const int tile = plane ? PT : TILE, brick = plane ? PB : BRICK;
    I.cper[a] = I.n[a] - 3;
    I.nt[a] = (I.n[a] - 2 + tile - 1) / tile;
    I.nb[a] = (I.cper[a] + brick - 1) / brick;
    I.wrapq[a] = (int)llround((g->L[a] - 2 * h) / h * (double)(1 << F));
  }
  I.h = (float)h;
  I.inv_h = (float)(1.0 / h);
  I.dt = (float)H.dt;
  I.qscale = (float)(H.dt * (double)(1 << F) / h);
  I.damping = (float)p->damping;
  I.inv_2F = (float)(1.0 / (double)(1 << F));
  I.mu = (float)p->mu;
  I.nmat = nmat;
  return GY_OK;
}

int alloc_grid(gy_ctx *c) {
  // global numbering across instances
  long long node_base = 0;
  int tile_base = 0, brick_base = 0, mat_base = 0;
  c->mats.clear();
  std::vector<int> tile_inst, brick_inst;
  for (size_t i = 0; i < c->inst.size(); i++) {
    InstDesc &I = c->inst[i];
    I.node_base = node_base;
    I.tile_base = tile_base;
    I.brick_base = brick_base;
    I.mat_base = mat_base;
    const int nt = I.nt[0] * I.nt[1] * I.nt[2], nb = I.nb[0] * I.nb[1] * I.nb[2];
    tile_inst.insert(tile_inst.end(), nt, (int)i);
    brick_inst.insert(brick_inst.end(), nb, (int)i);
    node_base += (long long)I.n[0] * I.n[1] * I.n[2];
    tile_base += nt;
    brick_base += nb;
    for (auto &m : c->ih[i].mats) c->mats.push_back(make_mat(m));
    mat_base += (int)c->ih[i].mats.size();
  }
  c->nnodes = node_base;
  DevPtrs &d = c->d;
  d.nnodes = node_base;
  d.ninst = (int)c->inst.size();
  d.ntiles = tile_base;
  d.nbricks = brick_base;
  d.nkeys = tile_base * TILE3;
  d.nblocks_scan = (d.nkeys + scan_items_per_block() - 1) / scan_items_per_block();
  int rc;
  if ((rc = dalloc(c, &d.acc, (size_t)node_base * NODE_W))) return rc;
  if ((rc = dalloc(c, &d.vel, (size_t)node_base * NODE_W))) return rc;
  if ((rc = dalloc(c, &d.bc_type, (size_t)node_base))) return rc;
  if ((rc = dalloc(c, &d.bc_val, (size_t)node_base * 2))) return rc;
  if ((rc = dalloc(c, &d.brick_flag, (size_t)brick_base))) return rc;
  if ((rc = dalloc(c, &d.brick_list, (size_t)brick_base))) return rc;
  if ((rc = dalloc(c, &d.tile_inst, (size_t)tile_base))) return rc;
  if ((rc = dalloc(c, &d.brick_inst, (size_t)brick_base))) return rc;
  if ((rc = dalloc(c, &d.count, (size_t)d.nkeys))) return rc;
  if ((rc = dalloc(c, &d.offsets, (size_t)d.nkeys + 1))) return rc;
  if ((rc = dalloc(c, &d.blocksum, (size_t)d.nblocks_scan))) return rc;
  if ((rc = dalloc(c, &d.blocktiles, (size_t)d.nblocks_scan))) return rc;
  if ((rc = dalloc(c, &d.blockbox, (size_t)d.nblocks_scan * 6))) return rc;
  if ((rc = dalloc(c, &d.active, (size_t)tile_base + 1))) return rc;
  if ((rc = dalloc(c, &d.tile_epoch, (size_t)tile_base))) return rc;
  if ((rc = dalloc(c, &d.halo_epoch, (size_t)tile_base))) return rc;
  if ((rc = dalloc(c, &d.nodemask, (size_t)tile_base * (P2G_NODES / 64)))) return rc;
  if ((rc = dalloc(c, &d.pcore, (size_t)tile_base * CORE_W))) return rc;
  if ((rc = dalloc(c, &d.partial, (size_t)tile_base * PARTIAL_W))) return rc;
  {
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device) != hipSuccess || cus <= 0) cus = 256;
    d.persist_wgs = 8 * ((cus * 3 + 7) / 8);  // three resident 256-thread workgroups per CU
    d.fused_waves = 3;
    d.grid_factor = 2;
  }
  if ((rc = dalloc(c, &d.inst, c->inst.size()))) return rc;
  if ((rc = dalloc(c, &d.mats, c->mats.size()))) return rc;
  if ((rc = dalloc(c, &d.ctrl, 1))) return rc;
  if ((rc = dalloc(c, &d.t_last, c->inst.size()))) return rc;
  if ((rc = dalloc(c, &d.t_cur, c->inst.size()))) return rc;
  HIPCHECK(c, hipMemcpyAsync(d.tile_inst, tile_inst.data(), tile_inst.size() * sizeof(int), hipMemcpyHostToDevice, c->stream));
  HIPCHECK(c, hipMemcpyAsync(d.brick_inst, brick_inst.data(), brick_inst.size() * sizeof(int), hipMemcpyHostToDevice, c->stream));
  HIPCHECK(c, hipMemcpyAsync(d.mats, c->mats.data(), c->mats.size() * sizeof(MatDev), hipMemcpyHostToDevice, c->stream));
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  return GY_OK;
}

int push_inst(gy_ctx *c) {
  HIPCHECK(c, hipMemcpyAsync(c->d.inst, c->inst.data(), c->inst.size() * sizeof(InstDesc), hipMemcpyHostToDevice, c->stream));
  HIPCHECK(c, hipMemcpyAsync(c->d.mats, c->mats.data(), c->mats.size() * sizeof(MatDev), hipMemcpyHostToDevice, c->stream));
  return GY_OK;
}

int check_status(gy_ctx *c) {
  Ctrl h{};
  HIPCHECK(c, hipMemcpy(&h, c->d.ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost));
  if (h.status & 2) return fail(c, GY_ERR_STATE, "internal binning inconsistency (status %d)", h.status);
  if (h.status & 1)
    return fail(c, GY_ERR_RANGE, "%d particle(s) had a GIMP stencil outside the grid (clamped)", h.oor);
  return GY_OK;
}

// How a step is enqueued. STEP_FULL: grid update, particle kernel, binning. Within a call
// of several steps of a single-body context (3D or plane) without heat transfer, the grid update of
// the next step runs beside the second scan kernel (they share no data): the call starts
// with the first step's grid update, then come STEP_AHEAD steps (particle kernel, binning
// with the next step's grid update) and one STEP_LAST step (particle kernel, binning), so
// that between calls the state is the same as after STEP_FULL steps. Every mode runs the
// same kernels on the same inputs.
enum StepMode { STEP_FULL = 0, STEP_AHEAD, STEP_LAST };

bool grid_ahead_ok(const gy_ctx *c) { return !c->d.multibody && !c->d.thermal; }

// One step after the binning of the current buffer is in place.
void enqueue_step(gy_ctx *c, StepMode mode = STEP_FULL) {
  void *s = c->stream;
  const bool th = c->d.thermal != 0;
  if (c->plane && c->d.multibody) {
    timed(c, K_GRID, [&] { launch_plane_mb_grid(c->d, s); });
    if (th) timed(c, K_THERM_G2P, [&] { launch_therm_g2p(c->d, s); });
    timed(c, K_G2P2G, [&] { launch_plane_mb_g2p2g(c->d, s); });  // with the centre-of-mass P2G
    timed(c, K_SCAN_PARTIAL, [&] { launch_scan_partial(c->d, s, 1); });
    timed(c, K_SCAN_FINAL, [&] { launch_scan_final(c->d, s); });
    timed(c, K_PLACE, [&] { launch_place(c->d, s); });
    timed(c, K_MB_CLAIM, [&] { launch_mb_cells(c->d, s); launch_plane_mb_claim(c->d, s); });
    timed(c, K_MB_P2G, [&] { launch_plane_mb_p2g(c->d, s); });
    if (th) timed(c, K_THERM_P2G, [&] { launch_therm_p2g(c->d, s); });
    return;
  }
  if (c->plane && mode != STEP_FULL && grid_ahead_ok(c)) {  // as the 3D path below
    const bool ahead = mode == STEP_AHEAD;
    timed(c, K_G2P2G, [&] { launch_plane_g2p2g(c->d, s); });
    timed(c, K_SCAN_PARTIAL, [&] { launch_scan_partial(c->d, s, 1, ahead ? 1 : 0); });
    if (ahead) timed(c, K_SCAN_GRID, [&] { launch_plane_scan_final_grid(c->d, s); });
    else timed(c, K_SCAN_FINAL, [&] { launch_scan_final(c->d, s); });
    timed(c, K_PLACE, [&] { launch_place(c->d, s); });
    return;
  }
  if (c->plane) {
    timed(c, K_GRID, [&] { launch_plane_grid(c->d, s); });
    if (th) timed(c, K_THERM_G2P, [&] { launch_therm_g2p(c->d, s); });
    timed(c, K_G2P2G, [&] { launch_plane_g2p2g(c->d, s); });
    timed(c, K_SCAN_PARTIAL, [&] { launch_scan_partial(c->d, s, 1); });
    timed(c, K_SCAN_FINAL, [&] { launch_scan_final(c->d, s); });
    timed(c, K_PLACE, [&] { launch_place(c->d, s); });
    if (th) timed(c, K_THERM_P2G, [&] { launch_therm_p2g(c->d, s); });
    return;
  }
  if (c->d.multibody) {
    timed(c, K_GRID, [&] { launch_mb_grid(c->d, s); });
    if (th) timed(c, K_THERM_G2P, [&] { launch_therm_g2p(c->d, s); });
    timed(c, K_G2P2G, [&] { launch_mb_g2p2g(c->d, s); });  // with the centre-of-mass P2G
    timed(c, K_SCAN_PARTIAL, [&] { launch_scan_partial(c->d, s, 1); });
    timed(c, K_SCAN_FINAL, [&] { launch_scan_final(c->d, s); });
    timed(c, K_PLACE, [&] { launch_place(c->d, s); });
    timed(c, K_MB_CLAIM, [&] { launch_mb_claim(c->d, s); });
    timed(c, K_MB_P2G, [&] { launch_mb_p2g(c->d, s); });
    if (th) timed(c, K_THERM_P2G, [&] { launch_therm_p2g(c->d, s); });
    return;
  }
  if (mode != STEP_FULL && grid_ahead_ok(c)) {
    // the grid update of this step has run (in the previous step's second scan kernel, or
    // on its own at the start of the call)
    const bool ahead = mode == STEP_AHEAD;
    timed(c, K_G2P2G, [&] { launch_g2p2g(c->d, s); });
    timed(c, K_SCAN_PARTIAL, [&] { launch_scan_partial(c->d, s, 1, ahead ? 1 : 0); });
    if (ahead) timed(c, K_SCAN_GRID, [&] { launch_scan_final_grid(c->d, s); });
    else timed(c, K_SCAN_FINAL, [&] { launch_scan_final(c->d, s); });
    timed(c, K_PLACE, [&] { launch_place(c->d, s); });
    return;
  }
  timed(c, K_GRID, [&] { launch_grid(c->d, s); });
  if (th) timed(c, K_THERM_G2P, [&] { launch_therm_g2p(c->d, s); });
  timed(c, K_G2P2G, [&] { launch_g2p2g(c->d, s); });
  timed(c, K_SCAN_PARTIAL, [&] { launch_scan_partial(c->d, s, 1); });  // with the step bookkeeping
  timed(c, K_SCAN_FINAL, [&] { launch_scan_final(c->d, s); });
  timed(c, K_PLACE, [&] { launch_place(c->d, s); });
  if (th) timed(c, K_THERM_P2G, [&] { launch_therm_p2g(c->d, s); });
}

// First step after an upload: bin the uploaded state and scatter it to the grid.
void enqueue_init(gy_ctx *c) {
  void *s = c->stream;
  if (c->plane) {
    timed(c, K_KEYS, [&] { launch_plane_keys(c->d, s); });
    timed(c, K_SCAN_PARTIAL, [&] { launch_scan_partial(c->d, s); });
    timed(c, K_SCAN_FINAL, [&] { launch_scan_final(c->d, s); });
    timed(c, K_PLACE, [&] { launch_place(c->d, s); });
    if (c->d.multibody) {
      timed(c, K_MB_CLAIM, [&] { launch_mb_cells(c->d, s); launch_plane_mb_claim(c->d, s); });
      timed(c, K_P2G, [&] { launch_plane_p2g_only(c->d, s); });
      timed(c, K_MB_P2G, [&] { launch_plane_mb_p2g(c->d, s); });
      launch_brick_list(c->d, s);
      if (c->d.thermal) timed(c, K_THERM_P2G, [&] { launch_therm_p2g(c->d, s); });
      return;
    }
    timed(c, K_P2G, [&] { launch_plane_p2g_only(c->d, s); });
    launch_brick_list(c->d, s);
    if (c->d.thermal) timed(c, K_THERM_P2G, [&] { launch_therm_p2g(c->d, s); });
    return;
  }
  timed(c, K_KEYS, [&] { launch_keys(c->d, s); });
  timed(c, K_SCAN_PARTIAL, [&] { launch_scan_partial(c->d, s); });
  timed(c, K_SCAN_FINAL, [&] { launch_scan_final(c->d, s); });
  timed(c, K_PLACE, [&] { launch_place(c->d, s); });
  if (c->d.multibody) {
    timed(c, K_MB_CLAIM, [&] { launch_mb_claim(c->d, s); });
    timed(c, K_P2G, [&] { launch_p2g_only(c->d, s); });
    timed(c, K_MB_P2G, [&] { launch_mb_p2g(c->d, s); });
    launch_brick_list(c->d, s);
    if (c->d.thermal) timed(c, K_THERM_P2G, [&] { launch_therm_p2g(c->d, s); });
    return;
  }
  timed(c, K_P2G, [&] { launch_p2g_only(c->d, s); });
  launch_brick_list(c->d, s);  // within a step the scan builds it
  if (c->d.thermal) timed(c, K_THERM_P2G, [&] { launch_therm_p2g(c->d, s); });
}

void drop_graph(gy_ctx *c) {
  for (hipGraphExec_t *g : {&c->graph, &c->graph_ahead, &c->graph_last}) {
    if (*g) (void)hipGraphExecDestroy(*g);
    *g = nullptr;
  }
  c->graph_valid = false;
}

int capture_step(gy_ctx *c, StepMode mode, hipGraphExec_t *out) {
  hipGraph_t g;
  HIPCHECK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  const bool prof = c->profile;
  c->profile = false;
  enqueue_step(c, mode);
  c->profile = prof;
  HIPCHECK(c, hipStreamEndCapture(c->stream, &g));
  hipError_t e = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) return fail(c, GY_ERR_HIP, "hipGraphInstantiate: %s", hipGetErrorString(e));
  return GY_OK;
}

int build_graph(gy_ctx *c) {
  drop_graph(c);
  int rc = capture_step(c, STEP_FULL, &c->graph);
  if (!rc && grid_ahead_ok(c)) rc = capture_step(c, STEP_AHEAD, &c->graph_ahead);
  if (!rc && grid_ahead_ok(c)) rc = capture_step(c, STEP_LAST, &c->graph_last);
  if (rc) return rc;
  c->graph_valid = true;
  return GY_OK;
}

}  // namespace

MatDev gy::material_constants(const gy_material &m) { return make_mat(m); }

// ---------------------------------------------------------------------------------------
extern "C" {

const char *gy_version(void) { return "graphyt_amd 0.1 (gfx950)"; }

const char *gy_last_error(const gy_ctx *c) { return c ? c->err.c_str() : "null context"; }

static int create_impl(gy_ctx **out, int32_t ninst, const gy_grid *grids, const gy_material *mats,
                       const int32_t *nmat, const gy_params *params) {
  if (!out) return GY_ERR_INVALID;
  *out = nullptr;
  if (ninst <= 0 || !grids || !mats || !nmat || !params) return GY_ERR_INVALID;
  gy_ctx *c = new gy_ctx();
  int rc;
  if (hipGetDevice(&c->device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return GY_ERR_HIP;
  }
  *out = c;  // kept on failure so gy_last_error can report
  c->ih.resize(ninst);
  c->inst.resize(ninst);
  int moff = 0;
  for (int i = 0; i < ninst; i++) {
    memset(&c->inst[i], 0, sizeof(InstDesc));
    if ((rc = describe(c, &grids[i], mats + moff, nmat[i], &params[i], c->ih[i], c->inst[i]))) {
      if (ninst > 1) c->err = "instance " + std::to_string(i) + ": " + c->err;
      return rc;
    }
    moff += nmat[i];
    if (c->inst[i].plane != c->inst[0].plane)
      return fail(c, GY_ERR_UNSUPPORTED, "instance %d: a context holds plane (is3D = false) or 3D instances, not both", i);
  }
  c->plane = c->inst[0].plane != 0;
  return alloc_grid(c);
}

int gy_create(gy_ctx **out, const gy_grid *grid, const gy_material *mats, int32_t nmat, const gy_params *params) {
  return create_impl(out, 1, grid, mats, &nmat, params);
}

int gy_create_batch(gy_ctx **out, int32_t ninst, const gy_grid *grids, const gy_material *mats,
                    const int32_t *nmat, const gy_params *params) {
  return create_impl(out, ninst, grids, mats, nmat, params);
}

int gy_instance_count(const gy_ctx *c) { return c ? (int)c->inst.size() : 0; }

static int node_bc_impl(gy_ctx *c, long long base, int32_t n_nodes, const int32_t *type, const double *val) {
  std::vector<uint8_t> t(n_nodes);
  std::vector<float> v(2 * (size_t)n_nodes, 0.f);
  for (int i = 0; i < n_nodes; i++) {
    if (type[i] < 0 || type[i] > GY_BC_MP_SYYSIN) return fail(c, GY_ERR_INVALID, "node %d: bad BC type %d", i, type[i]);
    t[i] = (uint8_t)type[i];
    if (val) {
      v[2 * i] = (float)val[4 * i];
      v[2 * i + 1] = (float)val[4 * i + 1];
    }
  }
  HIPCHECK(c, hipMemcpy(c->d.bc_type + base, t.data(), t.size(), hipMemcpyHostToDevice));
  HIPCHECK(c, hipMemcpy(c->d.bc_val + 2 * base, v.data(), v.size() * sizeof(float), hipMemcpyHostToDevice));
  return GY_OK;
}

int gy_set_node_bc(gy_ctx *c, int32_t n_nodes, const int32_t *type, const double *val) {
  if (!c || !type) return GY_ERR_INVALID;
  if (n_nodes != c->nnodes) return fail(c, GY_ERR_INVALID, "expected %lld node BCs, got %d", c->nnodes, n_nodes);
  return node_bc_impl(c, 0, n_nodes, type, val);
}

int gy_set_node_bc_inst(gy_ctx *c, int32_t inst, int32_t n_nodes, const int32_t *type, const double *val) {
  if (!c || !type) return GY_ERR_INVALID;
  if (inst < 0 || inst >= (int)c->inst.size()) return fail(c, GY_ERR_INVALID, "instance %d out of range", inst);
  const InstDesc &I = c->inst[inst];
  const long long nn = (long long)I.n[0] * I.n[1] * I.n[2];
  if (n_nodes != nn) return fail(c, GY_ERR_INVALID, "instance %d: expected %lld node BCs, got %d", inst, nn, n_nodes);
  return node_bc_impl(c, I.node_base, n_nodes, type, val);
}

static int upload_impl(gy_ctx *c, int32_t ninst, const gy_particles *views) {
  if (!c || !views) return GY_ERR_INVALID;
  if (ninst != (int)c->inst.size())
    return fail(c, GY_ERR_INVALID, "upload of %d instance(s) into a context of %d", ninst, (int)c->inst.size());
  long long total = 0;
  bool multibody = false, thermal = false;
  for (int k = 0; k < ninst; k++) {
    const gy_particles *in = &views[k];
    if (!in->pos || in->n <= 0) return fail(c, GY_ERR_INVALID, "instance %d: upload needs n > 0 and positions", k);
    const InstHost &H = c->ih[k];
    const int nmat = (int)H.mats.size();
    // several objects or rigid materials: the two-body contact step
    const int obj0 = in->obj_id ? in->obj_id[0] : 0;
    for (int i = 0; i < in->n; i++) {
      const int m = in->material ? in->material[i] : 0;
      if (m < 0 || m >= nmat) return fail(c, GY_ERR_INVALID, "instance %d particle %d: material %d out of range", k, i, m);
      if (H.mats[m].is_rigid) multibody = true;
      if (in->obj_id && (in->obj_id[i] < 0 || in->obj_id[i] > 0xffff))
        return fail(c, GY_ERR_INVALID, "instance %d particle %d: object id must be in [0, 65535]", k, i);
      if (in->obj_id && in->obj_id[i] != obj0) multibody = true;
      if (in->bc_type && (in->bc_type[i] < 0 || in->bc_type[i] > 127))
        return fail(c, GY_ERR_INVALID, "instance %d particle %d: bad BC type", k, i);
      if (in->temperature && in->temperature[i] != 0.0) thermal = true;
    }
    total += in->n;
  }
  if (total > 0x7fffffffLL - 1024) return fail(c, GY_ERR_UNSUPPORTED, "too many particles");
  const long long cap = total;
  DevPtrs &d = c->d;
  int rc;
  for (int b = 0; b < 2; b++) {
    if ((rc = dalloc(c, &d.pbuf[b], (size_t)NF * cap))) return rc;
    if ((rc = dalloc(c, &d.key[b], (size_t)cap))) return rc;
  }
  if ((rc = dalloc(c, &d.rank, (size_t)cap))) return rc;
  if ((rc = dalloc(c, &d.tmp, (size_t)cap))) return rc;
  if ((rc = dalloc(c, &d.pbc_val, 4 * (size_t)cap))) return rc;
  if ((rc = dalloc(c, &d.audit, 8 * ((size_t)(cap + 255) / 256 + 1)))) return rc;
  d.cap = cap;
  d.np = (int)total;
  // host staging in the device layout; instance k owns slots [pstart, pstart + n)
  std::vector<uint32_t> buf((size_t)NF * cap, 0u);
  auto F = [&](int f, long long s) -> uint32_t & { return buf[(size_t)f * cap + s]; };
  auto put = [&](int f, long long s, double x) { float y = (float)x; std::memcpy(&F(f, s), &y, 4); };
  std::vector<float> pbc(4 * (size_t)cap, 0.f);
  int kmask = 0;
  long long pstart = 0;
  for (int k = 0; k < ninst; k++) {
    const gy_particles *in = &views[k];
    InstHost &H = c->ih[k];
    InstDesc &I = c->inst[k];
    const int n = in->n;
    const int nmat = (int)H.mats.size();
    const double h = H.grid.cellsize, scale = (double)(1 << I.fbits) / h;
    H.carry_plastic.assign(nmat, 0);
    H.carry_damage.assign(nmat, 0);
    H.pos0.resize(3 * (size_t)n);
    const double psep3 = H.params.psep * H.params.psep * H.params.psep;
    const double psep2 = H.params.psep * H.params.psep;
    for (int i = 0; i < n; i++) {
      const long long s_ = pstart + i;
      for (int a = 0; a < 3; a++) {
        const double x = in->pos[3 * i + a];
        // 48-bit fixed point: q = the quantum count, lo = the fraction below it in 2^-16
        const double Q = std::nearbyint(x * scale * 65536.0);
        if (!(std::fabs(Q) < 2.0e9 * 65536.0)) return fail(c, GY_ERR_RANGE, "instance %d particle %d: position out of the fixed-point range", k, i);
        const long long Qi = (long long)Q;
        F(F_QX + a, s_) = (uint32_t)(int32_t)(Qi >> 16);
        uint32_t xv = (uint32_t)(Qi & 0xffff);
        H.pos0[3 * i + a] = in->pos0 ? in->pos0[3 * i + a] : x;
        if (in->vel) {  // v = v_hi + residual, the residual as a bfloat16
          const double v = in->vel[3 * i + a];
          const float vh = (float)v;
          put(F_VX + a, s_, v);
          xv |= bf16_bits_rn((float)(v - (double)vh));
        }
        F(F_XV0 + a, s_) = xv;
      }
      const int m = in->material ? in->material[i] : 0;
      for (int kk = 0; kk < 6; kk++) {
        if (in->stress) put(F_S0 + kk, s_, in->stress[6 * i + kk]);
        if (in->strain) put(F_E0 + kk, s_, in->strain[6 * i + kk]);
        if (in->plastic_strain) {
          put(F_P0 + kk, s_, in->plastic_strain[6 * i + kk]);
          if (in->plastic_strain[6 * i + kk] != 0.0) H.carry_plastic[m] = 1;
        }
      }
      if (in->acc_plastic_strain) {
        put(F_EPB, s_, in->acc_plastic_strain[i]);
        if (in->acc_plastic_strain[i] != 0.0) H.carry_plastic[m] = 1;
      }
      // Solver::initialize (Solver.cpp:40-67) when density/mass are not given: per unit
      // depth in a plane, where rigid particles keep rho*V (P2G weighs them with unit mass)
      const double rho = in->density ? in->density[i] : H.mats[m].density;
      const double mass = in->mass ? in->mass[i]
                                   : I.plane ? rho * psep2 : (H.mats[m].is_rigid ? 1.0 : rho * psep3);
      put(F_RHO, s_, rho);
      put(F_MASS, s_, mass);
      if (in->damage) { put(F_DMG, s_, in->damage[i]); if (in->damage[i] != 0.0) H.carry_damage[m] = 1; }
      if (in->temperature) put(F_T, s_, in->temperature[i]);
      if (in->temperature_grad)
        for (int a = 0; a < 3; a++) put(F_TG0 + a, s_, in->temperature_grad[3 * i + a]);
      if (in->smax) { put(F_SMAX, s_, in->smax[i]); if (in->smax[i] != 0.0) H.carry_damage[m] = 1; }
      const int bct = in->bc_type ? in->bc_type[i] : 0;
      F(F_TAG, s_) = make_tag(m, in->hold ? (in->hold[i] ? 1 : 0) : 0, bct, in->obj_id ? in->obj_id[i] : 0);
      F(F_PID, s_) = (uint32_t)s_;
      if (in->bc_val)
        for (int kk = 0; kk < 4; kk++) pbc[4 * (size_t)s_ + kk] = (float)in->bc_val[4 * i + kk];
    }
    for (int m = 0; m < nmat; m++) {
      MatDev &M = c->mats[I.mat_base + m];
      const MatDev base = make_mat(H.mats[m]);
      M.carry_plastic = base.carry_plastic || H.carry_plastic[m];
      M.carry_damage = base.carry_damage || H.carry_damage[m];
      H.carry_plastic[m] = (uint8_t)M.carry_plastic;
      H.carry_damage[m] = (uint8_t)M.carry_damage;
    }
    // model-set bits for the fused kernel (elastic 1, Newtonian 2, perfect 4, DP 8, GK 16,
    // particle BCs 32); see k_g2p2g in gy_kernels.hip
    std::vector<uint8_t> used(nmat, 0);
    for (int i = 0; i < n; i++) used[in->material ? in->material[i] : 0] = 1;
    for (int m = 0; m < nmat; m++) {
      if (!used[m]) continue;
      const gy_material &gm = H.mats[m];
      kmask |= gm.model == 1 ? 1 : gm.model == 2 ? 2 : gm.model == 3 ? 4 : gm.model == 4 ? 8 : 64;
      if (gm.is_damage && gm.damage_model == 1) kmask |= 16;
      if (c->mats[I.mat_base + m].carry_damage && !(gm.is_damage && gm.damage_model == 1)) kmask |= 16;
      // other Drucker-Prager multipliers: the reference's loop 128, the network 256
      if (gm.model == 4 && gm.dp_return == GY_DP_RETURN_REFERENCE) kmask |= 128;
      if (gm.model == 4 && gm.dp_return == GY_DP_RETURN_LEARNED) kmask |= 256;
    }
    for (int i = 0; in->bc_type && i < n; i++)
      if (in->bc_type[i]) { kmask |= 32; break; }
    H.nparticles = n;
    I.pstart = (int)pstart;
    I.pcount = n;
    pstart += n;
  }
  if ((kmask & (128 | 256)) && (c->plane || multibody))
    return fail(c, GY_ERR_UNSUPPORTED, "dp_return %s runs in 3D single-body contexts only",
                (kmask & 256) ? "GY_DP_RETURN_LEARNED" : "GY_DP_RETURN_REFERENCE");
  d.kmask = kmask;
  d.multibody = multibody ? 1 : 0;
  d.thermal = thermal ? 1 : 0;
  if (thermal && (rc = dalloc(c, &d.therm, 2 * (size_t)c->nnodes))) return rc;
  if (multibody) {
    if ((rc = dalloc(c, &d.acc2, (size_t)c->nnodes * MB_ACC2_W))) return rc;
    if ((rc = dalloc(c, &d.nrm, (size_t)c->nnodes * 6))) return rc;
    if ((rc = dalloc(c, &d.claim, (size_t)c->nnodes))) return rc;
    if ((rc = dalloc(c, &d.velb, (size_t)c->nnodes * MB_VELB_W))) return rc;
    if ((rc = dalloc(c, &d.cellobj, (size_t)d.nkeys))) return rc;
    if ((rc = dalloc(c, &d.twotile, (size_t)d.ntiles))) return rc;
  }
  HIPCHECK(c, hipMemcpyAsync(d.pbuf[0], buf.data(), buf.size() * 4, hipMemcpyHostToDevice, c->stream));
  HIPCHECK(c, hipMemcpyAsync(d.pbc_val, pbc.data(), pbc.size() * 4, hipMemcpyHostToDevice, c->stream));
  if ((rc = push_inst(c))) return rc;
  launch_set_ctrl(d, 0, 0.0, 1, c->stream);
  HIPCHECK(c, hipGetLastError());
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  c->uploaded = true;
  c->acc_valid = false;
  drop_graph(c);
  return GY_OK;
}

int gy_upload(gy_ctx *c, const gy_particles *in) {
  if (c && c->inst.size() != 1) return fail(c, GY_ERR_INVALID, "batched context: use gy_upload_batch");
  return upload_impl(c, 1, in);
}

int gy_upload_batch(gy_ctx *c, int32_t ninst, const gy_particles *views) { return upload_impl(c, ninst, views); }

int gy_set_dt(gy_ctx *c, double dt) {
  if (!c || !(dt > 0)) return GY_ERR_INVALID;
  for (size_t i = 0; i < c->inst.size(); i++) {
    c->ih[i].dt = dt;
    InstDesc &I = c->inst[i];
    I.dt = (float)dt;
    I.qscale = (float)(dt * (double)(1 << I.fbits) / c->ih[i].grid.cellsize);
  }
  if (c->d.inst) {
    int rc = push_inst(c);
    if (rc) return rc;
  }
  return GY_OK;
}

double gy_get_dt(const gy_ctx *c) { return c && !c->ih.empty() ? c->ih[0].dt : 0.0; }

double gy_get_dt_inst(const gy_ctx *c, int32_t inst) {
  return (c && inst >= 0 && inst < (int)c->ih.size()) ? c->ih[inst].dt : 0.0;
}

static int iterate_impl(gy_ctx *c, double t0, int32_t step0, int32_t nsteps) {
  if (!c) return GY_ERR_INVALID;
  if (!c->uploaded) return fail(c, GY_ERR_STATE, "no particles uploaded");
  if (nsteps <= 0) return GY_OK;
  if ((c->d.kmask & 256) && !c->d.nn.frag)
    return fail(c, GY_ERR_STATE, "a material uses the learned return map: gy_set_return_network first");
  launch_set_ctrl(c->d, step0, t0, 0, c->stream);
  if (!c->acc_valid) {
    enqueue_init(c);
    c->acc_valid = true;
  }
  const bool graph = c->use_graph && !c->profile && nsteps > 1;
  if (graph && !c->graph_valid) {
    int rc = build_graph(c);
    if (rc) return rc;
  }
  if (nsteps > 1 && grid_ahead_ok(c)) {
    // this step's grid update, then each step's binning beside the next step's grid update
    timed(c, K_GRID, [&] {
      if (c->plane) launch_plane_grid(c->d, c->stream);
      else launch_grid(c->d, c->stream);
    });
    for (int i = 0; i < nsteps; i++) {
      const StepMode m = i + 1 < nsteps ? STEP_AHEAD : STEP_LAST;
      if (graph) HIPCHECK(c, hipGraphLaunch(m == STEP_AHEAD ? c->graph_ahead : c->graph_last, c->stream));
      else enqueue_step(c, m);
    }
  } else if (graph) {
    for (int i = 0; i < nsteps; i++) HIPCHECK(c, hipGraphLaunch(c->graph, c->stream));
  } else {
    for (int i = 0; i < nsteps; i++) enqueue_step(c);
  }
  HIPCHECK(c, hipGetLastError());
  c->host_step = step0 + nsteps;
  return GY_OK;
}

int gy_iterate(gy_ctx *c, double t, int32_t step) { return iterate_impl(c, t, step, 1); }

int gy_iterate_n(gy_ctx *c, double t0, int32_t step0, int32_t nsteps) {
  return iterate_impl(c, t0, step0, nsteps);
}

int gy_synchronize(gy_ctx *c) {
  if (!c) return GY_ERR_INVALID;
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  collect_timing(c);
  if (!c->uploaded) return GY_OK;
  return check_status(c);
}

static int download_impl(gy_ctx *c, int inst, gy_particles *o) {
  if (!c || !o) return GY_ERR_INVALID;
  if (!c->uploaded) return fail(c, GY_ERR_STATE, "no particles uploaded");
  if (inst < 0 || inst >= (int)c->inst.size()) return fail(c, GY_ERR_INVALID, "instance %d out of range", inst);
  int rc = gy_synchronize(c);
  if (rc && rc != GY_ERR_RANGE) return rc;
  Ctrl h{};
  HIPCHECK(c, hipMemcpy(&h, c->d.ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost));
  const long long cap = c->d.cap;
  const InstHost &H = c->ih[inst];
  const InstDesc &I = c->inst[inst];
  const int n = I.pcount;
  if (o->n != n) return fail(c, GY_ERR_INVALID, "download view has n=%d, instance %d has %d", o->n, inst, n);
  // the instance's particles occupy slots [pstart, pstart + pcount) in every binning
  std::vector<uint32_t> buf((size_t)NF * n);
  for (int f = 0; f < NF; f++)
    HIPCHECK(c, hipMemcpy(buf.data() + (size_t)f * n, c->d.pbuf[h.parity] + (size_t)f * cap + I.pstart,
                          (size_t)n * 4, hipMemcpyDeviceToHost));
  const double h_ = H.grid.cellsize, inv = h_ / (double)(1 << I.fbits);
  auto U = [&](int f, int s) { return buf[(size_t)f * n + s]; };
  auto Fv = [&](int f, int s) { float y; uint32_t u = U(f, s); std::memcpy(&y, &u, 4); return (double)y; };
  const int Cx = H.n[0] - 1, Cy = H.n[1] - 1;
  for (int s = 0; s < n; s++) {
    const int i = (int)U(F_PID, s) - I.pstart;
    if (i < 0 || i >= n) return fail(c, GY_ERR_STATE, "corrupt particle id at slot %d", I.pstart + s);
    const uint32_t tag = U(F_TAG, s);
    const int m = tag_mat(tag);
    double x[3];
    for (int a = 0; a < 3; a++)
      x[a] = ((double)(int32_t)U(F_QX + a, s) + (double)(U(F_XV0 + a, s) & 0xffffu) * (1.0 / 65536.0)) * inv;
    for (int a = 0; a < 3; a++) {
      if (o->pos) o->pos[3 * i + a] = x[a];
      if (o->pos0) o->pos0[3 * i + a] = H.pos0[3 * i + a];
      if (o->displacement) o->displacement[3 * i + a] = x[a] - H.pos0[3 * i + a];
      if (o->vel) {
        const uint32_t r = U(F_XV0 + a, s) & 0xffff0000u;
        float vr;
        std::memcpy(&vr, &r, 4);
        o->vel[3 * i + a] = Fv(F_VX + a, s) + (double)vr;
      }
    }
    const bool cp = H.carry_plastic[m], cd = H.carry_damage[m];
    for (int k = 0; k < 6; k++) {
      if (o->stress) o->stress[6 * i + k] = Fv(F_S0 + k, s);
      if (o->strain) o->strain[6 * i + k] = Fv(F_E0 + k, s);
      if (o->plastic_strain) o->plastic_strain[6 * i + k] = cp ? Fv(F_P0 + k, s) : 0.0;
    }
    const double rho = Fv(F_RHO, s), mass = Fv(F_MASS, s);
    if (o->acc_plastic_strain) o->acc_plastic_strain[i] = cp ? Fv(F_EPB, s) : 0.0;
    if (o->density) o->density[i] = rho;
    if (o->mass) o->mass[i] = mass;
    if (o->volume) o->volume[i] = mass / rho;
    if (o->damage) o->damage[i] = cd ? Fv(F_DMG, s) : 0.0;
    if (o->smax) o->smax[i] = cd ? Fv(F_SMAX, s) : 0.0;
    if (o->temperature) o->temperature[i] = c->d.thermal ? Fv(F_T, s) : 0.0;
    if (o->temperature_grad)
      for (int a = 0; a < 3; a++) o->temperature_grad[3 * i + a] = c->d.thermal ? Fv(F_TG0 + a, s) : 0.0;
    if (o->material) o->material[i] = m;
    if (o->obj_id) o->obj_id[i] = tag_obj(tag);
    if (o->hold) o->hold[i] = (uint8_t)tag_hold(tag);
    if (o->bc_type) o->bc_type[i] = tag_bct(tag);
    if (o->cell_id) {  // TimeIntegration.cpp:110-115
      const int ix = (int)((x[0] - 0.25 * h_) / h_), iy = (int)((x[1] - 0.25 * h_) / h_), iz = (int)((x[2] - 0.25 * h_) / h_);
      o->cell_id[i] = ix + iy * Cx + iz * Cx * Cy;
    }
  }
  return rc;
}

int gy_download(gy_ctx *c, gy_particles *o) {
  if (c && c->inst.size() != 1) return fail(c, GY_ERR_INVALID, "batched context: use gy_download_inst");
  return download_impl(c, 0, o);
}

int gy_download_inst(gy_ctx *c, int32_t inst, gy_particles *o) { return download_impl(c, inst, o); }

int gy_download_nodes(gy_ctx *c, int32_t n_nodes, double *mass, double *vel, double *force) {
  if (!c) return GY_ERR_INVALID;
  if (n_nodes != c->nnodes) return fail(c, GY_ERR_INVALID, "expected %lld nodes, got %d", c->nnodes, n_nodes);
  int rc = gy_synchronize(c);
  if (rc && rc != GY_ERR_RANGE) return rc;
  std::vector<float> rec((size_t)n_nodes * NODE_W);
  HIPCHECK(c, hipMemcpy(rec.data(), c->d.vel, rec.size() * sizeof(float), hipMemcpyDeviceToHost));
  for (int i = 0; i < n_nodes; i++) {
    const float *r = &rec[(size_t)i * NODE_W];  // vx vy vz ax ay az m -
    if (mass) mass[i] = r[6];
    for (int a = 0; a < 3; a++) {
      if (vel) vel[3 * i + a] = r[a];
      if (force) force[3 * i + a] = (double)r[6] * (double)r[3 + a];
    }
  }
  return GY_OK;
}

int gy_audit(gy_ctx *c, gy_audit_out *out) {
  if (!c || !out) return GY_ERR_INVALID;
  if (!c->uploaded) return fail(c, GY_ERR_STATE, "no particles uploaded");
  timed(c, K_AUDIT, [&] { launch_audit(c->d, c->stream); });
  int rc = gy_synchronize(c);
  if (rc && rc != GY_ERR_RANGE) return rc;
  const int nb = (c->d.np + 255) / 256;
  double r[8];
  HIPCHECK(c, hipMemcpy(r, c->d.audit + (size_t)nb * 8, sizeof r, hipMemcpyDeviceToHost));
  Ctrl h{};
  HIPCHECK(c, hipMemcpy(&h, c->d.ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost));
  out->mass = r[0];
  out->momentum[0] = r[1];
  out->momentum[1] = r[2];
  out->momentum[2] = r[3];
  out->kinetic_energy = r[4];
  out->strain_energy_proxy = r[5];
  out->nonfinite = (int64_t)llround(r[6]);
  out->out_of_range = h.oor;
  return rc;
}

int gy_debug_binning(gy_ctx *c, int32_t *keys, int32_t *counts, int32_t *offsets, int32_t *order,
                     int64_t *nkeys, int64_t *nslots) {
  if (!c) return GY_ERR_INVALID;
  if (!c->uploaded) return fail(c, GY_ERR_STATE, "no particles uploaded");
  if (!c->acc_valid) {  // bin the uploaded state
    launch_set_ctrl(c->d, c->host_step, 0.0, 0, c->stream);
    enqueue_init(c);
    c->acc_valid = true;
  }
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  const DevPtrs &d = c->d;
  if (nkeys) *nkeys = d.nkeys;
  if (nslots) *nslots = d.np;
  Ctrl h{};
  HIPCHECK(c, hipMemcpy(&h, d.ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost));
  if (keys) HIPCHECK(c, hipMemcpy(keys, d.key[h.parity], (size_t)d.np * 4, hipMemcpyDeviceToHost));
  std::vector<int> off((size_t)d.nkeys + 1);
  HIPCHECK(c, hipMemcpy(off.data(), d.offsets, off.size() * 4, hipMemcpyDeviceToHost));
  if (offsets) std::memcpy(offsets, off.data(), (size_t)d.nkeys * 4);
  if (counts)
    for (int k = 0; k < d.nkeys; k++) counts[k] = off[k + 1] - off[k];
  if (order) {
    int *dord = nullptr;
    HIPCHECK(c, hipMalloc(&dord, (size_t)d.np * 4));
    launch_stable_order(d, dord, c->stream);
    HIPCHECK(c, hipMemcpyAsync(order, dord, (size_t)d.np * 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    (void)hipFree(dord);
  }
  return GY_OK;
}

int gy_counting_sort(const int32_t *keys, int64_t n, int64_t nkeys, int32_t *counts, int32_t *offsets, int32_t *order) {
  if (!keys || n <= 0 || nkeys <= 0 || n > 0x7fffffff || nkeys > 0x7ffffff0) return GY_ERR_INVALID;
  for (int64_t i = 0; i < n; i++)
    if (keys[i] < 0 || keys[i] >= nkeys) return GY_ERR_INVALID;
  gy_ctx tmp;
  if (hipStreamCreateWithFlags(&tmp.stream, hipStreamNonBlocking) != hipSuccess) return GY_ERR_HIP;
  DevPtrs &d = tmp.d;
  d.np = (int)n;
  // the scan kernels take whole tiles of keys: pad the key space (padded keys stay empty)
  d.nkeys = (int)((nkeys + TILE3 - 1) / TILE3 * TILE3);
  d.nblocks_scan = (int)((nkeys + scan_items_per_block() - 1) / scan_items_per_block());
  int rc = GY_OK;
  if ((rc = dalloc(&tmp, &d.key[0], (size_t)n)) || (rc = dalloc(&tmp, &d.rank, (size_t)n)) ||
      (rc = dalloc(&tmp, &d.tmp, (size_t)n)) || (rc = dalloc(&tmp, &d.count, (size_t)d.nkeys)) ||
      (rc = dalloc(&tmp, &d.offsets, (size_t)d.nkeys + 1)) || (rc = dalloc(&tmp, &d.blocksum, (size_t)d.nblocks_scan)) ||
      (rc = dalloc(&tmp, &d.blocktiles, (size_t)d.nblocks_scan)) ||
      (rc = dalloc(&tmp, &d.active, (size_t)d.nkeys / TILE3 + 1)) || (rc = dalloc(&tmp, &d.ctrl, 1))) {
    rc = GY_ERR_HIP;
  } else {
    int *dord = nullptr;
    if (hipMalloc(&dord, (size_t)n * 4) != hipSuccess ||
        hipMemcpyAsync(d.key[0], keys, (size_t)n * 4, hipMemcpyHostToDevice, tmp.stream) != hipSuccess) {
      rc = GY_ERR_HIP;
    } else {
      launch_hist(d, tmp.stream);
      launch_scan_partial(d, tmp.stream);
      launch_scan_final(d, tmp.stream);
      launch_place(d, tmp.stream);
      launch_stable_order(d, dord, tmp.stream);
      std::vector<int> off((size_t)nkeys + 1);
      if (hipMemcpyAsync(off.data(), d.offsets, off.size() * 4, hipMemcpyDeviceToHost, tmp.stream) != hipSuccess ||
          (order && hipMemcpyAsync(order, dord, (size_t)n * 4, hipMemcpyDeviceToHost, tmp.stream) != hipSuccess) ||
          hipStreamSynchronize(tmp.stream) != hipSuccess || hipGetLastError() != hipSuccess) {
        rc = GY_ERR_HIP;
      } else {
        if (offsets) std::memcpy(offsets, off.data(), (size_t)nkeys * 4);
        if (counts)
          for (int64_t k = 0; k < nkeys; k++) counts[k] = off[k + 1] - off[k];
      }
    }
    dfree(dord);
  }
  dfree(d.key[0]); dfree(d.rank); dfree(d.tmp); dfree(d.count); dfree(d.offsets); dfree(d.blocksum); dfree(d.ctrl);
  dfree(d.blocktiles); dfree(d.active);
  (void)hipStreamDestroy(tmp.stream);
  tmp.stream = nullptr;
  return rc;
}

}  // extern "C"

// ---- gy_counting_sort_noalloc: the counting sort without any allocation -------------
// Its scratch is one file-scope device array, `scratch`: every GPU that runs the entry
// point holds its own copy until the process ends. Host code never passes the array's
// name to a kernel (on the host it is only a stand-in for the symbol): the device address
// comes from hipGetSymbolAddress, once per device, and only that address is used.
namespace {
constexpr long long NA_MAX_N = 1LL << 20;      // keys per call
constexpr long long NA_MAX_KEYS = 1LL << 20;   // key space per call
constexpr long long na_align(long long w) { return (w + 63) / 64 * 64; }  // 256-byte segments
constexpr long long NA_KEYS_PAD = (NA_MAX_KEYS + TILE3 - 1) / TILE3 * TILE3;
constexpr int NA_WG = 256;
constexpr long long NA_SCAN_BLOCK = 4096;  // keys per scan block (scan_items_per_block)
constexpr long long NA_BLOCKS = (NA_KEYS_PAD + NA_SCAN_BLOCK - 1) / NA_SCAN_BLOCK;
// layout in 4-byte words: rank[n] | tmp[n] | count[keys] | offsets[keys + 1] | blocksum |
// blocktiles | active tiles | Ctrl
constexpr long long NA_RANK = 0;
constexpr long long NA_TMP = NA_RANK + na_align(NA_MAX_N);
constexpr long long NA_COUNT = NA_TMP + na_align(NA_MAX_N);
constexpr long long NA_OFF = NA_COUNT + na_align(NA_KEYS_PAD);
constexpr long long NA_BSUM = NA_OFF + na_align(NA_KEYS_PAD + 1);
constexpr long long NA_BTILES = NA_BSUM + na_align(NA_BLOCKS);
constexpr long long NA_ACTIVE = NA_BTILES + na_align(NA_BLOCKS);
constexpr long long NA_CTRL = NA_ACTIVE + na_align(NA_KEYS_PAD / TILE3 + 1);
constexpr long long NA_WORDS = NA_CTRL + na_align((long long)(sizeof(Ctrl) + 3) / 4);

__device__ __align__(256) uint32_t scratch[NA_WORDS];

std::mutex g_na_mu;
void *g_na_ptr[256] = {};  // device address of `scratch`, per device ordinal

// zero the histogram and the control block (no memset node: callers may capture this)
__global__ void k_na_clear(uint32_t *base, long long nkeys_pad) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nkeys_pad) base[NA_COUNT + i] = 0u;
  if (i < (long long)((sizeof(Ctrl) + 3) / 4)) base[NA_CTRL + i] = 0u;
}

// the caller's counts and offsets from the scratch offsets
__global__ void k_na_out(const uint32_t *base, long long nkeys, int32_t *counts, int32_t *offsets) {
  const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nkeys) return;
  const int32_t *off = reinterpret_cast<const int32_t *>(base + NA_OFF);
  if (offsets) offsets[k] = off[k];
  if (counts) counts[k] = off[k + 1] - off[k];
}
}  // namespace

extern "C" {

int gy_counting_sort_noalloc(const int32_t *keys, int64_t n, int64_t nkeys, int32_t *counts, int32_t *offsets,
                             int32_t *order, void *stream) {
  if (!keys || n <= 0 || nkeys <= 0) return GY_ERR_INVALID;
  if (n > NA_MAX_N || nkeys > NA_MAX_KEYS) return GY_ERR_INVALID;  // beyond the scratch
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 256) return GY_ERR_HIP;
  uint32_t *base = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_na_mu);
    if (!g_na_ptr[dev]) {
      void *p = nullptr;
      if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return GY_ERR_HIP;
      g_na_ptr[dev] = p;
    }
    base = static_cast<uint32_t *>(g_na_ptr[dev]);
  }
  hipStream_t st = static_cast<hipStream_t>(stream);
  DevPtrs d{};
  d.np = (int)n;
  d.nkeys = (int)((nkeys + TILE3 - 1) / TILE3 * TILE3);
  d.nblocks_scan = (int)((d.nkeys + scan_items_per_block() - 1) / scan_items_per_block());
  if (d.nblocks_scan > NA_BLOCKS) return GY_ERR_INVALID;
  d.key[0] = const_cast<int *>(keys);
  d.rank = reinterpret_cast<int *>(base + NA_RANK);
  d.tmp = reinterpret_cast<int *>(base + NA_TMP);
  d.count = reinterpret_cast<int *>(base + NA_COUNT);
  d.offsets = reinterpret_cast<int *>(base + NA_OFF);
  d.blocksum = reinterpret_cast<int *>(base + NA_BSUM);
  d.blocktiles = reinterpret_cast<int *>(base + NA_BTILES);
  d.active = reinterpret_cast<int *>(base + NA_ACTIVE);
  d.ctrl = reinterpret_cast<Ctrl *>(base + NA_CTRL);
  const long long clear = std::max<long long>(d.nkeys, (long long)((sizeof(Ctrl) + 3) / 4));
  hipLaunchKernelGGL(k_na_clear, dim3((unsigned)((clear + NA_WG - 1) / NA_WG)), dim3(NA_WG), 0, st, base,
                     (long long)d.nkeys);
  launch_hist(d, st);
  launch_scan_partial(d, st);
  launch_scan_final(d, st);
  launch_place(d, st);
  if (order) launch_stable_order(d, order, st);
  if (counts || offsets)
    hipLaunchKernelGGL(k_na_out, dim3((unsigned)((nkeys + NA_WG - 1) / NA_WG)), dim3(NA_WG), 0, st, base,
                       (long long)nkeys, counts, offsets);
  return hipGetLastError() == hipSuccess ? GY_OK : GY_ERR_HIP;
}

}  // extern "C"

// ---- snapshot of the device state between two steps ---------------------------------
// Everything the next step reads that earlier steps wrote: the current particle buffer,
// keys, ranks, the slot list, counts and offsets, scan partials and the tile list, the
// node records (direct-scatter sums, solved velocities), BCs, brick flags and list, the
// control block, the times, the contact and heat buffers, and the tile partials of the
// tiles that wrote them this step (the grid update reads no other tile's partials: their
// epoch stamps are stale). Loading it into a context set up and uploaded the same way
// puts that state back bit for bit.
namespace {
struct SnapHead {
  char magic[8];
  uint32_t version, ninst;
  int64_t np, cap, nnodes, ntiles, nkeys, nbricks, nmats;
  int32_t parity, acc_valid, multibody, thermal, kmask, host_step, ncore, nhalo;
};
struct Seg {
  void *dev;
  size_t bytes;
};
const char kSnapMagic[8] = {'G', 'Y', 'S', 'N', 'A', 'P', '\0', '\1'};

std::vector<Seg> snap_segments(gy_ctx *c, int parity) {
  DevPtrs &d = c->d;
  const size_t cap = (size_t)d.cap, nn = (size_t)c->nnodes, nt = (size_t)d.ntiles, nk = (size_t)d.nkeys,
               nb = (size_t)d.nbricks, ni = (size_t)d.ninst, ns = (size_t)d.nblocks_scan;
  std::vector<Seg> v = {{d.pbuf[parity], 4 * (size_t)NF * cap}, {d.key[0], 4 * cap}, {d.key[1], 4 * cap},
                        {d.rank, 4 * cap}, {d.tmp, 4 * cap}, {d.pbc_val, 16 * cap}, {d.count, 4 * nk},
                        {d.offsets, 4 * (nk + 1)}, {d.blocksum, 4 * ns}, {d.blocktiles, 4 * ns},
                        {d.active, 4 * (nt + 1)}, {d.tile_epoch, 4 * nt}, {d.halo_epoch, 4 * nt},
                        {d.acc, 4 * (size_t)NODE_W * nn}, {d.vel, 4 * (size_t)NODE_W * nn}, {d.bc_type, nn},
                        {d.bc_val, 8 * nn}, {d.brick_flag, nb}, {d.brick_list, 4 * nb}, {d.ctrl, sizeof(Ctrl)},
                        {d.t_last, 8 * ni}, {d.t_cur, 8 * ni}};
  if (d.multibody) {
    v.push_back({d.acc2, 4 * (size_t)MB_ACC2_W * nn});
    v.push_back({d.nrm, 24 * nn});
    v.push_back({d.claim, 16 * nn});
    v.push_back({d.velb, 4 * (size_t)MB_VELB_W * nn});
    v.push_back({d.cellobj, 16 * nk});
    v.push_back({d.twotile, 4 * nt});
  }
  if (d.thermal) v.push_back({d.therm, 8 * nn});
  return v;
}

constexpr size_t CORE_BYTES = 4 * (size_t)CORE_W + 8 * (P2G_NODES / 64);  // core partials + node mask
constexpr size_t HALO_BYTES = 4 * (size_t)PARTIAL_W;

// The tiles whose core (halo-shell) partials are current; the snapshot's size.
int snap_layout(gy_ctx *c, Ctrl &h, std::vector<int> &core, std::vector<int> &halo, size_t &bytes) {
  DevPtrs &d = c->d;
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  HIPCHECK(c, hipMemcpy(&h, d.ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost));
  std::vector<int> te((size_t)d.ntiles), he((size_t)d.ntiles);
  HIPCHECK(c, hipMemcpy(te.data(), d.tile_epoch, te.size() * 4, hipMemcpyDeviceToHost));
  HIPCHECK(c, hipMemcpy(he.data(), d.halo_epoch, he.size() * 4, hipMemcpyDeviceToHost));
  core.clear();
  halo.clear();
  for (int t = 0; t < d.ntiles; t++) {
    if (te[t] == h.epoch) core.push_back(t);
    if (he[t] == h.epoch) halo.push_back(t);
  }
  bytes = sizeof(SnapHead) + c->inst.size() * sizeof(InstDesc) + c->mats.size() * sizeof(MatDev);
  for (const auto &H : c->ih) bytes += sizeof(double) + H.carry_plastic.size() + H.carry_damage.size();
  for (const Seg &g : snap_segments(c, h.parity)) bytes += g.bytes;
  bytes += core.size() * (4 + CORE_BYTES) + halo.size() * (4 + HALO_BYTES);
  return GY_OK;
}
}  // namespace

extern "C" {

int gy_snapshot_size(gy_ctx *c, int64_t *bytes) {
  if (!c || !bytes) return GY_ERR_INVALID;
  if (!c->uploaded) return fail(c, GY_ERR_STATE, "no particles uploaded");
  Ctrl h{};
  std::vector<int> core, halo;
  size_t n = 0;
  int rc = snap_layout(c, h, core, halo, n);
  if (rc) return rc;
  *bytes = (int64_t)n;
  return GY_OK;
}

int gy_snapshot_save(gy_ctx *c, void *buf, int64_t bytes) {
  if (!c || !buf) return GY_ERR_INVALID;
  if (!c->uploaded) return fail(c, GY_ERR_STATE, "no particles uploaded");
  Ctrl h{};
  std::vector<int> core, halo;
  size_t n = 0;
  int rc = snap_layout(c, h, core, halo, n);
  if (rc) return rc;
  if (bytes != (int64_t)n) return fail(c, GY_ERR_INVALID, "snapshot buffer has %lld bytes, the state needs %zu", (long long)bytes, n);
  DevPtrs &d = c->d;
  char *p = static_cast<char *>(buf);
  SnapHead H{};
  std::memcpy(H.magic, kSnapMagic, 8);
  H.version = 1;
  H.ninst = (uint32_t)c->inst.size();
  H.np = d.np; H.cap = d.cap; H.nnodes = c->nnodes; H.ntiles = d.ntiles; H.nkeys = d.nkeys; H.nbricks = d.nbricks;
  H.nmats = (int64_t)c->mats.size();
  H.parity = h.parity; H.acc_valid = c->acc_valid ? 1 : 0; H.multibody = d.multibody; H.thermal = d.thermal;
  H.kmask = d.kmask; H.host_step = c->host_step; H.ncore = (int32_t)core.size(); H.nhalo = (int32_t)halo.size();
  std::memcpy(p, &H, sizeof H); p += sizeof H;
  std::memcpy(p, c->inst.data(), c->inst.size() * sizeof(InstDesc)); p += c->inst.size() * sizeof(InstDesc);
  std::memcpy(p, c->mats.data(), c->mats.size() * sizeof(MatDev)); p += c->mats.size() * sizeof(MatDev);
  for (const auto &I : c->ih) {
    std::memcpy(p, &I.dt, sizeof(double)); p += sizeof(double);
    std::memcpy(p, I.carry_plastic.data(), I.carry_plastic.size()); p += I.carry_plastic.size();
    std::memcpy(p, I.carry_damage.data(), I.carry_damage.size()); p += I.carry_damage.size();
  }
  for (const Seg &g : snap_segments(c, h.parity)) {
    HIPCHECK(c, hipMemcpy(p, g.dev, g.bytes, hipMemcpyDeviceToHost));
    p += g.bytes;
  }
  std::memcpy(p, core.data(), core.size() * 4); p += core.size() * 4;
  for (int t : core) {
    HIPCHECK(c, hipMemcpyAsync(p, d.pcore + (size_t)t * CORE_W, 4 * (size_t)CORE_W, hipMemcpyDeviceToHost, c->stream));
    HIPCHECK(c, hipMemcpyAsync(p + 4 * (size_t)CORE_W, d.nodemask + (size_t)t * (P2G_NODES / 64), 8 * (P2G_NODES / 64),
                               hipMemcpyDeviceToHost, c->stream));
    p += CORE_BYTES;
  }
  std::memcpy(p, halo.data(), halo.size() * 4); p += halo.size() * 4;
  for (int t : halo) {
    HIPCHECK(c, hipMemcpyAsync(p, d.partial + (size_t)t * PARTIAL_W, HALO_BYTES, hipMemcpyDeviceToHost, c->stream));
    p += HALO_BYTES;
  }
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  return (size_t)(p - static_cast<char *>(buf)) == n ? GY_OK : fail(c, GY_ERR_STATE, "snapshot layout mismatch");
}

int gy_snapshot_load(gy_ctx *c, const void *buf, int64_t bytes) {
  if (!c || !buf) return GY_ERR_INVALID;
  if (!c->uploaded) return fail(c, GY_ERR_STATE, "upload the particles of the same setup before loading a snapshot");
  DevPtrs &d = c->d;
  const char *p = static_cast<const char *>(buf), *end = p + bytes;
  SnapHead H;
  if (bytes < (int64_t)sizeof H) return fail(c, GY_ERR_INVALID, "snapshot too short");
  std::memcpy(&H, p, sizeof H); p += sizeof H;
  if (std::memcmp(H.magic, kSnapMagic, 8) != 0 || H.version != 1) return fail(c, GY_ERR_INVALID, "not a snapshot of this format");
  if (H.ninst != c->inst.size() || H.np != d.np || H.cap != d.cap || H.nnodes != c->nnodes || H.ntiles != d.ntiles ||
      H.nkeys != d.nkeys || H.nbricks != d.nbricks || H.nmats != (int64_t)c->mats.size())
    return fail(c, GY_ERR_INVALID, "the snapshot was taken of a different setup (grid, instances, materials or particle count)");
  if (H.multibody != d.multibody || H.thermal != d.thermal)
    return fail(c, GY_ERR_INVALID, "the snapshot's contact / heat-transfer mode differs from the uploaded particles'");
  if (H.parity != 0 && H.parity != 1) return fail(c, GY_ERR_INVALID, "corrupt snapshot");
  size_t need = c->inst.size() * sizeof(InstDesc) + c->mats.size() * sizeof(MatDev);
  for (const auto &I : c->ih) need += sizeof(double) + I.carry_plastic.size() + I.carry_damage.size();
  for (const Seg &g : snap_segments(c, H.parity)) need += g.bytes;
  need += (size_t)H.ncore * (4 + CORE_BYTES) + (size_t)H.nhalo * (4 + HALO_BYTES);
  if (H.ncore < 0 || H.nhalo < 0 || (size_t)(end - p) != need) return fail(c, GY_ERR_INVALID, "snapshot size mismatch");
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  drop_graph(c);
  std::memcpy(c->inst.data(), p, c->inst.size() * sizeof(InstDesc)); p += c->inst.size() * sizeof(InstDesc);
  std::memcpy(c->mats.data(), p, c->mats.size() * sizeof(MatDev)); p += c->mats.size() * sizeof(MatDev);
  for (auto &I : c->ih) {
    std::memcpy(&I.dt, p, sizeof(double)); p += sizeof(double);
    std::memcpy(I.carry_plastic.data(), p, I.carry_plastic.size()); p += I.carry_plastic.size();
    std::memcpy(I.carry_damage.data(), p, I.carry_damage.size()); p += I.carry_damage.size();
  }
  for (const Seg &g : snap_segments(c, H.parity)) {
    HIPCHECK(c, hipMemcpy(g.dev, p, g.bytes, hipMemcpyHostToDevice));
    p += g.bytes;
  }
  std::vector<int> core((size_t)H.ncore), halo((size_t)H.nhalo);
  std::memcpy(core.data(), p, core.size() * 4); p += core.size() * 4;
  for (int t : core) {
    if (t < 0 || t >= d.ntiles) return fail(c, GY_ERR_INVALID, "corrupt snapshot tile list");
    HIPCHECK(c, hipMemcpyAsync(d.pcore + (size_t)t * CORE_W, p, 4 * (size_t)CORE_W, hipMemcpyHostToDevice, c->stream));
    HIPCHECK(c, hipMemcpyAsync(d.nodemask + (size_t)t * (P2G_NODES / 64), p + 4 * (size_t)CORE_W, 8 * (P2G_NODES / 64),
                               hipMemcpyHostToDevice, c->stream));
    p += CORE_BYTES;
  }
  std::memcpy(halo.data(), p, halo.size() * 4); p += halo.size() * 4;
  for (int t : halo) {
    if (t < 0 || t >= d.ntiles) return fail(c, GY_ERR_INVALID, "corrupt snapshot tile list");
    HIPCHECK(c, hipMemcpyAsync(d.partial + (size_t)t * PARTIAL_W, p, HALO_BYTES, hipMemcpyHostToDevice, c->stream));
    p += HALO_BYTES;
  }
  int rc = push_inst(c);
  if (rc) return rc;
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  d.kmask = H.kmask;
  c->acc_valid = H.acc_valid != 0;
  c->host_step = H.host_step;
  return GY_OK;
}

}  // extern "C"

// ---- autotuner ----------------------------------------------------------------------
// Launch shapes tried per workload: persistent particle workgroups per CU (1, 2, 3, 4, 6),
// the fused kernel's build (3 waves per SIMD at 168 VGPRs, or 2 at 256) and the grid
// kernel's workgroups per persistent workgroup (1, 2, 4). The cell-tile edge is not among
// them: at edge 4 the FP64 P2G region of a tile (8^3 nodes x 7 sums) already takes
// 28.7 KB of the 160 KB LDS of a CU, and edge 6 would take 89.6 KB (one workgroup per
// CU). Each trial starts from the same device state (a snapshot), and the state is put
// back afterwards. The winner can be kept in a cache file, one line per workload shape.
namespace {
std::string tune_key(gy_ctx *c) {
  hipDeviceProp_t pr{};
  std::string arch = "unknown";
  int cus = 0;
  if (hipGetDeviceProperties(&pr, c->device) == hipSuccess) {
    arch = pr.gcnArchName;
    cus = pr.multiProcessorCount;
  }
  for (char &ch : arch)
    if (ch == ' ' || ch == '|') ch = '_';
  char buf[512];
  std::snprintf(buf, sizeof buf, "v1|%s|cu%d|inst%zu|np%d|nodes%lld|tiles%d|kmask%d|mb%d|th%d|plane%d", arch.c_str(), cus,
                c->inst.size(), c->d.np, c->nnodes, c->d.ntiles, c->d.kmask, c->d.multibody, c->d.thermal, c->plane ? 1 : 0);
  return buf;
}

bool tune_lookup(const char *path, const std::string &key, gy_tuning &t) {
  FILE *fp = std::fopen(path, "r");
  if (!fp) return false;
  char line[1024];
  bool found = false;
  while (std::fgets(line, sizeof line, fp)) {
    char k[600];
    long long wgs = 0;
    int waves = 0, factor = 0;
    double ms = 0;
    if (std::sscanf(line, "%599s %lld %d %d %lf", k, &wgs, &waves, &factor, &ms) == 5 && key == k) {
      t.persist_wgs = wgs;
      t.fused_waves = waves;
      t.grid_wgs_factor = factor;
      t.ms_per_step = ms;
      found = true;  // the last matching line wins
    }
  }
  std::fclose(fp);
  return found;
}

int tune_store(const char *path, const std::string &key, const gy_tuning &t) {
  std::vector<std::string> keep;
  if (FILE *fp = std::fopen(path, "r")) {
    char line[1024];
    while (std::fgets(line, sizeof line, fp)) {
      char k[600];
      if (std::sscanf(line, "%599s", k) == 1 && key != k) keep.push_back(line);
    }
    std::fclose(fp);
  }
  const std::string tmp = std::string(path) + ".tmp";
  FILE *fp = std::fopen(tmp.c_str(), "w");
  if (!fp) return GY_ERR_INVALID;
  for (const auto &l : keep) std::fputs(l.c_str(), fp);
  std::fprintf(fp, "%s %lld %d %d %.6f\n", key.c_str(), (long long)t.persist_wgs, t.fused_waves, t.grid_wgs_factor,
               t.ms_per_step);
  if (std::fclose(fp) != 0 || std::rename(tmp.c_str(), path) != 0) return GY_ERR_INVALID;
  return GY_OK;
}
}  // namespace

extern "C" {

int gy_get_tuning(const gy_ctx *c, gy_tuning *t) {
  if (!c || !t) return GY_ERR_INVALID;
  *t = gy_tuning();
  t->persist_wgs = c->d.persist_wgs;
  t->fused_waves = c->d.fused_waves;
  t->grid_wgs_factor = c->d.grid_factor;
  return GY_OK;
}

int gy_set_tuning(gy_ctx *c, const gy_tuning *t) {
  if (!c || !t) return GY_ERR_INVALID;
  if (t->persist_wgs < 8 || t->persist_wgs > (1 << 20) || (t->fused_waves != 2 && t->fused_waves != 3) ||
      (t->grid_wgs_factor != 1 && t->grid_wgs_factor != 2 && t->grid_wgs_factor != 4))
    return fail(c, GY_ERR_INVALID, "tuning out of range (workgroups >= 8, waves 2 or 3, grid factor 1, 2 or 4)");
  c->d.persist_wgs = (int)(8 * ((t->persist_wgs + 7) / 8));
  c->d.fused_waves = t->fused_waves;
  c->d.grid_factor = t->grid_wgs_factor;
  drop_graph(c);
  return GY_OK;
}

int gy_autotune_shapes(gy_ctx *c, int32_t trial_steps, const char *cache_path, gy_tuning *out) {
  if (!c || !out) return GY_ERR_INVALID;
  if (!c->uploaded) return fail(c, GY_ERR_STATE, "no particles uploaded");
  if (trial_steps < 1 || trial_steps > 1000) return fail(c, GY_ERR_INVALID, "trial_steps must be in [1, 1000]");
  int rc = gy_synchronize(c);
  if (rc) return rc;
  const std::string key = tune_key(c);
  gy_tuning t = gy_tuning();
  if (cache_path && *cache_path && tune_lookup(cache_path, key, t)) {
    if ((rc = gy_set_tuning(c, &t))) return rc;
    t.from_cache = 1;
    *out = t;
    return GY_OK;
  }
  Ctrl h{};
  HIPCHECK(c, hipMemcpy(&h, c->d.ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost));
  int64_t nbytes = 0;
  if ((rc = gy_snapshot_size(c, &nbytes))) return rc;
  std::vector<uint8_t> snap((size_t)nbytes);
  if ((rc = gy_snapshot_save(c, snap.data(), nbytes))) return rc;
  int cus = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device) != hipSuccess || cus <= 0) cus = 256;
  const bool saved_prof = c->profile;
  c->profile = false;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  // one trial: from the snapshot, an untimed step, then trial_steps timed steps
  auto trial = [&](const gy_tuning &cand, double &ms_out) -> int {
    int r = gy_set_tuning(c, &cand);
    if (r) return r;
    if ((r = gy_snapshot_load(c, snap.data(), nbytes))) return r;
    if ((r = iterate_impl(c, 0.0, h.step, 1))) return r;
    (void)hipEventRecord(e0, c->stream);
    if ((r = iterate_impl(c, 0.0, h.step + 1, trial_steps))) return r;
    (void)hipEventRecord(e1, c->stream);
    if (hipStreamSynchronize(c->stream) != hipSuccess) return fail(c, GY_ERR_HIP, "autotune trial failed");
    if ((r = check_status(c))) return r;
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    ms_out = ms / trial_steps;
    return GY_OK;
  };
  gy_tuning best = gy_tuning();
  best.ms_per_step = 1e300;
  const int per_cu[] = {1, 2, 3, 4, 6};
  const bool single_body = !c->d.multibody && !c->plane;  // the 2-wave build exists for the 3D single-body kernel
  for (int waves : {3, 2}) {
    if (waves == 2 && !single_body) break;
    for (int k : per_cu) {
      gy_tuning cand = gy_tuning();
      cand.persist_wgs = 8 * ((cus * k + 7) / 8);
      cand.fused_waves = waves;
      cand.grid_wgs_factor = 2;
      double ms = 0;
      if ((rc = trial(cand, ms))) break;
      if (ms < best.ms_per_step) { best = cand; best.ms_per_step = ms; }
    }
    if (rc) break;
  }
  if (!rc)
    for (int f : {1, 4}) {
      gy_tuning cand = best;
      cand.grid_wgs_factor = f;
      double ms = 0;
      if ((rc = trial(cand, ms))) break;
      if (ms < best.ms_per_step) { best = cand; best.ms_per_step = ms; }
    }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  c->profile = saved_prof;
  const int rc2 = gy_set_tuning(c, rc ? &t : &best);  // on failure: back to the defaults
  const int rc3 = gy_snapshot_load(c, snap.data(), nbytes);
  if (rc) return rc;
  if (rc2) return rc2;
  if (rc3) return rc3;
  best.from_cache = 0;
  if (cache_path && *cache_path && tune_store(cache_path, key, best) != GY_OK)
    return fail(c, GY_ERR_INVALID, "cannot write the tuning cache %s", cache_path);
  *out = best;
  return GY_OK;
}

int gy_autotune(gy_ctx *c, int32_t trial_steps, int64_t *best_wgs, double *best_ms) {
  gy_tuning t = gy_tuning();
  const int rc = gy_autotune_shapes(c, trial_steps, nullptr, &t);
  if (rc) return rc;
  if (best_wgs) *best_wgs = t.persist_wgs;
  if (best_ms) *best_ms = t.ms_per_step;
  return GY_OK;
}

int gy_set_option(gy_ctx *c, int32_t option, int64_t value) {
  if (!c) return GY_ERR_INVALID;
  switch (option) {
    case GY_OPT_USE_GRAPH: c->use_graph = value != 0; if (!c->use_graph) drop_graph(c); return GY_OK;
    case GY_OPT_PROFILE: c->profile = value != 0; return GY_OK;
    case GY_OPT_TILE: return value == TILE ? GY_OK : fail(c, GY_ERR_UNSUPPORTED, "tile edge %lld not built", (long long)value);
    case GY_OPT_PERSIST_WGS:
      if (value < 8 || value > 65536) return fail(c, GY_ERR_INVALID, "persistent workgroups must be in [8, 65536]");
      c->d.persist_wgs = (int)(8 * ((value + 7) / 8));
      drop_graph(c);
      return GY_OK;
    default: return fail(c, GY_ERR_INVALID, "unknown option %d", option);
  }
}

int gy_kernel_count(void) { return K_COUNT_; }
const char *gy_kernel_name(int32_t i) { return (i >= 0 && i < K_COUNT_) ? kKernelNames[i] : ""; }

int gy_kernel_stats(gy_ctx *c, int32_t i, int64_t *launches, double *total_ms) {
  if (!c || i < 0 || i >= K_COUNT_) return GY_ERR_INVALID;
  if (launches) *launches = c->launches[i];
  if (total_ms) *total_ms = c->total_ms[i];
  return GY_OK;
}

int gy_reset_kernel_stats(gy_ctx *c) {
  if (!c) return GY_ERR_INVALID;
  (void)hipStreamSynchronize(c->stream);
  collect_timing(c);
  for (int i = 0; i < K_COUNT_; i++) { c->launches[i] = 0; c->total_ms[i] = 0; }
  return GY_OK;
}

void *gy_stream(gy_ctx *c) { return c ? (void *)c->stream : nullptr; }

int gy_set_return_network(gy_ctx *c, gy_mlp *m) {
  if (!c || !m) return GY_ERR_INVALID;
  std::vector<uint2> frag(NN_TILES * 64);
  std::vector<float> bias(NN_NBIAS), norm(NN_NORM_N);
  int rc = mlp_export(m, frag.data(), bias.data(), norm.data());
  if (rc) return fail(c, rc, "network export failed: %s", gy_mlp_last_error(m));
  const size_t fb = frag.size() * sizeof(uint2), bb = bias.size() * sizeof(float), nb = norm.size() * sizeof(float);
  if (!c->nnbuf) {
    HIPCHECK(c, hipMalloc(&c->nnbuf, fb + bb + nb));
    char *base = (char *)c->nnbuf;
    c->d.nn.frag = (const uint2 *)base;
    c->d.nn.bias = (const float *)(base + fb);
    c->d.nn.norm = (const float *)(base + fb + bb);
    drop_graph(c);  // the captured launches hold the old pointers
  }
  HIPCHECK(c, hipMemcpyAsync((void *)c->d.nn.frag, frag.data(), fb, hipMemcpyHostToDevice, c->stream));
  HIPCHECK(c, hipMemcpyAsync((void *)c->d.nn.bias, bias.data(), bb, hipMemcpyHostToDevice, c->stream));
  HIPCHECK(c, hipMemcpyAsync((void *)c->d.nn.norm, norm.data(), nb, hipMemcpyHostToDevice, c->stream));
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  return GY_OK;
}

int gy_return_map_counts(gy_ctx *c, int64_t *network, int64_t *fallback, int32_t reset) {
  if (!c) return GY_ERR_INVALID;
  if (!c->d.ctrl) return fail(c, GY_ERR_STATE, "no particles uploaded");
  unsigned long long n[2] = {0, 0};
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  HIPCHECK(c, hipMemcpy(n, c->d.ctrl->nn_count, sizeof n, hipMemcpyDeviceToHost));
  if (network) *network = (int64_t)n[0];
  if (fallback) *fallback = (int64_t)n[1];
  if (reset) {
    const unsigned long long z[2] = {0, 0};
    HIPCHECK(c, hipMemcpy(c->d.ctrl->nn_count, z, sizeof z, hipMemcpyHostToDevice));
  }
  return GY_OK;
}

void gy_destroy(gy_ctx *c) {
  if (!c) return;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  collect_timing(c);
  drop_graph(c);
  DevPtrs &d = c->d;
  for (int b = 0; b < 2; b++) { dfree(d.pbuf[b]); dfree(d.key[b]); }
  dfree(d.rank); dfree(d.tmp); dfree(d.count); dfree(d.offsets); dfree(d.blocksum);
  dfree(d.acc); dfree(d.vel); dfree(d.brick_flag); dfree(d.brick_list); dfree(d.bc_type); dfree(d.bc_val);
  dfree(d.pbc_val); dfree(d.tile_inst); dfree(d.brick_inst); dfree(d.inst); dfree(d.mats);
  dfree(d.ctrl); dfree(d.t_last); dfree(d.t_cur); dfree(d.audit);
  dfree(d.blocktiles); dfree(d.blockbox); dfree(d.active); dfree(d.tile_epoch); dfree(d.halo_epoch); dfree(d.nodemask); dfree(d.pcore); dfree(d.partial);
  dfree(d.acc2); dfree(d.nrm); dfree(d.claim); dfree(d.velb); dfree(d.cellobj); dfree(d.twotile);
  dfree(d.therm);
  dfree(c->nnbuf);
  for (auto e : c->event_pool) (void)hipEventDestroy(e);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

}  // extern "C"
