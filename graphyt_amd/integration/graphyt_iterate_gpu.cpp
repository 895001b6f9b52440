// graphyt_iterate_gpu.cpp — the Graphyt-side binding of the MI355X step.
//
// Bodies of Graphyt's GPU slot, written against Graphyt's own headers and the C ABI of
// include/graphyt_amd.h:
//   Solver::iterate_GPU(double t, int step, Parameters*)  (Solver.hpp:21; stub Solver.cpp:293-302)
//   Solver::Synchronise()                                   (Solver.hpp:26; empty at Solver.cpp:305)
//   Solver::sync_GPU()                                      (Solver.hpp:27; Solver.cu:5-8)
// A maintainer adds this file to Graphyt's build, deletes the two empty stubs from
// Solver.cpp and links libgraphyt_amd.so. No Graphyt header changes.
//
// Protocol. The first iterate_GPU of a Solver creates its device context from the
// Solver's Nodes (L, cellsize, node BCs), materials, Parameters and dt, and uploads the
// particle arrays of its MaterialPoints. Each iterate_GPU then only enqueues one step on
// the context's stream (Solver::iterate_CPU's sequence, Solver.cpp:86-292). The host
// arrays are stale from then on, until Synchronise() copies the device state back into
// MaterialPoints (positions, velocities, stress and strain fields, density, volume,
// damage, displacement, cell ids, temperature) and Nodes (mass, velocity, force): a
// lazy, device-dirty download, one PCIe transfer per output step instead of one per
// step. After a Synchronise the host arrays are authoritative again (setters may change
// them) and the next iterate_GPU uploads them. iterate_CPU between GPU steps needs a
// Synchronise before it (it reads the host arrays) and leaves the host authoritative.
// The particle order of the host arrays never changes on the GPU path.
//
// Errors are thrown as std::runtime_error with gy_last_error's text (SWIG's %exception
// turns them into Python's RuntimeError).
#include <cstdint>
#include <mutex>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "Solver.hpp"
#include "graphyt_amd.h"

namespace {

struct GpuState {
  gy_ctx *ctx = nullptr;
  bool device_valid = false;  // the device holds the current particle state
  bool host_stale = false;    // the device has stepped since the last download
  double dt = 0.0;            // the dt the context runs with (Solver::setDt may change it)
  std::vector<uint8_t> hold;  // MaterialPoints::hold is a vector<bool>: bytes for the ABI
};

// Per-Solver device state. Kept beside the Solver so that Graphyt's headers stay as they
// are; a maintainer may equally make these Solver members.
std::mutex g_mu;
std::unordered_map<const Solver *, GpuState> g_state;

GpuState &state_of(const Solver *s) {
  std::lock_guard<std::mutex> lk(g_mu);
  return g_state[s];
}

void check(const GpuState &st, int rc, const char *what) {
  if (rc == GY_OK) return;
  std::string msg = std::string("graphyt_amd: ") + what + " failed (code " + std::to_string(rc) + ")";
  if (st.ctx) msg += std::string(": ") + gy_last_error(st.ctx);
  throw std::runtime_error(msg);
}

// MaterialModels (MaterialModels.hpp:5-48) -> gy_material, field for field.
gy_material to_gy(const MaterialModels &m) {
  gy_material g = gy_material();
  g.model = m.model;
  g.is_rigid = m.isRigid ? 1 : 0;
  g.is_damage = m.isDamage ? 1 : 0;
  g.damage_model = m.damageModel;
  g.bulk_mod = m.bulkMod;
  g.shear_mod = m.shearMod;
  g.density = m.density;
  g.viscosity = m.viscosity;
  g.yield_stress = m.yieldStress;
  g.cohesion = m.cohesion;
  g.critical_strain = m.criticalStrain;
  g.phi_deg = m.phi;
  g.psi_deg = m.psi;
  g.crack_speed = m.Cg;
  g.damage_m = m.damageM;
  g.damage_k = m.damageK;
  g.thermal_conductivity = m.thermalConductivity;
  g.specific_heat = m.specificHeatCapacity;
  g.turb_stress = m.turbStress;
  g.dp_apex_guard = 0;  // the reference's apex return
  return g;
}

// The particle arrays of MaterialPoints (MaterialPoints.hpp:84-137) as an ABI view.
gy_particles upload_view(MaterialPoints *mp, std::vector<uint8_t> &hold) {
  gy_particles v = gy_particles();
  v.n = mp->numParticles;
  v.pos = mp->position.data();
  v.pos0 = mp->pos0.data();
  v.vel = mp->velocity.data();
  v.stress = mp->stress.data();
  v.strain = mp->strain.data();
  v.plastic_strain = mp->plastic_strain.data();
  v.acc_plastic_strain = mp->acc_plastic_strain.data();
  v.density = mp->density.data();
  v.mass = mp->mass.data();
  v.damage = mp->damage.data();
  v.smax = mp->smax.data();
  v.material = mp->material.data();
  v.obj_id = mp->objID.data();
  hold.assign(mp->hold.begin(), mp->hold.end());
  v.hold = hold.data();
  v.bc_type = mp->BCTYPE.data();
  v.bc_val = mp->BCVAL.data();
  v.temperature = mp->temperature.data();
  v.temperature_grad = mp->temperatureGrad.data();
  return v;
}

// The fields a step changes, written back into MaterialPoints.
gy_particles download_view(MaterialPoints *mp) {
  gy_particles v = gy_particles();
  v.n = mp->numParticles;
  v.pos = mp->position.data();
  v.vel = mp->velocity.data();
  v.stress = mp->stress.data();
  v.strain = mp->strain.data();
  v.plastic_strain = mp->plastic_strain.data();
  v.acc_plastic_strain = mp->acc_plastic_strain.data();
  v.density = mp->density.data();
  v.damage = mp->damage.data();
  v.smax = mp->smax.data();
  v.volume = mp->volume.data();
  v.displacement = mp->displacement.data();
  v.cell_id = mp->cellID.data();
  v.temperature = mp->temperature.data();
  v.temperature_grad = mp->temperatureGrad.data();
  return v;
}

}  // namespace

void Solver::iterate_GPU(double t, int step, Parameters *prm) {
  GpuState &st = state_of(this);
  if (!st.ctx) {
    gy_grid g = gy_grid();
    for (int k = 0; k < 3; k++) g.L[k] = nds->L[k];
    g.cellsize = nds->cellsize;
    std::vector<gy_material> mats;
    for (size_t i = 0; i < materials.size(); i++) mats.push_back(to_gy(materials[i]));
    gy_params p = gy_params();
    for (int k = 0; k < 3; k++) p.gravity[k] = prm->gravity[k];
    p.damping = prm->velocityDampingCoef;
    p.mu = prm->mu;
    p.cfl = prm->CFL;
    p.psep = mpts->psep;
    p.dt = this->dt;  // Solver.cpp:15-32 or setDt
    p.is3D = prm->is3D ? 1 : 0;
    check(st, gy_create(&st.ctx, &g, mats.data(), (int32_t)mats.size(), &p), "gy_create");
    check(st, gy_set_node_bc(st.ctx, nds->numNodes, nds->BCTYPE, nds->BCVAL), "gy_set_node_bc");
    st.dt = this->dt;
  }
  if (this->dt != st.dt) {
    check(st, gy_set_dt(st.ctx, this->dt), "gy_set_dt");
    st.dt = this->dt;
  }
  if (!st.device_valid) {
    gy_particles v = upload_view(mpts, st.hold);
    check(st, gy_upload(st.ctx, &v), "gy_upload");
    st.device_valid = true;
  }
  check(st, gy_iterate(st.ctx, t, step), "gy_iterate");
  st.host_stale = true;
  mpts->t = t;  // Solver.cpp:264-265
  nds->t = t;
}

void Solver::Synchronise() {
  GpuState &st = state_of(this);
  if (!st.ctx || !st.host_stale) return;
  gy_particles v = download_view(mpts);
  check(st, gy_download(st.ctx, &v), "gy_download");
  check(st, gy_download_nodes(st.ctx, nds->numNodes, nds->mass, nds->velocity, nds->force), "gy_download_nodes");
  st.host_stale = false;
  st.device_valid = false;  // the host arrays are authoritative again
}

void Solver::sync_GPU() {
  GpuState &st = state_of(this);
  if (st.ctx) check(st, gy_synchronize(st.ctx), "gy_synchronize");
}

// Releases the device context of a Solver (Solver::~Solver is empty in Graphyt,
// Solver.cpp:38; a maintainer calls this from it).
extern "C" void graphyt_gpu_release(const Solver *s) {
  std::lock_guard<std::mutex> lk(g_mu);
  auto it = g_state.find(s);
  if (it == g_state.end()) return;
  if (it->second.ctx) gy_destroy(it->second.ctx);
  g_state.erase(it);
}
