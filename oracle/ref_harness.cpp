// ref_harness.cpp — C interface around the compiled reference (Graphyt) classes.
//
// TEST INFRASTRUCTURE ONLY. oracle/Makefile compiles this file together with the
// reference's own sources (read in place from the reference checkout, never copied) into
// oracle/_ref/libgraphyt_ref.so. The tests use it to pin the FP64 restatement
// (cpu_ref.cpp) and bench.py may time it as the CPU baseline. It exposes the same
// signatures as cpu_ref.cpp with the `rref_` prefix.
//
// Built a second time with -DGY_REF_GPU into oracle/_ref/libgraphyt_ref_gpu.so, together
// with the Graphyt-side binding graphyt_amd/integration/graphyt_iterate_gpu.cpp (the
// bodies of Solver::iterate_GPU / Synchronise / sync_GPU) and libgraphyt_amd.so: then the
// reference's own Solver runs the MI355X step through its iterate_GPU slot
// (rref_iterate_gpu, rref_synchronise), which the GPU tests drive like a Graphyt script.
//
// Scenario set-up follows the reference's Python usage (README.md:180-294):
// Nodes(L, h); MaterialPoints(nodes); psep; add(i, x, y, z) per particle; numParticles;
// per-particle setters; Parameters(tmax, is3D) + setters; Solver(...); optional setDt.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <unistd.h>
#include <unordered_map>
#include <vector>

#include "Solver.hpp"
#include "../include/graphyt_amd.h"
#ifdef GY_REF_GPU
#include <exception>
#include <string>
extern "C" void graphyt_gpu_release(const Solver *s);
#endif

namespace {

struct Key {
  double x, y, z;
  bool operator==(const Key &o) const { return x == o.x && y == o.y && z == o.z; }
};
struct KeyHash {
  size_t operator()(const Key &k) const {
    uint64_t a, b, c;
    std::memcpy(&a, &k.x, 8);
    std::memcpy(&b, &k.y, 8);
    std::memcpy(&c, &k.z, 8);
    return (size_t)(a * 0x9E3779B97F4A7C15ull ^ (b + 0x632BE59BD9B4E019ull) * 31 ^ c * 1315423911ull);
  }
};

struct Ref {
  Nodes *nodes = nullptr;
  MaterialPoints *mp = nullptr;
  Parameters *par = nullptr;
  Solver *sol = nullptr;
  std::unordered_map<Key, int, KeyHash> orig;  // pos0 -> caller index
};

// The reference prints set-up chatter on stdout; keep the caller's stdout clean.
struct Quiet {
  int saved = -1;
  Quiet() {
    fflush(stdout);
    saved = dup(1);
    FILE *nul = fopen("/dev/null", "w");
    if (nul) { dup2(fileno(nul), 1); fclose(nul); }
  }
  ~Quiet() {
    fflush(stdout);
    if (saved >= 0) { dup2(saved, 1); close(saved); }
  }
};

}  // namespace

extern "C" {

void *rref_create(const gy_grid *g, const gy_material *mats, int32_t nmat, const gy_params *p,
                  const gy_particles *in) {
  if (!g || !mats || !p || !in || !in->pos || in->n <= 0 || nmat <= 0) return nullptr;
  Quiet q;
  Ref *r = new Ref();
  std::vector<double> L = {g->L[0], g->L[1], g->L[2]};
  r->nodes = new Nodes(L, g->cellsize);
  r->mp = new MaterialPoints(r->nodes);
  r->mp->psep = p->psep;
  const int n = in->n;
  for (int i = 0; i < n; i++) r->mp->add(i, in->pos[3 * i], in->pos[3 * i + 1], in->pos[3 * i + 2]);
  r->mp->numParticles = n;
  for (int i = 0; i < n; i++) {
    if (in->vel) r->mp->setVel(i, in->vel[3 * i], in->vel[3 * i + 1], in->vel[3 * i + 2]);
    if (in->stress) {
      const double *s = in->stress + 6 * i;
      r->mp->setStress(i, s[0], s[1], s[2], s[3], s[4], s[5]);
    }
    if (in->material) r->mp->setMaterialID(i, in->material[i]);
    if (in->obj_id) r->mp->setObjectID(i, in->obj_id[i]);
    if (in->hold && in->hold[i]) r->mp->holdParticle(i);
    if (in->damage) r->mp->setDamage(i, in->damage[i]);
    if (in->temperature) r->mp->setTemp(i, in->temperature[i]);
    if (in->bc_type && in->bc_type[i] != 0) {
      std::vector<double> v(4, 0.0);
      if (in->bc_val)
        for (int k = 0; k < 4; k++) v[k] = in->bc_val[4 * i + k];
      r->mp->setBC(i, (BCTypes::function)in->bc_type[i], v);
    }
    Key k{in->pos[3 * i], in->pos[3 * i + 1], in->pos[3 * i + 2]};
    r->orig[k] = i;
  }
  r->par = new Parameters(1.0, p->is3D != 0);
  r->par->setGravity(p->gravity[0], p->gravity[1], p->gravity[2]);
  r->par->setDampingCoef(p->damping);
  r->par->setFriction(p->mu);
  r->par->setCFL(p->cfl);
  std::vector<MaterialModels> mv(nmat);
  for (int i = 0; i < nmat; i++) {
    MaterialModels &m = mv[i];
    const gy_material &s = mats[i];
    m.setModel(s.model);
    m.setIsRigid(s.is_rigid != 0);
    m.setIsDamage(s.is_damage != 0);
    m.setDamageModel(s.damage_model);
    m.setBulkMod(s.bulk_mod);
    m.setShearMod(s.shear_mod);
    m.setDensity(s.density);
    m.setViscosity(s.viscosity);
    m.setYieldStress(s.yield_stress);
    m.setCohesion(s.cohesion);
    m.setCriticalStrain(s.critical_strain);
    m.setFrictionAngle(s.phi_deg);
    m.setDilatancyAngle(s.psi_deg);
    m.setCrackSpeed(s.crack_speed);
    m.setDamageM(s.damage_m);
    m.setDamageK(s.damage_k);
    m.setThermalConductivity(s.thermal_conductivity);
    m.setSpecificHeatCapacity(s.specific_heat);
    m.setTurbulentStress(s.turb_stress);
  }
  r->sol = new Solver(r->nodes, r->mp, r->par, mv);
  if (p->dt > 0) r->sol->setDt(p->dt);
  return r;
}

int32_t rref_set_node_bc(void *h, int32_t n, const int32_t *type, const double *val4) {
  Ref *r = (Ref *)h;
  if (n != r->nodes->numNodes) return GY_ERR_INVALID;
  for (int i = 0; i < n; i++) {
    std::vector<double> v(4, 0.0);
    if (val4)
      for (int k = 0; k < 4; k++) v[k] = val4[4 * i + k];
    r->nodes->setBC(i, (BCTypes::function)type[i], v);
  }
  return GY_OK;
}

double rref_dt(void *h) { return ((Ref *)h)->sol->dt; }
void rref_set_dt(void *h, double dt) { ((Ref *)h)->sol->setDt(dt); }
int32_t rref_num_nodes(void *h) { return ((Ref *)h)->nodes->numNodes; }

void rref_iterate(void *h, double t, int32_t step) {
  Ref *r = (Ref *)h;
  Quiet q;
  r->sol->iterate_CPU(t, step, r->par);
}

void rref_iterate_n(void *h, double t0, int32_t step0, int32_t n) {
  Ref *r = (Ref *)h;
  Quiet q;
  for (int i = 0; i < n; i++) r->sol->iterate_CPU(t0 + i * r->sol->dt, step0 + i, r->par);
}

// Outputs in the caller's original order, matched through the (permuted) pos0 array.
void rref_get(void *h, gy_particles *o) {
  Ref *r = (Ref *)h;
  MaterialPoints *m = r->mp;
  const int n = m->numParticles;
  for (int p = 0; p < n; p++) {
    Key k{m->pos0[3 * p], m->pos0[3 * p + 1], m->pos0[3 * p + 2]};
    auto it = r->orig.find(k);
    if (it == r->orig.end()) continue;
    const int i = it->second;
    for (int c = 0; c < 3; c++) {
      if (o->pos) o->pos[3 * i + c] = m->position[3 * p + c];
      if (o->pos0) o->pos0[3 * i + c] = m->pos0[3 * p + c];
      if (o->vel) o->vel[3 * i + c] = m->velocity[3 * p + c];
      if (o->displacement) o->displacement[3 * i + c] = m->displacement[3 * p + c];
    }
    for (int c = 0; c < 6; c++) {
      if (o->stress) o->stress[6 * i + c] = m->stress[6 * p + c];
      if (o->strain) o->strain[6 * i + c] = m->strain[6 * p + c];
      if (o->plastic_strain) o->plastic_strain[6 * i + c] = m->plastic_strain[6 * p + c];
    }
    if (o->acc_plastic_strain) o->acc_plastic_strain[i] = m->acc_plastic_strain[p];
    if (o->density) o->density[i] = m->density[p];
    if (o->mass) o->mass[i] = m->mass[p];
    if (o->damage) o->damage[i] = m->damage[p];
    if (o->smax) o->smax[i] = m->smax[p];
    if (o->material) o->material[i] = m->material[p];
    if (o->obj_id) o->obj_id[i] = m->objID[p];
    if (o->volume) o->volume[i] = m->volume[p];
    if (o->cell_id) o->cell_id[i] = m->cellID[p];
    if (o->temperature) o->temperature[i] = m->temperature[p];
  }
}

void rref_get_nodes(void *h, double *mass, double *vel, double *vb1, double *ab1, double *force) {
  Ref *r = (Ref *)h;
  Nodes *nd = r->nodes;
  const size_t N = nd->numNodes;
  if (mass) std::memcpy(mass, nd->mass, N * sizeof(double));
  if (vel) std::memcpy(vel, nd->velocity, 3 * N * sizeof(double));
  if (vb1) std::memcpy(vb1, nd->velocity_body1, 3 * N * sizeof(double));
  if (ab1) std::memcpy(ab1, nd->acceleration_body1, 3 * N * sizeof(double));
  if (force) std::memcpy(force, nd->force, 3 * N * sizeof(double));
}

// MaterialPoints::Serialise and Nodes::Serialise (MaterialPoints.cpp:341-364,
// Nodes.cpp:165-177): output/<name>_mpts_<step>.csv and output/<name>_nds_<step>.csv under
// the current directory, particles in the reference's current (sorted) order.
void rref_serialise(void *h, int32_t step, const char *name) {
  Ref *r = (Ref *)h;
  std::vector<char> buf(name, name + std::strlen(name) + 1);
  r->mp->Serialise(step, buf.data());
  r->nodes->Serialise(step, buf.data());
}

// The reference's own eigen-decomposition (eig3.cpp:258-267), for pinning the oracle's.
void rref_eig3(const double *A9, double *V9, double *d3) {
  double A[3][3], V[3][3];
  for (int i = 0; i < 9; i++) A[i / 3][i % 3] = A9[i];
  eigen_decomposition(A, V, d3);
  for (int i = 0; i < 9; i++) V9[i] = V[i / 3][i % 3];
}

#ifdef GY_REF_GPU
static std::string g_gpu_error;

// solver.iterate_GPU(t, step, parameters) and solver.Synchronise() of the reference's
// Solver with the MI355X binding; 0 or -1 (message in rref_gpu_error) on an exception.
int32_t rref_iterate_gpu(void *h, double t, int32_t step) {
  Ref *r = (Ref *)h;
  try {
    r->sol->iterate_GPU(t, step, r->par);
  } catch (const std::exception &e) {
    g_gpu_error = e.what();
    return -1;
  }
  return 0;
}

int32_t rref_iterate_gpu_n(void *h, double t0, int32_t step0, int32_t n) {
  Ref *r = (Ref *)h;
  for (int i = 0; i < n; i++)
    if (rref_iterate_gpu(h, t0 + i * r->sol->dt, step0 + i) != 0) return -1;
  return 0;
}

int32_t rref_synchronise(void *h) {
  Ref *r = (Ref *)h;
  try {
    r->sol->Synchronise();
  } catch (const std::exception &e) {
    g_gpu_error = e.what();
    return -1;
  }
  return 0;
}

const char *rref_gpu_error() { return g_gpu_error.c_str(); }
#endif

void rref_destroy(void *h) {
  Ref *r = (Ref *)h;
#ifdef GY_REF_GPU
  graphyt_gpu_release(r->sol);
#endif
  delete r->sol;
  delete r->par;
  delete r->mp;
  delete r->nodes;  // Nodes arrays are never freed by the reference (Nodes.cpp:143-145)
  delete r;
}

}  // extern "C"
