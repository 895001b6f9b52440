// graphyt_module.cpp — the graphyt Python surface (Nodes, MaterialPoints, MaterialModels,
// Parameters, Solver, BCTypes) as C++ host classes whose Solver runs the MPM step on the
// GPU through the C ABI of libgraphyt_amd.so.
//
// The names, argument order and meaning follow the reference's SWIG interface
// (Nodes.i:8-23, MaterialPoints.i:8-66, MaterialModels.i, Parameters.i:7-15,
// Solver.i:10-19, BoundaryConditions.hpp:6-77), so a script written for
// `graphyt.Solver.iterate_CPU` runs with `iterate_GPU`. Host arrays use the reference
// layouts (vectors interleaved per particle, tensors xx yy zz xy yz xz) and keep the
// caller's particle order across steps.
//
// Device state is synchronised lazily: a setter after a step first pulls the device
// state, and the next iterate_GPU uploads it again; getters pull the device state when
// a step ran since the last pull. Errors raise RuntimeError with the library's message.
// There is no CPU step here: iterate_CPU is the reference's own.
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "graphyt_amd.h"

namespace py = pybind11;

namespace gyhost {

// ---- BCTypes (BoundaryConditions.hpp:6-77): the integer values are the ABI ----------
static const char *const kBCNames[] = {
    "none", "grid_vx", "grid_vy", "grid_vz", "grid_temp", "grid_tflux_x", "grid_tflux_y", "grid_tflux_z",
    "grid_fx", "grid_fy", "grid_fz", "grid_vx_vy", "grid_vx_vz", "grid_vy_vz", "grid_fx_fy", "grid_fy_fz",
    "grid_fx_fz", "grid_fx_fy_fz", "grid_vx_vy_vz", "grid_vx_temp", "grid_vy_temp", "grid_vz_temp",
    "grid_fx_temp", "grid_fy_temp", "grid_fz_temp", "grid_vx_vy_temp", "grid_vx_vz_temp", "grid_vy_vz_temp",
    "grid_fx_fy_temp", "grid_fy_fz_temp", "grid_fx_fz_temp", "grid_fx_fy_fz_temp", "grid_vx_vy_vz_temp",
    "mp_vx", "mp_vy", "mp_vz", "mp_temp", "mp_fx", "mp_fy", "mp_fz", "mp_vx_vy", "mp_vx_vz", "mp_vy_vz",
    "mp_fx_fy", "mp_fy_fz", "mp_fx_fz", "mp_fx_fy_fz", "mp_vx_vy_vz", "mp_vx_temp", "mp_vy_temp",
    "mp_vz_temp", "mp_fx_temp", "mp_fy_temp", "mp_fz_temp", "mp_vx_vy_temp", "mp_vx_vz_temp",
    "mp_vy_vz_temp", "mp_fx_fy_temp", "mp_fy_fz_temp", "mp_fx_fz_temp", "mp_fx_fy_fz_temp",
    "mp_vx_vy_vz_temp", "mp_stress", "mp_stress_xx_yy", "mpstrain", "mp_vxsin", "mp_vysin", "mp_syysin"};
constexpr int kNumBC = (int)(sizeof(kBCNames) / sizeof(kBCNames[0]));
static_assert(kNumBC == 68, "68 boundary-condition codes");
struct BCTypes {};  // namespace object: graphyt.BCTypes.<name> is the integer code

inline void check(gy_ctx *ctx, int rc, const char *what) {
  if (rc == GY_OK) return;
  std::string msg = std::string(what) + " failed (graphyt_amd status " + std::to_string(rc) + ")";
  const char *e = ctx ? gy_last_error(ctx) : nullptr;
  if (e && *e) msg += ": " + std::string(e);
  throw std::runtime_error(msg);
}

// ---- Parameters (Parameters.cpp:8-35) ------------------------------------------------
struct Parameters {
  Parameters(double tmax_, bool is3D_) : tmax(tmax_), is3D(is3D_) {}
  void setGravity(double gx, double gy, double gz) { gravity = {gx, gy, gz}; }
  void setDampingCoef(double c) { velocityDampingCoef = c; }
  void setCFL(double c) { CFL = c; }
  void setFriction(double m) { mu = m; }
  double dt = 0.0;
  double tmax;
  double CFL = 0.1;
  bool is3D;
  double mu = 0.0;
  std::vector<double> gravity{0.0, 0.0, 0.0};  // the reference leaves it unset until setGravity
  double velocityDampingCoef = 0.0;
};

// ---- MaterialModels (defaults MaterialModels.cpp:7-25) --------------------------------
struct MaterialModels {
  void setBulkMod(double v) { bulkMod = v; }
  void setShearMod(double v) { shearMod = v; }
  void setDensity(double v) { density = v; }
  void setIsRigid(double v) { isRigid = v != 0.0; }  // double in MaterialModels.i:14
  void setViscosity(double v) { viscosity = v; }
  void setThermalConductivity(double v) { thermalConductivity = v; }
  void setSpecificHeatCapacity(double v) { specificHeatCapacity = v; }
  void setYieldStress(double v) { yieldStress = v; }
  void setCohesion(double v) { cohesion = v; }
  void setCriticalStrain(double v) { criticalStrain = v; }
  void setFrictionAngle(double v) { phi = v; }
  void setDilatancyAngle(double v) { psi = v; }
  void setIsDamage(bool v) { isDamage = v; }
  void setDamageModel(int v) { damageModel = v; }
  void setCrackSpeed(double v) { Cg = v; }
  void setDamageM(double v) { damageM = v; }
  void setDamageK(double v) { damageK = v; }
  void setTurbulentStress(double v) { turbStress = v; }
  void setModel(int v) { model = v; }
  // not in the reference: the Drucker-Prager apex guard of gy_material
  void setApexGuard(bool v) { apexGuard = v; }

  gy_material to_gy() const {
    gy_material g{};
    g.model = model;
    g.is_rigid = isRigid;
    g.is_damage = isDamage;
    g.damage_model = damageModel;
    g.bulk_mod = bulkMod;
    g.shear_mod = shearMod;
    g.density = density;
    g.viscosity = viscosity;
    g.yield_stress = yieldStress;
    g.cohesion = cohesion;
    g.critical_strain = criticalStrain;
    g.phi_deg = phi;
    g.psi_deg = psi;
    g.crack_speed = Cg;
    g.damage_m = damageM;
    g.damage_k = damageK;
    g.thermal_conductivity = thermalConductivity;
    g.specific_heat = specificHeatCapacity;
    g.turb_stress = turbStress;
    g.dp_apex_guard = apexGuard;
    return g;
  }

  double bulkMod = 1e7, shearMod = 5000.0, density = 1000.0;
  bool isRigid = false;
  double viscosity = 1.0, thermalConductivity = 391.0, specificHeatCapacity = 385.2;
  double yieldStress = 1e7, cohesion = 1e5, criticalStrain = 0.01, phi = 40.0, psi = 10.0;
  bool isDamage = false;
  int damageModel = 1;
  double Cg = 586.52, damageM = 9.0, damageK = 5.39e36, turbStress = 1e10;
  int model = 1;
  bool apexGuard = false;
};

class Solver;

// ---- Nodes (Nodes.cpp:8-141): uniform grid from the origin, n = i + j*Nx + k*Nx*Ny ---
struct Nodes {
  Nodes(std::vector<double> L_, double cellsize_) : L(L_), cellsize(cellsize_) {
    if (L.size() != 3 || !(cellsize > 0)) throw std::invalid_argument("Nodes(L[3], cellsize > 0)");
    Nx = (int)(1 + L[0] / cellsize);  // Nodes.cpp:71-73
    Ny = (int)(1 + L[1] / cellsize);
    Nz = (int)(1 + L[2] / cellsize);
    numNodes = Nx * Ny * Nz;
    position.resize(3 * (size_t)numNodes);
    for (int k = 0; k < Nz; k++)
      for (int j = 0; j < Ny; j++)
        for (int i = 0; i < Nx; i++) {
          const size_t n = i + (size_t)Nx * (j + (size_t)Ny * k);
          position[3 * n] = i * cellsize;
          position[3 * n + 1] = j * cellsize;
          position[3 * n + 2] = k * cellsize;
        }
    mass.assign(numNodes, 0.0);
    velocity.assign(3 * (size_t)numNodes, 0.0);
    force.assign(3 * (size_t)numNodes, 0.0);
    BCTYPE.assign(numNodes, 0);
    BCVAL.assign(4 * (size_t)numNodes, 0.0);
  }
  // Nodes.cpp:155-163
  void setBC(int n, int type, std::vector<double> val) {
    if (n < 0 || n >= numNodes) throw std::out_of_range("node index");
    if (type < 0 || type >= kNumBC) throw std::invalid_argument("BC type");
    if (val.size() > 4) throw std::invalid_argument("at most 4 BC values");
    BCTYPE[n] = type;
    for (size_t i = 0; i < val.size(); i++) BCVAL[4 * (size_t)n + i] = val[i];
    bc_dirty = true;
  }
  // Nodes.cpp:165-177: output/<name>_nds_<step>.csv
  void Serialise(int step, const std::string &name) const {
    const std::string path = "output/" + name + "_nds_" + std::to_string(step) + ".csv";
    FILE *fp = std::fopen(path.c_str(), "w+");
    if (!fp) throw std::runtime_error("cannot open " + path);
    std::fprintf(fp, "Pos_x,Pos_y,Pos_z\n");
    for (int i = 0; i < numNodes; i++)
      std::fprintf(fp, "%f,%f,%f\n", position[3 * (size_t)i], position[3 * (size_t)i + 1], position[3 * (size_t)i + 2]);
    std::fclose(fp);
  }
  double getPosX(int n) const { return position.at(3 * (size_t)n); }
  double getPosY(int n) const { return position.at(3 * (size_t)n + 1); }
  double getPosZ(int n) const { return position.at(3 * (size_t)n + 2); }
  void pull();  // node fields of the last step from the device

  std::vector<double> L;
  double cellsize;
  int Nx, Ny, Nz, numNodes;
  std::vector<double> position, mass, velocity, force, BCVAL;
  std::vector<int32_t> BCTYPE;
  bool bc_dirty = true;
  Solver *solver = nullptr;
};

// ---- MaterialPoints (MaterialPoints.cpp:95-364) ----------------------------------------
struct MaterialPoints {
  explicit MaterialPoints(Nodes *n) : nodes(n) {}

  // MaterialPoints::add(i, x, y, z): append one particle at (x, y, z); `i` names the
  // new particle's index as in the reference (its cell id is computed from it).
  void add(int i, double x, double y, double z);
  // addPyckGeom(pos, n): n particles from a flat array of 3n coordinates
  void addPyckGeom(py::array_t<double, py::array::c_style | py::array::forcecast> pos, int n) {
    if (pos.size() < 3 * (py::ssize_t)n) throw std::invalid_argument("positions hold fewer than 3n values");
    const double *p = pos.data();
    for (int i = 0; i < n; i++) add((int)size(), p[3 * i], p[3 * i + 1], p[3 * i + 2]);
    numParticles = (int)size();
  }
  // addPyckIntField(field) (MaterialPoints.cpp:32-53): a field named "material" or
  // "objectID" sets that per-particle array from field.data (pyck's IntField, or
  // graphyt.IntField); other names are ignored. addPyckDoubleField is a no-op there too.
  void addPyckIntField(py::object field) {
    const std::string name = py::str(field.attr("name"));
    auto data = py::array_t<int32_t, py::array::c_style | py::array::forcecast>(field.attr("data"));
    const size_t n = size();
    if ((size_t)data.size() < n) throw std::invalid_argument("field holds fewer values than particles");
    const int32_t *v = data.data();
    if (name == "material") {
      touch();
      for (size_t i = 0; i < n; i++) material[i] = v[i];
    } else if (name == "objectID") {
      touch();
      for (size_t i = 0; i < n; i++) objID[i] = v[i];
    }
  }
  void addPyckDoubleField(py::object) {}
  size_t size() const { return position.size() / 3; }

  void pull();     // device -> host when a step ran since the last pull
  void touch();    // before a host-side change: pull, then mark for upload
  void idx(int p) const {
    if (p < 0 || (size_t)p >= size()) throw std::out_of_range("particle index");
  }

  // scalar getters (MaterialPoints.cpp:237-260); getStress* read stress[3p + k] exactly
  // as the reference does (its tests depend on it)
  double getPosX(int p) { pull(); idx(p); return position[3 * (size_t)p]; }
  double getPosY(int p) { pull(); idx(p); return position[3 * (size_t)p + 1]; }
  double getPosZ(int p) { pull(); idx(p); return position[3 * (size_t)p + 2]; }
  double getDispX(int p) { pull(); idx(p); return displacement[3 * (size_t)p]; }
  double getDispY(int p) { pull(); idx(p); return displacement[3 * (size_t)p + 1]; }
  double getDispZ(int p) { pull(); idx(p); return displacement[3 * (size_t)p + 2]; }
  double getVelX(int p) { pull(); idx(p); return velocity[3 * (size_t)p]; }
  double getVelY(int p) { pull(); idx(p); return velocity[3 * (size_t)p + 1]; }
  double getVelZ(int p) { pull(); idx(p); return velocity[3 * (size_t)p + 2]; }
  double stress3(int p, int k) { pull(); idx(p); return stress[3 * (size_t)p + k]; }
  int getObjID(int p) { idx(p); return objID[p]; }
  int getMatID(int p) { idx(p); return material[p]; }
  int getCellID(int p) { pull(); idx(p); return cellID[p]; }

  // setters (MaterialPoints.cpp:88-93, 302-339)
  void setBC(int p, int type, std::vector<double> val) {
    idx(p);
    if (type < 0 || type >= kNumBC) throw std::invalid_argument("BC type");
    if (val.size() > 4) throw std::invalid_argument("at most 4 BC values");
    touch();
    BCTYPE[p] = type;
    for (size_t i = 0; i < val.size(); i++) BCVAL[4 * (size_t)p + i] = val[i];
  }
  void setDamage(int p, double D) { idx(p); touch(); damage[p] = D; }
  void setVel(int p, double vx, double vy, double vz) {
    idx(p); touch();
    velocity[3 * (size_t)p] = vx; velocity[3 * (size_t)p + 1] = vy; velocity[3 * (size_t)p + 2] = vz;
  }
  void setStress(int p, double sxx, double syy, double szz, double sxy, double syz, double sxz) {
    idx(p); touch();
    double *s = &stress[6 * (size_t)p];
    s[0] = sxx; s[1] = syy; s[2] = szz; s[3] = sxy; s[4] = syz; s[5] = sxz;
  }
  void setTemp(int p, double T) { idx(p); touch(); temperature[p] = T; }
  void setMaterialID(int p, int m) { idx(p); touch(); material[p] = m; }
  void setObjectID(int p, int o) { idx(p); touch(); objID[p] = o; }
  void holdParticle(int p) { idx(p); touch(); hold[p] = 1; }
  // MaterialPoints.cpp:58-85: materials by cell id
  void setMaterialByCellId(int cell, int m) {
    touch();
    for (size_t p = 0; p < size(); p++)
      if (cellID[p] == cell) material[p] = m;
  }
  void setMaterialByCellArray(const std::vector<int> &mats) {
    touch();
    for (size_t p = 0; p < size(); p++) {
      const int c = cellID[p];
      if (c < 0 || (size_t)c >= mats.size()) throw std::out_of_range("cell id beyond the material array");
      material[p] = mats[c];
    }
  }
  // MaterialPoints.cpp:274-296 (the 2D rotation about a z axis)
  void setRotatingBodyVel(int objectID, double cx, double cy, double cz, double omega) {
    (void)cz;
    touch();
    for (size_t p = 0; p < size(); p++) {
      if (objID[p] != objectID) continue;
      const double x = position[3 * p], y = position[3 * p + 1];
      const double r = std::sqrt((x - cx) * (x - cx) + (y - cy) * (y - cy));
      const double ct = (y - cy) / r, st = (x - cx) / r;
      velocity[3 * p] = -omega * r * ct;
      velocity[3 * p + 1] = omega * r * st;
      velocity[3 * p + 2] = 0.0;
      for (int k = 0; k < 6; k++) stress[6 * p + k] = 0.0;
    }
  }
  // MaterialPoints.cpp:341-364: output/<name>_mpts_<step>.csv
  void Serialise(int step, const std::string &name) {
    pull();
    const std::string path = "output/" + name + "_mpts_" + std::to_string(step) + ".csv";
    FILE *fp = std::fopen(path.c_str(), "w+");
    if (!fp) throw std::runtime_error("cannot open " + path);
    std::fprintf(fp, "x,y,z,vm,P,T,vx,vy,vz,rho\n");
    for (size_t i = 0; i < (size_t)numParticles && i < size(); i++) {
      const double *s = &stress[6 * i];
      const double vm = std::pow(std::pow(s[0], 2) + std::pow(s[1], 2) + std::pow(s[2], 2) + std::pow(2.0 * s[3], 2) +
                                     std::pow(2.0 * s[4], 2) + std::pow(2.0 * s[5], 2),
                                 0.5);
      const double P = 0.33 * (s[0] + s[1] + s[2]);
      std::fprintf(fp, "%f,%f,%f,%f,%f,%f,%f,%f,%f,%f\n", position[3 * i], position[3 * i + 1], position[3 * i + 2],
                   vm, P, temperature[i], velocity[3 * i], velocity[3 * i + 1], velocity[3 * i + 2], density[i]);
    }
    std::fclose(fp);
  }

  gy_particles view() {
    gy_particles v{};
    v.n = (int32_t)size();
    v.pos = position.data(); v.pos0 = pos0.data(); v.vel = velocity.data();
    v.stress = stress.data(); v.strain = strain.data(); v.plastic_strain = plastic_strain.data();
    v.acc_plastic_strain = acc_plastic_strain.data(); v.density = density.data(); v.mass = mass.data();
    v.damage = damage.data(); v.smax = smax.data(); v.material = material.data(); v.obj_id = objID.data();
    v.hold = hold.data(); v.bc_type = BCTYPE.data(); v.bc_val = BCVAL.data(); v.volume = volume.data();
    v.displacement = displacement.data(); v.cell_id = cellID.data(); v.temperature = temperature.data();
    v.temperature_grad = temperatureGrad.data();
    return v;
  }

  Nodes *nodes;
  int numParticles = 0;
  double psep = 0.0;
  std::vector<double> position, pos0, displacement, velocity, acceleration, stress, strain, plastic_strain,
      acc_plastic_strain, damage, smax, density, volume, mass, temperature, normals, BCVAL;
  std::vector<double> temperatureGrad;  // MaterialPoints::temperatureGrad (3 per particle)
  std::vector<int32_t> cellID, material, objID, BCTYPE;
  std::vector<uint8_t> hold;
  bool host_dirty = true;    // host arrays changed since the last upload
  bool device_newer = false; // a step ran since the last pull
  Solver *solver = nullptr;
};

void MaterialPoints::add(int i, double x, double y, double z) {
  const double p3[3] = {x, y, z};
  for (int a = 0; a < 3; a++) {
    position.push_back(p3[a]);
    pos0.push_back(p3[a]);
    displacement.push_back(0.0);
    velocity.push_back(0.0);
    acceleration.push_back(0.0);
    normals.push_back(0.0);
    temperatureGrad.push_back(0.0);
  }
  for (int k = 0; k < 6; k++) {
    stress.push_back(0.0);
    strain.push_back(0.0);
    plastic_strain.push_back(0.0);
  }
  for (int k = 0; k < 4; k++) BCVAL.push_back(0.0);
  acc_plastic_strain.push_back(0.0);
  damage.push_back(0.0);
  smax.push_back(0.0);
  density.push_back(0.0);
  volume.push_back(0.0);
  mass.push_back(0.0);
  temperature.push_back(0.0);
  material.push_back(0);
  objID.push_back(0);
  BCTYPE.push_back(0);
  hold.push_back(0);
  // cell id from the position of particle i (MaterialPoints.cpp:220-231)
  const int Nx = nodes->Nx, Ny = nodes->Ny;
  const double h = nodes->cellsize;
  const size_t q = (i >= 0 && (size_t)i < size()) ? (size_t)i : size() - 1;
  const int ix = (int)(position[3 * q] / h), iy = (int)(position[3 * q + 1] / h), iz = (int)(position[3 * q + 2] / h);
  cellID.push_back(ix + iy * (Nx - 1) + iz * (Nx - 1) * (Ny - 1));
  numParticles = (int)size();
  host_dirty = true;
}

// ---- Solver (Solver.cpp:7-84, iterate_GPU slot Solver.hpp:21) ------------------------
class Solver {
 public:
  Solver(Nodes *nds_, MaterialPoints *mpts_, Parameters *prm_, std::vector<MaterialModels> mats)
      : nds(nds_), mpts(mpts_), prm(prm_), materials(std::move(mats)) {
    if (!nds || !mpts || !prm) throw std::invalid_argument("Solver(nodes, matpoints, parameters, materials)");
    if (materials.empty()) throw std::invalid_argument("at least one material");
    // dt = CFL * h / sqrt(max K / min rho) over every material (Solver.cpp:15-32)
    double maxK = 1.0, minrho = 1e15;
    for (const auto &m : materials) {
      if (m.bulkMod > maxK) maxK = m.bulkMod;
      if (m.density < minrho) minrho = m.density;
    }
    dt = prm->CFL * (nds->cellsize / std::sqrt(maxK / minrho));
    tmax = prm->tmax;
    startTimeLoop = std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
    initialize();
    mpts->solver = this;
    nds->solver = this;
  }
  ~Solver() {
    if (mpts && mpts->solver == this) mpts->solver = nullptr;
    if (nds && nds->solver == this) nds->solver = nullptr;
    if (ctx) gy_destroy(ctx);
  }

  // Solver::initialize (Solver.cpp:39-67): per unit depth in the plane path, where rigid
  // particles keep rho * V
  void initialize() {
    const bool d3 = prm->is3D;
    const double v = d3 ? mpts->psep * mpts->psep * mpts->psep : mpts->psep * mpts->psep;
    for (size_t p = 0; p < mpts->size(); p++) {
      const int m = mpts->material[p];
      if (m < 0 || (size_t)m >= materials.size()) throw std::invalid_argument("particle material out of range");
      mpts->density[p] = materials[m].density;
      mpts->volume[p] = v;
      mpts->mass[p] = (d3 && materials[m].isRigid) ? 1.0 : mpts->density[p] * v;
    }
    mpts->host_dirty = true;
  }
  void setDt(double d) {
    dt = d;
    if (ctx) check(ctx, gy_set_dt(ctx, d), "gy_set_dt");
  }

  // one step, == iterate_CPU(t, step, parameters)
  void iterate_GPU(double t, int step, Parameters *p) {
    prepare(p);
    int rc;
    {
      py::gil_scoped_release nogil;
      rc = gy_iterate(ctx, t, step);
    }
    check(ctx, rc, "iterate_GPU");
    mpts->device_newer = true;
  }
  // nsteps steps from t = 0, step 0 (iterate_CPU_steps, Solver.cpp:75-84, with the
  // solver's dt), replayed from a captured hipGraph
  void iterate_GPU_steps(int nsteps, Parameters *p) {
    prepare(p);
    int rc;
    {
      py::gil_scoped_release nogil;
      rc = gy_iterate_n(ctx, 0.0, 0, nsteps);
    }
    check(ctx, rc, "iterate_GPU_steps");
    mpts->device_newer = true;
  }
  // wait for the device and pull particle and node state into the host arrays
  void Synchronise() {
    if (!ctx) return;
    check(ctx, gy_synchronize(ctx), "gy_synchronize");
    mpts->pull();
    nds->pull();
  }

  void download_particles() {
    gy_particles v = mpts->view();
    check(ctx, gy_download(ctx, &v), "gy_download");
  }
  void download_nodes() {
    check(ctx, gy_download_nodes(ctx, nds->numNodes, nds->mass.data(), nds->velocity.data(), nds->force.data()),
          "gy_download_nodes");
  }
  bool has_context() const { return ctx != nullptr; }

  // Binary checkpoint of the state of the last step (the format of
  // graphyt_amd/checkpoint.py): particle arrays, node BCs, grid, t, step, dt.
  void checkpoint(const std::string &path, double t, int64_t step) {
    Synchronise();
    struct Arr { const char *name; int code; size_t count; const void *data; };
    const auto &P = *mpts;
    std::vector<Arr> arrs = {
        {"pos", 0, P.position.size(), P.position.data()}, {"pos0", 0, P.pos0.size(), P.pos0.data()},
        {"vel", 0, P.velocity.size(), P.velocity.data()}, {"stress", 0, P.stress.size(), P.stress.data()},
        {"strain", 0, P.strain.size(), P.strain.data()},
        {"plastic_strain", 0, P.plastic_strain.size(), P.plastic_strain.data()},
        {"acc_plastic_strain", 0, P.acc_plastic_strain.size(), P.acc_plastic_strain.data()},
        {"density", 0, P.density.size(), P.density.data()}, {"mass", 0, P.mass.size(), P.mass.data()},
        {"volume", 0, P.volume.size(), P.volume.data()}, {"damage", 0, P.damage.size(), P.damage.data()},
        {"smax", 0, P.smax.size(), P.smax.data()}, {"temperature", 0, P.temperature.size(), P.temperature.data()},
        {"temperature_grad", 0, P.temperatureGrad.size(), P.temperatureGrad.data()},
        {"material", 1, P.material.size(), P.material.data()},
        {"obj_id", 1, P.objID.size(), P.objID.data()}, {"hold", 2, P.hold.size(), P.hold.data()},
        {"bc_type", 1, P.BCTYPE.size(), P.BCTYPE.data()}, {"bc_val", 0, P.BCVAL.size(), P.BCVAL.data()},
        {"cell_id", 1, P.cellID.size(), P.cellID.data()},
        {"node_bc_type", 1, nds->BCTYPE.size(), nds->BCTYPE.data()},
        {"node_bc_val", 0, nds->BCVAL.size(), nds->BCVAL.data()},
        {"grid_L", 0, nds->L.size(), nds->L.data()}, {"grid_h", 0, 1, &nds->cellsize},
        {"psep", 0, 1, &P.psep}};
    // the device state of the last step (gy_snapshot_save), when there is a device context
    std::vector<uint8_t> snap;
    if (ctx) {
      int64_t nb = 0;
      check(ctx, gy_snapshot_size(ctx, &nb), "gy_snapshot_size");
      snap.resize((size_t)nb);
      check(ctx, gy_snapshot_save(ctx, snap.data(), nb), "gy_snapshot_save");
      arrs.push_back({"device_snapshot", 2, snap.size(), snap.data()});
    }
    FILE *fp = std::fopen(path.c_str(), "wb");
    if (!fp) throw std::runtime_error("cannot open " + path);
    const size_t esize[3] = {8, 4, 1};
    bool ok = std::fwrite("GYCKPT\x00\x01", 1, 8, fp) == 8;
    const uint32_t head[2] = {1u, (uint32_t)arrs.size()};
    const double td[2] = {t, dt};
    ok = ok && std::fwrite(head, 4, 2, fp) == 2 && std::fwrite(td, 8, 2, fp) == 2 && std::fwrite(&step, 8, 1, fp) == 1;
    for (const auto &a : arrs) {
      const uint32_t ln = (uint32_t)std::strlen(a.name);
      const uint8_t code = (uint8_t)a.code;
      const uint64_t count = a.count;
      ok = ok && std::fwrite(&ln, 4, 1, fp) == 1 && std::fwrite(a.name, 1, ln, fp) == ln &&
           std::fwrite(&code, 1, 1, fp) == 1 && std::fwrite(&count, 8, 1, fp) == 1 &&
           (count == 0 || std::fwrite(a.data, esize[a.code], count, fp) == count);
    }
    ok = (std::fclose(fp) == 0) && ok;
    if (!ok) throw std::runtime_error("writing " + path + " failed");
  }

  // Restore a checkpoint written for the same setup (grid, particle count, materials):
  // the host arrays are replaced; with the checkpoint's device snapshot the device context
  // is set up from them and given back the checkpointed device state bit for bit (binning,
  // node records, the previous step's time), else they are uploaded before the next step.
  // Returns (t, step).
  std::pair<double, int64_t> restore(const std::string &path) {
    FILE *fp = std::fopen(path.c_str(), "rb");
    if (!fp) throw std::runtime_error("cannot open " + path);
    std::vector<char> data;
    {
      char buf[1 << 16];
      size_t k;
      while ((k = std::fread(buf, 1, sizeof buf, fp)) > 0) data.insert(data.end(), buf, buf + k);
      std::fclose(fp);
    }
    size_t off = 0;
    auto take = [&](void *dst, size_t n) {
      if (off + n > data.size()) throw std::runtime_error(path + ": truncated checkpoint");
      std::memcpy(dst, data.data() + off, n);
      off += n;
    };
    char magic[8];
    take(magic, 8);
    if (std::memcmp(magic, "GYCKPT\x00\x01", 8) != 0) throw std::runtime_error(path + ": not a graphyt checkpoint");
    uint32_t head[2];
    take(head, 8);
    if (head[0] != 1u) throw std::runtime_error(path + ": unsupported checkpoint version");
    double td[2];
    int64_t step = 0;
    take(td, 16);
    take(&step, 8);
    Synchronise();  // nothing the device holds is newer than what the checkpoint replaces
    std::vector<uint8_t> snap;
    auto &P = *mpts;
    const size_t n = P.size();
    const size_t esize[3] = {8, 4, 1};
    for (uint32_t a = 0; a < head[1]; a++) {
      uint32_t ln;
      take(&ln, 4);
      if (ln > 256) throw std::runtime_error(path + ": bad array name");
      std::string name(ln, '\0');
      take(&name[0], ln);
      uint8_t code;
      uint64_t count;
      take(&code, 1);
      take(&count, 8);
      if (code > 2) throw std::runtime_error(path + ": bad array type");
      const size_t nbytes = (size_t)count * esize[code];
      if (off + nbytes > data.size()) throw std::runtime_error(path + ": truncated array " + name);
      const char *src = data.data() + off;
      off += nbytes;
      auto fill = [&](auto &vec, int want_code, size_t want) {
        if (code != want_code || count != want) throw std::runtime_error(path + ": " + name + " does not fit this setup");
        std::memcpy(vec.data(), src, nbytes);
      };
      if (name == "pos") fill(P.position, 0, 3 * n);
      else if (name == "pos0") fill(P.pos0, 0, 3 * n);
      else if (name == "vel") fill(P.velocity, 0, 3 * n);
      else if (name == "stress") fill(P.stress, 0, 6 * n);
      else if (name == "strain") fill(P.strain, 0, 6 * n);
      else if (name == "plastic_strain") fill(P.plastic_strain, 0, 6 * n);
      else if (name == "acc_plastic_strain") fill(P.acc_plastic_strain, 0, n);
      else if (name == "density") fill(P.density, 0, n);
      else if (name == "mass") fill(P.mass, 0, n);
      else if (name == "volume") fill(P.volume, 0, n);
      else if (name == "damage") fill(P.damage, 0, n);
      else if (name == "smax") fill(P.smax, 0, n);
      else if (name == "temperature") fill(P.temperature, 0, n);
      else if (name == "temperature_grad") fill(P.temperatureGrad, 0, 3 * n);
      else if (name == "material") fill(P.material, 1, n);
      else if (name == "obj_id") fill(P.objID, 1, n);
      else if (name == "hold") fill(P.hold, 2, n);
      else if (name == "bc_type") fill(P.BCTYPE, 1, n);
      else if (name == "bc_val") fill(P.BCVAL, 0, 4 * n);
      else if (name == "cell_id") fill(P.cellID, 1, n);
      else if (name == "node_bc_type") fill(nds->BCTYPE, 1, (size_t)nds->numNodes);
      else if (name == "node_bc_val") fill(nds->BCVAL, 0, 4 * (size_t)nds->numNodes);
      else if (name == "device_snapshot") {
        if (code != 2) throw std::runtime_error(path + ": bad device snapshot");
        snap.assign(reinterpret_cast<const uint8_t *>(src), reinterpret_cast<const uint8_t *>(src) + nbytes);
      } else if (name == "grid_h") {
        double h;
        std::memcpy(&h, src, 8);
        if (code != 0 || count != 1 || h != nds->cellsize) throw std::runtime_error(path + ": grid spacing differs");
      }
    }
    for (size_t p = 0; p < n; p++)
      for (int k = 0; k < 3; k++) P.displacement[3 * p + k] = P.position[3 * p + k] - P.pos0[3 * p + k];
    dt = td[1];
    if (ctx) check(ctx, gy_set_dt(ctx, dt), "gy_set_dt");
    P.host_dirty = true;
    P.device_newer = false;
    nds->bc_dirty = true;
    if (!snap.empty()) {
      prepare(prm);  // context, node BCs and the restored arrays on the device
      check(ctx, gy_snapshot_load(ctx, snap.data(), (int64_t)snap.size()), "gy_snapshot_load");
    }
    return {td[0], step};
  }

  Nodes *nds;
  MaterialPoints *mpts;
  Parameters *prm;
  std::vector<MaterialModels> materials;
  double dt = 0.0, tmax = 0.0, startTimeLoop = 0.0;

 private:
  void prepare(Parameters *p) {
    if (p && p != prm) {
      if (ctx && p->is3D != prm->is3D) throw std::runtime_error("iterate_GPU: is3D changed after the first step");
      prm = p;
    }
    if (!ctx) {
      gy_grid g{};
      for (int a = 0; a < 3; a++) g.L[a] = nds->L[a];
      g.cellsize = nds->cellsize;
      std::vector<gy_material> mats;
      for (const auto &m : materials) mats.push_back(m.to_gy());
      gy_params gp{};
      for (int a = 0; a < 3; a++) gp.gravity[a] = prm->gravity[a];
      gp.damping = prm->velocityDampingCoef;
      gp.mu = prm->mu;
      gp.cfl = prm->CFL;
      gp.psep = mpts->psep;
      gp.dt = dt;
      gp.is3D = prm->is3D ? 1 : 0;
      gy_ctx *c = nullptr;
      const int rc = gy_create(&c, &g, mats.data(), (int)mats.size(), &gp);
      if (rc != GY_OK) {
        std::string msg = "gy_create failed (graphyt_amd status " + std::to_string(rc) + ")";
        if (c) {
          const char *e = gy_last_error(c);
          if (e && *e) msg += ": " + std::string(e);
          gy_destroy(c);
        }
        throw std::runtime_error(msg);
      }
      ctx = c;
    }
    if (nds->bc_dirty) {
      check(ctx, gy_set_node_bc(ctx, nds->numNodes, nds->BCTYPE.data(), nds->BCVAL.data()), "gy_set_node_bc");
      nds->bc_dirty = false;
    }
    if (mpts->host_dirty) {
      gy_particles v = mpts->view();
      check(ctx, gy_upload(ctx, &v), "gy_upload");
      mpts->host_dirty = false;
      mpts->device_newer = false;
    }
  }

  gy_ctx *ctx = nullptr;
};

void MaterialPoints::pull() {
  if (device_newer && solver && solver->has_context()) {
    device_newer = false;
    solver->download_particles();
  }
}
void MaterialPoints::touch() {
  pull();
  host_dirty = true;
}
void Nodes::pull() {
  if (solver && solver->has_context()) solver->download_nodes();
}

// numpy view of a host vector, kept alive by its owner
template <typename T>
py::array_t<T> view_of(std::vector<T> &v, py::handle owner) {
  return py::array_t<T>({(py::ssize_t)v.size()}, {(py::ssize_t)sizeof(T)}, v.data(), owner);
}

}  // namespace gyhost

PYBIND11_MODULE(graphyt, m) {
  using namespace gyhost;
  m.doc() = "graphyt surface (Nodes, MaterialPoints, MaterialModels, Parameters, Solver) over the MI355X step";

  // IntField(name, n, dim) (intField.h): a named integer field with n*dim values in
  // `data` (a numpy view), for addPyckIntField where pyck is not installed
  struct IntField {
    std::string name;
    long n = 0;
    int dim = 1;
    std::vector<int32_t> data;
  };
  py::class_<IntField>(m, "IntField")
      .def(py::init([](const std::string &name, long n, int dim) {
             if (n < 0 || dim < 1) throw std::invalid_argument("IntField(name, n >= 0, dim >= 1)");
             IntField f;
             f.name = name;
             f.n = n;
             f.dim = dim;
             f.data.assign((size_t)n * dim, 0);
             return f;
           }),
           py::arg("name"), py::arg("n"), py::arg("dim") = 1)
      .def_readonly("name", &IntField::name)
      .def_readonly("n", &IntField::n)
      .def_readonly("dim", &IntField::dim)
      .def_property_readonly("data", [](py::object self) {
        auto &f = self.cast<IntField &>();
        return py::array_t<int32_t>({(py::ssize_t)f.data.size()}, {(py::ssize_t)sizeof(int32_t)}, f.data.data(), self);
      });

  // BCTypes: graphyt.BCTypes.grid_vx etc. (integers, the reference's codes)
  {
    py::class_<BCTypes> bc(m, "BCTypes");
    for (int i = 0; i < kNumBC; i++) bc.attr(kBCNames[i]) = i;
  }

  py::class_<Parameters>(m, "Parameters")
      .def(py::init<double, bool>(), py::arg("tmax"), py::arg("is3D"))
      .def("setGravity", &Parameters::setGravity)
      .def("setDampingCoef", &Parameters::setDampingCoef)
      .def("setCFL", &Parameters::setCFL)
      .def("setFriction", &Parameters::setFriction)
      .def_readwrite("dt", &Parameters::dt)
      .def_readwrite("tmax", &Parameters::tmax)
      .def_readwrite("CFL", &Parameters::CFL)
      .def_readwrite("is3D", &Parameters::is3D)
      .def_readwrite("mu", &Parameters::mu)
      .def_readwrite("gravity", &Parameters::gravity)
      .def_readwrite("velocityDampingCoef", &Parameters::velocityDampingCoef);

  py::class_<MaterialModels>(m, "MaterialModels")
      .def(py::init<>())
      .def("setBulkMod", &MaterialModels::setBulkMod)
      .def("setShearMod", &MaterialModels::setShearMod)
      .def("setDensity", &MaterialModels::setDensity)
      .def("setIsRigid", &MaterialModels::setIsRigid)
      .def("setViscosity", &MaterialModels::setViscosity)
      .def("setThermalConductivity", &MaterialModels::setThermalConductivity)
      .def("setSpecificHeatCapacity", &MaterialModels::setSpecificHeatCapacity)
      .def("setYieldStress", &MaterialModels::setYieldStress)
      .def("setCohesion", &MaterialModels::setCohesion)
      .def("setCriticalStrain", &MaterialModels::setCriticalStrain)
      .def("setFrictionAngle", &MaterialModels::setFrictionAngle)
      .def("setDilatancyAngle", &MaterialModels::setDilatancyAngle)
      .def("setIsDamage", &MaterialModels::setIsDamage)
      .def("setDamageModel", &MaterialModels::setDamageModel)
      .def("setCrackSpeed", &MaterialModels::setCrackSpeed)
      .def("setDamageM", &MaterialModels::setDamageM)
      .def("setDamageK", &MaterialModels::setDamageK)
      .def("setTurbulentStress", &MaterialModels::setTurbulentStress)
      .def("setModel", &MaterialModels::setModel)
      .def("setApexGuard", &MaterialModels::setApexGuard)
      .def_readwrite("bulkMod", &MaterialModels::bulkMod)
      .def_readwrite("shearMod", &MaterialModels::shearMod)
      .def_readwrite("density", &MaterialModels::density)
      .def_readwrite("isRigid", &MaterialModels::isRigid)
      .def_readwrite("viscosity", &MaterialModels::viscosity)
      .def_readwrite("yieldStress", &MaterialModels::yieldStress)
      .def_readwrite("cohesion", &MaterialModels::cohesion)
      .def_readwrite("criticalStrain", &MaterialModels::criticalStrain)
      .def_readwrite("phi", &MaterialModels::phi)
      .def_readwrite("psi", &MaterialModels::psi)
      .def_readwrite("isDamage", &MaterialModels::isDamage)
      .def_readwrite("damageModel", &MaterialModels::damageModel)
      .def_readwrite("Cg", &MaterialModels::Cg)
      .def_readwrite("damageM", &MaterialModels::damageM)
      .def_readwrite("damageK", &MaterialModels::damageK)
      .def_readwrite("model", &MaterialModels::model);

  py::class_<Nodes>(m, "Nodes")
      .def(py::init<std::vector<double>, double>(), py::arg("L"), py::arg("cellsize"))
      .def("_sync", &Nodes::pull)  // node fields of the last step into the host arrays (writers call it)
      .def("setBC", &Nodes::setBC, py::arg("nIdx"), py::arg("type"), py::arg("val"))
      .def("Serialise", &Nodes::Serialise, py::arg("step"), py::arg("filename"))
      .def("getPosX", &Nodes::getPosX)
      .def("getPosY", &Nodes::getPosY)
      .def("getPosZ", &Nodes::getPosZ)
      .def("getPosArr", [](py::object self) { return view_of(self.cast<Nodes &>().position, self); })
      .def("getMassArr", [](py::object self) { auto &n = self.cast<Nodes &>(); n.pull(); return view_of(n.mass, self); })
      .def("getVelArr", [](py::object self) { auto &n = self.cast<Nodes &>(); n.pull(); return view_of(n.velocity, self); })
      .def("getForceArr", [](py::object self) { auto &n = self.cast<Nodes &>(); n.pull(); return view_of(n.force, self); })
      .def_readonly("cellsize", &Nodes::cellsize)
      .def_readonly("numNodes", &Nodes::numNodes)
      .def_readonly("L", &Nodes::L)
      .def_readonly("Nx", &Nodes::Nx)
      .def_readonly("Ny", &Nodes::Ny)
      .def_readonly("Nz", &Nodes::Nz);

  auto arr = [](std::vector<double> MaterialPoints::*f) {
    return [f](py::object self) {
      auto &mp = self.cast<MaterialPoints &>();
      mp.pull();
      return view_of(mp.*f, self);
    };
  };
  auto iarr = [](std::vector<int32_t> MaterialPoints::*f) {
    return [f](py::object self) {
      auto &mp = self.cast<MaterialPoints &>();
      mp.pull();
      return view_of(mp.*f, self);
    };
  };
  py::class_<MaterialPoints>(m, "MaterialPoints")
      .def(py::init<Nodes *>(), py::keep_alive<1, 2>())
      .def("_sync", &MaterialPoints::pull)  // particle state of the last step into the host arrays
      .def("add", &MaterialPoints::add, py::arg("i"), py::arg("x"), py::arg("y"), py::arg("z"))
      .def("addPyckGeom", &MaterialPoints::addPyckGeom, py::arg("pos"), py::arg("numParticle"))
      .def("addPyckIntField", &MaterialPoints::addPyckIntField, py::arg("intField"))
      .def("addPyckDoubleField", &MaterialPoints::addPyckDoubleField, py::arg("doubleField"))
      .def("getPosX", &MaterialPoints::getPosX)
      .def("getPosY", &MaterialPoints::getPosY)
      .def("getPosZ", &MaterialPoints::getPosZ)
      .def("getDispX", &MaterialPoints::getDispX)
      .def("getDispY", &MaterialPoints::getDispY)
      .def("getDispZ", &MaterialPoints::getDispZ)
      .def("getVelX", &MaterialPoints::getVelX)
      .def("getVelY", &MaterialPoints::getVelY)
      .def("getVelZ", &MaterialPoints::getVelZ)
      .def("getStressXX", [](MaterialPoints &mp, int p) { return mp.stress3(p, 0); })
      .def("getStressYY", [](MaterialPoints &mp, int p) { return mp.stress3(p, 1); })
      .def("getStressZZ", [](MaterialPoints &mp, int p) { return mp.stress3(p, 2); })
      .def("getStressXY", [](MaterialPoints &mp, int p) { return mp.stress3(p, 3); })
      .def("getStressYZ", [](MaterialPoints &mp, int p) { return mp.stress3(p, 4); })
      .def("getStressZX", [](MaterialPoints &mp, int p) { return mp.stress3(p, 5); })
      .def("getObjID", &MaterialPoints::getObjID)
      .def("getMatID", &MaterialPoints::getMatID)
      .def("getCellID", &MaterialPoints::getCellID)
      .def("getPosArr", arr(&MaterialPoints::position))
      .def("getMassArr", arr(&MaterialPoints::mass))
      .def("getMatArr", iarr(&MaterialPoints::material))
      .def("getCellIDArr", iarr(&MaterialPoints::cellID))
      .def("getVelArr", arr(&MaterialPoints::velocity))
      .def("getAccArr", arr(&MaterialPoints::acceleration))
      .def("getDispArr", arr(&MaterialPoints::displacement))
      .def("getStressArr", arr(&MaterialPoints::stress))
      .def("getStrainArr", arr(&MaterialPoints::strain))
      .def("getTemperatureArr", arr(&MaterialPoints::temperature))
      .def("getPlasticStrainArr", arr(&MaterialPoints::plastic_strain))
      .def("getAccPlasticStrainArr", arr(&MaterialPoints::acc_plastic_strain))
      .def("getDamageArr", arr(&MaterialPoints::damage))
      .def("getSMaxArr", arr(&MaterialPoints::smax))
      .def("getNormalsArr", arr(&MaterialPoints::normals))
      .def("getVolumeArr", arr(&MaterialPoints::volume))
      .def("getDensityArr", arr(&MaterialPoints::density))
      .def("getObjectIDArr", iarr(&MaterialPoints::objID))
      .def("setBC", &MaterialPoints::setBC, py::arg("mp"), py::arg("type"), py::arg("val"))
      .def("setDamage", &MaterialPoints::setDamage)
      .def("setVel", &MaterialPoints::setVel)
      .def("setStress", &MaterialPoints::setStress)
      .def("setTemp", &MaterialPoints::setTemp)
      .def("setRotatingBodyVel", &MaterialPoints::setRotatingBodyVel)
      .def("setMaterialID", &MaterialPoints::setMaterialID)
      .def("setObjectID", &MaterialPoints::setObjectID)
      .def("holdParticle", &MaterialPoints::holdParticle)
      .def("setMaterialByCell", &MaterialPoints::setMaterialByCellArray, py::arg("materialArray"))
      .def("setMaterialByCell", &MaterialPoints::setMaterialByCellId, py::arg("idx"), py::arg("materialID"))
      .def("Serialise", &MaterialPoints::Serialise)
      .def_readwrite("numParticles", &MaterialPoints::numParticles)
      .def_readwrite("psep", &MaterialPoints::psep);

  py::class_<Solver>(m, "Solver")
      .def(py::init<Nodes *, MaterialPoints *, Parameters *, std::vector<MaterialModels>>(), py::keep_alive<1, 2>(),
           py::keep_alive<1, 3>(), py::keep_alive<1, 4>())
      .def("iterate_GPU", &Solver::iterate_GPU, py::arg("t"), py::arg("step"), py::arg("parameters"))
      .def("iterate_GPU_steps", &Solver::iterate_GPU_steps, py::arg("nsteps"), py::arg("parameters"))
      .def("setDt", &Solver::setDt)
      .def("Synchronise", &Solver::Synchronise)
      .def("checkpoint", &Solver::checkpoint, py::arg("path"), py::arg("t"), py::arg("step"))
      .def("restore", &Solver::restore, py::arg("path"))
      .def("initialize", &Solver::initialize)
      .def_readonly("dt", &Solver::dt)
      .def_readwrite("tmax", &Solver::tmax)
      .def_readonly("startTimeLoop", &Solver::startTimeLoop);
}
