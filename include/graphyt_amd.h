/*
 * graphyt_amd.h — C ABI of the MI355X (gfx950) MPM time step.
 *
 * This is the drop-in boundary for Graphyt's `Solver::iterate_GPU(double t, int step,
 * Parameters*)` slot (Solver.hpp:21, Solver.i:16, stub at Solver.cpp:293-302): plain C,
 * POD arguments, int status codes, no torch or HIP types in any signature. Host code
 * (the C++ `graphyt` classes, a ctypes/cgo/JNI binding) owns FP64 host arrays; the
 * context owns FP32 device state and replays the step on its own HIP stream.
 *
 * Field layouts follow the reference: vectors interleaved per particle
 * (`pos[3p+k]`), symmetric tensors as [xx, yy, zz, xy, yz, xz]
 * (MaterialPoints.cpp:312-318, DamageModels.cpp:73-81), nodes numbered
 * n = i + j*Nx + k*Nx*Ny (Nodes.cpp:77-89).
 */
#ifndef GRAPHYT_AMD_H
#define GRAPHYT_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes --------------------------------------------------------------- */
#define GY_OK              0
#define GY_ERR_INVALID     1   /* bad argument or inconsistent sizes                   */
#define GY_ERR_UNSUPPORTED 2   /* a request outside the implemented path: more than 16
                                  materials, psep > cellsize, a model id outside 1..5,
                                  mixed 3D and plane instances, > 2^31 particles      */
#define GY_ERR_HIP         3   /* a HIP runtime call failed                            */
#define GY_ERR_RANGE       4   /* a particle's GIMP stencil left the grid              */
#define GY_ERR_STATE       5   /* call order violated (e.g. iterate before upload)     */

/* ---- boundary-condition codes: BoundaryConditions.hpp:6-77 (values are ABI) ------- */
enum gy_bc_type {
  GY_BC_NONE = 0,
  GY_BC_GRID_VX = 1, GY_BC_GRID_VY = 2, GY_BC_GRID_VZ = 3, GY_BC_GRID_TEMP = 4,
  GY_BC_GRID_TFLUX_X = 5, GY_BC_GRID_TFLUX_Y = 6, GY_BC_GRID_TFLUX_Z = 7,
  GY_BC_GRID_FX = 8, GY_BC_GRID_FY = 9, GY_BC_GRID_FZ = 10,
  GY_BC_GRID_VX_VY = 11, GY_BC_GRID_VX_VZ = 12, GY_BC_GRID_VY_VZ = 13,
  GY_BC_GRID_FX_FY = 14, GY_BC_GRID_FY_FZ = 15, GY_BC_GRID_FX_FZ = 16, GY_BC_GRID_FX_FY_FZ = 17,
  GY_BC_GRID_VX_VY_VZ = 18,
  /* 19..32: grid_*_temp variants (no-ops in the reference) */
  GY_BC_MP_VX = 33, GY_BC_MP_VY = 34, GY_BC_MP_VZ = 35, GY_BC_MP_TEMP = 36,
  GY_BC_MP_FX = 37, GY_BC_MP_FY = 38, GY_BC_MP_FZ = 39,
  GY_BC_MP_VX_VY = 40, GY_BC_MP_VX_VZ = 41, GY_BC_MP_VY_VZ = 42,
  GY_BC_MP_FX_FY = 43, GY_BC_MP_FY_FZ = 44, GY_BC_MP_FX_FZ = 45, GY_BC_MP_FX_FY_FZ = 46,
  GY_BC_MP_VX_VY_VZ = 47,
  /* 48..61: mp_*_temp variants (no-ops in the reference) */
  GY_BC_MP_STRESS = 62, GY_BC_MP_STRESS_XX_YY = 63, GY_BC_MPSTRAIN = 64,
  GY_BC_MP_VXSIN = 65, GY_BC_MP_VYSIN = 66, GY_BC_MP_SYYSIN = 67
};

/* ---- plain-old-data descriptions ------------------------------------------------- */

/* Background grid: Nodes(L, cellsize) (Nodes.cpp:8-141). Node counts follow
 * Nx = int(1 + L/h) (Nodes.cpp:71-73); L/h must be an integer (Nodes.cpp:12 vs 71). */
typedef struct gy_grid {
  double L[3];
  double cellsize;
} gy_grid;

/* One MaterialModels entry (MaterialModels.hpp:5-48, defaults MaterialModels.cpp:7-25). */
typedef struct gy_material {
  int32_t model;          /* 1 elastic, 2 Newtonian, 3 perfect plastic, 4 Drucker-Prager,
                             5 turbulent liquid (its flag lives in the damage field);
                             all five run in 3D and plane contexts */
  int32_t is_rigid;
  int32_t is_damage;
  int32_t damage_model;   /* 1 = Grady-Kipp */
  double  bulk_mod;
  double  shear_mod;
  double  density;
  double  viscosity;
  double  yield_stress;
  double  cohesion;
  double  critical_strain;
  double  phi_deg;        /* friction angle, degrees */
  double  psi_deg;        /* dilatancy angle, degrees */
  double  crack_speed;    /* Cg */
  double  damage_m;
  double  damage_k;
  double  thermal_conductivity;   /* heat transfer (Extrapolation.cpp:281-283, 359-362) runs
                                     when any uploaded particle temperature is nonzero */
  double  specific_heat;
  double  turb_stress;
  /* Drucker-Prager apex guard (not in the reference): when nonzero, the products
   * cval*(xi/eta_bar) and dcval*(xi/eta_bar) of the apex return
   * (ConstitutiveModels.cpp:364-365) are taken as 0 when cval (dcval) is 0, so a
   * cohesionless, non-dilatant material (c = 0, psi = 0) returns to sigma = 0 instead of
   * the reference's 0*inf = NaN. 0 = reference behaviour. */
  int32_t dp_apex_guard;
  /* Drucker-Prager multipliers (not in the reference; model 4 only, 0 for others):
   * GY_DP_RETURN_CLOSED (0) Newton to its fixed point, GY_DP_RETURN_REFERENCE (1) the
   * reference's loop control (ConstitutiveModels.cpp:321-343, 353-374: up to 201
   * passes), GY_DP_RETURN_LEARNED (2) the network of gy_set_return_network (3D
   * single-body contexts, is_damage = 0). Modes 1 and 2 run in 3D single-body contexts. */
  int32_t dp_return;
} gy_material;

/* Parameters (Parameters.hpp:16-32) plus the particle spacing MaterialPoints::psep and
 * the step size Solver::dt (Solver.cpp:32 or Solver::setDt). */
typedef struct gy_params {
  double  gravity[3];
  double  damping;        /* velocityDampingCoef */
  double  mu;             /* contact friction coefficient */
  double  cfl;
  double  psep;
  double  dt;
  int32_t is3D;           /* Parameters::is3D: 1 = the 3D step; 0 = the plane step per unit
                             depth (L[2] = 0, one node layer, 9-node stencil, plane model
                             branches, Extrapolation.cpp:90-124) */
  int32_t reserved;
} gy_params;

/* Host view of particle fields; any pointer may be NULL (field skipped on upload or
 * download; defaults on upload are the values MaterialPoints::add pushes, i.e. zero,
 * material 0, object 0, no hold, no BC). Arrays are indexed by the caller's particle
 * order, which the context preserves across steps. */
typedef struct gy_particles {
  int32_t  n;
  int32_t  reserved;
  double  *pos;                /* 3n */
  double  *pos0;               /* 3n: reference position for displacement       */
  double  *vel;                /* 3n */
  double  *stress;             /* 6n */
  double  *strain;             /* 6n */
  double  *plastic_strain;     /* 6n */
  double  *acc_plastic_strain; /* n  */
  double  *density;            /* n  */
  double  *mass;               /* n  */
  double  *damage;             /* n  */
  double  *smax;               /* n  */
  int32_t *material;           /* n  */
  int32_t *obj_id;             /* n  */
  uint8_t *hold;               /* n  */
  int32_t *bc_type;            /* n  */
  double  *bc_val;             /* 4n */
  /* download-only derived fields */
  double  *volume;             /* n  : mass / density                       */
  double  *displacement;       /* 3n : pos - pos0                           */
  int32_t *cell_id;            /* n  : reference cellID (TimeIntegration.cpp:110-115) */
  /* appended after the derived fields (layout of the fields above unchanged) */
  double  *temperature;        /* n  : MaterialPoints::temperature (setTemp); heat
                                *      transfer runs when any uploaded value is nonzero */
  double  *temperature_grad;   /* 3n : MaterialPoints::temperatureGrad of the last G2P
                                *      (0 after MaterialPoints::add); kept for restarts */
} gy_particles;

/* FP64 audit of the device state (sums over all particles of every instance). */
typedef struct gy_audit_out {
  double  mass;
  double  momentum[3];
  double  kinetic_energy;
  double  strain_energy_proxy;  /* 0.5 * V * sigma:sigma / (2G): bounded check value */
  int64_t nonfinite;            /* particles with a NaN/Inf in x, v or sigma */
  int64_t out_of_range;         /* particles whose stencil left the grid (clamped) */
} gy_audit_out;

typedef struct gy_ctx gy_ctx;

/* ---- lifecycle -------------------------------------------------------------------- */

/* Solver::Solver + Solver::initialize (Solver.cpp:7-69) for one instance. `params->dt`
 * <= 0 selects the reference rule dt = CFL*h/sqrt(maxK/minRho) (Solver.cpp:15-32). */
int  gy_create(gy_ctx **out, const gy_grid *grid, const gy_material *mats, int32_t nmat,
               const gy_params *params);

/* Nodes::setBC for every node at once (Nodes.cpp:155-163): bc_type[n], bc_val[4n..4n+3]. */
int  gy_set_node_bc(gy_ctx *ctx, int32_t n_nodes, const int32_t *bc_type, const double *bc_val);

/* Upload particle state (MaterialPoints::add + setters, MaterialPoints.cpp:95-339).
 * If `mass` or `density` is NULL they are set as Solver::initialize does
 * (rho = rho_mat, V = psep^3, m = rho*V, rigid m = 1; Solver.cpp:42-56). */
int  gy_upload(gy_ctx *ctx, const gy_particles *host);

/* Solver::setDt (Solver.cpp:70-73). */
int  gy_set_dt(gy_ctx *ctx, double dt);
double gy_get_dt(const gy_ctx *ctx);

/* One Solver::iterate_CPU-equivalent step (Solver.cpp:86-292), enqueued on the context
 * stream; returns without waiting. */
int  gy_iterate(gy_ctx *ctx, double t, int32_t step);

/* `nsteps` steps with t = t0 + i*dt, step = step0 + i (Solver::iterate_CPU_steps,
 * Solver.cpp:75-84, with the time passed explicitly). Within one call of a single-body
 * context without heat transfer, each later step's grid update runs in the same launch as
 * the step before's binning scan; between calls the state is the same as after as many
 * gy_iterate calls, and the results agree with them to rounding. */
int  gy_iterate_n(gy_ctx *ctx, double t0, int32_t step0, int32_t nsteps);

/* Wait for the context stream; reports a deferred device-side error (GY_ERR_RANGE). */
int  gy_synchronize(gy_ctx *ctx);

/* Download particle fields in the caller's order: FP32 device state widened to FP64;
 * positions are 48-bit fixed point (h/2^(F+16)) and velocities FP32 plus a bfloat16
 * residual, both exact in FP64. */
int  gy_download(gy_ctx *ctx, gy_particles *host);

/* Download the node fields of the last step (Nodes::getMassArr / getVelArr /
 * getForceArr, Nodes.cpp:146-149): mass (n_nodes), centre-of-mass velocity (3 n_nodes)
 * and force (3 n_nodes) as m * a, with a the nodal acceleration after the velocity BCs
 * (the reference keeps the force from before them). Any pointer may be NULL. */
int  gy_download_nodes(gy_ctx *ctx, int32_t n_nodes, double *mass, double *vel, double *force);

/* FP64 mass / momentum / energy audit reduction on the device. */
int  gy_audit(gy_ctx *ctx, gy_audit_out *out);

/* Binning of the current particle positions (the counting sort the next step uses):
 * keys[s] for device slot s, counts/offsets per key (nkeys entries), and the stable
 * order (order[i] = slot). Any pointer may be NULL; *nkeys receives the key count. */
int  gy_debug_binning(gy_ctx *ctx, int32_t *keys, int32_t *counts, int32_t *offsets,
                      int32_t *order, int64_t *nkeys, int64_t *nslots);

/* Standalone stable counting sort of host keys in [0, nkeys) with the same kernels the
 * step uses: counts, exclusive offsets and order (order[i] = index of the i-th key). */
int  gy_counting_sort(const int32_t *keys, int64_t n, int64_t nkeys,
                      int32_t *counts, int32_t *offsets, int32_t *order);

/* Zero-allocation variant of gy_counting_sort. No counterpart in the reference: it is for
 * callers that cannot allocate device memory — code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc. `keys` (n), `counts`, `offsets`
 * (nkeys each) and `order` (n) are DEVICE pointers (counts, offsets, order may be NULL);
 * `stream` is a hipStream_t (NULL: the null stream). The call only enqueues work on
 * `stream`: no allocation, no free, no synchronisation; results are ready when the stream
 * is. Keys outside [0, nkeys) are skipped (neither counted nor placed). Limits:
 * n <= 1,048,576 and nkeys <= 1,048,576 (GY_ERR_INVALID beyond them). Its scratch is a
 * library-owned device array of 16.8 MB that each GPU it runs on holds until the process
 * ends; every call on a GPU reuses that GPU's array, so calls that overlap on one GPU
 * (two streams, two threads) race on it and must be serialised by the caller. */
int  gy_counting_sort_noalloc(const int32_t *keys, int64_t n, int64_t nkeys, int32_t *counts,
                              int32_t *offsets, int32_t *order, void *stream);

/* Snapshot of the device state between two steps (no counterpart in the reference, which
 * has no checkpoint): every buffer the next step reads that earlier steps wrote — the
 * current particle buffer, the binning, the node records, the tile partials of this step,
 * the control block and the previous step's time (read by the sine particle BCs), the
 * contact and heat buffers. gy_snapshot_load puts it back bit for bit into a context that
 * was created and uploaded with the same setup (grid, materials, parameters, particle
 * count and modes); gy_iterate then continues from it. The step itself is reproducible to
 * rounding only (the order of its atomic adds varies between runs), so a continuation
 * follows the original run as closely as two runs of it follow each other. */
int  gy_snapshot_size(gy_ctx *ctx, int64_t *bytes);
int  gy_snapshot_save(gy_ctx *ctx, void *buf, int64_t bytes);
int  gy_snapshot_load(gy_ctx *ctx, const void *buf, int64_t bytes);

/* ---- batched ensembles ----------------------------------------------------------
 * No counterpart in the reference, which runs one Solver per problem: `ninst`
 * independent problems advanced together by every gy_iterate / gy_iterate_n call, in
 * the same launches (one persistent particle kernel over the tiles of all instances,
 * one captured graph per step). Instance i has grid grids[i], the nmat[i] materials that
 * follow the previous instances' in `mats`, and parameters params[i] (its own dt and
 * gravity). Material ids in a particle view are local to its instance. */
int  gy_create_batch(gy_ctx **out, int32_t ninst, const gy_grid *grids, const gy_material *mats,
                     const int32_t *nmat, const gy_params *params);
int  gy_instance_count(const gy_ctx *ctx);
int  gy_set_node_bc_inst(gy_ctx *ctx, int32_t inst, int32_t n_nodes, const int32_t *type, const double *val);
int  gy_upload_batch(gy_ctx *ctx, int32_t ninst, const gy_particles *views);
int  gy_download_inst(gy_ctx *ctx, int32_t inst, gy_particles *host);
double gy_get_dt_inst(const gy_ctx *ctx, int32_t inst);

/* Autotuner (no counterpart in the reference, whose OpenMP schedule is fixed). Launch
 * shapes of the step, per workload: */
typedef struct gy_tuning {
  int64_t persist_wgs;      /* workgroups of the persistent particle kernel               */
  int32_t fused_waves;      /* build of the 3D single-body fused kernel: 3 resident waves
                               per SIMD (168 VGPRs) or 2 (256 VGPRs)                       */
  int32_t grid_wgs_factor;  /* grid-kernel workgroups per persistent workgroup: 1, 2 or 4  */
  double  ms_per_step;      /* the trial time of this choice (tuning runs)                 */
  int32_t from_cache;       /* 1: read from the cache file, no trials run                  */
  int32_t reserved;
} gy_tuning;
/* Times `trial_steps` steps for each candidate (workgroups 1, 2, 3, 4, 6 per CU x both
 * builds, then the grid factor), every trial from the same device state (a snapshot that
 * is put back afterwards), and applies the fastest. With a cache file (may be NULL) a
 * previous result for the same workload shape (device, instances, particles, nodes, tiles,
 * model set, contact and heat modes) is applied without trials, and a new result is added
 * to it (one text line per shape). */
int  gy_autotune_shapes(gy_ctx *ctx, int32_t trial_steps, const char *cache_path, gy_tuning *out);
int  gy_get_tuning(const gy_ctx *ctx, gy_tuning *out);
int  gy_set_tuning(gy_ctx *ctx, const gy_tuning *t);
/* The same tuner without a cache, reporting only the workgroup count and time per step. */
int  gy_autotune(gy_ctx *ctx, int32_t trial_steps, int64_t *best_persist_wgs, double *best_ms_per_step);

/* ---- learned Drucker-Prager return map -------------------------------------------
 * No counterpart in the reference. A 14-64-64-64-1 SiLU network (BF16 matrix-core
 * products, FP32 accumulation), trained on the GPU on samples of the return map
 * (druckerPrager_Damage, ConstitutiveModels.cpp:238-428), gives the multiplier of the
 * cone return; the new stress (6) and accumulated plastic strain (1) are rebuilt from it
 * along the exact return direction, and the apex return is exact.
 * Constitutive sample (GY_DP_IN floats): strain increment xx yy zz xy yz xz, prior stress
 * (same order), accumulated plastic strain. Map output (GY_DP_OUT floats): new stress,
 * new accumulated plastic strain.
 * Network input (GY_MLP_IN floats): the sample, then the material's friction angle in
 * degrees, each normalised as (x - mean) / std. Network output (GY_MLP_OUT): y, with the
 * cone multiplier dL = (out_scale * y + out_offset) * F / (G + K eta eta_bar), F the
 * trial yield function: the multiplier relative to that of a cone without hardening. */
#define GY_DP_IN       13
#define GY_DP_OUT      7
#define GY_MLP_IN      14
#define GY_MLP_OUT     1
#define GY_MLP_PARAMS  10448   /* padded layout: W1 16x64, W2 64x64, W3 64x64, W4 64x16, biases */
typedef struct gy_mlp gy_mlp;
int  gy_mlp_create(gy_mlp **out, uint64_t seed);
int  gy_mlp_param_count(void);
int  gy_mlp_set_params(gy_mlp *m, const float *params);
int  gy_mlp_get_params(gy_mlp *m, float *params);
/* normalisation (GY_MLP_IN each), the per-input range [lo, hi] (normalised units) of the
 * gate, the output's scale and offset, and the range [out_lo, out_hi] of the scaled output
 * the gate accepts (outside it, or where the return's apex/cone branch would differ
 * across it, the sample takes the closed form) */
int  gy_mlp_set_norm(gy_mlp *m, const float *in_mean, const float *in_std, const float *in_lo,
                     const float *in_hi, float out_scale, float out_offset, float out_lo, float out_hi);
/* network on normalised inputs x[n][GY_MLP_IN] -> y[n][GY_MLP_OUT] */
int  gy_mlp_forward(gy_mlp *m, const float *x, int64_t n, float *y);
/* gradient of mean((y - t)^2) over the n samples w.r.t. the GY_MLP_PARAMS parameters */
int  gy_mlp_gradient(gy_mlp *m, const float *x, const float *t, int64_t n, float *grad, double *loss);
/* training set (normalised) kept on the device; Adam over mini-batches */
int  gy_mlp_load_data(gy_mlp *m, const float *x, const float *t, int64_t n);
int  gy_mlp_train(gy_mlp *m, int32_t epochs, int32_t batch, double lr, double lr_decay, double *last_loss);
/* Return-map modes (gy_dp_map and gy_material::dp_return). */
#define GY_DP_RETURN_CLOSED     0
#define GY_DP_RETURN_REFERENCE  1
#define GY_DP_RETURN_LEARNED    2
/* The return map over n raw samples of material mat (model 4; strain increments, so
 * dt = 1): in[n][GY_DP_IN] -> out[n][GY_DP_OUT]. multiplier (optional, n floats): the
 * cone multiplier each sample used, 0 for elastic samples. Mode GY_DP_RETURN_LEARNED
 * needs m: elastic trial states exactly, yielding samples through the network when the
 * gate accepts them (gy_mlp_set_norm), the others in closed form; n_network and
 * n_fallback count them.
 * kernel_ms: the kernel's device time. */
int  gy_dp_map(const gy_material *mat, gy_mlp *m, int32_t mode, const float *in, float *out, float *multiplier,
               int64_t n, int64_t *n_network, int64_t *n_fallback, double *kernel_ms);
const char *gy_mlp_last_error(const gy_mlp *m);
void gy_mlp_destroy(gy_mlp *m);

/* The network of the materials with dp_return = GY_DP_RETURN_LEARNED, in the step: a copy
 * of m's parameters and normalisation as they are now. Needed before gy_iterate when
 * particles of such a material are uploaded. No counterpart in the reference. */
int  gy_set_return_network(gy_ctx *ctx, gy_mlp *m);
/* Yielding samples of learned materials that took the network, and those the gate sent
 * to the closed form, since the upload or the last reset. */
int  gy_return_map_counts(gy_ctx *ctx, int64_t *network, int64_t *fallback, int32_t reset);

/* Options: see GY_OPT_*. */
#define GY_OPT_USE_GRAPH   1   /* replay gy_iterate_n through a captured hipGraph (default 1) */
#define GY_OPT_PROFILE     2   /* record HIP events around every kernel (default 0)          */
#define GY_OPT_TILE        3   /* cell-tile edge of the particle kernels (4 is built)         */
/* option 4 is not assigned */
#define GY_OPT_PERSIST_WGS 5   /* workgroups of the persistent particle kernel (default 3/CU) */
int  gy_set_option(gy_ctx *ctx, int32_t option, int64_t value);

/* Per-kernel accumulated device time since the last reset (GY_OPT_PROFILE on).
 * `name` is one of the kernel names listed by gy_kernel_name(i). */
int  gy_kernel_count(void);
const char *gy_kernel_name(int32_t i);
int  gy_kernel_stats(gy_ctx *ctx, int32_t i, int64_t *launches, double *total_ms);
int  gy_reset_kernel_stats(gy_ctx *ctx);

/* The context's HIP stream (a hipStream_t), for callers that time on it. */
void *gy_stream(gy_ctx *ctx);

const char *gy_last_error(const gy_ctx *ctx);
const char *gy_version(void);
void gy_destroy(gy_ctx *ctx);

#ifdef __cplusplus
}
#endif
#endif /* GRAPHYT_AMD_H */
