// gy_models.hpp — per-particle constitutive and damage models on the device, shared by
// the MPM step kernels and the return-map surrogate kernels.
#pragma once
#include <hip/hip_runtime.h>

#include "gy_internal.hpp"

namespace gy {
namespace {

// ---------------------------------------------------------------------------------------
// Constitutive models (FP32), following ConstitutiveModels.cpp operation by operation.

// elastic, ConstitutiveModels.cpp:3-30
__device__ __forceinline__ void cm_elastic(const MatDev &M, const float er[6], float s[6], float dt) {
  const float vol = 0.3333f * (er[0] + er[1] + er[2]);
  const float G2 = 2.0f * M.G, K3 = 3.0f * M.K;
  s[0] += (G2 * (er[0] - vol) + K3 * vol) * dt;
  s[1] += (G2 * (er[1] - vol) + K3 * vol) * dt;
  s[2] += (G2 * (er[2] - vol) + K3 * vol) * dt;
  s[3] += (G2 * er[3]) * dt;
  s[4] += (G2 * er[4]) * dt;
  s[5] += (G2 * er[5]) * dt;
}

// newtonianLiquid, ConstitutiveModels.cpp:33-64. PLANE (Parameters::is3D false): sigma_zz,
// _yz, _zx keep their values (ConstitutiveModels.cpp:59-63).
template <bool PLANE = false>
__device__ __forceinline__ void cm_newtonian(const MatDev &M, const float er[6], const float e[6],
                                             float s[6], float dt) {
  const float vol = 0.33333f * (er[0] + er[1] + er[2]);
  const float P = M.K * (e[0] + e[1] + e[2]);
  const float c = M.visc * (2.0f - (2.0f / 3.0f));
  s[0] = P + c * (er[0] - vol) * dt;
  s[1] = P + c * (er[1] - vol) * dt;
  if (!PLANE) s[2] = P + c * (er[2] - vol) * dt;
  s[3] = 2.0f * M.visc * er[3] * dt;
  if (PLANE) return;
  s[4] = 2.0f * M.visc * er[4] * dt;
  s[5] = 2.0f * M.visc * er[5] * dt;
}

// turbulentLiquid, ConstitutiveModels.cpp:66-126. As there, the deviatoric rate is formed
// while the volumetric part is still zero, so it is the full rate; the damage field holds
// the turbulence flag: a set flag halves the viscosity of this update, then the flag is
// set when KE - VE - turbStress > 0 and cleared (with the viscosity restored) otherwise.
// KE = m |v|^2 / 2 with the velocity after this step's update, VE = mu (shear rates^2) / 2.
// PLANE: sigma_zz, _yz, _zx keep their values (ConstitutiveModels.cpp:121-125).
template <bool PLANE = false>
__device__ __forceinline__ void cm_turbulent(const MatDev &M, const float er[6], const float e[6],
                                             const float v[3], float mass, float &dmg, float s[6], float dt) {
  float mu = M.visc;
  if (dmg == 1.0f) mu = M.visc / 2.0f;
  const float KE = 0.5f * mass * ((v[0] * v[0]) + (v[1] * v[1]) + (v[2] * v[2]));
  const float VE = 0.5f * M.visc * (er[3] * er[3] + er[4] * er[4] + er[5] * er[5]);
  const float crit = KE - VE - M.turb;
  if (crit > 0.0f) dmg = 1.0f;
  if (crit <= 0.0f) {
    dmg = 0.0f;
    mu = M.visc;
  }
  const float P = M.K * (e[0] + e[1] + e[2]);
  const float c = mu * (2.0f - (2.0f / 3.0f));
  s[0] = P + c * er[0] * dt;
  s[1] = P + c * er[1] * dt;
  s[3] = 2.0f * mu * er[3] * dt;
  if (PLANE) return;
  s[2] = P + c * er[2] * dt;
  s[4] = 2.0f * mu * er[4] * dt;
  s[5] = 2.0f * mu * er[5] * dt;
}

// perfectPlasticity, ConstitutiveModels.cpp:131-234. PLANE: the trial sigma_zz, _yz, _zx
// are zero (ConstitutiveModels.cpp:154-160), so the elastic branch writes zeros there and
// the plastic branch leaves them (ConstitutiveModels.cpp:217-222).
template <bool PLANE = false>
__device__ __forceinline__ void cm_perfect(const MatDev &M, const float er[6], float s[6], float p[6],
                                           float &epb, float dt) {
  const float G = M.G, K = M.K;
  const float vol = 0.33f * (er[0] + er[1] + er[2]);
  float tr[6];
  tr[0] = s[0] + (2.0f * G * (er[0] - vol) + 3.0f * K * vol) * dt;
  tr[1] = s[1] + (2.0f * G * (er[1] - vol) + 3.0f * K * vol) * dt;
  tr[2] = PLANE ? 0.0f : s[2] + (2.0f * G * (er[2] - vol) + 3.0f * K * vol) * dt;
  tr[3] = s[3] + (2.0f * G * er[3]) * dt;
  tr[4] = PLANE ? 0.0f : s[4] + (2.0f * G * er[4]) * dt;
  tr[5] = PLANE ? 0.0f : s[5] + (2.0f * G * er[5]) * dt;
  const float a01 = tr[0] - tr[1], a12 = tr[1] - tr[2], a20 = tr[2] - tr[0];
  const float q = sqrtf(0.5f * a01 * a01 + 0.5f * a12 * a12 + 0.5f * a20 * a20 + tr[3] * tr[3] +
                        tr[4] * tr[4] + tr[5] * tr[5]);
  if (!(q > M.sy)) {
    for (int k = 0; k < 6; k++) s[k] = tr[k];
    return;
  }
  const float dL = (q - M.sy) / (3.0f * G);
  const float f = dL * (0.5f / q);
  float pr[6];
  pr[0] = f * (2.0f * tr[0] - tr[1] - tr[2]);
  pr[1] = f * (2.0f * tr[1] - tr[2] - tr[0]);
  pr[2] = f * (2.0f * tr[2] - tr[0] - tr[1]);
  pr[3] = f * (6.0f * tr[3]);
  pr[4] = f * (6.0f * tr[4]);
  pr[5] = f * (6.0f * tr[5]);
  const float b01 = pr[0] - pr[1], b12 = pr[1] - pr[2], b20 = pr[2] - pr[0];
  epb += dt * sqrtf(0.5f * b01 * b01 + 0.5f * b12 * b12 + 0.5f * b20 * b20 + pr[3] * pr[3] +
                    pr[4] * pr[4] + pr[5] * pr[5]);
  const float inv_dt = 1.0f / dt;
  float el[6];
  for (int k = 0; k < 6; k++) {
    p[k] += pr[k];
    el[k] = er[k] - pr[k] * inv_dt;
  }
  s[0] += (2.0f * G * (el[0] - vol) + 3.0f * K * vol) * dt;
  s[1] += (2.0f * G * (el[1] - vol) + 3.0f * K * vol) * dt;
  s[3] += (2.0f * G * el[3]) * dt;
  if (PLANE) return;
  s[2] += (2.0f * G * (el[2] - vol) + 3.0f * K * vol) * dt;
  s[4] += (2.0f * G * el[4]) * dt;
  s[5] += (2.0f * G * el[5]) * dt;
}

// druckerPrager_Damage, ConstitutiveModels.cpp:238-428, in three parts: the elastic
// trial state and yield test (dp_trial), the multipliers of the cone and apex returns
// (Newton on piecewise-linear consistency conditions), and the return from the trial
// state given the cone multiplier (dp_return). Multipliers, per material
// (gy_material::dp_return):
//   GY_DP_RETURN_CLOSED     Newton to the iterate's fixed point (at most 16 passes): the
//                           reference's root to rounding;
//   GY_DP_RETURN_REFERENCE  the reference's loop control, while |Fp/cval| > 1e-8 and at
//                           most 201 passes (with c = 0 the test divides by cval = 0 and
//                           every pass runs);
//   GY_DP_RETURN_LEARNED    the cone multiplier from the network (gy_kernels.hip), the
//                           apex multiplier as GY_DP_RETURN_CLOSED.
struct DpTrial {
  float tr[6];     // elastic trial stress (ConstitutiveModels.cpp:270-276)
  float I1, J2, sJ2, F;
  bool yield;
};

__device__ __forceinline__ void dp_trial(const MatDev &M, const float er[6], const float s[6], float dt,
                                         DpTrial &T) {
  const float G = M.G, K = M.K;
  const float vol = 0.3333f * (er[0] + er[1] + er[2]);
  T.tr[0] = s[0] + (2.0f * G * (er[0] - vol) + 3.0f * K * vol) * dt;
  T.tr[1] = s[1] + (2.0f * G * (er[1] - vol) + 3.0f * K * vol) * dt;
  T.tr[2] = s[2] + (2.0f * G * (er[2] - vol) + 3.0f * K * vol) * dt;
  T.tr[3] = s[3] + (2.0f * G * er[3]) * dt;
  T.tr[4] = s[4] + (2.0f * G * er[4]) * dt;
  T.tr[5] = s[5] + (2.0f * G * er[5]) * dt;
  T.I1 = (1.0f / 3.0f) * (T.tr[0] + T.tr[1] + T.tr[2]);
  const float sxx = T.tr[0] - T.I1, syy = T.tr[1] - T.I1, szz = T.tr[2] - T.I1;
  T.J2 = 0.5f * (sxx * sxx + syy * syy + szz * szz + 2.0f * T.tr[3] * T.tr[3] + 2.0f * T.tr[4] * T.tr[4] +
                 2.0f * T.tr[5] * T.tr[5]);
  T.sJ2 = sqrtf(T.J2);
  T.F = T.sJ2 + M.eta * T.I1 - M.xi * M.c;
  T.yield = M.damage ? ((T.F > 0.0f) && (T.I1 < 0.0f)) : (T.F > 0.0f);
}

// Cone multiplier dL (ConstitutiveModels.cpp:313-343); ep0: accumulated plastic strain
// before the step.
template <bool REF_LOOP>
__device__ __forceinline__ float dp_cone_multiplier(const MatDev &M, const DpTrial &T, float ep0) {
  const float G = M.G, K = M.K;
  const float slope0 = -G - K * M.eta * M.eta_bar;
  float dL = 0.0f;
  if (REF_LOOP) {
    int it = 0;
    float err = 1.0f;
    while (err > 1e-8f) {
      it++;
      float cv = M.c - (ep0 + M.xi * dL) * M.c_over_ecr;
      float dcv = -M.c_over_ecr;
      if (cv < M.c_floor) { cv = M.c_floor; dcv = 0.0f; }
      const float Fp = T.sJ2 - G * dL + M.eta * (T.I1 - M.eta_bar * K * dL) - M.xi * cv;
      const float dFp = slope0 - M.xi * M.xi * dcv;
      dL = dL - Fp / dFp;
      err = fabsf(Fp / cv);
      if (it > 200) break;
    }
    return dL;
  }
  for (int it = 0; it < 16; it++) {
    float cv = M.c - (ep0 + M.xi * dL) * M.c_over_ecr;
    float dcv = -M.c_over_ecr;
    if (cv < M.c_floor) { cv = M.c_floor; dcv = 0.0f; }
    const float Fp = T.sJ2 - G * dL + M.eta * (T.I1 - M.eta_bar * K * dL) - M.xi * cv;
    const float dFp = slope0 - M.xi * M.xi * dcv;
    const float dLn = dL - Fp / dFp;
    const bool done = (dLn == dL) || !(dLn == dLn);
    dL = dLn;
    if (done) break;
  }
  return dL;
}

// Apex multiplier deps (ConstitutiveModels.cpp:347-375).
template <bool REF_LOOP>
__device__ __forceinline__ float dp_apex_multiplier(const MatDev &M, const DpTrial &T, float ep0) {
  const float K = M.K;
  float de = 0.0f;
  int it = 0;
  float err = 1.0f;
  for (;;) {
    if (REF_LOOP && !(err > 1e-8f)) break;
    if (!REF_LOOP && it == 16) break;
    it++;
    float cv = M.c - (ep0 + M.xi_over_eta * de) * M.c_over_ecr;
    float dcv = -M.c_over_ecr;
    if (cv < M.c_floor) { cv = M.c_floor; dcv = 0.0f; }
    float Gc = cv * M.xi_over_etab, dGc = M.xi_over_eta * dcv * M.xi_over_etab;
    if (M.apex_guard) {
      if (cv == 0.0f) Gc = 0.0f;
      if (dcv == 0.0f) dGc = 0.0f;
    }
    const float Gp = Gc - T.I1 + K * de;
    const float den = de - Gp / (dGc + K);
    if (REF_LOOP) {
      de = den;
      err = fabsf(Gp / cv);
      if (it > 200) break;
    } else {
      const bool done = (den == de) || !(den == den);
      de = den;
      if (done) break;
    }
  }
  return de;
}

// The return from a yielding trial state with cone multiplier dL
// (ConstitutiveModels.cpp:345-424): the apex return when sqrt(J2) - G dL < 0 (not for a
// damage material), the cone return otherwise. s holds the stress before the step and
// receives the new one; p and epb are updated. sqrt(J2) = 0 takes the apex return: a
// yielding state has F > 0 and so dL > 0, which makes the test negative, but a trial
// stress of ~1e-31 Pa (the velocity gradient of a body at rest, to rounding) leaves J2
// and F / G below the FP32 normal range, flushed to zero, and the cone return would
// divide 0 by sqrt(J2) = 0.
// Returns false when the apex multiplier is negative: s, p and epb are then left as they
// were (the only case in which s is read).
template <bool REF_LOOP>
__device__ __forceinline__ bool dp_return(const MatDev &M, const DpTrial &T, float dL, float s[6], float p[6],
                                          float &epb) {
  const float G = M.G, K = M.K;
  if ((T.sJ2 - G * dL < 0.0f || T.sJ2 == 0.0f) && !M.damage) {
    const float de = dp_apex_multiplier<REF_LOOP>(M, T, epb);
    if (de < 0.0f) return false;  // deps < 0: stress left as it was (ConstitutiveModels.cpp:376-378)
    const float sm = T.I1 - K * de;
    s[0] = sm; s[1] = sm; s[2] = sm; s[3] = 0.0f; s[4] = 0.0f; s[5] = 0.0f;
    epb += de * M.xi_over_eta;
    const float pv = de * (M.eta_bar / 3.0f);
    p[0] += pv; p[1] += pv; p[2] += pv;
    return true;
  }
  const float sxx = T.tr[0] - T.I1, syy = T.tr[1] - T.I1, szz = T.tr[2] - T.I1;
  const float r = 1.0f - G * dL / T.sJ2;
  const float sm = T.I1 - K * M.eta_bar * dL;
  s[0] = r * sxx + sm;
  s[1] = r * syy + sm;
  s[2] = r * szz + sm;
  s[3] = r * T.tr[3];
  s[4] = r * T.tr[4];
  s[5] = r * T.tr[5];
  epb += dL * M.xi;
  const float hj = 0.5f / T.J2, pv = M.eta_bar / 3.0f;
  p[0] += dL * (hj * sxx + pv);
  p[1] += dL * (hj * syy + pv);
  p[2] += dL * (hj * szz + pv);
  p[3] += dL * hj * T.tr[3];
  p[4] += dL * hj * T.tr[4];
  p[5] += dL * hj * T.tr[5];
  return true;
}

// The cone multiplier of a learned material from the network's scaled output yr (the
// multiplier over its linear estimate k = F / (G + K eta eta_bar)); negative when the gate
// sends the sample to the closed form:
//  * yr outside the trained output range [ylo, yhi] (an input combination the network
//    extrapolates on);
//  * the branch test sqrt(J2) - G dL < 0 of the return (ConstitutiveModels.cpp:346)
//    giving different answers at the two ends of that range: near the switch an error of
//    the network's size picks the other branch (at near-zero pressure a multiplier too
//    large by 1e-5 sends a cone return to the apex, which keeps the old stress).
__device__ __forceinline__ float dp_learned_multiplier(const MatDev &M, const DpTrial &T, float yr, float ylo,
                                                       float yhi) {
  if (!(yr >= ylo && yr <= yhi)) return -1.0f;
  const float k = T.F / (M.G + M.K * M.eta * M.eta_bar);
  const bool apex_lo = T.sJ2 - M.G * fmaxf(ylo * k, 0.0f) < 0.0f;
  const bool apex_hi = T.sJ2 - M.G * fmaxf(yhi * k, 0.0f) < 0.0f;
  if (apex_lo != apex_hi) return -1.0f;
  return fmaxf(yr * k, 0.0f);
}

// The whole map with the closed or (LOOP and the material asks for it) the reference's
// multipliers. A learned material reaching this function (outside the kernels that
// evaluate the network) takes the closed form.
template <bool LOOP = false>
__device__ __forceinline__ void cm_drucker_prager(const MatDev &M, const float er[6], float s[6], float p[6],
                                                  float &epb, float dt) {
  DpTrial T;
  dp_trial(M, er, s, dt, T);
  if (!T.yield) {
#pragma unroll
    for (int k = 0; k < 6; k++) s[k] = T.tr[k];
    return;
  }
  if (LOOP && M.dp_mode == GY_DP_RETURN_REFERENCE) {
    const float dL = dp_cone_multiplier<true>(M, T, epb);
    dp_return<true>(M, T, dL, s, p, epb);
  } else {
    const float dL = dp_cone_multiplier<false>(M, T, epb);
    dp_return<false>(M, T, dL, s, p, epb);
  }
}

// Symmetric 3x3 eigen-decomposition in FP64 by cyclic Jacobi rotations (Golub & Van Loan,
// Alg. 8.5.3); eigenvectors in the columns of V.
__device__ void jacobi3(double a[3][3], double V[3][3]) {
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) V[i][j] = (i == j) ? 1.0 : 0.0;
  for (int sweep = 0; sweep < 12; sweep++) {
    const double off = a[0][1] * a[0][1] + a[0][2] * a[0][2] + a[1][2] * a[1][2];
    const double dg = a[0][0] * a[0][0] + a[1][1] * a[1][1] + a[2][2] * a[2][2];
    if (off <= 1e-36 * dg || off == 0.0) break;
#pragma unroll
    for (int pq = 0; pq < 3; pq++) {
      const int p = (pq == 2) ? 1 : 0, q = (pq == 0) ? 1 : 2;
      if (a[p][q] == 0.0) continue;
      const double th = (a[q][q] - a[p][p]) / (2.0 * a[p][q]);
      const double t = (th >= 0.0 ? 1.0 : -1.0) / (fabs(th) + sqrt(th * th + 1.0));
      const double c = 1.0 / sqrt(t * t + 1.0), sn = t * c;
      for (int k = 0; k < 3; k++) {
        const double akp = a[k][p], akq = a[k][q];
        a[k][p] = c * akp - sn * akq;
        a[k][q] = sn * akp + c * akq;
      }
      for (int k = 0; k < 3; k++) {
        const double apk = a[p][k], aqk = a[q][k];
        a[p][k] = c * apk - sn * aqk;
        a[q][k] = sn * apk + c * aqk;
      }
      for (int k = 0; k < 3; k++) {
        const double vkp = V[k][p], vkq = V[k][q];
        V[k][p] = c * vkp - sn * vkq;
        V[k][q] = sn * vkp + c * vkq;
      }
    }
  }
}

// gradyKippDamage 3D branch (DamageModels.cpp:71-150), the whole block in FP64.
__device__ void gk_damage(const MatDev &M, float s[6], float &D, float &smax, float mass, float rho,
                          float dt) {
  double a[3][3] = {{s[0], s[3], s[5]}, {s[3], s[1], s[4]}, {s[5], s[4], s[2]}};
  const double check = fabs(a[0][0] + a[1][1] + a[2][2] + a[0][1] + a[1][2] + a[0][2]);
  if (!(check > 1e-3)) return;
  double V[3][3];
  jacobi3(a, V);
  double ev[3] = {a[0][0], a[1][1], a[2][2]};
  const double smx = fmax(ev[0], fmax(ev[1], fmax(ev[2], 0.0)));
  smax = (float)smx;
  if (smx > 1e-3) {
    const double eff = smx / M.gk_K43;
    const double vol = (double)mass / (double)rho;
    const double emin = pow(vol * M.gk_k, -1.0 / M.gk_m);
    if (eff > emin) {
      const double rate = M.gk_m3 * M.gk_alpha13 * pow(eff, M.gk_m / 3.0);
      double Dn = pow(pow((double)D, 1.0 / 3.0) + (double)dt * rate, 3.0);
      if (Dn > 1.0) Dn = 1.0;
      D = (float)Dn;
      for (int k = 0; k < 3; k++)
        if (ev[k] > 0.0) ev[k] = (1.0 - Dn) * ev[k];
    }
  }
  double o[6] = {0, 0, 0, 0, 0, 0};
  for (int c = 0; c < 3; c++) {
    const double n = sqrt(V[0][c] * V[0][c] + V[1][c] * V[1][c] + V[2][c] * V[2][c]);
    const double x = V[0][c] / n, y = V[1][c] / n, z = V[2][c] / n, l = ev[c];
    o[0] += x * l * x; o[1] += y * l * y; o[2] += z * l * z;
    o[3] += x * l * y; o[4] += y * l * z; o[5] += x * l * z;
  }
  for (int k = 0; k < 6; k++) s[k] = (float)o[k];
}

// gradyKippDamage 2D branch (DamageModels.cpp:11-69), in FP64: principal stresses of the
// in-plane tensor by the rotation angle theta (pi/4 when |sxx - syy| < 1e-5, with
// pi = 3.14159 as there), damage growth of the tensile ones, and sigma_zz = 0 whenever the
// stress is rebuilt (growth or complete damage); otherwise the stress is left as read.
__device__ __noinline__ void gk_damage_plane(const MatDev &M, float s[6], float &D, float &smax, float mass, float rho,
                                float dt) {
  const double sxx = s[0], syy = s[1], sxy = s[3];
  const double th = (fabs(sxx - syy) < 1e-5) ? 3.14159 / 4.0 : 0.5 * atan(2.0 * sxy / (sxx - syy));
  const double c = cos(th), sn = sin(th);
  double s1 = sxx * c * c + syy * sn * sn + 2.0 * c * sn * sxy;
  double s2 = sxx * sn * sn + syy * c * c - 2.0 * c * sn * sxy;
  smax = (float)fmax(s1, fmax(s2, 0.0));
  const double smx = fmax(s1, fmax(s2, 0.0));
  double Dn = D;
  bool rebuild = false;
  if (smx > 1e-3) {
    const double eff = smx / M.gk_K43;
    const double vol = (double)mass / (double)rho;
    const double emin = pow(vol * M.gk_k, -1.0 / M.gk_m);
    if (eff > emin) {
      const double rate = M.gk_m3 * M.gk_alpha13 * pow(eff, M.gk_m / 3.0);
      Dn = pow(pow(Dn, 1.0 / 3.0) + (double)dt * rate, 3.0);
      if (Dn > 1.0) Dn = 1.0;
      rebuild = true;
    }
  }
  if (Dn == 1.0) rebuild = true;
  D = (float)Dn;
  if (!rebuild) return;
  // the reference rebuilds once more when D is 1 after the growth; that second pass scales
  // principal stresses already scaled to zero and changes nothing
  if (s1 > 0.0) s1 = (1.0 - Dn) * s1;
  if (s2 > 0.0) s2 = (1.0 - Dn) * s2;
  s[0] = (float)(s1 * c * c + s2 * sn * sn);
  s[1] = (float)(s1 * sn * sn + s2 * c * c);
  s[3] = (float)((s1 - s2) * c * sn);
  s[2] = 0.0f;
}

}  // namespace
}  // namespace gy
