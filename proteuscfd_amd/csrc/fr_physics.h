// Finite-rate (FR) reacting-Euler state physics, per state / per face, FP64.
//
// State per node (non-dimensional): rho_1..rho_NS, u, v, w, T (NE = NS + 4
// values, the reference's native variables, ucs/compressibleFR.tcc:12-31).
// The reference also stores P, rho, cv_i and conc_i per node; of those only
// P and rho are read on this path, and both are pure functions of the native
// values (ComputeAuxiliaryVariables, :754-814), so they are recomputed where
// needed instead of being stored.
//
// Every function restates the reference's arithmetic in its order
// (citations per function); with FP contraction off the results are the
// reference's to the bit, transcendental-free as they are (square roots
// and divisions are correctly rounded on both sides).  The functions compile
// for the host as well, so the same code can be checked on the CPU.
#pragma once

#include <cmath>

#ifdef __HIPCC__
#include <hip/hip_runtime.h>
#define PCFD_HD __host__ __device__ __forceinline__
#else
#define PCFD_HD inline
#endif

namespace pcfd {
namespace fr {

#ifndef __HIPCC__
using std::fabs;
using std::isnan;
using std::sqrt;
#endif

constexpr double kUnivR = 8.31447215;  // J/(mol K), ucs/chem_constants.h:5
constexpr int kMaxS = 8;
constexpr int kSubit = 10;             // N_SUBIT, ucs/compressibleFR.tcc:32

// Model constants of one FR case (species thermo + the case's scales).
struct Model {
  int ns;
  double mw[kMaxS];      // kg/mol
  double R[kMaxS];       // UNIV_R / MW (Species::R, ucs/species.tcc:315)
  double lo[kMaxS][7];   // NASA-7, T <= 1000 K
  double hi[kMaxS][7];   // NASA-7, T > 1000 K
  double ref_density, ref_temperature, ref_pressure, ref_velocity;
  double vref2;          // ref_velocity^2 (ref_specific_enthalpy)
  double s_ref;          // ref_velocity^2 / ref_temperature
  double Pref;           // freestream pressure (UpdateQinf)
  int sym2d;
  int local_dt;          // localTimeStepping
  double dt_param;       // timeStep parameter
  int no_cvbc;           // noCVBC: upwind farfield, mirrored walls
};

template <int NS>
struct State {
  static constexpr int NE = NS + 4;
  double q[NE];
  double P, rho;  // aux
};

// Species::GetThermoCoeff (ucs/species.tcc:98-137): outside 200-6000 K the
// pinned set is the one the range test would pick anyway
PCFD_HD const double* coeffs(const Model& m, int s, double T) { return T > 1000.0 ? m.hi[s] : m.lo[s]; }

// Species::GetCp (ucs/species.tcc:43-53), J/(kg K)
PCFD_HD double cp_species(const Model& m, int s, double T) {
  const double* a = coeffs(m, s, T);
  const double cp_R = a[0] + T * (a[1] + T * (a[2] + T * (a[3] + T * a[4])));
  return cp_R * m.R[s];
}

// Species::GetH (ucs/species.tcc:56-70, href = 0), J/kg
PCFD_HD double h_species(const Model& m, int s, double T) {
  const double* a = coeffs(m, s, T);
  const double h_R = a[5] + T * (a[0] + T * (a[1] / 2.0 + T * (a[2] / 3.0 + T * (a[3] / 4.0 + T * a[4] / 5.0))));
  return h_R * m.R[s];
}

// Species::GetdHdT (ucs/species.tcc:337-348)
PCFD_HD double dhdt_species(const Model& m, int s, double T) {
  const double* a = coeffs(m, s, T);
  const double r = a[0] + T * (a[1] + T * (a[2] + T * (a[3] + T * a[4])));
  return r * m.R[s];
}

// ComputeAuxiliaryVariables (ucs/compressibleFR.tcc:754-814): P (Dalton's law
// over ideal-gas species, ChemModel::GetP, ucs/chem.tcc:988-997) and rho
template <int NS>
PCFD_HD void aux(const Model& m, State<NS>& s) {
  double rho = 0.0, P = 0.0;
  const double TDim = s.q[NS + 3] * m.ref_temperature;
#pragma unroll
  for (int i = 0; i < NS; ++i) rho += s.q[i];
#pragma unroll
  for (int i = 0; i < NS; ++i) P += (s.q[i] * m.ref_density) * m.R[i] * TDim;
  s.P = P / m.ref_pressure;
  s.rho = rho;
}

// GetFluidProperties (ucs/compressibleFR.tcc:1521-1570 over
// ChemModel::GetFluidProperties/GetCp/GetCv, ucs/chem.tcc:544-574,1001-1041)
template <int NS>
PCFD_HD void props(const Model& m, const double* rhoi, double T, double& cv, double& cp, double& R, double& gamma,
                   double& c2) {
  double rd[NS], rho = 0.0, Rd = 0.0;
#pragma unroll
  for (int i = 0; i < NS; ++i) rd[i] = rhoi[i] * m.ref_density;
  const double Td = T * m.ref_temperature;
#pragma unroll
  for (int i = 0; i < NS; ++i) rho += rd[i];
#pragma unroll
  for (int i = 0; i < NS; ++i) Rd += rd[i] * m.R[i];
  Rd /= rho;
  double cpd = 0.0, cvd = 0.0, cpi[NS];
#pragma unroll
  for (int i = 0; i < NS; ++i) cpi[i] = cp_species(m, i, Td);
#pragma unroll
  for (int i = 0; i < NS; ++i) cpd += cpi[i] * (rd[i] / rho);
#pragma unroll
  for (int i = 0; i < NS; ++i) cvd += (rd[i] / rho) * (cpi[i] - m.R[i]);
  gamma = cpd / cvd;
  const double c2d = gamma * Rd * Td;
  cv = cvd / m.s_ref;
  cp = cpd / m.s_ref;
  R = Rd / m.s_ref;
  c2 = c2d / m.vref2;
}

// GetTotalEnthalpy's specific enthalpy h / ref_specific_enthalpy
// (ucs/compressibleFR.tcc:1248-1273, ChemModel::GetSpecificEnthalpy ucs/chem.tcc:586-594)
template <int NS>
PCFD_HD double hmix(const Model& m, const State<NS>& s) {
  const double Td = s.q[NS + 3] * m.ref_temperature;
  double h = 0.0;
#pragma unroll
  for (int i = 0; i < NS; ++i) h += h_species(m, i, Td) * (s.q[i] / s.rho);
  return h / m.vref2;
}

template <int NS>
PCFD_HD double total_enthalpy(const Model& m, const State<NS>& s) {
  const double u = s.q[NS], v = s.q[NS + 1], w = s.q[NS + 2];
  return hmix(m, s) * s.rho + 0.5 * s.rho * (u * u + v * v + w * w);
}

// GetTheta (ucs/compressibleFR.tcc:1640-1643), vdotn = 0
template <int NS>
PCFD_HD double theta_of(const State<NS>& s, const double* a) {
  return s.q[NS] * a[0] + s.q[NS + 1] * a[1] + s.q[NS + 2] * a[2] + 0.0;
}

// HLLCFlux (ucs/compressibleFR.tcc:300-548) with vdotn = 0, NaN components
// zeroed as NumericalFlux does (ucs/eqnset.tcc:74-88).  Returns true when a
// component was zeroed.
template <int NS>
PCFD_HD bool hllc(const Model& m, const State<NS>& L, const State<NS>& Rs, const double* a, double beta,
                  double* flux) {
  constexpr int NE = NS + 4;
  double cvL, cpL, RL, gL, c2L, cvR, cpR, RR, gR, c2R;
  const double uL = L.q[NS], vL = L.q[NS + 1], wL = L.q[NS + 2], TL = L.q[NS + 3], pL = L.P, rhoL = L.rho;
  const double pgL = pL - m.Pref;
  props<NS>(m, L.q, TL, cvL, cpL, RL, gL, c2L);
  const double HTL = total_enthalpy(m, L), ETL = HTL - pL;
  const double thetaL = theta_of(L, a);
  const double uR = Rs.q[NS], vR = Rs.q[NS + 1], wR = Rs.q[NS + 2], TR = Rs.q[NS + 3], pR = Rs.P, rhoR = Rs.rho;
  const double pgR = pR - m.Pref;
  props<NS>(m, Rs.q, TR, cvR, cpR, RR, gR, c2R);
  const double HTR = total_enthalpy(m, Rs), ETR = HTR - pR;
  const double thetaR = theta_of(Rs, a);
  const double rho = sqrt(rhoL * rhoR);
  const double sigma = rho / (rhoL + rho);
  State<NS> roe;
#pragma unroll
  for (int i = 0; i < NS; ++i) roe.q[i] = L.q[i] + sigma * (Rs.q[i] - L.q[i]);
  roe.q[NS] = uL + sigma * (uR - uL);
  roe.q[NS + 1] = vL + sigma * (vR - vL);
  roe.q[NS + 2] = wL + sigma * (wR - wL);
  roe.q[NS + 3] = TL + sigma * (TR - TL);
  aux(m, roe);
  const double theta = theta_of(roe, a);
  double cv, cp, R, gt, c2;
  props<NS>(m, roe.q, roe.q[NS + 3], cv, cp, R, gt, c2);
  const double oneMBeta = 1.0 - beta;
  const double thetaPrime = theta * (1.0 + beta) * 0.5;
  const double cPrime = 0.5 * sqrt(theta * theta * (oneMBeta * oneMBeta) + 4.0 * beta * c2);
  const double thetaLPrime = thetaL * (1.0 + beta) * 0.5;
  const double thetaRPrime = thetaR * (1.0 + beta) * 0.5;
  const double cLPrime = 0.5 * sqrt(thetaL * thetaL * (oneMBeta * oneMBeta) + 4.0 * beta * c2L);
  const double cRPrime = 0.5 * sqrt(thetaR * thetaR * (oneMBeta * oneMBeta) + 4.0 * beta * c2R);
  const double eig5L = thetaLPrime - cLPrime;
  const double eig4R = thetaRPrime + cRPrime;
  const double eig4 = thetaPrime + cPrime;
  const double eig5 = thetaPrime - cPrime;
  const double SL = eig5L < eig5 ? eig5L : eig5;  // MIN, ucs/macros.h
  const double SR = eig4R > eig4 ? eig4R : eig4;  // MAX
  double SM = (pgR - pgL + rhoL * thetaL * (SL - thetaL) - rhoR * thetaR * (SR - thetaR)) /
              (rhoL * (SL - thetaL) - rhoR * (SR - thetaR));
  double Qf[NE];
  double pStar = 0.0;
  bool died = false;
  if (SL >= 0.0) {
#pragma unroll
    for (int i = 0; i < NS; ++i) Qf[i] = L.q[i];
    Qf[NS] = rhoL * uL;
    Qf[NS + 1] = rhoL * vL;
    Qf[NS + 2] = rhoL * wL;
    Qf[NS + 3] = ETL;
    pStar = pgL;
    SM = thetaL;
  } else if (SR <= 0.0) {
#pragma unroll
    for (int i = 0; i < NS; ++i) Qf[i] = Rs.q[i];
    Qf[NS] = rhoR * uR;
    Qf[NS + 1] = rhoR * vR;
    Qf[NS + 2] = rhoR * wR;
    Qf[NS + 3] = ETR;
    pStar = pgR;
    SM = thetaR;
  } else if (SL <= 0.0 && SM >= 0.0) {
    pStar = pgL + rhoL * (thetaL - SL) * (thetaL - SM);
    const double omega = 1.0 / (SL - SM), const1 = SL - thetaL, const2 = pStar - pgL;
#pragma unroll
    for (int i = 0; i < NS; ++i) Qf[i] = omega * const1 * L.q[i];
    Qf[NS] = omega * (const1 * rhoL * uL + const2 * a[0]);
    Qf[NS + 1] = omega * (const1 * rhoL * vL + const2 * a[1]);
    Qf[NS + 2] = omega * (const1 * rhoL * wL + const2 * a[2]);
    Qf[NS + 3] = omega * (const1 * ETL - pgL * thetaL + (pStar * SM)) + m.Pref * (SM - thetaL) * omega;
  } else if (SM <= 0.0 && SR >= 0.0) {
    pStar = pgR + rhoR * (thetaR - SR) * (thetaR - SM);
    const double omega = 1.0 / (SR - SM), const1 = SR - thetaR, const2 = pStar - pgR;
#pragma unroll
    for (int i = 0; i < NS; ++i) Qf[i] = omega * const1 * Rs.q[i];
    Qf[NS] = omega * (const1 * rhoR * uR + const2 * a[0]);
    Qf[NS + 1] = omega * (const1 * rhoR * vR + const2 * a[1]);
    Qf[NS + 2] = omega * (const1 * rhoR * wR + const2 * a[2]);
    Qf[NS + 3] = omega * (const1 * ETR - pgR * thetaR + (pStar * SM)) + m.Pref * (SM - thetaR) * omega;
  } else {
    // "HLLC died" (a NaN wave speed): the reference's flux is then NaN in
    // every component and NumericalFlux zeroes it
    died = true;
  }
  if (died) {
#pragma unroll
    for (int i = 0; i < NE; ++i) flux[i] = 0.0;
    return true;
  }
  const double area = a[3];
  const double Et = Qf[NS + 3];
  const double th = SM;
#pragma unroll
  for (int i = 0; i < NS; ++i) flux[i] = area * Qf[i] * th;
  flux[NS] = area * (Qf[NS] * th + pStar * a[0]);
  flux[NS + 1] = area * (Qf[NS + 1] * th + pStar * a[1]);
  flux[NS + 2] = area * (Qf[NS + 2] * th + pStar * a[2]);
  flux[NS + 3] = area * (Et * th + pStar * th) + m.Pref * th * area;
  if (m.sym2d == 1) flux[NS] = 0.0;
  else if (m.sym2d == 2) flux[NS + 1] = 0.0;
  else if (m.sym2d == 3) flux[NS + 2] = 0.0;
  bool z = false;
#pragma unroll
  for (int i = 0; i < NE; ++i)
    if (isnan(flux[i])) {
      flux[i] = 0.0;
      z = true;
    }
  return z;
}

// MaxEigenvalue (ucs/compressibleFR.tcc:1602-1637)
template <int NS>
PCFD_HD double max_eig(const Model& m, const State<NS>& s, const double* a, double beta) {
  double cv, cp, R, g, c2;
  props<NS>(m, s.q, s.q[NS + 3], cv, cp, R, g, c2);
  const double theta = theta_of(s, a);
  const double oneMBeta = 1.0 - beta;
  const double thetaPrime = theta * (1.0 + beta) * 0.5;
  const double cPrime = 0.5 * sqrt(theta * theta * (oneMBeta * oneMBeta) + 4.0 * beta * c2);
  const double e4 = fabs(thetaPrime + cPrime), e5 = fabs(thetaPrime - cPrime);
  return e4 > e5 ? e4 : e5;
}

// PerpVectors (ucs/geometry.h:100-128)
PCFD_HD void perp_vectors(const double* n, double* v1, double* v2) {
  double dot;
  v1[0] = v1[1] = v1[2] = 0.0;
  if (fabs(dot = n[0]) < 0.95) v1[0] = 1.0;
  else if (fabs(dot = n[1]) < 0.95) v1[1] = 1.0;
  else {
    dot = n[2];
    v1[2] = 1.0;
  }
  v1[0] -= dot * n[0];
  v1[1] -= dot * n[1];
  v1[2] -= dot * n[2];
  double mg = sqrt(v1[0] * v1[0] + v1[1] * v1[1] + v1[2] * v1[2]);
  v1[0] /= mg;
  v1[1] /= mg;
  v1[2] /= mg;
  v2[0] = n[1] * v1[2] - n[2] * v1[1];
  v2[1] = n[2] * v1[0] - n[0] * v1[2];
  v2[2] = n[0] * v1[1] - n[1] * v1[0];
  mg = sqrt(v2[0] * v2[0] + v2[1] * v2[1] + v2[2] * v2[2]);
  v2[0] /= mg;
  v2[1] /= mg;
  v2[2] /= mg;
}

// Eigensystem (ucs/compressibleFR.tcc:149-290): eigenvalues and right/left
// eigenvectors of the preconditioned [rho_i, u, v, w, P] system; the species
// sound speeds are set to the mixture's, as the reference does
template <int NS>
PCFD_HD void eigensystem(const Model& m, const State<NS>& s, const double* avec, double beta, double* eig, double* T,
                         double* Tinv) {
  constexpr int NE = NS + 4;
  const double nx = avec[0], ny = avec[1], nz = avec[2];
  const double bm1 = beta - 1.0, bp1 = beta + 1.0, oneMBeta = 1.0 - beta;
  const double theta = theta_of(s, avec);
  const double rho = s.rho;
  double cv, cp, R, g, c2;
  props<NS>(m, s.q, s.q[NS + 3], cv, cp, R, g, c2);
  const double thetaPrime = 0.5 * bp1 * theta;
  const double cPrime = 0.5 * sqrt(theta * theta * (oneMBeta * oneMBeta) + 4.0 * beta * c2);
  double l[3], mm[3];
  perp_vectors(avec, l, mm);
  const double lx = l[0], ly = l[1], lz = l[2], mx = mm[0], my = mm[1], mz = mm[2];
  constexpr int uloc = NS, vloc = NS + 1, wloc = NS + 2, tloc = NS + 3;
#pragma unroll
  for (int i = 0; i < NS + 2; ++i) eig[i] = theta;
  eig[wloc] = thetaPrime + cPrime;
  eig[tloc] = thetaPrime - cPrime;
  const double betam = oneMBeta * 0.5;
  const double Xp = theta * betam + cPrime;
  const double Xm = theta * betam - cPrime;
#pragma unroll
  for (int k = 0; k < NE * NE; ++k) T[k] = 0.0;
#pragma unroll
  for (int i = 0; i < NS; ++i) {
    const double rhoi = s.q[i];
    T[NE * i + i] = 1.0;
    T[NE * i + wloc] = -(rhoi * (c2 + bm1 * c2 - theta * bm1 * Xm)) / (c2 * Xm);
    T[NE * i + tloc] = (rhoi * (c2 + bm1 * c2 - theta * bm1 * Xp)) / (c2 * Xp);
  }
  T[NE * uloc + uloc] = lx;
  T[NE * uloc + vloc] = mx;
  T[NE * uloc + wloc] = nx;
  T[NE * uloc + tloc] = -nx;
  T[NE * vloc + uloc] = ly;
  T[NE * vloc + vloc] = my;
  T[NE * vloc + wloc] = ny;
  T[NE * vloc + tloc] = -ny;
  T[NE * wloc + uloc] = lz;
  T[NE * wloc + vloc] = mz;
  T[NE * wloc + wloc] = nz;
  T[NE * wloc + tloc] = -nz;
  T[NE * tloc + wloc] = -rho * Xm;
  T[NE * tloc + tloc] = rho * Xp;
#pragma unroll
  for (int k = 0; k < NE * NE; ++k) Tinv[k] = 0.0;
#pragma unroll
  for (int i = 0; i < NS; ++i) {
    const double rhoi = s.q[i];
    Tinv[i * NE + i] = 1.0;
    const double KK = -rhoi * (c2 * (Xm + Xp) - bm1 * Xm * Xp * theta + bm1 * c2 * (Xm + Xp));
    Tinv[i * NE + uloc] = -((ly * mz - lz * my) * KK) / (c2 * Xm * Xp);
    Tinv[i * NE + vloc] = ((lx * mz - lz * mx) * KK) / (c2 * Xm * Xp);
    Tinv[i * NE + wloc] = -((lx * my - ly * mx) * KK) / (c2 * Xm * Xp);
    Tinv[i * NE + tloc] = (rhoi * (c2 + bm1 * c2)) / (rho * c2 * Xm * Xp);
  }
  Tinv[NE * uloc + uloc] = my * nz - mz * ny;
  Tinv[NE * uloc + vloc] = -(mx * nz - mz * nx);
  Tinv[NE * uloc + wloc] = mx * ny - my * nx;
  Tinv[NE * uloc + tloc] = 0.0;
  Tinv[NE * vloc + uloc] = -(ly * nz - lz * ny);
  Tinv[NE * vloc + vloc] = lx * nz - lz * nx;
  Tinv[NE * vloc + wloc] = -(lx * ny - ly * nx);
  Tinv[NE * vloc + tloc] = 0.0;
  Tinv[NE * wloc + uloc] = ((Xp) * (ly * mz - lz * my)) / (2.0 * cPrime);
  Tinv[NE * wloc + vloc] = -((Xp) * (lx * mz - lz * mx)) / (2.0 * cPrime);
  Tinv[NE * wloc + wloc] = ((Xp) * (lx * my - ly * mx)) / (2.0 * cPrime);
  Tinv[NE * wloc + tloc] = 1.0 / (2.0 * rho * cPrime);
  Tinv[NE * tloc + uloc] = ((Xm) * (ly * mz - lz * my)) / (2.0 * cPrime);
  Tinv[NE * tloc + vloc] = -((Xm) * (lx * mz - lz * mx)) / (2.0 * cPrime);
  Tinv[NE * tloc + wloc] = ((Xm) * (lx * my - ly * mx)) / (2.0 * cPrime);
  Tinv[NE * tloc + tloc] = 1.0 / (2.0 * rho * cPrime);
}

// NewtonFindTGivenP (ucs/compressibleFR.tcc:2342-2378): T (non-dim) at which
// the species densities give pressure p (non-dim), from Tinit
template <int NS>
PCFD_HD double t_from_p(const Model& m, const double* rhoi, double p, double Tinit) {
  const double tol = 1.0e-15, Tref = m.ref_temperature;
  double rd[NS];
#pragma unroll
  for (int i = 0; i < NS; ++i) rd[i] = rhoi[i] * m.ref_density;
  double TDim = Tinit * Tref;
  const double PgoalDim = p * m.ref_pressure;
  for (int j = 0; j < 28; ++j) {
    const double TpDim = TDim + 1.0e-8;
    double P0 = 0.0, P1 = 0.0;
#pragma unroll
    for (int i = 0; i < NS; ++i) P0 += rd[i] * m.R[i] * TDim;
#pragma unroll
    for (int i = 0; i < NS; ++i) P1 += rd[i] * m.R[i] * TpDim;
    const double z = PgoalDim - P0;
    const double zp = PgoalDim - P1;
    const double dzdT = (zp - z) / (TpDim - TDim);
    const double dT = -z / dzdT;
    if (fabs(dT / Tref) < tol) break;
    TDim += dT;
  }
  return TDim / Tref;
}

// LU with partial (logical) pivoting and LuSolve (ucs/matrix.h:184-264):
// p holds the row order; b is overwritten with the solution
template <int N>
PCFD_HD void lu(double* a, int* p) {
  int row = 0;
#pragma unroll
  for (int i = 0; i < N; ++i) p[i] = i;
  for (int i = 0; i < N; ++i) {
    double large = 0.0;
    for (int j = i; j < N; ++j)
      if (fabs(a[p[j] * N + i]) > fabs(large)) {
        large = a[p[j] * N + i];
        row = j;
      }
    const int t = p[i];
    p[i] = p[row];
    p[row] = t;
    large = 1.0 / large;
    for (int j = i + 1; j < N; ++j) a[p[j] * N + i] *= large;
    for (int j = i + 1; j < N; ++j)
      for (int k = i + 1; k < N; ++k) a[p[j] * N + k] -= a[p[j] * N + i] * a[p[i] * N + k];
  }
}

template <int N, typename T>
PCFD_HD void lu_solve(const T* a, T* b, const int* p) {
  T x[N];
  for (int i = 0; i < N; ++i) {
    T sum = 0.0;
    for (int j = 0; j < i; ++j) sum += a[p[i] * N + j] * x[j];
    x[i] = b[p[i]] - sum;
  }
  for (int i = N - 1; i >= 0; --i) {
    T sum = 0.0;
    for (int j = N - 1; j > i; --j) sum += a[p[i] * N + j] * b[j];
    b[i] = (x[i] - sum) / a[p[i] * N + i];
  }
}

// MatVecMult (ucs/matrix.h:63-74)
template <int N>
PCFD_HD void matvec(const double* a, const double* v, double* out) {
  for (int i = 0; i < N; ++i) {
    out[i] = a[i * N] * v[0];
    for (int j = 1; j < N; ++j) out[i] += a[i * N + j] * v[j];
  }
}

// GetFarfieldBoundaryVariables (ucs/compressibleFR.tcc:940-1040), the
// characteristic (CVBC) branch, vdotn = 0
template <int NS>
PCFD_HD void farfield(const Model& m, const State<NS>& L, State<NS>& R, const State<NS>& Qinf, const double* a,
                      double beta) {
  constexpr int NE = NS + 4;
  if (m.no_cvbc) {
    // noCVBC: the interior or the freestream state by the sign of the first
    // eigenvalue (theta of the average state), 10 times over
    for (int sub = 0; sub < kSubit; ++sub) {
      State<NS> avg;
#pragma unroll
      for (int i = 0; i < NE; ++i) avg.q[i] = 0.5 * (L.q[i] + R.q[i]);
      aux(m, avg);
      R = theta_of(avg, a) >= 0.0 ? L : Qinf;
    }
    return;
  }
  double eig[NE], T[NE * NE], Ti[NE * NE], rhs[NE], ql[NE], qi[NE];
  for (int sub = 0; sub < kSubit; ++sub) {
    State<NS> avg;
#pragma unroll
    for (int i = 0; i < NE; ++i) avg.q[i] = 0.5 * (L.q[i] + R.q[i]);
    aux(m, avg);
    eigensystem(m, avg, a, beta, eig, T, Ti);
#pragma unroll
    for (int i = 0; i < NE; ++i) {
      ql[i] = L.q[i];
      qi[i] = Qinf.q[i];
    }
    const double Tguess = ql[NE - 1];
    ql[NE - 1] = L.P;
    qi[NE - 1] = Qinf.P;
    for (int i = 0; i < NE; ++i) {
      rhs[i] = 0.0;
      const double* src = eig[i] >= 0.0 ? ql : qi;
      for (int j = 0; j < NE; ++j) rhs[i] += Ti[i * NE + j] * src[j];
    }
    matvec<NE>(T, rhs, R.q);
#pragma unroll
    for (int i = 0; i < NS; ++i)
      if (R.q[i] < 0.0) R.q[i] = 0.0;
    R.q[NE - 1] = t_from_p<NS>(m, R.q, R.q[NE - 1], Tguess);
  }
}

// GetInviscidWallBoundaryVariables (ucs/compressibleFR.tcc:1042-1154), the
// characteristic branch, vdotn = 0: the last characteristic equation is
// replaced by theta = 0 and the 9x9 system solved by LU
template <int NS>
PCFD_HD void wall(const Model& m, const State<NS>& L, State<NS>& R, const double* a, double beta) {
  constexpr int NE = NS + 4;
  if (m.no_cvbc) {
    // noCVBC: the interior state with its velocity mirrored in the wall
    // (MirrorVector, ucs/geometry.h:469-476), vdotn = 0
    State<NS> Lm = L;
    const double vdotn = 0.0;
    Lm.q[NS] += vdotn * a[0];
    Lm.q[NS + 1] += vdotn * a[1];
    Lm.q[NS + 2] += vdotn * a[2];
    aux(m, Lm);
#pragma unroll
    for (int i = 0; i < NE; ++i) R.q[i] = Lm.q[i];
    const double dot = 2.0 * (Lm.q[NS] * a[0] + Lm.q[NS + 1] * a[1] + Lm.q[NS + 2] * a[2]);
    R.q[NS] = Lm.q[NS] - dot * a[0];
    R.q[NS + 1] = Lm.q[NS + 1] - dot * a[1];
    R.q[NS + 2] = Lm.q[NS + 2] - dot * a[2];
    return;
  }
  double eig[NE], T[NE * NE], Ti[NE * NE], rhs[NE], ql[NE];
  int p[NE];
  for (int sub = 0; sub < kSubit; ++sub) {
    State<NS> avg;
#pragma unroll
    for (int i = 0; i < NE; ++i) avg.q[i] = 0.5 * (L.q[i] + R.q[i]);
    aux(m, avg);
    eigensystem(m, avg, a, beta, eig, T, Ti);
#pragma unroll
    for (int i = 0; i < NE; ++i) ql[i] = L.q[i];
    const double Tguess = ql[NE - 1];
    ql[NE - 1] = L.P;
    for (int i = 0; i < NE; ++i) {
      rhs[i] = 0.0;
      for (int j = 0; j < NE; ++j) rhs[i] += Ti[i * NE + j] * ql[j];
    }
#pragma unroll
    for (int i = 0; i < NS; ++i) Ti[(NE - 1) * NE + i] = 0.0;
    Ti[(NE - 1) * NE + NS] = a[0];
    Ti[(NE - 1) * NE + NS + 1] = a[1];
    Ti[(NE - 1) * NE + NS + 2] = a[2];
    Ti[(NE - 1) * NE + NS + 3] = 0.0;
    rhs[NE - 1] = 0.0;
    lu<NE>(Ti, p);
    lu_solve<NE>(Ti, rhs, p);
#pragma unroll
    for (int i = 0; i < NE; ++i) R.q[i] = rhs[i];
    R.q[NE - 1] = t_from_p<NS>(m, R.q, R.q[NE - 1], Tguess);
  }
}

// BCTypes ids (ucs/bc_defines.h)
enum { kParallel = 0, kDirichlet = 1, kNeumann = 2, kImpermeable = 3, kFarfield = 6, kSonicInflow = 7,
       kSonicOutflow = 8, kSymmetry = 9 };

// CalculateBoundaryVariables (ucs/bc.tcc:1058-1398) for the types this path
// supports; the aux of QR then QL recomputed at the end.  Returns false for a
// type it does not handle (left unchanged, as the reference's default branch).
template <int NS>
PCFD_HD bool bc_state(const Model& m, int type, State<NS>& L, State<NS>& R, const State<NS>& Qref, const double* a,
                      double beta) {
  constexpr int NE = NS + 4;
  if (type == kParallel) return true;
  if (type == kSonicInflow || type == kDirichlet) {
    R = Qref;
    L = Qref;
  } else if (type == kSonicOutflow || type == kNeumann) {
#pragma unroll
    for (int i = 0; i < NE; ++i) R.q[i] = L.q[i];
  } else if (type == kFarfield) {
    farfield(m, L, R, Qref, a, beta);
  } else if (type == kImpermeable || type == kSymmetry) {
    wall(m, L, R, a, beta);
  } else {
    return false;
  }
  aux(m, R);
  aux(m, L);
  return true;
}

// ApplyDQ (ucs/compressibleFR.tcc:885-937)
template <int NS>
PCFD_HD void apply_dq(const Model& m, const double* dQ, State<NS>& s) {
#pragma unroll
  for (int i = 0; i < NS; ++i) {
    const double r = s.q[i] + dQ[i];
    if (!(r < 0.0)) s.q[i] += dQ[i];
  }
  const double T = s.q[NS + 3] + dQ[NS + 3];
  s.q[NS + 3] = T < 0.0 ? 1.0e-10 : T;
  s.q[NS] += dQ[NS];
  s.q[NS + 1] += dQ[NS + 1];
  s.q[NS + 2] += dQ[NS + 2];
  aux(m, s);
}

// ContributeTemporalTerms (ucs/compressibleFR.tcc:1318-1464) with cnp1 = 1,
// added to the diagonal block A (row-major NE x NE)
template <int NS>
PCFD_HD void temporal(const Model& m, const State<NS>& s, double vol, double dtau, double beta, double* A) {
  constexpr int NE = NS + 4;
  const double cnp1 = 1.0;
  const double vOverDt = m.local_dt ? cnp1 * vol / m.dt_param + vol / dtau : cnp1 * vol / dtau;
  constexpr int uloc = NS, vloc = NS + 1, wloc = NS + 2, tloc = NS + 3;
  const double rho = s.rho, T = s.q[tloc];
  const double u = s.q[uloc], v = s.q[vloc], w = s.q[wloc];
  const double rvOverDt = rho * vOverDt;
  double v2 = u * u + v * v + w * w;
  double cv, cp, R, gamma, c2;
  props<NS>(m, s.q, T, cv, cp, R, gamma, c2);
  const double rhoDim = rho * m.ref_density, TDim = T * m.ref_temperature;
  v2 *= m.ref_velocity * m.ref_velocity;
  double rd[NS], Yi[NS], dEtdRhoi[NS];
#pragma unroll
  for (int i = 0; i < NS; ++i) {
    rd[i] = s.q[i] * m.ref_density;
    Yi[i] = s.q[i] / rho;
  }
  // ChemModel::dEtdP_dEtdRhoi (ucs/chem.tcc:828-858) with the ideal-gas
  // derivatives (ucs/EOS.tcc:42-70) and dRmixdRhoi (ucs/chem.tcc:861-870)
  double dEtdP;
  {
    const double hv2 = 0.5 * v2;
    double rhomix = 0.0, Rmix = 0.0;
#pragma unroll
    for (int i = 0; i < NS; ++i) {
      rhomix += rd[i];
      Rmix += rd[i] * m.R[i];
    }
    Rmix /= rhomix;
    const double dPdrhomix = Rmix * TDim;
    const double dPdRmix = rhomix * TDim;
#pragma unroll
    for (int i = 0; i < NS; ++i) {
      double d = m.R[i] * (rhomix - rd[i]) / (rhomix * rhomix);
#pragma unroll
      for (int j = 0; j < NS; ++j) {
        if (j == i) continue;
        d -= m.R[j] * rd[j] / (rhomix * rhomix);
      }
      dEtdRhoi[i] = hv2 + h_species(m, i, TDim) - (dPdrhomix + dPdRmix * d);
    }
    dEtdP = 0.0;
    const double dTdP = 1.0 / (rhomix * Rmix);
#pragma unroll
    for (int i = 0; i < NS; ++i) dEtdP += rd[i] * (dhdt_species(m, i, TDim)) * dTdP;
    dEtdP -= 1.0;
  }
  const double ref_detdrho = m.vref2 * m.ref_density / m.ref_density;
#pragma unroll
  for (int i = 0; i < NS; ++i) dEtdRhoi[i] /= ref_detdrho;
  const double ref_detdP = m.vref2 * m.ref_density / m.ref_pressure;
  dEtdP /= ref_detdP;
  double dPdT = rhoDim * (R * m.s_ref);
  dPdT /= (m.ref_pressure / m.ref_temperature);
  const double oneOBeta = 1.0 / beta, oneMbeta = 1.0 - beta;
  double tob[NS];
#pragma unroll
  for (int i = 0; i < NS; ++i) tob[i] = (Yi[i] * oneMbeta / c2) * oneOBeta;
#pragma unroll
  for (int i = 0; i < NS; ++i) {
    A[NE * i + i] += vOverDt;
    A[NE * i + tloc] += tob[i] * vOverDt * dPdT;
  }
  const double vel[3] = {u, v, w};
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const int row = NS + k;
    double precond = 0.0;
#pragma unroll
    for (int j = 0; j < NS; ++j) {
      A[NE * row + j] += vel[k] * vOverDt;
      precond += tob[j] * vel[k];
    }
    A[NE * row + row] += rvOverDt;
    A[NE * row + tloc] += precond * vOverDt * dPdT;
  }
  double precond = 0.0;
#pragma unroll
  for (int j = 0; j < NS; ++j) {
    A[NE * tloc + j] += dEtdRhoi[j] * vOverDt;
    precond += dEtdRhoi[j] * tob[j];
  }
  A[NE * tloc + uloc] += u * rvOverDt;
  A[NE * tloc + vloc] += v * rvOverDt;
  A[NE * tloc + wloc] += w * rvOverDt;
  precond += dEtdP * (oneOBeta);
  A[NE * tloc + tloc] += precond * vOverDt * dPdT;
}

}  // namespace fr
}  // namespace pcfd
