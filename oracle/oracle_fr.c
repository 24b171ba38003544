/*
 * CPU oracle, finite-rate (FR) equation set -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C FP64 restatement of the reference's CompressibleFREqnSet on the
 * first-order implicit path, in the reference's order of operations:
 *   native state rho_1..rho_ns, u, v, w, T; aux P, rho, cv_i, conc_i
 *                                       ucs/compressibleFR.tcc:12-31
 *   GetFluidProperties                  ucs/compressibleFR.tcc:1521-1570,
 *                                       ucs/chem.tcc:544-574, 1001-1041
 *   GetTotalEnergy / GetTotalEnthalpy   ucs/compressibleFR.tcc:1218-1273
 *   HLLCFlux (+ NumericalFlux's NaN zeroing)
 *                                       ucs/compressibleFR.tcc:300-548,
 *                                       ucs/eqnset.tcc:36-90
 *   Eigensystem + PerpVectors           ucs/compressibleFR.tcc:149-290,
 *                                       ucs/geometry.h:100-128
 *   GetFarfieldBoundaryVariables        ucs/compressibleFR.tcc:940-1040
 *   GetInviscidWallBoundaryVariables    ucs/compressibleFR.tcc:1042-1154
 *   CalculateBoundaryVariables          ucs/bc.tcc:1058-1398 (types on this path)
 *   MaxEigenvalue                       ucs/compressibleFR.tcc:1602-1637
 *   ApplyDQ                             ucs/compressibleFR.tcc:885-937
 *   ContributeTemporalTerms             ucs/compressibleFR.tcc:1318-1464,
 *                                       ucs/chem.tcc:828-858, ucs/EOS.tcc:42-70
 *   UpdateQinf                          ucs/compressibleFR.tcc:1465-1518
 *   SpatialResidual (first order)       ucs/residual.tcc:64-114
 *   ComputeTimesteps                    ucs/timestep.tcc:6-143
 *   ComputeSpatialJacobian + temporal   ucs/jacobian.tcc:12-250, 253-304,
 *                                       436-535
 *   CRS::SGS on neqn x neqn blocks      ucs/crs.tcc:61-173
 *   NewtonIterate (first order)         ucs/solutionSpace.tcc:572-804
 * The thermodynamics (NASA-7 Cp and H, ideal-gas EOS, T from P by Newton)
 * are those of oracle_chem.c.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define UNIV_R 8.31447215
#define FR_NSUBIT 10

static double fmax2(double a, double b) { return a > b ? a : b; } /* MAX, ucs/macros.h */
static double fmin2(double a, double b) { return a < b ? a : b; } /* MIN */

static const double* coeffs(const o_chem* c, int s, double T) {
  /* Species::GetThermoCoeff (ucs/species.tcc:98-137): the coefficient set
   * is pinned outside 200-6000 K, the polynomial still takes T */
  return (T < 200.0) ? c->lo[s] : (T > 6000.0) ? c->hi[s] : (T > 1000.0 ? c->hi[s] : c->lo[s]);
}

/* Species::GetH (ucs/species.tcc:56-70; href = 0, :327): J/kg */
static double species_h(const o_chem* c, int s, double T) {
  const double* a = coeffs(c, s, T);
  double h_R = a[5] + T * (a[0] + T * (a[1] / 2.0 + T * (a[2] / 3.0 + T * (a[3] / 4.0 + T * a[4] / 5.0))));
  return h_R * (UNIV_R / c->mw[s]);
}

/* Species::GetdHdT (ucs/species.tcc:337-348) */
static double species_dhdt(const o_chem* c, int s, double T) {
  const double* a = coeffs(c, s, T);
  double r = a[0] + T * (a[1] + T * (a[2] + T * (a[3] + T * a[4])));
  return r * (UNIV_R / c->mw[s]);
}

static double species_R(const o_chem* c, int s) { return UNIV_R / c->mw[s]; }

int o_fr_neqn(const o_fr* f) { return f->chem->ns + 4; }
int o_fr_nvars(const o_fr* f) { return 3 * f->chem->ns + 6; }

/* ComputeAuxiliaryVariables (ucs/compressibleFR.tcc:754-814) in place */
void o_fr_aux_q(const o_fr* f, double* Q) {
  const int ns = f->chem->ns;
  o_fr_aux(f->chem, Q, f->ref_pressure, f->ref_velocity, &Q[ns + 4]);
}

/* GetFluidProperties (ucs/compressibleFR.tcc:1521-1570 over ucs/chem.tcc:
 * 544-574): mixture cv, cp, R, gamma and c^2 = gamma R T, non-dimensional */
static void fr_props(const o_fr* f, const double* rhoi, double T, double* cv, double* cp, double* R,
                     double* gamma, double* c2) {
  const o_chem* c = f->chem;
  const int ns = c->ns;
  const double s_ref = f->ref_velocity * f->ref_velocity / c->ref_temperature;
  double rd[O_CHEM_MAXS], rho = 0.0, Rd = 0.0;
  for (int i = 0; i < ns; i++) rd[i] = rhoi[i] * c->ref_density;
  const double Td = T * c->ref_temperature;
  for (int i = 0; i < ns; i++) rho += rd[i];
  for (int i = 0; i < ns; i++) Rd += rd[i] * species_R(c, i);
  Rd /= rho;
  /* GetCp / GetCv (ucs/chem.tcc:1001-1041): mass-fraction weighted */
  double X[O_CHEM_MAXS], cpd = 0.0, cvd = 0.0;
  for (int i = 0; i < ns; i++) X[i] = rd[i] / rho;
  for (int i = 0; i < ns; i++) cpd += o_chem_cp(c, i, Td) * X[i];
  for (int i = 0; i < ns; i++) cvd += X[i] * (o_chem_cp(c, i, Td) - species_R(c, i));
  const double g = cpd / cvd;
  const double c2d = g * Rd * Td;
  *gamma = g;
  *cv = cvd / s_ref;
  *cp = cpd / s_ref;
  *R = Rd / s_ref;
  *c2 = c2d / (f->ref_velocity * f->ref_velocity);
}

/* GetTotalEnthalpy / GetTotalEnergy (ucs/compressibleFR.tcc:1218-1273) */
static double fr_h_mix(const o_fr* f, const double* Q) {
  const o_chem* c = f->chem;
  const int ns = c->ns;
  const double rho = Q[ns + 5];
  const double Td = Q[ns + 3] * c->ref_temperature;
  double h = 0.0;
  for (int i = 0; i < ns; i++) h += species_h(c, i, Td) * (Q[i] / rho);
  return h / (f->ref_velocity * f->ref_velocity);
}

double o_fr_total_enthalpy(const o_fr* f, const double* Q) {
  const int ns = f->chem->ns;
  const double u = Q[ns], v = Q[ns + 1], w = Q[ns + 2], rho = Q[ns + 5];
  return fr_h_mix(f, Q) * rho + 0.5 * rho * (u * u + v * v + w * w);
}

double o_fr_total_energy(const o_fr* f, const double* Q) {
  const int ns = f->chem->ns;
  const double u = Q[ns], v = Q[ns + 1], w = Q[ns + 2], P = Q[ns + 4], rho = Q[ns + 5];
  return (fr_h_mix(f, Q) * rho - P) + 0.5 * rho * (u * u + v * v + w * w);
}

/* GetTheta (ucs/compressibleFR.tcc:1640-1643) with vdotn = 0 */
static double fr_theta(const o_fr* f, const double* Q, const double* a) {
  const int ns = f->chem->ns;
  return Q[ns] * a[0] + Q[ns + 1] * a[1] + Q[ns + 2] * a[2] + 0.0;
}

/* MatVecMult (ucs/matrix.h:63-74) */
static void matvec(const double* a, const double* v, double* out, int n) {
  for (int i = 0; i < n; i++) {
    out[i] = a[i * n] * v[0];
    for (int j = 1; j < n; j++) out[i] += a[i * n + j] * v[j];
  }
}

/* HLLCFlux (ucs/compressibleFR.tcc:300-548), vdotn = 0; NaN components
 * zeroed as NumericalFlux does (ucs/eqnset.tcc:74-88).  Returns 1 when a
 * component was zeroed. */
int o_fr_hllc(const o_fr* f, const double* QL, const double* QR, const double* avec, double beta, double* flux) {
  const o_chem* c = f->chem;
  const int ns = c->ns, neqn = ns + 4, nv = 3 * ns + 6;
  double cvL, cpL, RL, gammaL, c2L, cvR, cpR, RR, gammaR, c2R;
  const double uL = QL[ns], vL = QL[ns + 1], wL = QL[ns + 2], TL = QL[ns + 3], pL = QL[ns + 4], rhoL = QL[ns + 5];
  const double pgL = pL - f->Pref;
  fr_props(f, QL, TL, &cvL, &cpL, &RL, &gammaL, &c2L);
  const double HTL = o_fr_total_enthalpy(f, QL), ETL = HTL - pL;
  const double cL = sqrt(c2L);
  const double thetaL = fr_theta(f, QL, avec);
  const double uR = QR[ns], vR = QR[ns + 1], wR = QR[ns + 2], TR = QR[ns + 3], pR = QR[ns + 4], rhoR = QR[ns + 5];
  const double pgR = pR - f->Pref;
  fr_props(f, QR, TR, &cvR, &cpR, &RR, &gammaR, &c2R);
  const double HTR = o_fr_total_enthalpy(f, QR), ETR = HTR - pR;
  const double cR = sqrt(c2R);
  const double thetaR = fr_theta(f, QR, avec);
  (void)cL;
  (void)cR;
  double rho = sqrt(rhoL * rhoR);
  const double sigma = rho / (rhoL + rho);
  double roeQ[3 * O_CHEM_MAXS + 6];
  for (int i = 0; i < ns; i++) roeQ[i] = QL[i] + sigma * (QR[i] - QL[i]);
  roeQ[ns + 0] = uL + sigma * (uR - uL);
  roeQ[ns + 1] = vL + sigma * (vR - vL);
  roeQ[ns + 2] = wL + sigma * (wR - wL);
  roeQ[ns + 3] = TL + sigma * (TR - TL);
  o_fr_aux_q(f, roeQ);
  double theta = fr_theta(f, roeQ, avec);
  double cv, cp, R, gtrash, c2;
  fr_props(f, roeQ, roeQ[ns + 3], &cv, &cp, &R, &gtrash, &c2);
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
  const double SL = fmin2(eig5L, eig5);
  const double SR = fmax2(eig4R, eig4);
  double SM = (pgR - pgL + rhoL * thetaL * (SL - thetaL) - rhoR * thetaR * (SR - thetaR)) /
              (rhoL * (SL - thetaL) - rhoR * (SR - thetaR));
  double Qf[3 * O_CHEM_MAXS + 6];
  double pStar = 0.0;
  (void)nv;
  if (SL >= 0.0) {
    for (int i = 0; i < ns; i++) Qf[i] = QL[i];
    Qf[ns] = rhoL * uL;
    Qf[ns + 1] = rhoL * vL;
    Qf[ns + 2] = rhoL * wL;
    Qf[ns + 3] = ETL;
    pStar = pgL;
    SM = thetaL;
  } else if (SR <= 0.0) {
    for (int i = 0; i < ns; i++) Qf[i] = QR[i];
    Qf[ns] = rhoR * uR;
    Qf[ns + 1] = rhoR * vR;
    Qf[ns + 2] = rhoR * wR;
    Qf[ns + 3] = ETR;
    pStar = pgR;
    SM = thetaR;
  } else if (SL <= 0.0 && SM >= 0.0) {
    pStar = pgL + rhoL * (thetaL - SL) * (thetaL - SM);
    const double omega = 1.0 / (SL - SM), const1 = SL - thetaL, const2 = pStar - pgL;
    for (int i = 0; i < ns; i++) Qf[i] = omega * const1 * QL[i];
    Qf[ns] = omega * (const1 * rhoL * uL + const2 * avec[0]);
    Qf[ns + 1] = omega * (const1 * rhoL * vL + const2 * avec[1]);
    Qf[ns + 2] = omega * (const1 * rhoL * wL + const2 * avec[2]);
    Qf[ns + 3] = omega * (const1 * ETL - pgL * thetaL + (pStar * SM)) + f->Pref * (SM - thetaL) * omega;
  } else if (SM <= 0.0 && SR >= 0.0) {
    pStar = pgR + rhoR * (thetaR - SR) * (thetaR - SM);
    const double omega = 1.0 / (SR - SM), const1 = SR - thetaR, const2 = pStar - pgR;
    for (int i = 0; i < ns; i++) Qf[i] = omega * const1 * QR[i];
    Qf[ns] = omega * (const1 * rhoR * uR + const2 * avec[0]);
    Qf[ns + 1] = omega * (const1 * rhoR * vR + const2 * avec[1]);
    Qf[ns + 2] = omega * (const1 * rhoR * wR + const2 * avec[2]);
    Qf[ns + 3] = omega * (const1 * ETR - pgR * thetaR + (pStar * SM)) + f->Pref * (SM - thetaR) * omega;
  } else {
    /* the reference soft-aborts ("HLLC died"); the flux stays undefined */
    for (int i = 0; i < neqn; i++) flux[i] = NAN;
    int z = 0;
    for (int i = 0; i < neqn; i++)
      if (isnan(flux[i])) {
        flux[i] = 0.0;
        z = 1;
      }
    return z;
  }
  const double area = avec[3];
  const double Et = Qf[ns + 3];
  theta = SM;
  const double thetabar = theta;  /* theta - vdotn */
  for (int i = 0; i < ns; i++) flux[i] = area * Qf[i] * theta;
  flux[ns] = area * (Qf[ns] * theta + pStar * avec[0]);
  flux[ns + 1] = area * (Qf[ns + 1] * theta + pStar * avec[1]);
  flux[ns + 2] = area * (Qf[ns + 2] * theta + pStar * avec[2]);
  flux[ns + 3] = area * (Et * theta + pStar * thetabar) + f->Pref * thetabar * area;
  if (f->sym2d == 1) flux[ns] = 0.0;
  else if (f->sym2d == 2) flux[ns + 1] = 0.0;
  else if (f->sym2d == 3) flux[ns + 2] = 0.0;
  (void)rho;
  int z = 0;
  for (int i = 0; i < neqn; i++)
    if (isnan(flux[i])) {
      flux[i] = 0.0;
      z = 1;
    }
  return z;
}

/* MaxEigenvalue (ucs/compressibleFR.tcc:1602-1637) */
double o_fr_max_eig(const o_fr* f, const double* Q, const double* avec, double beta) {
  const int ns = f->chem->ns;
  double cv, cp, R, g, c2;
  fr_props(f, Q, Q[ns + 3], &cv, &cp, &R, &g, &c2);
  const double theta = fr_theta(f, Q, avec);
  const double oneMBeta = 1.0 - beta;
  const double thetaPrime = theta * (1.0 + beta) * 0.5;
  const double cPrime = 0.5 * sqrt(theta * theta * (oneMBeta * oneMBeta) + 4.0 * beta * c2);
  return fmax2(fabs(thetaPrime + cPrime), fabs(thetaPrime - cPrime));
}

/* PerpVectors (ucs/geometry.h:100-128) */
static void perp_vectors(const double* n, double* v1, double* v2) {
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
  double m = sqrt(v1[0] * v1[0] + v1[1] * v1[1] + v1[2] * v1[2]);
  v1[0] /= m;
  v1[1] /= m;
  v1[2] /= m;
  v2[0] = n[1] * v1[2] - n[2] * v1[1];
  v2[1] = n[2] * v1[0] - n[0] * v1[2];
  v2[2] = n[0] * v1[1] - n[1] * v1[0];
  m = sqrt(v2[0] * v2[0] + v2[1] * v2[1] + v2[2] * v2[2]);
  v2[0] /= m;
  v2[1] /= m;
  v2[2] /= m;
}

/* Eigensystem (ucs/compressibleFR.tcc:149-290): eigenvalues and the right /
 * left eigenvectors of the [rho_i, u, v, w, P] system */
void o_fr_eigensystem(const o_fr* f, const double* Q, const double* avec, double beta, double* eig, double* T,
                      double* Tinv) {
  const int ns = f->chem->ns, neqn = ns + 4;
  const double nx = avec[0], ny = avec[1], nz = avec[2];
  const double bm1 = beta - 1.0, bp1 = beta + 1.0, oneMBeta = 1.0 - beta;
  const double theta = fr_theta(f, Q, avec);
  const double rho = Q[ns + 5];
  double cv, cp, R, g, c2;
  fr_props(f, Q, Q[ns + 3], &cv, &cp, &R, &g, &c2);
  const double thetaPrime = 0.5 * bp1 * theta;
  const double cPrime = 0.5 * sqrt(theta * theta * (oneMBeta * oneMBeta) + 4.0 * beta * c2);
  double c2i[O_CHEM_MAXS];
  for (int i = 0; i < ns; i++) c2i[i] = c2;  /* the reference resets the species values to the bulk one */
  double l[3], m[3];
  perp_vectors(avec, l, m);
  const double lx = l[0], ly = l[1], lz = l[2], mx = m[0], my = m[1], mz = m[2];
  const int uloc = ns, vloc = ns + 1, wloc = ns + 2, tloc = ns + 3;
  for (int i = 0; i < ns + 2; i++) eig[i] = theta;
  eig[wloc] = thetaPrime + cPrime;
  eig[tloc] = thetaPrime - cPrime;
  const double betam = oneMBeta * 0.5;
  const double Xp = theta * betam + cPrime;
  const double Xm = theta * betam - cPrime;
  memset(T, 0, sizeof(double) * neqn * neqn);
  for (int i = 0; i < ns; i++) {
    const double rhoi = Q[i];
    T[neqn * i + i] = 1.0;
    T[neqn * i + wloc] = -(rhoi * (c2i[i] + bm1 * c2 - theta * bm1 * Xm)) / (c2i[i] * Xm);
    T[neqn * i + tloc] = (rhoi * (c2i[i] + bm1 * c2 - theta * bm1 * Xp)) / (c2i[i] * Xp);
  }
  T[neqn * uloc + uloc] = lx;
  T[neqn * uloc + vloc] = mx;
  T[neqn * uloc + wloc] = nx;
  T[neqn * uloc + tloc] = -nx;
  T[neqn * vloc + uloc] = ly;
  T[neqn * vloc + vloc] = my;
  T[neqn * vloc + wloc] = ny;
  T[neqn * vloc + tloc] = -ny;
  T[neqn * wloc + uloc] = lz;
  T[neqn * wloc + vloc] = mz;
  T[neqn * wloc + wloc] = nz;
  T[neqn * wloc + tloc] = -nz;
  T[neqn * tloc + wloc] = -rho * Xm;
  T[neqn * tloc + tloc] = rho * Xp;
  memset(Tinv, 0, sizeof(double) * neqn * neqn);
  for (int i = 0; i < ns; i++) {
    const double rhoi = Q[i];
    Tinv[i * neqn + i] = 1.0;
    const double KK = -rhoi * (c2i[i] * (Xm + Xp) - bm1 * Xm * Xp * theta + bm1 * c2 * (Xm + Xp));
    Tinv[i * neqn + uloc] = -((ly * mz - lz * my) * KK) / (c2i[i] * Xm * Xp);
    Tinv[i * neqn + vloc] = ((lx * mz - lz * mx) * KK) / (c2i[i] * Xm * Xp);
    Tinv[i * neqn + wloc] = -((lx * my - ly * mx) * KK) / (c2i[i] * Xm * Xp);
    Tinv[i * neqn + tloc] = (rhoi * (c2i[i] + bm1 * c2)) / (rho * c2i[i] * Xm * Xp);
  }
  Tinv[neqn * uloc + uloc] = my * nz - mz * ny;
  Tinv[neqn * uloc + vloc] = -(mx * nz - mz * nx);
  Tinv[neqn * uloc + wloc] = mx * ny - my * nx;
  Tinv[neqn * uloc + tloc] = 0.0;
  Tinv[neqn * vloc + uloc] = -(ly * nz - lz * ny);
  Tinv[neqn * vloc + vloc] = lx * nz - lz * nx;
  Tinv[neqn * vloc + wloc] = -(lx * ny - ly * nx);
  Tinv[neqn * vloc + tloc] = 0.0;
  Tinv[neqn * wloc + uloc] = ((Xp) * (ly * mz - lz * my)) / (2.0 * cPrime);
  Tinv[neqn * wloc + vloc] = -((Xp) * (lx * mz - lz * mx)) / (2.0 * cPrime);
  Tinv[neqn * wloc + wloc] = ((Xp) * (lx * my - ly * mx)) / (2.0 * cPrime);
  Tinv[neqn * wloc + tloc] = 1.0 / (2.0 * rho * cPrime);
  Tinv[neqn * tloc + uloc] = ((Xm) * (ly * mz - lz * my)) / (2.0 * cPrime);
  Tinv[neqn * tloc + vloc] = -((Xm) * (lx * mz - lz * mx)) / (2.0 * cPrime);
  Tinv[neqn * tloc + wloc] = ((Xm) * (lx * my - ly * mx)) / (2.0 * cPrime);
  Tinv[neqn * tloc + tloc] = 1.0 / (2.0 * rho * cPrime);
}

/* GetFarfieldBoundaryVariables (ucs/compressibleFR.tcc:940-1040), vdotn = 0:
 * the characteristic (CVBC) branch, or with noCVBC the upwind copy of the
 * interior or freestream state by the sign of the first eigenvalue */
static void fr_farfield(const o_fr* f, const double* QL, double* QR, const double* Qinf, const double* avec,
                        double beta) {
  const int ns = f->chem->ns, neqn = ns + 4, nv = 3 * ns + 6;
  if (f->no_cvbc) {
    double qavg[3 * O_CHEM_MAXS + 6], eig[O_CHEM_MAXS + 4];
    double T[(O_CHEM_MAXS + 4) * (O_CHEM_MAXS + 4)], Tinv[(O_CHEM_MAXS + 4) * (O_CHEM_MAXS + 4)];
    for (int subit = 0; subit < FR_NSUBIT; subit++) {
      for (int i = 0; i < neqn; i++) qavg[i] = 0.5 * (QL[i] + QR[i]);
      o_fr_aux_q(f, qavg);
      o_fr_eigensystem(f, qavg, avec, beta, eig, T, Tinv);
      memcpy(QR, eig[0] >= 0.0 ? QL : Qinf, sizeof(double) * nv);
    }
    return;
  }
  double qavg[3 * O_CHEM_MAXS + 6], eig[O_CHEM_MAXS + 4];
  double T[(O_CHEM_MAXS + 4) * (O_CHEM_MAXS + 4)], Tinv[(O_CHEM_MAXS + 4) * (O_CHEM_MAXS + 4)];
  double rhs[O_CHEM_MAXS + 4], ql[O_CHEM_MAXS + 4], qinf[O_CHEM_MAXS + 4];
  for (int subit = 0; subit < FR_NSUBIT; subit++) {
    for (int i = 0; i < neqn; i++) qavg[i] = 0.5 * (QL[i] + QR[i]);
    o_fr_aux_q(f, qavg);
    o_fr_eigensystem(f, qavg, avec, beta, eig, T, Tinv);
    memcpy(ql, QL, sizeof(double) * neqn);
    memcpy(qinf, Qinf, sizeof(double) * neqn);
    const double Tguess = ql[neqn - 1];
    ql[neqn - 1] = QL[ns + 4];
    qinf[neqn - 1] = Qinf[ns + 4];
    for (int i = 0; i < neqn; i++) {
      rhs[i] = 0.0;
      for (int j = 0; j < neqn; j++) rhs[i] += Tinv[i * neqn + j] * (eig[i] >= 0.0 ? ql[j] : qinf[j]);
    }
    matvec(T, rhs, QR, neqn);
    for (int i = 0; i < ns; i++)
      if (QR[i] < 0.0) QR[i] = 0.0;
    const double pgoal = QR[neqn - 1];
    int it;
    QR[neqn - 1] = o_fr_t_from_p(f->chem, QR, pgoal, Tguess, f->ref_pressure, &it);
  }
}

/* GetInviscidWallBoundaryVariables (ucs/compressibleFR.tcc:1042-1154),
 * vdotn = 0: the characteristic (CVBC) branch, or with noCVBC the interior
 * state with its velocity mirrored in the wall (MirrorVector,
 * ucs/geometry.h:469-476) */
static void fr_wall(const o_fr* f, const double* QL, double* QR, const double* avec, double beta) {
  const int ns = f->chem->ns, neqn = ns + 4;
  if (f->no_cvbc) {
    double QLmod[3 * O_CHEM_MAXS + 6];
    memcpy(QLmod, QL, sizeof(double) * (3 * ns + 6));
    const double vdotn = 0.0;
    QLmod[ns] += vdotn * avec[0];
    QLmod[ns + 1] += vdotn * avec[1];
    QLmod[ns + 2] += vdotn * avec[2];
    o_fr_aux_q(f, QLmod);
    for (int i = 0; i < neqn; i++) QR[i] = QLmod[i];
    const double dot = 2.0 * (QLmod[ns] * avec[0] + QLmod[ns + 1] * avec[1] + QLmod[ns + 2] * avec[2]);
    QR[ns] = QLmod[ns] - dot * avec[0];
    QR[ns + 1] = QLmod[ns + 1] - dot * avec[1];
    QR[ns + 2] = QLmod[ns + 2] - dot * avec[2];
    return;
  }
  double qavg[3 * O_CHEM_MAXS + 6], eig[O_CHEM_MAXS + 4];
  double T[(O_CHEM_MAXS + 4) * (O_CHEM_MAXS + 4)], Tinv[(O_CHEM_MAXS + 4) * (O_CHEM_MAXS + 4)];
  double rhs[O_CHEM_MAXS + 4], scr[O_CHEM_MAXS + 4], ql[O_CHEM_MAXS + 4];
  int p[O_CHEM_MAXS + 4];
  for (int subit = 0; subit < FR_NSUBIT; subit++) {
    for (int i = 0; i < neqn; i++) qavg[i] = 0.5 * (QL[i] + QR[i]);
    o_fr_aux_q(f, qavg);
    o_fr_eigensystem(f, qavg, avec, beta, eig, T, Tinv);
    memcpy(ql, QL, sizeof(double) * neqn);
    const double Tguess = ql[neqn - 1];
    ql[neqn - 1] = QL[ns + 4];
    for (int i = 0; i < neqn; i++) {
      rhs[i] = 0.0;
      for (int j = 0; j < neqn; j++) rhs[i] += Tinv[i * neqn + j] * ql[j];
    }
    for (int i = 0; i < ns; i++) Tinv[(neqn - 1) * neqn + i] = 0.0;
    Tinv[(neqn - 1) * neqn + ns] = avec[0];
    Tinv[(neqn - 1) * neqn + ns + 1] = avec[1];
    Tinv[(neqn - 1) * neqn + ns + 2] = avec[2];
    Tinv[(neqn - 1) * neqn + ns + 3] = 0.0;
    rhs[neqn - 1] = 0.0;
    o_lu(Tinv, p, neqn);
    o_lu_solve(Tinv, rhs, p, scr, neqn);
    memcpy(QR, rhs, sizeof(double) * neqn);
    const double pgoal = QR[neqn - 1];
    int it;
    QR[neqn - 1] = o_fr_t_from_p(f->chem, QR, pgoal, Tguess, f->ref_pressure, &it);
  }
}

/* CalculateBoundaryVariables (ucs/bc.tcc:1058-1398) for the FR state */
void o_fr_bc_state(const o_fr* f, int bctype, double* QL, double* QR, const double* Qref, const double* avec,
                   double beta) {
  const int ns = f->chem->ns, neqn = ns + 4, nv = 3 * ns + 6;
  if (bctype == O_BC_PARALLEL) return;
  if (bctype == O_BC_SONIC_INFLOW || bctype == O_BC_DIRICHLET) {
    for (int i = 0; i < nv; i++) QR[i] = QL[i] = Qref[i];
  } else if (bctype == O_BC_SONIC_OUTFLOW || bctype == O_BC_NEUMANN) {
    for (int i = 0; i < neqn; i++) QR[i] = QL[i];
  } else if (bctype == O_BC_FARFIELD) {
    fr_farfield(f, QL, QR, Qref, avec, beta);
  } else if (bctype == O_BC_IMPERMEABLE || bctype == O_BC_SYMMETRY) {
    fr_wall(f, QL, QR, avec, beta);
  } else {
    return;
  }
  o_fr_aux_q(f, QR);
  o_fr_aux_q(f, QL);
}

/* ApplyDQ (ucs/compressibleFR.tcc:885-937) */
void o_fr_apply_dq(const o_fr* f, const double* dQ, double* Q) {
  const int ns = f->chem->ns;
  const double minT = 1.0e-10;
  for (int i = 0; i < ns; i++) {
    const double rho = Q[i] + dQ[i];
    if (rho < 0.0) {
      if (Q[i] == 0.0) Q[i] = 0.0;
    } else {
      Q[i] += dQ[i];
    }
  }
  const double projectedT = Q[ns + 3] + dQ[ns + 3];
  Q[ns + 3] = projectedT < 0.0 ? minT : projectedT;
  Q[ns] += dQ[ns];
  Q[ns + 1] += dQ[ns + 1];
  Q[ns + 2] += dQ[ns + 2];
  o_fr_aux_q(f, Q);
}

/* ContributeTemporalTerms (ucs/compressibleFR.tcc:1318-1464) with cnp1 = 1;
 * dt is the timeStep parameter, dtau the local pseudo time step */
void o_fr_temporal(const o_fr* f, const double* Q, double vol, double dtau, double beta, double* A) {
  const o_chem* c = f->chem;
  const int ns = c->ns, neqn = ns + 4;
  const double cnp1 = 1.0;
  const double vOverDt = f->local_dt ? cnp1 * vol / f->dt_param + vol / dtau : cnp1 * vol / dtau;
  const int uloc = ns, vloc = ns + 1, wloc = ns + 2, tloc = ns + 3;
  const double rho = Q[ns + 5], T = Q[tloc], P = Q[ns + 4];
  const double u = Q[uloc], v = Q[vloc], w = Q[wloc];
  const double rvOverDt = rho * vOverDt;
  double v2 = u * u + v * v + w * w;
  double cv, cp, R, gamma, c2;
  fr_props(f, Q, T, &cv, &cp, &R, &gamma, &c2);
  const double s_ref = f->ref_velocity * f->ref_velocity / c->ref_temperature;
  const double rhoDim = rho * c->ref_density, TDim = T * c->ref_temperature, PDim = P * f->ref_pressure;
  v2 *= f->ref_velocity * f->ref_velocity;
  double rhoiDim[O_CHEM_MAXS], Yi[O_CHEM_MAXS], dEtdRhoi[O_CHEM_MAXS];
  for (int i = 0; i < ns; i++) {
    rhoiDim[i] = Q[i] * c->ref_density;
    Yi[i] = Q[i] / rho;
  }
  /* ChemModel::dEtdP_dEtdRhoi (ucs/chem.tcc:828-858) with the ideal-gas
   * derivatives of ucs/EOS.tcc:42-70 */
  double dEtdP;
  {
    const double hv2 = 0.5 * v2;
    double rhomix = 0.0, Rmix = 0.0;
    for (int i = 0; i < ns; i++) {
      rhomix += rhoiDim[i];
      Rmix += rhoiDim[i] * species_R(c, i);
    }
    Rmix /= rhomix;
    const double dPdrhomix = Rmix * TDim;  /* GetdP_dRho */
    const double dPdRmix = rhomix * TDim;  /* GetdP_dR */
    for (int i = 0; i < ns; i++) {
      /* dRmixdRhoi (ucs/chem.tcc:861-870) */
      double d = species_R(c, i) * (rhomix - rhoiDim[i]) / (rhomix * rhomix);
      for (int j = 0; j < ns; j++) {
        if (j == i) continue;
        d -= species_R(c, j) * rhoiDim[j] / (rhomix * rhomix);
      }
      dEtdRhoi[i] = hv2 + species_h(c, i, TDim) - (dPdrhomix + dPdRmix * d);
    }
    dEtdP = 0.0;
    const double dTdP = 1.0 / (rhomix * Rmix);  /* GetdT_dP */
    for (int i = 0; i < ns; i++) dEtdP += rhoiDim[i] * (species_dhdt(c, i, TDim)) * dTdP;
    dEtdP -= 1.0;
  }
  (void)PDim;
  const double ref_detdrho = f->ref_velocity * f->ref_velocity * c->ref_density / c->ref_density;
  for (int i = 0; i < ns; i++) dEtdRhoi[i] /= ref_detdrho;
  const double ref_detdP = f->ref_velocity * f->ref_velocity * c->ref_density / f->ref_pressure;
  dEtdP /= ref_detdP;
  double dPdT = rhoDim * (R * s_ref);  /* GetdP_dT(R*s_ref, rhoDim, PDim, TDim) */
  dPdT /= (f->ref_pressure / c->ref_temperature);
  double c2i[O_CHEM_MAXS], thetaOBetai[O_CHEM_MAXS];
  for (int i = 0; i < ns; i++) c2i[i] = c2;
  const double oneOBeta = 1.0 / beta, oneMbeta = 1.0 - beta;
  for (int i = 0; i < ns; i++) thetaOBetai[i] = (Yi[i] * oneMbeta / c2i[i]) * oneOBeta;
  for (int i = 0; i < ns; i++) {
    A[neqn * i + i] += vOverDt;
    A[neqn * i + tloc] += thetaOBetai[i] * vOverDt * dPdT;
  }
  const int rows[3] = {uloc, vloc, wloc};
  const double vel[3] = {u, v, w};
  for (int k = 0; k < 3; k++) {
    double precond = 0.0;
    for (int j = 0; j < ns; j++) {
      A[neqn * rows[k] + j] += vel[k] * vOverDt;
      precond += thetaOBetai[j] * vel[k];
    }
    A[neqn * rows[k] + rows[k]] += rvOverDt;
    A[neqn * rows[k] + tloc] += precond * vOverDt * dPdT;
  }
  double precond = 0.0;
  for (int j = 0; j < ns; j++) {
    A[neqn * tloc + j] += dEtdRhoi[j] * vOverDt;
    precond += dEtdRhoi[j] * thetaOBetai[j];
  }
  A[neqn * tloc + uloc] += u * rvOverDt;
  A[neqn * tloc + vloc] += v * rvOverDt;
  A[neqn * tloc + wloc] += w * rvOverDt;
  precond += dEtdP * (oneOBeta);
  A[neqn * tloc + tloc] += precond * vOverDt * dPdT;
}

/* UpdateQinf (ucs/compressibleFR.tcc:1465-1518): Qinf from the mass
 * fractions, the initial pressure and temperature (SI) and the velocity
 * vector (non-dimensional); also sets Pref */
void o_fr_qinf(o_fr* f, const double* Y, double pDim, double TDim, const double* vel, double* Qinf) {
  const o_chem* c = f->chem;
  const int ns = c->ns;
  double rDim = 0.0;
  for (int i = 0; i < ns; i++) {
    const double piDim = pDim * Y[i];
    rDim += piDim / (species_R(c, i) * TDim);  /* IdealGasEOS::GetRho */
  }
  const double r = rDim / c->ref_density;
  for (int i = 0; i < ns; i++) Qinf[i] = Y[i] * r;
  Qinf[ns] = vel[0];
  Qinf[ns + 1] = vel[1];
  Qinf[ns + 2] = vel[2];
  Qinf[ns + 3] = TDim / c->ref_temperature;
  o_fr_aux_q(f, Qinf);
  f->Pref = Qinf[ns + 4];
}

/* ---- mesh-level phases ---------------------------------------------------- */

/* UpdateBCs (ucs/bc.tcc:1399-1457): qref per face tag, nv values each */
void o_fr_update_bcs(const o_fr* f, const o_mesh* m, const double* qref, double* q) {
  const int nv = o_fr_nvars(f);
  for (int k = 0; k < m->nbnode; k++) {
    const int L = m->bedges[2 * k], R = m->bedges[2 * k + 1], tag = m->bfactag[k];
    o_fr_bc_state(f, m->bc_type[tag], &q[(size_t)nv * L], &q[(size_t)nv * R], &qref[(size_t)nv * tag],
                  &m->bedge_a[4 * k], f->beta);
  }
}

/* SpatialResidual, first order (ucs/residual.tcc:64-114): interior HLLC
 * fluxes (b[L] -= F, b[R] += F), boundary fluxes (b[L] -= F), then
 * b += vol S; returns residGlobal and the per-equation norms */
double o_fr_residual(const o_fr* f, const o_mesh* m, const double* q, double* b, double* norms) {
  const int neqn = o_fr_neqn(f), nv = o_fr_nvars(f);
  const int nn = m->nnode;
  double F[O_CHEM_MAXS + 4];
  memset(b, 0, sizeof(double) * (size_t)neqn * nn);
  for (int e = 0; e < m->nedge; e++) {
    const int L = m->edges[2 * e], R = m->edges[2 * e + 1];
    o_fr_hllc(f, &q[(size_t)nv * L], &q[(size_t)nv * R], &m->edge_a[4 * e], f->beta, F);
    for (int k = 0; k < neqn; k++) {
      b[(size_t)neqn * L + k] += -F[k];
      b[(size_t)neqn * R + k] += F[k];
    }
  }
  for (int e = 0; e < m->nbnode; e++) {
    const int L = m->bedges[2 * e], R = m->bedges[2 * e + 1];
    o_fr_hllc(f, &q[(size_t)nv * L], &q[(size_t)nv * R], &m->bedge_a[4 * e], f->beta, F);
    for (int k = 0; k < neqn; k++) b[(size_t)neqn * L + k] += -F[k];
  }
  {
    /* SourceTerm is added for every node; it is zero without reactions */
    double S[O_CHEM_MAXS + 4];
    for (int i = 0; i < nn; i++) {
      if (f->rxn_on) o_chem_source(f->chem, &q[(size_t)nv * i], m->vol[i], S);
      else
        for (int k = 0; k < neqn; k++) S[k] = 0.0;
      for (int k = 0; k < neqn; k++) b[(size_t)neqn * i + k] += S[k];
    }
  }
  double tot = 0.0;
  for (int k = 0; k < neqn; k++) norms[1 + k] = 0.0;
  for (size_t i = 0; i < (size_t)nn * neqn; i++) tot += b[i] * b[i];
  for (int i = 0; i < nn; i++)
    for (int k = 0; k < neqn; k++) norms[1 + k] += b[(size_t)neqn * i + k] * b[(size_t)neqn * i + k];
  norms[0] = sqrt(tot) / ((double)nn * neqn);
  for (int k = 0; k < neqn; k++) norms[1 + k] = sqrt(norms[1 + k]) / (double)nn;
  return norms[0];
}

/* ComputeTimesteps (ucs/timestep.tcc:6-143), local time stepping */
double o_fr_timestep(const o_fr* f, const o_mesh* m, const double* q, double cfl, double* dt) {
  const int neqn = o_fr_neqn(f), nv = o_fr_nvars(f);
  double t[3 * O_CHEM_MAXS + 6];
  for (int i = 0; i < m->nnode; i++) dt[i] = 0.0;
  for (int e = 0; e < m->nedge; e++) {
    const int L = m->edges[2 * e], R = m->edges[2 * e + 1];
    for (int k = 0; k < neqn; k++) t[k] = 0.5 * (q[(size_t)nv * L + k] + q[(size_t)nv * R + k]);
    o_fr_aux_q(f, t);
    const double me = o_fr_max_eig(f, t, &m->edge_a[4 * e], f->beta) * m->edge_a[4 * e + 3];
    dt[L] += me;
    dt[R] += me;
  }
  for (int e = 0; e < m->nbnode; e++) {
    const int L = m->bedges[2 * e], R = m->bedges[2 * e + 1];
    for (int k = 0; k < neqn; k++) t[k] = 0.5 * (q[(size_t)nv * L + k] + q[(size_t)nv * R + k]);
    o_fr_aux_q(f, t);
    dt[L] += o_fr_max_eig(f, t, &m->bedge_a[4 * e], f->beta) * m->bedge_a[4 * e + 3];
  }
  dt[0] = cfl * (m->vol[0] / dt[0]);
  double dtmin = dt[0];
  for (int i = 1; i < m->nnode; i++) {
    dt[i] = cfl * (m->vol[i] / dt[i]);
    dtmin = fmin2(dtmin, dt[i]);
  }
  return dtmin;
}

/* ---- block-CRS with neqn x neqn blocks ------------------------------------ */

int o_fr_crs_init(const o_mesh* m, int neqn, o_fr_crs* A) {
  o_crs base;
  if (o_crs_init(m, &base) != 0) return -1;
  A->nnode = base.nnode;
  A->nblocks = base.nblocks;
  A->neqn = neqn;
  A->ia = base.ia;
  A->ja = base.ja;
  A->iau = base.iau;
  A->e_lr = base.e_lr;
  A->e_rl = base.e_rl;
  free(base.pv);
  free(base.M);
  A->pv = (int*)malloc(sizeof(int) * neqn * (size_t)A->nnode);
  A->M = (double*)calloc((size_t)A->nblocks * neqn * neqn, sizeof(double));
  A->ludiag = 0;
  return 0;
}

void o_fr_crs_free(o_fr_crs* A) {
  free(A->ia); free(A->ja); free(A->iau); free(A->e_lr); free(A->e_rl); free(A->pv); free(A->M);
  memset(A, 0, sizeof(*A));
}

static double* blk(o_fr_crs* A, int k) { return &A->M[(size_t)k * A->neqn * A->neqn]; }

static int find_block(const o_fr_crs* A, int row, int col) {
  for (int k = A->ia[row]; k < A->ia[row + 1]; k++)
    if (A->ja[k] == col) return k;
  return -1;
}

/* ComputeJacobians (ucs/jacobian.tcc:12-250): one-sided FD blocks of the
 * HLLC flux (h = 1e-8, aux recomputed for every perturbed state), the
 * boundary part with the BC states re-evaluated under perturbation
 * (jacobianBoundaryEval 0; q's phantom and hard-set states are updated as the
 * reference's does), column sums to the diagonal, minus the source
 * Jacobian, plus the temporal terms on dt. */
void o_fr_jacobian(const o_fr* f, const o_mesh* m, const double* qref, double* q, const double* dt, o_fr_crs* A) {
  const int neqn = o_fr_neqn(f), nv = o_fr_nvars(f), n2 = neqn * neqn;
  const double h = 1.0e-8;
  double QPL[3 * O_CHEM_MAXS + 6], QPR[3 * O_CHEM_MAXS + 6];
  double fS[O_CHEM_MAXS + 4], fL[O_CHEM_MAXS + 4], fR[O_CHEM_MAXS + 4];
  double tL[(O_CHEM_MAXS + 4) * (O_CHEM_MAXS + 4)], tR[(O_CHEM_MAXS + 4) * (O_CHEM_MAXS + 4)];
  memset(A->M, 0, sizeof(double) * (size_t)A->nblocks * n2);
  A->ludiag = 0; /* CRSMatrix::Blank (ucs/crsmatrix.tcc:506-510) */
  /* Kernel_NumJac (ucs/jacobian.tcc:253-304) */
  for (int e = 0; e < m->nedge; e++) {
    const int L = m->edges[2 * e], R = m->edges[2 * e + 1];
    const double* QL = &q[(size_t)nv * L];
    const double* QR = &q[(size_t)nv * R];
    const double* a = &m->edge_a[4 * e];
    o_fr_hllc(f, QL, QR, a, f->beta, fS);
    for (int i = 0; i < neqn; i++) {
      memcpy(QPL, QL, sizeof(double) * nv);
      memcpy(QPR, QR, sizeof(double) * nv);
      QPL[i] += h;
      QPR[i] += h;
      o_fr_aux_q(f, QPL);
      o_fr_aux_q(f, QPR);
      o_fr_hllc(f, QPL, QR, a, f->beta, fL);
      o_fr_hllc(f, QL, QPR, a, f->beta, fR);
      for (int j = 0; j < neqn; j++) tL[j * neqn + i] = (fS[j] - fL[j]) / h;
      for (int j = 0; j < neqn; j++) tR[j * neqn + i] = (fR[j] - fS[j]) / h;
    }
    double* pR = blk(A, find_block(A, L, R));
    double* pL = blk(A, find_block(A, R, L));
    for (int k = 0; k < n2; k++) {
      pR[k] += tR[k];
      pL[k] += tL[k];
    }
  }
  /* Bkernel_NumJac (ucs/jacobian.tcc:458-535) */
  for (int e = 0; e < m->nbnode; e++) {
    const int L = m->bedges[2 * e], R = m->bedges[2 * e + 1], tag = m->bfactag[e];
    double* QL = &q[(size_t)nv * L];
    double* QR = &q[(size_t)nv * R];
    const double* a = &m->bedge_a[4 * e];
    const double* Qref = &qref[(size_t)nv * tag];
    const int bct = m->bc_type[tag];
    o_fr_bc_state(f, bct, QL, QR, Qref, a, f->beta);
    o_fr_hllc(f, QL, QR, a, f->beta, fS);
    for (int i = 0; i < neqn; i++) {
      memcpy(QPL, QL, sizeof(double) * nv);
      memcpy(QPR, QR, sizeof(double) * nv);
      QPL[i] += h;
      QPR[i] += h;
      o_fr_aux_q(f, QPL);
      o_fr_aux_q(f, QPR);
      o_fr_hllc(f, QL, QPR, a, f->beta, fR);
      memcpy(QPR, QR, sizeof(double) * nv);
      o_fr_aux_q(f, QPR);
      o_fr_bc_state(f, bct, QPL, QPR, Qref, a, f->beta);
      o_fr_hllc(f, QPL, QPR, a, f->beta, fL);
      for (int j = 0; j < neqn; j++) tL[j * neqn + i] = (fL[j] - fS[j]) / h;
    }
    double* pD = blk(A, A->iau[L]);
    for (int k = 0; k < n2; k++) pD[k] += tL[k];
  }
  /* Kernel_Diag_NumJac (ucs/jacobian.tcc:436-456) */
  for (int e = 0; e < m->nedge; e++) {
    const int L = m->edges[2 * e], R = m->edges[2 * e + 1];
    const double* jacL = blk(A, find_block(A, R, L));
    const double* jacR = blk(A, find_block(A, L, R));
    double* dL = blk(A, A->iau[L]);
    double* dR = blk(A, A->iau[R]);
    for (int k = 0; k < n2; k++) {
      dL[k] += -jacL[k];
      dR[k] += -jacR[k];
    }
  }
  /* source Jacobian (ucs/jacobian.tcc:198-209) */
  if (f->rxn_on) {
    double S[(O_CHEM_MAXS + 4) * (O_CHEM_MAXS + 4)];
    for (int i = 0; i < m->nnode; i++) {
      o_chem_source_jacobian(f->chem, &q[(size_t)nv * i], m->vol[i], S);
      double* d = blk(A, A->iau[i]);
      for (int k = 0; k < n2; k++) d[k] -= S[k];
    }
  }
  /* ContributeTemporalTerms (ucs/jacobian.tcc:213-250) */
  for (int i = 0; i < m->nnode; i++) o_fr_temporal(f, &q[(size_t)nv * i], m->vol[i], dt[i], f->beta, blk(A, A->iau[i]));
}

/* PrepareSGS + SGS (ucs/crsmatrix.tcc:839-876, ucs/crs.tcc:61-173), x0 = 0;
 * order: forward row order (NULL = natural), the backward sweep reverses it.
 * Returns |xNorm - xNormOld| as the reference's SGS does. */
static void fr_sgs_row(o_fr_crs* A, int node, const double* b, double* x) {
  const int n = A->neqn;
  double rhs[O_CHEM_MAXS + 4], vout[O_CHEM_MAXS + 4], tmp[O_CHEM_MAXS + 4];
  memcpy(rhs, &b[(size_t)n * node], sizeof(double) * n);
  for (int kb = A->ia[node] + 1; kb < A->ia[node + 1]; kb++) {
    matvec(blk(A, kb), &x[(size_t)n * A->ja[kb]], vout, n);
    for (int j = 0; j < n; j++) rhs[j] -= vout[j];
  }
  o_lu_solve(blk(A, A->iau[node]), rhs, &A->pv[(size_t)n * node], tmp, n);
  memcpy(&x[(size_t)n * node], rhs, sizeof(double) * n);
}

double o_fr_sgs(o_fr_crs* A, int nsgs, const double* b, double* x, const int* order) {
  const int n = A->neqn, nn = A->nnode;
  /* PrepareSGS factors once per assembly (ludiag, ucs/crsmatrix.tcc:840-876) */
  if (!A->ludiag)
    for (int i = 0; i < nn; i++) o_lu(blk(A, A->iau[i]), &A->pv[(size_t)n * i], n);
  A->ludiag = 1;
  memset(x, 0, sizeof(double) * (size_t)n * nn);
  double xOld = 0.0, xNorm = 0.0;
  for (int it = 0; it < nsgs; it++) {
    for (int i = 0; i < nn; i++) fr_sgs_row(A, order ? order[i] : i, b, x);
    for (int i = nn - 1; i >= 0; i--) fr_sgs_row(A, order ? order[i] : i, b, x);
    xOld = xNorm;
    double s = 0.0;
    for (size_t i = 0; i < (size_t)nn * n; i++) s += x[i] * x[i];
    xNorm = sqrt(s) / (double)(nn * n);
  }
  return fabs(xOld - xNorm);
}

/* One PreTimeAdvance + NewtonIterate + iter++ at first order
 * (ucs/solutionSpace.tcc:572-804); returns residGlobal.  dt holds the time
 * step of the previous call (the Jacobian is built on it first). */
double o_fr_step(o_fr_solver* s, double* norms, int* nzeroed) {
  const o_fr* f = &s->f;
  const o_mesh* m = s->m;
  const int neqn = o_fr_neqn(f), nv = o_fr_nvars(f), nn = m->nnode;
  if (s->nsgs > 0 && (s->iter % s->jac_freq == 0 || s->iter == 1)) o_fr_jacobian(f, m, s->qref, s->q, s->dt, &s->A);
  s->cfl = (s->cfl_ramp > 0 && s->iter < s->cfl_ramp)
               ? (s->cfl_end - s->cfl_start) / (double)s->cfl_ramp * (double)s->iter + s->cfl_start
               : s->cfl_end;
  o_fr_timestep(f, m, s->q, s->cfl, s->dt);
  o_fr_update_bcs(f, m, s->qref, s->q);
  double nrm[1 + O_CHEM_MAXS + 4];
  const double r = o_fr_residual(f, m, s->q, s->b, nrm);
  o_fr_sgs(&s->A, s->nsgs, s->b, s->x, s->order);
  int zc = 0;
  for (int i = 0; i < nn; i++) {
    int z = 0;
    for (int k = 0; k < neqn; k++)
      if (isnan(s->x[(size_t)neqn * i + k]) || isinf(s->x[(size_t)neqn * i + k])) z = 1;
    if (z) {
      zc++;
      for (int k = 0; k < neqn; k++) s->x[(size_t)neqn * i + k] = 0.0;
    }
  }
  for (int i = 0; i < nn; i++) o_fr_apply_dq(f, &s->x[(size_t)neqn * i], &s->q[(size_t)nv * i]);
  if (nzeroed) *nzeroed = zc;
  if (norms) memcpy(norms, nrm, sizeof(double) * (1 + neqn));
  s->iter++;
  return r;
}

int o_fr_solver_init(o_fr_solver* s, const o_mesh* m, const o_fr* f, const double* qref, const double* q0,
                     double cfl, double cfl_start, int cfl_ramp, int nsgs, int jac_freq) {
  memset(s, 0, sizeof(*s));
  s->m = m;
  s->f = *f;
  const int neqn = o_fr_neqn(f), nv = o_fr_nvars(f);
  const size_t nt = (size_t)m->nnode + m->nbnode;
  s->q = (double*)malloc(sizeof(double) * nv * nt);
  memcpy(s->q, q0, sizeof(double) * nv * nt);
  for (size_t i = 0; i < nt; i++) o_fr_aux_q(f, &s->q[nv * i]);
  s->qref = (double*)malloc(sizeof(double) * nv * (size_t)m->nfactag);
  memcpy(s->qref, qref, sizeof(double) * nv * (size_t)m->nfactag);
  s->b = (double*)calloc((size_t)neqn * m->nnode, sizeof(double));
  s->x = (double*)calloc((size_t)neqn * m->nnode, sizeof(double));
  s->dt = (double*)calloc((size_t)m->nnode, sizeof(double));
  s->cfl_end = cfl;
  s->cfl_start = cfl_start;
  s->cfl_ramp = cfl_ramp;
  s->nsgs = nsgs;
  s->jac_freq = jac_freq > 0 ? jac_freq : 1;
  o_fr_crs_init(m, neqn, &s->A);
  /* PreIterate (ucs/solutionSpace.tcc:509-569): the counter starts at 0 */
  s->iter = 0;
  s->cfl = (cfl_ramp > 0) ? cfl_start : cfl;
  o_fr_timestep(f, m, s->q, s->cfl, s->dt);
  return 0;
}

void o_fr_solver_free(o_fr_solver* s) {
  free(s->q); free(s->qref); free(s->b); free(s->x); free(s->dt);
  o_fr_crs_free(&s->A);
  memset(s, 0, sizeof(*s));
}
