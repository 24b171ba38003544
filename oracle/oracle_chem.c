/*
 * CPU oracle, finite-rate chemistry (test infrastructure only).
 *
 * Restates the reference's mass production rates and the finite-rate
 * source term with its finite-difference Jacobian, in the reference's
 * loop order:
 *   Species::GetThermoCoeff             ucs/species.tcc:98-137
 *   Reaction::GetForwardReactionRate    ucs/reaction.tcc:607-624 (+ rate laws :708-760)
 *   Reaction::GetEquilibriumReactionRate ucs/reaction.tcc:628-679
 *   Reaction::GetBackwardReactionRate   ucs/reaction.tcc:682-706
 *   Reaction::GetMassProductionRate     ucs/reaction.tcc:762-856
 *   ChemModel::GetMassProductionRates   ucs/chem.tcc:575-583
 *   CompressibleFREqnSet::SourceTerm    ucs/compressibleFR.tcc:1276-1312
 *   EqnSet::SourceTermJacobian          ucs/eqnset.tcc:162-187
 * Constants: UNIV_R = 8.31447215 J/(mol K) (ucs/chem_constants.h:5),
 * Pref = 101325 Pa (ucs/reaction.tcc:667).  Molecular weights in kg/mol
 * (ucs/species.tcc:167-171), concentrations in mol/m^3.
 */
#include <math.h>
#include <string.h>

#include "oracle.h"

#define UNIV_R 8.31447215

/* isWholeNumber: ucs/macros.h:24-30 */
static int whole(double x) {
  int up = (int)(x + 0.5);
  int flat = (int)x;
  return up == flat;
}

/* pow(Type, Int) when whole, else pow(Type, Type): both are the C pow of
 * doubles once the integer is promoted */
static double powr(double x, double e) { return whole(e) ? pow(x, (double)(int)e) : pow(x, e); }

static void thermo_coeff(const o_chem* c, int s, double T, double* a) {
  const double* src = (T < 200.0) ? c->lo[s] : (T > 6000.0) ? c->hi[s] : (T > 1000.0 ? c->hi[s] : c->lo[s]);
  for (int i = 0; i < 7; i++) a[i] = src[i];
}

double o_chem_kf(const o_chem* c, int r, double T) {
  double A = c->A[r], EA = c->EA[r], n = c->n[r];
  switch (c->type[r]) {
    case 0: return (A * exp(-EA / (UNIV_R * T)));
    case 1: return (A * pow(T, n) * exp(-EA / (UNIV_R * T)));
    case 2: return (A * pow(T, n) * exp(-EA / T));
    case 3: return (A * pow(T, n));
  }
  return -999.0;
}

double o_chem_kc(const o_chem* c, int r, double T) {
  double a[7];
  double nu = 0.0, d1 = 0.0, d2 = 0.0, d3 = 0.0, d4 = 0.0, d5 = 0.0, d6 = 0.0, d7 = 0.0;
  for (int i = 0; i < c->nloc[r]; i++) {
    int ii = c->loc[r][i];
    thermo_coeff(c, ii, T, a);
    double dnu = c->nupp[r][i] - c->nup[r][i];
    nu += dnu;
    d1 += dnu * a[0];
    d2 += dnu * a[1];
    d3 += dnu * a[2];
    d4 += dnu * a[3];
    d5 += dnu * a[4];
    d6 += dnu * a[5];
    d7 += dnu * a[6];
  }
  double Kp = exp(d1 * (log(T) - 1.0) + T * (d2 / 2.0 + T * (d3 / 6.0 + T * (d4 / 12.0 + d5 / 20.0 * T))) - d6 / T + d7);
  double Kc = Kp;
  if (!(fabs(nu) < 1.0e-15)) {
    double Pref = 101325.0;
    if (whole(nu)) Kc *= pow(Pref / (UNIV_R * T), (double)(int)nu);
    else Kc *= pow(Pref / (UNIV_R * T), nu);
  }
  return Kc;
}

/* GetMassProductionRate of one reaction for one global species (kg/m^3/s) */
static double mass_production(const o_chem* c, int r, const double* rhoi, double T, int sid) {
  int kk = -1;
  for (int i = 0; i < c->nloc[r]; i++)
    if (c->loc[r][i] == sid) {
      kk = i;
      break;
    }
  double nu_ir = kk >= 0 ? c->nupp[r][kk] - c->nup[r][kk] : 0.0;
  /* AlmostEqualRelative(nu, 0, 1e-15) holds only for nu == 0 */
  if (fabs(nu_ir - 0.0) <= fabs(nu_ir) * 1.0e-15) return 0.0;
  double X[O_CHEM_MAXS];
  for (int k = 0; k < c->ns; k++) X[k] = rhoi[k] / c->mw[k];
  double Kf = o_chem_kf(c, r, T);
  double Kb = Kf / o_chem_kc(c, r, T);
  double prod_form = 1.0, prod_dest = 1.0;
  for (int k = 0; k < c->nloc[r]; k++) {
    int g = c->loc[r][k];
    prod_form *= powr(X[g], c->nup[r][k]);
    prod_dest *= powr(X[g], c->nupp[r][k]);
  }
  double Mconc = 1.0;
  if (c->tb[r]) {
    Mconc = 0.0;
    for (int k = 0; k < c->ns; k++) Mconc += X[k];
    for (int k = 0; k < c->nloc[r]; k++) Mconc += (c->eff[r][k] - 1.0) * X[c->loc[r][k]];
  }
  double net = Kf * prod_form - Kb * prod_dest;
  double wdot = nu_ir * Mconc * net;
  wdot *= c->mw[sid];
  return wdot;
}

void o_chem_wdot(const o_chem* c, const double* rhoi, double T, double* wdot) {
  for (int i = 0; i < c->ns; i++) {
    wdot[i] = 0.0;
    for (int j = 0; j < c->nr; j++) wdot[i] += mass_production(c, j, rhoi, T, i);
  }
}

/* FR state: rho_i (ns), u, v, w, T -- non-dimensional */
void o_chem_source(const o_chem* c, const double* Q, double vol, double* source) {
  int ns = c->ns, neqn = ns + 4;
  double rhoi[O_CHEM_MAXS], wdot[O_CHEM_MAXS];
  for (int i = 0; i < neqn; i++) source[i] = 0.0;
  double T = Q[ns + 3] * c->ref_temperature;
  for (int i = 0; i < ns; i++) rhoi[i] = Q[i] * c->ref_density;
  o_chem_wdot(c, rhoi, T, wdot);
  for (int i = 0; i < ns; i++) {
    wdot[i] /= (c->ref_density / c->ref_time);
    source[i] = vol * wdot[i];
  }
}

void o_chem_source_jacobian(const o_chem* c, const double* Q, double vol, double* A) {
  int neqn = c->ns + 4;
  double s[O_CHEM_MAXS + 4], sp[O_CHEM_MAXS + 4], QP[O_CHEM_MAXS + 4];
  const double h = 1.0e-8;
  for (int i = 0; i < neqn * neqn; i++) A[i] = 0.0;
  o_chem_source(c, Q, vol, s);
  for (int i = 0; i < neqn; i++) {
    memcpy(QP, Q, sizeof(double) * neqn);
    QP[i] += h;
    o_chem_source(c, QP, vol, sp);
    for (int j = 0; j < neqn; j++) A[j * neqn + i] = (sp[j] - s[j]) / h;
  }
}

/* Species::GetCp (ucs/species.tcc:43-55), J/(kg K) */
double o_chem_cp(const o_chem* c, int s, double T) {
  double a[7];
  thermo_coeff(c, s, T, a);
  double cp_R = a[0] + T * (a[1] + T * (a[2] + T * (a[3] + T * a[4])));
  return cp_R * (UNIV_R / c->mw[s]);
}

/* ChemModel::GetP (ucs/chem.tcc:988-997) with IdealGasEOS::GetP
 * (ucs/EOS.tcc:36-38): P = sum_i (rho_i R_i) T, SI */
static double fr_p(const o_chem* c, const double* rhoiDim, double TDim) {
  double P = 0.0;
  for (int i = 0; i < c->ns; i++) P += rhoiDim[i] * (UNIV_R / c->mw[i]) * TDim;
  return P;
}

/* CompressibleFREqnSet::ComputeAuxiliaryVariables (ucs/compressibleFR.tcc:
 * 754-814): aux = p, rho, cv_1..cv_ns (Mayer, ucs/EOS.tcc:79-82, scaled by
 * ref_velocity^2 / T_ref), conc_1..conc_ns (MassToMole, ucs/chem.tcc:980-986,
 * on the non-dimensional rho_i as the reference passes them) */
void o_fr_aux(const o_chem* c, const double* Q, double ref_pressure, double ref_velocity, double* aux) {
  int ns = c->ns;
  double rho = 0.0, rd[O_CHEM_MAXS];
  for (int i = 0; i < ns; i++) {
    rho += Q[i];
    rd[i] = Q[i] * c->ref_density;
  }
  double TDim = Q[ns + 3] * c->ref_temperature;
  aux[0] = fr_p(c, rd, TDim) / ref_pressure;
  aux[1] = rho;
  double s_ref = ref_velocity * ref_velocity / c->ref_temperature;
  for (int i = 0; i < ns; i++) {
    aux[2 + i] = (o_chem_cp(c, i, TDim) - UNIV_R / c->mw[i]) / s_ref;
    aux[2 + ns + i] = Q[i] / c->mw[i] / 1000.0;
  }
}

/* CompressibleFREqnSet::NewtonFindTGivenP (ucs/compressibleFR.tcc:2342-2378) */
double o_fr_t_from_p(const o_chem* c, const double* rhoi, double p, double Tinit, double ref_pressure, int* iters) {
  const int maxit = 28;
  const double tol = 1.0e-15, Tref = c->ref_temperature;
  double rd[O_CHEM_MAXS];
  for (int i = 0; i < c->ns; i++) rd[i] = rhoi[i] * c->ref_density;
  double TDim = Tinit * Tref, PgoalDim = p * ref_pressure;
  int j;
  for (j = 0; j < maxit; j++) {
    double TpDim = TDim + 1.0e-8;
    double z = PgoalDim - fr_p(c, rd, TDim);
    double zp = PgoalDim - fr_p(c, rd, TpDim);
    double dzdT = (zp - z) / (TpDim - TDim);
    double dT = -z / dzdT;
    if (fabs(dT / Tref) < tol) break;
    TDim += dT;
  }
  if (iters) *iters = j;
  return TDim / Tref;
}
