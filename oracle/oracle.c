/*
 * CPU oracle -- TEST INFRASTRUCTURE ONLY.  Clean-room FP64 restatement of the
 * reference ucs.x implicit edge-loop path (see oracle.h).  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it, as the
 * checker; the product (proteuscfd_amd/) never links or calls it.
 *
 * Parity pins (see tests/test_oracle.py): Roe flux consistency F(Q,Q)=F_phys,
 * T*Tinv=I, supersonic upwinding, freestream preservation, conservation,
 * LSQ exactness on linear fields, Venkat closed forms, dense LU vs numpy,
 * SGS vs a textbook block Gauss-Seidel, FD Jacobian vs complex-step.
 */
#define _POSIX_C_SOURCE 199309L
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#define NE O_NEQN
#define NV O_NVARS
#define NT O_NTERMS

static inline double omax(double x, double y) { return x > y ? x : y; }  /* ucs/macros.h:32-36 */
static inline double omin(double x, double y) { return x < y ? x : y; }  /* ucs/macros.h:38-42 */

static int cmp_int(const void* a, const void* b);

static double now_s(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

/* ComputeAuxiliaryVariables: ucs/compressible.tcc:1229-1243, with
 * ComputeTemperature T = gamma*P/rho (ucs/compressible.tcc:1103-1110) */
void o_aux(double* Q, double gamma) {
  double gm1 = gamma - 1.0;
  double u = Q[1] / Q[0];
  double v = Q[2] / Q[0];
  double w = Q[3] / Q[0];
  double V2 = u * u + v * v + w * w;
  double P = gm1 * (Q[4] - 0.5 * Q[0] * V2);
  Q[6] = P;
  {
    /* ComputeTemperature recomputes P through ComputePressure */
    double r = Q[0];
    double uu = Q[1] / r, vv = Q[2] / r, ww = Q[3] / r;
    double v2h = 0.5 * (uu * uu + vv * vv + ww * ww);
    double p = gm1 * (Q[4] - r * v2h);
    Q[5] = gamma * p / r;
  }
  Q[7] = u;
  Q[8] = v;
  Q[9] = w;
}

static double pressure_of(const double* Q, double gamma) { /* ucs/compressible.tcc:1090-1101 */
  double r = Q[0];
  double u = Q[1] / r, v = Q[2] / r, w = Q[3] / r;
  double v2h = 0.5 * (u * u + v * v + w * w);
  return (gamma - 1.0) * (Q[4] - r * v2h);
}

static double theta_of(const double* Q, const double* a, double vdotn) { /* ucs/compressible.tcc:995-1006 */
  double u = Q[1] / Q[0], v = Q[2] / Q[0], w = Q[3] / Q[0];
  return u * a[0] + v * a[1] + w * a[2] + vdotn;
}

/* RoeVariables: ucs/compressible.tcc:533-578 */
static void roe_variables(const double* QL, const double* QR, double gamma, double* Qroe) {
  double gm1 = gamma - 1.0;
  double rhoL = QL[0], rhoR = QR[0];
  double uL = QL[1] / QL[0], uR = QR[1] / QR[0];
  double vL = QL[2] / QL[0], vR = QR[2] / QR[0];
  double wL = QL[3] / QL[0], wR = QR[3] / QR[0];
  double EL = QL[4], ER = QR[4];
  double v2L = uL * uL + vL * vL + wL * wL;
  double v2R = uR * uR + vR * vR + wR * wR;
  double PL = gm1 * (EL - 0.5 * rhoL * v2L);
  double PR = gm1 * (ER - 0.5 * rhoR * v2R);
  double hL = (EL + PL) / rhoL;
  double hR = (ER + PR) / rhoR;
  double rho = sqrt(rhoL * rhoR);
  double sigma = rho / (rhoL + rho);
  double u = uL + sigma * (uR - uL);
  double v = vL + sigma * (vR - vL);
  double w = wL + sigma * (wR - wL);
  double h = hL + sigma * (hR - hL);
  double v2h = 0.5 * (u * u + v * v + w * w);
  Qroe[0] = rho;
  Qroe[1] = rho * u;
  Qroe[2] = rho * v;
  Qroe[3] = rho * w;
  Qroe[4] = rho / gamma * (h + gm1 * v2h);
}

/* Eigensystem: ucs/compressible.tcc:580-684 (right T and left Tinv eigenvectors) */
static void eigensystem(const double* Q, const double* avec, double vdotn, double gamma, double* eig,
                        double* T, double* Ti) {
  double nx = avec[0], ny = avec[1], nz = avec[2];
  double rho = Q[0], ru = Q[1], rv = Q[2], rw = Q[3], rE = Q[4];
  double u = ru / rho, v = rv / rho, w = rw / rho;
  double gm1 = gamma - 1.0;
  double thetaf = u * nx + v * ny + w * nz;
  double theta = thetaf + vdotn;
  double v2h = 0.5 * (u * u + v * v + w * w);
  double P = gm1 * (rE - rho * v2h);
  double c2 = gamma * P / rho;
  double c = sqrt(c2);
  T[0] = nx;  T[5] = u * nx;  T[10] = v * nx + rho * nz;  T[15] = w * nx - rho * ny;  T[20] = v2h * nx + rho * (v * nz - w * ny);
  T[1] = ny;  T[6] = u * ny - rho * nz;  T[11] = v * ny;  T[16] = w * ny + rho * nx;  T[21] = v2h * ny + rho * (w * nx - u * nz);
  T[2] = nz;  T[7] = u * nz + rho * ny;  T[12] = v * nz - rho * nx;  T[17] = w * nz;  T[22] = v2h * nz + rho * (u * ny - v * nx);
  T[3] = rho / c;  T[8] = rho * (u / c + nx);  T[13] = rho * (v / c + ny);  T[18] = rho * (w / c + nz);
  T[23] = rho * (v2h / c + thetaf + c / gm1);
  T[4] = rho / c;  T[9] = rho * (u / c - nx);  T[14] = rho * (v / c - ny);  T[19] = rho * (w / c - nz);
  T[24] = rho * (v2h / c - thetaf + c / gm1);

  Ti[0] = nx - nz * v / rho + ny * w / rho - nx / c2 * v2h * gm1;
  Ti[1] = nx / c2 * u * gm1;
  Ti[2] = nz / rho + nx / c2 * v * gm1;
  Ti[3] = -ny / rho + nx / c2 * w * gm1;
  Ti[4] = -nx / c2 * gm1;
  Ti[5] = ny + nz * u / rho - nx * w / rho - ny / c2 * v2h * gm1;
  Ti[6] = -nz / rho + ny / c2 * u * gm1;
  Ti[7] = ny / c2 * v * gm1;
  Ti[8] = nx / rho + ny / c2 * w * gm1;
  Ti[9] = -ny / c2 * gm1;
  Ti[10] = nz - ny * u / rho + nx * v / rho - nz / c2 * v2h * gm1;
  Ti[11] = ny / rho + nz / c2 * u * gm1;
  Ti[12] = -nx / rho + nz / c2 * v * gm1;
  Ti[13] = nz / c2 * w * gm1;
  Ti[14] = -nz / c2 * gm1;
  Ti[15] = -0.5 / rho * (thetaf - gm1 * v2h / c);
  Ti[16] = 0.5 / rho * (nx - gm1 * u / c);
  Ti[17] = 0.5 / rho * (ny - gm1 * v / c);
  Ti[18] = 0.5 / rho * (nz - gm1 * w / c);
  Ti[19] = 0.5 / rho * (gm1 / c);
  Ti[20] = 0.5 / rho * (thetaf + gm1 * v2h / c);
  Ti[21] = -0.5 / rho * (nx + gm1 * u / c);
  Ti[22] = -0.5 / rho * (ny + gm1 * v / c);
  Ti[23] = -0.5 / rho * (nz + gm1 * w / c);
  Ti[24] = +0.5 / rho * (gm1 / c);
  eig[0] = theta;
  eig[1] = theta;
  eig[2] = theta;
  eig[3] = theta + c;
  eig[4] = theta - c;
}

/* Flux: ucs/compressible.tcc:686-710 */
static void phys_flux(const double* Q, const double* avec, double vdotn, double gamma, double* flux) {
  double rho = Q[0], ru = Q[1], rv = Q[2], rw = Q[3], rEt = Q[4];
  double u = ru / rho, v = rv / rho, w = rw / rho;
  double v2h = 0.5 * (u * u + v * v + w * w);
  double gm1 = gamma - 1.0;
  double P = gm1 * (rEt - rho * v2h);
  double ht = (rEt + P) / rho;
  double rhotheta = rho * (avec[0] * u + avec[1] * v + avec[2] * w + vdotn);
  flux[0] = rhotheta;
  flux[1] = (u * rhotheta + P * avec[0]);
  flux[2] = (v * rhotheta + P * avec[1]);
  flux[3] = (w * rhotheta + P * avec[2]);
  flux[4] = (ht * rhotheta - vdotn * P);
}

static void matvec5(const double* a, const double* v, double* out) { /* ucs/matrix.h:63-74 */
  for (int i = 0; i < 5; i++) {
    out[i] = a[i * 5 + 0] * v[0];
    for (int j = 1; j < 5; j++) out[i] += a[i * 5 + j] * v[j];
  }
}

/* RoeFlux: ucs/compressible.tcc:92-230, with EqnSet::NumericalFlux NaN
 * kneecapping (ucs/eqnset.tcc:54-90).  Returns 1 when a NaN was zeroed. */
int o_roe_flux(const double* QL, const double* QR, const double* avec, double vdotn, double gamma,
               int symmetry2d, double* flux) {
  double Qroe[5], T[25], Ti[25], eig[5], fL[5], fR[5], dQ[5], dv[5], dr[5];
  double area = avec[3];
  roe_variables(QL, QR, gamma, Qroe);
  eigensystem(Qroe, avec, vdotn, gamma, eig, T, Ti);
  double gm1 = gamma - 1.0;
  double rhoL = QL[0], uL = QL[1] / rhoL, vL = QL[2] / rhoL, wL = QL[3] / rhoL, EL = QL[4];
  double PL = gm1 * (EL - 0.5 * rhoL * (uL * uL + vL * vL + wL * wL));
  double rhoR = QR[0], uR = QR[1] / rhoR, vR = QR[2] / rhoR, wR = QR[3] / rhoR, ER = QR[4];
  double PR = gm1 * (ER - 0.5 * rhoR * (uR * uR + vR * vR + wR * wR));
  double thetaL = uL * avec[0] + vL * avec[1] + wL * avec[2] + vdotn;
  double thetaR = uR * avec[0] + vR * avec[1] + wR * avec[2] + vdotn;
  double cR = sqrt(gamma * PR / rhoR);
  double cL = sqrt(gamma * PL / rhoL);
  double eigL, eigR, e, eps;
  /* Harten-Hyman type-2 entropy fix, ucs/compressible.tcc:148-189 */
  eigL = thetaL; eigR = thetaR; e = eig[0];
  eps = omax(0.0, omax(e - eigL, eigR - e));
  if (fabs(eig[0]) < eps) {
    eig[0] = 0.5 * (eig[0] * eig[0] / eps + eps);
    eig[1] = eig[0];
    eig[2] = eig[0];
  } else {
    eig[0] = eig[1] = eig[2] = fabs(eig[0]);
  }
  eigL = thetaL + cL; eigR = thetaR + cR; e = eig[3];
  eps = omax(0.0, omax(e - eigL, eigR - e));
  if (fabs(eig[3]) < eps) eig[3] = 0.5 * (eig[3] * eig[3] / eps + eps);
  else eig[3] = fabs(eig[3]);
  eigL = thetaL - cL; eigR = thetaR - cR; e = eig[4];
  eps = omax(0.0, omax(e - eigL, eigR - e));
  if (fabs(eig[4]) < eps) eig[4] = 0.5 * (eig[4] * eig[4] / eps + eps);
  else eig[4] = fabs(eig[4]);
  for (int i = 0; i < 5; i++) dQ[i] = QR[i] - QL[i];
  matvec5(Ti, dQ, dv);
  for (int i = 0; i < 5; i++) dv[i] *= fabs(eig[i]);
  matvec5(T, dv, dr);
  phys_flux(QL, avec, vdotn, gamma, fL);
  phys_flux(QR, avec, vdotn, gamma, fR);
  for (int i = 0; i < 5; i++) flux[i] = 0.5 * area * (fL[i] + fR[i] - dr[i]);
  if (symmetry2d == 1) flux[1] = 0.0;
  else if (symmetry2d == 2) flux[2] = 0.0;
  else if (symmetry2d == 3) flux[3] = 0.0;
  int err = 0;
  for (int i = 0; i < 5; i++)
    if (isnan(flux[i])) {
      flux[i] = 0.0;
      err = 1;
    }
  return err;
}

/* GetFarfieldBoundaryVariables: ucs/compressible.tcc:1245-1372 (static mesh,
 * vdotn = 0), 10 fixed-point sub-iterations using the updated QR. */
static void farfield_state(const double* QL, double* QR, const double* Qinf, const double* avec,
                           double vdotn, double gamma) {
  double tmp[5];
  for (int subit = 0; subit < 10; subit++) {
    double u = vdotn * avec[0], v = vdotn * avec[1], w = vdotn * avec[2];
    for (int i = 0; i < NE; i++) tmp[i] = (QL[i] + QR[i]) / 2.0;
    double theta = theta_of(tmp, avec, vdotn);
    double rhob, ub, vb, wb, pb;
    double rhoi = QL[0];
    double ui = QL[1] / QL[0] + u, vi = QL[2] / QL[0] + v, wi = QL[3] / QL[0] + w;
    double pi = QL[6];
    double rhoinf = Qinf[0];
    double uinf = Qinf[1] / Qinf[0] + u, vinf = Qinf[2] / Qinf[0] + v, winf = Qinf[3] / Qinf[0] + w;
    double pinf = Qinf[6];
    double pavg = pressure_of(tmp, gamma);
    double rhoavg = tmp[0];
    double c2avg = gamma * (pavg / rhoavg);
    double cavg = sqrt(c2avg);
    double nx = avec[0], ny = avec[1], nz = avec[2];
    if (theta == 0.0) {
      return;
    } else if (theta > 0.0 && fabs(theta / cavg) >= 1.0) {
      for (int i = 0; i < NE; i++) QR[i] = QL[i];
    } else if (theta < 0.0 && fabs(theta / cavg) >= 1.0) {
      for (int i = 0; i < NE; i++) QR[i] = Qinf[i];
    } else if (theta > 0.0 && fabs(theta / cavg) < 1.0) {
      pb = pinf;
      rhob = rhoi + (pb - pi) / c2avg;
      double t = (pb - pi) / (rhoavg * cavg);
      ub = ui - nx * t; vb = vi - ny * t; wb = wi - nz * t;
      QR[0] = rhob; QR[1] = rhob * ub; QR[2] = rhob * vb; QR[3] = rhob * wb;
      QR[4] = pb / (gamma - 1.0) + 0.5 * rhob * (ub * ub + vb * vb + wb * wb);
    } else if (theta < 0.0 && fabs(theta / cavg) < 1.0) {
      pb = 0.5 * (pinf + pi + rhoavg * cavg * (nx * (uinf - ui) + ny * (vinf - vi) + nz * (winf - wi)));
      rhob = rhoinf + (pb - pinf) / c2avg;
      double t = (pb - pinf) / (rhoavg * cavg);
      ub = uinf + nx * t; vb = vinf + ny * t; wb = winf + nz * t;
      QR[0] = rhob; QR[1] = rhob * ub; QR[2] = rhob * vb; QR[3] = rhob * wb;
      QR[4] = pb / (gamma - 1.0) + 0.5 * rhob * (ub * ub + vb * vb + wb * wb);
    } else {
      return;
    }
  }
}

/* GetInviscidWallBoundaryVariables: ucs/compressible.tcc:1374-1472 (CVBC branch) */
static void wall_state(const double* QL, double* QR, const double* avec, double vdotn, double gamma) {
  double tmp[5];
  for (int subit = 0; subit < 10; subit++) {
    double u = vdotn * avec[0], v = vdotn * avec[1], w = vdotn * avec[2];
    double rhoi = QL[0];
    double ru = u * rhoi, rv = v * rhoi, rw = w * rhoi;
    for (int i = 0; i < NE; i++) tmp[i] = (QL[i] + QR[i]) / 2.0;
    double ui = QL[1] / QL[0] + ru, vi = QL[2] / QL[0] + rv, wi = QL[3] / QL[0] + rw;
    double pi = QL[6];
    double rhoavg = tmp[0];
    tmp[1] += rhoavg * u;
    tmp[2] += rhoavg * v;
    tmp[3] += rhoavg * w;
    double pavg = pressure_of(tmp, gamma);
    double c2avg = gamma * (pavg / rhoavg);
    double cavg = sqrt(c2avg);
    double nx = avec[0], ny = avec[1], nz = avec[2];
    double pb = pi + rhoavg * cavg * theta_of(QL, avec, vdotn);
    double rhob = rhoi + (pb - pi) / c2avg;
    double t = (pb - pi) / (rhoavg * cavg);
    double ub = ui - nx * t, vb = vi - ny * t, wb = wi - nz * t;
    QR[0] = rhob; QR[1] = rhob * ub; QR[2] = rhob * vb; QR[3] = rhob * wb;
    QR[4] = pb / (gamma - 1.0) + 0.5 * rhob * (ub * ub + vb * vb + wb * wb);
  }
}

/* GetViscousWallBoundaryVariables (ucs/compressible.tcc:1544-1575) for a
 * wall at rest (vel = 0, no bleed steps): isothermal walls set the energy
 * from Twall, adiabatic walls take density and energy from the wall-normal
 * neighbour; both sides hard-set the interior node too. */
static void noslip_state(double* QL, double* QR, double twall, const double* normalQ, double gamma) {
  const double vel[3] = {0.0, 0.0, 0.0};
  if (twall < 0.0) {
    QR[0] = QL[0] = normalQ[0];
    QR[4] = QL[4] = normalQ[4];
  } else {
    double v2 = vel[0] * vel[0] + vel[1] * vel[1] + vel[2] * vel[2];
    double gm1 = gamma - 1.0;
    double rhoEt = (twall * QL[0] / (gamma * gm1) + 0.5 * QR[0] * v2);
    QR[0] = QL[0];
    QR[4] = QL[4] = rhoEt;
  }
  QR[1] = QL[1] = QL[0] * vel[0];
  QR[2] = QL[2] = QL[0] * vel[1];
  QR[3] = QL[3] = QL[0] * vel[2];
}

void o_bc_state(int bctype, double* QL, double* QR, const double* Qref, const double* avec, double gamma) {
  o_bc_state_ex(bctype, QL, QR, Qref, avec, gamma, -1.0, NULL);
}

/* CalculateBoundaryVariables: ucs/bc.tcc:1058-1398 (types on this path) */
void o_bc_state_ex(int bctype, double* QL, double* QR, const double* Qref, const double* avec, double gamma,
                   double twall, const double* normalQ) {
  const double vdotn = 0.0;
  if (bctype == O_BC_PARALLEL) return;
  if (bctype == O_BC_NOSLIP) {
    if (twall < 0.0 && !normalQ) return;
    noslip_state(QL, QR, twall, normalQ, gamma);
    o_aux(QR, gamma);
    o_aux(QL, gamma);
    return;
  }
  if (bctype == O_BC_SONIC_INFLOW || bctype == O_BC_DIRICHLET) {
    for (int i = 0; i < NV; i++) QR[i] = QL[i] = Qref[i];
  } else if (bctype == O_BC_SONIC_OUTFLOW || bctype == O_BC_NEUMANN) {
    for (int i = 0; i < NE; i++) QR[i] = QL[i];
  } else if (bctype == O_BC_FARFIELD) {
    farfield_state(QL, QR, Qref, avec, vdotn, gamma);
  } else if (bctype == O_BC_IMPERMEABLE || bctype == O_BC_SYMMETRY) {
    wall_state(QL, QR, avec, vdotn, gamma);
  } else {
    return; /* unsupported on this path: leave the phantom untouched */
  }
  o_aux(QR, gamma);
  o_aux(QL, gamma);
}

static double tag_twall(const o_mesh* m, int tag) { return m->twall ? m->twall[tag] : -1.0; }

static const double* normal_state(const o_mesh* m, const double* q, int e) {
  int nn = m->bnormal ? m->bnormal[e] : -1;
  return nn >= 0 ? &q[NV * nn] : NULL;
}

/* UpdateBCs -> BdriverNoScatter(BC_Kernel): ucs/bc.tcc:1399-1457, 723-745 */
void o_update_bcs(const o_mesh* m, const o_params* p, double* q) {
  for (int e = 0; e < m->nbnode; e++) {
    int L = m->bedges[2 * e], R = m->bedges[2 * e + 1];
    int tag = m->bfactag[e];
    double nQ[NV];
    const double* ns = normal_state(m, q, e);
    if (ns) memcpy(nQ, ns, sizeof(nQ));
    o_bc_state_ex(m->bc_type[tag], &q[L * NV], &q[R * NV], &m->qref[tag * NV], &m->bedge_a[4 * e], p->gamma,
                  tag_twall(m, tag), ns ? nQ : NULL);
  }
}

/* Wall-normal neighbour: over the node's neighbours in ascending id, the
 * last one maximising -unit(x_pt - x_wall).a with a start of 0
 * (ucs/bc.tcc:1182-1210). */
void o_wall_normal_nodes(const o_mesh* m, int* bnormal) {
  int nn = m->nnode;
  int* cnt = (int*)calloc((size_t)nn + 1, sizeof(int));
  for (int e = 0; e < m->nedge; e++) {
    cnt[m->edges[2 * e] + 1]++;
    cnt[m->edges[2 * e + 1] + 1]++;
  }
  for (int i = 0; i < nn; i++) cnt[i + 1] += cnt[i];
  int* adj = (int*)malloc(sizeof(int) * (size_t)cnt[nn]);
  int* fill = (int*)malloc(sizeof(int) * (size_t)nn);
  memcpy(fill, cnt, sizeof(int) * (size_t)nn);
  for (int e = 0; e < m->nedge; e++) {
    int a = m->edges[2 * e], b = m->edges[2 * e + 1];
    adj[fill[a]++] = b;
    adj[fill[b]++] = a;
  }
  for (int i = 0; i < nn; i++) qsort(&adj[cnt[i]], (size_t)(cnt[i + 1] - cnt[i]), sizeof(int), cmp_int);
  for (int e = 0; e < m->nbnode; e++) {
    bnormal[e] = -1;
    if (m->bc_type[m->bfactag[e]] != O_BC_NOSLIP) continue;
    int L = m->bedges[2 * e];
    const double* wx = &m->xyz[3 * L];
    const double* a = &m->bedge_a[4 * e];
    double dotmax = 0.0;
    for (int k = cnt[L]; k < cnt[L + 1]; k++) {
      int pt = adj[k];
      double dx[3];
      for (int i = 0; i < 3; i++) dx[i] = m->xyz[3 * pt + i] - wx[i];
      double mag = sqrt(dx[0] * dx[0] + dx[1] * dx[1] + dx[2] * dx[2]);
      for (int i = 0; i < 3; i++) dx[i] = dx[i] / mag;
      double dot = 0.0;
      for (int i = 0; i < 3; i++) dot -= dx[i] * a[i];
      if (dot >= dotmax) {
        bnormal[e] = pt;
        dotmax = dot;
      }
    }
  }
  free(cnt);
  free(adj);
  free(fill);
}

/* ComputeNodeLSQCoefficients (weighted set sw): ucs/gradient.tcc:114-138, 460-503 */
void o_lsq_moments(const o_mesh* m, double* sw) {
  memset(sw, 0, sizeof(double) * 6 * (size_t)m->nnode);
  for (int e = 0; e < m->nedge; e++) {
    int L = m->edges[2 * e], R = m->edges[2 * e + 1];
    const double* x1 = &m->xyz[3 * L];
    const double* x2 = &m->xyz[3 * R];
    double dx[3] = {x1[0] - x2[0], x1[1] - x2[1], x1[2] - x2[2]};
    double ds2 = 1.0 / (dx[0] * dx[0] + dx[1] * dx[1] + dx[2] * dx[2]);
    double tl[6], tr[6];
    tl[0] = dx[0] * dx[0] * ds2;
    tl[1] = dx[0] * dx[1] * ds2;
    tl[2] = dx[0] * dx[2] * ds2;
    tl[3] = dx[1] * dx[1] * ds2;
    tl[4] = dx[1] * dx[2] * ds2;
    tl[5] = dx[2] * dx[2] * ds2;
    dx[0] = -dx[0]; dx[1] = -dx[1]; dx[2] = -dx[2];
    tr[0] = tl[0];
    tr[1] = (dx[0] * dx[1]) * ds2;
    tr[2] = (dx[0] * dx[2]) * ds2;
    tr[3] = tl[3];
    tr[4] = (dx[1] * dx[2]) * ds2;
    tr[5] = tl[5];
    for (int k = 0; k < 6; k++) {
      sw[6 * L + k] += tl[k];
      sw[6 * R + k] += tr[k];
    }
  }
}

/* ComputeLSQCoefficients: ucs/gradient.tcc:140-168 */
static void lsq_coeff(const double* s, const double* dxbar, double* we) {
  double s11 = s[0], s12 = s[1], s13 = s[2], s22 = s[3], s23 = s[4], s33 = s[5];
  double r11 = s11, r12 = s12, r13 = s13;
  double r12_r11 = (r11 == 0.0) ? 0.0 : r12 / r11;
  double r22 = s22 - r12 * r12_r11;
  double r23 = s23 - r12_r11 * r13;
  double r13_r11 = (r11 == 0.0) ? 0.0 : r13 / r11;
  double r23_r22 = (r22 == 0.0) ? 0.0 : r23 / r22;
  double r33 = s33 - r13 * r13_r11 - r23 * r23_r22;
  double dykdx = (dxbar[1] - (r12_r11)*dxbar[0]);
  we[2] = (r33 == 0.0) ? 0.0 : (dxbar[2] - r13_r11 * dxbar[0] - r23_r22 * dykdx) / r33;
  we[1] = (r22 == 0.0) ? 0.0 : (dykdx - r23 * we[2]) / r22;
  we[0] = (r11 == 0.0) ? 0.0 : (dxbar[0] - r12 * we[1] - r13 * we[2]) / r11;
}

static const int kGradList[NT] = {0, 1, 2, 3, 4, 5, 7, 8, 9}; /* ucs/compressible.tcc:1008-1027 */

/* Gradient::Compute (weighted LSQ): ucs/gradient.tcc:56-112, 250-316, 544-565 */
void o_gradient(const o_mesh* m, const double* q, const double* sw, double* grad) {
  memset(grad, 0, sizeof(double) * NT * 3 * (size_t)m->nnode);
  for (int e = 0; e < m->nedge; e++) {
    int L = m->edges[2 * e], R = m->edges[2 * e + 1];
    const double* xL = &m->xyz[3 * L];
    const double* xR = &m->xyz[3 * R];
    double dx[3] = {xL[0] - xR[0], xL[1] - xR[1], xL[2] - xR[2]};
    const double* qL = &q[NV * L];
    const double* qR = &q[NV * R];
    double weL[3], weR[3];
    double dx2 = dx[0] * dx[0] + dx[1] * dx[1] + dx[2] * dx[2];
    double weight = 1.0 / sqrt(dx2);
    dx[0] *= weight; dx[1] *= weight; dx[2] *= weight;
    lsq_coeff(&sw[6 * L], dx, weL);
    dx[0] = -dx[0]; dx[1] = -dx[1]; dx[2] = -dx[2];
    lsq_coeff(&sw[6 * R], dx, weR);
    for (int i = 0; i < NT; i++) {
      int k = kGradList[i];
      double dq = weight * (qR[k] - qL[k]);
      for (int j = 0; j < 3; j++) {
        grad[NT * 3 * L + 3 * i + j] += -weL[j] * dq;
        grad[NT * 3 * R + 3 * i + j] += +weR[j] * dq;
      }
    }
  }
  /* boundary half-edges contribute only across parallel ghosts (none here);
   * symmetry planes: remove the normal component, in half-edge order */
  for (int e = 0; e < m->nbnode; e++) {
    int tag = m->bfactag[e];
    if (m->bc_type[tag] != O_BC_SYMMETRY) continue;
    int L = m->bedges[2 * e];
    const double* a = &m->bedge_a[4 * e];
    double* g = &grad[NT * 3 * L];
    for (int i = 0; i < NT; i++) {
      double d = g[3 * i] * a[0] + g[3 * i + 1] * a[1] + g[3 * i + 2] * a[2];
      for (int j = 0; j < 3; j++) g[3 * i + j] -= d * a[j];
    }
  }
}

/* EqnSet::NumericalFlux (ucs/eqnset.tcc:56-88): fluxType 1 (algebraicFlux)
 * is the average of the two physical fluxes, 0.5*|a|*(F(QL)+F(QR)), with the
 * NaN components zeroed; otherwise the Roe flux. */
static void numflux(const o_params* p, const double* QL, const double* QR, const double* a, double* F) {
  if (p->flux_type == 1) {
    double f2[5];
    phys_flux(QL, a, 0.0, p->gamma, F);
    phys_flux(QR, a, 0.0, p->gamma, f2);
    for (int i = 0; i < NE; i++) {
      F[i] = 0.5 * a[3] * (F[i] + f2[i]);
      if (isnan(F[i])) F[i] = 0.0;
    }
    return;
  }
  o_roe_flux(QL, QR, a, 0.0, p->gamma, p->symmetry2d, F);
}

/* Green-Gauss gradient (gradientType 1): Kernel_Green_Gauss_Gradient and
 * Bkernel_Green_Gauss_Gradient (ucs/gradient.tcc:172-250) add the face
 * average times a*|a| to the left node and subtract it from the right one
 * (boundary half-edges: to the node, with the phantom as the other state),
 * then Gradient::Compute divides by the volume and applies the symmetry
 * fix (:68-95, 544-565). */
void o_gradient_gg(const o_mesh* m, const double* q, double* grad) {
  memset(grad, 0, sizeof(double) * NT * 3 * (size_t)m->nnode);
  for (int e = 0; e < m->nedge; e++) {
    int L = m->edges[2 * e], R = m->edges[2 * e + 1];
    const double* a = &m->edge_a[4 * e];
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < NT; j++) {
        int k = kGradList[j];
        double f = 0.5 * (q[NV * L + k] + q[NV * R + k]);
        double t = f * a[i] * a[3];
        grad[NT * 3 * L + 3 * j + i] += t;
        grad[NT * 3 * R + 3 * j + i] += -t;
      }
  }
  for (int e = 0; e < m->nbnode; e++) {
    int L = m->bedges[2 * e], R = m->bedges[2 * e + 1];
    const double* a = &m->bedge_a[4 * e];
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < NT; j++) {
        int k = kGradList[j];
        double f = 0.5 * (q[NV * L + k] + q[NV * R + k]);
        grad[NT * 3 * L + 3 * j + i] += f * a[i] * a[3];
      }
  }
  for (int n = 0; n < m->nnode; n++)
    for (int j = 0; j < NT * 3; j++) grad[NT * 3 * n + j] /= m->vol[n];
  for (int e = 0; e < m->nbnode; e++) {
    int tag = m->bfactag[e];
    if (m->bc_type[tag] != O_BC_SYMMETRY) continue;
    int L = m->bedges[2 * e];
    const double* a = &m->bedge_a[4 * e];
    double* g = &grad[NT * 3 * L];
    for (int i = 0; i < NT; i++) {
      double d = g[3 * i] * a[0] + g[3 * i + 1] * a[1] + g[3 * i + 2] * a[2];
      for (int j = 0; j < 3; j++) g[3 * i + j] -= d * a[j];
    }
  }
}

/* ExtrapolateVariables + ExtrapolateCorrection: ucs/compressible.tcc:1038-1052,
 * ucs/eqnset.h:231-238 */
static void extrapolate(double* Qho, const double* q, const double* dQ, const double* g, const double* dx,
                        const double* lim, double chi) {
  for (int i = 0; i < NE; i++) {
    double corr = 0.5 * chi * dQ[i] + (1.0 - chi) * (g[3 * i] * dx[0] + g[3 * i + 1] * dx[1] + g[3 * i + 2] * dx[2]);
    Qho[i] = q[i] + corr * lim[i];
  }
}

/* BadExtrapolation: ucs/compressible.tcc:1054-1078 */
static int bad_extrapolation(const double* Q, double gamma) {
  double r = Q[0];
  double u = Q[1] / r, v = Q[2] / r, w = Q[3] / r, E = Q[4];
  double v2h = 0.5 * (u * u + v * v + w * w);
  double pr = (gamma - 1.0) * (E - r * v2h);
  if (pr < 1.0e-10) return 1;
  if (r < 0.0) return 1;
  if (E < 1.0e-10) return 1;
  return 0;
}

static double venkat_phi(double y) { return (y * y + 2.0 * y) / (y * y + y + 2.0); }

/* Limiter::Compute: ucs/limiters.tcc:52-132 (min/max init 0, neighbours only),
 * Barth :193-295, Venkat :370-462, modified Venkat :533-648 (K configurable,
 * reference K=1), pressure clip :736-815, negatives -> 0 :118-124 */
void o_limiter(const o_mesh* m, const o_params* p, const double* q, const double* grad, double* l) {
  int nn = m->nnode;
  double* qmin = (double*)calloc((size_t)nn * NE, sizeof(double));
  double* qmax = (double*)calloc((size_t)nn * NE, sizeof(double));
  const double ones[5] = {1, 1, 1, 1, 1};
  for (int i = 0; i < nn * NE; i++) l[i] = 1.0;
  for (int e = 0; e < m->nedge; e++) {
    int L = m->edges[2 * e], R = m->edges[2 * e + 1];
    for (int i = 0; i < NE; i++) {
      qmax[NE * L + i] = omax(qmax[NE * L + i], q[NV * R + i]);
      qmin[NE * L + i] = omin(qmin[NE * L + i], q[NV * R + i]);
    }
    for (int i = 0; i < NE; i++) {
      qmax[NE * R + i] = omax(qmax[NE * R + i], q[NV * L + i]);
      qmin[NE * R + i] = omin(qmin[NE * R + i], q[NV * L + i]);
    }
  }
  const double Pi = 3.141592653589793;
  double K3 = p->venkat_K * p->venkat_K * p->venkat_K;
  if (p->limiter >= 1 && p->limiter <= 3) {
    for (int e = 0; e < m->nedge; e++) {
      int L = m->edges[2 * e], R = m->edges[2 * e + 1];
      const double* qL = &q[NV * L];
      const double* qR = &q[NV * R];
      const double* xL = &m->xyz[3 * L];
      const double* xR = &m->xyz[3 * R];
      double QL[5], QR[5], dQ[5], dx[3];
      for (int i = 0; i < NE; i++) dQ[i] = qR[i] - qL[i];
      dx[0] = 0.5 * (xR[0] - xL[0]);
      dx[1] = 0.5 * (xR[1] - xL[1]);
      dx[2] = 0.5 * (xR[2] - xL[2]);
      extrapolate(QL, qL, dQ, &grad[NT * 3 * L], dx, ones, p->chi);
      dx[0] = -dx[0]; dx[1] = -dx[1]; dx[2] = -dx[2];
      for (int i = 0; i < NE; i++) dQ[i] = -dQ[i];
      extrapolate(QR, qR, dQ, &grad[NT * 3 * R], dx, ones, p->chi);
      double* lL = &l[NE * L];
      double* lR = &l[NE * R];
      if (p->limiter == 1 || p->limiter == 2) {
        for (int side = 0; side < 2; side++) {
          const double* Qh = side ? QR : QL;
          const double* qq = side ? qR : qL;
          const double* mx = side ? &qmax[NE * R] : &qmax[NE * L];
          const double* mn = side ? &qmin[NE * R] : &qmin[NE * L];
          double* ll = side ? lR : lL;
          for (int j = 0; j < NE; j++) {
            double t = 1.0;
            if (Qh[j] > qq[j]) t = (mx[j] - qq[j]) / (Qh[j] - qq[j]);
            else if (Qh[j] < qq[j]) t = (mn[j] - qq[j]) / (Qh[j] - qq[j]);
            if (p->limiter == 2) {
              t = venkat_phi(t);
            } else {
              t = omax(0.0, t);
              t = omin(1.0, t);
            }
            ll[j] = omin(ll[j], t);
          }
        }
      } else {
        double Ll3 = 6.0 * Pi * m->vol[L];
        double Lr3 = 6.0 * Pi * m->vol[R];
        double DP = 0.0, DM;
        for (int j = 0; j < NE; j++) {
          DM = QL[j] - qL[j];
          if (QL[j] > qL[j]) DP = qmax[NE * L + j] - qL[j];
          else if (QL[j] < qL[j]) DP = qmin[NE * L + j] - qL[j];
          double ep2 = Ll3 * K3;
          double t = (DP * DP + ep2 + 2.0 * DM * DP) / (DP * DP + 2.0 * DM * DM + DM * DP + ep2);
          lL[j] = omin(lL[j], t);
        }
        for (int j = 0; j < NE; j++) {
          DM = QR[j] - qR[j];
          if (QR[j] > qR[j]) DP = qmax[NE * R + j] - qR[j];
          if (QR[j] <= qR[j]) DP = qmin[NE * R + j] - qR[j];
          double ep2 = Lr3 * K3;
          double t = (DP * DP + ep2 + 2.0 * DM * DP) / (DP * DP + 2.0 * DM * DM + DM * DP + ep2);
          lR[j] = omin(lR[j], t);
        }
      }
    }
  }
  if (p->limiter != 0) {
    /* pressure clip */
    double* lsnap = NULL;
    if (p->clip_jacobi) {
      lsnap = (double*)malloc(sizeof(double) * NE * (size_t)nn);
      memcpy(lsnap, l, sizeof(double) * NE * (size_t)nn);
    }
    const double* lread = p->clip_jacobi ? lsnap : l;
    for (int e = 0; e < m->nedge; e++) {
      int L = m->edges[2 * e], R = m->edges[2 * e + 1];
      const double* qL = &q[NV * L];
      const double* qR = &q[NV * R];
      const double* xL = &m->xyz[3 * L];
      const double* xR = &m->xyz[3 * R];
      double QL[5], QR[5], dQ[5], dx[3], Qroe[5];
      dx[0] = 0.5 * (xR[0] - xL[0]);
      dx[1] = 0.5 * (xR[1] - xL[1]);
      dx[2] = 0.5 * (xR[2] - xL[2]);
      for (int j = 0; j < NE; j++) dQ[j] = qR[j] - qL[j];
      extrapolate(QL, qL, dQ, &grad[NT * 3 * L], dx, &lread[NE * L], p->chi);
      if (bad_extrapolation(QL, p->gamma))
        for (int i = 0; i < NE; i++) l[NE * L + i] = 0.0;
      dx[0] = -dx[0]; dx[1] = -dx[1]; dx[2] = -dx[2];
      for (int j = 0; j < NE; j++) dQ[j] = -dQ[j];
      extrapolate(QR, qR, dQ, &grad[NT * 3 * R], dx, &lread[NE * R], p->chi);
      if (bad_extrapolation(QR, p->gamma))
        for (int i = 0; i < NE; i++) l[NE * R + i] = 0.0;
      roe_variables(QL, QR, p->gamma, Qroe);
      if (bad_extrapolation(Qroe, p->gamma))
        for (int i = 0; i < NE; i++) l[NE * L + i] = l[NE * R + i] = 0.0;
    }
    free(lsnap);
    for (int i = 0; i < nn * NE; i++)
      if (l[i] < 0.0) l[i] = 0.0;
  }
  free(qmin);
  free(qmax);
}

/* ComputeViscosity, Sutherland's law on the non-dimensional T = Q[5]
 * (ucs/eqnset.h:251-268) */
static double viscosity(const o_params* p, const double* Q) {
  double T = Q[5];
  double S = p->sutherland_s;
  double mu = (1.0 + S) * pow(T, 1.5) / (T + S);
  return mu;
}

/* ViscousFlux: ucs/compressible.tcc:712-795 (written on the right-hand side) */
void o_viscous_flux(const o_params* p, const double* Q, const double* grad, const double* avec, double* flux) {
  const double mut = 0.0;
  double Tx = grad[15], Ty = grad[16], Tz = grad[17];
  double ux = grad[18], uy = grad[19], uz = grad[20];
  double vx = grad[21], vy = grad[22], vz = grad[23];
  double wx = grad[24], wy = grad[25], wz = grad[26];
  double rho = Q[0];
  double u = Q[1] / rho, v = Q[2] / rho, w = Q[3] / rho;
  double mu = viscosity(p, Q);
  double tmut = (mu + mut);
  double fact = -2.0 / 3.0 * (ux + vy + wz);
  double tauxx = 2.0 * ux + fact;
  double tauyy = 2.0 * vy + fact;
  double tauzz = 2.0 * wz + fact;
  double tauxy = uy + vx;
  double tauxz = uz + wx;
  double tauyz = vz + wy;
  double RK = avec[3] / p->re_tilde;
  double RKT = RK * tmut;
  double cp = 1.0 / (p->gamma - 1.0);
  double k = mu / p->pr * cp;
  double kT = mut / p->prt * cp;
  double c1 = -(k + kT);
  double Tn = Tx * avec[0] + Ty * avec[1] + Tz * avec[2];
  double tauxn = tauxx * avec[0] + tauxy * avec[1] + tauxz * avec[2];
  double tauyn = tauxy * avec[0] + tauyy * avec[1] + tauyz * avec[2];
  double tauzn = tauxz * avec[0] + tauyz * avec[1] + tauzz * avec[2];
  flux[0] = 0.0;
  flux[1] = -RKT * (tauxn);
  flux[2] = -RKT * (tauyn);
  flux[3] = -RKT * (tauzn);
  flux[4] = -RKT * (tauxn * u + tauyn * v + tauzn * w) + RK * c1 * Tn;
  if (p->symmetry2d == 1) flux[1] = 0.0;
  else if (p->symmetry2d == 2) flux[2] = 0.0;
  else if (p->symmetry2d == 3) flux[3] = 0.0;
}

/* one side (Dx = the directional-derivative weights of that side) of
 * ViscousJacobian: ucs/compressible.tcc:1662-1857 */
static void viscous_jac_side(const double* Dx, const double* avec, double u, double v, double w, double RK,
                             double RKT, double Tn, double dP_dr, double P_over_rho, double dP_dru, double dP_drv,
                             double dP_drw, double dP_dret, double* a) {
  const double c43 = 4.0 / 3.0, mc23 = -2.0 / 3.0;
  double dux = -u * Dx[0], duy = -u * Dx[1], duz = -u * Dx[2];
  double dvx = -v * Dx[0], dvy = -v * Dx[1], dvz = -v * Dx[2];
  double dwx = -w * Dx[0], dwy = -w * Dx[1], dwz = -w * Dx[2];
  double dfact = -2.0 / 3.0 * (dux + dvy + dwz);
  double dtauxx = (2.0 * dux + dfact);
  double dtauyy = (2.0 * dvy + dfact);
  double dtauzz = (2.0 * dwz + dfact);
  double dtauxy = duy + dvx;
  double dtauxz = duz + dwx;
  double dtauyz = dvz + dwy;
  double dtauxn = dtauxx * avec[0] + dtauxy * avec[1] + dtauxz * avec[2];
  double dtauyn = dtauxy * avec[0] + dtauyy * avec[1] + dtauyz * avec[2];
  double dtauzn = dtauxz * avec[0] + dtauyz * avec[1] + dtauzz * avec[2];
  double d2u = (c43 * Dx[0] * avec[0] + Dx[1] * avec[1] + Dx[2] * avec[2]);
  double d2v = (mc23 * Dx[1] * avec[0] + Dx[0] * avec[1]);
  double d2w = (mc23 * Dx[2] * avec[0] + Dx[0] * avec[2]);
  double d3u = (mc23 * Dx[0] * avec[1] + Dx[1] * avec[0]);
  double d3v = (Dx[0] * avec[0] + c43 * Dx[1] * avec[1] + Dx[2] * avec[2]);
  double d3w = (mc23 * Dx[2] * avec[1] + Dx[1] * avec[2]);
  double d4u = (mc23 * Dx[0] * avec[2] + Dx[2] * avec[0]);
  double d4v = (mc23 * Dx[1] * avec[2] + Dx[2] * avec[1]);
  double d4w = (Dx[0] * avec[0] + Dx[1] * avec[1] + c43 * Dx[2] * avec[2]);
  for (int i = 0; i < 5; i++) a[i] = 0.0;
  a[5] = -RKT * (dtauxn);
  a[6] = -RKT * d2u;
  a[7] = -RKT * d2v;
  a[8] = -RKT * d2w;
  a[9] = 0.0;
  a[10] = -RKT * (dtauyn);
  a[11] = -RKT * d3u;
  a[12] = -RKT * d3v;
  a[13] = -RKT * d3w;
  a[14] = 0.0;
  a[15] = -RKT * (dtauzn);
  a[16] = -RKT * d4u;
  a[17] = -RKT * d4v;
  a[18] = -RKT * d4w;
  a[19] = 0.0;
  a[20] = -RKT * (dtauxn * u + dtauyn * v + dtauzn * w) + RK * Tn * (dP_dr - P_over_rho);
  a[21] = -RKT * (d2u * u + d3u * v + d4u * w) + RK * Tn * dP_dru;
  a[22] = -RKT * (d2v * u + d3v * v + d4v * w) + RK * Tn * dP_drv;
  a[23] = -RKT * (d2w * u + d3w * v + d4w * w) + RK * Tn * dP_drw;
  a[24] = RK * Tn * dP_dret;
}

/* ViscousJacobian: ucs/compressible.tcc:1632-1893.  aL = dR/dQL with the
 * sign flipped for the left node, aR = dR/dQR. */
void o_viscous_jacobian(const o_params* p, const double* QL, const double* QR, const double* dx, double s2,
                        const double* avec, double* aL, double* aR) {
  const double mut = 0.0;
  double Qavg[NV];
  for (int i = 0; i < NE; i++) Qavg[i] = 0.5 * (QL[i] + QR[i]);
  o_aux(Qavg, p->gamma);
  double mu = viscosity(p, Qavg);
  double tmut = (mu + mut);
  double RK = avec[3] / p->re_tilde;
  double RKT = RK * tmut;
  double rhoL = QL[0], uL = QL[1] / rhoL, vL = QL[2] / rhoL, wL = QL[3] / rhoL, PL = QL[6];
  double rhoR = QR[0], uR = QR[1] / rhoR, vR = QR[2] / rhoR, wR = QR[3] / rhoR, PR = QR[6];
  double u = 0.5 * (uL + uR), v = 0.5 * (vL + vR), w = 0.5 * (wL + wR);
  double DxL[3], DxR[3];
  for (int i = 0; i < 3; i++) {
    DxL[i] = -dx[i] / (rhoL * s2);
    DxR[i] = dx[i] / (rhoR * s2);
  }
  double gamma = p->gamma, gm1 = (gamma - 1.0);
  double v2L = (uL * uL + vL * vL + wL * wL);
  double v2R = (uR * uR + vR * vR + wR * wR);
  double dT_dPL = gamma / rhoL, dT_dPR = gamma / rhoR;
  double cp = 1.0 / (p->gamma - 1.0);
  double k = mu / p->pr * cp;
  double kT = mut / p->prt * cp;
  double c1 = -(k + kT);
  double TnL = (DxL[0] * avec[0] + DxL[1] * avec[1] + DxL[2] * avec[2]) * dT_dPL * c1;
  double TnR = (DxR[0] * avec[0] + DxR[1] * avec[1] + DxR[2] * avec[2]) * dT_dPR * c1;
  viscous_jac_side(DxR, avec, u, v, w, RK, RKT, TnR, +gm1 * 0.5 * v2R, PR / rhoR, -gm1 * uR, -gm1 * vR, -gm1 * wR,
                   +gm1, aR);
  viscous_jac_side(DxL, avec, u, v, w, RK, RKT, TnL, +gm1 * 0.5 * v2L, PL / rhoL, -gm1 * uL, -gm1 * vL, -gm1 * wL,
                   +gm1, aL);
  for (int i = 0; i < 25; i++) aL[i] = -aL[i];
  if (p->symmetry2d >= 1 && p->symmetry2d <= 3) {
    int r = p->symmetry2d;
    for (int i = 0; i < NE; i++) aL[r * NE + i] = aR[r * NE + i] = 0.0;
  }
}

/* Kernel_Viscous_Flux / Bkernel_Viscous_Flux (ucs/residual.tcc:389-562):
 * averaged state, averaged gradient with the directional correction along
 * the edge at second order (terms at gradientsLoc {0..5,7,8,9},
 * ucs/compressible.tcc:1009-1027); the boundary side uses the node's own
 * gradient.  b[R] += F, b[L] -= F; boundary b[L] -= F. */
static void viscous_residual(const o_mesh* m, const o_params* p, int iter, const double* q, const double* grad,
                             double* b) {
  static const int gloc[O_NTERMS] = {0, 1, 2, 3, 4, 5, 7, 8, 9};
  int ho = (p->sorder > 1 && iter > p->nfirst);
  for (int e = 0; e < m->nedge; e++) {
    int L = m->edges[2 * e], R = m->edges[2 * e + 1];
    const double* qL = &q[NV * L];
    const double* qR = &q[NV * R];
    double Qavg[NV], g[NT * 3], F[5];
    for (int i = 0; i < NE; i++) Qavg[i] = (qL[i] + qR[i]) / 2.0;
    o_aux(Qavg, p->gamma);
    const double* gL = &grad[NT * 3 * L];
    const double* gR = &grad[NT * 3 * R];
    for (int i = 0; i < NT * 3; i++) g[i] = 0.5 * (gL[i] + gR[i]);
    if (ho) {
      double dx[3], s2 = 0.0;
      for (int i = 0; i < 3; i++) {
        dx[i] = (m->xyz[3 * R + i] - m->xyz[3 * L + i]);
        s2 += dx[i] * dx[i];
      }
      for (int j = 0; j < NT; j++) {
        double qdots = dx[0] * g[j * 3] + dx[1] * g[j * 3 + 1] + dx[2] * g[j * 3 + 2];
        int loc = gloc[j];
        double dq = (qR[loc] - qL[loc] - qdots) / s2;
        for (int i = 0; i < 3; i++) g[j * 3 + i] += dq * dx[i];
      }
    }
    o_viscous_flux(p, Qavg, g, &m->edge_a[4 * e], F);
    for (int i = 0; i < NE; i++) {
      b[NE * L + i] += -F[i];
      b[NE * R + i] += F[i];
    }
  }
  for (int e = 0; e < m->nbnode; e++) {
    int L = m->bedges[2 * e], R = m->bedges[2 * e + 1];
    double Qavg[NV], F[5];
    for (int i = 0; i < NE; i++) Qavg[i] = 0.5 * (q[NV * L + i] + q[NV * R + i]);
    o_aux(Qavg, p->gamma);
    o_viscous_flux(p, Qavg, &grad[NT * 3 * L], &m->bedge_a[4 * e], F);
    for (int i = 0; i < NE; i++) b[NE * L + i] += -F[i];
  }
}

/* Bkernel_BC_Res_Modify -> ModifyViscousWallResidual (ucs/bc.tcc:902-1030,
 * ucs/compressible.tcc:1608-1630): hard-set rows of noSlip nodes */
static void noslip_residual(const o_mesh* m, double* b) {
  for (int e = 0; e < m->nbnode; e++) {
    int tag = m->bfactag[e];
    if (m->bc_type[tag] != O_BC_NOSLIP) continue;
    double* r = &b[NE * m->bedges[2 * e]];
    r[1] = 0.0;
    r[2] = 0.0;
    r[3] = 0.0;
    r[4] = 0.0;
    if (tag_twall(m, tag) < 0.0) r[0] = 0.0;
  }
}

/* ComputeResiduals/SpatialResidual: ucs/residual.tcc:12-122, interior kernel
 * :191-296 (b[L] -= F, b[R] += F), boundary kernel :298-387 (b[L] -= F with the
 * unextrapolated state against the phantom), residGlobal/strided norms
 * ucs/parallel.h:120-182.  norms[0] = residGlobal, norms[1..5] per equation. */
void o_residual(const o_mesh* m, const o_params* p, int iter, const double* q, const double* grad,
                const double* l, double* b, double* norms) {
  int nn = m->nnode;
  memset(b, 0, sizeof(double) * NE * (size_t)nn);
  int ho = (p->sorder > 1 && iter > p->nfirst);
  for (int e = 0; e < m->nedge; e++) {
    int L = m->edges[2 * e], R = m->edges[2 * e + 1];
    const double* qL = &q[NV * L];
    const double* qR = &q[NV * R];
    double QL[NV], QR[NV], F[5];
    memcpy(QL, qL, sizeof(QL));
    memcpy(QR, qR, sizeof(QR));
    if (ho) {
      const double* xL = &m->xyz[3 * L];
      const double* xR = &m->xyz[3 * R];
      double dx[3], dQ[5];
      dx[0] = 0.5 * (xR[0] - xL[0]);
      dx[1] = 0.5 * (xR[1] - xL[1]);
      dx[2] = 0.5 * (xR[2] - xL[2]);
      for (int j = 0; j < NE; j++) dQ[j] = qR[j] - qL[j];
      extrapolate(QL, qL, dQ, &grad[NT * 3 * L], dx, &l[NE * L], p->chi);
      for (int j = 0; j < NE; j++) dQ[j] = -dQ[j];
      dx[0] = -dx[0]; dx[1] = -dx[1]; dx[2] = -dx[2];
      extrapolate(QR, qR, dQ, &grad[NT * 3 * R], dx, &l[NE * R], p->chi);
      o_aux(QL, p->gamma);
      o_aux(QR, p->gamma);
    }
    numflux(p, QL, QR, &m->edge_a[4 * e], F);
    for (int i = 0; i < NE; i++) {
      b[NE * L + i] += -F[i];
      b[NE * R + i] += F[i];
    }
  }
  for (int e = 0; e < m->nbnode; e++) {
    int L = m->bedges[2 * e], R = m->bedges[2 * e + 1];
    double F[5];
    numflux(p, &q[NV * L], &q[NV * R], &m->bedge_a[4 * e], F);
    for (int i = 0; i < NE; i++) b[NE * L + i] += -F[i];
  }
  if (p->viscous) viscous_residual(m, p, iter, q, grad, b);
  noslip_residual(m, b);
  if (norms) {
    double s = 0.0, se[5] = {0, 0, 0, 0, 0};
    for (int i = 0; i < nn * NE; i++) s += b[i] * b[i];
    for (int i = 0; i < nn; i++)
      for (int j = 0; j < NE; j++) se[j] += b[NE * i + j] * b[NE * i + j];
    norms[0] = sqrt(s) / (double)(nn * NE);
    for (int j = 0; j < NE; j++) norms[1 + j] = sqrt(se[j]) / (double)nn;
  }
}

/* MaxEigenvalue: ucs/compressible.tcc:798-821 */
static double max_eig(const double* Q, const double* a, double gamma) {
  double rho = Q[0];
  double u = Q[1] / rho, v = Q[2] / rho, w = Q[3] / rho;
  double v2h = 0.5 * (u * u + v * v + w * w);
  double P = (gamma - 1.0) * (Q[4] - rho * v2h);
  double c = sqrt(gamma * P / rho);
  double theta = theta_of(Q, a, 0.0);
  return omax(fabs(theta + c), fabs(theta - c));
}

/* ComputeTimesteps (local time stepping): ucs/timestep.tcc:6-143 */
double o_timestep(const o_mesh* m, const o_params* p, const double* q, double cfl, double* dt) {
  int nn = m->nnode;
  for (int i = 0; i < nn; i++) dt[i] = 0.0;
  if (p->global_dt && p->fixed_dt > 0.0) { /* ucs/timestep.tcc:48-55 */
    for (int i = 0; i < nn; i++) dt[i] = p->fixed_dt;
    return p->fixed_dt;
  }
  for (int e = 0; e < m->nedge; e++) {
    int L = m->edges[2 * e], R = m->edges[2 * e + 1];
    double t[NV];
    for (int i = 0; i < NE; i++) t[i] = 0.5 * (q[NV * L + i] + q[NV * R + i]);
    o_aux(t, p->gamma);
    double me = max_eig(t, &m->edge_a[4 * e], p->gamma);
    double area = m->edge_a[4 * e + 3];
    dt[L] += me * area;
    dt[R] += me * area;
  }
  for (int e = 0; e < m->nbnode; e++) {
    int L = m->bedges[2 * e], R = m->bedges[2 * e + 1];
    double t[NV];
    for (int i = 0; i < NE; i++) t[i] = 0.5 * (q[NV * L + i] + q[NV * R + i]);
    o_aux(t, p->gamma);
    dt[L] += max_eig(t, &m->bedge_a[4 * e], p->gamma) * m->bedge_a[4 * e + 3];
  }
  if (p->global_dt) {
    /* one dt, the minimum over nodes (ucs/timestep.tcc:56-74); as in the
     * reference, the VNN limit is applied from node 1 on */
    double dtmin = cfl * (m->vol[0] / dt[0]);
    for (int i = 1; i < nn; i++) {
      dtmin = omin(dtmin, cfl * (m->vol[i] / dt[i]));
      if (p->enable_vnn) dtmin = omin(dtmin, p->vnn * pow(m->vol[i], 2.0 / 3.0));
    }
    for (int i = 0; i < nn; i++) dt[i] = dtmin;
    return dtmin;
  }
  /* local time stepping (ucs/timestep.tcc:30-47); the VNN limit is applied
   * from node 1 on, as in the reference */
  dt[0] = cfl * (m->vol[0] / dt[0]);
  double dtmin = dt[0];
  for (int i = 1; i < nn; i++) {
    dt[i] = cfl * (m->vol[i] / dt[i]);
    if (p->enable_vnn) dt[i] = omin(dt[i], p->vnn * pow(m->vol[i], 2.0 / 3.0));
    dtmin = omin(dtmin, dt[i]);
  }
  return dtmin;
}

/* Param::UpdateCFL, prescribed ramp branch: ucs/param.tcc:718-725 */
double o_cfl(const o_params* p, int iter) {
  if (p->cfl_ramp > 0 && iter < p->cfl_ramp)
    return (p->cfl - p->cfl_start) / (double)p->cfl_ramp * (double)iter + p->cfl_start;
  return p->cfl;
}

/* Param::UpdateCFL (ucs/param.tcc:695-726) with the residual trend of
 * PreTimeAdvance (ucs/solutionSpace.tcc:629-633): res = residGlobal of the
 * previous iteration, res_nm1 of the one before (0 until they exist). */
double o_cfl_update(const o_params* p, int iter, double res, double res_nm1, double cfl_prev) {
  if (!p->dynamic_cfl) return o_cfl(p, iter);
  double delta = 9999.0;
  if (res_nm1 > 1.0e-21) delta = (res_nm1 - res) / res_nm1;
  if (iter > 1) {
    double sgn = delta > 0.0 ? 1.0 : -1.0;
    double A = omax(0.0, cfl_prev + 0.1 * sgn * cfl_prev);
    return omin(A, p->cfl);
  }
  return p->cfl_start;
}

static int cmp_int(const void* a, const void* b) {
  int x = *(const int*)a, y = *(const int*)b;
  return (x > y) - (x < y);
}

/* CRSMatrix::Init: ucs/crsmatrix.tcc:47-97; row i = [diag, neighbours
 * ascending]; per-edge block positions replace GetPointer's row search
 * (ucs/crsmatrix.tcc:739-781). */
int o_crs_init(const o_mesh* m, o_crs* A) {
  int nn = m->nnode;
  memset(A, 0, sizeof(*A));
  A->nnode = nn;
  int* deg = (int*)calloc(nn + 1, sizeof(int));
  for (int e = 0; e < m->nedge; e++) {
    deg[m->edges[2 * e]]++;
    deg[m->edges[2 * e + 1]]++;
  }
  A->ia = (int*)malloc(sizeof(int) * (nn + 1));
  A->iau = (int*)malloc(sizeof(int) * nn);
  A->ia[0] = 0;
  for (int i = 0; i < nn; i++) A->ia[i + 1] = A->ia[i] + deg[i] + 1;
  A->nblocks = A->ia[nn];
  A->ja = (int*)malloc(sizeof(int) * A->nblocks);
  int* fill = (int*)malloc(sizeof(int) * nn);
  for (int i = 0; i < nn; i++) {
    A->iau[i] = A->ia[i];
    A->ja[A->ia[i]] = i;
    fill[i] = A->ia[i] + 1;
  }
  for (int e = 0; e < m->nedge; e++) {
    int a = m->edges[2 * e], b = m->edges[2 * e + 1];
    A->ja[fill[a]++] = b;
    A->ja[fill[b]++] = a;
  }
  for (int i = 0; i < nn; i++) qsort(&A->ja[A->ia[i] + 1], A->ia[i + 1] - A->ia[i] - 1, sizeof(int), cmp_int);
  A->e_lr = (int*)malloc(sizeof(int) * (m->nedge > 0 ? m->nedge : 1));
  A->e_rl = (int*)malloc(sizeof(int) * (m->nedge > 0 ? m->nedge : 1));
  for (int e = 0; e < m->nedge; e++) {
    int a = m->edges[2 * e], b = m->edges[2 * e + 1];
    int k;
    for (k = A->ia[a] + 1; k < A->ia[a + 1]; k++)
      if (A->ja[k] == b) break;
    A->e_lr[e] = k;
    for (k = A->ia[b] + 1; k < A->ia[b + 1]; k++)
      if (A->ja[k] == a) break;
    A->e_rl[e] = k;
  }
  A->pv = (int*)malloc(sizeof(int) * NE * (size_t)nn);
  A->M = (double*)calloc((size_t)A->nblocks * 25, sizeof(double));
  free(deg);
  free(fill);
  return 0;
}

void o_crs_free(o_crs* A) {
  free(A->ia); free(A->ja); free(A->iau); free(A->pv); free(A->M); free(A->e_lr); free(A->e_rl);
  memset(A, 0, sizeof(*A));
}

/* CRSMatrix::BlankSubRow: ucs/crsmatrix.tcc:525-541 */
static void blank_subrow(o_crs* A, int row, int sub) {
  for (int k = A->ia[row]; k < A->ia[row + 1]; k++)
    for (int i = 0; i < NE; i++) A->M[25 * (size_t)k + sub * NE + i] = 0.0;
  A->M[25 * (size_t)A->iau[row] + sub * NE + sub] = 1.0;
}

/* ComputeJacobians: ComputeSpatialJacobian (ucs/jacobian.tcc:129-210) with
 * Kernel_NumJac (:253-304, one-sided h=1e-8 on the unextrapolated states),
 * Bkernel_NumJac (:458-535, boundary states re-evaluated under perturbation,
 * jacobianBoundaryEval=0), Kernel_Diag_NumJac (:436-456, diagonal = -sum of
 * the column's off-diagonal blocks), then ContributeTemporalTerms
 * (:213-250, A_ii += vol/dt_i).  Bkernel_NumJac writes the recomputed
 * phantom (and Dirichlet interior) states back into q, as the reference does. */
void o_jacobian(const o_mesh* m, const o_params* p, double* q, const double* dt, o_crs* A) {
  const double h = 1.0e-8;
  double g = p->gamma;
  memset(A->M, 0, sizeof(double) * 25 * (size_t)A->nblocks);
  for (int e = 0; e < m->nedge; e++) {
    int L = m->edges[2 * e], R = m->edges[2 * e + 1];
    const double* QL = &q[NV * L];
    const double* QR = &q[NV * R];
    const double* a = &m->edge_a[4 * e];
    double fS[5], fL[5], fR[5], QPL[NV], QPR[NV];
    double* bRL = &A->M[25 * (size_t)A->e_rl[e]];
    double* bLR = &A->M[25 * (size_t)A->e_lr[e]];
    if (p->jac_field_type == 1) {
      /* Kernel_NumJac_Centered (ucs/jacobian.tcc:307-367) */
      double fLd[5], fRd[5];
      for (int i = 0; i < NE; i++) {
        memcpy(QPL, QL, sizeof(QPL));
        memcpy(QPR, QR, sizeof(QPR));
        QPL[i] += h;
        QPR[i] += h;
        o_aux(QPL, g);
        o_aux(QPR, g);
        numflux(p, QPL, QR, a, fL);
        numflux(p, QL, QPR, a, fR);
        memcpy(QPL, QL, sizeof(QPL));
        memcpy(QPR, QR, sizeof(QPR));
        QPL[i] -= h;
        QPR[i] -= h;
        o_aux(QPL, g);
        o_aux(QPR, g);
        numflux(p, QPL, QR, a, fLd);
        numflux(p, QL, QPR, a, fRd);
        for (int j = 0; j < NE; j++) bRL[j * NE + i] += (fLd[j] - fL[j]) / (2.0 * h);
        for (int j = 0; j < NE; j++) bLR[j * NE + i] += (fR[j] - fRd[j]) / (2.0 * h);
      }
      continue;
    }
    numflux(p, QL, QR, a, fS);
    for (int i = 0; i < NE; i++) {
      memcpy(QPL, QL, sizeof(QPL));
      memcpy(QPR, QR, sizeof(QPR));
      QPL[i] += h;
      QPR[i] += h;
      o_aux(QPL, g);
      o_aux(QPR, g);
      numflux(p, QPL, QR, a, fL);
      numflux(p, QL, QPR, a, fR);
      for (int j = 0; j < NE; j++) bRL[j * NE + i] += (fS[j] - fL[j]) / h;
      for (int j = 0; j < NE; j++) bLR[j * NE + i] += (fR[j] - fS[j]) / h;
    }
  }
  for (int e = 0; e < m->nbnode; e++) {
    int L = m->bedges[2 * e], R = m->bedges[2 * e + 1];
    int tag = m->bfactag[e];
    int bct = m->bc_type[tag];
    const double* Qref = &m->qref[NV * tag];
    const double* a = &m->bedge_a[4 * e];
    double* QL = &q[NV * L];
    double* QR = &q[NV * R];
    double fS[5], fL[5], QPL[NV], QPR[NV];
    double tw = tag_twall(m, tag);
    const double* ns = normal_state(m, q, e);
    o_bc_state_ex(bct, QL, QR, Qref, a, g, tw, ns);
    double* D = &A->M[25 * (size_t)A->iau[L]];
    if (p->jac_boundary_type == 1) {
      /* Bkernel_NumJac_Centered (ucs/jacobian.tcc:538-640) with jacobianBoundaryEval = 0
       * (ucs/param.tcc:174): the +h side re-evaluates the boundary state from the
       * perturbed node state (:577-584), the -h side uses the unperturbed boundary
       * state (:606-608), so even a hard-set (Dirichlet) node gives half a one-sided
       * column */
      for (int i = 0; i < NE; i++) {
        memcpy(QPL, QL, sizeof(QPL));
        QPL[i] += h;
        o_aux(QPL, g);
        memcpy(QPR, QR, sizeof(QPR));
        o_aux(QPR, g);
        o_bc_state_ex(bct, QPL, QPR, Qref, a, g, tw, ns);
        numflux(p, QPL, QPR, a, fL);
        memcpy(QPL, QL, sizeof(QPL));
        QPL[i] -= h;
        o_aux(QPL, g);
        numflux(p, QPL, QR, a, fS);
        for (int j = 0; j < NE; j++) D[j * NE + i] += (fL[j] - fS[j]) / (2.0 * h);
      }
      continue;
    }
    numflux(p, QL, QR, a, fS);
    for (int i = 0; i < NE; i++) {
      memcpy(QPL, QL, sizeof(QPL));
      QPL[i] += h;
      o_aux(QPL, g);
      memcpy(QPR, QR, sizeof(QPR));
      o_aux(QPR, g);
      o_bc_state_ex(bct, QPL, QPR, Qref, a, g, tw, ns);
      numflux(p, QPL, QPR, a, fL);
      for (int j = 0; j < NE; j++) D[j * NE + i] += (fL[j] - fS[j]) / h;
    }
  }
  if (p->viscous) {
    /* Driver(Kernel_Viscous_Jac): ucs/jacobian.tcc:734-767; the boundary
     * variant (:771-816) ends with size 0 and contributes nothing */
    for (int e = 0; e < m->nedge; e++) {
      int L = m->edges[2 * e], R = m->edges[2 * e + 1];
      double dx[3], s2 = 0.0, aL[25], aR[25];
      for (int i = 0; i < 3; i++) {
        dx[i] = (m->xyz[3 * R + i] - m->xyz[3 * L + i]);
        s2 += dx[i] * dx[i];
      }
      o_viscous_jacobian(p, &q[NV * L], &q[NV * R], dx, s2, &m->edge_a[4 * e], aL, aR);
      double* bRL = &A->M[25 * (size_t)A->e_rl[e]];
      double* bLR = &A->M[25 * (size_t)A->e_lr[e]];
      for (int k = 0; k < 25; k++) {
        bRL[k] += aL[k];
        bLR[k] += aR[k];
      }
    }
  }
  for (int e = 0; e < m->nedge; e++) {
    int L = m->edges[2 * e], R = m->edges[2 * e + 1];
    double* DL = &A->M[25 * (size_t)A->iau[L]];
    double* DR = &A->M[25 * (size_t)A->iau[R]];
    const double* jL = &A->M[25 * (size_t)A->e_rl[e]];
    const double* jR = &A->M[25 * (size_t)A->e_lr[e]];
    for (int k = 0; k < 25; k++) {
      DL[k] += -jL[k];
      DR[k] += -jR[k];
    }
  }
  for (int i = 0; i < m->nnode; i++) {
    double* D = &A->M[25 * (size_t)A->iau[i]];
    for (int k = 0; k < NE; k++) D[k * NE + k] += 1.0 * m->vol[i] / dt[i];
  }
  /* Bkernel_BC_Jac_Modify -> ModifyViscousWallJacobian (ucs/bc.tcc:747-900,
   * ucs/compressible.tcc:1577-1606), per noSlip half-edge in order */
  for (int e = 0; e < m->nbnode; e++) {
    int tag = m->bfactag[e];
    if (m->bc_type[tag] != O_BC_NOSLIP) continue;
    int row = m->bedges[2 * e];
    double tw = tag_twall(m, tag);
    blank_subrow(A, row, 1);
    blank_subrow(A, row, 2);
    blank_subrow(A, row, 3);
    double* diag = &A->M[25 * (size_t)A->iau[row]];
    if (tw < 0.0) {
      blank_subrow(A, row, 0);
      blank_subrow(A, row, 4);
      int nn = m->bnormal ? m->bnormal[e] : -1;
      for (int k = A->ia[row]; nn >= 0 && k < A->ia[row + 1]; k++)
        if (A->ja[k] == nn) {
          A->M[25 * (size_t)k + 0] = -1.0;
          A->M[25 * (size_t)k + 24] = -1.0;
        }
    } else {
      double v2 = 0.0;
      blank_subrow(A, row, 4);
      diag[4 * NE + 0] = -(tw / (g * (g - 1.0)) + 0.5 * v2);
    }
  }
}

/* LU with partial pivoting, in place, rows addressed through p:
 * ucs/matrix.h:110-190 */
int o_lu(double* a, int* p, int n) {
  int condBad = 0, row = 0;
  const double smallnum = 1.0e-15;
  for (int i = 0; i < n; i++) p[i] = i;
  for (int i = 0; i < n; i++) {
    double large = 0.0;
    for (int j = i; j < n; j++) {
      if (fabs(a[p[j] * n + i]) > fabs(large)) {
        large = a[p[j] * n + i];
        row = j;
      }
    }
    if (large == 0.0) condBad++;
    else if (fabs(large) < smallnum) condBad++;
    int t = p[i];
    p[i] = p[row];
    p[row] = t;
    large = 1.0 / large;
    for (int j = i + 1; j < n; j++) a[p[j] * n + i] *= large;
    for (int j = i + 1; j < n; j++)
      for (int k = i + 1; k < n; k++) a[p[j] * n + k] -= a[p[j] * n + i] * a[p[i] * n + k];
  }
  return condBad > 0;
}

/* LuSolve: ucs/matrix.h:237-264 (b is overwritten with the solution) */
void o_lu_solve(const double* a, double* b, const int* p, double* x, int n) {
  for (int i = 0; i < n; i++) {
    double sum = 0.0;
    for (int j = 0; j < i; j++) sum += a[p[i] * n + j] * x[j];
    x[i] = b[p[i]] - sum;
  }
  for (int i = n - 1; i >= 0; i--) {
    double sum = 0.0;
    for (int j = n - 1; j > i; j--) sum += a[p[i] * n + j] * b[j];
    b[i] = (x[i] - sum) / a[p[i] * n + i];
  }
}

/* CRSMatrix::PrepareSGS: ucs/crsmatrix.tcc:839-876 */
void o_prepare_sgs(o_crs* A) {
  for (int i = 0; i < A->nnode; i++) o_lu(&A->M[25 * (size_t)A->iau[i]], &A->pv[NE * i], NE);
}

static void sgs_row(const o_crs* A, int node, const double* b, double* x) {
  double rhs[5], vout[5], tmp[5];
  memcpy(rhs, &b[NE * node], sizeof(rhs));
  for (int k = A->ia[node] + 1; k < A->ia[node + 1]; k++) {
    matvec5(&A->M[25 * (size_t)k], &x[NE * A->ja[k]], vout);
    for (int j = 0; j < NE; j++) rhs[j] -= vout[j];
  }
  o_lu_solve(&A->M[25 * (size_t)A->iau[node]], rhs, &A->pv[NE * node], tmp, NE);
  memcpy(&x[NE * node], rhs, sizeof(rhs));
}

/* CRS::SGS: ucs/crs.tcc:61-173.  order: forward sweep row order (NULL =
 * natural order); the backward sweep visits the reverse order.  Returns
 * |xNorm - xNormOld| with xNorm = sqrt(sum x^2)/(nnode*neqn). */
double o_sgs(const o_crs* A, int nsgs, const double* b, double* x, const int* order) {
  int nn = A->nnode;
  double xOld = 0.0, xNorm = 0.0;
  for (int it = 0; it < nsgs; it++) {
    for (int i = 0; i < nn; i++) sgs_row(A, order ? order[i] : i, b, x);
    for (int i = nn - 1; i >= 0; i--) sgs_row(A, order ? order[i] : i, b, x);
    xOld = xNorm;
    double s = 0.0;
    for (int i = 0; i < nn * NE; i++) s += x[i] * x[i];
    xNorm = sqrt(s) / (double)(nn * NE);
  }
  return fabs(xOld - xNorm);
}

/* CompressibleEqnSet::ApplyDQ: ucs/compressible.tcc:928-993 */
void o_apply_dq(const double* dQ, double* Q, double gamma) {
  double gm1 = gamma - 1.0;
  const double minP = 1.0e-10, minRho = 1.0e-10, minE = 1.0e-10;
  if (Q[0] + dQ[0] < 0.0) Q[0] = minRho;
  else Q[0] += dQ[0];
  if (Q[4] + dQ[4] < 0.0) Q[4] = minE;
  else Q[4] += dQ[4];
  double rho = Q[0];
  double u = Q[1] / rho, v = Q[2] / rho, w = Q[3] / rho, E = Q[4];
  double v2 = u * u + v * v + w * w;
  if (E < 0.5 * rho * v2) {
    double v2mod = 2.0 * (E - minP / gm1);
    double frac = 0.0;
    if (v2mod > 0.0) frac = sqrt(v2mod / v2);
    u *= frac; v *= frac; w *= frac;
    Q[1] = rho * u; Q[2] = rho * v; Q[3] = rho * w;
  } else {
    Q[1] += dQ[1]; Q[2] += dQ[2]; Q[3] += dQ[3];
  }
  o_aux(Q, gamma);
}

int o_solver_init(o_solver* s, const o_mesh* m, const o_params* p, const double* q0) {
  memset(s, 0, sizeof(*s));
  s->m = m;
  s->p = *p;
  size_t nt = (size_t)m->nnode + m->nbnode;
  s->q = (double*)malloc(sizeof(double) * NV * nt);
  memcpy(s->q, q0, sizeof(double) * NV * nt);
  for (size_t i = 0; i < nt; i++) o_aux(&s->q[NV * i], p->gamma);
  s->grad = (double*)calloc((size_t)m->nnode * NT * 3, sizeof(double));
  s->l = (double*)malloc(sizeof(double) * NE * (size_t)m->nnode);
  for (size_t i = 0; i < (size_t)NE * m->nnode; i++) s->l[i] = 1.0; /* Limiter ctor: ucs/limiters.tcc:39-43 */
  s->sw = (double*)malloc(sizeof(double) * 6 * (size_t)m->nnode);
  s->b = (double*)calloc((size_t)NE * m->nnode, sizeof(double));
  s->x = (double*)calloc((size_t)NE * m->nnode, sizeof(double));
  s->dt = (double*)calloc((size_t)m->nnode, sizeof(double));
  o_lsq_moments(m, s->sw);
  if (p->nsgs > 0) o_crs_init(m, &s->A);
  /* the iteration counter starts at 0 (SolutionSpaceBase, ucs/solutionSpaceBase.h:45):
   * the first Newton iteration runs at iter 0 -- CFL = cflStart and, since
   * iter > firstOrderSteps fails, first order without gradient or limiter */
  s->iter = 0;
  /* PreIterate: UpdateCFL + ComputeTimesteps, ucs/solutionSpace.tcc:509-569 */
  s->cfl = o_cfl_update(p, s->iter, 0.0, 0.0, p->cfl_start); /* PreIterate (ucs/solutionSpace.tcc:546-551) */
  s->res = s->res_nm1 = 0.0;
  o_timestep(m, &s->p, s->q, s->cfl, s->dt);
  return 0;
}

/* PreTimeAdvance (ucs/solutionSpace.tcc:572-636) + NewtonIterate
 * (:639-904) + PostTimeAdvance iter++ (:907-916). */
double o_solver_step(o_solver* s, double* norms, int* nzeroed) {
  const o_mesh* m = s->m;
  const o_params* p = &s->p;
  int nn = m->nnode;
  double t0 = now_s();
  if (p->nsgs > 0 && (s->iter % (p->jac_freq > 0 ? p->jac_freq : 1) == 0 || s->iter == 1))
    o_jacobian(m, p, s->q, s->dt, &s->A);
  double t1 = now_s();
  s->t_jac += t1 - t0;
  s->cfl = o_cfl_update(p, s->iter, s->res, s->res_nm1, s->cfl);
  o_timestep(m, p, s->q, s->cfl, s->dt);
  o_update_bcs(m, p, s->q);
  double t2 = now_s();
  if ((p->sorder > 1 && s->iter > p->nfirst) || p->viscous) {
    if (p->gradient_type == 1) o_gradient_gg(m, s->q, s->grad);
    else o_gradient(m, s->q, s->sw, s->grad);
  }
  int lr = p->limiter_refresh > 0 ? p->limiter_refresh : 1;
  if ((s->iter % lr == 0) && p->sorder > 1 && p->limiter && s->iter > p->nfirst && s->iter < p->limiter_freeze)
    o_limiter(m, p, s->q, s->grad, s->l);
  double t3 = now_s();
  s->t_grad += t3 - t2;
  double nrm[6];
  o_residual(m, p, s->iter, s->q, s->grad, s->l, s->b, nrm);
  s->res_nm1 = s->res; /* ucs/solutionSpace.tcc:705-706 */
  s->res = nrm[0];
  double t4 = now_s();
  s->t_res += t4 - t3;
  if (p->nsgs == 0 && p->rk_stages == 4) {
    /* Four-stage Runge-Kutta with Jameson's coefficients (an extension: the
     * reference's explicit path is the single stage below): stage k sets
     * Q = Q0 + a_k dt/vol b(Q_{k-1}) through ApplyDQ, rows with a NaN/Inf
     * correction keep Q0 (counted); boundary states, and at second order the
     * gradient and limiter, are refreshed before each later stage's
     * residual; dt is the step's.  The residual norms are the first stage's. */
    static const double a[4] = {0.25, 1.0 / 3.0, 0.5, 1.0};
    const size_t nq = (size_t)NV * (size_t)(nn + m->nbnode);
    double* q0 = (double*)malloc(sizeof(double) * nq);
    memcpy(q0, s->q, sizeof(double) * nq);
    int zc = 0;
    for (int k = 0; k < 4; k++) {
      if (k > 0) {
        o_update_bcs(m, p, s->q);
        if ((p->sorder > 1 && s->iter > p->nfirst) || p->viscous) {
          if (p->gradient_type == 1) o_gradient_gg(m, s->q, s->grad);
          else o_gradient(m, s->q, s->sw, s->grad);
        }
        if ((s->iter % lr == 0) && p->sorder > 1 && p->limiter && s->iter > p->nfirst &&
            s->iter < p->limiter_freeze)
          o_limiter(m, p, s->q, s->grad, s->l);
        double tmp[6];
        o_residual(m, p, s->iter, s->q, s->grad, s->l, s->b, tmp);
      }
      for (int i = 0; i < nn; i++) {
        double x[NE];
        int z = 0;
        for (int j = 0; j < NE; j++) {
          x[j] = a[k] * (s->b[NE * i + j] * s->dt[i] / m->vol[i]);
          if (isnan(x[j]) || isinf(x[j])) z = 1;
        }
        if (z) {
          zc++;
          for (int j = 0; j < NE; j++) x[j] = 0.0;
        }
        memcpy(&s->q[NV * i], &q0[NV * i], sizeof(double) * NV);
        o_apply_dq(x, &s->q[NV * i], p->gamma);
      }
    }
    free(q0);
    s->t_lin += now_s() - t4;
    if (nzeroed) *nzeroed = zc;
    if (norms) memcpy(norms, nrm, sizeof(nrm));
    s->iter++;
    return nrm[0];
  }
  if (p->nsgs > 0) {
    o_prepare_sgs(&s->A);
    memset(s->x, 0, sizeof(double) * NE * (size_t)nn);
    o_sgs(&s->A, p->nsgs, s->b, s->x, s->order);
  } else {
    /* ExplicitSolve: ucs/solve.tcc:70-101 -- x = b dt/vol and ApplyDQ; the
     * update below applies x a second time, as NewtonIterate does. */
    for (int i = 0; i < nn; i++) {
      for (int j = 0; j < NE; j++) s->x[NE * i + j] = s->b[NE * i + j] * s->dt[i] / m->vol[i];
      o_apply_dq(&s->x[NE * i], &s->q[NV * i], p->gamma);
    }
  }
  double t5 = now_s();
  s->t_lin += t5 - t4;
  int zc = 0;
  for (int i = 0; i < nn; i++) {
    int z = 0;
    for (int k = 0; k < NE; k++) {
      double v = s->x[NE * i + k];
      if (isnan(v) || isinf(v)) {
        s->x[NE * i + k] = 0.0;
        z = 1;
      }
    }
    if (z) {
      zc++;
      for (int k = 0; k < NE; k++) s->x[NE * i + k] = 0.0;
    }
  }
  for (int i = 0; i < nn; i++) o_apply_dq(&s->x[NE * i], &s->q[NV * i], p->gamma);
  s->t_upd += now_s() - t5;
  if (nzeroed) *nzeroed = zc;
  if (norms) memcpy(norms, nrm, sizeof(nrm));
  s->iter++;
  return nrm[0];
}

void o_solver_free(o_solver* s) {
  free(s->q); free(s->grad); free(s->l); free(s->sw); free(s->b); free(s->x); free(s->dt);
  if (s->A.M) o_crs_free(&s->A);
  memset(s, 0, sizeof(*s));
}
