// Device physics for the compressible (perfect-gas) equation set.
// Each function follows the reference formula named in its comment; the
// operation order is kept so results agree with the reference to rounding.
#pragma once
#include <hip/hip_runtime.h>

namespace pcfd {

// host+device so the same formulas can be exercised in CPU builds
#define PCFD_DEV __host__ __device__ __forceinline__

// ComputeAuxiliaryVariables pressure (ucs/compressible.tcc:1229-1236):
// P = (g-1)(E - 0.5*rho*|u|^2) -- the value GetPressure() later returns.
PCFD_DEV double aux_pressure(const double* Q, double gm1) {
  double u = Q[1] / Q[0], v = Q[2] / Q[0], w = Q[3] / Q[0];
  double V2 = u * u + v * v + w * w;
  return gm1 * (Q[4] - 0.5 * Q[0] * V2);
}

// ComputePressure (ucs/compressible.tcc:1090-1101): P = (g-1)(E - rho*v2h)
PCFD_DEV double comp_pressure(const double* Q, double gm1) {
  double r = Q[0];
  double u = Q[1] / r, v = Q[2] / r, w = Q[3] / r;
  double v2h = 0.5 * (u * u + v * v + w * w);
  return gm1 * (Q[4] - r * v2h);
}

PCFD_DEV double theta_of(const double* Q, const double* a) {  // GetTheta, static mesh
  double u = Q[1] / Q[0], v = Q[2] / Q[0], w = Q[3] / Q[0];
  return u * a[0] + v * a[1] + w * a[2] + 0.0;
}

// Flux: ucs/compressible.tcc:686-710 (vdotn = 0)
PCFD_DEV void phys_flux(const double* Q, const double* a, double gm1, double* f) {
  double rho = Q[0];
  double u = Q[1] / rho, v = Q[2] / rho, w = Q[3] / rho;
  double v2h = 0.5 * (u * u + v * v + w * w);
  double P = gm1 * (Q[4] - rho * v2h);
  double ht = (Q[4] + P) / rho;
  double rt = rho * (a[0] * u + a[1] * v + a[2] * w + 0.0);
  f[0] = rt;
  f[1] = (u * rt + P * a[0]);
  f[2] = (v * rt + P * a[1]);
  f[3] = (w * rt + P * a[2]);
  f[4] = (ht * rt - 0.0 * P);
}

PCFD_DEV double dmax(double x, double y) { return x > y ? x : y; }
PCFD_DEV double dmin(double x, double y) { return x < y ? x : y; }

// ---------------------------------------------------------------------------
// Roe flux in wave-strength form.  Same Roe average, eigenvalues,
// Harten-Hyman type-2 entropy fix (ucs/compressible.tcc:92-230) and NaN
// zeroing (ucs/eqnset.tcc:74-88) as the reference's RoeFlux, whose dissipation
// T*|L|*Tinv*dQ is two explicit 5x5 mat-vecs; here it is assembled
// from the characteristic wave strengths of the primitive jumps (the Roe
// average makes the two forms identical; results differ by rounding only)
// and every division by rho, c, c^2 is a multiplication by a reciprocal.
// Per-state quantities live in RoeSide so a finite-difference Jacobian can
// reuse the unperturbed side.
struct RoeSide {
  double ir, u, v, w, p, h, th, c;
};

PCFD_DEV void roe_side(const double* Q, const double* n, double gamma, RoeSide& s) {
  const double gm1 = gamma - 1.0;
  s.ir = 1.0 / Q[0];
  s.u = Q[1] * s.ir;
  s.v = Q[2] * s.ir;
  s.w = Q[3] * s.ir;
  const double v2 = s.u * s.u + s.v * s.v + s.w * s.w;
  s.p = gm1 * (Q[4] - 0.5 * Q[0] * v2);
  s.h = (Q[4] + s.p) * s.ir;
  s.th = s.u * n[0] + s.v * n[1] + s.w * n[2];
  s.c = sqrt(gamma * s.p * s.ir);
}

// The normal-independent part of roe_side, and its normal velocity: for a
// state that meets many faces (a node's own state in an edge loop).
PCFD_DEV void roe_side_state(const double* Q, double gamma, RoeSide& s) {
  const double gm1 = gamma - 1.0;
  s.ir = 1.0 / Q[0];
  s.u = Q[1] * s.ir;
  s.v = Q[2] * s.ir;
  s.w = Q[3] * s.ir;
  const double v2 = s.u * s.u + s.v * s.v + s.w * s.w;
  s.p = gm1 * (Q[4] - 0.5 * Q[0] * v2);
  s.h = (Q[4] + s.p) * s.ir;
  s.th = 0.0;
  s.c = sqrt(gamma * s.p * s.ir);
}

PCFD_DEV void roe_side_th(RoeSide& s, const double* n) { s.th = s.u * n[0] + s.v * n[1] + s.w * n[2]; }

PCFD_DEV double hh_fix(double e, double eL, double eR) {
  const double eps = dmax(0.0, dmax(e - eL, eR - e));
  return (fabs(e) < eps) ? 0.5 * (e * e / eps + eps) : fabs(e);
}

// Returns true for an unphysical state: a NaN component was zeroed, or a
// side's sound speed is NaN (negative pressure), where the entropy fix's
// max() keeps or drops the NaN by argument position.
PCFD_DEV bool roe_flux_s(const double* QL, const RoeSide& L, const double* QR, const RoeSide& R, const double* a,
                         double gamma, int sym2d, double* F) {
  const double gm1 = gamma - 1.0;
  const double nx = a[0], ny = a[1], nz = a[2];
  const double rr = sqrt(QL[0] * QR[0]);
  const double sigma = rr / (QL[0] + rr);
  const double u = L.u + sigma * (R.u - L.u);
  const double v = L.v + sigma * (R.v - L.v);
  const double w = L.w + sigma * (R.w - L.w);
  const double h = L.h + sigma * (R.h - L.h);
  const double v2h = 0.5 * (u * u + v * v + w * w);
  const double c2 = gm1 * (h - v2h);
  const double c = sqrt(c2);
  const double ic2 = 1.0 / c2;
  const double th = u * nx + v * ny + w * nz;
  const double l0 = hh_fix(th, L.th, R.th);
  const double l3 = hh_fix(th + c, L.th + L.c, R.th + R.c);
  const double l4 = hh_fix(th - c, L.th - L.c, R.th - R.c);
  const double dr = QR[0] - QL[0];
  const double du = R.u - L.u, dv = R.v - L.v, dw = R.w - L.w;
  const double dp = R.p - L.p;
  const double dV = du * nx + dv * ny + dw * nz;
  const double rc = rr * c;
  const double t4 = l4 * (0.5 * (dp - rc * dV) * ic2);  // theta - c wave
  const double t3 = l3 * (0.5 * (dp + rc * dV) * ic2);  // theta + c wave
  const double t0 = l0 * (dr - dp * ic2);               // entropy wave
  const double ts = l0 * rr;                            // shear waves
  double d[5];
  d[0] = t4 + t3 + t0;
  d[1] = t4 * (u - c * nx) + t3 * (u + c * nx) + t0 * u + ts * (du - dV * nx);
  d[2] = t4 * (v - c * ny) + t3 * (v + c * ny) + t0 * v + ts * (dv - dV * ny);
  d[3] = t4 * (w - c * nz) + t3 * (w + c * nz) + t0 * w + ts * (dw - dV * nz);
  d[4] = t4 * (h - c * th) + t3 * (h + c * th) + t0 * v2h + ts * (u * du + v * dv + w * dw - th * dV);
  const double rtL = QL[0] * L.th, rtR = QR[0] * R.th;
  const double hA = 0.5 * a[3];
  F[0] = hA * (rtL + rtR - d[0]);
  F[1] = hA * ((L.u * rtL + L.p * nx) + (R.u * rtR + R.p * nx) - d[1]);
  F[2] = hA * ((L.v * rtL + L.p * ny) + (R.v * rtR + R.p * ny) - d[2]);
  F[3] = hA * ((L.w * rtL + L.p * nz) + (R.w * rtR + R.p * nz) - d[3]);
  F[4] = hA * (L.h * rtL + R.h * rtR - d[4]);
  if (sym2d == 1) F[1] = 0.0;
  else if (sym2d == 2) F[2] = 0.0;
  else if (sym2d == 3) F[3] = 0.0;
  bool nan = __builtin_isnan(L.c) || __builtin_isnan(R.c);
#pragma unroll
  for (int i = 0; i < 5; ++i)
    if (__builtin_isnan(F[i])) {
      F[i] = 0.0;
      nan = true;
    }
  return nan;
}

PCFD_DEV bool roe_flux_fast(const double* QL, const double* QR, const double* a, double gamma, int sym2d,
                            double* F) {
  RoeSide L, R;
  roe_side(QL, a, gamma, L);
  roe_side(QR, a, gamma, R);
  return roe_flux_s(QL, L, QR, R, a, gamma, sym2d, F);
}

// F(L, R, a) = -F(R, L, -a) holds for the Roe flux of physical states, but
// not for an unphysical one: the entropy fix's max() keeps or drops a NaN
// sound speed by argument position, and NaN components are zeroed
// afterwards.  A flux evaluated from the right node's side on such a state is
// therefore re-evaluated in the reference's (left, right) order:
// G = -F(q_left, q_right, a_canonical).
PCFD_DEV void roe_flux_canonical(const double* q_left, const double* q_right, const double* a_canon, double gamma,
                                 int sym2d, double* G) {
  double F[5];
  roe_flux_fast(q_left, q_right, a_canon, gamma, sym2d, F);
#pragma unroll
  for (int i = 0; i < 5; ++i) G[i] = -F[i];
}

// EqnSet::NumericalFlux with fluxType algebraicFlux (ucs/eqnset.tcc:56-63,
// 74-88): the average of the two physical fluxes times |a|, NaN components
// zeroed (no symmetry2d zeroing on this branch).
PCFD_DEV void alg_flux(const double* QL, const double* QR, const double* a, double gamma, double* F) {
  double f1[5], f2[5];
  phys_flux(QL, a, gamma - 1.0, f1);
  phys_flux(QR, a, gamma - 1.0, f2);
#pragma unroll
  for (int i = 0; i < 5; ++i) {
    F[i] = 0.5 * a[3] * (f1[i] + f2[i]);
    if (__builtin_isnan(F[i])) F[i] = 0.0;
  }
}

// EqnSet::NumericalFlux: the Roe flux (roeFlux, the reference default) or the
// algebraic flux
template <bool ALG>
PCFD_DEV void num_flux(const double* QL, const double* QR, const double* a, double gamma, int sym2d, double* F) {
  if constexpr (ALG) alg_flux(QL, QR, a, gamma, F);
  else roe_flux_fast(QL, QR, a, gamma, sym2d, F);
}

// MaxEigenvalue on the edge-averaged state (ucs/compressible.tcc:798-821)
PCFD_DEV double max_eig(const double* Q, const double* a, double gamma) {
  const double gm1 = gamma - 1.0;
  double rho = Q[0];
  double u = Q[1] / rho, v = Q[2] / rho, w = Q[3] / rho;
  double v2h = 0.5 * (u * u + v * v + w * w);
  double P = gm1 * (Q[4] - rho * v2h);
  double c = sqrt(gamma * P / rho);
  double th = theta_of(Q, a);
  double e4 = th + c, e5 = th - c;
  return dmax(fabs(e4), fabs(e5));
}

// BC type ids (ucs/bc_defines.h:4-30)
enum : int {
  BC_PARALLEL = 0, BC_DIRICHLET = 1, BC_NEUMANN = 2, BC_IMPERMEABLE = 3, BC_NOSLIP = 4,
  BC_FARFIELD_VISCOUS = 5, BC_FARFIELD = 6, BC_SONIC_INFLOW = 7, BC_SONIC_OUTFLOW = 8, BC_SYMMETRY = 9
};

// GetFarfieldBoundaryVariables (ucs/compressible.tcc:1245-1372): 10 fixed-point
// sub-iterations on the phantom QR using the updated average state.
// pL / pinf are the auxiliary pressures of QL / Qinf.
PCFD_DEV void farfield_state(const double* QL, double pL, double* QR, const double* Qinf, double pinf,
                             const double* a, double gamma) {
  const double gm1 = gamma - 1.0;
  const double nx = a[0], ny = a[1], nz = a[2];
  for (int subit = 0; subit < 10; ++subit) {
    double t[5];
#pragma unroll
    for (int i = 0; i < 5; ++i) t[i] = (QL[i] + QR[i]) / 2.0;
    double theta = theta_of(t, a);
    double rhoi = QL[0];
    double ui = QL[1] / QL[0] + 0.0, vi = QL[2] / QL[0] + 0.0, wi = QL[3] / QL[0] + 0.0;
    double rhoinf = Qinf[0];
    double uinf = Qinf[1] / Qinf[0] + 0.0, vinf = Qinf[2] / Qinf[0] + 0.0, winf = Qinf[3] / Qinf[0] + 0.0;
    double pavg = comp_pressure(t, gm1);
    double rhoavg = t[0];
    double c2avg = gamma * (pavg / rhoavg);
    double cavg = sqrt(c2avg);
    if (theta == 0.0) return;
    double mn = fabs(theta / cavg);
    if (theta > 0.0 && mn >= 1.0) {
#pragma unroll
      for (int i = 0; i < 5; ++i) QR[i] = QL[i];
    } else if (theta < 0.0 && mn >= 1.0) {
#pragma unroll
      for (int i = 0; i < 5; ++i) QR[i] = Qinf[i];
    } else if (theta > 0.0 && mn < 1.0) {
      double pb = pinf;
      double rhob = rhoi + (pb - pL) / c2avg;
      double tt = (pb - pL) / (rhoavg * cavg);
      double ub = ui - nx * tt, vb = vi - ny * tt, wb = wi - nz * tt;
      QR[0] = rhob; QR[1] = rhob * ub; QR[2] = rhob * vb; QR[3] = rhob * wb;
      QR[4] = pb / (gamma - 1.0) + 0.5 * rhob * (ub * ub + vb * vb + wb * wb);
    } else if (theta < 0.0 && mn < 1.0) {
      double pb = 0.5 * (pinf + pL + rhoavg * cavg * (nx * (uinf - ui) + ny * (vinf - vi) + nz * (winf - wi)));
      double rhob = rhoinf + (pb - pinf) / c2avg;
      double tt = (pb - pinf) / (rhoavg * cavg);
      double ub = uinf + nx * tt, vb = vinf + ny * tt, wb = winf + nz * tt;
      QR[0] = rhob; QR[1] = rhob * ub; QR[2] = rhob * vb; QR[3] = rhob * wb;
      QR[4] = pb / (gamma - 1.0) + 0.5 * rhob * (ub * ub + vb * vb + wb * wb);
    } else {
      return;
    }
  }
}

// GetInviscidWallBoundaryVariables, CVBC branch (ucs/compressible.tcc:1374-1472)
PCFD_DEV void wall_state(const double* QL, double pL, double* QR, const double* a, double gamma) {
  const double gm1 = gamma - 1.0;
  const double nx = a[0], ny = a[1], nz = a[2];
  const double thL = theta_of(QL, a);
  for (int subit = 0; subit < 10; ++subit) {
    double t[5];
#pragma unroll
    for (int i = 0; i < 5; ++i) t[i] = (QL[i] + QR[i]) / 2.0;
    double rhoi = QL[0];
    double ui = QL[1] / QL[0] + 0.0, vi = QL[2] / QL[0] + 0.0, wi = QL[3] / QL[0] + 0.0;
    double rhoavg = t[0];
    t[1] += rhoavg * 0.0;
    t[2] += rhoavg * 0.0;
    t[3] += rhoavg * 0.0;
    double pavg = comp_pressure(t, gm1);
    double c2avg = gamma * (pavg / rhoavg);
    double cavg = sqrt(c2avg);
    double pb = pL + rhoavg * cavg * thL;
    double rhob = rhoi + (pb - pL) / c2avg;
    double tt = (pb - pL) / (rhoavg * cavg);
    double ub = ui - nx * tt, vb = vi - ny * tt, wb = wi - nz * tt;
    QR[0] = rhob; QR[1] = rhob * ub; QR[2] = rhob * vb; QR[3] = rhob * wb;
    QR[4] = pb / (gamma - 1.0) + 0.5 * rhob * (ub * ub + vb * vb + wb * wb);
  }
}

// Dirichlet-type BCs set both the phantom and the interior node to Qref
// (ucs/bc.tcc:1076-1083).  Callers handle them explicitly: merging that
// path with the QL-dependent ones below was mis-compiled for gfx950 (the
// phantom store picked up the pre-update interior state).
PCFD_DEV bool is_dirichlet(int type) { return type == BC_SONIC_INFLOW || type == BC_DIRICHLET; }

// noSlip wall at rest (GetViscousWallBoundaryVariables, ucs/compressible.tcc:
// 1544-1575, vel = 0, no bleed steps): the wall node itself is hard-set.
// Isothermal (twall >= 0): energy from the wall temperature; adiabatic
// (twall < 0): density and energy of the wall-normal neighbour nQ.
PCFD_DEV void noslip_state(double* QL, double* QR, double twall, const double* nQ, double gamma) {
  const double vel = 0.0;
  if (twall < 0.0) {
    QR[0] = QL[0] = nQ[0];
    QR[4] = QL[4] = nQ[4];
  } else {
    const double v2 = vel * vel + vel * vel + vel * vel;
    const double gm1 = gamma - 1.0;
    const double rhoEt = (twall * QL[0] / (gamma * gm1) + 0.5 * QR[0] * v2);
    QR[0] = QL[0];
    QR[4] = QL[4] = rhoEt;
  }
  QR[1] = QL[1] = QL[0] * vel;
  QR[2] = QL[2] = QL[0] * vel;
  QR[3] = QL[3] = QL[0] * vel;
}

// T, u, v, w of ComputeAuxiliaryVariables (ucs/compressible.tcc:1229-1243,
// T = gamma*P/rho through ComputePressure, :1090-1110), same operation order
PCFD_DEV void aux_tuvw(const double* Q, double gamma, double* t) {
  const double gm1 = gamma - 1.0;
  const double r = Q[0];
  const double uu = Q[1] / r, vv = Q[2] / r, ww = Q[3] / r;
  const double v2h = 0.5 * (uu * uu + vv * vv + ww * ww);
  const double p = gm1 * (Q[4] - r * v2h);
  t[0] = gamma * p / r;
  t[1] = uu;
  t[2] = vv;
  t[3] = ww;
}

// Sutherland's law, non-dimensional (ComputeViscosity, ucs/eqnset.h:251-268);
// T^1.5 is evaluated as T*sqrt(T) (two roundings instead of pow's one)
PCFD_DEV double sutherland(double T, double S) { return (1.0 + S) * (T * sqrt(T)) / (T + S); }

struct ViscParams {
  double gamma, re_tilde, pr, prt, suth;
  int sym2d;
};

// ViscousFlux (ucs/compressible.tcc:712-795), laminar (mut = 0), written on
// the right-hand side.  Q: averaged conservative state, T its temperature,
// g[12] = gradients of T, u, v, w (x, y, z each).
PCFD_DEV void viscous_flux(const double* Q, double T, const double* g, const double* a, const ViscParams& vp,
                           double* F) {
  const double mut = 0.0;
  const double Tx = g[0], Ty = g[1], Tz = g[2];
  const double ux = g[3], uy = g[4], uz = g[5];
  const double vx = g[6], vy = g[7], vz = g[8];
  const double wx = g[9], wy = g[10], wz = g[11];
  const double rho = Q[0];
  const double u = Q[1] / rho, v = Q[2] / rho, w = Q[3] / rho;
  const double mu = sutherland(T, vp.suth);
  const double tmut = (mu + mut);
  const double fact = -2.0 / 3.0 * (ux + vy + wz);
  const double tauxx = 2.0 * ux + fact;
  const double tauyy = 2.0 * vy + fact;
  const double tauzz = 2.0 * wz + fact;
  const double tauxy = uy + vx;
  const double tauxz = uz + wx;
  const double tauyz = vz + wy;
  const double RK = a[3] / vp.re_tilde;
  const double RKT = RK * tmut;
  const double cp = 1.0 / (vp.gamma - 1.0);
  const double k = mu / vp.pr * cp;
  const double kT = mut / vp.prt * cp;
  const double c1 = -(k + kT);
  const double Tn = Tx * a[0] + Ty * a[1] + Tz * a[2];
  const double tauxn = tauxx * a[0] + tauxy * a[1] + tauxz * a[2];
  const double tauyn = tauxy * a[0] + tauyy * a[1] + tauyz * a[2];
  const double tauzn = tauxz * a[0] + tauyz * a[1] + tauzz * a[2];
  F[0] = 0.0;
  F[1] = -RKT * (tauxn);
  F[2] = -RKT * (tauyn);
  F[3] = -RKT * (tauzn);
  F[4] = -RKT * (tauxn * u + tauyn * v + tauzn * w) + RK * c1 * Tn;
  if (vp.sym2d == 1) F[1] = 0.0;
  else if (vp.sym2d == 2) F[2] = 0.0;
  else if (vp.sym2d == 3) F[3] = 0.0;
}

// Column c of one side of ViscousJacobian (ucs/compressible.tcc:1632-1893)
// before the left side's sign flip.  Dx: that side's directional weights
// (+-dx/(rho s2)); u, v, w: edge-averaged velocity; Tn: that side's
// temperature-gradient weight; uS, vS, wS, PS, irS: that side's velocity,
// pressure and 1/rho.  The columns of rho u, rho v, rho w (c = 1..3, m =
// c - 1) share one form: row k is -2/3 Dx[m] a[k] + Dx[k] a[m], plus Dx.a on
// row m (which gives the reference's 4/3 Dx[m] a[m] + the other two terms).
PCFD_DEV void visc_jac_col(int c, const double* Dx, const double* a, double u, double v, double w, double RK,
                           double RKT, double Tn, double gm1, double uS, double vS, double wS, double PS, double irS,
                           double* col) {
  const double mc23 = -2.0 / 3.0;
  double d0, d1, d2, e4;
  if (c == 0) {
    const double dux = -u * Dx[0], duy = -u * Dx[1], duz = -u * Dx[2];
    const double dvx = -v * Dx[0], dvy = -v * Dx[1], dvz = -v * Dx[2];
    const double dwx = -w * Dx[0], dwy = -w * Dx[1], dwz = -w * Dx[2];
    const double dfact = mc23 * (dux + dvy + dwz);
    const double dtauxx = (2.0 * dux + dfact);
    const double dtauyy = (2.0 * dvy + dfact);
    const double dtauzz = (2.0 * dwz + dfact);
    const double dtauxy = duy + dvx;
    const double dtauxz = duz + dwx;
    const double dtauyz = dvz + dwy;
    d0 = dtauxx * a[0] + dtauxy * a[1] + dtauxz * a[2];
    d1 = dtauxy * a[0] + dtauyy * a[1] + dtauyz * a[2];
    d2 = dtauxz * a[0] + dtauyz * a[1] + dtauzz * a[2];
    const double v2 = (uS * uS + vS * vS + wS * wS);
    e4 = RK * Tn * (gm1 * 0.5 * v2 - PS * irS);
  } else if (c == 4) {
    d0 = d1 = d2 = 0.0;
    e4 = RK * Tn * gm1;
  } else {
    const int m = c - 1;
    const double Dm = m == 0 ? Dx[0] : (m == 1 ? Dx[1] : Dx[2]);
    const double am = m == 0 ? a[0] : (m == 1 ? a[1] : a[2]);
    const double sm = m == 0 ? uS : (m == 1 ? vS : wS);
    const double DA = Dx[0] * a[0] + Dx[1] * a[1] + Dx[2] * a[2];
    d0 = mc23 * Dm * a[0] + Dx[0] * am + (m == 0 ? DA : 0.0);
    d1 = mc23 * Dm * a[1] + Dx[1] * am + (m == 1 ? DA : 0.0);
    d2 = mc23 * Dm * a[2] + Dx[2] * am + (m == 2 ? DA : 0.0);
    e4 = RK * Tn * (-gm1 * sm);
  }
  col[0] = 0.0;
  col[1] = -RKT * d0;
  col[2] = -RKT * d1;
  col[3] = -RKT * d2;
  col[4] = -RKT * (d0 * u + d1 * v + d2 * w) + e4;
}

// Columns c of both sides of ViscousJacobian for the edge (QL, QR), left
// side already sign-flipped; symmetry2D rows zeroed.  The quantities of
// ucs/compressible.tcc:1632-1893 with each division by rho or s2 taken as a
// product with one reciprocal and T^1.5 as T sqrt(T).
PCFD_DEV void visc_jac_cols(int c, const double* QL, const double* QR, const double* xL, const double* xR,
                            const double* a, const ViscParams& vp, double* colL, double* colR) {
  const double gamma = vp.gamma, gm1 = gamma - 1.0;
  // edge-averaged state -> temperature -> Sutherland viscosity
  const double ra = 0.5 * (QL[0] + QR[0]), ira = 1.0 / ra;
  const double ua = 0.5 * (QL[1] + QR[1]) * ira, va = 0.5 * (QL[2] + QR[2]) * ira, wa = 0.5 * (QL[3] + QR[3]) * ira;
  const double pa = gm1 * (0.5 * (QL[4] + QR[4]) - ra * (0.5 * (ua * ua + va * va + wa * wa)));
  const double mu = sutherland(gamma * pa * ira, vp.suth);
  const double RK = a[3] / vp.re_tilde;
  const double RKT = RK * mu;
  const double irL = 1.0 / QL[0], irR = 1.0 / QR[0];
  const double uL = QL[1] * irL, vL = QL[2] * irL, wL = QL[3] * irL;
  const double uR = QR[1] * irR, vR = QR[2] * irR, wR = QR[3] * irR;
  const double PL = gm1 * (QL[4] - 0.5 * QL[0] * (uL * uL + vL * vL + wL * wL));
  const double PR = gm1 * (QR[4] - 0.5 * QR[0] * (uR * uR + vR * vR + wR * wR));
  const double u = 0.5 * (uL + uR), v = 0.5 * (vL + vR), w = 0.5 * (wL + wR);
  double dx[3], s2 = 0.0;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    dx[k] = (xR[k] - xL[k]);
    s2 += dx[k] * dx[k];
  }
  const double is2 = 1.0 / s2;
  const double fL = -irL * is2, fR = irR * is2;
  const double DxL[3] = {dx[0] * fL, dx[1] * fL, dx[2] * fL};
  const double DxR[3] = {dx[0] * fR, dx[1] * fR, dx[2] * fR};
  // conductivity k = mu Cp / Pr with Cp = 1/(gamma-1); turbulent part 0
  const double c1 = -(mu / (vp.pr * gm1));
  const double TnL = (DxL[0] * a[0] + DxL[1] * a[1] + DxL[2] * a[2]) * (gamma * irL) * c1;
  const double TnR = (DxR[0] * a[0] + DxR[1] * a[1] + DxR[2] * a[2]) * (gamma * irR) * c1;
  visc_jac_col(c, DxR, a, u, v, w, RK, RKT, TnR, gm1, uR, vR, wR, PR, irR, colR);
  visc_jac_col(c, DxL, a, u, v, w, RK, RKT, TnL, gm1, uL, vL, wL, PL, irL, colL);
#pragma unroll
  for (int k = 0; k < 5; ++k) colL[k] = -colL[k];
  if (vp.sym2d >= 1 && vp.sym2d <= 3) {
    colL[vp.sym2d] = 0.0;
    colR[vp.sym2d] = 0.0;
  }
}

// CalculateBoundaryVariables for the non-Dirichlet BC types on this path
// (ucs/bc.tcc:1058-1398).  QL is read only.  Returns false for types it does
// not handle (the phantom is then left untouched).
PCFD_DEV bool bc_state(int type, const double* QL, double* QR, const double* Qref, double pref, const double* a,
                       double gamma) {
  const double gm1 = gamma - 1.0;
  if (type == BC_SONIC_OUTFLOW || type == BC_NEUMANN) {
#pragma unroll
    for (int i = 0; i < 5; ++i) QR[i] = QL[i];
    return true;
  } else if (type == BC_FARFIELD) {
    farfield_state(QL, aux_pressure(QL, gm1), QR, Qref, pref, a, gamma);
    return true;
  } else if (type == BC_IMPERMEABLE || type == BC_SYMMETRY) {
    wall_state(QL, aux_pressure(QL, gm1), QR, a, gamma);
    return true;
  }
  return false;
}

// ExtrapolateVariables with ExtrapolateCorrection (ucs/compressible.tcc:1038-1052,
// ucs/eqnset.h:231-238): Qho = q + (chi/2 dQ + (1-chi) grad.dx) * phi
PCFD_DEV double extrap1(double q, double dQ, double gx, double gy, double gz, const double* dx, double chi,
                        double phi) {
  double corr = 0.5 * chi * dQ + (1.0 - chi) * (gx * dx[0] + gy * dx[1] + gz * dx[2]);
  return q + corr * phi;
}

// BadExtrapolation (ucs/compressible.tcc:1054-1078)
PCFD_DEV bool bad_state(const double* Q, double gamma) {
  double r = Q[0];
  double u = Q[1] / r, v = Q[2] / r, w = Q[3] / r, E = Q[4];
  double v2h = 0.5 * (u * u + v * v + w * w);
  double p = (gamma - 1.0) * (E - r * v2h);
  if (p < 1.0e-10) return true;
  if (r < 0.0) return true;
  if (E < 1.0e-10) return true;
  return false;
}

// RoeVariables state only (ucs/compressible.tcc:533-578)
PCFD_DEV void roe_state(const double* QL, const double* QR, double gamma, double* Qr) {
  const double gm1 = gamma - 1.0;
  double rhoL = QL[0], rhoR = QR[0];
  double uL = QL[1] / QL[0], uR = QR[1] / QR[0];
  double vL = QL[2] / QL[0], vR = QR[2] / QR[0];
  double wL = QL[3] / QL[0], wR = QR[3] / QR[0];
  double v2L = uL * uL + vL * vL + wL * wL;
  double v2R = uR * uR + vR * vR + wR * wR;
  double PL = gm1 * (QL[4] - 0.5 * rhoL * v2L);
  double PR = gm1 * (QR[4] - 0.5 * rhoR * v2R);
  double hL = (QL[4] + PL) / rhoL;
  double hR = (QR[4] + PR) / rhoR;
  double rr = sqrt(rhoL * rhoR);
  double sigma = rr / (rhoL + rr);
  double ur = uL + sigma * (uR - uL);
  double vr = vL + sigma * (vR - vL);
  double wr = wL + sigma * (wR - wL);
  double hr = hL + sigma * (hR - hL);
  double v2hr = 0.5 * (ur * ur + vr * vr + wr * wr);
  Qr[0] = rr;
  Qr[1] = rr * ur;
  Qr[2] = rr * vr;
  Qr[3] = rr * wr;
  Qr[4] = rr / gamma * (hr + gm1 * v2hr);
}

// ComputeLSQCoefficients (ucs/gradient.tcc:140-168)
PCFD_DEV void lsq_coeff(const double* s, const double* d, double* we) {
  double r11 = s[0], r12 = s[1], r13 = s[2];
  double r12_r11 = (r11 == 0.0) ? 0.0 : r12 / r11;
  double r22 = s[3] - r12 * r12_r11;
  double r23 = s[4] - r12_r11 * r13;
  double r13_r11 = (r11 == 0.0) ? 0.0 : r13 / r11;
  double r23_r22 = (r22 == 0.0) ? 0.0 : r23 / r22;
  double r33 = s[5] - r13 * r13_r11 - r23 * r23_r22;
  double dykdx = (d[1] - (r12_r11)*d[0]);
  we[2] = (r33 == 0.0) ? 0.0 : (d[2] - r13_r11 * d[0] - r23_r22 * dykdx) / r33;
  we[1] = (r22 == 0.0) ? 0.0 : (dykdx - r23 * we[2]) / r22;
  we[0] = (r11 == 0.0) ? 0.0 : (d[0] - r12 * we[1] - r13 * we[2]) / r11;
}

// CompressibleEqnSet::ApplyDQ clipping (ucs/compressible.tcc:928-993)
PCFD_DEV void apply_dq(const double* dQ, double* Q, double gamma) {
  const double gm1 = gamma - 1.0;
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
}

}  // namespace pcfd
