// Finite-rate chemistry model on the device (shared by the exact kernels
// and the surrogate's fallback): ChemModel / Reaction arithmetic of
// ucs/chem.tcc:575-583 and ucs/reaction.tcc:607-856 in FP64.
#pragma once
#include <hip/hip_runtime.h>

#include "../../include/pcfd_chem.h"

namespace pcfd {

constexpr int kChemBlock = 128;
constexpr double kUnivR = 8.31447215;  // J/(mol K), ucs/chem_constants.h:5

struct ChemDev {
  int ns, nr;
  double mw[PCFD_CHEM_MAXS];
  double lo[PCFD_CHEM_MAXS][7], hi[PCFD_CHEM_MAXS][7];
  int type[PCFD_CHEM_MAXR], tb[PCFD_CHEM_MAXR], nloc[PCFD_CHEM_MAXR], loc[PCFD_CHEM_MAXR][PCFD_CHEM_MAXS];
  double A[PCFD_CHEM_MAXR], EA[PCFD_CHEM_MAXR], n[PCFD_CHEM_MAXR];
  double nup[PCFD_CHEM_MAXR][PCFD_CHEM_MAXS], nupp[PCFD_CHEM_MAXR][PCFD_CHEM_MAXS],
      eff[PCFD_CHEM_MAXR][PCFD_CHEM_MAXS];
  double ref_density, ref_time, ref_temperature;
};

// isWholeNumber (ucs/macros.h:24-30); integer exponents go through pow of
// the promoted integer like the reference's pow(Type, Int)
__device__ __forceinline__ double powr(double x, double e) {
  const int up = (int)(e + 0.5), flat = (int)e;
  return pow(x, up == flat ? (double)flat : e);
}

__device__ __forceinline__ double kf_of(const ChemDev& m, int r, double T) {
  const double A = m.A[r], EA = m.EA[r], n = m.n[r];
  switch (m.type[r]) {
    case 0: return (A * exp(-EA / (kUnivR * T)));
    case 1: return (A * pow(T, n) * exp(-EA / (kUnivR * T)));
    case 2: return (A * pow(T, n) * exp(-EA / T));
    case 3: return (A * pow(T, n));
  }
  return -999.0;
}

// GetEquilibriumReactionRate (ucs/reaction.tcc:628-679)
__device__ __forceinline__ double kc_of(const ChemDev& m, int r, double T) {
  double nu = 0.0, d[7] = {0, 0, 0, 0, 0, 0, 0};
  const bool high = T > 1000.0;  // GetThermoCoeff: pinned ranges pick the same set
  for (int i = 0; i < m.nloc[r]; ++i) {
    const int s = m.loc[r][i];
    const double* a = high ? m.hi[s] : m.lo[s];
    const double dnu = m.nupp[r][i] - m.nup[r][i];
    nu += dnu;
#pragma unroll
    for (int k = 0; k < 7; ++k) d[k] += dnu * a[k];
  }
  const double Kp =
      exp(d[0] * (log(T) - 1.0) + T * (d[1] / 2.0 + T * (d[2] / 6.0 + T * (d[3] / 12.0 + d[4] / 20.0 * T))) -
          d[5] / T + d[6]);
  double Kc = Kp;
  if (!(fabs(nu) < 1.0e-15)) Kc *= powr(101325.0 / (kUnivR * T), nu);
  return Kc;
}

// GetMassProductionRates for one state: rhoi (kg/m^3), T (K) -> wdot (kg/m^3/s)
__device__ inline void wdot_of(const ChemDev& m, const double* rhoi, double T, double* wdot) {
  double X[PCFD_CHEM_MAXS];
  for (int k = 0; k < m.ns; ++k) {
    X[k] = rhoi[k] / m.mw[k];
    wdot[k] = 0.0;
  }
  for (int r = 0; r < m.nr; ++r) {
    const double Kf = kf_of(m, r, T);
    const double Kb = Kf / kc_of(m, r, T);
    double pf = 1.0, pd = 1.0;
    for (int k = 0; k < m.nloc[r]; ++k) {
      const double x = X[m.loc[r][k]];
      pf *= powr(x, m.nup[r][k]);
      pd *= powr(x, m.nupp[r][k]);
    }
    double Mc = 1.0;
    if (m.tb[r]) {
      Mc = 0.0;
      for (int k = 0; k < m.ns; ++k) Mc += X[k];
      for (int k = 0; k < m.nloc[r]; ++k) Mc += (m.eff[r][k] - 1.0) * X[m.loc[r][k]];
    }
    const double net = Kf * pf - Kb * pd;
    // species order; the first local occurrence of a species carries its nu
    for (int i = 0; i < m.ns; ++i) {
      int kk = -1;
      for (int k = 0; k < m.nloc[r]; ++k)
        if (m.loc[r][k] == i) {
          kk = k;
          break;
        }
      const double nu = kk >= 0 ? m.nupp[r][kk] - m.nup[r][kk] : 0.0;
      if (nu == 0.0) continue;  // AlmostEqualRelative(nu, 0, 1e-15): contributes +0
      double w = nu * Mc * net;
      w *= m.mw[i];
      wdot[i] += w;
    }
  }
}

// SourceTerm (ucs/compressibleFR.tcc:1276-1312) on the FR state
__device__ inline void source_of(const ChemDev& m, const double* Q, double vol, double* src) {
  double rhoi[PCFD_CHEM_MAXS], w[PCFD_CHEM_MAXS];
  const double T = Q[m.ns + 3] * m.ref_temperature;
  for (int i = 0; i < m.ns; ++i) rhoi[i] = Q[i] * m.ref_density;
  wdot_of(m, rhoi, T, w);
  for (int i = 0; i < m.ns; ++i) {
    w[i] /= (m.ref_density / m.ref_time);
    src[i] = vol * w[i];
  }
}

// wdot_of of a state and of its ns copies with one species density raised
// (copy c holds rhop[c] in place of rhoi[c]).  The rate constants are
// functions of T alone and a reaction's mass-action factors of the species
// a copy leaves alone are the state's, so both are evaluated once per
// reaction and shared; every copy's products and sums, in their order, are
// wdot_of's, so w[(1 + c) * ns + i] equals wdot_of of copy c bit for bit
// and w[i] that of the state.  w [(ns + 1) * ns].
__device__ inline void wdot_rho_columns(const ChemDev& m, const double* rhoi, const double* rhop, double T,
                                        double* w) {
  const int ns = m.ns;
  double X[PCFD_CHEM_MAXS], Xp[PCFD_CHEM_MAXS];
  for (int k = 0; k < ns; ++k) {
    X[k] = rhoi[k] / m.mw[k];
    Xp[k] = rhop[k] / m.mw[k];
  }
  for (int k = 0; k < (ns + 1) * ns; ++k) w[k] = 0.0;
  for (int r = 0; r < m.nr; ++r) {
    const double Kf = kf_of(m, r, T);
    const double Kb = Kf / kc_of(m, r, T);
    const int nl = m.nloc[r];
    double ff[PCFD_CHEM_MAXS], fd[PCFD_CHEM_MAXS], nu[PCFD_CHEM_MAXS];
    for (int k = 0; k < nl; ++k) {
      const double x = X[m.loc[r][k]];
      ff[k] = powr(x, m.nup[r][k]);
      fd[k] = powr(x, m.nupp[r][k]);
    }
    for (int i = 0; i < ns; ++i) {
      int kk = -1;
      for (int k = 0; k < nl; ++k)
        if (m.loc[r][k] == i) {
          kk = k;
          break;
        }
      nu[i] = kk >= 0 ? m.nupp[r][kk] - m.nup[r][kk] : 0.0;
    }
    for (int s = 0; s <= ns; ++s) {
      const int c = s - 1;  // -1: the state itself
      double pf = 1.0, pd = 1.0;
      for (int k = 0; k < nl; ++k) {
        if (m.loc[r][k] == c) {
          pf *= powr(Xp[c], m.nup[r][k]);
          pd *= powr(Xp[c], m.nupp[r][k]);
        } else {
          pf *= ff[k];
          pd *= fd[k];
        }
      }
      double Mc = 1.0;
      if (m.tb[r]) {
        Mc = 0.0;
        for (int k = 0; k < ns; ++k) Mc += k == c ? Xp[k] : X[k];
        for (int k = 0; k < nl; ++k) {
          const int l = m.loc[r][k];
          Mc += (m.eff[r][k] - 1.0) * (l == c ? Xp[l] : X[l]);
        }
      }
      const double net = Kf * pf - Kb * pd;
      double* ws = w + s * ns;
      for (int i = 0; i < ns; ++i) {
        if (nu[i] == 0.0) continue;
        double v = nu[i] * Mc * net;
        v *= m.mw[i];
        ws[i] += v;
      }
    }
  }
}

// source_of of the FR state Q and of its ns copies with Q[c] raised by h
// (the species columns of the forward-difference source Jacobian):
// src[i] is the state's, src[(1 + c) * ns + i] copy c's, each bit for bit
// what source_of gives.  src [(ns + 1) * ns].
__device__ inline void source_rho_columns(const ChemDev& m, const double* Q, double h, double vol, double* src) {
  const int ns = m.ns;
  double rhoi[PCFD_CHEM_MAXS], rhop[PCFD_CHEM_MAXS];
  const double T = Q[ns + 3] * m.ref_temperature;
  for (int i = 0; i < ns; ++i) {
    rhoi[i] = Q[i] * m.ref_density;
    rhop[i] = (Q[i] + h) * m.ref_density;
  }
  wdot_rho_columns(m, rhoi, rhop, T, src);
  for (int k = 0; k < (ns + 1) * ns; ++k) {
    double v = src[k];
    v /= (m.ref_density / m.ref_time);
    src[k] = vol * v;
  }
}


}  // namespace pcfd
