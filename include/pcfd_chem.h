/*
 * Finite-rate chemistry on the GPU: mass production rates, the finite-rate
 * source term and its Jacobian block, and the learned surrogate (a small
 * MLP) that stands in for them with an exact fallback.
 *
 * Reference interfaces replaced:
 *   ChemModel::GetMassProductionRates   ucs/chem.tcc:575-583
 *   Reaction::GetMassProductionRate     ucs/reaction.tcc:762-856
 *     (GetForwardReactionRate :607-624, GetEquilibriumReactionRate
 *      :628-679, GetBackwardReactionRate :682-706)
 *   CompressibleFREqnSet::SourceTerm     ucs/compressibleFR.tcc:1276-1312
 *   EqnSet::SourceTermJacobian           ucs/eqnset.tcc:162-187
 * The model arrives as plain arrays (what ChemModel holds after reading
 * the .rxn file and the thermo database), in SI with molecular weights in
 * kg/mol (ucs/species.tcc:167-171).  FR state per cell: rho_1..rho_ns, u,
 * v, w, T (non-dimensional), neqn = ns + 4.
 */
#ifndef PCFD_CHEM_H
#define PCFD_CHEM_H

#include "pcfd.h" /* PCFD_OK / PCFD_ERR_* return codes */

#ifdef __cplusplus
extern "C" {
#endif

#define PCFD_CHEM_MAXS 8   /* species */
#define PCFD_CHEM_MAXR 16  /* reactions */

typedef struct pcfd_chem pcfd_chem;

typedef struct {
  int nspecies;
  int nreactions;
  const double* mw;         /* [ns] kg/mol */
  const double* nasa_lo;    /* [ns*7] NASA-7 coefficients used for T <= 1000 K */
  const double* nasa_hi;    /* [ns*7] NASA-7 coefficients used for T > 1000 K */
  const int* rxn_type;      /* [nr] 0 Arrhenius, 1 ModArrhenius, 2 GuptaModArrhenius, 3 Power */
  const int* third_body;    /* [nr] 1: third-body concentration multiplies the rate */
  const double* A;          /* [nr] */
  const double* EA;         /* [nr] */
  const double* n;          /* [nr] temperature exponent */
  const int* nloc;          /* [nr] species taking part (reaction-local list length) */
  const int* loc;           /* [nr*8] global species index of local slot k, parse order */
  const double* nup;        /* [nr*8] reactant coefficients nu' */
  const double* nupp;       /* [nr*8] product coefficients nu'' */
  const double* tb_eff;     /* [nr*8] third-body efficiencies (1 where unlisted) */
  double ref_density;       /* kg/m^3 */
  double ref_time;          /* s */
  double ref_temperature;   /* K */
} pcfd_chem_desc;

int pcfd_chem_create(const pcfd_chem_desc* desc, pcfd_chem** out);
int pcfd_chem_destroy(pcfd_chem* c);
const char* pcfd_chem_last_error(const pcfd_chem* c);

/* GetMassProductionRates for n states (host arrays): rhoi[n*ns] kg/m^3,
 * T[n] K -> wdot[n*ns] kg/(m^3 s). */
int pcfd_chem_wdot(pcfd_chem* c, int n, const double* rhoi, const double* T, double* wdot);

/* SourceTerm (and SourceTermJacobian when jac != NULL) for n cells (host
 * arrays): q[n*neqn] FR state, vol[n] -> src[n*neqn], jac[n*neqn*neqn]
 * (row-major dS_j/dQ_i at [j*neqn + i], one-sided h = 1e-8). */
int pcfd_chem_source(pcfd_chem* c, int n, const double* q, const double* vol, double* src, double* jac);

/* The same on device-resident arrays (device pointers, no copies): the
 * call an FR residual / Jacobian assembly makes per Newton step. */
int pcfd_chem_source_device(pcfd_chem* c, int n, const double* q, const double* vol, double* src, double* jac);

/* Device time (ms) of the last pcfd_chem_* launch, from HIP events on the
 * chemistry stream. */
double pcfd_chem_last_kernel_ms(const pcfd_chem* c);

#ifdef __cplusplus
}
#endif
#endif
