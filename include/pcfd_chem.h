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

/* CompressibleFREqnSet::ComputeAuxiliaryVariables (ucs/compressibleFR.tcc:754-814) for n states,
 * ideal-gas EOS per species (ucs/EOS.tcc:36-38, 79-82).  q [n*(ns+4)]: rho_1..rho_ns, u, v, w, T
 * (non-dimensional); aux [n*(2+2ns)]: p, rho, cv_1..cv_ns, conc_1..conc_ns with
 *   p = sum_i (rho_i rho_ref) R_i (T T_ref) / ref_pressure,  rho = sum_i rho_i,
 *   cv_i = (Cp_i(T T_ref) - R_i) / (ref_velocity^2 / T_ref),  conc_i = rho_i / MW_i / 1000
 * (MassToMole, ucs/chem.tcc:980-986, applied to the non-dimensional rho_i as the reference does).
 * ref_pressure (Pa) and ref_velocity (m/s) are the case's Param scales (ucs/param.tcc:355-375). */
int pcfd_chem_fr_aux(pcfd_chem* c, int n, const double* q, double ref_pressure, double ref_velocity, double* aux);

/* CompressibleFREqnSet::NewtonFindTGivenP (ucs/compressibleFR.tcc:2342-2378): the temperature at
 * which the species densities rhoi [n*ns] give the pressure p [n] (both non-dimensional), by
 * Newton iteration from Tinit [n] with a forward difference (1e-8 K) for dP/dT, at most 28 steps,
 * stopping when |dT| / T_ref < 1e-15.  T [n] non-dimensional; iters [n] (may be NULL): the step
 * at which it stopped, 28 when it did not converge (the reference warns and returns the last T). */
int pcfd_chem_t_from_p(pcfd_chem* c, int n, const double* rhoi, const double* p, const double* Tinit,
                       double ref_pressure, double* T, int* iters);

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

/* ---- learned surrogate (no counterpart in the reference) -------------
 * MLP 7 -> 64 -> 64 -> 5, tanh hidden layers, linear output; BF16 MFMA
 * products with FP32 accumulation; FP32 master parameters, Adam.
 * Parameter vector (float32, PCFD_MLP_NPARAM), tile-padded rows:
 *   W1[64][16] (cols 7..15 zero), b1[64], W2[64][64], b2[64],
 *   W3[16][64] (rows 5..15 zero), b3[16] (5..15 zero). */
#define PCFD_MLP_NPARAM 6288
typedef struct pcfd_mlp pcfd_mlp;

int pcfd_mlp_create(unsigned long long seed, pcfd_mlp** out);  /* Glorot-uniform weights, zero biases */
int pcfd_mlp_destroy(pcfd_mlp* m);
const char* pcfd_mlp_last_error(const pcfd_mlp* m);
double pcfd_mlp_last_kernel_ms(const pcfd_mlp* m);
int pcfd_mlp_set_params(pcfd_mlp* m, const float* params);  /* also resets the Adam state */
/* Error network of the production-rate surrogate m: a second network of
 * the same shape trained on m's own errors (the targets
 * ln(|sigma_i (y_i - target_i)| + 1e-3), normalised by the desc's
 * err_mid/err_scale).  With one set, the pcfd_chem_surrogate_* calls on m
 * (and an FR context's surrogate) evaluate both networks and apply the
 * desc's err_tol rule; partner = NULL removes it.  Borrowed: it must outlive
 * its use. */
int pcfd_mlp_set_partner(pcfd_mlp* m, pcfd_mlp* partner);
int pcfd_mlp_get_params(pcfd_mlp* m, float* params);
/* y[n*5] = net(x[n*7]); host arrays / device arrays */
int pcfd_mlp_forward(pcfd_mlp* m, int n, const float* x, float* y);
int pcfd_mlp_forward_device(pcfd_mlp* m, int n, const float* x, float* y);
/* mean squared error over n samples x 5 outputs and its gradient[NPARAM] */
int pcfd_mlp_loss_grad(pcfd_mlp* m, int n, const float* x, const float* t, float* loss, float* grad);
/* Adam on mini-batches drawn with replacement (splitmix64 from seed),
 * cosine-decayed step size; the data set is copied to the device once.
 * loss_history[steps] (may be NULL) receives each batch loss. */
int pcfd_mlp_train(pcfd_mlp* m, int n, const float* x, const float* t, int steps, int batch, float lr,
                   unsigned long long seed, float* loss_history);

/* Feature/target transforms of the production-rate surrogate and its
 * fallback rule (see pcfd_chem_surrogate_wdot_device). */
typedef struct {
  double lnT_mid, lnT_half;      /* f0 = (ln T - mid) / half */
  double lnrho_mid, lnrho_half;  /* f1 = (ln rho - mid) / half */
  double y_mid, y_scale;         /* f2..f6 = (g(Y_i) - y_mid) / y_scale, */
  double y_eps;                  /*   g(Y) = ln(Y + y_eps) when y_eps > 0, else Y */
  double rate_scale;             /* s0 [1/s]: target_i = asinh(wdot_i / (rho s0)) / sigma_i */
  double sigma[5];
  double T_min, T_max, rho_min, rho_max;  /* training box; outside it the exact rates are used */
  double tol, floor;             /* element-balance bound: fall back when
                                    imbalance > tol * (sum|w_i/M_i| + rho*floor) */
  double atoms[2][5];            /* atoms of the two conserved elements (N, O
                                    for air) in each species, in the model's
                                    species numbering; imbalance =
                                    sum_e |sum_i atoms[e][i] w_i / M_i| */
  /* with an error network (pcfd_mlp_set_partner): its output i, g_i, gives
     the predicted error of output i in asinh units, e_i = exp(g_i *
     err_scale[i] + err_mid[i]); fall back when
       max_i rho s0 (sinh(|z_i| + e_i) - sinh(|z_i|)) > err_tol * (max_i |w_i| + rho * rate_floor),
     z_i = sigma_i * output_i, rate_floor in 1/s; ignored without one */
  double err_tol, rate_floor;
  double err_mid[5], err_scale[5];
} pcfd_surrogate_desc;

/* Production rates by the surrogate for n cells (device arrays rhoi[n*5]
 * kg/m^3, T[n] K -> wdot[n*5] kg/(m^3 s)), with the exact rates of `chem`
 * for every cell the rule sends back; *nfallback counts those cells and
 * *ms the device time of both passes. */
int pcfd_chem_surrogate_wdot_device(pcfd_chem* chem, pcfd_mlp* mlp, const pcfd_surrogate_desc* d, int n,
                                    const double* rhoi, const double* T, double* wdot, int* nfallback, double* ms);

/* The same as pcfd_chem_surrogate_wdot_device without allocating device
 * memory, for callers that cannot call hipMalloc (code under HIP graph
 * capture, tight serving loops, harnesses that forbid it).  No counterpart in
 * the reference.  The fallback list, per-pass counts and their offsets live
 * in a fixed library-owned device array (one per GPU, 17.3 MB, held until the
 * process ends) that limits a call to PCFD_SURROGATE_NOALLOC_MAX_CELLS cells;
 * larger n is refused with PCFD_ERR_ARG.  Every call on one GPU shares that
 * array: calls that overlap on one GPU (on different mlp objects or streams)
 * race and must be serialised by the caller. */
#define PCFD_SURROGATE_NOALLOC_MAX_CELLS (1 << 22)
int pcfd_chem_surrogate_wdot_device_noalloc(pcfd_chem* chem, pcfd_mlp* mlp, const pcfd_surrogate_desc* d, int n,
                                            const double* rhoi, const double* T, double* wdot, int* nfallback,
                                            double* ms);

/* the same on host arrays (copied in and out; *ms is device time only) */
int pcfd_chem_surrogate_wdot(pcfd_chem* chem, pcfd_mlp* mlp, const pcfd_surrogate_desc* d, int n, const double* rhoi,
                             const double* T, double* wdot, int* nfallback, double* ms);

/* exact production rates on device arrays (the method without the surrogate) */
int pcfd_chem_wdot_device(pcfd_chem* c, int n, const double* rhoi, const double* T, double* wdot);

#ifdef __cplusplus
}
#endif
#endif
