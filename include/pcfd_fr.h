/*
 * Finite-rate (FR) reacting Euler on one GPU: the first-order implicit path
 * ucs.x runs for equationSet = compressibleEulerFR.  It replaces, for one
 * partition:
 *   CompressibleFREqnSet<Type> (ucs/compressibleFR.h, ucs/compressibleFR.tcc):
 *     HLLCFlux :300-548, Eigensystem :149-290, GetFarfieldBoundaryVariables
 *     :940-1040, GetInviscidWallBoundaryVariables :1042-1154, ApplyDQ
 *     :885-937, ContributeTemporalTerms :1318-1464, MaxEigenvalue :1602-1637,
 *     SourceTerm :1276-1312, ComputeAuxiliaryVariables :754-814
 *   UpdateBCs / CalculateBoundaryVariables   ucs/bc.tcc:1058-1457
 *   SpatialResidual (first order)            ucs/residual.tcc:12-122
 *   ComputeTimesteps                         ucs/timestep.tcc:6-143
 *   ComputeJacobians (+ temporal, source)    ucs/jacobian.tcc:12-250,253-304,436-535
 *   CRSMatrix::PrepareSGS + CRS::SGS         ucs/crsmatrix.tcc:839-876, ucs/crs.tcc:61-173
 *   SolutionSpace::PreTimeAdvance / NewtonIterate / PostTimeAdvance
 *                                            ucs/solutionSpace.tcc:572-916
 * State per node (non-dimensional): rho_1..rho_ns, u, v, w, T; neqn = ns + 4.
 * The reference's auxiliary values (P, rho, cv_i, conc_i: nvars = 3 ns + 6
 * per node in its q) are functions of these and are recomputed, not stored.
 * The chemistry model (species thermo, reactions, reference density, time
 * and temperature) comes from a pcfd_chem object (include/pcfd_chem.h).
 * Boundary types handled: sonicInflow, dirichlet, sonicOutflow, neumann,
 * farField, impermeableWall, symmetry (the characteristic states, or with
 * noCVBC the upwind / mirrored ones, ucs/compressibleFR.tcc:962-973,1109-1128).
 */
#ifndef PCFD_FR_H
#define PCFD_FR_H

#include "pcfd.h"
#include "pcfd_chem.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pcfd_fr pcfd_fr;

/* Param<Type> members of an FR case (ucs/param.h), after PostCompute. */
typedef struct {
  double ref_pressure;  /* Pa (refPressure) */
  double ref_velocity;  /* m/s (PostCompute's speed of sound of the MW/Cp gas) */
  double Pref;          /* non-dimensional freestream pressure P(Qinf) (UpdateQinf, compressibleFR.tcc:1465-1518) */
  double beta;          /* the "beta" field: max(betaMin, M^2) below Mach 1, else 1 (solutionSpace.tcc:236-248) */
  int symmetry2d;       /* symmetry2D */
  int reactions;        /* reactionsOn */
  int local_dt;         /* localTimeStepping */
  double dt_param;      /* timeStep (-1 by default; enters the temporal Jacobian as vol / timeStep) */
  double cfl, cfl_start;
  int cfl_ramp;         /* cflRampingSteps */
  int nsgs;             /* numberSGS */
  int jac_freq;         /* jacobianFreq */
  int precision;        /* PCFD_PLAIN_FP64 (drop-in) or PCFD_CHOSEN_FP32 (Jacobian and SGS in FP32) */
  int no_cvbc;          /* noCVBC: upwind farfield and mirrored-velocity walls instead of the characteristic states */
} pcfd_fr_params;

/* mesh: as for pcfd_create (its qref/twall are not used here); qref: per face
 * tag the reference state's ns + 4 native values (BCObj::Qref, the freestream
 * Qinf unless the .bc file sets one).  chem is borrowed and must outlive the
 * context.  Supported species counts: 5. */
int pcfd_fr_create(const pcfd_mesh_desc* mesh, pcfd_chem* chem, const double* qref, pcfd_fr** out);
int pcfd_fr_destroy(pcfd_fr* f);
const char* pcfd_fr_last_error(const pcfd_fr* f);
int pcfd_fr_set_params(pcfd_fr* f, const pcfd_fr_params* p);
/* ncolor, nblocks, device bytes held */
int pcfd_fr_info(const pcfd_fr* f, int* ncolor, long long* nblocks, long long* device_bytes);

/* State transfer, (nnode + nbnode) rows of `stride` doubles (host arrays):
 * upload reads the first ns + 4 values of each row; download writes them and,
 * when stride >= 3 ns + 6, the auxiliary values P, rho, cv_i, conc_i after
 * them (the reference's q layout). */
int pcfd_fr_upload_q(pcfd_fr* f, const double* q, int stride);
int pcfd_fr_download_q(pcfd_fr* f, double* q, int stride);

/* Single phases (host output arrays may be NULL) */
int pcfd_fr_update_bcs(pcfd_fr* f);
/* b_out [nnode*neqn]; norms [1+neqn]: residGlobal and the per-equation norms */
int pcfd_fr_residual(pcfd_fr* f, double* b_out, double* norms);
int pcfd_fr_timestep(pcfd_fr* f, double cfl, double* dt_out, double* dtmin);
int pcfd_fr_set_timestep(pcfd_fr* f, const double* dt);
/* FD Jacobian on the current dt: off-diagonals, boundary part with the BC
 * states re-evaluated (the phantoms / hard-set states are updated as the
 * reference's does), column sums to the diagonal, minus the source
 * Jacobian, plus the temporal terms */
int pcfd_fr_jacobian(pcfd_fr* f);
/* blocks of the assembled (unfactored) matrix in the caller's block-CRS
 * layout: ia [nnode+1], ja [ia[nnode]] with every row's diagonal and
 * neighbours in any order -> M [ia[nnode]*neqn*neqn] */
int pcfd_fr_download_jacobian(pcfd_fr* f, const int* ia, const int* ja, double* M);
/* the block-CRS pattern of the device matrix: each row's diagonal, then its
 * neighbours in the order SGS sums them; ia [nnode+1], ja [nnode + 2 nedge] */
int pcfd_fr_pattern(pcfd_fr* f, int* ia, int* ja);
/* PrepareSGS + nsgs symmetric multicolour sweeps from x = 0 on the current
 * residual; x_out [nnode*neqn]; delta = |xNorm - xNormOld| of the last sweep */
int pcfd_fr_sgs(pcfd_fr* f, int nsgs, double* x_out, double* delta);
/* the row order of one forward sweep (colour by colour) [nnode] */
int pcfd_fr_sgs_order(pcfd_fr* f, int* order);
/* NaN/Inf rows of the correction zeroed (counted), then ApplyDQ on every node */
int pcfd_fr_apply_dq(pcfd_fr* f, int* nzeroed);

/* The steady loop: PreIterate (iteration counter 0, first time step), then
 * per step PreTimeAdvance + NewtonIterate + PostTimeAdvance.  norms
 * [1+neqn] of the step's residual (may be NULL). */
int pcfd_fr_init_iterations(pcfd_fr* f);
int pcfd_fr_step(pcfd_fr* f, double* norms, int* nzeroed);
/* up to maxiter steps, stopping when residGlobal <= drop * its first value;
 * history [maxiter] (may be NULL); *iters steps taken */
int pcfd_fr_solve(pcfd_fr* f, int maxiter, double drop, double* history, int* iters);

/* The learned part (no counterpart in the reference): with a surrogate set,
 * the residual's finite-rate source takes its production rates from the
 * MLP surrogate with the exact-rate fallback of
 * pcfd_chem_surrogate_wdot_device (include/pcfd_chem.h); the source
 * Jacobian in the preconditioner stays exact.  mlp = NULL restores the exact
 * rates.  mlp and d are borrowed (d is copied). */
int pcfd_fr_set_surrogate(pcfd_fr* f, pcfd_mlp* mlp, const pcfd_surrogate_desc* d);
/* cells sent to the exact rates by the last residual; residuals evaluated
 * and cells sent back since the surrogate was set; device ms of the
 * surrogate passes */
int pcfd_fr_surrogate_stats(const pcfd_fr* f, int* last_fallback, long long* calls, long long* fallback_total,
                            double* ms_total);

/* Device milliseconds per phase since the last reset: [0] Jacobian, [1] time
 * step, [2] boundary states, [3] residual, [4] SGS, [5] update, [6] the
 * Jacobian's edge kernel alone, [7] the residual's edge flux kernel alone */
#define PCFD_FR_NPHASE 8
int pcfd_fr_phase_times(pcfd_fr* f, double* ms, int reset);

#ifdef __cplusplus
}
#endif
#endif
