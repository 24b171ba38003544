/*
 * pcfd.h -- C ABI of the MI355X (gfx950) implicit edge-based finite-volume
 * hot path.  Plain pointers and sizes only; every call is synchronous with
 * respect to the host buffers it takes, returns 0 on success and a negative
 * code on failure (pcfd_last_error() gives the message).  No C++ exception
 * crosses this boundary.  One context = one GPU stream; calls on one
 * context must not overlap (the reference is single-threaded per rank too).
 *
 * Reference interfaces replaced (paths relative to the reference tree):
 *   Driver/Bdriver edge loops + Kernel ABI   ucs/driver.tcc:5-306, ucs/kernel.h:8-45
 *   UpdateBCs / BC_Kernel                    ucs/bc.tcc:1399-1457, 723-745
 *   ComputeNodeLSQCoefficients               ucs/gradient.tcc:114-138
 *   Gradient::Compute                        ucs/gradient.tcc:56-112
 *   Limiter::Compute                         ucs/limiters.tcc:52-132
 *   ComputeResiduals / SpatialResidual       ucs/residual.tcc:12-122
 *   ComputeTimesteps                         ucs/timestep.tcc:6-77
 *   ComputeJacobians                         ucs/jacobian.tcc:12-250
 *   CRSMatrix::PrepareSGS                    ucs/crsmatrix.tcc:839-876
 *   CRS::SGS                                 ucs/crs.tcc:61-173
 *   NewtonIterate update (NaN rows, ApplyDQ) ucs/solutionSpace.tcc:772-804,
 *                                            ucs/compressible.tcc:928-993
 *   ExplicitSolve                            ucs/solve.tcc:70-140
 *   SolutionSpace::NewtonIterate loop        ucs/solutionSpace.tcc:572-904
 */
#ifndef PCFD_H
#define PCFD_H

#ifdef __cplusplus
extern "C" {
#endif

#define PCFD_OK 0
#define PCFD_ERR_ARG (-1)
#define PCFD_ERR_HIP (-2)
#define PCFD_ERR_STATE (-3)
#define PCFD_ERR_UNSUPPORTED (-4)

/* precision modes */
#define PCFD_PLAIN_FP64 0   /* everything FP64: the drop-in */
#define PCFD_CHOSEN_FP32 1  /* Jacobian storage + SGS in FP32, residual/gradients/norms FP64 */

typedef struct pcfd_ctx pcfd_ctx;

/* Mesh description: the arrays ucs.x's Mesh<Real> holds after BuildMaps +
 * CalcMetrics (ucs/mesh.h:199-254, ucs/uns_base.h:12-37).  Node ids: interior
 * [0,nnode), boundary phantoms [nnode, nnode+nbnode) with phantom id of
 * half-edge k equal to nnode+k (ucs/mesh.tcc:1102-1133).  Copied at create. */
typedef struct {
  int nnode;
  int nbnode;               /* boundary half-edges == phantom nodes */
  int nedge;
  const int* edges;         /* [nedge*2] n0<n1 */
  const double* edge_a;     /* [nedge*4] unit normal n0->n1, a[3] = |area| */
  const int* bedges;        /* [nbnode*2] (interior node, phantom node) */
  const double* bedge_a;    /* [nbnode*4] outward unit normal, a[3] = |area| */
  const int* bfactag;       /* [nbnode] face tag of each half-edge */
  const double* xyz;        /* [nnode*3] */
  const double* vol;        /* [nnode] */
  int nfactag;              /* face tags are 0..nfactag-1 */
  const int* bc_type;       /* [nfactag] BCTypes ids (ucs/bc_defines.h:4-30) */
  const double* qref;       /* [nfactag*5] conservative reference state per tag (BCObj Qref) */
  const double* twall;      /* [nfactag] non-dimensional wall temperature of noSlip tags, <0 adiabatic
                               (BCObj twall / ref_temperature, ucs/bc.tcc:1284-1290); NULL: adiabatic */
} pcfd_mesh_desc;

/* Solver parameters (ucs/param.tcc keywords) */
typedef struct {
  double gamma;
  int sorder;            /* spatialOrder 1|2|3 */
  int nfirst;            /* firstOrderSteps */
  double chi;            /* 0 for sorder 2, 0.5 for sorder 3 (ucs/param.tcc:410-417) */
  int limiter;           /* 0 off, 1 Barth, 2 Venkat, 3 modified Venkat */
  double venkat_K;       /* modified Venkat K (reference hard-codes 1) */
  int limiter_refresh;   /* limiterRefresh */
  int limiter_freeze;    /* limiterFreeze */
  double cfl, cfl_start; /* CFL, startingCFL */
  int cfl_ramp;          /* rampCFL */
  int nsgs;              /* numberSGS (0 -> explicit) */
  int symmetry2d;        /* -1 off */
  int jac_freq;          /* jacobianUpdateFrequency */
  int precision;         /* PCFD_PLAIN_FP64 | PCFD_CHOSEN_FP32 */
  int gradient_terms;    /* 9: full reference set; 5: conservative only (inviscid fast path) */
  int viscous;           /* 1: laminar Navier-Stokes terms (CompressibleNS, Laminar model, mut = 0) */
  double re_tilde;       /* Re / Mach (ucs/compressible.tcc:753-755) */
  double prandtl;        /* prandtlNumber (0.72) */
  double prandtl_t;      /* PrT (0.85, ucs/param.tcc:89) */
  double sutherland_s;   /* 110.4 K / ref_temperature (ucs/eqnset.h:251-268) */
  /* Options that are off in the reference's defaults; 0 selects the default. */
  int gradient_type;     /* gradientType: 0 weighted LSQ, 1 Green-Gauss (ucs/gradient.tcc:56-112, 172-250) */
  int global_dt;         /* 1: localTimeStepping off -- one dt everywhere: timeStep if fixed_dt > 0, else the
                            minimum over nodes (ucs/timestep.tcc:48-74) */
  double fixed_dt;       /* timeStep (non-dimensional); used with global_dt when > 0 */
  int enable_vnn;        /* enableVNN: dt = min(dt, VNN vol^(2/3)) (ucs/timestep.tcc:37-45) */
  double vnn;            /* VNN (reference default 20) */
  int dynamic_cfl;       /* dynamicCFL: CFL changes by +-10% per iteration with the residual trend
                            (ucs/param.tcc:697-713, ucs/solutionSpace.tcc:629-633) */
  int jac_field_type;    /* jacobianFieldType: 0 one-sided FD (h = 1e-8), 1 centred FD (ucs/jacobian.tcc:307-367);
                            2 (complex step) gives PCFD_ERR_UNSUPPORTED */
  int flux_type;         /* PCFD_FLUX_ROE (roeFlux, the default) or PCFD_FLUX_ALGEBRAIC (algebraicFlux,
                            ucs/eqnset.tcc:56-77); 2 (HLLC) gives PCFD_ERR_UNSUPPORTED */
  int jac_boundary_type; /* jacobianBoundaryType: 0 one-sided FD, 1 centred FD (Bkernel_NumJac_Centered,
                            ucs/jacobian.tcc:538-640, with jacobianBoundaryEval 0: the +h evaluation
                            re-evaluates the boundary state, the -h one keeps the unperturbed one);
                            2 (complex step) gives PCFD_ERR_UNSUPPORTED */
  int rk_stages;         /* explicit path (nsgs 0): 0 or 1 -- the reference's single-stage ExplicitSolve
                            (ucs/solve.tcc:70-140, with NewtonIterate's second application of the
                            correction); 4 -- Jameson's four stages, Q_k = Q_0 + a_k dt/vol b(Q_{k-1}) with
                            a = 1/4, 1/3, 1/2, 1, boundary states (and gradient, limiter) refreshed per
                            stage, dt of the step (no counterpart in the reference) */
} pcfd_params;

#define PCFD_FLUX_ROE 0
#define PCFD_FLUX_ALGEBRAIC 1

const char* pcfd_last_error(const pcfd_ctx* ctx);
const char* pcfd_version(void);

/* Context lifetime.  Selects the current HIP device; builds the node-gather
 * CSR, the half-edge lists, the block-CRS pattern and the node colouring. */
int pcfd_create(const pcfd_mesh_desc* mesh, pcfd_ctx** out);
int pcfd_destroy(pcfd_ctx* ctx);
int pcfd_set_params(pcfd_ctx* ctx, const pcfd_params* params);
int pcfd_info(const pcfd_ctx* ctx, int* ncolor, long long* nblocks, long long* device_bytes);

/* State transfer in the reference layout q[node*nvars + k] (AoS, nvars >= 5,
 * conservative variables first).  Covers interior + phantom nodes.  On
 * download with nvars >= 10 the auxiliary T,P,u,v,w are filled as
 * ComputeAuxiliaryVariables does (ucs/compressible.tcc:1229-1243). */
int pcfd_upload_q(pcfd_ctx* ctx, const double* q, int nvars);
int pcfd_download_q(pcfd_ctx* ctx, double* q, int nvars);

/* Phases.  Optional host outputs (NULL = keep on device) use the reference
 * layouts: qgrad[node*27 + term*3 + dir], l[node*5+k], b[node*5+k], dt[node]. */
int pcfd_update_bcs(pcfd_ctx* ctx);
int pcfd_gradient(pcfd_ctx* ctx, double* qgrad_out);
int pcfd_limiter(pcfd_ctx* ctx, double* l_out);
int pcfd_set_limiter(pcfd_ctx* ctx, const double* l);  /* l[node*5+k] */
int pcfd_residual(pcfd_ctx* ctx, int iter, double* b_out, double norms[6]);
int pcfd_timestep(pcfd_ctx* ctx, double cfl, double* dt_out, double* dtmin);
int pcfd_set_timestep(pcfd_ctx* ctx, const double* dt);
/* ComputeJacobians (spatial FD Jacobian + temporal term with the current dt).
 * Like the reference's Bkernel_NumJac it re-evaluates the boundary states,
 * so phantom (and Dirichlet interior) states change. */
int pcfd_jacobian(pcfd_ctx* ctx);
/* Download Jacobian blocks for an arbitrary block-CRS pattern (ia[nnode+1],
 * ja[nblocks]); M[blk*25 + r*5 + c].  Valid after pcfd_jacobian and before
 * pcfd_prepare_sgs (which factorises the diagonal blocks in place). */
int pcfd_download_jacobian(pcfd_ctx* ctx, const int* ia, const int* ja, double* M);
int pcfd_prepare_sgs(pcfd_ctx* ctx);
/* nSgs symmetric multicolour sweeps from x0 = 0 on b; x_out[node*5+k].
 * *delta_xnorm gets |xnorm - xnorm_prev| of the last iteration. */
int pcfd_sgs(pcfd_ctx* ctx, int nsgs, double* x_out, double* delta_xnorm);
/* colour-sorted forward sweep order (a permutation of 0..nnode-1) */
int pcfd_sgs_order(pcfd_ctx* ctx, int* order);
/* NaN/Inf rows of x zeroed, q += x with ApplyDQ clipping. */
int pcfd_apply_dq(pcfd_ctx* ctx, int* nzeroed);
/* ExplicitSolve (x = b dt/vol, ApplyDQ) followed by the NewtonIterate update. */
int pcfd_explicit_update(pcfd_ctx* ctx, int* nzeroed);

/* One outer iteration = PreTimeAdvance + NewtonIterate + iter++, reference
 * order (ucs/solutionSpace.tcc:572-904).  resid[6] = residGlobal + per-eqn. */
int pcfd_iterate(pcfd_ctx* ctx, double resid[6]);
int pcfd_reset_iteration(pcfd_ctx* ctx);  /* iter = 1, limiter = 1, PreIterate dt */
int pcfd_get_iteration(const pcfd_ctx* ctx);

/* Steady solve: iterate until residGlobal <= drop * residGlobal(iter 1) or
 * max_iter.  history[max_iter] (nullable) receives residGlobal per iteration.
 * Returns the number of iterations done in *niter and wall seconds in *secs. */
int pcfd_solve(pcfd_ctx* ctx, int max_iter, double drop, double* history, int* niter, double* secs);

/* Per-phase device time (ms) accumulated since the last reset, measured with
 * HIP events on the context stream: [0] bcs [1] gradient [2] limiter
 * [3] residual [4] timestep [5] jacobian [6] prepare_sgs [7] sgs [8] update,
 * and kernel-only times inside phases: [9] SGS colour sweeps (no packing or
 * norms) [10] residual flux kernel (no norm reduction) [11] Jacobian fill
 * kernels (boundary part, rows, Dirichlet rows). */
#define PCFD_NPHASE 12
int pcfd_phase_times(pcfd_ctx* ctx, double ms[PCFD_NPHASE], int reset);
int pcfd_enable_phase_timing(pcfd_ctx* ctx, int on);
int pcfd_synchronize(pcfd_ctx* ctx);

#ifdef __cplusplus
}
#endif
#endif
