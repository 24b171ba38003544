/*
 * CPU oracle (test infrastructure only -- never linked into the product).
 *
 * A clean-room FP64 restatement of the reference ucs.x implicit edge-loop
 * hot path in the reference's own layouts: AoS state q[node*10 + k]
 * (5 conservative + 5 auxiliary T,P,u,v,w), gradients
 * qgrad[node*27 + term*3 + dir], limiter l[node*5+k], residual b[node*5+k],
 * block-CRS M[blk*25 + r*5 + c] with the diagonal first in each row, and the
 * reference's edge-ordered scatter.  Each function cites the reference
 * file:line it follows.  It is the parity checker for the HIP path and the
 * CPU baseline timed by bench.py.  It must not be used by the product path.
 */
#ifndef PCFD_ORACLE_H
#define PCFD_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

#define O_NEQN 5
#define O_NVARS 10
#define O_NTERMS 9

/* BC type ids: ucs/bc_defines.h:4-30 */
enum {
  O_BC_PARALLEL = 0, O_BC_DIRICHLET = 1, O_BC_NEUMANN = 2, O_BC_IMPERMEABLE = 3, O_BC_NOSLIP = 4,
  O_BC_FARFIELD_VISCOUS = 5, O_BC_FARFIELD = 6, O_BC_SONIC_INFLOW = 7, O_BC_SONIC_OUTFLOW = 8,
  O_BC_SYMMETRY = 9
};

typedef struct {
  int nnode, nbnode, nedge;
  const int* edges;      /* [nedge*2] */
  const double* edge_a;  /* [nedge*4] */
  const int* bedges;     /* [nbnode*2] (node, phantom) */
  const double* bedge_a; /* [nbnode*4] */
  const int* bfactag;    /* [nbnode] */
  const double* xyz;     /* [nnode*3] */
  const double* vol;     /* [nnode] */
  int nfactag;
  const int* bc_type;    /* [nfactag] */
  const double* qref;    /* [nfactag*10] reference state (with aux vars) */
  const double* twall;   /* [nfactag] non-dimensional wall temperature of noSlip tags (<0: adiabatic); may be NULL */
  const int* bnormal;    /* [nbnode] most wall-normal neighbour of a noSlip half-edge's node (o_wall_normal_nodes) */
} o_mesh;

typedef struct {
  double gamma;
  int sorder, nfirst;           /* spatialOrder, firstOrderSteps */
  double chi;
  int limiter;                  /* 0 off, 1 Barth, 2 Venkat, 3 modified Venkat */
  double venkat_K;              /* K of the modified Venkat limiter (reference: 1) */
  int limiter_refresh, limiter_freeze;
  double cfl, cfl_start;
  int cfl_ramp;
  int nsgs;                     /* 0 -> explicit */
  int symmetry2d;
  int jac_freq;
  int clip_jacobi;              /* 0: reference sequential pressure clip; 1: all edges see the pre-clip limiter */
  int viscous;                  /* laminar Navier-Stokes terms (CompressibleNS, mut = 0) */
  double re_tilde;              /* Re / Mach (ReTilde, ucs/compressible.tcc:753-755) */
  double pr, prt;               /* Prandtl and turbulent Prandtl numbers */
  double sutherland_s;          /* 110.4 K / T_ref (ucs/eqnset.h:251-268) */
  int gradient_type;            /* gradientType: 0 weighted LSQ, 1 Green-Gauss */
  int global_dt;                /* 1: localTimeStepping off */
  double fixed_dt;              /* timeStep (> 0: fixed dt with global_dt) */
  int enable_vnn;               /* enableVNN */
  double vnn;                   /* VNN */
  int dynamic_cfl;              /* dynamicCFL */
  int jac_field_type;           /* jacobianFieldType: 0 one-sided FD, 1 centred FD */
  int flux_type;                /* 0 Roe, 1 algebraic */
  int jac_boundary_type;        /* jacobianBoundaryType: 0 one-sided FD, 1 centred FD */
  int rk_stages;                /* explicit path: 0/1 the reference's single stage, 4 Jameson's four stages */
} o_params;

typedef struct {
  int nnode, nblocks;
  int* ia;   /* [nnode+1] */
  int* ja;   /* [nblocks] */
  int* iau;  /* [nnode] */
  int* pv;   /* [nnode*5] */
  double* M; /* [nblocks*25] */
  int* e_lr; /* per edge: block index of (n0 row, n1 col) */
  int* e_rl; /* per edge: block index of (n1 row, n0 col) */
} o_crs;

void o_aux(double* Q, double gamma);
int o_roe_flux(const double* QL, const double* QR, const double* avec, double vdotn, double gamma,
               int symmetry2d, double* flux);
void o_bc_state(int bctype, double* QL, double* QR, const double* Qref, const double* avec, double gamma);
/* as o_bc_state, plus noSlip walls: twall (non-dimensional, <0 adiabatic) and
 * the state of the wall-normal neighbour (adiabatic walls only) */
void o_bc_state_ex(int bctype, double* QL, double* QR, const double* Qref, const double* avec, double gamma,
                   double twall, const double* normalQ);
/* wall-normal neighbour of each noSlip half-edge's node (-1 elsewhere):
 * ucs/bc.tcc:1182-1210 */
void o_wall_normal_nodes(const o_mesh* m, int* bnormal);
/* CompressibleEqnSet::ViscousFlux (ucs/compressible.tcc:712-795), mut = 0 */
void o_viscous_flux(const o_params* p, const double* Q, const double* grad, const double* avec, double* flux);
/* CompressibleEqnSet::ViscousJacobian (ucs/compressible.tcc:1632-1893), mut = 0 */
void o_viscous_jacobian(const o_params* p, const double* QL, const double* QR, const double* dx, double s2,
                        const double* avec, double* aL, double* aR);

void o_update_bcs(const o_mesh* m, const o_params* p, double* q);
void o_lsq_moments(const o_mesh* m, double* sw);
void o_gradient(const o_mesh* m, const double* q, const double* sw, double* grad);
void o_gradient_gg(const o_mesh* m, const double* q, double* grad);
void o_limiter(const o_mesh* m, const o_params* p, const double* q, const double* grad, double* l);
void o_residual(const o_mesh* m, const o_params* p, int iter, const double* q, const double* grad,
                const double* l, double* b, double* norms /* [6] */);
double o_timestep(const o_mesh* m, const o_params* p, const double* q, double cfl, double* dt);
double o_cfl(const o_params* p, int iter);
double o_cfl_update(const o_params* p, int iter, double res, double res_nm1, double cfl_prev);

int o_crs_init(const o_mesh* m, o_crs* A);
void o_crs_free(o_crs* A);
void o_jacobian(const o_mesh* m, const o_params* p, double* q, const double* dt, o_crs* A);
void o_prepare_sgs(o_crs* A);
double o_sgs(const o_crs* A, int nsgs, const double* b, double* x, const int* order);
int o_lu(double* a, int* piv, int n);
void o_lu_solve(const double* a, double* b, const int* piv, double* x, int n);
void o_apply_dq(const double* dq, double* Q, double gamma);

/* Finite-rate chemistry model (ucs/chem.tcc, ucs/reaction.tcc,
 * ucs/species.tcc): per reaction the local species list in the reaction
 * parser's order with stoichiometric coefficients and third-body
 * efficiencies. */
#define O_CHEM_MAXS 8
#define O_CHEM_MAXR 16
typedef struct {
  int ns, nr;
  double mw[O_CHEM_MAXS];                       /* kg/mol */
  double lo[O_CHEM_MAXS][7], hi[O_CHEM_MAXS][7];  /* NASA-7, T <= 1000 K / T > 1000 K */
  int type[O_CHEM_MAXR];                        /* 0 Arrhenius 1 ModArrhenius 2 GuptaModArrhenius 3 Power */
  int tb[O_CHEM_MAXR];                          /* third bodies present */
  double A[O_CHEM_MAXR], EA[O_CHEM_MAXR], n[O_CHEM_MAXR];
  int nloc[O_CHEM_MAXR];
  int loc[O_CHEM_MAXR][O_CHEM_MAXS];            /* global species of local slot */
  double nup[O_CHEM_MAXR][O_CHEM_MAXS], nupp[O_CHEM_MAXR][O_CHEM_MAXS], eff[O_CHEM_MAXR][O_CHEM_MAXS];
  double ref_density, ref_time, ref_temperature;
} o_chem;

double o_chem_kf(const o_chem* c, int r, double T);
double o_chem_kc(const o_chem* c, int r, double T);
/* GetMassProductionRates: rhoi (kg/m^3), T (K) -> wdot (kg/m^3/s) */
void o_chem_wdot(const o_chem* c, const double* rhoi, double T, double* wdot);
/* SourceTerm on the FR state (rho_i.., u, v, w, T; non-dimensional) */
void o_chem_source(const o_chem* c, const double* Q, double vol, double* source);
/* SourceTermJacobian by one-sided differences, h = 1e-8 */
void o_chem_source_jacobian(const o_chem* c, const double* Q, double vol, double* A);
double o_chem_cp(const o_chem* c, int s, double T);
void o_fr_aux(const o_chem* c, const double* Q, double ref_pressure, double ref_velocity, double* aux);
double o_fr_t_from_p(const o_chem* c, const double* rhoi, double p, double Tinit, double ref_pressure, int* iters);

/* Finite-rate (FR) equation set, first-order implicit path (oracle_fr.c).
 * State per node: rho_1..rho_ns, u, v, w, T (neqn = ns + 4), then aux
 * P, rho, cv_1..cv_ns, conc_1..conc_ns (nvars = 3 ns + 6). */
typedef struct {
  const o_chem* chem;  /* with ref_density, ref_time, ref_temperature set */
  double ref_pressure, ref_velocity;
  double Pref;         /* gauge reference: the freestream pressure (UpdateQinf) */
  double beta;         /* preconditioning beta (1 at and above Mach 1) */
  int sym2d;           /* symmetry2D */
  int rxn_on;          /* reactionsOn */
  int local_dt;        /* localTimeStepping */
  double dt_param;     /* timeStep parameter (default -1) */
  int no_cvbc;         /* noCVBC: upwind farfield and mirror walls instead of the characteristic states */
} o_fr;

typedef struct {
  int nnode, nblocks, neqn;
  int *ia, *ja, *iau, *e_lr, *e_rl, *pv;
  double* M; /* [nblocks*neqn*neqn] */
  int ludiag; /* diagonal blocks hold their LU factors (CRSMatrix::ludiag) */
} o_fr_crs;

typedef struct {
  const o_mesh* m;
  o_fr f;
  double* q;    /* (nnode+nbnode)*nvars */
  double* qref; /* nfactag*nvars */
  double *b, *x, *dt;
  o_fr_crs A;
  int iter, nsgs, jac_freq, cfl_ramp;
  double cfl, cfl_start, cfl_end;
  const int* order;
} o_fr_solver;

int o_fr_neqn(const o_fr* f);
int o_fr_nvars(const o_fr* f);
void o_fr_aux_q(const o_fr* f, double* Q);
double o_fr_total_enthalpy(const o_fr* f, const double* Q);
double o_fr_total_energy(const o_fr* f, const double* Q);
int o_fr_hllc(const o_fr* f, const double* QL, const double* QR, const double* avec, double beta, double* flux);
double o_fr_max_eig(const o_fr* f, const double* Q, const double* avec, double beta);
void o_fr_eigensystem(const o_fr* f, const double* Q, const double* avec, double beta, double* eig, double* T,
                      double* Tinv);
void o_fr_bc_state(const o_fr* f, int bctype, double* QL, double* QR, const double* Qref, const double* avec,
                   double beta);
void o_fr_apply_dq(const o_fr* f, const double* dQ, double* Q);
void o_fr_temporal(const o_fr* f, const double* Q, double vol, double dtau, double beta, double* A);
void o_fr_qinf(o_fr* f, const double* Y, double pDim, double TDim, const double* vel, double* Qinf);
void o_fr_update_bcs(const o_fr* f, const o_mesh* m, const double* qref, double* q);
double o_fr_residual(const o_fr* f, const o_mesh* m, const double* q, double* b, double* norms);
double o_fr_timestep(const o_fr* f, const o_mesh* m, const double* q, double cfl, double* dt);
int o_fr_crs_init(const o_mesh* m, int neqn, o_fr_crs* A);
void o_fr_crs_free(o_fr_crs* A);
void o_fr_jacobian(const o_fr* f, const o_mesh* m, const double* qref, double* q, const double* dt, o_fr_crs* A);
double o_fr_sgs(o_fr_crs* A, int nsgs, const double* b, double* x, const int* order);
int o_fr_solver_init(o_fr_solver* s, const o_mesh* m, const o_fr* f, const double* qref, const double* q0,
                     double cfl, double cfl_start, int cfl_ramp, int nsgs, int jac_freq);
double o_fr_step(o_fr_solver* s, double* norms, int* nzeroed);
void o_fr_solver_free(o_fr_solver* s);

/* Full solver state for the steady Newton / explicit loop. */
typedef struct {
  const o_mesh* m;
  o_params p;
  double* q;     /* (nnode+nbnode)*10 */
  double* grad;  /* nnode*27 */
  double* l;     /* nnode*5 */
  double* sw;    /* nnode*6 */
  double* b;     /* nnode*5 */
  double* x;     /* nnode*5 */
  double* dt;    /* nnode */
  o_crs A;
  int iter;
  double cfl;
  const int* order; /* SGS row order (NULL = natural) */
  double t_jac, t_grad, t_res, t_lin, t_upd; /* phase timers (s) */
  double res, res_nm1; /* residGlobal of the last two iterations (dynamic CFL) */
} o_solver;

int o_solver_init(o_solver* s, const o_mesh* m, const o_params* p, const double* q0);
/* one outer iteration: PreTimeAdvance + NewtonIterate + iter++ ; returns residGlobal */
double o_solver_step(o_solver* s, double* norms, int* nzeroed);
void o_solver_free(o_solver* s);

#ifdef __cplusplus
}
#endif
#endif
