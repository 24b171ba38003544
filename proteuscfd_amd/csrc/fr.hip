// Finite-rate (FR) reacting Euler, first-order implicit path, on one GPU:
// the C ABI of include/pcfd_fr.h, its device layout and its kernels.
//
// Layout in HBM (ns = NS species, NE = NS + 4):
//   q      [ntot][NE]   native state per node, interior nodes then one
//                       phantom per half-edge (AoS: an edge kernel's two
//                       endpoint loads are two runs of NE*8 bytes)
//   eflux  [nedge][NE]  HLLC flux of every edge, n0 -> n1 orientation
//   elam   [nedge]      max eigenvalue x |a| of every edge (time step)
//   b      [nnode][NE]  residual;  dt [nnode]
//   node lists (CSR in edge order): nptr [nnode+1], nslot [2 nedge] =
//                       2 e + side (side 0: the node is n0 of edge e)
//   block matrix (SELL, rows grouped by colour, 64-row chunks): off-diagonal
//                       block of row r slot s at ((sbase[chunk]+s)*NE2 + k)*64
//                       + lane, k = 0..NE2-1 (row-major entry); diagonal at
//                       (chunk*NE2 + k)*64 + lane; a row's slots follow its
//                       CSR (edge) order.  h_eoff[2e], h_eoff[2e+1] (host):
//                       where edge e's blocks (n0,n1) and (n1,n0) start.
//
// Every flux and residual sum runs in the reference's order (edges in edge
// order, then half-edges, then the source), as node gathers with no atomics,
// and FP contraction is off in this unit: the FP64 results are the
// reference's arithmetic.
#pragma clang fp contract(off)

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/pcfd_fr.h"
#include "../../include/pcfd_mesh.h"
#include "chem_model.h"
#include "fr_physics.h"

namespace pcfd {
const ChemDev* chem_device_model(const pcfd_chem* c);
int chem_nspecies(const pcfd_chem* c);
// the production-rate surrogate with exact fallback (kernels_mlp.hip): runs on
// the surrogate's stream and returns with it synchronised
int surrogate_launch(pcfd_chem* chem, pcfd_mlp& m, const pcfd_surrogate_desc* d, int n, const double* rhoi,
                     const double* T, double* wdot, int* lst, int* cnt, int* off, int* nfallback, double* ms);
}  // namespace pcfd

using namespace pcfd;
using fr::Model;
using fr::State;

namespace {

constexpr int kFB = 128;  // threads per block for the FR kernels
constexpr int kLane = 64;

thread_local std::string g_fr_create_err;

void hcheck(hipError_t e, const char* what) {
  if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

inline int nblocks(long long n, int bs = kFB) { return (int)((n + bs - 1) / bs); }

template <int NS>
__device__ __forceinline__ void load_state(const double* __restrict__ q, long long i, State<NS>& s,
                                           const Model& m) {
  constexpr int NE = NS + 4;
#pragma unroll
  for (int k = 0; k < NE; ++k) s.q[k] = q[i * NE + k];
  fr::aux(m, s);
}

template <int NS>
__device__ __forceinline__ void store_state(double* __restrict__ q, long long i, const State<NS>& s) {
  constexpr int NE = NS + 4;
#pragma unroll
  for (int k = 0; k < NE; ++k) q[i * NE + k] = s.q[k];
}

// ---- boundary states (UpdateBCs, ucs/bc.tcc:1399-1457) ---------------------
// One thread per boundary node: its half-edges in half-edge order (a
// Dirichlet / sonic-inflow half-edge resets the node's own state, which the
// node's later half-edges then see, as in the reference's half-edge loop).
template <int NS>
__global__ void __launch_bounds__(kFB) k_fr_bcs(const Model* __restrict__ mp, int nbnd, const int* __restrict__ bnd,
                                               const int* __restrict__ he_ptr, const int* __restrict__ he_id,
                                               const int* __restrict__ he_tag, const double* __restrict__ he_a,
                                               const int* __restrict__ bc_type, const double* __restrict__ qref,
                                               int nnode, double beta, double* __restrict__ q) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nbnd) return;
  const Model& m = *mp;
  const int i = bnd[t];
  State<NS> L;
  load_state<NS>(q, i, L, m);
  bool lchanged = false;
  for (int p = he_ptr[i]; p < he_ptr[i + 1]; ++p) {
    const int h = he_id[p], tag = he_tag[h];
    const int ty = bc_type[tag];
    State<NS> R, Qr;
    load_state<NS>(q, (long long)nnode + h, R, m);
    load_state<NS>(qref, tag, Qr, m);
    double a[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) a[k] = he_a[4LL * h + k];
    if (fr::bc_state<NS>(m, ty, L, R, Qr, a, beta)) {
      store_state<NS>(q, (long long)nnode + h, R);
      if (ty == fr::kSonicInflow || ty == fr::kDirichlet) lchanged = true;
    }
  }
  if (lchanged) store_state<NS>(q, i, L);
}

// ---- edge passes -----------------------------------------------------------
// HLLC flux of every edge (Kernel_Inviscid_Flux at first order,
// ucs/residual.tcc:191-296), n0 -> n1.
template <int NS>
__global__ void __launch_bounds__(kFB) k_fr_edge_flux(const Model* __restrict__ mp, int nedge,
                                                     const int* __restrict__ edges, const double* __restrict__ ea,
                                                     const double* __restrict__ q, double beta,
                                                     double* __restrict__ eflux) {
  constexpr int NE = NS + 4;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nedge) return;
  const Model& m = *mp;
  State<NS> L, R;
  load_state<NS>(q, edges[2 * e], L, m);
  load_state<NS>(q, edges[2 * e + 1], R, m);
  double a[4], F[NE];
#pragma unroll
  for (int k = 0; k < 4; ++k) a[k] = ea[4LL * e + k];
  fr::hllc<NS>(m, L, R, a, beta, F);
#pragma unroll
  for (int k = 0; k < NE; ++k) eflux[(long long)e * NE + k] = F[k];
}

// max eigenvalue of the edge-averaged state times |a| (Kernel_Timestep,
// ucs/timestep.tcc:72-108)
template <int NS>
__global__ void __launch_bounds__(kFB) k_fr_edge_lam(const Model* __restrict__ mp, int nedge,
                                                    const int* __restrict__ edges, const double* __restrict__ ea,
                                                    const double* __restrict__ q, double beta,
                                                    double* __restrict__ elam) {
  constexpr int NE = NS + 4;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nedge) return;
  const Model& m = *mp;
  const long long L = edges[2 * e], R = edges[2 * e + 1];
  State<NS> s;
#pragma unroll
  for (int k = 0; k < NE; ++k) s.q[k] = 0.5 * (q[L * NE + k] + q[R * NE + k]);
  fr::aux(m, s);
  double a[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) a[k] = ea[4LL * e + k];
  elam[e] = fr::max_eig<NS>(m, s, a, beta) * a[3];
}

// Per-block partial sums of b^2 (total and per equation) with a fixed tree
template <int NE>
__device__ __forceinline__ void norm_partials(const double* bi, bool valid, double* red) {
  __shared__ double sh[NE + 1][kFB];
  double tot = 0.0;
#pragma unroll
  for (int k = 0; k < NE; ++k) {
    const double v = valid ? bi[k] * bi[k] : 0.0;
    sh[1 + k][threadIdx.x] = v;
    tot += v;
  }
  sh[0][threadIdx.x] = tot;
  __syncthreads();
  for (int s = kFB / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
#pragma unroll
      for (int k = 0; k <= NE; ++k) sh[k][threadIdx.x] += sh[k][threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k <= NE; ++k) red[(long long)k * gridDim.x + blockIdx.x] = sh[k][0];
  }
}

// Residual of every node: edge fluxes in edge order (b[n0] -= F, b[n1] += F),
// then the node's boundary fluxes (Bkernel_Inviscid_Flux, ucs/residual.tcc:
// 298-387: b -= F of the node against its phantom), then the source term
// (SpatialResidual, :100-114)
template <int NS>
__global__ void __launch_bounds__(kFB) k_fr_residual(const Model* __restrict__ mp, const ChemDev* __restrict__ cm,
                                                    int nnode, const int* __restrict__ nptr,
                                                    const int* __restrict__ nslot, const double* __restrict__ eflux,
                                                    const int* __restrict__ he_ptr, const int* __restrict__ he_id,
                                                    const double* __restrict__ he_a, const double* __restrict__ q,
                                                    const double* __restrict__ vol, int rxn,
                                                    const double* __restrict__ wrates, double beta,
                                                    double* __restrict__ b, double* __restrict__ red) {
  constexpr int NE = NS + 4;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const bool valid = i < nnode;
  double acc[NE];
#pragma unroll
  for (int k = 0; k < NE; ++k) acc[k] = 0.0;
  if (valid) {
    const Model& m = *mp;
    for (int p = nptr[i]; p < nptr[i + 1]; ++p) {
      const int s = nslot[p];
      const double* F = eflux + (long long)(s >> 1) * NE;
      if (s & 1) {
#pragma unroll
        for (int k = 0; k < NE; ++k) acc[k] += F[k];
      } else {
#pragma unroll
        for (int k = 0; k < NE; ++k) acc[k] += -F[k];
      }
    }
    if (he_ptr[i + 1] > he_ptr[i]) {
      State<NS> L;
      load_state<NS>(q, i, L, m);
      for (int p = he_ptr[i]; p < he_ptr[i + 1]; ++p) {
        const int h = he_id[p];
        State<NS> R;
        load_state<NS>(q, (long long)nnode + h, R, m);
        double a[4], F[NE];
#pragma unroll
        for (int k = 0; k < 4; ++k) a[k] = he_a[4LL * h + k];
        fr::hllc<NS>(m, L, R, a, beta, F);
#pragma unroll
        for (int k = 0; k < NE; ++k) acc[k] += -F[k];
      }
    }
    // the source vector is added whole (zero without reactions, and zero in
    // the momentum and energy entries), as SpatialResidual does
    double S[NS];
    if (rxn && wrates) {
      // production rates from the surrogate (kg/m^3/s), scaled as SourceTerm does
#pragma unroll
      for (int k = 0; k < NS; ++k) {
        double w = wrates[(long long)i * NS + k];
        w /= (cm->ref_density / cm->ref_time);
        S[k] = vol[i] * w;
      }
    } else if (rxn) {
      double Q[NE];
#pragma unroll
      for (int k = 0; k < NE; ++k) Q[k] = q[(long long)i * NE + k];
      source_of(*cm, Q, vol[i], S);
    } else {
#pragma unroll
      for (int k = 0; k < NS; ++k) S[k] = 0.0;
    }
#pragma unroll
    for (int k = 0; k < NS; ++k) acc[k] += S[k];
#pragma unroll
    for (int k = NS; k < NE; ++k) acc[k] += 0.0;
#pragma unroll
    for (int k = 0; k < NE; ++k) b[(long long)i * NE + k] = acc[k];
  }
  norm_partials<NE>(acc, valid, red);
}

// Dimensional species densities and temperature of every node, the
// production-rate surrogate's inputs (SourceTerm's scaling,
// ucs/compressibleFR.tcc:1284-1290)
template <int NS>
__global__ void __launch_bounds__(kFB) k_fr_chem_inputs(const ChemDev* __restrict__ cm, int nnode,
                                                       const double* __restrict__ q, double* __restrict__ rhoi,
                                                       double* __restrict__ T) {
  constexpr int NE = NS + 4;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nnode) return;
#pragma unroll
  for (int k = 0; k < NS; ++k) rhoi[(long long)i * NS + k] = q[(long long)i * NE + k] * cm->ref_density;
  T[i] = q[(long long)i * NE + NS + 3] * cm->ref_temperature;
}

// Time step of every node (ComputeTimesteps with local time stepping,
// ucs/timestep.tcc:6-40): edge eigenvalues in edge order, then the node's
// half-edges against their phantoms; dt = CFL vol / sum
template <int NS>
__global__ void __launch_bounds__(kFB) k_fr_timestep(const Model* __restrict__ mp, int nnode,
                                                    const int* __restrict__ nptr, const int* __restrict__ nslot,
                                                    const double* __restrict__ elam, const int* __restrict__ he_ptr,
                                                    const int* __restrict__ he_id, const double* __restrict__ he_a,
                                                    const double* __restrict__ q, const double* __restrict__ vol,
                                                    double cfl, double beta, double* __restrict__ dt) {
  constexpr int NE = NS + 4;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nnode) return;
  const Model& m = *mp;
  double s = 0.0;
  for (int p = nptr[i]; p < nptr[i + 1]; ++p) s += elam[nslot[p] >> 1];
  for (int p = he_ptr[i]; p < he_ptr[i + 1]; ++p) {
    const int h = he_id[p];
    const long long R = (long long)nnode + h;
    State<NS> t;
#pragma unroll
    for (int k = 0; k < NE; ++k) t.q[k] = 0.5 * (q[(long long)i * NE + k] + q[R * NE + k]);
    fr::aux(m, t);
    double a[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) a[k] = he_a[4LL * h + k];
    s += fr::max_eig<NS>(m, t, a, beta) * a[3];
  }
  dt[i] = cfl * (vol[i] / s);
}

// ---- Jacobian --------------------------------------------------------------
// Kernel_NumJac (ucs/jacobian.tcc:253-304): one-sided differences (h = 1e-8)
// of the edge's HLLC flux in each native variable of either side, aux
// recomputed for every perturbed state.  Block (n0,n1) = dF/dq_n1, block
// (n1,n0) = -dF/dq_n0; their negations, the two diagonal contributions of
// Kernel_Diag_NumJac (:436-456; the n0 row's diagonal gets minus block
// (n1,n0)), are written to ediag, contiguously per edge side, and
// k_fr_jac_diag places the blocks in their SELL slots (scattered 8-byte
// stores from here cost more than the second read there).
// One thread per (edge, column): thread 9e + i perturbs variable i on
// either side (two fluxes) against the edge's unperturbed flux, computed
// once beforehand; nine threads share an edge's loads, at twice the
// occupancy of one thread per edge.
template <int NS, typename TM>
__global__ void __launch_bounds__(kFB) k_fr_jac_edge(const Model* __restrict__ mp, int nedge,
                                                    const int* __restrict__ edges, const double* __restrict__ ea,
                                                    const double* __restrict__ q, double beta,
                                                    const double* __restrict__ fs,
                                                    TM* __restrict__ ediag) {
  constexpr int NE = NS + 4;
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)nedge * NE) return;
  const int e = (int)(t / NE), i = (int)(t % NE);
  const Model& m = *mp;
  State<NS> L, R;
  load_state<NS>(q, edges[2 * e], L, m);
  load_state<NS>(q, edges[2 * e + 1], R, m);
  double a[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) a[k] = ea[4LL * e + k];
  // the unperturbed flux, evaluated once per edge by k_fr_edge_flux
  double fS[NE], fX[NE];
#pragma unroll
  for (int k = 0; k < NE; ++k) fS[k] = fs[(long long)e * NE + k];
  const double h = 1.0e-8;
  // in the storage precision: the FP32 form's diagonal sums FP32 terms
  TM* dL = ediag + (long long)e * 2 * NE * NE;
  TM* dR = dL + NE * NE;
  {
    State<NS> P = L;
    // P.q[i] += h without a dynamic register index
#pragma unroll
    for (int k = 0; k < NE; ++k) P.q[k] = (k == i) ? P.q[k] + h : P.q[k];
    fr::aux(m, P);
    fr::hllc<NS>(m, P, R, a, beta, fX);
#pragma unroll
    for (int j = 0; j < NE; ++j) {
      // added to a zeroed block, as the reference's scatter does; the
      // off-diagonal block A[R][L] is -dL (written by k_fr_jac_diag)
      const double tl = 0.0 + (fS[j] - fX[j]) / h;
      dL[j * NE + i] = (TM)(-tl);
    }
  }
  {
    State<NS> P = R;
#pragma unroll
    for (int k = 0; k < NE; ++k) P.q[k] = (k == i) ? P.q[k] + h : P.q[k];
    fr::aux(m, P);
    fr::hllc<NS>(m, L, P, a, beta, fX);
#pragma unroll
    for (int j = 0; j < NE; ++j) {
      const double tr = 0.0 + (fX[j] - fS[j]) / h;
      dR[j * NE + i] = (TM)(-tr);  // A[L][R] = -dR
    }
  }
}

// Bkernel_NumJac (ucs/jacobian.tcc:458-535) with jacobianBoundaryEval 0:
// per boundary node, its half-edges in order; the BC state is set on the
// stored states first (q updated, as the reference's in-place call does),
// then each native variable of the node perturbed, the BC re-evaluated from
// the perturbed state and the boundary flux differenced.  The node's
// boundary blocks are summed in half-edge order into jdb.
// Two passes: k_fr_jac_bnd_base walks each node's half-edges once, applies
// the BC updates to q and records, per half-edge, the node's and the
// boundary state's native values right after that half-edge's BC call and
// the unperturbed flux (bhe, 3 NE values); k_fr_jac_bnd_cols gives each
// (node, column) a thread that differences from those records and sums its
// column over the half-edges in their order.  A State's auxiliary values are
// always aux() of its native ones (load_state and bc_state recompute them),
// so the records carry every bit the perturbed evaluations read.
template <int NS>
__global__ void __launch_bounds__(kFB) k_fr_jac_bnd_base(const Model* __restrict__ mp, int nbnd,
                                                        const int* __restrict__ bnd, const int* __restrict__ he_ptr,
                                                        const int* __restrict__ he_id, const int* __restrict__ he_tag,
                                                        const double* __restrict__ he_a,
                                                        const int* __restrict__ bc_type,
                                                        const double* __restrict__ qref, int nnode, double beta,
                                                        double* __restrict__ q, double* __restrict__ bhe) {
  constexpr int NE = NS + 4;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nbnd) return;
  const Model& m = *mp;
  const int i = bnd[t];
  State<NS> L;
  load_state<NS>(q, i, L, m);
  bool lchanged = false;
  for (int p = he_ptr[i]; p < he_ptr[i + 1]; ++p) {
    const int h = he_id[p], tag = he_tag[h];
    const int ty = bc_type[tag];
    State<NS> R, Qr;
    load_state<NS>(q, (long long)nnode + h, R, m);
    load_state<NS>(qref, tag, Qr, m);
    double a[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) a[k] = he_a[4LL * h + k];
    if (fr::bc_state<NS>(m, ty, L, R, Qr, a, beta)) {
      store_state<NS>(q, (long long)nnode + h, R);
      if (ty == fr::kSonicInflow || ty == fr::kDirichlet) lchanged = true;
    }
    double fS[NE];
    fr::hllc<NS>(m, L, R, a, beta, fS);
    double* rec = bhe + (long long)p * 3 * NE;
#pragma unroll
    for (int k = 0; k < NE; ++k) {
      rec[k] = L.q[k];
      rec[NE + k] = R.q[k];
      rec[2 * NE + k] = fS[k];
    }
  }
  if (lchanged) store_state<NS>(q, i, L);
}

template <int NS>
__global__ void __launch_bounds__(kFB) k_fr_jac_bnd_cols(const Model* __restrict__ mp, int nbnd,
                                                        const int* __restrict__ bnd, const int* __restrict__ he_ptr,
                                                        const int* __restrict__ he_id, const int* __restrict__ he_tag,
                                                        const double* __restrict__ he_a,
                                                        const int* __restrict__ bc_type,
                                                        const double* __restrict__ qref, double beta,
                                                        const double* __restrict__ bhe, double* __restrict__ jdb) {
  constexpr int NE = NS + 4, NE2 = NE * NE;
  const long long t2 = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t2 >= (long long)nbnd * NE) return;
  const int t = (int)(t2 / NE), c = (int)(t2 % NE);
  const Model& m = *mp;
  const int i = bnd[t];
  const double hh = 1.0e-8;
  double acc[NE];
#pragma unroll
  for (int j = 0; j < NE; ++j) acc[j] = 0.0;
  for (int p = he_ptr[i]; p < he_ptr[i + 1]; ++p) {
    const int h = he_id[p], tag = he_tag[h];
    const int ty = bc_type[tag];
    const double* rec = bhe + (long long)p * 3 * NE;
    State<NS> PL, PR, Qr;
    load_state<NS>(rec, 0, PL, m);
    load_state<NS>(rec + NE, 0, PR, m);
    load_state<NS>(qref, tag, Qr, m);
    double a[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) a[k] = he_a[4LL * h + k];
    // PL.q[c] += hh without a dynamic register index
#pragma unroll
    for (int k = 0; k < NE; ++k) PL.q[k] = (k == c) ? PL.q[k] + hh : PL.q[k];
    fr::aux(m, PL);
    fr::bc_state<NS>(m, ty, PL, PR, Qr, a, beta);
    double fL[NE];
    fr::hllc<NS>(m, PL, PR, a, beta, fL);
#pragma unroll
    for (int j = 0; j < NE; ++j) acc[j] += (fL[j] - rec[2 * NE + j]) / hh;
  }
  double* J = jdb + (long long)t * NE2;
#pragma unroll
  for (int j = 0; j < NE; ++j) J[j * NE + c] = acc[j];
}

// Per node: the temporal block (ContributeTemporalTerms applied to a zero
// block, ucs/jacobian.tcc:213-250) and the source Jacobian's species rows
// (EqnSet::SourceTermJacobian, ucs/eqnset.tcc:162-187: forward differences,
// h = 1e-8; the velocity columns are exactly zero, the source not reading
// u, v, w).  nterm[i] = [temporal NE*NE | source NS x (NS+1), columns
// rho_1..rho_NS, T].
template <int NS>
__global__ void __launch_bounds__(kFB) k_fr_node_terms(const Model* __restrict__ mp, const ChemDev* __restrict__ cm,
                                                      int nnode, const double* __restrict__ q,
                                                      const double* __restrict__ vol, const double* __restrict__ dt,
                                                      int rxn, double beta, double* __restrict__ nterm) {
  constexpr int NE = NS + 4, NE2 = NE * NE, NT = NE2 + NS * (NS + 1);
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nnode) return;
  const Model& m = *mp;
  double* o = nterm + (long long)i * NT;
  State<NS> S;
  load_state<NS>(q, i, S, m);
  const double v = vol[i];
  double T[NE2];
#pragma unroll
  for (int k = 0; k < NE2; ++k) T[k] = 0.0;
  fr::temporal<NS>(m, S, v, dt[i], beta, T);
#pragma unroll
  for (int k = 0; k < NE2; ++k) o[k] = T[k];
  if (!rxn) return;
  double sc[(NS + 1) * NS], s1[NS], Q[NE];
#pragma unroll
  for (int k = 0; k < NE; ++k) Q[k] = S.q[k];
  const double h = 1.0e-8;
  // the state and its species columns (Q[c] + h) share the T-only rate
  // constants: the same values as NS + 1 separate source evaluations
  source_rho_columns(*cm, Q, h, v, sc);
  const double* s0 = sc;
  for (int c = 0; c < NS; ++c)
    for (int j = 0; j < NS; ++j) o[NE2 + j * (NS + 1) + c] = (sc[(1 + c) * NS + j] - s0[j]) / h;
  // the T column
  Q[NS + 3] += h;
  source_of(*cm, Q, v, s1);
  for (int j = 0; j < NS; ++j) o[NE2 + j * (NS + 1) + NS] = (s1[j] - s0[j]) / h;
}

// whether ContributeTemporalTerms adds to entry (r, c) of the diagonal block
template <int NS>
__device__ __forceinline__ bool temporal_entry(int r, int c) {
  constexpr int tloc = NS + 3;
  if (r < NS) return c == r || c == tloc;
  if (r < NS + 3) return c < NS || c == r || c == tloc;
  return true;
}

// Diagonal block of every row: the boundary part, then minus the transposed
// off-diagonal blocks in edge order (Kernel_Diag_NumJac, ucs/jacobian.tcc:
// 436-456), minus the source Jacobian (:198-209), plus the temporal terms
// (:213-250); and the row's off-diagonal blocks from the other sides of
// its edges.  A workgroup takes one 64-row SELL chunk; wave w builds block
// row w of those 64 rows (so every store is 64 consecutive rows' entry).
template <int NS, typename TM>
__global__ void __launch_bounds__(64 * (NS + 4)) k_fr_jac_diag(int nrow_pad, const int* __restrict__ srow,
                                                             const int* __restrict__ nptr,
                                                             const int* __restrict__ nslot,
                                                             const TM* __restrict__ ediag,
                                                             const int* __restrict__ bidx,
                                                             const double* __restrict__ jdb,
                                                             const double* __restrict__ nterm, int rxn,
                                                             const int* __restrict__ sbase, TM* __restrict__ M,
                                                             TM* __restrict__ D) {
  constexpr int NE = NS + 4, NE2 = NE * NE, NT = NE2 + NS * (NS + 1);
  const int lane = threadIdx.x & 63, rr = threadIdx.x >> 6;
  const int r = blockIdx.x * 64 + lane;
  if (r >= nrow_pad) return;
  const int i = srow[r];
  TM* Dr = D + (long long)(r >> 6) * NE2 * kLane + (r & 63);
  if (i < 0) {
    // padding rows: identity, so the factorisation of the chunk stays finite
#pragma unroll
    for (int c = 0; c < NE; ++c) Dr[(long long)(rr * NE + c) * kLane] = (TM)(c == rr ? 1.0 : 0.0);
    return;
  }
  double A[NE];
  const int bi = bidx[i];
#pragma unroll
  for (int c = 0; c < NE; ++c) A[c] = bi >= 0 ? jdb[(long long)bi * NE2 + rr * NE + c] : 0.0;
  // slot s of the row is its CSR slot nptr[i] + s: the row's own side of
  // the edge adds to the diagonal, the other side, negated, is the row's
  // off-diagonal block (A[L][R] = -dR, A[R][L] = -dL), written here with
  // contiguous lanes
  TM* mo = M + (long long)sbase[r >> 6] * NE2 * kLane + (long long)(rr * NE) * kLane + (r & 63);
  for (int p = nptr[i]; p < nptr[i + 1]; ++p) {
    const int sd = nslot[p];
    const TM* e = ediag + (long long)sd * NE2 + rr * NE;
    const TM* o = ediag + (long long)(sd ^ 1) * NE2 + rr * NE;
    TM* mb = mo + (long long)(p - nptr[i]) * NE2 * kLane;
#pragma unroll
    for (int c = 0; c < NE; ++c) {
      A[c] += (double)e[c];
      mb[(long long)c * kLane] = -o[c];
    }
  }
  const double* nt = nterm + (long long)i * NT;
  if (rxn && rr < NS) {
#pragma unroll
    for (int c = 0; c < NS; ++c) A[c] -= nt[NE2 + rr * (NS + 1) + c];
    A[NS + 3] -= nt[NE2 + rr * (NS + 1) + NS];
  }
#pragma unroll
  for (int c = 0; c < NE; ++c)
    if (temporal_entry<NS>(rr, c)) A[c] += nt[rr * NE + c];
#pragma unroll
  for (int c = 0; c < NE; ++c) Dr[(long long)(rr * NE + c) * kLane] = (TM)A[c];
}

// fr::lu (LU with logical pivoting, ucs/matrix.h:184-235) with the rows
// kept physically in pivot order: on return b[k * N + c] is lu's
// a[p[k] * N + c].  Every element meets the same operations in the same
// order, so the factor is the same bits, while all indices are static and
// the block stays in registers (lu's a[p[j] * N + i] sends it to scratch);
// a row swap is a select over the static row indices.  As in lu, the pivot
// row index carries over from the previous column when a column holds no
// nonzero, so the swap partner may be an earlier row.
template <int N, typename T>
__device__ __forceinline__ void lu_rows(T* b, int* p) {
  int row = 0;
#pragma unroll
  for (int i = 0; i < N; ++i) p[i] = i;
#pragma unroll
  for (int i = 0; i < N; ++i) {
    T large = 0.0;
#pragma unroll
    for (int j = i; j < N; ++j)
      if (fabs(b[j * N + i]) > fabs(large)) {
        large = b[j * N + i];
        row = j;
      }
#pragma unroll
    for (int j = 0; j < N; ++j) {
      if (j == i) continue;
      const bool s = (j == row);
      const int pi = p[i], pj = p[j];
      p[i] = s ? pj : pi;
      p[j] = s ? pi : pj;
#pragma unroll
      for (int k = 0; k < N; ++k) {
        const T x = b[i * N + k], y = b[j * N + k];
        b[i * N + k] = s ? y : x;
        b[j * N + k] = s ? x : y;
      }
    }
    large = (T)1.0 / large;
#pragma unroll
    for (int j = i + 1; j < N; ++j) b[j * N + i] *= large;
#pragma unroll
    for (int j = i + 1; j < N; ++j)
#pragma unroll
      for (int k = i + 1; k < N; ++k) b[j * N + k] -= b[j * N + i] * b[i * N + k];
  }
}

// PrepareSGS (ucs/crsmatrix.tcc:839-876): LU with logical pivoting of every
// diagonal block, in place; the row order is packed 4 bits per entry.
template <int NS, typename TM>
__global__ void __launch_bounds__(64) k_fr_prepare(int nrow_pad, TM* __restrict__ D, long long* __restrict__ piv) {
  constexpr int NE = NS + 4, NE2 = NE * NE;
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nrow_pad) return;
  TM* Dr = D + (long long)(r >> 6) * NE2 * kLane + (r & 63);
  // in the storage precision: the FP32 form factors in FP32 arithmetic
  TM A[NE2];
#pragma unroll
  for (int k = 0; k < NE2; ++k) A[k] = Dr[(long long)k * kLane];
  int p[NE];
  lu_rows<NE, TM>(A, p);
  long long pk = 0;
#pragma unroll
  for (int k = 0; k < NE; ++k) pk |= (long long)p[k] << (4 * k);
  piv[r] = pk;
  // stored with the rows in pivot order (row k holds factor row p[k]), so the
  // solve indexes the factor statically
#pragma unroll
  for (int k = 0; k < NE2; ++k) Dr[(long long)k * kLane] = A[k];
}

// LuSolve (ucs/matrix.h:237-264) on a factor stored with its rows in pivot
// order: the same products and sums as the logical-pivot form, with only
// the right-hand side picked by the pivot
template <int N, typename T>
__device__ __forceinline__ void lu_solve_rows(const T* a, T* b, const int* p) {
  T x[N];
#pragma unroll
  for (int i = 0; i < N; ++i) {
    T sum = 0.0;
#pragma unroll
    for (int j = 0; j < i; ++j) sum += a[i * N + j] * x[j];
    T bp = b[0];
#pragma unroll
    for (int k = 1; k < N; ++k) bp = (p[i] == k) ? b[k] : bp;
    x[i] = bp - sum;
  }
#pragma unroll
  for (int i = N - 1; i >= 0; --i) {
    T sum = 0.0;
#pragma unroll
    for (int j = N - 1; j > i; --j) sum += a[i * N + j] * b[j];
    b[i] = (x[i] - sum) / a[i * N + i];
  }
}

// One colour of a symmetric Gauss-Seidel sweep (CRS::SGS, ucs/crs.tcc:61-173):
// rows of a colour share no block, so they update together; per row the
// neighbours' products are subtracted in slot (edge) order, then LuSolve.
// A workgroup takes one 64-row SELL chunk; wave w forms component w of the
// 64 rows' right-hand sides (block row w of every slot: each block-entry load
// is 64 consecutive rows' entry), the nine components meet in LDS, and wave
// 0 solves the diagonal blocks.  Per component the arithmetic is the
// reference's matrix-vector product row and subtraction, in its order.
template <int NS, typename TM>
__global__ void __launch_bounds__(64 * (NS + 4)) k_fr_sgs_color(int chunk0, const int* __restrict__ srow,
                                                              const int* __restrict__ swidth,
                                                              const int* __restrict__ sbase,
                                                              const int* __restrict__ scol, const TM* __restrict__ Moff,
                                                              const TM* __restrict__ D,
                                                              const long long* __restrict__ piv,
                                                              const double* __restrict__ b, TM* __restrict__ x) {
  constexpr int NE = NS + 4, NE2 = NE * NE;
  __shared__ TM sh[NE][64];
  // the neighbours' x for a group of SB slots, double-buffered by group
  // parity: wave rr gathers component rr for the 64 rows, every wave reads
  // all NE (one gather per component instead of one per wave)
  constexpr int SB = 4;
  __shared__ TM xs[2][SB][NE][64];
  const int lane = threadIdx.x & 63, rr = threadIdx.x >> 6;
  const int chunk = chunk0 + blockIdx.x;
  const int r = chunk * 64 + lane;
  const int i = srow[r];
  TM rhs = i >= 0 ? (TM)b[(long long)i * NE + rr] : (TM)0;
  const int w = swidth[chunk];
  const long long cb = (long long)sbase[chunk] * kLane + lane;
  const TM* mb = Moff + (long long)sbase[chunk] * NE2 * kLane + (long long)(rr * NE) * kLane + lane;
  // every wave runs the chunk's w slots (the barriers need all of them);
  // a row's padding slots (j < 0) contribute nothing
  for (int s0 = 0; s0 < w; s0 += SB) {
    TM* xg = &xs[(s0 / SB) & 1][0][0][0];
    int jj[SB];
#pragma unroll
    for (int t = 0; t < SB; ++t) {
      jj[t] = (i >= 0 && s0 + t < w) ? scol[cb + (long long)(s0 + t) * kLane] : -1;
      xg[(t * NE + rr) * 64 + lane] = jj[t] >= 0 ? x[(long long)jj[t] * NE + rr] : (TM)0;
    }
    __syncthreads();
#pragma unroll
    for (int t = 0; t < SB; ++t) {
      if (jj[t] >= 0) {
        const TM* blk = mb + (long long)(s0 + t) * NE2 * kLane;
        const TM* xb = xg + t * NE * 64;
        TM v = blk[0] * xb[lane];
#pragma unroll
        for (int c = 1; c < NE; ++c) v += blk[(long long)c * kLane] * xb[c * 64 + lane];
        rhs -= v;
      }
    }
  }
  if (i >= 0) sh[rr][lane] = rhs;
  __syncthreads();
  if (rr != 0 || i < 0) return;
  TM rv[NE];
#pragma unroll
  for (int k = 0; k < NE; ++k) rv[k] = sh[k][lane];
  const TM* dg = D + (long long)chunk * NE2 * kLane + lane;
  TM A[NE2];
#pragma unroll
  for (int k = 0; k < NE2; ++k) A[k] = dg[(long long)k * kLane];
  const long long pk = piv[r];
  int p[NE];
#pragma unroll
  for (int k = 0; k < NE; ++k) p[k] = (int)((pk >> (4 * k)) & 15);
  lu_solve_rows<NE, TM>(A, rv, p);
#pragma unroll
  for (int k = 0; k < NE; ++k) x[(long long)i * NE + k] = rv[k];
}

template <typename T>
__global__ void __launch_bounds__(kFB) k_fr_zero(long long n, T* __restrict__ x) {
  const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) x[k] = (T)0;
}

template <typename T>
__global__ void __launch_bounds__(kFB) k_fr_sumsq(long long n, const T* __restrict__ x, double* __restrict__ red) {
  __shared__ double sh[kFB];
  double s = 0.0;
  for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (long long)gridDim.x * blockDim.x) {
    const double v = (double)x[k];
    s += v * v;
  }
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int w = kFB / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) red[blockIdx.x] = sh[0];
}

// NaN/Inf rows of the correction zeroed and counted (NewtonIterate,
// ucs/solutionSpace.tcc:772-800), then ApplyDQ on the node
template <int NS, typename TX>
__global__ void __launch_bounds__(kFB) k_fr_apply(const Model* __restrict__ mp, int nnode, const TX* __restrict__ x,
                                                 double* __restrict__ q, int* __restrict__ nzero) {
  constexpr int NE = NS + 4;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nnode) return;
  const Model& m = *mp;
  double dq[NE];
  bool bad = false;
#pragma unroll
  for (int k = 0; k < NE; ++k) {
    dq[k] = (double)x[(long long)i * NE + k];
    if (isnan(dq[k]) || isinf(dq[k])) bad = true;
  }
  if (bad) {
#pragma unroll
    for (int k = 0; k < NE; ++k) dq[k] = 0.0;
    atomicAdd(nzero, 1);
  }
  State<NS> s;
  load_state<NS>(q, i, s, m);
  fr::apply_dq<NS>(m, dq, s);
  store_state<NS>(q, i, s);
}

// ComputeAuxiliaryVariables in full for a download: P, rho, cv_i, conc_i
template <int NS>
__global__ void __launch_bounds__(kFB) k_fr_full_aux(const Model* __restrict__ mp, long long n,
                                                    const double* __restrict__ q, double* __restrict__ aux) {
  constexpr int NE = NS + 4, NA = 2 + 2 * NS;
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const Model& m = *mp;
  State<NS> s;
  load_state<NS>(q, i, s, m);
  double* o = aux + i * NA;
  o[0] = s.P;
  o[1] = s.rho;
  const double TDim = s.q[NS + 3] * m.ref_temperature;
#pragma unroll
  for (int k = 0; k < NS; ++k) {
    o[2 + k] = (fr::cp_species(m, k, TDim) - m.R[k]) / m.s_ref;
    o[2 + NS + k] = s.q[k] / m.mw[k] / 1000.0;
  }
}

}  // namespace

// ---- context -----------------------------------------------------------------

struct pcfd_fr {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr, e2 = nullptr, e3 = nullptr;
  std::string err;
  const pcfd_chem* chem = nullptr;
  int ns = 0, ne = 0, nnode = 0, nbnode = 0, nedge = 0, ntot = 0, ntag = 0, nbnd = 0;
  int ncolor = 0, nchunk = 0, nrow_pad = 0;
  long long nslot = 0;
  std::vector<int> color_chunk, h_nptr, h_nslot, h_edges, h_srow, h_spos, h_sbase;
  std::vector<long long> h_eoff;
  Model model{};
  Model* dmodel = nullptr;
  pcfd_fr_params prm{};
  bool have_params = false;
  int iter = 0;
  double cfl = 0.0;
  bool lu_done = false;
  double phase_ms[PCFD_FR_NPHASE] = {0};
  long long bytes = 0;
  // device arrays
  double *q = nullptr, *qref = nullptr, *vol = nullptr, *edge_a = nullptr, *he_a = nullptr, *eflux = nullptr,
         *elam = nullptr, *b = nullptr, *dt = nullptr, *jdb = nullptr, *bhe = nullptr, *red = nullptr, *aux = nullptr;
  int *edges = nullptr, *nptr = nullptr, *nslot_d = nullptr, *he_ptr = nullptr, *he_id = nullptr, *he_tag = nullptr,
      *bc_type = nullptr, *bnd = nullptr, *bidx = nullptr, *srow = nullptr, *swidth = nullptr, *sbase = nullptr,
      *scol = nullptr, *icount = nullptr;
  long long* piv = nullptr;
  double *ediag = nullptr, *nterm = nullptr;  // Jacobian diagonal parts: per edge side, per node
  double *Md = nullptr, *Dd = nullptr, *xd = nullptr;
  float *Mf = nullptr, *Df = nullptr, *xf = nullptr;
  // production-rate surrogate in the residual (null: exact rates)
  pcfd_mlp* mlp = nullptr;
  pcfd_surrogate_desc sdesc{};
  double *s_rhoi = nullptr, *s_T = nullptr, *s_w = nullptr;
  int *s_lst = nullptr, *s_cnt = nullptr, *s_off = nullptr;
  int s_last_fallback = 0;
  long long s_calls = 0, s_fallback_total = 0;
  double s_ms = 0.0;
};

namespace {

template <typename T>
void dalloc(pcfd_fr& f, T** p, size_t n) {
  hcheck(hipMalloc((void**)p, sizeof(T) * std::max<size_t>(n, 1)), "hipMalloc");
  f.bytes += (long long)(sizeof(T) * std::max<size_t>(n, 1));
}

template <typename T>
void dput(pcfd_fr& f, T** p, const std::vector<T>& h) {
  dalloc(f, p, h.size());
  if (!h.empty()) hcheck(hipMemcpy(*p, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice), "hipMemcpy");
}

template <typename T>
void dfree(T*& p) {
  if (p) (void)hipFree(p);
  p = nullptr;
}

void free_all(pcfd_fr& f) {
  dfree(f.q); dfree(f.qref); dfree(f.vol); dfree(f.edge_a); dfree(f.he_a); dfree(f.eflux); dfree(f.elam);
  dfree(f.b); dfree(f.dt); dfree(f.jdb); dfree(f.bhe); dfree(f.red); dfree(f.aux); dfree(f.edges); dfree(f.nptr);
  dfree(f.nslot_d); dfree(f.he_ptr); dfree(f.he_id); dfree(f.he_tag); dfree(f.bc_type); dfree(f.bnd);
  dfree(f.bidx); dfree(f.srow); dfree(f.swidth); dfree(f.sbase); dfree(f.scol); dfree(f.icount);
  dfree(f.piv); dfree(f.Md); dfree(f.Dd); dfree(f.xd); dfree(f.Mf); dfree(f.Df); dfree(f.xf); dfree(f.dmodel);
  dfree(f.s_rhoi); dfree(f.s_T); dfree(f.s_w); dfree(f.s_lst); dfree(f.s_cnt); dfree(f.s_off);
  dfree(f.ediag); dfree(f.nterm);
  if (f.e0) (void)hipEventDestroy(f.e0);
  if (f.e1) (void)hipEventDestroy(f.e1);
  if (f.e2) (void)hipEventDestroy(f.e2);
  if (f.e3) (void)hipEventDestroy(f.e3);
  if (f.stream) (void)hipStreamDestroy(f.stream);
  f.e0 = f.e1 = f.e2 = f.e3 = nullptr;
  f.stream = nullptr;
}

template <typename F>
int guarded(pcfd_fr* f, F&& fn) {
  if (!f) return PCFD_ERR_ARG;
  try {
    hcheck(hipSetDevice(f->device), "hipSetDevice");
    return fn(*f);
  } catch (const std::exception& e) {
    f->err = e.what();
    return PCFD_ERR_HIP;
  }
}

int need_params(pcfd_fr& f) {
  if (!f.have_params) {
    f.err = "parameters not set (pcfd_fr_set_params)";
    return PCFD_ERR_STATE;
  }
  return PCFD_OK;
}

bool fp32(const pcfd_fr& f) { return f.prm.precision == PCFD_CHOSEN_FP32; }

// phase timers: [k] += device time between e0/e1 (or e2/e3 for a sub-timer)
void tbeg(pcfd_fr& f) { hcheck(hipEventRecord(f.e0, f.stream), "hipEventRecord"); }
void tend(pcfd_fr& f, int k) {
  hcheck(hipEventRecord(f.e1, f.stream), "hipEventRecord");
  hcheck(hipEventSynchronize(f.e1), "hipEventSynchronize");
  float ms = 0.f;
  hcheck(hipEventElapsedTime(&ms, f.e0, f.e1), "hipEventElapsedTime");
  f.phase_ms[k] += ms;
}
void sbeg(pcfd_fr& f) { hcheck(hipEventRecord(f.e2, f.stream), "hipEventRecord"); }
void send(pcfd_fr& f, int k) {
  hcheck(hipEventRecord(f.e3, f.stream), "hipEventRecord");
  hcheck(hipEventSynchronize(f.e3), "hipEventSynchronize");
  float ms = 0.f;
  hcheck(hipEventElapsedTime(&ms, f.e2, f.e3), "hipEventElapsedTime");
  f.phase_ms[k] += ms;
}

constexpr int NS5 = 5;

void run_bcs(pcfd_fr& f) {
  if (f.nbnd == 0) return;
  hipLaunchKernelGGL(k_fr_bcs<NS5>, dim3(nblocks(f.nbnd)), dim3(kFB), 0, f.stream, f.dmodel, f.nbnd, f.bnd, f.he_ptr,
                     f.he_id, f.he_tag, f.he_a, f.bc_type, f.qref, f.nnode, f.prm.beta, f.q);
  hcheck(hipGetLastError(), "k_fr_bcs");
}

// residual into f.b; norms[1+neqn] (residGlobal, per-equation) when non-null
void run_residual(pcfd_fr& f, double* norms) {
  if (f.nedge > 0) {
    sbeg(f);
    hipLaunchKernelGGL(k_fr_edge_flux<NS5>, dim3(nblocks(f.nedge)), dim3(kFB), 0, f.stream, f.dmodel, f.nedge, f.edges,
                       f.edge_a, f.q, f.prm.beta, f.eflux);
    hcheck(hipGetLastError(), "k_fr_edge_flux");
    send(f, 7);
  }
  const double* wr = nullptr;
  if (f.mlp && f.prm.reactions) {
    hipLaunchKernelGGL(k_fr_chem_inputs<NS5>, dim3(nblocks(f.nnode)), dim3(kFB), 0, f.stream,
                       chem_device_model(f.chem), f.nnode, f.q, f.s_rhoi, f.s_T);
    hcheck(hipGetLastError(), "k_fr_chem_inputs");
    hcheck(hipStreamSynchronize(f.stream), "sync");
    int nf = 0;
    double ms = 0.0;
    const int rc = surrogate_launch(const_cast<pcfd_chem*>(f.chem), *f.mlp, &f.sdesc, f.nnode, f.s_rhoi, f.s_T,
                                    f.s_w, f.s_lst, f.s_cnt, f.s_off, &nf, &ms);
    if (rc != PCFD_OK) throw std::runtime_error("surrogate production rates failed");
    f.s_last_fallback = nf;
    f.s_fallback_total += nf;
    f.s_calls++;
    f.s_ms += ms;
    wr = f.s_w;
  }
  const int nb = nblocks(f.nnode);
  hipLaunchKernelGGL(k_fr_residual<NS5>, dim3(nb), dim3(kFB), 0, f.stream, f.dmodel, chem_device_model(f.chem),
                     f.nnode, f.nptr, f.nslot_d, f.eflux, f.he_ptr, f.he_id, f.he_a, f.q, f.vol, f.prm.reactions,
                     wr, f.prm.beta, f.b, f.red);
  hcheck(hipGetLastError(), "k_fr_residual");
  if (norms) {
    const int ne = f.ne;
    std::vector<double> h((size_t)(ne + 1) * nb);
    hcheck(hipMemcpyAsync(h.data(), f.red, sizeof(double) * h.size(), hipMemcpyDeviceToHost, f.stream), "copy");
    hcheck(hipStreamSynchronize(f.stream), "sync");
    for (int k = 0; k <= ne; ++k) {
      double s = 0.0;
      for (int j = 0; j < nb; ++j) s += h[(size_t)k * nb + j];
      norms[k] = k == 0 ? std::sqrt(s) / ((double)f.nnode * ne) : std::sqrt(s) / (double)f.nnode;
    }
  }
}

double run_timestep(pcfd_fr& f, double cfl, bool want_min) {
  if (f.nedge > 0) {
    hipLaunchKernelGGL(k_fr_edge_lam<NS5>, dim3(nblocks(f.nedge)), dim3(kFB), 0, f.stream, f.dmodel, f.nedge, f.edges,
                       f.edge_a, f.q, f.prm.beta, f.elam);
    hcheck(hipGetLastError(), "k_fr_edge_lam");
  }
  hipLaunchKernelGGL(k_fr_timestep<NS5>, dim3(nblocks(f.nnode)), dim3(kFB), 0, f.stream, f.dmodel, f.nnode, f.nptr,
                     f.nslot_d, f.elam, f.he_ptr, f.he_id, f.he_a, f.q, f.vol, cfl, f.prm.beta, f.dt);
  hcheck(hipGetLastError(), "k_fr_timestep");
  if (!want_min) return 0.0;
  std::vector<double> h(f.nnode);
  hcheck(hipMemcpyAsync(h.data(), f.dt, sizeof(double) * f.nnode, hipMemcpyDeviceToHost, f.stream), "copy");
  hcheck(hipStreamSynchronize(f.stream), "sync");
  double m = h[0];
  for (int i = 1; i < f.nnode; ++i) m = std::min(m, h[i]);
  return m;
}

template <typename TM>
void run_jacobian_t(pcfd_fr& f, TM* M, TM* D) {
  if (f.nedge > 0) {
    sbeg(f);
    hipLaunchKernelGGL(k_fr_edge_flux<NS5>, dim3(nblocks(f.nedge)), dim3(kFB), 0, f.stream, f.dmodel, f.nedge,
                       f.edges, f.edge_a, f.q, f.prm.beta, f.eflux);
    hcheck(hipGetLastError(), "k_fr_edge_flux");
    hipLaunchKernelGGL((k_fr_jac_edge<NS5, TM>), dim3(nblocks((long long)f.nedge * (NS5 + 4))), dim3(kFB), 0,
                       f.stream, f.dmodel, f.nedge, f.edges, f.edge_a, f.q, f.prm.beta, f.eflux,
                       reinterpret_cast<TM*>(f.ediag));
    hcheck(hipGetLastError(), "k_fr_jac_edge");
    send(f, 6);
  }
  if (f.nbnd > 0) {
    hipLaunchKernelGGL(k_fr_jac_bnd_base<NS5>, dim3(nblocks(f.nbnd)), dim3(kFB), 0, f.stream, f.dmodel, f.nbnd,
                       f.bnd, f.he_ptr, f.he_id, f.he_tag, f.he_a, f.bc_type, f.qref, f.nnode, f.prm.beta, f.q,
                       f.bhe);
    hcheck(hipGetLastError(), "k_fr_jac_bnd_base");
    hipLaunchKernelGGL(k_fr_jac_bnd_cols<NS5>, dim3(nblocks((long long)f.nbnd * (NS5 + 4))), dim3(kFB), 0,
                       f.stream, f.dmodel, f.nbnd, f.bnd, f.he_ptr, f.he_id, f.he_tag, f.he_a, f.bc_type, f.qref,
                       f.prm.beta, f.bhe, f.jdb);
    hcheck(hipGetLastError(), "k_fr_jac_bnd_cols");
  }
  hipLaunchKernelGGL(k_fr_node_terms<NS5>, dim3(nblocks(f.nnode)), dim3(kFB), 0, f.stream, f.dmodel,
                     chem_device_model(f.chem), f.nnode, f.q, f.vol, f.dt, f.prm.reactions, f.prm.beta, f.nterm);
  hcheck(hipGetLastError(), "k_fr_node_terms");
  hipLaunchKernelGGL((k_fr_jac_diag<NS5, TM>), dim3(f.nrow_pad / kLane), dim3(kLane * (NS5 + 4)), 0, f.stream,
                     f.nrow_pad, f.srow, f.nptr, f.nslot_d, reinterpret_cast<const TM*>(f.ediag), f.bidx, f.jdb,
                     f.nterm, f.prm.reactions,
                     f.sbase, M, D);
  hcheck(hipGetLastError(), "k_fr_jac_diag");
  f.lu_done = false;
}

void run_jacobian(pcfd_fr& f) {
  if (fp32(f)) run_jacobian_t<float>(f, f.Mf, f.Df);
  else run_jacobian_t<double>(f, f.Md, f.Dd);
}

void run_prepare(pcfd_fr& f) {
  if (fp32(f))
    hipLaunchKernelGGL((k_fr_prepare<NS5, float>), dim3(nblocks(f.nrow_pad, 64)), dim3(64), 0, f.stream, f.nrow_pad,
                       f.Df, f.piv);
  else
    hipLaunchKernelGGL((k_fr_prepare<NS5, double>), dim3(nblocks(f.nrow_pad, 64)), dim3(64), 0, f.stream, f.nrow_pad,
                       f.Dd, f.piv);
  hcheck(hipGetLastError(), "k_fr_prepare");
  f.lu_done = true;
}

template <typename TM>
double xnorm(pcfd_fr& f, const TM* x) {
  const int nb = 256;
  const long long n = (long long)f.nnode * f.ne;
  hipLaunchKernelGGL(k_fr_sumsq<TM>, dim3(nb), dim3(kFB), 0, f.stream, n, x, f.red);
  hcheck(hipGetLastError(), "k_fr_sumsq");
  double h[nb];
  hcheck(hipMemcpyAsync(h, f.red, sizeof(h), hipMemcpyDeviceToHost, f.stream), "copy");
  hcheck(hipStreamSynchronize(f.stream), "sync");
  double s = 0.0;
  for (int k = 0; k < nb; ++k) s += h[k];
  return std::sqrt(s) / (double)n;
}

template <typename TM>
void run_sgs_t(pcfd_fr& f, int nsgs, const TM* M, const TM* D, TM* x, double* delta) {
  const long long n = (long long)f.nnode * f.ne;
  hipLaunchKernelGGL(k_fr_zero<TM>, dim3(nblocks(n)), dim3(kFB), 0, f.stream, n, x);
  hcheck(hipGetLastError(), "k_fr_zero");
  double xold = 0.0, xn = 0.0;
  // A colour pass directly after a pass over the same colour (the turn of a
  // symmetric sweep, and the start of the next) reads the same neighbour x
  // and rewrites the same values to the bit -- a colour's rows never read
  // each other -- so it is skipped.
  int last = -1;
  for (int it = 0; it < nsgs; ++it) {
    for (int dir = 0; dir < 2; ++dir)
      for (int cc = 0; cc < f.ncolor; ++cc) {
        const int col = dir == 0 ? cc : f.ncolor - 1 - cc;
        const int nch = f.color_chunk[col + 1] - f.color_chunk[col];
        if (nch == 0 || col == last) continue;
        last = col;
        hipLaunchKernelGGL((k_fr_sgs_color<NS5, TM>), dim3(nch), dim3(kLane * (NS5 + 4)), 0, f.stream,
                           f.color_chunk[col], f.srow, f.swidth, f.sbase, f.scol, M, D, f.piv, f.b, x);
        hcheck(hipGetLastError(), "k_fr_sgs_color");
      }
    if (delta && it >= nsgs - 2) {
      xold = xn;
      xn = xnorm(f, x);
    }
  }
  if (delta) *delta = std::fabs(xold - xn);
}

void run_sgs(pcfd_fr& f, int nsgs, double* delta) {
  if (fp32(f)) run_sgs_t<float>(f, nsgs, f.Mf, f.Df, f.xf, delta);
  else run_sgs_t<double>(f, nsgs, f.Md, f.Dd, f.xd, delta);
}

int run_apply(pcfd_fr& f) {
  hcheck(hipMemsetAsync(f.icount, 0, sizeof(int), f.stream), "hipMemsetAsync");
  if (fp32(f))
    hipLaunchKernelGGL((k_fr_apply<NS5, float>), dim3(nblocks(f.nnode)), dim3(kFB), 0, f.stream, f.dmodel, f.nnode,
                       f.xf, f.q, f.icount);
  else
    hipLaunchKernelGGL((k_fr_apply<NS5, double>), dim3(nblocks(f.nnode)), dim3(kFB), 0, f.stream, f.dmodel, f.nnode,
                       f.xd, f.q, f.icount);
  hcheck(hipGetLastError(), "k_fr_apply");
  int nz = 0;
  hcheck(hipMemcpyAsync(&nz, f.icount, sizeof(int), hipMemcpyDeviceToHost, f.stream), "copy");
  hcheck(hipStreamSynchronize(f.stream), "sync");
  return nz;
}

// Param::UpdateCFL without dynamic CFL (ucs/param.tcc:695-726)
double cfl_of(const pcfd_fr_params& p, int iter) {
  if (p.cfl_ramp > 0 && iter < p.cfl_ramp)
    return (p.cfl - p.cfl_start) / (double)p.cfl_ramp * (double)iter + p.cfl_start;
  return p.cfl;
}

// PreTimeAdvance + NewtonIterate + PostTimeAdvance (ucs/solutionSpace.tcc:572-916)
double step(pcfd_fr& f, double* norms, int* nzeroed) {
  const int jf = f.prm.jac_freq > 0 ? f.prm.jac_freq : 1;
  if (f.prm.nsgs > 0 && (f.iter % jf == 0 || f.iter == 1)) {
    tbeg(f);
    run_jacobian(f);
    tend(f, 0);
  }
  f.cfl = cfl_of(f.prm, f.iter);
  tbeg(f);
  run_timestep(f, f.cfl, false);
  tend(f, 1);
  tbeg(f);
  run_bcs(f);
  tend(f, 2);
  std::vector<double> nrm(f.ne + 1);
  tbeg(f);
  run_residual(f, nrm.data());
  tend(f, 3);
  int nz = 0;
  if (f.prm.nsgs > 0) {
    tbeg(f);
    // PrepareSGS factors the diagonal blocks once per assembly (its ludiag
    // flag, reset when the matrix is rebuilt, ucs/crsmatrix.tcc:505-510,839-876)
    if (!f.lu_done) run_prepare(f);
    run_sgs(f, f.prm.nsgs, nullptr);
    tend(f, 4);
    tbeg(f);
    nz = run_apply(f);
    tend(f, 5);
  } else {
    throw std::runtime_error("pcfd_fr_step: the explicit (nsgs = 0) update is not provided for FR");
  }
  if (norms) std::copy(nrm.begin(), nrm.end(), norms);
  if (nzeroed) *nzeroed = nz;
  f.iter++;
  return nrm[0];
}

}  // namespace

extern "C" {

int pcfd_fr_create(const pcfd_mesh_desc* m, pcfd_chem* chem, const double* qref, pcfd_fr** out) {
  if (!m || !chem || !qref || !out) {
    g_fr_create_err = "pcfd_fr_create: null argument";
    return PCFD_ERR_ARG;
  }
  *out = nullptr;
  const int ns = chem_nspecies(chem);
  if (ns != NS5) {
    g_fr_create_err = "pcfd_fr_create: " + std::to_string(ns) + " species; this build supports 5";
    return PCFD_ERR_UNSUPPORTED;
  }
  if (m->nnode <= 0 || m->nbnode < 0 || m->nedge < 0 || m->nfactag <= 0 || !m->vol || !m->bc_type ||
      (m->nedge > 0 && (!m->edges || !m->edge_a)) || (m->nbnode > 0 && (!m->bedges || !m->bedge_a || !m->bfactag))) {
    g_fr_create_err = "pcfd_fr_create: invalid mesh description";
    return PCFD_ERR_ARG;
  }
  const int nn = m->nnode, nb = m->nbnode, ne = m->nedge, nt = m->nfactag, NE = ns + 4, NE2 = NE * NE;
  for (long long e = 0; e < ne; ++e) {
    const int a = m->edges[2 * e], c = m->edges[2 * e + 1];
    if (a < 0 || c < 0 || a >= nn || c >= nn || a == c) {
      g_fr_create_err = "pcfd_fr_create: edge " + std::to_string(e) + " has invalid nodes";
      return PCFD_ERR_ARG;
    }
  }
  for (int h = 0; h < nb; ++h) {
    const int a = m->bedges[2 * h], ph = m->bedges[2 * h + 1], tag = m->bfactag[h];
    if (a < 0 || a >= nn || ph != nn + h || tag < 0 || tag >= nt) {
      g_fr_create_err = "pcfd_fr_create: half-edge " + std::to_string(h) + " invalid";
      return PCFD_ERR_ARG;
    }
  }
  for (int t = 0; t < nt; ++t) {
    const int ty = m->bc_type[t];
    if (!(ty == 0 || ty == 1 || ty == 2 || ty == 3 || ty == 6 || ty == 7 || ty == 8 || ty == 9)) {
      bool used = false;
      for (int h = 0; h < nb && !used; ++h) used = m->bfactag[h] == t;
      if (used) {
        g_fr_create_err = "pcfd_fr_create: face tag " + std::to_string(t) + " has a boundary type (" +
                          std::to_string(ty) + ") the FR path does not provide";
        return PCFD_ERR_UNSUPPORTED;
      }
    }
  }
  pcfd_fr* f = new pcfd_fr();
  try {
    hcheck(hipGetDevice(&f->device), "hipGetDevice");
    hcheck(hipStreamCreateWithFlags(&f->stream, hipStreamNonBlocking), "hipStreamCreate");
    hcheck(hipEventCreate(&f->e0), "hipEventCreate");
    hcheck(hipEventCreate(&f->e1), "hipEventCreate");
    hcheck(hipEventCreate(&f->e2), "hipEventCreate");
    hcheck(hipEventCreate(&f->e3), "hipEventCreate");
    f->chem = chem;
    f->ns = ns;
    f->ne = NE;
    f->nnode = nn;
    f->nbnode = nb;
    f->nedge = ne;
    f->ntot = nn + nb;
    f->ntag = nt;
    // node lists in edge order (the order the reference's edge loop adds to a node)
    std::vector<int> nptr(nn + 1, 0);
    for (int e = 0; e < ne; ++e) {
      nptr[m->edges[2 * e] + 1]++;
      nptr[m->edges[2 * e + 1] + 1]++;
    }
    for (int i = 0; i < nn; ++i) nptr[i + 1] += nptr[i];
    std::vector<int> nslot(2 * (size_t)ne), nbr(2 * (size_t)ne);
    {
      std::vector<int> fill(nptr.begin(), nptr.end() - 1);
      for (int e = 0; e < ne; ++e) {
        const int a = m->edges[2 * e], c = m->edges[2 * e + 1];
        nbr[fill[a]] = c;
        nslot[fill[a]++] = 2 * e;
        nbr[fill[c]] = a;
        nslot[fill[c]++] = 2 * e + 1;
      }
      std::vector<int> stamp(nn, -1);
      for (int i = 0; i < nn; ++i)
        for (int p = nptr[i]; p < nptr[i + 1]; ++p) {
          if (stamp[nbr[p]] == i) throw std::runtime_error("pcfd_fr_create: duplicate edge between nodes");
          stamp[nbr[p]] = i;
        }
    }
    // half-edges per node, in half-edge order
    std::vector<int> he_ptr(nn + 1, 0), he_id(nb);
    for (int h = 0; h < nb; ++h) he_ptr[m->bedges[2 * h] + 1]++;
    for (int i = 0; i < nn; ++i) he_ptr[i + 1] += he_ptr[i];
    {
      std::vector<int> fill(he_ptr.begin(), he_ptr.end() - 1);
      for (int h = 0; h < nb; ++h) he_id[fill[m->bedges[2 * h]]++] = h;
    }
    std::vector<int> bnd, bidx(nn, -1);
    for (int i = 0; i < nn; ++i)
      if (he_ptr[i + 1] > he_ptr[i]) {
        bidx[i] = (int)bnd.size();
        bnd.push_back(i);
      }
    f->nbnd = (int)bnd.size();
    // colouring and the SELL layout of the block matrix
    std::vector<int> colour(nn);
    const int ncol = pcfd_color_nodes(nn, ne, m->edges, colour.data());
    if (ncol < 0) throw std::runtime_error(pcfd_mesh_error());
    f->ncolor = ncol;
    std::vector<std::vector<int>> members(ncol);
    for (int i = 0; i < nn; ++i) members[colour[i]].push_back(i);
    auto deg = [&](int i) { return nptr[i + 1] - nptr[i]; };
    std::vector<int> srow;
    f->color_chunk.assign(ncol + 1, 0);
    for (int k = 0; k < ncol; ++k) {
      auto& v = members[k];
      std::stable_sort(v.begin(), v.end(), [&](int a, int c) { return deg(a) > deg(c); });
      srow.insert(srow.end(), v.begin(), v.end());
      while (srow.size() % kLane) srow.push_back(-1);
      f->color_chunk[k + 1] = (int)(srow.size() / kLane);
    }
    f->nrow_pad = (int)srow.size();
    f->nchunk = f->nrow_pad / kLane;
    std::vector<int> swidth(f->nchunk, 0), sbase(f->nchunk, 0), spos(nn, -1);
    long long slots = 0;
    for (int k = 0; k < f->nchunk; ++k) {
      int w = 0;
      for (int l = 0; l < kLane; ++l) {
        const int i = srow[(size_t)k * kLane + l];
        if (i >= 0) {
          w = std::max(w, deg(i));
          spos[i] = k * kLane + l;
        }
      }
      swidth[k] = w;
      sbase[k] = (int)slots;
      slots += w;
    }
    f->nslot = slots;
    std::vector<int> scol((size_t)slots * kLane, -1);
    std::vector<long long> eoff(2 * (size_t)ne);
    for (int i = 0; i < nn; ++i) {
      const int r = spos[i], k = r / kLane, l = r % kLane;
      for (int p = nptr[i]; p < nptr[i + 1]; ++p) {
        const long long s = (long long)sbase[k] + (p - nptr[i]);
        scol[(size_t)s * kLane + l] = nbr[p];
        // row i's block for this neighbour: edge e's (n0,n1) block when i is n0
        eoff[2 * (size_t)(nslot[p] >> 1) + (nslot[p] & 1)] = s * NE2 * kLane + l;
      }
    }
    f->h_nptr = nptr;
    f->h_nslot = nslot;
    f->h_edges.assign(m->edges, m->edges + 2 * (size_t)ne);
    f->h_srow = srow;
    f->h_spos = spos;
    f->h_sbase = sbase;
    f->h_eoff = eoff;
    // model constants (scales set with the parameters)
    Model md{};
    md.ns = ns;
    // species data from the chemistry model
    ChemDev cd{};
    hcheck(hipMemcpy(&cd, chem_device_model(chem), sizeof(ChemDev), hipMemcpyDeviceToHost), "hipMemcpy");
    for (int s = 0; s < ns; ++s) {
      md.mw[s] = cd.mw[s];
      md.R[s] = fr::kUnivR / cd.mw[s];
      for (int k = 0; k < 7; ++k) {
        md.lo[s][k] = cd.lo[s][k];
        md.hi[s][k] = cd.hi[s][k];
      }
    }
    md.ref_density = cd.ref_density;
    md.ref_temperature = cd.ref_temperature;
    f->model = md;
    // device arrays
    std::vector<double> he_a(4 * (size_t)nb), vol(m->vol, m->vol + nn), ea(m->edge_a, m->edge_a + 4 * (size_t)ne);
    for (int h = 0; h < nb; ++h)
      for (int k = 0; k < 4; ++k) he_a[4 * (size_t)h + k] = m->bedge_a[4 * (size_t)h + k];
    std::vector<int> edges(m->edges, m->edges + 2 * (size_t)ne), he_tag(m->bfactag, m->bfactag + nb),
        bct(m->bc_type, m->bc_type + nt);
    std::vector<double> qr(qref, qref + (size_t)nt * NE);
    hcheck(hipMalloc((void**)&f->dmodel, sizeof(Model)), "hipMalloc");
    dput(*f, &f->edges, edges);
    dput(*f, &f->edge_a, ea);
    dput(*f, &f->vol, vol);
    dput(*f, &f->he_a, he_a);
    dput(*f, &f->he_ptr, he_ptr);
    dput(*f, &f->he_id, he_id);
    dput(*f, &f->he_tag, he_tag);
    dput(*f, &f->bc_type, bct);
    dput(*f, &f->qref, qr);
    dput(*f, &f->bnd, bnd);
    dput(*f, &f->bidx, bidx);
    dput(*f, &f->nptr, nptr);
    dput(*f, &f->nslot_d, nslot);
    dput(*f, &f->srow, srow);
    dput(*f, &f->swidth, swidth);
    dput(*f, &f->sbase, sbase);
    dput(*f, &f->scol, scol);
    dalloc(*f, &f->q, (size_t)f->ntot * NE);
    dalloc(*f, &f->eflux, (size_t)ne * NE);
    dalloc(*f, &f->elam, (size_t)ne);
    dalloc(*f, &f->b, (size_t)nn * NE);
    dalloc(*f, &f->dt, (size_t)nn);
    dalloc(*f, &f->jdb, (size_t)std::max(f->nbnd, 1) * NE2);
    dalloc(*f, &f->bhe, (size_t)std::max(nb, 1) * 3 * NE);
    dalloc(*f, &f->red, (size_t)(NE + 1) * std::max(nblocks(nn), 256));
    dalloc(*f, &f->icount, 4);
    dalloc(*f, &f->piv, (size_t)f->nrow_pad);
    dalloc(*f, &f->ediag, 2 * (size_t)ne * NE2);
    dalloc(*f, &f->nterm, (size_t)nn * (NE2 + ns * (ns + 1)));
    hcheck(hipMemset(f->q, 0, sizeof(double) * (size_t)f->ntot * NE), "hipMemset");
  } catch (const std::exception& e) {
    g_fr_create_err = e.what();
    free_all(*f);
    delete f;
    return PCFD_ERR_HIP;
  }
  *out = f;
  return PCFD_OK;
}

int pcfd_fr_destroy(pcfd_fr* f) {
  if (!f) return PCFD_ERR_ARG;
  free_all(*f);
  delete f;
  return PCFD_OK;
}

const char* pcfd_fr_last_error(const pcfd_fr* f) { return f ? f->err.c_str() : g_fr_create_err.c_str(); }

int pcfd_fr_set_params(pcfd_fr* fp, const pcfd_fr_params* p) {
  return guarded(fp, [&](pcfd_fr& f) -> int {
    if (!p || !(p->ref_pressure > 0.0) || !(p->ref_velocity > 0.0) || !(p->beta > 0.0) ||
        (p->precision != PCFD_PLAIN_FP64 && p->precision != PCFD_CHOSEN_FP32) || p->nsgs < 0 ||
        !(p->cfl > 0.0) || (p->cfl_ramp > 0 && !(p->cfl_start > 0.0))) {
      f.err = "pcfd_fr_set_params: invalid parameters";
      return PCFD_ERR_ARG;
    }
    const bool was32 = f.have_params && fp32(f);
    f.prm = *p;
    f.have_params = true;
    Model& md = f.model;
    md.ref_pressure = p->ref_pressure;
    md.ref_velocity = p->ref_velocity;
    md.vref2 = p->ref_velocity * p->ref_velocity;
    md.s_ref = p->ref_velocity * p->ref_velocity / md.ref_temperature;
    md.Pref = p->Pref;
    md.sym2d = p->symmetry2d;
    md.local_dt = p->local_dt;
    md.dt_param = p->dt_param;
    md.no_cvbc = p->no_cvbc;
    hcheck(hipMemcpy(f.dmodel, &md, sizeof(Model), hipMemcpyHostToDevice), "hipMemcpy");
    const size_t NE2 = (size_t)f.ne * f.ne;
    const size_t noff = (size_t)f.nslot * kLane * NE2, ndg = (size_t)f.nchunk * kLane * NE2;
    if (fp32(f)) {
      if (!f.Mf) {
        dalloc(f, &f.Mf, noff);
        dalloc(f, &f.Df, ndg);
        dalloc(f, &f.xf, (size_t)f.nnode * f.ne);
      }
    } else if (!f.Md) {
      dalloc(f, &f.Md, noff);
      dalloc(f, &f.Dd, ndg);
      dalloc(f, &f.xd, (size_t)f.nnode * f.ne);
    }
    if (was32 != fp32(f)) f.lu_done = false;
    return PCFD_OK;
  });
}

int pcfd_fr_info(const pcfd_fr* f, int* ncolor, long long* nblk, long long* device_bytes) {
  if (!f) return PCFD_ERR_ARG;
  if (ncolor) *ncolor = f->ncolor;
  if (nblk) *nblk = (long long)f->nnode + 2LL * f->nedge;
  if (device_bytes) *device_bytes = f->bytes;
  return PCFD_OK;
}

int pcfd_fr_upload_q(pcfd_fr* fp, const double* q, int stride) {
  return guarded(fp, [&](pcfd_fr& f) -> int {
    if (!q || stride < f.ne) {
      f.err = "pcfd_fr_upload_q: stride must be at least ns + 4";
      return PCFD_ERR_ARG;
    }
    std::vector<double> h((size_t)f.ntot * f.ne);
    for (size_t i = 0; i < (size_t)f.ntot; ++i)
      for (int k = 0; k < f.ne; ++k) h[i * f.ne + k] = q[i * stride + k];
    hcheck(hipMemcpyAsync(f.q, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice, f.stream), "copy");
    hcheck(hipStreamSynchronize(f.stream), "sync");
    return PCFD_OK;
  });
}

int pcfd_fr_download_q(pcfd_fr* fp, double* q, int stride) {
  return guarded(fp, [&](pcfd_fr& f) -> int {
    if (!q || stride < f.ne) {
      f.err = "pcfd_fr_download_q: stride must be at least ns + 4";
      return PCFD_ERR_ARG;
    }
    std::vector<double> h((size_t)f.ntot * f.ne);
    hcheck(hipMemcpyAsync(h.data(), f.q, sizeof(double) * h.size(), hipMemcpyDeviceToHost, f.stream), "copy");
    const int na = 2 + 2 * f.ns;
    std::vector<double> ha;
    const bool full = stride >= f.ne + na && f.have_params;
    if (full) {
      if (!f.aux) dalloc(f, &f.aux, (size_t)f.ntot * na);
      hipLaunchKernelGGL(k_fr_full_aux<NS5>, dim3(nblocks(f.ntot)), dim3(kFB), 0, f.stream, f.dmodel,
                         (long long)f.ntot, f.q, f.aux);
      hcheck(hipGetLastError(), "k_fr_full_aux");
      ha.resize((size_t)f.ntot * na);
      hcheck(hipMemcpyAsync(ha.data(), f.aux, sizeof(double) * ha.size(), hipMemcpyDeviceToHost, f.stream), "copy");
    }
    hcheck(hipStreamSynchronize(f.stream), "sync");
    for (size_t i = 0; i < (size_t)f.ntot; ++i) {
      for (int k = 0; k < f.ne; ++k) q[i * stride + k] = h[i * f.ne + k];
      if (full)
        for (int k = 0; k < na; ++k) q[i * stride + f.ne + k] = ha[i * na + k];
    }
    return PCFD_OK;
  });
}

int pcfd_fr_update_bcs(pcfd_fr* fp) {
  return guarded(fp, [&](pcfd_fr& f) -> int {
    if (int rc = need_params(f)) return rc;
    run_bcs(f);
    hcheck(hipStreamSynchronize(f.stream), "sync");
    return PCFD_OK;
  });
}

int pcfd_fr_residual(pcfd_fr* fp, double* b_out, double* norms) {
  return guarded(fp, [&](pcfd_fr& f) -> int {
    if (int rc = need_params(f)) return rc;
    std::vector<double> nrm(f.ne + 1);
    run_residual(f, nrm.data());
    if (norms) std::copy(nrm.begin(), nrm.end(), norms);
    if (b_out)
      hcheck(hipMemcpyAsync(b_out, f.b, sizeof(double) * (size_t)f.nnode * f.ne, hipMemcpyDeviceToHost, f.stream),
             "copy");
    hcheck(hipStreamSynchronize(f.stream), "sync");
    return PCFD_OK;
  });
}

int pcfd_fr_timestep(pcfd_fr* fp, double cfl, double* dt_out, double* dtmin) {
  return guarded(fp, [&](pcfd_fr& f) -> int {
    if (int rc = need_params(f)) return rc;
    const double m = run_timestep(f, cfl, true);
    if (dtmin) *dtmin = m;
    if (dt_out)
      hcheck(hipMemcpyAsync(dt_out, f.dt, sizeof(double) * f.nnode, hipMemcpyDeviceToHost, f.stream), "copy");
    hcheck(hipStreamSynchronize(f.stream), "sync");
    return PCFD_OK;
  });
}

int pcfd_fr_set_timestep(pcfd_fr* fp, const double* dt) {
  return guarded(fp, [&](pcfd_fr& f) -> int {
    if (!dt) return PCFD_ERR_ARG;
    hcheck(hipMemcpyAsync(f.dt, dt, sizeof(double) * f.nnode, hipMemcpyHostToDevice, f.stream), "copy");
    hcheck(hipStreamSynchronize(f.stream), "sync");
    return PCFD_OK;
  });
}

int pcfd_fr_jacobian(pcfd_fr* fp) {
  return guarded(fp, [&](pcfd_fr& f) -> int {
    if (int rc = need_params(f)) return rc;
    run_jacobian(f);
    hcheck(hipStreamSynchronize(f.stream), "sync");
    return PCFD_OK;
  });
}

int pcfd_fr_download_jacobian(pcfd_fr* fp, const int* ia, const int* ja, double* M) {
  return guarded(fp, [&](pcfd_fr& f) -> int {
    if (int rc = need_params(f)) return rc;
    if (!ia || !ja || !M) return PCFD_ERR_ARG;
    if (f.lu_done) {
      f.err = "pcfd_fr_download_jacobian: the diagonal blocks are factored (call pcfd_fr_jacobian again)";
      return PCFD_ERR_STATE;
    }
    const size_t NE2 = (size_t)f.ne * f.ne;
    const size_t noff = (size_t)f.nslot * kLane * NE2, ndg = (size_t)f.nchunk * kLane * NE2;
    std::vector<double> off(noff), dg(ndg);
    if (fp32(f)) {
      std::vector<float> a(noff), c(ndg);
      hcheck(hipMemcpy(a.data(), f.Mf, sizeof(float) * noff, hipMemcpyDeviceToHost), "copy");
      hcheck(hipMemcpy(c.data(), f.Df, sizeof(float) * ndg, hipMemcpyDeviceToHost), "copy");
      std::copy(a.begin(), a.end(), off.begin());
      std::copy(c.begin(), c.end(), dg.begin());
    } else {
      hcheck(hipMemcpy(off.data(), f.Md, sizeof(double) * noff, hipMemcpyDeviceToHost), "copy");
      hcheck(hipMemcpy(dg.data(), f.Dd, sizeof(double) * ndg, hipMemcpyDeviceToHost), "copy");
    }
    for (int i = 0; i < f.nnode; ++i) {
      if (ia[i + 1] - ia[i] != f.h_nptr[i + 1] - f.h_nptr[i] + 1) {
        f.err = "pcfd_fr_download_jacobian: row " + std::to_string(i) + " has a different block count";
        return PCFD_ERR_ARG;
      }
      const int r = f.h_spos[i];
      for (int k = ia[i]; k < ia[i + 1]; ++k) {
        const int j = ja[k];
        double* out = M + (size_t)k * NE2;
        if (j == i) {
          const size_t base = (size_t)(r / kLane) * NE2 * kLane + r % kLane;
          for (size_t c = 0; c < NE2; ++c) out[c] = dg[base + c * kLane];
          continue;
        }
        int p = f.h_nptr[i];
        const int pe = f.h_nptr[i + 1];
        for (; p < pe; ++p) {
          const int e = f.h_nslot[p] >> 1, side = f.h_nslot[p] & 1;
          if (f.h_edges[2 * (size_t)e + (1 - side)] == j) break;
        }
        if (p == pe) {
          f.err = "pcfd_fr_download_jacobian: (" + std::to_string(i) + "," + std::to_string(j) + ") is not a block";
          return PCFD_ERR_ARG;
        }
        const long long base = f.h_eoff[2 * (size_t)(f.h_nslot[p] >> 1) + (f.h_nslot[p] & 1)];
        for (size_t c = 0; c < NE2; ++c) out[c] = off[(size_t)base + c * kLane];
      }
    }
    return PCFD_OK;
  });
}

int pcfd_fr_pattern(pcfd_fr* fp, int* ia, int* ja) {
  return guarded(fp, [&](pcfd_fr& f) -> int {
    if (!ia || !ja) return PCFD_ERR_ARG;
    int k = 0;
    for (int i = 0; i < f.nnode; ++i) {
      ia[i] = k;
      ja[k++] = i;
      for (int p = f.h_nptr[i]; p < f.h_nptr[i + 1]; ++p) {
        const int e = f.h_nslot[p] >> 1, side = f.h_nslot[p] & 1;
        ja[k++] = f.h_edges[2 * (size_t)e + (1 - side)];
      }
    }
    ia[f.nnode] = k;
    return PCFD_OK;
  });
}

int pcfd_fr_sgs(pcfd_fr* fp, int nsgs, double* x_out, double* delta) {
  return guarded(fp, [&](pcfd_fr& f) -> int {
    if (int rc = need_params(f)) return rc;
    if (nsgs < 0) return PCFD_ERR_ARG;
    if (!f.lu_done) run_prepare(f);
    run_sgs(f, nsgs, delta);
    if (x_out) {
      const size_t n = (size_t)f.nnode * f.ne;
      if (fp32(f)) {
        std::vector<float> h(n);
        hcheck(hipMemcpyAsync(h.data(), f.xf, sizeof(float) * n, hipMemcpyDeviceToHost, f.stream), "copy");
        hcheck(hipStreamSynchronize(f.stream), "sync");
        std::copy(h.begin(), h.end(), x_out);
      } else {
        hcheck(hipMemcpyAsync(x_out, f.xd, sizeof(double) * n, hipMemcpyDeviceToHost, f.stream), "copy");
      }
    }
    hcheck(hipStreamSynchronize(f.stream), "sync");
    return PCFD_OK;
  });
}

int pcfd_fr_sgs_order(pcfd_fr* fp, int* order) {
  return guarded(fp, [&](pcfd_fr& f) -> int {
    if (!order) return PCFD_ERR_ARG;
    int k = 0;
    for (int r : f.h_srow)
      if (r >= 0) order[k++] = r;
    return PCFD_OK;
  });
}

int pcfd_fr_apply_dq(pcfd_fr* fp, int* nzeroed) {
  return guarded(fp, [&](pcfd_fr& f) -> int {
    if (int rc = need_params(f)) return rc;
    const int nz = run_apply(f);
    if (nzeroed) *nzeroed = nz;
    return PCFD_OK;
  });
}

int pcfd_fr_init_iterations(pcfd_fr* fp) {
  return guarded(fp, [&](pcfd_fr& f) -> int {
    if (int rc = need_params(f)) return rc;
    // PreIterate (ucs/solutionSpace.tcc:509-569): the counter starts at 0,
    // the first time step at iteration 0's CFL
    f.iter = 0;
    f.cfl = cfl_of(f.prm, 0);
    run_timestep(f, f.cfl, false);
    hcheck(hipStreamSynchronize(f.stream), "sync");
    return PCFD_OK;
  });
}

int pcfd_fr_step(pcfd_fr* fp, double* norms, int* nzeroed) {
  return guarded(fp, [&](pcfd_fr& f) -> int {
    if (int rc = need_params(f)) return rc;
    step(f, norms, nzeroed);
    return PCFD_OK;
  });
}

int pcfd_fr_solve(pcfd_fr* fp, int maxiter, double drop, double* history, int* iters) {
  return guarded(fp, [&](pcfd_fr& f) -> int {
    if (int rc = need_params(f)) return rc;
    if (maxiter < 0) return PCFD_ERR_ARG;
    double r0 = 0.0;
    int it = 0;
    for (; it < maxiter;) {
      const double r = step(f, nullptr, nullptr);
      if (history) history[it] = r;
      if (it == 0) r0 = r;
      ++it;
      if (!(r > drop * r0)) break;
    }
    if (iters) *iters = it;
    return PCFD_OK;
  });
}

int pcfd_fr_set_surrogate(pcfd_fr* fp, pcfd_mlp* mlp, const pcfd_surrogate_desc* d) {
  return guarded(fp, [&](pcfd_fr& f) -> int {
    if (mlp && !d) {
      f.err = "pcfd_fr_set_surrogate: a surrogate needs its descriptor";
      return PCFD_ERR_ARG;
    }
    f.mlp = mlp;
    if (!mlp) return PCFD_OK;
    f.sdesc = *d;
    if (!f.s_rhoi) {
      const size_t npass = ((size_t)f.nnode + 63) / 64;
      dalloc(f, &f.s_rhoi, (size_t)f.nnode * f.ns);
      dalloc(f, &f.s_T, (size_t)f.nnode);
      dalloc(f, &f.s_w, (size_t)f.nnode * f.ns);
      dalloc(f, &f.s_lst, npass * 64);
      dalloc(f, &f.s_cnt, npass);
      dalloc(f, &f.s_off, npass + 1);
    }
    f.s_calls = f.s_fallback_total = 0;
    f.s_ms = 0.0;
    return PCFD_OK;
  });
}

int pcfd_fr_surrogate_stats(const pcfd_fr* f, int* last_fallback, long long* calls, long long* fallback_total,
                            double* ms_total) {
  if (!f) return PCFD_ERR_ARG;
  if (last_fallback) *last_fallback = f->s_last_fallback;
  if (calls) *calls = f->s_calls;
  if (fallback_total) *fallback_total = f->s_fallback_total;
  if (ms_total) *ms_total = f->s_ms;
  return PCFD_OK;
}

int pcfd_fr_phase_times(pcfd_fr* f, double* ms, int reset) {
  if (!f) return PCFD_ERR_ARG;
  if (ms)
    for (int k = 0; k < PCFD_FR_NPHASE; ++k) ms[k] = f->phase_ms[k];
  if (reset)
    for (int k = 0; k < PCFD_FR_NPHASE; ++k) f->phase_ms[k] = 0.0;
  return PCFD_OK;
}

}  // extern "C"
