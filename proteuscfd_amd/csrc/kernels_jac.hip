// First-order finite-difference block Jacobian and diagonal LU.
//
// Reference: ComputeSpatialJacobian (ucs/jacobian.tcc:129-210) scatters per
// edge A[R][L] += -dF/dQL and A[L][R] += dF/dQR (Kernel_NumJac :253-304),
// adds boundary blocks to the diagonal with the boundary state re-evaluated
// under perturbation (Bkernel_NumJac :458-535), then sums each column's
// off-diagonal blocks onto the diagonal (Kernel_Diag_NumJac :436-456), and
// ContributeTemporalTerms adds vol/dt (:213-250).
//
// Here one thread owns one block row of the SELL matrix (threads run in
// SELL row order, so 64 lanes write 64 consecutive addresses): it writes
// the row's off-diagonal blocks and accumulates its own diagonal (boundary
// part first, then edges in ascending neighbour order, then the temporal
// term), so every block has a single writer.  Storage type T is double
// (plain) or float (chosen form); the differences are always taken in FP64.
#include <stdexcept>

#include "common.h"
#include "physics.h"

namespace pcfd {

struct JacArgs {
  int nnode, ntot, ntag, nbnode, nrow_pad, sym2d;
  double gamma;
  const int *gwidth, *gbase, *gcol;
  const double* gnrm;
  long long gstride;
  const int *he_ptr, *he_id, *he_tag;
  const double* he_a;
  const int* bc_type;
  const double* qref;
  const double *vol, *dt;
  double* q;
  const int *srow, *sbase;
  const double* xyz;       // SoA, for the viscous directional weights
  const double* twall;     // per tag, noSlip walls
  const int* he_nrm;       // per half-edge, wall-normal neighbour (noSlip)
  int viscous;
  ViscParams vp;
  int flux_alg;            // 1: algebraic flux (generic form only)
  int centred;             // 1: centred differences, jacobianFieldType 1 (generic form only)
  int bcentred;            // 1: centred boundary differences, jacobianBoundaryType 1
};

// One-sided differences with h = 1e-8 (ucs/jacobian.tcc:268): the quotient
// is taken as a product with 1/h = 1e8, which is exact in binary64, so it
// differs from the division by h at most in the last bit.
constexpr double kInvH = 1.0e8;
// centred differences (Kernel_NumJac_Centered, ucs/jacobian.tcc:307-367):
// the quotient by 2h as a product with 1/(2h) = 5e7, exact in binary64
constexpr double kInv2H = 5.0e7;

// Five lanes per block row: lane c of a row's group owns perturbation
// column c -- it writes column c of every block of the row and accumulates
// column c of the diagonal.  Each block entry is produced by exactly the
// same finite difference as a one-thread-per-row evaluation, with a fifth
// of the live state per lane.  A wavefront holds 12 rows (lanes 60-63 idle).
__device__ __forceinline__ void perturb(const double* Q, int c, double h, double* QP) {
#pragma unroll
  for (int k = 0; k < 5; ++k) QP[k] = Q[k] + (k == c ? h : 0.0);
}

// Boundary half-edges (Bkernel_NumJac), sequential per boundary node, five
// lanes per node as in k_jacobian.  Writes the node's boundary part of the
// diagonal block to Db[t*25 + r*5 + c] (t = boundary-node index) and the
// re-evaluated phantom states (lane 0).
//
// With A.bcentred (Bkernel_NumJac_Centered, ucs/jacobian.tcc:538-640, at the
// default jacobianBoundaryEval 0) column c is (F(q+h e_c, BC(q+h e_c)) -
// F(q-h e_c, QR)) / 2h: the +h side re-evaluates the boundary state, the -h
// side keeps the unperturbed one QR.  bnd_down() forms that -h flux.
template <bool ALG>
__device__ __forceinline__ void bnd_down(const double* QL, const double* QR, int c, const double* a, double gamma,
                                         int sym2d, double* fD) {
  double QPD[5];
  perturb(QL, c, -1.0e-8, QPD);
  num_flux<ALG>(QPD, QR, a, gamma, sym2d, fD);
}
template <bool ALG>
__global__ void __launch_bounds__(kBlock) k_jac_boundary(JacArgs A, const int* __restrict__ bnd_nodes, int nbnd,
                                                        double* __restrict__ Db, double* __restrict__ Qb) {
  const int lane64 = threadIdx.x & 63;
  if (lane64 >= 60) return;
  const int t = ((blockIdx.x * blockDim.x + threadIdx.x) >> 6) * 12 + lane64 / 5;
  const int c = lane64 % 5;
  if (t >= nbnd) return;
  const int i = bnd_nodes[t];
  const double h = 1.0e-8;
  const double gamma = A.gamma;
  const bool cen = A.bcentred != 0;
  const double w = cen ? kInv2H : kInvH;
  double Dc[5] = {0, 0, 0, 0, 0};
  const int p0 = A.he_ptr[i], p1 = A.he_ptr[i + 1];
  // A node whose half-edges are all of types that leave its own state alone
  // (no noSlip / Dirichlet re-set) sees the same unperturbed phantom state
  // and flux in its five lanes: at every fifth half-edge lane c evaluates
  // them for half-edge p + c (and writes that phantom), and each lane reads
  // half-edge p's from lane (p - p0) % 5 of its group.
  bool plain = true;
  for (int p = p0; p < p1; ++p) {
    const int type = A.bc_type[A.he_tag[A.he_id[p]]];
    if (type == BC_NOSLIP || is_dirichlet(type)) plain = false;
  }
  if (plain) {
    double QL[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) QL[k] = A.q[k * A.ntot + i];
    const int g0 = lane64 - c;
    double fSg[5] = {0, 0, 0, 0, 0}, QRg[5] = {0, 0, 0, 0, 0};
    for (int p = p0; p < p1; ++p) {
      const int k5 = (p - p0) % 5;
      if (k5 == 0 && p + c < p1) {
        const int hc = A.he_id[p + c];
        const int tc = A.he_tag[hc];
        const double ac[4] = {A.he_a[hc], A.he_a[A.nbnode + hc], A.he_a[2 * A.nbnode + hc],
                              A.he_a[3 * A.nbnode + hc]};
        double Qr[5];
#pragma unroll
        for (int k = 0; k < 5; ++k) {
          QRg[k] = A.q[k * A.ntot + A.nnode + hc];
          Qr[k] = A.qref[k * A.ntag + tc];
        }
        if (bc_state(A.bc_type[tc], QL, QRg, Qr, A.qref[5 * A.ntag + tc], ac, gamma)) {
#pragma unroll
          for (int k = 0; k < 5; ++k) A.q[k * A.ntot + A.nnode + hc] = QRg[k];
        }
        if (!cen) num_flux<ALG>(QL, QRg, ac, gamma, A.sym2d, fSg);
      }
      double fS[5], QR[5];
#pragma unroll
      for (int k = 0; k < 5; ++k) {
        fS[k] = __shfl(fSg[k], g0 + k5);
        QR[k] = __shfl(QRg[k], g0 + k5);
      }
      const int hh = A.he_id[p];
      const int tag = A.he_tag[hh];
      const double a[4] = {A.he_a[hh], A.he_a[A.nbnode + hh], A.he_a[2 * A.nbnode + hh], A.he_a[3 * A.nbnode + hh]};
      double Qr[5];
#pragma unroll
      for (int k = 0; k < 5; ++k) Qr[k] = A.qref[k * A.ntag + tag];
      double QPL[5];
      perturb(QL, c, h, QPL);
      double QPR[5] = {QR[0], QR[1], QR[2], QR[3], QR[4]};
      bc_state(A.bc_type[tag], QPL, QPR, Qr, A.qref[5 * A.ntag + tag], a, gamma);
      double fL[5];
      num_flux<ALG>(QPL, QPR, a, gamma, A.sym2d, fL);
      if (cen) bnd_down<ALG>(QL, QR, c, a, gamma, A.sym2d, fS);
#pragma unroll
      for (int rr = 0; rr < 5; ++rr) Dc[rr] += (fL[rr] - fS[rr]) * w;
    }
    if (c == 0) {
#pragma unroll
      for (int k = 0; k < 5; ++k) Qb[(long long)t * 5 + k] = QL[k];
    }
  } else {
    double QL[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) QL[k] = A.q[k * A.ntot + i];
    for (int p = p0; p < p1; ++p) {
      const int hh = A.he_id[p];
      const int tag = A.he_tag[hh];
      const int type = A.bc_type[tag];
      const int ph = A.nnode + hh;
      const double a[4] = {A.he_a[hh], A.he_a[A.nbnode + hh], A.he_a[2 * A.nbnode + hh],
                           A.he_a[3 * A.nbnode + hh]};
      double QR[5], Qr[5];
#pragma unroll
      for (int k = 0; k < 5; ++k) {
        QR[k] = A.q[k * A.ntot + ph];
        Qr[k] = A.qref[k * A.ntag + tag];
      }
      const double pref = A.qref[5 * A.ntag + tag];
      if (type == BC_NOSLIP) {
        // CalculateBoundaryVariables hard-sets the wall node for the base
        // and for every perturbed state (ucs/jacobian.tcc:484-512)
        const double tw = A.twall[tag];
        const int nn = A.he_nrm[hh];
        if (tw < 0.0 && nn < 0) continue;
        double nQ[5] = {0, 0, 0, 0, 0};
        if (tw < 0.0) {
#pragma unroll
          for (int k = 0; k < 5; ++k) nQ[k] = A.q[k * A.ntot + nn];
        }
        noslip_state(QL, QR, tw, nQ, gamma);
        if (c == 0) {
#pragma unroll
          for (int k = 0; k < 5; ++k) A.q[k * A.ntot + ph] = QR[k];
        }
        double fS[5], fL[5], QPL[5];
        if (cen) bnd_down<ALG>(QL, QR, c, a, gamma, A.sym2d, fS);
        else num_flux<ALG>(QL, QR, a, gamma, A.sym2d, fS);
        perturb(QL, c, h, QPL);
        double QPR[5] = {QR[0], QR[1], QR[2], QR[3], QR[4]};
        noslip_state(QPL, QPR, tw, nQ, gamma);
        num_flux<ALG>(QPL, QPR, a, gamma, A.sym2d, fL);
#pragma unroll
        for (int rr = 0; rr < 5; ++rr) Dc[rr] += (fL[rr] - fS[rr]) * w;
        continue;
      }
      if (is_dirichlet(type)) {
        // QL = QR = Qref and every perturbed evaluation resets both to Qref,
        // so the one-sided contribution is exactly zero; the centred one keeps
        // its -h side: (F(Qref, Qref) - F(Qref - h e_c, Qref)) / 2h
#pragma unroll
        for (int k = 0; k < 5; ++k) {
          QL[k] = Qr[k];
          if (c == 0) A.q[k * A.ntot + ph] = Qr[k];
        }
        if (cen) {
          double fU[5], fD[5];
          num_flux<ALG>(Qr, Qr, a, gamma, A.sym2d, fU);
          bnd_down<ALG>(Qr, Qr, c, a, gamma, A.sym2d, fD);
#pragma unroll
          for (int rr = 0; rr < 5; ++rr) Dc[rr] += (fU[rr] - fD[rr]) * w;
        }
        continue;
      }
      if (bc_state(type, QL, QR, Qr, pref, a, gamma) && c == 0) {
#pragma unroll
        for (int k = 0; k < 5; ++k) A.q[k * A.ntot + ph] = QR[k];
      }
      double fS[5];
      if (cen) bnd_down<ALG>(QL, QR, c, a, gamma, A.sym2d, fS);
      else num_flux<ALG>(QL, QR, a, gamma, A.sym2d, fS);
      double QPL[5];
      perturb(QL, c, h, QPL);
      double QPR[5] = {QR[0], QR[1], QR[2], QR[3], QR[4]};
      bc_state(type, QPL, QPR, Qr, pref, a, gamma);
      double fL[5];
      num_flux<ALG>(QPL, QPR, a, gamma, A.sym2d, fL);
#pragma unroll
      for (int rr = 0; rr < 5; ++rr) Dc[rr] += (fL[rr] - fS[rr]) * w;
    }
    // the node state after the re-evaluation: the viscous Jacobian, which
    // the reference runs after Bkernel_NumJac, sees it
    if (c == 0) {
#pragma unroll
      for (int k = 0; k < 5; ++k) Qb[(long long)t * 5 + k] = QL[k];
    }
  }
#pragma unroll
  for (int rr = 0; rr < 5; ++rr) Db[(long long)t * 25 + rr * 5 + c] = Dc[rr];
}

// the edge flux of the generic form, evaluated from node i's side (x is
// node i's state); a is the canonical n0 -> n1 normal, `left` whether i is n0
__device__ __forceinline__ void jac_flux(const JacArgs& A, const double* x, const double* y, const double* ai,
                                         const double* a, bool left, double* F) {
  if (A.flux_alg) alg_flux(x, y, ai, A.gamma, F);
  else if (roe_flux_fast(x, y, ai, A.gamma, A.sym2d, F) && !left) roe_flux_canonical(y, x, a, A.gamma, A.sym2d, F);
}

// GENERIC = false: Roe flux, one-sided differences (the reference defaults),
// with the own node's Roe-side quantities formed once per row.
// GENERIC = true: the flux (Roe or algebraic) and the difference form
// (one-sided or centred) chosen at run time, each evaluation from scratch.
template <typename T, bool VISC, bool GENERIC>
__global__ void __launch_bounds__(kBlock) k_jacobian(JacArgs A, const int* __restrict__ bidx,
                                                    const double* __restrict__ Db, const double* __restrict__ Qb,
                                                    T* __restrict__ Moff, T* __restrict__ Dg) {
  const int lane64 = threadIdx.x & 63;
  if (lane64 >= 60) return;
  const int r = ((blockIdx.x * blockDim.x + threadIdx.x) >> 6) * 12 + lane64 / 5;
  const int c = lane64 % 5;
  if (r >= A.nrow_pad) return;
  const int i = A.srow[r];
  if (i < 0) return;
  const int chunk = r >> 6, lane = r & 63;
  const double h = 1.0e-8;
  const double gamma = A.gamma;
  double Dc[5] = {0, 0, 0, 0, 0};  // column c of the diagonal block
  const int bt = bidx[i];
  if (bt >= 0) {
    // boundary contributions come first, as Bdriver runs before Kernel_Diag_NumJac
#pragma unroll
    for (int rr = 0; rr < 5; ++rr) Dc[rr] = Db[(long long)bt * 25 + rr * 5 + c];
  }
  double qi[5], xi[3] = {0.0, 0.0, 0.0};
#pragma unroll
  for (int k = 0; k < 5; ++k) qi[k] = A.q[k * A.ntot + i];
  if (VISC) {
    xi[0] = A.xyz[i];
    xi[1] = A.xyz[A.nnode + i];
    xi[2] = A.xyz[2 * A.nnode + i];
  }
  // --- interior edges (Kernel_NumJac): row blocks + diagonal column sums ---
  // Each edge flux is evaluated from this node's side: G(x, y) = F(x, y, n_i)
  // with n_i the face normal pointing out of node i.  The Roe flux with the
  // Harten-Hyman fix is antisymmetric, F(L, R, a) = -F(R, L, -a), so the
  // reference's F(L, R, a) is +G when i is the left node and -G when it is
  // the right node, and the two signs cancel in what the row needs:
  //   diagonal column c:   -A[j][i] column c = (G(q_i + h e_c, q_j) - G(q_i, q_j)) / h
  //   off-diagonal column c:  A[i][j] column c = (G(q_i, q_j + h e_c) - G(q_i, q_j)) / h
  // (the same one-sided differences as the reference's (L,R) evaluation up
  // to the rounding of the flux).  The own node's states, unperturbed and
  // perturbed in component c, are formed once per row.
  // (with centred differences: dcol = (G(q_i + h e_c, q_j) - G(q_i - h e_c, q_j)) / 2h,
  //  offc = (G(q_i, q_j + h e_c) - G(q_i, q_j - h e_c)) / 2h)
  double qiP[5], qiM[5];
  perturb(qi, c, h, qiP);
  perturb(qi, c, -h, qiM);
  RoeSide S0, SP;
  if constexpr (!GENERIC) {
    roe_side_state(qi, gamma, S0);
    roe_side_state(qiP, gamma, SP);
  }
  const long long gb = (long long)A.gbase[i >> 6] * kW + (i & 63);
  const int w = A.gwidth[i >> 6];
  T* ob = Moff + (long long)A.sbase[chunk] * 25 * kW + lane + (long long)c * kW;
  const int g0 = lane64 - c;  // first lane of this row's group of five
  double fSg[5] = {0, 0, 0, 0, 0};
  for (int s = 0; s < w; ++s) {
    const long long gi = gb + (long long)s * kW;
    const int j = A.gcol[gi];
    if (j < 0) break;
    const bool left = i < j;
    const double a[4] = {A.gnrm[gi], A.gnrm[A.gstride + gi], A.gnrm[2 * A.gstride + gi], A.gnrm[3 * A.gstride + gi]};
    const double sg = left ? 1.0 : -1.0;
    const double ai[4] = {sg * a[0], sg * a[1], sg * a[2], a[3]};  // out of node i
    double qj[5], qjP[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) qj[k] = A.q[k * A.ntot + j];
    perturb(qj, c, h, qjP);
    T* blk = ob + (long long)s * 25 * kW;
    // offc: column c of this row's block A[i][j]; dcol: this edge's part of
    // column c of the diagonal block
    double offc[5], dcol[5];
    if constexpr (!GENERIC) {
      // The unperturbed flux G(q_i, q_j) is the same for the row's five
      // lanes: at every fifth slot, lane c evaluates it for slot s + c, and
      // each lane then takes slot s's value from lane s % 5 of its group.
      // (Inviscid rows only: with the viscous columns the extra registers
      // cost more than the four flux evaluations they save.)
      if (!VISC && s % 5 == 0) {
        const int sc = s + c;
        const long long gic = gb + (long long)sc * kW;
        const int jc = sc < w ? A.gcol[gic] : -1;
#pragma unroll
        for (int k = 0; k < 5; ++k) fSg[k] = 0.0;
        if (jc >= 0) {
          const double sgc = i < jc ? 1.0 : -1.0;
          const double aic[4] = {sgc * A.gnrm[gic], sgc * A.gnrm[A.gstride + gic], sgc * A.gnrm[2 * A.gstride + gic],
                                 A.gnrm[3 * A.gstride + gic]};
          double qc[5];
#pragma unroll
          for (int k = 0; k < 5; ++k) qc[k] = A.q[k * A.ntot + jc];
          RoeSide Sc = S0, Sjc;
          roe_side_th(Sc, aic);
          roe_side(qc, aic, gamma, Sjc);
          if (roe_flux_s(qi, Sc, qc, Sjc, aic, gamma, A.sym2d, fSg) && i > jc) {
            const double ac[4] = {-aic[0], -aic[1], -aic[2], aic[3]};
            roe_flux_canonical(qc, qi, ac, gamma, A.sym2d, fSg);
          }
        }
      }
      roe_side_th(S0, ai);
      roe_side_th(SP, ai);
      RoeSide Sj, SjP;
      roe_side(qj, ai, gamma, Sj);
      double fS[5], fP[5];
      if constexpr (VISC) {
        if (roe_flux_s(qi, S0, qj, Sj, ai, gamma, A.sym2d, fS) && !left)
          roe_flux_canonical(qj, qi, a, gamma, A.sym2d, fS);
      } else {
#pragma unroll
        for (int k = 0; k < 5; ++k) fS[k] = __shfl(fSg[k], g0 + s % 5);
      }
      __builtin_amdgcn_sched_barrier(0);
      if (roe_flux_s(qiP, SP, qj, Sj, ai, gamma, A.sym2d, fP) && !left)
        roe_flux_canonical(qj, qiP, a, gamma, A.sym2d, fP);
#pragma unroll
      for (int rr = 0; rr < 5; ++rr) dcol[rr] = (fP[rr] - fS[rr]) * kInvH;
      __builtin_amdgcn_sched_barrier(0);
      roe_side(qjP, ai, gamma, SjP);
      if (roe_flux_s(qi, S0, qjP, SjP, ai, gamma, A.sym2d, fP) && !left)
        roe_flux_canonical(qjP, qi, a, gamma, A.sym2d, fP);
#pragma unroll
      for (int rr = 0; rr < 5; ++rr) offc[rr] = (fP[rr] - fS[rr]) * kInvH;
      __builtin_amdgcn_sched_barrier(0);
    } else {
      double fA[5], fB[5];
      if (A.centred) {
        double qjM[5];
        perturb(qj, c, -h, qjM);
        jac_flux(A, qiP, qj, ai, a, left, fA);
        jac_flux(A, qiM, qj, ai, a, left, fB);
#pragma unroll
        for (int rr = 0; rr < 5; ++rr) dcol[rr] = (fA[rr] - fB[rr]) * kInv2H;
        jac_flux(A, qi, qjP, ai, a, left, fA);
        jac_flux(A, qi, qjM, ai, a, left, fB);
#pragma unroll
        for (int rr = 0; rr < 5; ++rr) offc[rr] = (fA[rr] - fB[rr]) * kInv2H;
      } else {
        jac_flux(A, qi, qj, ai, a, left, fB);
        jac_flux(A, qiP, qj, ai, a, left, fA);
#pragma unroll
        for (int rr = 0; rr < 5; ++rr) dcol[rr] = (fA[rr] - fB[rr]) * kInvH;
        jac_flux(A, qi, qjP, ai, a, left, fA);
#pragma unroll
        for (int rr = 0; rr < 5; ++rr) offc[rr] = (fA[rr] - fB[rr]) * kInvH;
      }
    }
    if (VISC) {
      // Kernel_Viscous_Jac (ucs/jacobian.tcc:734-767): A[R][L] += aL,
      // A[L][R] += aR, added before the column sums; it runs after the
      // boundary Jacobian has hard-set wall / Dirichlet nodes, so boundary
      // nodes enter with their re-evaluated states
      const double xj[3] = {A.xyz[j], A.xyz[A.nnode + j], A.xyz[2 * A.nnode + j]};
      double vi[5], vj[5];
      const int bj = bidx[j];
#pragma unroll
      for (int k = 0; k < 5; ++k) {
        vi[k] = bt >= 0 ? Qb[(long long)bt * 5 + k] : qi[k];
        vj[k] = bj >= 0 ? Qb[(long long)bj * 5 + k] : qj[k];
      }
      double vL[5], vR[5];
      visc_jac_cols(c, left ? vi : vj, left ? vj : vi, left ? xi : xj, left ? xj : xi, a, A.vp, vL, vR);
#pragma unroll
      for (int rr = 0; rr < 5; ++rr) {
        offc[rr] += left ? vR[rr] : vL[rr];
        dcol[rr] -= left ? vL[rr] : vR[rr];
      }
      __builtin_amdgcn_sched_barrier(0);
    }
#pragma unroll
    for (int rr = 0; rr < 5; ++rr) {
      blk[rr * 5 * kW] = (T)offc[rr];
      Dc[rr] += dcol[rr];
    }
  }
  // --- temporal term, steady pseudo-time: A_ii += 1*vol/dtau ---
  const double vt = 1.0 * A.vol[i] / A.dt[i];
#pragma unroll
  for (int rr = 0; rr < 5; ++rr)
    if (rr == c) Dc[rr] += vt;
  T* dg = Dg + (long long)chunk * 25 * kW + lane + (long long)c * kW;
#pragma unroll
  for (int rr = 0; rr < 5; ++rr) dg[rr * 5 * kW] = (T)Dc[rr];
}

// Dirichlet / sonic-inflow and noSlip overwrites of boundary node states,
// deferred until all rows have read the pre-update states (the reference
// applies them inside Bkernel_NumJac after Kernel_NumJac has run): the
// states k_jac_boundary ended with are copied into q.
__global__ void __launch_bounds__(kBlock) k_commit_hardset(int nbnd, int ntot, const int* __restrict__ bnd_nodes,
                                                          const double* __restrict__ Qb, double* __restrict__ q) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nbnd) return;
  const int i = bnd_nodes[t];
#pragma unroll
  for (int k = 0; k < 5; ++k) q[k * ntot + i] = Qb[(long long)t * 5 + k];
}

// Bkernel_BC_Jac_Modify -> ModifyViscousWallJacobian (ucs/bc.tcc:747-900,
// ucs/compressible.tcc:1577-1606) after the temporal terms: per noSlip
// half-edge, BlankSubRow (ucs/crsmatrix.tcc:525-541) zeroes sub-row r of
// every block of the row and puts 1 on the diagonal; isothermal walls then
// tie the energy change to the density change, adiabatic walls tie density
// and energy to the wall-normal neighbour.  One thread per boundary node.
template <typename T>
__device__ __forceinline__ void blank_subrow(T* Moff, T* Dg, const int* scol, long long sb, int w, int chunk,
                                             int lane, int sub) {
  for (int s = 0; s < w; ++s) {
    if (scol[(sb + s) * kW + lane] < 0) break;
    T* blk = Moff + (sb + s) * 25 * kW + lane;
#pragma unroll
    for (int k = 0; k < 5; ++k) blk[(sub * 5 + k) * kW] = (T)0;
  }
  T* dg = Dg + (long long)chunk * 25 * kW + lane;
#pragma unroll
  for (int k = 0; k < 5; ++k) dg[(sub * 5 + k) * kW] = (T)0;
  dg[(sub * 5 + sub) * kW] = (T)1;
}

template <typename T>
__global__ void __launch_bounds__(kBlock) k_noslip_rows(JacArgs A, const int* __restrict__ bnd_nodes, int nbnd,
                                                       const int* __restrict__ spos, const int* __restrict__ swidth,
                                                       const int* __restrict__ scol, T* __restrict__ Moff,
                                                       T* __restrict__ Dg) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nbnd) return;
  const int i = bnd_nodes[t];
  const int r = spos[i];
  const int chunk = r >> 6, lane = r & 63;
  const long long sb = A.sbase[chunk];
  const int w = swidth[chunk];
  const double g = A.gamma;
  for (int p = A.he_ptr[i]; p < A.he_ptr[i + 1]; ++p) {
    const int hh = A.he_id[p];
    const int tag = A.he_tag[hh];
    if (A.bc_type[tag] != BC_NOSLIP) continue;
    const double tw = A.twall[tag];
    blank_subrow(Moff, Dg, scol, sb, w, chunk, lane, 1);
    blank_subrow(Moff, Dg, scol, sb, w, chunk, lane, 2);
    blank_subrow(Moff, Dg, scol, sb, w, chunk, lane, 3);
    if (tw < 0.0) {
      blank_subrow(Moff, Dg, scol, sb, w, chunk, lane, 0);
      blank_subrow(Moff, Dg, scol, sb, w, chunk, lane, 4);
      const int nn = A.he_nrm[hh];
      for (int s = 0; nn >= 0 && s < w; ++s) {
        const int j = scol[(sb + s) * kW + lane];
        if (j < 0) break;
        if (j != nn) continue;
        T* blk = Moff + (sb + s) * 25 * kW + lane;
        blk[0 * kW] = (T)-1.0;
        blk[24 * kW] = (T)-1.0;
      }
    } else {
      const double v2 = 0.0;
      blank_subrow(Moff, Dg, scol, sb, w, chunk, lane, 4);
      Dg[(long long)chunk * 25 * kW + lane + 20 * kW] = (T)(-(tw / (g * (g - 1.0)) + 0.5 * v2));
    }
  }
}

static JacArgs jac_args(Ctx& c) {
  const pcfd_params& p = c.prm;
  return JacArgs{c.nnode, c.ntot, c.ntag, c.nbnode, c.nrow_pad, p.symmetry2d, p.gamma, c.d.gwidth, c.d.gbase,
                 c.d.gcol, c.d.gnrm, c.nslot_g * kW, c.d.he_ptr, c.d.he_id, c.d.he_tag, c.d.he_a, c.d.bc_type,
                 c.d.qref, c.d.vol, c.d.dt, c.d.q, c.d.srow, c.d.sbase, c.d.xyz, c.d.twall, c.d.he_nrm, p.viscous,
                 ViscParams{p.gamma, p.re_tilde, p.prandtl, p.prandtl_t, p.sutherland_s, p.symmetry2d},
                 p.flux_type == PCFD_FLUX_ALGEBRAIC ? 1 : 0, p.jac_field_type == 1 ? 1 : 0,
                 p.jac_boundary_type == 1 ? 1 : 0};
}

static void jac_boundary(Ctx& c, const JacArgs& A, hipStream_t s) {
  if (c.nbnd == 0) return;
  const long long bw = (c.nbnd + 11) / 12;
  hipLaunchKernelGGL(A.flux_alg ? k_jac_boundary<true> : k_jac_boundary<false>, dim3(nblk(bw * kW)), dim3(kBlock),
                     0, s, A, c.d.bnd_nodes, c.nbnd, c.d.jdb, c.d.jqb);
  PCFD_HIP(hipGetLastError());
}

static void jac_commit(Ctx& c, hipStream_t s) {
  if (c.nbnd == 0) return;
  hipLaunchKernelGGL(k_commit_hardset, dim3(nblk(c.nbnd)), dim3(kBlock), 0, s, c.nbnd, c.ntot, c.d.bnd_nodes,
                     c.d.jqb, c.d.q);
  PCFD_HIP(hipGetLastError());
}

static void jac_rows(Ctx& c, const JacArgs& A, hipStream_t s) {
  const pcfd_params& p = c.prm;
  const long long waves = (c.nrow_pad + 11) / 12;  // 12 rows x 5 lanes per wavefront
  const int grid = nblk(waves * kW);
  const bool f32 = p.precision == PCFD_CHOSEN_FP32;
  const bool gen = A.flux_alg || A.centred;
  if (f32) {
    auto kern = p.viscous ? (gen ? k_jacobian<float, true, true> : k_jacobian<float, true, false>)
                          : (gen ? k_jacobian<float, false, true> : k_jacobian<float, false, false>);
    hipLaunchKernelGGL(kern, dim3(grid), dim3(kBlock), 0, s, A, c.d.bidx, c.d.jdb, c.d.jqb, c.d.Mf, c.d.Df);
  } else {
    auto kern = p.viscous ? (gen ? k_jacobian<double, true, true> : k_jacobian<double, true, false>)
                          : (gen ? k_jacobian<double, false, true> : k_jacobian<double, false, false>);
    hipLaunchKernelGGL(kern, dim3(grid), dim3(kBlock), 0, s, A, c.d.bidx, c.d.jdb, c.d.jqb, c.d.Md, c.d.Dd);
  }
  PCFD_HIP(hipGetLastError());
}

static void jac_noslip(Ctx& c, const JacArgs& A, hipStream_t s) {
  if (!c.has_noslip) return;
  if (c.prm.precision == PCFD_CHOSEN_FP32)
    hipLaunchKernelGGL(k_noslip_rows<float>, dim3(nblk(c.nbnd)), dim3(kBlock), 0, s, A, c.d.bnd_nodes, c.nbnd,
                       c.d.spos, c.d.swidth, c.d.scol, c.d.Mf, c.d.Df);
  else
    hipLaunchKernelGGL(k_noslip_rows<double>, dim3(nblk(c.nbnd)), dim3(kBlock), 0, s, A, c.d.bnd_nodes, c.nbnd,
                       c.d.spos, c.d.swidth, c.d.scol, c.d.Md, c.d.Dd);
  PCFD_HIP(hipGetLastError());
}

// In reference order on the library stream: boundary part (phantom states,
// boundary diagonal contributions), rows, then the hard-set commit.
void launch_jacobian(Ctx& c) {
  const JacArgs A = jac_args(c);
  sub_begin(c);
  jac_boundary(c, A, c.stream);
  jac_rows(c, A, c.stream);
  jac_commit(c, c.stream);
  jac_noslip(c, A, c.stream);
  sub_end(c, 11);
}

// The rows on the second stream, overlapping the time step, BCs, gradient,
// limiter and residual that follow on the library stream (they touch
// neither the matrix nor, with the states the rows read kept apart, the rows'
// inputs).  The rows read the states from before the hard-set commit -- a
// snapshot when the case has hard-set nodes, q itself otherwise (then
// nothing on the library stream writes interior states before the join) --
// and the dt array the caller then retires (the new time step goes to the
// other buffer).  jacobian_join() orders the library stream after the rows.
void launch_jacobian_overlapped(Ctx& c) {
  JacArgs A = jac_args(c);
  jac_boundary(c, A, c.stream);
  if (c.has_hardset && c.nbnd > 0) {
    PCFD_HIP(hipMemcpyAsync(c.d.qjac, c.d.q, sizeof(double) * 5 * (size_t)c.ntot, hipMemcpyDeviceToDevice,
                            c.stream));
    A.q = c.d.qjac;
    jac_commit(c, c.stream);
  }
  PCFD_HIP(hipEventRecord(c.ev_fork, c.stream));
  PCFD_HIP(hipStreamWaitEvent(c.stream2, c.ev_fork, 0));
  if (c.timing) PCFD_HIP(hipEventRecord(c.ev4, c.stream2));
  jac_rows(c, A, c.stream2);
  jac_noslip(c, A, c.stream2);
  if (c.timing) PCFD_HIP(hipEventRecord(c.ev5, c.stream2));
  PCFD_HIP(hipEventRecord(c.ev_join, c.stream2));
  c.jac_pending = true;
}

void jacobian_join(Ctx& c) {
  if (!c.jac_pending) return;
  PCFD_HIP(hipStreamWaitEvent(c.stream, c.ev_join, 0));
  c.jac_pending = false;
  if (c.timing) {
    PCFD_HIP(hipEventSynchronize(c.ev5));
    float ms = 0.f;
    PCFD_HIP(hipEventElapsedTime(&ms, c.ev4, c.ev5));
    c.phase_ms[5] += ms;
    c.phase_ms[11] += ms;
  }
}

// In-place LU with partial pivoting of each diagonal block (LU,
// ucs/matrix.h:110-190; PrepareSGS, ucs/crsmatrix.tcc:839-876).  The
// reference leaves rows in place and addresses them through the pivot
// vector; here the pivot rows are swapped physically (compile-time register
// indices, no scratch), which performs the same arithmetic.  Row r of the
// stored factor is original row perm[r]; perm is packed 3 bits per entry.
template <typename T>
__global__ void __launch_bounds__(kBlock) k_prepare_sgs(int nrow_pad, const int* __restrict__ srow, T* __restrict__ Dg,
                                                       int* __restrict__ piv) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nrow_pad || srow[r] < 0) return;
  T* a = Dg + (long long)(r >> 6) * 25 * kW + (r & 63);
  T A[25];
#pragma unroll
  for (int k = 0; k < 25; ++k) A[k] = a[k * kW];
  int p[5] = {0, 1, 2, 3, 4};
  int row = 0;  // persists across columns, like the reference's
#pragma unroll
  for (int c = 0; c < 5; ++c) {
    T large = 0;
#pragma unroll
    for (int j = c; j < 5; ++j) {
      T v = A[j * 5 + c];
      if (fabs(v) > fabs(large)) {
        large = v;
        row = j;
      }
    }
#pragma unroll
    for (int j = c + 1; j < 5; ++j) {
      if (j == row) {
#pragma unroll
        for (int k = 0; k < 5; ++k) {
          T t = A[c * 5 + k];
          A[c * 5 + k] = A[j * 5 + k];
          A[j * 5 + k] = t;
        }
        int t = p[c];
        p[c] = p[j];
        p[j] = t;
      }
    }
    large = (T)1.0 / large;
#pragma unroll
    for (int j = c + 1; j < 5; ++j) A[j * 5 + c] *= large;
#pragma unroll
    for (int j = c + 1; j < 5; ++j)
#pragma unroll
      for (int k = c + 1; k < 5; ++k) A[j * 5 + k] -= A[j * 5 + c] * A[c * 5 + k];
  }
#pragma unroll
  for (int k = 0; k < 25; ++k) a[k * kW] = A[k];
  piv[r] = p[0] | (p[1] << 3) | (p[2] << 6) | (p[3] << 9) | (p[4] << 12);
}

void launch_prepare_sgs(Ctx& c) {
  if (c.prm.precision == PCFD_CHOSEN_FP32)
    hipLaunchKernelGGL(k_prepare_sgs<float>, dim3(nblk(c.nrow_pad)), dim3(kBlock), 0, c.stream, c.nrow_pad, c.d.srow,
                       c.d.Df, c.d.piv);
  else
    hipLaunchKernelGGL(k_prepare_sgs<double>, dim3(nblk(c.nrow_pad)), dim3(kBlock), 0, c.stream, c.nrow_pad,
                       c.d.srow, c.d.Dd, c.d.piv);
  PCFD_HIP(hipGetLastError());
}

}  // namespace pcfd
