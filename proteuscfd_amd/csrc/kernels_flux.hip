// Edge-loop kernels of the residual path, written as node gathers.
//
// The reference loops over edges and scatters both edge contributions into
// node arrays (ucs/driver.tcc:5-67, 274-306).  Here one thread owns one node
// and visits its incident edges in ascending neighbour order -- the same
// order in which the reference's edge loop (edges sorted by (n0, n1))
// reaches that node -- evaluating every edge function with the canonical
// (n0, n1) argument order.  Both ends of an edge therefore see bit-identical
// edge values, sums come out in the reference's order, and no
// floating-point atomics are needed.  Neighbour lists use the gELL layout
// (common.h): the s-th neighbours of 64 consecutive nodes are contiguous.
#include <stdexcept>

#include "common.h"
#include "physics.h"
#include "scan.h"

namespace pcfd {



static Gell gell(const Ctx& c) { return Gell{c.d.gwidth, c.d.gbase, c.d.gcol, c.d.gnrm, c.nslot_g * kW}; }

// ---------------------------------------------------------------------------
// weighted LSQ moments: Kernel_LSQ_CoefficientsWeighted (ucs/gradient.tcc:460-503)
__global__ void __launch_bounds__(kBlock) k_lsq_moments(int nnode, Gell g, const double* __restrict__ xyz,
                                                       double* __restrict__ sw) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nnode) return;
  double s[6] = {0, 0, 0, 0, 0, 0};
  double xi = xyz[i], yi = xyz[nnode + i], zi = xyz[2 * nnode + i];
  const long long b = gbase_of(g, i);
  const int w = g.width[i >> 6];
  for (int k = 0; k < w; ++k) {
    const int j = g.col[b + (long long)k * kW];
    if (j < 0) break;
    double xj = xyz[j], yj = xyz[nnode + j], zj = xyz[2 * nnode + j];
    double dx0, dx1, dx2;
    if (i < j) { dx0 = xi - xj; dx1 = yi - yj; dx2 = zi - zj; }
    else       { dx0 = xj - xi; dx1 = yj - yi; dx2 = zj - zi; }
    double ds2 = 1.0 / (dx0 * dx0 + dx1 * dx1 + dx2 * dx2);
    s[0] += dx0 * dx0 * ds2;
    s[1] += dx0 * dx1 * ds2;
    s[2] += dx0 * dx2 * ds2;
    s[3] += dx1 * dx1 * ds2;
    s[4] += dx1 * dx2 * ds2;
    s[5] += dx2 * dx2 * ds2;
  }
#pragma unroll
  for (int k = 0; k < 6; ++k) sw[k * nnode + i] = s[k];
}

void launch_lsq_moments(Ctx& c) {
  hipLaunchKernelGGL(k_lsq_moments, dim3(nblk(c.nnode)), dim3(kBlock), 0, c.stream, c.nnode, gell(c), c.d.xyz,
                     c.d.sw);
  PCFD_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// UpdateBCs (ucs/bc.tcc:1399-1457): one thread per boundary node walks its
// half-edges in half-edge order, so a Dirichlet/sonic-inflow overwrite of the
// node state is seen by the node's later half-edges exactly as in the
// reference's sequential loop.  Each phantom is written by its owner only.
__global__ void __launch_bounds__(kBlock) k_update_bcs(int nbnd, int nnode, int ntot, int ntag,
                                                      const int* __restrict__ bnd_nodes,
                                                      const int* __restrict__ he_ptr, const int* __restrict__ he_id,
                                                      const int* __restrict__ he_tag,
                                                      const double* __restrict__ he_a, int nbnode,
                                                      const int* __restrict__ bc_type,
                                                      const double* __restrict__ qref,
                                                      const double* __restrict__ twall,
                                                      const int* __restrict__ he_nrm, double gamma,
                                                      double* __restrict__ q) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nbnd) return;
  int i = bnd_nodes[t];
  double QL[5];
#pragma unroll
  for (int k = 0; k < 5; ++k) QL[k] = q[k * ntot + i];
  double Q0[5] = {QL[0], QL[1], QL[2], QL[3], QL[4]};
  for (int p = he_ptr[i]; p < he_ptr[i + 1]; ++p) {
    int h = he_id[p];
    int tag = he_tag[h];
    int ph = nnode + h;
    double a[4] = {he_a[h], he_a[nbnode + h], he_a[2 * nbnode + h], he_a[3 * nbnode + h]};
    double QR[5], Qr[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) {
      QR[k] = q[k * ntot + ph];
      Qr[k] = qref[k * ntag + tag];
    }
    double pref = qref[5 * ntag + tag];
    const int type = bc_type[tag];
    if (type == BC_NOSLIP) {
      // hard-set wall node (adiabatic walls read the wall-normal neighbour
      // as it stands before this update)
      const double tw = twall[tag];
      const int nn = he_nrm[h];
      if (tw < 0.0 && nn < 0) continue;
      double nQ[5] = {0, 0, 0, 0, 0};
      if (tw < 0.0) {
#pragma unroll
        for (int k = 0; k < 5; ++k) nQ[k] = q[k * ntot + nn];
      }
      noslip_state(QL, QR, tw, nQ, gamma);
#pragma unroll
      for (int k = 0; k < 5; ++k) q[k * ntot + ph] = QR[k];
    } else if (is_dirichlet(type)) {
#pragma unroll
      for (int k = 0; k < 5; ++k) {
        QL[k] = Qr[k];
        q[k * ntot + ph] = Qr[k];
      }
    } else if (bc_state(type, QL, QR, Qr, pref, a, gamma)) {
#pragma unroll
      for (int k = 0; k < 5; ++k) q[k * ntot + ph] = QR[k];
    }
  }
#pragma unroll
  for (int k = 0; k < 5; ++k)
    if (QL[k] != Q0[k]) q[k * ntot + i] = QL[k];
}

void launch_update_bcs(Ctx& c) {
  if (c.nbnd == 0) return;
  hipLaunchKernelGGL(k_update_bcs, dim3(nblk(c.nbnd)), dim3(kBlock), 0, c.stream, c.nbnd, c.nnode, c.ntot,
                     c.ntag, c.d.bnd_nodes, c.d.he_ptr, c.d.he_id, c.d.he_tag, c.d.he_a, c.nbnode, c.d.bc_type,
                     c.d.qref, c.d.twall, c.d.he_nrm, c.prm.gamma, c.d.q);
  PCFD_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// gradient terms of the compressible set: rho, rho u, rho v, rho w, E, T, u, v, w
// (GetGradientsLocation, ucs/compressible.tcc:1008-1027)
template <int NT>
__device__ __forceinline__ void grad_terms(const double* Q, double gm1, double gamma, double* t) {
  t[0] = Q[0]; t[1] = Q[1]; t[2] = Q[2]; t[3] = Q[3]; t[4] = Q[4];
  if (NT > 5) {
    const double ir = 1.0 / Q[0];
    const double u = Q[1] * ir, v = Q[2] * ir, w = Q[3] * ir;
    t[5] = gamma * (gm1 * (Q[4] - Q[0] * (0.5 * (u * u + v * v + w * w)))) * ir;
    t[6] = u;
    t[7] = v;
    t[8] = w;
  }
}

// Gradient::Compute, weighted LSQ (ucs/gradient.tcc:56-112, 250-316) and the
// symmetry-plane projection Bkernel_Symmetry_Fix (:544-565).  The per-node
// Gram-Schmidt factors of ComputeLSQCoefficients (:140-168) depend only on
// the node's moments and are formed once per node.
// Limiter::Compute (ucs/limiters.tcc:52-132): neighbour min/max with both
// bounds starting at 0 (:59-63), then the per-edge-end limiter value, the
// minimum over the node's edges.  limit_node is the second part, given the
// node's state, its first five gradient terms and the neighbour bounds.
__device__ __forceinline__ void limit_node(int i, int nnode, int ntot, int type, double K3, double chi,
                                           const Gell& g, const double* __restrict__ xyz,
                                           const double* __restrict__ vol, const double* __restrict__ q,
                                           const double* qi, const double* gi, const double* qmin,
                                           const double* qmax, double* phi) {
#pragma unroll
  for (int k = 0; k < 5; ++k) phi[k] = 1.0;
  const long long b = gbase_of(g, i);
  const int w = g.width[i >> 6];
  double xi[3] = {xyz[i], xyz[nnode + i], xyz[2 * nnode + i]};
  const double ep2 = 6.0 * 3.141592653589793 * vol[i] * K3;
  for (int s = 0; s < w; ++s) {
    const int j = g.col[b + (long long)s * kW];
    if (j < 0) break;
    const bool left = i < j;
    double xj[3] = {xyz[j], xyz[nnode + j], xyz[2 * nnode + j]};
    double dx[3];
    // dx = 0.5 (xR - xL) for the left end, its negation for the right end
    if (left) { dx[0] = 0.5 * (xj[0] - xi[0]); dx[1] = 0.5 * (xj[1] - xi[1]); dx[2] = 0.5 * (xj[2] - xi[2]); }
    else { dx[0] = -(0.5 * (xi[0] - xj[0])); dx[1] = -(0.5 * (xi[1] - xj[1])); dx[2] = -(0.5 * (xi[2] - xj[2])); }
#pragma unroll
    for (int k = 0; k < 5; ++k) {
      double qj = q[k * ntot + j];
      double dQ = left ? (qj - qi[k]) : -(qi[k] - qj);
      double Qh = extrap1(qi[k], dQ, gi[3 * k], gi[3 * k + 1], gi[3 * k + 2], dx, chi, 1.0);
      double t;
      if (type == 3) {
        double DM = Qh - qi[k];
        double DP;
        if (left) DP = (Qh > qi[k]) ? qmax[k] - qi[k] : ((Qh < qi[k]) ? qmin[k] - qi[k] : 0.0);
        else DP = (Qh > qi[k]) ? qmax[k] - qi[k] : qmin[k] - qi[k];
        t = (DP * DP + ep2 + 2.0 * DM * DP) / (DP * DP + 2.0 * DM * DM + DM * DP + ep2);
      } else {
        t = 1.0;
        if (Qh > qi[k]) t = (qmax[k] - qi[k]) / (Qh - qi[k]);
        else if (Qh < qi[k]) t = (qmin[k] - qi[k]) / (Qh - qi[k]);
        if (type == 2) {
          t = (t * t + 2.0 * t) / (t * t + t + 2.0);
        } else {
          t = dmax(0.0, t);
          t = dmin(1.0, t);
        }
      }
      phi[k] = dmin(phi[k], t);
    }
  }
}

struct LimArgs {  // the fused limiter's inputs and output (LIM = true)
  int type;
  double K3, chi;
  const double* vol;
  double* lim2;
};

// With LIM the same kernel also forms the limiter (k_limiter's arithmetic):
// the neighbour bounds come from the states the gradient loop reads, and the
// per-edge limiter loop runs on the gradient still in registers.
template <int NT, bool LIM>
__global__ void __launch_bounds__(kBlock) k_gradient(int nnode, int ntot, Gell g, const double* __restrict__ xyz,
                                                    const double* __restrict__ sw, const double* __restrict__ q,
                                                    const int* __restrict__ he_ptr, const int* __restrict__ he_id,
                                                    const int* __restrict__ he_tag,
                                                    const double* __restrict__ he_a, int nbnode,
                                                    const int* __restrict__ bc_type, double gamma,
                                                    double* __restrict__ grad, LimArgs L) {
  int i = xcd_block() * blockDim.x + threadIdx.x;
  if (i >= nnode) return;
  const double gm1 = gamma - 1.0;
  double gr[NT * 3];
#pragma unroll
  for (int k = 0; k < NT * 3; ++k) gr[k] = 0.0;
  // ComputeLSQCoefficients factors
  const double r11 = sw[i], r12 = sw[nnode + i], r13 = sw[2 * nnode + i];
  const double r12_r11 = (r11 == 0.0) ? 0.0 : r12 / r11;
  const double r22 = sw[3 * nnode + i] - r12 * r12_r11;
  const double r23 = sw[4 * nnode + i] - r12_r11 * r13;
  const double r13_r11 = (r11 == 0.0) ? 0.0 : r13 / r11;
  const double r23_r22 = (r22 == 0.0) ? 0.0 : r23 / r22;
  const double r33 = sw[5 * nnode + i] - r13 * r13_r11 - r23 * r23_r22;
  const double i11 = (r11 == 0.0) ? 0.0 : 1.0 / r11;
  const double i22 = (r22 == 0.0) ? 0.0 : 1.0 / r22;
  const double i33 = (r33 == 0.0) ? 0.0 : 1.0 / r33;
  double Qi[5];
#pragma unroll
  for (int k = 0; k < 5; ++k) Qi[k] = q[k * ntot + i];
  double ti[NT];
  grad_terms<NT>(Qi, gm1, gamma, ti);
  double xi = xyz[i], yi = xyz[nnode + i], zi = xyz[2 * nnode + i];
  double qmin[5] = {0, 0, 0, 0, 0}, qmax[5] = {0, 0, 0, 0, 0};
  const long long b = gbase_of(g, i);
  const int w = g.width[i >> 6];
  for (int s = 0; s < w; ++s) {
    const int j = g.col[b + (long long)s * kW];
    if (j < 0) break;
    double Qj[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) Qj[k] = q[k * ntot + j];
    if constexpr (LIM) {
#pragma unroll
      for (int k = 0; k < 5; ++k) {
        qmax[k] = dmax(qmax[k], Qj[k]);
        qmin[k] = dmin(qmin[k], Qj[k]);
      }
    }
    double tj[NT];
    grad_terms<NT>(Qj, gm1, gamma, tj);
    double xj = xyz[j], yj = xyz[nnode + j], zj = xyz[2 * nnode + j];
    const bool left = i < j;
    double d[3];
    if (left) { d[0] = xi - xj; d[1] = yi - yj; d[2] = zi - zj; }
    else      { d[0] = xj - xi; d[1] = yj - yi; d[2] = zj - zi; }
    double dx2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
    double weight = 1.0 / sqrt(dx2);
    d[0] *= weight; d[1] *= weight; d[2] *= weight;
    if (!left) { d[0] = -d[0]; d[1] = -d[1]; d[2] = -d[2]; }
    const double dykdx = d[1] - r12_r11 * d[0];
    const double we2 = (d[2] - r13_r11 * d[0] - r23_r22 * dykdx) * i33;
    const double we1 = (dykdx - r23 * we2) * i22;
    const double we0 = (d[0] - r12 * we1 - r13 * we2) * i11;
    const double we[3] = {we0, we1, we2};
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      double dq = left ? weight * (tj[t] - ti[t]) : weight * (ti[t] - tj[t]);
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        if (left) gr[3 * t + k] += -we[k] * dq;
        else      gr[3 * t + k] += +we[k] * dq;
      }
    }
  }
  for (int p = he_ptr[i]; p < he_ptr[i + 1]; ++p) {
    int h = he_id[p];
    if (bc_type[he_tag[h]] != BC_SYMMETRY) continue;
    double a0 = he_a[h], a1 = he_a[nbnode + h], a2 = he_a[2 * nbnode + h];
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      double dt = gr[3 * t] * a0 + gr[3 * t + 1] * a1 + gr[3 * t + 2] * a2;
      gr[3 * t] -= dt * a0;
      gr[3 * t + 1] -= dt * a1;
      gr[3 * t + 2] -= dt * a2;
    }
  }
#pragma unroll
  for (int k = 0; k < NT * 3; ++k) grad[k * nnode + i] = gr[k];
  if constexpr (LIM) {
    double phi[5];
    limit_node(i, nnode, ntot, L.type, L.K3, L.chi, g, xyz, L.vol, q, Qi, gr, qmin, qmax, phi);
#pragma unroll
    for (int k = 0; k < 5; ++k) L.lim2[k * nnode + i] = phi[k];
  }
}

// Green-Gauss gradient, gradientType 1 (Kernel_/Bkernel_Green_Gauss_Gradient,
// ucs/gradient.tcc:172-250): each face adds its averaged terms times a*|a|
// to the left node and subtracts them from the right one, every boundary
// half-edge adds the average with its phantom state; the sum is divided by
// the volume and the symmetry projection applied (:68-95).  Gathered per
// node: the contributions arrive in ascending neighbour order (the order of
// the reference's edge loop), then in half-edge order.
template <int NT>
__global__ void __launch_bounds__(kBlock) k_gradient_gg(int nnode, int ntot, Gell g, const double* __restrict__ vol,
                                                       const double* __restrict__ q,
                                                       const int* __restrict__ he_ptr, const int* __restrict__ he_id,
                                                       const int* __restrict__ he_tag,
                                                       const double* __restrict__ he_a, int nbnode,
                                                       const int* __restrict__ bc_type, double gamma,
                                                       double* __restrict__ grad) {
  int i = xcd_block() * blockDim.x + threadIdx.x;
  if (i >= nnode) return;
  const double gm1 = gamma - 1.0;
  double gr[NT * 3];
#pragma unroll
  for (int k = 0; k < NT * 3; ++k) gr[k] = 0.0;
  double Qi[5];
#pragma unroll
  for (int k = 0; k < 5; ++k) Qi[k] = q[k * ntot + i];
  double ti[NT];
  grad_terms<NT>(Qi, gm1, gamma, ti);
  const long long b = gbase_of(g, i);
  const int w = g.width[i >> 6];
  for (int s = 0; s < w; ++s) {
    const long long idx = b + (long long)s * kW;
    const int j = g.col[idx];
    if (j < 0) break;
    double Qj[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) Qj[k] = q[k * ntot + j];
    double tj[NT];
    grad_terms<NT>(Qj, gm1, gamma, tj);
    const double a[4] = {g.nrm[idx], g.nrm[g.stride + idx], g.nrm[2 * g.stride + idx], g.nrm[3 * g.stride + idx]};
    const bool left = i < j;
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      const double f = 0.5 * (ti[t] + tj[t]);
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const double v = f * a[k] * a[3];
        gr[3 * t + k] += left ? v : -v;
      }
    }
  }
  for (int p = he_ptr[i]; p < he_ptr[i + 1]; ++p) {
    const int h = he_id[p];
    double Qp[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) Qp[k] = q[k * ntot + nnode + h];
    double tp[NT];
    grad_terms<NT>(Qp, gm1, gamma, tp);
    const double a[4] = {he_a[h], he_a[nbnode + h], he_a[2 * nbnode + h], he_a[3 * nbnode + h]};
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      const double f = 0.5 * (ti[t] + tp[t]);
#pragma unroll
      for (int k = 0; k < 3; ++k) gr[3 * t + k] += f * a[k] * a[3];
    }
  }
  const double vi = vol[i];
#pragma unroll
  for (int k = 0; k < NT * 3; ++k) gr[k] /= vi;
  for (int p = he_ptr[i]; p < he_ptr[i + 1]; ++p) {
    int h = he_id[p];
    if (bc_type[he_tag[h]] != BC_SYMMETRY) continue;
    double a0 = he_a[h], a1 = he_a[nbnode + h], a2 = he_a[2 * nbnode + h];
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      double dt = gr[3 * t] * a0 + gr[3 * t + 1] * a1 + gr[3 * t + 2] * a2;
      gr[3 * t] -= dt * a0;
      gr[3 * t + 1] -= dt * a1;
      gr[3 * t + 2] -= dt * a2;
    }
  }
#pragma unroll
  for (int k = 0; k < NT * 3; ++k) grad[k * nnode + i] = gr[k];
}

void launch_gradient(Ctx& c, int nterms) {
  if (c.prm.gradient_type == 1) {
    if (nterms == 9)
      hipLaunchKernelGGL(k_gradient_gg<9>, dim3(nblk(c.nnode)), dim3(kBlock), 0, c.stream, c.nnode, c.ntot, gell(c),
                         c.d.vol, c.d.q, c.d.he_ptr, c.d.he_id, c.d.he_tag, c.d.he_a, c.nbnode, c.d.bc_type,
                         c.prm.gamma, c.d.grad);
    else
      hipLaunchKernelGGL(k_gradient_gg<5>, dim3(nblk(c.nnode)), dim3(kBlock), 0, c.stream, c.nnode, c.ntot, gell(c),
                         c.d.vol, c.d.q, c.d.he_ptr, c.d.he_id, c.d.he_tag, c.d.he_a, c.nbnode, c.d.bc_type,
                         c.prm.gamma, c.d.grad);
    PCFD_HIP(hipGetLastError());
    return;
  }
  const LimArgs none{0, 0.0, 0.0, nullptr, nullptr};
  auto kern = nterms == 9 ? k_gradient<9, false> : k_gradient<5, false>;
  hipLaunchKernelGGL(kern, dim3(nblk(c.nnode)), dim3(kBlock), 0, c.stream, c.nnode, c.ntot, gell(c), c.d.xyz,
                     c.d.sw, c.d.q, c.d.he_ptr, c.d.he_id, c.d.he_tag, c.d.he_a, c.nbnode, c.d.bc_type, c.prm.gamma,
                     c.d.grad, none);
  PCFD_HIP(hipGetLastError());
}


// ---------------------------------------------------------------------------
// Writes the pre-clip limiter to lim2.
__global__ void __launch_bounds__(kBlock) k_limiter(int nnode, int ntot, int type, double K3, double chi, Gell g,
                                                   const double* __restrict__ xyz, const double* __restrict__ vol,
                                                   const double* __restrict__ q, const double* __restrict__ grad,
                                                   double* __restrict__ lim2) {
  int i = xcd_block() * blockDim.x + threadIdx.x;
  if (i >= nnode) return;
  double qi[5], qmin[5], qmax[5], phi[5], gi[15];
#pragma unroll
  for (int k = 0; k < 5; ++k) {
    qi[k] = q[k * ntot + i];
    qmin[k] = 0.0;
    qmax[k] = 0.0;
  }
#pragma unroll
  for (int k = 0; k < 15; ++k) gi[k] = grad[k * nnode + i];
  const long long b = gbase_of(g, i);
  const int w = g.width[i >> 6];
  for (int s = 0; s < w; ++s) {
    const int j = g.col[b + (long long)s * kW];
    if (j < 0) break;
#pragma unroll
    for (int k = 0; k < 5; ++k) {
      double v = q[k * ntot + j];
      qmax[k] = dmax(qmax[k], v);
      qmin[k] = dmin(qmin[k], v);
    }
  }
  limit_node(i, nnode, ntot, type, K3, chi, g, xyz, vol, q, qi, gi, qmin, qmax, phi);
#pragma unroll
  for (int k = 0; k < 5; ++k) lim2[k * nnode + i] = phi[k];
}

// BadExtrapolation (ucs/compressible.tcc:1054-1078) with one reciprocal
__device__ __forceinline__ bool bad_state_r(const double* Q, double gamma) {
  const double r = Q[0], ir = 1.0 / r;
  const double u = Q[1] * ir, v = Q[2] * ir, w = Q[3] * ir, E = Q[4];
  const double p = (gamma - 1.0) * (E - r * (0.5 * (u * u + v * v + w * w)));
  return (p < 1.0e-10) || (r < 0.0) || (E < 1.0e-10);
}

// Kernel_PressureClip (ucs/limiters.tcc:736-815) in the reference's edge
// order, then negative limiters -> 0 (:118-124).  The reference walks the
// edges in order and zeroes a node's limiters as soon as one of its
// extrapolations, or the Roe state of the pair, is unphysical, so a later
// edge sees the zeroed values.  An edge can only zero anything when one of
// its ends extrapolates to an unphysical state: the Roe state of two
// physical states is physical (its pressure is at least the smaller end
// pressure), and a zeroed node extrapolates to its own state.  Hence:
//  1. k_clip_flag: every node checks its edges with the pre-clip limiters and
//     is flagged when one of its own extrapolations or a Roe state is
//     unphysical, or its own state is; it writes its clamped pre-clip
//     limiter as the result for the case no edge touches it;
//  2. k_clip_count / k_scan_counts / k_clip_list: the edges with a flagged
//     end, in edge order (per-64-edge ballots and a fixed-order scan: no
//     atomics);
//  3. k_clip_replay: one thread runs the reference's loop over that list on
//     the current limiters.  Edges off the list cannot fire, so the result
//     is the reference's full loop.
__global__ void __launch_bounds__(kBlock) k_clip_flag(int nnode, int ntot, double chi, double gamma, Gell g,
                                                     const double* __restrict__ xyz, const double* __restrict__ q,
                                                     const double* __restrict__ grad,
                                                     const double* __restrict__ lim2, double* __restrict__ lim,
                                                     unsigned char* __restrict__ flag) {
  int i = xcd_block() * blockDim.x + threadIdx.x;
  if (i >= nnode) return;
  // The node's own state, gradient and limiter stay in registers; only the
  // neighbour side is gathered.  Both extrapolations are written relative to
  // node i: dQ = qj - qi and dx = (xj - xi)/2 give bit-identical states to
  // the reference's (L,R) form, because a - b == -(b - a) and 0.5*(-d) ==
  // -(0.5*d) exactly in IEEE arithmetic.
  double qi[5], li[5], gi[15];
#pragma unroll
  for (int k = 0; k < 5; ++k) {
    qi[k] = q[k * ntot + i];
    li[k] = lim2[k * nnode + i];
  }
#pragma unroll
  for (int k = 0; k < 15; ++k) gi[k] = grad[k * nnode + i];
  bool fire = bad_state(qi, gamma);
  const double xi[3] = {xyz[i], xyz[nnode + i], xyz[2 * nnode + i]};
  const long long b = gbase_of(g, i);
  const int w = g.width[i >> 6];
  for (int s = 0; s < w && !fire; ++s) {
    const int j = g.col[b + (long long)s * kW];
    if (j < 0) break;
    const double dx[3] = {0.5 * (xyz[j] - xi[0]), 0.5 * (xyz[nnode + j] - xi[1]), 0.5 * (xyz[2 * nnode + j] - xi[2])};
    const double mdx[3] = {-dx[0], -dx[1], -dx[2]};
    double Qi[5], Qj[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) {
      const double qj = q[k * ntot + j];
      const double dQ = qj - qi[k];
      Qi[k] = extrap1(qi[k], dQ, gi[3 * k], gi[3 * k + 1], gi[3 * k + 2], dx, chi, li[k]);
      Qj[k] = extrap1(qj, -dQ, grad[(3 * k) * nnode + j], grad[(3 * k + 1) * nnode + j],
                      grad[(3 * k + 2) * nnode + j], mdx, chi, lim2[k * nnode + j]);
    }
    if (bad_state(Qi, gamma)) fire = true;
    double Qr[5];
    if (i < j) roe_state(Qi, Qj, gamma, Qr);
    else roe_state(Qj, Qi, gamma, Qr);
    if (bad_state(Qr, gamma)) fire = true;
  }
  flag[i] = fire ? 1 : 0;
#pragma unroll
  for (int k = 0; k < 5; ++k) lim[k * nnode + i] = li[k] < 0.0 ? 0.0 : li[k];
}

// cnt[p] = number of edges p*64 .. p*64+63 with a flagged end
__global__ void __launch_bounds__(kBlock) k_clip_count(int nedge, int npass, const int* __restrict__ edges,
                                                      const unsigned char* __restrict__ flag,
                                                      int* __restrict__ cnt) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int pass = (int)(t >> 6), lane = threadIdx.x & 63;
  const long long e = t;
  bool cand = false;
  if (e < nedge) cand = (flag[edges[2 * e]] | flag[edges[2 * e + 1]]) != 0;
  const unsigned long long m = __ballot(cand);
  if (lane == 0 && pass < npass) cnt[pass] = __popcll(m);
}

// list[off[p] + r] = the r-th flagged edge of pass p (edge order)
__global__ void __launch_bounds__(kBlock) k_clip_list(int nedge, int npass, const int* __restrict__ edges,
                                                     const unsigned char* __restrict__ flag,
                                                     const int* __restrict__ cnt, const int* __restrict__ off,
                                                     int* __restrict__ list) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int pass = (int)(t >> 6), lane = threadIdx.x & 63;
  if (pass >= npass || cnt[pass] == 0) return;  // uniform across the wavefront
  const long long e = t;
  bool cand = false;
  if (e < nedge) cand = (flag[edges[2 * e]] | flag[edges[2 * e + 1]]) != 0;
  const unsigned long long m = __ballot(cand);
  if (cand) list[off[pass] + __popcll(m & ((1ull << lane) - 1ull))] = (int)e;
}

// The reference's PressureClip loop over the listed edges, in order, on the
// current (raw) limiters lim2; zeroed nodes are written to lim2 and lim.
// One thread: the list is empty unless a clip fires, and short when it does.
__global__ void __launch_bounds__(64) k_clip_replay(int nnode, int ntot, int npass, double chi, double gamma,
                                                   const int* __restrict__ edges, const int* __restrict__ off,
                                                   const int* __restrict__ list, const double* __restrict__ xyz,
                                                   const double* __restrict__ q, const double* __restrict__ grad,
                                                   double* lim2, double* lim) {
  if (threadIdx.x != 0) return;
  const int total = off[npass];
  for (int t = 0; t < total; ++t) {
    const int e = list[t];
    const int L = edges[2 * e], R = edges[2 * e + 1];
    double dx[3], mdx[3], dQ[5], QL[5], QR[5];
    for (int d = 0; d < 3; ++d) {
      dx[d] = 0.5 * (xyz[d * nnode + R] - xyz[d * nnode + L]);
      mdx[d] = -dx[d];
    }
    for (int k = 0; k < 5; ++k) {
      dQ[k] = q[k * ntot + R] - q[k * ntot + L];
      QL[k] = extrap1(q[k * ntot + L], dQ[k], grad[(3 * k) * nnode + L], grad[(3 * k + 1) * nnode + L],
                      grad[(3 * k + 2) * nnode + L], dx, chi, lim2[k * nnode + L]);
    }
    bool zl = bad_state(QL, gamma), zr = false;
    if (zl)
      for (int k = 0; k < 5; ++k) lim2[k * nnode + L] = 0.0;
    for (int k = 0; k < 5; ++k)
      QR[k] = extrap1(q[k * ntot + R], -dQ[k], grad[(3 * k) * nnode + R], grad[(3 * k + 1) * nnode + R],
                      grad[(3 * k + 2) * nnode + R], mdx, chi, lim2[k * nnode + R]);
    if (bad_state(QR, gamma)) zr = true;
    double Qr[5];
    roe_state(QL, QR, gamma, Qr);
    if (bad_state(Qr, gamma)) zl = zr = true;
    for (int k = 0; k < 5; ++k) {
      if (zl) lim2[k * nnode + L] = lim[k * nnode + L] = 0.0;
      if (zr) lim2[k * nnode + R] = lim[k * nnode + R] = 0.0;
    }
  }
}

void launch_pressure_clip(Ctx& c) {
  const pcfd_params& p = c.prm;
  hipLaunchKernelGGL(k_clip_flag, dim3(nblk(c.nnode)), dim3(kBlock), 0, c.stream, c.nnode, c.ntot, p.chi, p.gamma,
                     gell(c), c.d.xyz, c.d.q, c.d.grad, c.d.lim2, c.d.lim, c.d.clip_flag);
  PCFD_HIP(hipGetLastError());
  if (c.nedge == 0) return;
  const int npass = (c.nedge + 63) / 64;
  const int nb = nblk((long long)npass * 64);
  hipLaunchKernelGGL(k_clip_count, dim3(nb), dim3(kBlock), 0, c.stream, c.nedge, npass, c.d.edges, c.d.clip_flag,
                     c.d.clip_cnt);
  PCFD_HIP(hipGetLastError());
  launch_scan_counts(c.stream, npass, c.d.clip_cnt, c.d.clip_off);
  hipLaunchKernelGGL(k_clip_list, dim3(nb), dim3(kBlock), 0, c.stream, c.nedge, npass, c.d.edges, c.d.clip_flag,
                     c.d.clip_cnt, c.d.clip_off, c.d.clip_list);
  PCFD_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_clip_replay, dim3(1), dim3(64), 0, c.stream, c.nnode, c.ntot, npass, p.chi, p.gamma,
                     c.d.edges, c.d.clip_off, c.d.clip_list, c.d.xyz, c.d.q, c.d.grad, c.d.lim2, c.d.lim);
  PCFD_HIP(hipGetLastError());
}

void launch_limiter(Ctx& c) {
  const pcfd_params& p = c.prm;
  double K3 = p.venkat_K * p.venkat_K * p.venkat_K;
  hipLaunchKernelGGL(k_limiter, dim3(nblk(c.nnode)), dim3(kBlock), 0, c.stream, c.nnode, c.ntot, p.limiter, K3,
                     p.chi, gell(c), c.d.xyz, c.d.vol, c.d.q, c.d.grad, c.d.lim2);
  PCFD_HIP(hipGetLastError());
  launch_pressure_clip(c);
}

// weighted LSQ gradient and the limiter in one pass, then the pressure clip
void launch_gradient_limiter(Ctx& c, int nterms) {
  const pcfd_params& p = c.prm;
  const LimArgs L{p.limiter, p.venkat_K * p.venkat_K * p.venkat_K, p.chi, c.d.vol, c.d.lim2};
  auto kern = nterms == 9 ? k_gradient<9, true> : k_gradient<5, true>;
  hipLaunchKernelGGL(kern, dim3(nblk(c.nnode)), dim3(kBlock), 0, c.stream, c.nnode, c.ntot, gell(c), c.d.xyz,
                     c.d.sw, c.d.q, c.d.he_ptr, c.d.he_id, c.d.he_tag, c.d.he_a, c.nbnode, c.d.bc_type, p.gamma,
                     c.d.grad, L);
  PCFD_HIP(hipGetLastError());
  launch_pressure_clip(c);
}

// ---------------------------------------------------------------------------
// SpatialResidual (ucs/residual.tcc:65-122): Kernel_Inviscid_Flux (:191-296,
// b[L] -= F, b[R] += F) gathered per node, then Bkernel_Inviscid_Flux
// (:298-387, b[L] -= F on the unextrapolated state vs the phantom).
// Also writes per-block partial sums of b^2 (total + per equation).
template <bool HO, bool ALG>
__global__ void __launch_bounds__(kBlock) k_residual(int nnode, int ntot, double gamma, double chi, int sym2d, Gell g,
                                                    const double* __restrict__ xyz, const double* __restrict__ q,
                                                    const double* __restrict__ grad, const double* __restrict__ lim,
                                                    const int* __restrict__ he_ptr, const int* __restrict__ he_id,
                                                    const double* __restrict__ he_a, int nbnode,
                                                    double* __restrict__ b, double* __restrict__ red) {
  const int blk = xcd_block();
  int i = blk * blockDim.x + threadIdx.x;
  double bi[5] = {0, 0, 0, 0, 0};
  if (i < nnode) {
    double qi[5], gi[15], li[5], xi[3];
#pragma unroll
    for (int k = 0; k < 5; ++k) qi[k] = q[k * ntot + i];
    if (HO) {
#pragma unroll
      for (int k = 0; k < 15; ++k) gi[k] = grad[k * nnode + i];
#pragma unroll
      for (int k = 0; k < 5; ++k) li[k] = lim[k * nnode + i];
      xi[0] = xyz[i]; xi[1] = xyz[nnode + i]; xi[2] = xyz[2 * nnode + i];
    }
    const long long base = gbase_of(g, i);
    const int w = g.width[i >> 6];
    for (int s = 0; s < w; ++s) {
      const long long idx = base + (long long)s * kW;
      const int j = g.col[idx];
      if (j < 0) break;
      const bool left = i < j;
      double a[4] = {g.nrm[idx], g.nrm[g.stride + idx], g.nrm[2 * g.stride + idx], g.nrm[3 * g.stride + idx]};
      double qj[5];
#pragma unroll
      for (int k = 0; k < 5; ++k) qj[k] = q[k * ntot + j];
      double QL[5], QR[5];
      if (HO) {
        // both ends extrapolated relative to node i (dQ = qj - qi, dx =
        // (xj - xi)/2), then assigned to L/R: the same bits as the (L,R)
        // form because a - b == -(b - a) and 0.5*(-d) == -(0.5*d) in IEEE
        // arithmetic, with two selects per variable instead of eight
        const double dx[3] = {0.5 * (xyz[j] - xi[0]), 0.5 * (xyz[nnode + j] - xi[1]),
                              0.5 * (xyz[2 * nnode + j] - xi[2])};
        const double mdx[3] = {-dx[0], -dx[1], -dx[2]};
#pragma unroll
        for (int k = 0; k < 5; ++k) {
          const double dQ = qj[k] - qi[k];
          const double Qi = extrap1(qi[k], dQ, gi[3 * k], gi[3 * k + 1], gi[3 * k + 2], dx, chi, li[k]);
          const double Qj = extrap1(qj[k], -dQ, grad[(3 * k) * nnode + j], grad[(3 * k + 1) * nnode + j],
                                    grad[(3 * k + 2) * nnode + j], mdx, chi, lim[k * nnode + j]);
          QL[k] = left ? Qi : Qj;
          QR[k] = left ? Qj : Qi;
        }
      } else {
#pragma unroll
        for (int k = 0; k < 5; ++k) {
          QL[k] = left ? qi[k] : qj[k];
          QR[k] = left ? qj[k] : qi[k];
        }
      }
      double F[5];
      num_flux<ALG>(QL, QR, a, gamma, sym2d, F);
#pragma unroll
      for (int k = 0; k < 5; ++k) bi[k] += left ? -F[k] : F[k];
    }
    for (int p = he_ptr[i]; p < he_ptr[i + 1]; ++p) {
      int h = he_id[p];
      double a[4] = {he_a[h], he_a[nbnode + h], he_a[2 * nbnode + h], he_a[3 * nbnode + h]};
      double QR[5];
#pragma unroll
      for (int k = 0; k < 5; ++k) QR[k] = q[k * ntot + nnode + h];
      double F[5];
      num_flux<ALG>(qi, QR, a, gamma, sym2d, F);
#pragma unroll
      for (int k = 0; k < 5; ++k) bi[k] += -F[k];
    }
#pragma unroll
    for (int k = 0; k < 5; ++k) b[k * nnode + i] = bi[k];
  }
  block_norm_partials(bi, red, blk);
}

// final reduction of the per-block partial sums (single block, fixed order)
__global__ void __launch_bounds__(kBlock) k_reduce6(int nparts, const double* __restrict__ red,
                                                   double* __restrict__ out) {
  __shared__ double sh[kBlock];
  for (int k = 0; k < 6; ++k) {
    double s = 0.0;
    for (int p = threadIdx.x; p < nparts; p += blockDim.x) s += red[k * nparts + p];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
      if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
      __syncthreads();
    }
    if (threadIdx.x == 0) out[k] = sh[0];
    __syncthreads();
  }
}

void launch_residual(Ctx& c, int ho) {
  const pcfd_params& p = c.prm;
  int nb = nblk(c.nnode);
  sub_begin(c);
  const bool alg = p.flux_type == PCFD_FLUX_ALGEBRAIC;
  auto kern = ho ? (alg ? k_residual<true, true> : k_residual<true, false>)
                 : (alg ? k_residual<false, true> : k_residual<false, false>);
  hipLaunchKernelGGL(kern, dim3(nb), dim3(kBlock), 0, c.stream, c.nnode, c.ntot, p.gamma, p.chi, p.symmetry2d,
                     gell(c), c.d.xyz, c.d.q, c.d.grad, c.d.lim, c.d.he_ptr, c.d.he_id, c.d.he_a, c.nbnode, c.d.b,
                     c.d.red);
  PCFD_HIP(hipGetLastError());
  c.nred = nb;
  launch_viscous(c, ho);
  sub_end(c, 10);
}

// residGlobal = sqrt(sum b^2)/(nnode*neqn), per-equation sqrt(sum b_k^2)/nnode
// (ParallelL2Norm / StridedParallelL2Norm, ucs/parallel.h:120-182)
void launch_norms(Ctx& c, double* host6) {
  hipLaunchKernelGGL(k_reduce6, dim3(1), dim3(kBlock), 0, c.stream, c.nred, c.d.red, c.d.red_out);
  PCFD_HIP(hipGetLastError());
  if (host6) {
    double s[6];
    PCFD_HIP(hipMemcpyAsync(s, c.d.red_out, sizeof(s), hipMemcpyDeviceToHost, c.stream));
    PCFD_HIP(hipStreamSynchronize(c.stream));
    host6[0] = std::sqrt(s[0]) / (double)((long long)c.nnode * 5);
    for (int k = 0; k < 5; ++k) host6[1 + k] = std::sqrt(s[1 + k]) / (double)c.nnode;
  }
}

// ---------------------------------------------------------------------------
// MaxEigenvalue (ucs/compressible.tcc:798-821) with one reciprocal
__device__ __forceinline__ double max_eig_r(const double* Q, const double* a, double gamma) {
  const double rho = Q[0], ir = 1.0 / rho;
  const double u = Q[1] * ir, v = Q[2] * ir, w = Q[3] * ir;
  const double P = (gamma - 1.0) * (Q[4] - rho * (0.5 * (u * u + v * v + w * w)));
  const double c = sqrt(gamma * P * ir);
  const double th = u * a[0] + v * a[1] + w * a[2];
  return dmax(fabs(th + c), fabs(th - c));
}

// ComputeTimesteps, local time stepping (ucs/timestep.tcc:6-47).  With
// vol23 set, the von Neumann limit dt = min(dt, VNN vol^(2/3)) is applied
// from node 1 on, as the reference's loop does (:37-45).
__global__ void __launch_bounds__(kBlock) k_timestep(int nnode, int ntot, double gamma, double cfl, Gell g,
                                                    const double* __restrict__ q, const double* __restrict__ vol,
                                                    const int* __restrict__ he_ptr, const int* __restrict__ he_id,
                                                    const double* __restrict__ he_a, int nbnode,
                                                    const double* __restrict__ vol23, double vnn,
                                                    double* __restrict__ dt) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nnode) return;
  double qi[5];
#pragma unroll
  for (int k = 0; k < 5; ++k) qi[k] = q[k * ntot + i];
  double s = 0.0;
  const long long base = gbase_of(g, i);
  const int w = g.width[i >> 6];
  for (int p = 0; p < w; ++p) {
    const long long idx = base + (long long)p * kW;
    const int j = g.col[idx];
    if (j < 0) break;
    double a[4] = {g.nrm[idx], g.nrm[g.stride + idx], g.nrm[2 * g.stride + idx], g.nrm[3 * g.stride + idx]};
    double t[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) t[k] = 0.5 * (qi[k] + q[k * ntot + j]);
    s += max_eig_r(t, a, gamma) * a[3];
  }
  for (int p = he_ptr[i]; p < he_ptr[i + 1]; ++p) {
    int h = he_id[p];
    double a[4] = {he_a[h], he_a[nbnode + h], he_a[2 * nbnode + h], he_a[3 * nbnode + h]};
    double t[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) t[k] = 0.5 * (qi[k] + q[k * ntot + nnode + h]);
    s += max_eig_r(t, a, gamma) * a[3];
  }
  double d = cfl * (vol[i] / s);
  if (vol23 && i > 0) d = dmin(d, vnn * vol23[i]);
  dt[i] = d;
}

// global time step: every node gets the same value, timeStep when given
// (ucs/timestep.tcc:48-55), else the minimum of the local values
// (:56-74; with the VNN limit that minimum is the same, node 0 excepted
// from the limit in both).  The minimum is reduced in two fixed stages.
__global__ void __launch_bounds__(kBlock) k_fill(int n, const double* __restrict__ src, double value,
                                                double* __restrict__ v) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) v[i] = src ? src[0] : value;
}

__global__ void __launch_bounds__(kBlock) k_min_final(int nparts, const double* __restrict__ red,
                                                     double* __restrict__ out) {
  __shared__ double sh[kBlock];
  double m = red[0];
  for (int p = threadIdx.x; p < nparts; p += blockDim.x) m = dmin(m, red[p]);
  sh[threadIdx.x] = m;
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sh[threadIdx.x] = dmin(sh[threadIdx.x], sh[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = sh[0];
}

__global__ void __launch_bounds__(kBlock) k_min_partial(int n, const double* __restrict__ v, double* __restrict__ red) {
  __shared__ double sh[kBlock];
  double m = 1.0e300;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) m = dmin(m, v[i]);
  sh[threadIdx.x] = m;
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sh[threadIdx.x] = dmin(sh[threadIdx.x], sh[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) red[blockIdx.x] = sh[0];
}

void launch_timestep(Ctx& c, double cfl) {
  const pcfd_params& p = c.prm;
  if (p.global_dt && p.fixed_dt > 0.0) {
    hipLaunchKernelGGL(k_fill, dim3(nblk(c.nnode)), dim3(kBlock), 0, c.stream, c.nnode, nullptr, p.fixed_dt, c.d.dt);
    PCFD_HIP(hipGetLastError());
    return;
  }
  hipLaunchKernelGGL(k_timestep, dim3(nblk(c.nnode)), dim3(kBlock), 0, c.stream, c.nnode, c.ntot, p.gamma, cfl,
                     gell(c), c.d.q, c.d.vol, c.d.he_ptr, c.d.he_id, c.d.he_a, c.nbnode,
                     p.enable_vnn ? c.d.vol23 : nullptr, p.vnn, c.d.dt);
  PCFD_HIP(hipGetLastError());
  if (p.global_dt) {
    const int nb = std::min(nblk(c.nnode), 1024);
    hipLaunchKernelGGL(k_min_partial, dim3(nb), dim3(kBlock), 0, c.stream, c.nnode, c.d.dt, c.d.red);
    PCFD_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_min_final, dim3(1), dim3(kBlock), 0, c.stream, nb, c.d.red, c.d.red_out + 7);
    PCFD_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_fill, dim3(nblk(c.nnode)), dim3(kBlock), 0, c.stream, c.nnode, c.d.red_out + 7, 0.0,
                       c.d.dt);
    PCFD_HIP(hipGetLastError());
  }
}

double launch_dtmin(Ctx& c) {
  int nb = std::min(nblk(c.nnode), 1024);
  hipLaunchKernelGGL(k_min_partial, dim3(nb), dim3(kBlock), 0, c.stream, c.nnode, c.d.dt, c.d.red);
  PCFD_HIP(hipGetLastError());
  std::vector<double> h(nb);
  PCFD_HIP(hipMemcpyAsync(h.data(), c.d.red, sizeof(double) * nb, hipMemcpyDeviceToHost, c.stream));
  PCFD_HIP(hipStreamSynchronize(c.stream));
  double m = h[0];
  for (double v : h) m = v < m ? v : m;
  return m;
}

}  // namespace pcfd
