// Laminar viscous residual and the noSlip residual edits.
//
// Reference: SpatialResidual runs Driver(Kernel_Viscous_Flux) and
// Bdriver(Bkernel_Viscous_Flux) after the inviscid drivers
// (ucs/residual.tcc:85-97, kernels :389-562); Bkernel_BC_Res_Modify then
// hard-sets the rows of noSlip wall nodes (ucs/residual.tcc:40-43,
// ucs/bc.tcc:902-1030, ucs/compressible.tcc:1608-1630) before the norms.
//
// One thread per node continues the node's inviscid sum in the reference's
// order: interior edges in ascending neighbour id, then the node's
// half-edges.  Each edge is evaluated with its arguments in canonical
// (n0 < n1) order, so both ends compute the same flux bit for bit.  Only the
// gradient terms T, u, v, w (terms 5-8) enter the flux; the other terms'
// face corrections are never read and are not formed.
#include <stdexcept>

#include "common.h"
#include "physics.h"

namespace pcfd {

__device__ __forceinline__ void load_g12(const double* __restrict__ grad, int nnode, int i, double* g) {
#pragma unroll
  for (int k = 0; k < 12; ++k) g[k] = grad[(15 + k) * nnode + i];
}

template <bool FLUX>
__global__ void __launch_bounds__(kBlock) k_viscous(int nnode, int ntot, int ntag, int ho, ViscParams vp, Gell g,
                                                   const double* __restrict__ xyz, const double* __restrict__ q,
                                                   const double* __restrict__ grad,
                                                   const int* __restrict__ he_ptr, const int* __restrict__ he_id,
                                                   const int* __restrict__ he_tag, const double* __restrict__ he_a,
                                                   int nbnode, const int* __restrict__ bc_type,
                                                   const double* __restrict__ twall, double* __restrict__ b,
                                                   double* __restrict__ red) {
  const int blk = xcd_block();
  const int i = blk * blockDim.x + threadIdx.x;
  double bi[5] = {0, 0, 0, 0, 0};
  if (i < nnode) {
#pragma unroll
    for (int k = 0; k < 5; ++k) bi[k] = b[k * nnode + i];
    double qi[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) qi[k] = q[k * ntot + i];
    if (FLUX) {
      double ti[4], gi[12], xi[3];
      aux_tuvw(qi, vp.gamma, ti);
      load_g12(grad, nnode, i, gi);
      xi[0] = xyz[i];
      xi[1] = xyz[nnode + i];
      xi[2] = xyz[2 * nnode + i];
      const long long base = gbase_of(g, i);
      const int w = g.width[i >> 6];
      for (int s = 0; s < w; ++s) {
        const long long idx = base + (long long)s * kW;
        const int j = g.col[idx];
        if (j < 0) break;
        const bool left = i < j;
        const double a[4] = {g.nrm[idx], g.nrm[g.stride + idx], g.nrm[2 * g.stride + idx],
                             g.nrm[3 * g.stride + idx]};
        double qj[5], tj[4], gj[12], xj[3];
#pragma unroll
        for (int k = 0; k < 5; ++k) qj[k] = q[k * ntot + j];
        aux_tuvw(qj, vp.gamma, tj);
        load_g12(grad, nnode, j, gj);
        xj[0] = xyz[j];
        xj[1] = xyz[nnode + j];
        xj[2] = xyz[2 * nnode + j];
        const double* qL = left ? qi : qj;
        const double* qR = left ? qj : qi;
        const double* tL = left ? ti : tj;
        const double* tR = left ? tj : ti;
        const double* gL = left ? gi : gj;
        const double* gR = left ? gj : gi;
        const double* xL = left ? xi : xj;
        const double* xR = left ? xj : xi;
        double Qavg[5], ta[4], gf[12];
#pragma unroll
        for (int k = 0; k < 5; ++k) Qavg[k] = (qL[k] + qR[k]) / 2.0;
        aux_tuvw(Qavg, vp.gamma, ta);
#pragma unroll
        for (int k = 0; k < 12; ++k) gf[k] = 0.5 * (gL[k] + gR[k]);
        if (ho) {
          double dx[3], s2 = 0.0;
#pragma unroll
          for (int k = 0; k < 3; ++k) {
            dx[k] = (xR[k] - xL[k]);
            s2 += dx[k] * dx[k];
          }
#pragma unroll
          for (int t = 0; t < 4; ++t) {
            const double qdots = dx[0] * gf[3 * t] + dx[1] * gf[3 * t + 1] + dx[2] * gf[3 * t + 2];
            const double dq = (tR[t] - tL[t] - qdots) / s2;
#pragma unroll
            for (int k = 0; k < 3; ++k) gf[3 * t + k] += dq * dx[k];
          }
        }
        double F[5];
        viscous_flux(Qavg, ta[0], gf, a, vp, F);
#pragma unroll
        for (int k = 0; k < 5; ++k) bi[k] += left ? -F[k] : F[k];
      }
      for (int p = he_ptr[i]; p < he_ptr[i + 1]; ++p) {
        const int h = he_id[p];
        const double a[4] = {he_a[h], he_a[nbnode + h], he_a[2 * nbnode + h], he_a[3 * nbnode + h]};
        double Qavg[5], ta[4], F[5];
#pragma unroll
        for (int k = 0; k < 5; ++k) Qavg[k] = 0.5 * (qi[k] + q[k * ntot + nnode + h]);
        aux_tuvw(Qavg, vp.gamma, ta);
        viscous_flux(Qavg, ta[0], gi, a, vp, F);
#pragma unroll
        for (int k = 0; k < 5; ++k) bi[k] += -F[k];
      }
    }
    // ModifyViscousWallResidual, once per noSlip half-edge (idempotent)
    for (int p = he_ptr[i]; p < he_ptr[i + 1]; ++p) {
      const int tag = he_tag[he_id[p]];
      if (bc_type[tag] != BC_NOSLIP) continue;
      bi[1] = 0.0;
      bi[2] = 0.0;
      bi[3] = 0.0;
      bi[4] = 0.0;
      if (twall[tag] < 0.0) bi[0] = 0.0;
    }
#pragma unroll
    for (int k = 0; k < 5; ++k) b[k * nnode + i] = bi[k];
  }
  block_norm_partials(bi, red, blk);
}

void launch_viscous(Ctx& c, int higher_order) {
  const pcfd_params& p = c.prm;
  if (!p.viscous && !c.has_noslip) return;
  const ViscParams vp{p.gamma, p.re_tilde, p.prandtl, p.prandtl_t, p.sutherland_s, p.symmetry2d};
  const int nb = nblk(c.nnode);
  const Gell g{c.d.gwidth, c.d.gbase, c.d.gcol, c.d.gnrm, c.nslot_g * kW};
  if (p.viscous)
    hipLaunchKernelGGL(k_viscous<true>, dim3(nb), dim3(kBlock), 0, c.stream, c.nnode, c.ntot, c.ntag, higher_order,
                       vp, g, c.d.xyz, c.d.q, c.d.grad, c.d.he_ptr, c.d.he_id, c.d.he_tag, c.d.he_a, c.nbnode,
                       c.d.bc_type, c.d.twall, c.d.b, c.d.red);
  else
    hipLaunchKernelGGL(k_viscous<false>, dim3(nb), dim3(kBlock), 0, c.stream, c.nnode, c.ntot, c.ntag,
                       higher_order, vp, g, c.d.xyz, c.d.q, c.d.grad, c.d.he_ptr, c.d.he_id, c.d.he_tag, c.d.he_a,
                       c.nbnode, c.d.bc_type, c.d.twall, c.d.b, c.d.red);
  PCFD_HIP(hipGetLastError());
  c.nred = nb;
}

}  // namespace pcfd
