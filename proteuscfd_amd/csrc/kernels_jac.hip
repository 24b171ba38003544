// First-order finite-difference block Jacobian and diagonal LU.
//
// Reference: ComputeSpatialJacobian (ucs/jacobian.tcc:129-210) scatters per
// edge A[R][L] += -dF/dQL and A[L][R] += dF/dQR (Kernel_NumJac :253-304),
// adds boundary blocks to the diagonal with the boundary state re-evaluated
// under perturbation (Bkernel_NumJac :458-535), then sums each column's
// off-diagonal blocks onto the diagonal (Kernel_Diag_NumJac :436-456), and
// ContributeTemporalTerms adds vol/dt (:213-250).
//
// Here one thread owns one block row: it writes the row's off-diagonal
// blocks and accumulates its own diagonal (boundary part first, then edges
// in ascending neighbour order, then the temporal term), so every block has
// a single writer.  Storage type T is double (plain) or float (chosen form);
// the finite differences are always evaluated in FP64.
#include <stdexcept>

#include "common.h"
#include "physics.h"

namespace pcfd {

template <typename T>
__global__ void __launch_bounds__(kBlock) k_jacobian(int nnode, int ntot, int ntag, double gamma, int sym2d,
                                                    const int* __restrict__ row_ptr, const int* __restrict__ col,
                                                    const double* __restrict__ nrm, int nnz,
                                                    const int* __restrict__ he_ptr, const int* __restrict__ he_id,
                                                    const int* __restrict__ he_tag, const double* __restrict__ he_a,
                                                    int nbnode, const int* __restrict__ bc_type,
                                                    const double* __restrict__ qref, const double* __restrict__ vol,
                                                    const double* __restrict__ dt, double* __restrict__ q,
                                                    const int* __restrict__ blk_ptr, T* __restrict__ M) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nnode) return;
  const double h = 1.0e-8;
  double D[25];
#pragma unroll
  for (int k = 0; k < 25; ++k) D[k] = 0.0;
  double qi[5];
#pragma unroll
  for (int k = 0; k < 5; ++k) qi[k] = q[k * ntot + i];

  // --- boundary half-edges (Bkernel_NumJac), sequential per node ---
  {
    double QL[5] = {qi[0], qi[1], qi[2], qi[3], qi[4]};
    for (int p = he_ptr[i]; p < he_ptr[i + 1]; ++p) {
      int hh = he_id[p];
      int tag = he_tag[hh];
      int type = bc_type[tag];
      int ph = nnode + hh;
      double a[4] = {he_a[hh], he_a[nbnode + hh], he_a[2 * nbnode + hh], he_a[3 * nbnode + hh]};
      double QR[5], Qr[5];
#pragma unroll
      for (int k = 0; k < 5; ++k) {
        QR[k] = q[k * ntot + ph];
        Qr[k] = qref[k * ntag + tag];
      }
      double pref = qref[5 * ntag + tag];
      if (is_dirichlet(type)) {
        // QL = QR = Qref, and every perturbed evaluation resets both to
        // Qref again, so the reference's contribution is exactly zero
#pragma unroll
        for (int k = 0; k < 5; ++k) {
          QL[k] = Qr[k];
          q[k * ntot + ph] = Qr[k];
        }
        continue;
      }
      bool handled = bc_state(type, QL, QR, Qr, pref, a, gamma);
      if (handled) {
#pragma unroll
        for (int k = 0; k < 5; ++k) q[k * ntot + ph] = QR[k];
      }
      double fS[5];
      roe_flux(QL, QR, a, gamma, sym2d, fS);
#pragma unroll
      for (int c = 0; c < 5; ++c) {
        double QPL[5] = {QL[0], QL[1], QL[2], QL[3], QL[4]};
        QPL[c] += h;
        double QPR[5] = {QR[0], QR[1], QR[2], QR[3], QR[4]};
        bc_state(type, QPL, QPR, Qr, pref, a, gamma);
        double fL[5];
        roe_flux(QPL, QPR, a, gamma, sym2d, fL);
#pragma unroll
        for (int r = 0; r < 5; ++r) D[r * 5 + c] += (fL[r] - fS[r]) / h;
      }
    }
  }
  // --- interior edges (Kernel_NumJac), row i blocks + diagonal column sums ---
  const int b0 = blk_ptr[i];
  for (int p = row_ptr[i]; p < row_ptr[i + 1]; ++p) {
    int j = col[p];
    const bool left = i < j;
    double a[4] = {nrm[p], nrm[nnz + p], nrm[2 * nnz + p], nrm[3 * nnz + p]};
    double qj[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) qj[k] = q[k * ntot + j];
    const double* QL = left ? qi : qj;
    const double* QR = left ? qj : qi;
    double fS[5];
    roe_flux(QL, QR, a, gamma, sym2d, fS);
    T* blk = M + (size_t)(b0 + 1 + (p - row_ptr[i])) * 25;
#pragma unroll
    for (int c = 0; c < 5; ++c) {
      double QPL[5] = {QL[0], QL[1], QL[2], QL[3], QL[4]};
      double QPR[5] = {QR[0], QR[1], QR[2], QR[3], QR[4]};
      QPL[c] += h;
      QPR[c] += h;
      double fL[5], fR[5];
      roe_flux(QPL, QR, a, gamma, sym2d, fL);
      roe_flux(QL, QPR, a, gamma, sym2d, fR);
#pragma unroll
      for (int r = 0; r < 5; ++r) {
        double aRL = (fS[r] - fL[r]) / h;  // A[R][L] = -dF/dQL
        double aLR = (fR[r] - fS[r]) / h;  // A[L][R] = +dF/dQR
        if (left) {
          blk[r * 5 + c] = (T)aLR;
          D[r * 5 + c] += -aRL;
        } else {
          blk[r * 5 + c] = (T)aRL;
          D[r * 5 + c] += -aLR;
        }
      }
    }
  }
  // --- temporal term, steady pseudo-time: A_ii += 1*vol/dtau ---
  const double vt = 1.0 * vol[i] / dt[i];
#pragma unroll
  for (int k = 0; k < 5; ++k) D[k * 5 + k] += vt;
  T* d = M + (size_t)b0 * 25;
#pragma unroll
  for (int k = 0; k < 25; ++k) d[k] = (T)D[k];
}

// Dirichlet / sonic-inflow overwrite of boundary node states, deferred until
// all rows have read the pre-update states (the reference applies it inside
// Bkernel_NumJac after Kernel_NumJac has run).
__global__ void __launch_bounds__(kBlock) k_commit_dirichlet(int nbnd, int ntot, int ntag, const int* __restrict__ bnd_nodes,
                                   const int* __restrict__ he_ptr, const int* __restrict__ he_id,
                                   const int* __restrict__ he_tag, const int* __restrict__ bc_type,
                                   const double* __restrict__ qref, double* __restrict__ q) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nbnd) return;
  int i = bnd_nodes[t];
  int last = -1;
  for (int p = he_ptr[i]; p < he_ptr[i + 1]; ++p) {
    int tag = he_tag[he_id[p]];
    int ty = bc_type[tag];
    if (ty == BC_SONIC_INFLOW || ty == BC_DIRICHLET) last = tag;
  }
  if (last >= 0) {
#pragma unroll
    for (int k = 0; k < 5; ++k) q[k * ntot + i] = qref[k * ntag + last];
  }
}

void launch_jacobian(Ctx& c) {
  const pcfd_params& p = c.prm;
  if (p.precision == PCFD_CHOSEN_FP32) {
    hipLaunchKernelGGL(k_jacobian<float>, dim3(nblk(c.nnode)), dim3(kBlock), 0, c.stream, c.nnode, c.ntot, c.ntag,
                       p.gamma, p.symmetry2d, c.d.row_ptr, c.d.col, c.d.nrm, c.nnz, c.d.he_ptr, c.d.he_id,
                       c.d.he_tag, c.d.he_a, c.nbnode, c.d.bc_type, c.d.qref, c.d.vol, c.d.dt, c.d.q, c.d.blk_ptr,
                       c.d.Mf);
  } else {
    hipLaunchKernelGGL(k_jacobian<double>, dim3(nblk(c.nnode)), dim3(kBlock), 0, c.stream, c.nnode, c.ntot,
                       c.ntag, p.gamma, p.symmetry2d, c.d.row_ptr, c.d.col, c.d.nrm, c.nnz, c.d.he_ptr, c.d.he_id,
                       c.d.he_tag, c.d.he_a, c.nbnode, c.d.bc_type, c.d.qref, c.d.vol, c.d.dt, c.d.q, c.d.blk_ptr,
                       c.d.Md);
  }
  PCFD_HIP(hipGetLastError());
  if (c.nbnd > 0) {
    hipLaunchKernelGGL(k_commit_dirichlet, dim3(nblk(c.nbnd)), dim3(kBlock), 0, c.stream, c.nbnd, c.ntot, c.ntag,
                       c.d.bnd_nodes, c.d.he_ptr, c.d.he_id, c.d.he_tag, c.d.bc_type, c.d.qref, c.d.q);
    PCFD_HIP(hipGetLastError());
  }
}

// In-place LU with partial pivoting of each diagonal block (LU,
// ucs/matrix.h:110-190; PrepareSGS, ucs/crsmatrix.tcc:839-876).  The
// reference leaves rows in place and addresses them through the pivot
// vector; here the pivot rows are swapped physically (compile-time register
// indices, no scratch), which performs the same arithmetic.  Row r of the
// stored factor is original row perm[r]; perm is packed 3 bits per entry.
template <typename T>
__global__ void __launch_bounds__(kBlock) k_prepare_sgs(int nnode, const int* __restrict__ blk_ptr, T* __restrict__ M, int* __restrict__ piv) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nnode) return;
  T* a = M + (size_t)blk_ptr[i] * 25;
  T A[25];
#pragma unroll
  for (int k = 0; k < 25; ++k) A[k] = a[k];
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
    // swap rows c and row (and their permutation entries)
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
  for (int k = 0; k < 25; ++k) a[k] = A[k];
  piv[i] = p[0] | (p[1] << 3) | (p[2] << 6) | (p[3] << 9) | (p[4] << 12);
}

void launch_prepare_sgs(Ctx& c) {
  if (c.prm.precision == PCFD_CHOSEN_FP32)
    hipLaunchKernelGGL(k_prepare_sgs<float>, dim3(nblk(c.nnode)), dim3(kBlock), 0, c.stream, c.nnode, c.d.blk_ptr,
                       c.d.Mf, c.d.piv);
  else
    hipLaunchKernelGGL(k_prepare_sgs<double>, dim3(nblk(c.nnode)), dim3(kBlock), 0, c.stream, c.nnode,
                       c.d.blk_ptr, c.d.Md, c.d.piv);
  PCFD_HIP(hipGetLastError());
}

}  // namespace pcfd
