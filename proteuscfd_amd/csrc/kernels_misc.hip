// Solution update: NaN/Inf rows of the correction zeroed, then ApplyDQ
// (ucs/solutionSpace.tcc:772-804, ucs/compressible.tcc:928-993), and the
// explicit update (ExplicitSolve, ucs/solve.tcc:70-101).
#include <stdexcept>

#include "common.h"
#include "physics.h"

namespace pcfd {

template <typename T>
__global__ void __launch_bounds__(kBlock) k_apply_dq(int nnode, int ntot, double gamma, T* __restrict__ x,
                                                    double* __restrict__ q, int* __restrict__ count) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nnode) return;
  double dq[5], Q[5];
  bool zero = false;
  T* xi = x + (long long)i * 5;
#pragma unroll
  for (int k = 0; k < 5; ++k) {
    dq[k] = (double)xi[k];
    if (__builtin_isnan(dq[k]) || __builtin_isinf(dq[k])) zero = true;
    Q[k] = q[k * ntot + i];
  }
  if (zero) {
#pragma unroll
    for (int k = 0; k < 5; ++k) {
      dq[k] = 0.0;
      xi[k] = (T)0;
    }
    atomicAdd(count, 1);
  }
  apply_dq(dq, Q, gamma);
#pragma unroll
  for (int k = 0; k < 5; ++k) q[k * ntot + i] = Q[k];
}

// ExplicitSolve x = b*dt/vol with ApplyDQ, then NewtonIterate's update, which
// applies the same x a second time (the reference calls ApplyDQ in both).
__global__ void __launch_bounds__(kBlock) k_explicit(int nnode, int ntot, double gamma, const double* __restrict__ b,
                                                    const double* __restrict__ dt, const double* __restrict__ vol,
                                                    double* __restrict__ q, int* __restrict__ count) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nnode) return;
  double dq[5], Q[5];
#pragma unroll
  for (int k = 0; k < 5; ++k) {
    dq[k] = b[k * nnode + i] * dt[i] / vol[i];
    Q[k] = q[k * ntot + i];
  }
  apply_dq(dq, Q, gamma);
  bool zero = false;
#pragma unroll
  for (int k = 0; k < 5; ++k)
    if (__builtin_isnan(dq[k]) || __builtin_isinf(dq[k])) zero = true;
  if (zero) {
#pragma unroll
    for (int k = 0; k < 5; ++k) dq[k] = 0.0;
    atomicAdd(count, 1);
  }
  apply_dq(dq, Q, gamma);
#pragma unroll
  for (int k = 0; k < 5; ++k) q[k * ntot + i] = Q[k];
}

__global__ void k_zero_int(int* p) { *p = 0; }

void launch_apply_dq(Ctx& c, int explicit_mode, int* nzeroed) {
  hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(1), 0, c.stream, c.d.icount);
  PCFD_HIP(hipGetLastError());
  if (explicit_mode) {
    hipLaunchKernelGGL(k_explicit, dim3(nblk(c.nnode)), dim3(kBlock), 0, c.stream, c.nnode, c.ntot, c.prm.gamma,
                       c.d.b, c.d.dt, c.d.vol, c.d.q, c.d.icount);
  } else if (c.prm.precision == PCFD_CHOSEN_FP32) {
    hipLaunchKernelGGL(k_apply_dq<float>, dim3(nblk(c.nnode)), dim3(kBlock), 0, c.stream, c.nnode, c.ntot,
                       c.prm.gamma, c.d.xf, c.d.q, c.d.icount);
  } else {
    hipLaunchKernelGGL(k_apply_dq<double>, dim3(nblk(c.nnode)), dim3(kBlock), 0, c.stream, c.nnode, c.ntot,
                       c.prm.gamma, c.d.xd, c.d.q, c.d.icount);
  }
  PCFD_HIP(hipGetLastError());
  if (nzeroed) {
    PCFD_HIP(hipMemcpyAsync(nzeroed, c.d.icount, sizeof(int), hipMemcpyDeviceToHost, c.stream));
    PCFD_HIP(hipStreamSynchronize(c.stream));
  }
}

}  // namespace pcfd
