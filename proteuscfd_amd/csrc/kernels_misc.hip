// Solution update: NaN/Inf rows of the correction zeroed, then ApplyDQ
// (ucs/solutionSpace.tcc:772-804, ucs/compressible.tcc:928-993), and the
// explicit update (ExplicitSolve, ucs/solve.tcc:70-101).
#include <stdexcept>

#include "common.h"
#include "physics.h"
#include "scan.h"

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

// One stage of the four-stage Runge-Kutta extension: Q = Q0 + a dt/vol b
// through ApplyDQ; a row whose correction is NaN/Inf keeps Q0 (counted).
__global__ void __launch_bounds__(kBlock) k_rk_stage(int nnode, int ntot, double gamma, double a,
                                                    const double* __restrict__ b, const double* __restrict__ dt,
                                                    const double* __restrict__ vol, const double* __restrict__ q0,
                                                    double* __restrict__ q, int* __restrict__ count) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nnode) return;
  double dq[5], Q[5];
  bool zero = false;
#pragma unroll
  for (int k = 0; k < 5; ++k) {
    dq[k] = a * (b[k * nnode + i] * dt[i] / vol[i]);
    if (__builtin_isnan(dq[k]) || __builtin_isinf(dq[k])) zero = true;
    Q[k] = q0[k * ntot + i];
  }
  if (zero) {
#pragma unroll
    for (int k = 0; k < 5; ++k) dq[k] = 0.0;
    atomicAdd(count, 1);
  }
  apply_dq(dq, Q, gamma);
#pragma unroll
  for (int k = 0; k < 5; ++k) q[k * ntot + i] = Q[k];
}

void launch_zero_count(Ctx& c) {
  hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(1), 0, c.stream, c.d.icount);
  PCFD_HIP(hipGetLastError());
}

void launch_rk_stage(Ctx& c, double a, const double* q0) {
  hipLaunchKernelGGL(k_rk_stage, dim3(nblk(c.nnode)), dim3(kBlock), 0, c.stream, c.nnode, c.ntot, c.prm.gamma, a,
                     c.d.b, c.d.dt, c.d.vol, q0, c.d.q, c.d.icount);
  PCFD_HIP(hipGetLastError());
}

// Exclusive prefix sum of per-pass counts (one workgroup of 1024, fixed
// order): off[p] = sum of cnt[0..p-1], off[npass] = total.  Used to compact
// lists in order without atomics (the surrogate's fallback cells, the
// pressure clip's candidate edges).
// Counts are read in coalesced tiles of 1024, up to kScanTiles tiles at a
// time so their loads are in flight together; each tile is scanned with
// wavefront shuffles plus one pass over the 16 wavefront totals.
constexpr int kScanBlock = 1024;
constexpr int kScanTiles = 16;

__global__ void __launch_bounds__(kScanBlock) k_scan_counts(int npass, const int* __restrict__ cnt,
                                                                int* __restrict__ off) {
  __shared__ int wsum[kScanBlock / 64];
  __shared__ int carry_sh;
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  if (t == 0) carry_sh = 0;
  __syncthreads();
  for (int base = 0; base < npass; base += kScanBlock * kScanTiles) {
    int v[kScanTiles];
#pragma unroll
    for (int k = 0; k < kScanTiles; ++k) {
      const int p = base + k * kScanBlock + t;
      v[k] = p < npass ? cnt[p] : 0;
    }
#pragma unroll
    for (int k = 0; k < kScanTiles; ++k) {
      // inclusive scan within the wavefront
      int x = v[k];
#pragma unroll
      for (int d = 1; d < 64; d <<= 1) {
        const int y = __shfl_up(x, d);
        if (lane >= d) x += y;
      }
      if (lane == 63) wsum[wv] = x;
      __syncthreads();
      int before = carry_sh;
      for (int u = 0; u < wv; ++u) before += wsum[u];
      const int p = base + k * kScanBlock + t;
      if (p < npass) off[p] = before + x - v[k];
      __syncthreads();
      if (t == kScanBlock - 1) carry_sh = before + x;
      __syncthreads();
    }
  }
  if (t == 0) off[npass] = carry_sh;
}

void launch_scan_counts(hipStream_t stream, int n, const int* cnt, int* off) {
  hipLaunchKernelGGL(k_scan_counts, dim3(1), dim3(kScanBlock), 0, stream, n, cnt, off);
  PCFD_HIP(hipGetLastError());
}


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
