// Multicolour symmetric block Gauss-Seidel (CRS::SGS, ucs/crs.tcc:61-173).
//
// The reference sweeps rows 0..n-1 then n-1..0 with the newest x.  Rows of
// one colour share no off-diagonal block, so sweeping colour by colour is
// exactly Gauss-Seidel in the colour-sorted row order (and its reverse on the
// way back); within a colour all rows update in parallel.
#include <stdexcept>

#include "common.h"
#include "physics.h"

namespace pcfd {

template <typename T>
__device__ __forceinline__ T sel5(const T* v, int k) {
  T r = v[0];
  r = (k == 1) ? v[1] : r;
  r = (k == 2) ? v[2] : r;
  r = (k == 3) ? v[3] : r;
  r = (k == 4) ? v[4] : r;
  return r;
}

// LuSolve on a physically row-permuted factor (ucs/matrix.h:237-264): forward
// sums ascend, backward sums descend, as in the reference.
template <typename T>
__device__ __forceinline__ void lu_solve5(const T* A, int pk, T* b) {
  T y[5];
#pragma unroll
  for (int r = 0; r < 5; ++r) {
    T s = 0;
#pragma unroll
    for (int c = 0; c < r; ++c) s += A[r * 5 + c] * y[c];
    y[r] = sel5(b, (pk >> (3 * r)) & 7) - s;
  }
#pragma unroll
  for (int r = 4; r >= 0; --r) {
    T s = 0;
#pragma unroll
    for (int c = 4; c > r; --c) s += A[r * 5 + c] * b[c];
    b[r] = (y[r] - s) / A[r * 5 + r];
  }
}

template <typename T>
__global__ void __launch_bounds__(kBlock) k_sgs_color(int n, const int* __restrict__ nodes, int nnode,
                                                     const int* __restrict__ row_ptr, const int* __restrict__ col,
                                                     const int* __restrict__ blk_ptr, const T* __restrict__ M,
                                                     const int* __restrict__ piv, const double* __restrict__ b,
                                                     T* __restrict__ x) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int i = nodes[t];
  T rhs[5];
#pragma unroll
  for (int k = 0; k < 5; ++k) rhs[k] = (T)b[k * nnode + i];
  const int b0 = blk_ptr[i] + 1 - row_ptr[i];
  for (int p = row_ptr[i]; p < row_ptr[i + 1]; ++p) {
    const int j = col[p];
    const T* blk = M + (size_t)(b0 + p) * 25;
    T xj[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) xj[k] = x[k * nnode + j];
#pragma unroll
    for (int r = 0; r < 5; ++r) {
      T v = blk[r * 5 + 0] * xj[0];
#pragma unroll
      for (int c = 1; c < 5; ++c) v += blk[r * 5 + c] * xj[c];
      rhs[r] -= v;
    }
  }
  const T* D = M + (size_t)blk_ptr[i] * 25;
  T A[25];
#pragma unroll
  for (int k = 0; k < 25; ++k) A[k] = D[k];
  lu_solve5(A, piv[i], rhs);
#pragma unroll
  for (int k = 0; k < 5; ++k) x[k * nnode + i] = rhs[k];
}

template <typename T>
__global__ void __launch_bounds__(kBlock) k_zero(long long n, T* __restrict__ x) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = (T)0;
}

template <typename T>
__global__ void __launch_bounds__(kBlock) k_sumsq_partial(long long n, const T* __restrict__ x, double* __restrict__ red) {
  __shared__ double sh[kBlock];
  double s = 0.0;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    double v = (double)x[i];
    s += v * v;
  }
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int w = kBlock / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) red[blockIdx.x] = sh[0];
}

template <typename T>
static double xnorm(Ctx& c, const T* x) {
  const int nb = 512;
  long long n = 5LL * c.nnode;
  hipLaunchKernelGGL(k_sumsq_partial<T>, dim3(nb), dim3(kBlock), 0, c.stream, n, x, c.d.red);
  PCFD_HIP(hipGetLastError());
  double h[nb];
  PCFD_HIP(hipMemcpyAsync(h, c.d.red, sizeof(h), hipMemcpyDeviceToHost, c.stream));
  PCFD_HIP(hipStreamSynchronize(c.stream));
  double s = 0.0;
  for (int k = 0; k < nb; ++k) s += h[k];
  return std::sqrt(s) / (double)n;
}

template <typename T>
static void sgs_impl(Ctx& c, int nsgs, const T* M, T* x, double* delta) {
  long long n5 = 5LL * c.nnode;
  hipLaunchKernelGGL(k_zero<T>, dim3(nblk(n5)), dim3(kBlock), 0, c.stream, n5, x);
  PCFD_HIP(hipGetLastError());
  double xold = 0.0, xn = 0.0;
  for (int it = 0; it < nsgs; ++it) {
    for (int dir = 0; dir < 2; ++dir) {
      for (int cc = 0; cc < c.ncolor; ++cc) {
        int col = dir == 0 ? cc : c.ncolor - 1 - cc;
        int beg = c.color_ptr[col], cnt = c.color_ptr[col + 1] - beg;
        if (cnt == 0) continue;
        hipLaunchKernelGGL(k_sgs_color<T>, dim3(nblk(cnt)), dim3(kBlock), 0, c.stream, cnt,
                           c.d.color_nodes + beg, c.nnode, c.d.row_ptr, c.d.col, c.d.blk_ptr, M, c.d.piv, c.d.b, x);
        PCFD_HIP(hipGetLastError());
      }
    }
    if (delta && it >= nsgs - 2) {
      xold = xn;
      xn = xnorm(c, x);
    }
  }
  if (delta) *delta = std::fabs(xold - xn);
}

void launch_sgs(Ctx& c, int nsgs, double* delta) {
  if (c.prm.precision == PCFD_CHOSEN_FP32)
    sgs_impl<float>(c, nsgs, c.d.Mf, c.d.xf, delta);
  else
    sgs_impl<double>(c, nsgs, c.d.Md, c.d.xd, delta);
}

}  // namespace pcfd
