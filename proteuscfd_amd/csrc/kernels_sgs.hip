// Multicolour symmetric block Gauss-Seidel (CRS::SGS, ucs/crs.tcc:61-173).
//
// The reference sweeps rows 0..n-1 then n-1..0 with the newest x.  Rows of
// one colour share no off-diagonal block, so sweeping colour by colour is
// exactly Gauss-Seidel in the colour-sorted row order (and its reverse on the
// way back); within a colour all rows update in parallel.  Blocks are read
// from the SELL layout (one lane per row, the 64 rows of a chunk read one
// 512-byte (FP64) / 256-byte (FP32) line per block entry), x is kept as
// 5-wide AoS (x[i*5+k]): a neighbour's values are 40 (FP64) / 20 (FP32)
// consecutive bytes, and the lines the gathers pull in carry no padding.
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

// b (node-order SoA, FP64) -> SELL row order [chunk][eq][lane] in type T
template <typename T>
__global__ void __launch_bounds__(kBlock) k_pack_b(int nrow_pad, int nnode, const int* __restrict__ srow,
                                                  const double* __restrict__ b, T* __restrict__ bs) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nrow_pad) return;
  const int i = srow[r];
  T* o = bs + (long long)(r >> 6) * 5 * kW + (r & 63);
#pragma unroll
  for (int k = 0; k < 5; ++k) o[k * kW] = (i >= 0) ? (T)b[k * nnode + i] : (T)0;
}

// Two lanes per row: lanes l and l+32 of a wavefront share a row and take
// alternate slots, so a colour launch has twice as many waves in flight;
// their partial sums are combined with one fixed exchange (both lanes end
// with the same value) before lane l solves the diagonal block.  A
// wavefront covers 32 consecutive SELL rows, so each block-entry load of a
// lane half reads 32 consecutive rows' values: a full 128-byte line in FP32.
// (Measured on the 1.03M-node bump: 2 lanes/row beats 4 by 4% in FP32 and
// 1.4% in FP64, and 8 lanes/row is 25% slower.)
constexpr int kTPR = 2;

// Block-matrix entries are read once per sweep.  FP64 blocks are loaded
// non-temporally (they then do not evict the gathered x lines from L2); for
// FP32 blocks the plain load measured faster.
template <typename T>
__device__ __forceinline__ T mload(const T* p) {
  if constexpr (sizeof(T) == 8) return __builtin_nontemporal_load(p);
  else return *p;
}

template <typename T>
__global__ void __launch_bounds__(kBlock) k_sgs_color(int row0, int nrows, const int* __restrict__ srow,
                                                     const int* __restrict__ swidth, const int* __restrict__ sbase,
                                                     const int* __restrict__ scol, const T* __restrict__ Moff,
                                                     const T* __restrict__ Dg, const int* __restrict__ piv,
                                                     const T* __restrict__ bs, T* __restrict__ x) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int sub = (threadIdx.x & 63) >> 5;
  const int local = (t >> 6) * 32 + (threadIdx.x & 31);
  const bool inr = local < nrows;
  const int r = row0 + (inr ? local : 0);
  const int i = inr ? srow[r] : -1;
  const int chunk = r >> 6, lane = r & 63;
  T acc[5] = {0, 0, 0, 0, 0};
  if (i >= 0) {
    const int w = swidth[chunk];
    const long long cb = (long long)sbase[chunk] * kW + lane;
    const T* mb = Moff + (long long)sbase[chunk] * 25 * kW + lane;
    // the next slot's column index is loaded one iteration ahead, so this
    // slot's x and block loads issue without waiting for it
    int jn = (sub < w) ? scol[cb + (long long)sub * kW] : -1;
    for (int s = sub; s < w; s += kTPR) {
      const int j = jn;
      if (j < 0) break;
      jn = (s + kTPR < w) ? scol[cb + (long long)(s + kTPR) * kW] : -1;
      const T* xj = x + (long long)j * 5;
      const T x0 = xj[0], x1 = xj[1], x2 = xj[2], x3 = xj[3], x4 = xj[4];
      const T* blk = mb + (long long)s * 25 * kW;
#pragma unroll
      for (int rr = 0; rr < 5; ++rr) {
        T v = mload(blk + (rr * 5 + 0) * kW) * x0;
        v += mload(blk + (rr * 5 + 1) * kW) * x1;
        v += mload(blk + (rr * 5 + 2) * kW) * x2;
        v += mload(blk + (rr * 5 + 3) * kW) * x3;
        v += mload(blk + (rr * 5 + 4) * kW) * x4;
        acc[rr] += v;
      }
    }
  }
#pragma unroll
  for (int k = 0; k < 5; ++k) {
    acc[k] += __shfl_xor(acc[k], 32);
  }
  if (i < 0 || sub != 0) return;
  T rhs[5];
  const T* bb = bs + (long long)chunk * 5 * kW + lane;
#pragma unroll
  for (int k = 0; k < 5; ++k) rhs[k] = bb[k * kW] - acc[k];
  const T* dg = Dg + (long long)chunk * 25 * kW + lane;
  T A[25];
#pragma unroll
  for (int k = 0; k < 25; ++k) A[k] = mload(dg + k * kW);
  lu_solve5(A, piv[r], rhs);
  T* xo = x + (long long)i * 5;
#pragma unroll
  for (int k = 0; k < 5; ++k) xo[k] = rhs[k];
}

template <typename T>
__global__ void __launch_bounds__(kBlock) k_zero(long long n, T* __restrict__ x) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = (T)0;
}

template <typename T>
__global__ void __launch_bounds__(kBlock) k_sumsq_x(int nnode, const T* __restrict__ x, double* __restrict__ red) {
  __shared__ double sh[kBlock];
  double s = 0.0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nnode; i += gridDim.x * blockDim.x) {
#pragma unroll
    for (int k = 0; k < 5; ++k) {
      double v = (double)x[(long long)i * 5 + k];
      s += v * v;
    }
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
  hipLaunchKernelGGL(k_sumsq_x<T>, dim3(nb), dim3(kBlock), 0, c.stream, c.nnode, x, c.d.red);
  PCFD_HIP(hipGetLastError());
  double h[nb];
  PCFD_HIP(hipMemcpyAsync(h, c.d.red, sizeof(h), hipMemcpyDeviceToHost, c.stream));
  PCFD_HIP(hipStreamSynchronize(c.stream));
  double s = 0.0;
  for (int k = 0; k < nb; ++k) s += h[k];
  return std::sqrt(s) / (double)(5LL * c.nnode);
}

template <typename T>
static void sgs_impl(Ctx& c, int nsgs, const T* Moff, const T* Dg, T* bs, T* x, double* delta) {
  long long n5 = 5LL * c.nnode;
  hipLaunchKernelGGL(k_zero<T>, dim3(nblk(n5)), dim3(kBlock), 0, c.stream, n5, x);
  PCFD_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_pack_b<T>, dim3(nblk(c.nrow_pad)), dim3(kBlock), 0, c.stream, c.nrow_pad, c.nnode, c.d.srow,
                     c.d.b, bs);
  PCFD_HIP(hipGetLastError());
  double xold = 0.0, xn = 0.0;
  // A colour pass directly after a pass over the same colour (the turn of
  // each symmetric sweep, and the start of the next forward sweep) reads the
  // same neighbour x and writes the same values bit for bit: the rows of a
  // colour never read each other.  Those passes are skipped.
  int last = -1;
  for (int it = 0; it < nsgs; ++it) {
    sub_begin(c);
    for (int dir = 0; dir < 2; ++dir) {
      for (int cc = 0; cc < c.ncolor; ++cc) {
        const int col = dir == 0 ? cc : c.ncolor - 1 - cc;
        const int row0 = c.color_chunk[col] * kW;
        const int nrows = (c.color_chunk[col + 1] - c.color_chunk[col]) * kW;
        if (nrows == 0 || col == last) continue;
        last = col;
        hipLaunchKernelGGL(k_sgs_color<T>, dim3(nblk((long long)nrows * kTPR)), dim3(kBlock), 0, c.stream, row0,
                           nrows, c.d.srow, c.d.swidth, c.d.sbase, c.d.scol, Moff, Dg, c.d.piv, bs, x);
        PCFD_HIP(hipGetLastError());
      }
    }
    sub_end(c, 9);
    if (delta && it >= nsgs - 2) {
      xold = xn;
      xn = xnorm(c, x);
    }
  }
  if (delta) *delta = std::fabs(xold - xn);
}

void launch_sgs(Ctx& c, int nsgs, double* delta) {
  if (c.prm.precision == PCFD_CHOSEN_FP32)
    sgs_impl<float>(c, nsgs, c.d.Mf, c.d.Df, c.d.bsf, c.d.xf, delta);
  else
    sgs_impl<double>(c, nsgs, c.d.Md, c.d.Dd, c.d.bsd, c.d.xd, delta);
}

}  // namespace pcfd
