// Chemistry surrogate: a 7 -> 64 -> 64 -> 5 tanh MLP on the matrix cores.
//
// Every matrix product runs on v_mfma_f32_16x16x16_bf16 (BF16 operands, FP32
// accumulation).  A wavefront owns a tile of 16 samples:
//   forward   Z1 = X W1^T (4 MFMA), Z2 = H1 W2^T (16), Y = H2 W3^T (4)
//   backward  dW3 += dY^T H2 (4), dH2 = dY W3 (4), dW2 += dZ2^T H1 (16),
//             dH1 = dZ2 W2 (16), dW1 += dZ1^T X (4)
// with the tile's activations and backward signals staged in LDS (FP32)
// between products, because the MFMA result layout (lane holds 4 rows of
// one column) differs from the operand layout (lane holds 4 consecutive K
// of one row).  Weight gradients accumulate in the MFMA accumulators over
// all tiles a wavefront processes and are written once per wavefront; a
// second kernel sums the per-wavefront partials in a fixed order, so the
// gradient is deterministic.  Adam updates FP32 master weights and
// refreshes the BF16 copies the products read.
//
// MFMA 16x16x16 operand layout (lane l, r = l % 16, q = l / 16):
//   A[m][k]: lane holds A[r][4q .. 4q+3]     B[k][n]: lane holds B[4q .. 4q+3][r]
//   D[m][n]: lane holds D[4q .. 4q+3][r]
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/pcfd_chem.h"
#include "chem_model.h"
#include "scan.h"

namespace pcfd {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kNin = 7, kNh = 64, kNout = 5, kKin = 16, kKout = 16;
// parameter vector layout (FP32)
constexpr int oW1 = 0, oB1 = oW1 + kNh * kKin, oW2 = oB1 + kNh, oB2 = oW2 + kNh * kNh, oW3 = oB2 + kNh,
              oB3 = oW3 + kKout * kNh, kNparam = oB3 + kKout;
static_assert(kNparam == PCFD_MLP_NPARAM, "parameter layout");
constexpr int kMlpBlock = 256;  // 4 wavefronts
constexpr int kWaves = kMlpBlock / 64;

// FP32 -> BF16, round to nearest even (the gfx950 conversion instruction)
__device__ __forceinline__ short f2bf(float f) { return __builtin_bit_cast(short, (__bf16)f); }

// tanh(x) = 1 - 2 / (exp(2x) + 1) on the hardware exp2 and reciprocal
// (relative error ~1e-6, far below the BF16 rounding of the activations)
__device__ __forceinline__ float tanh_fast(float x) {
  const float e = __expf(2.f * x);
  return 1.f - 2.f * __builtin_amdgcn_rcpf(e + 1.f);
}

__device__ __forceinline__ f32x4 mfma(bf16x4 a, bf16x4 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// LDS exchange inside one wavefront: make this lane's LDS writes visible to
// the other lanes of the wavefront before they read.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

struct MlpWeights {
  const short* W1;   // [64][16] bf16
  const short* W2;   // [64][64]
  const short* W3;   // [16][64]
  const float* p;    // FP32 parameters (biases read from here)
};

// per-wavefront LDS tiles
struct WaveTiles {
  float X[16][kKin];
  float H1[16][kNh];
  float H2[16][kNh];
  float D1[16][kNh];  // dZ1
  float D2[16][kNh];  // dZ2
  float DY[16][kKout];
};

__device__ __forceinline__ bf16x4 frag_rows(const float* base, int stride, int row, int k0) {
  // A operand from a row-major FP32 tile: row `row`, columns k0..k0+3
  const float* p = base + row * stride + k0;
  bf16x4 v;
  v[0] = f2bf(p[0]);
  v[1] = f2bf(p[1]);
  v[2] = f2bf(p[2]);
  v[3] = f2bf(p[3]);
  return v;
}

__device__ __forceinline__ bf16x4 frag_cols(const float* base, int stride, int k0, int col) {
  // operand built from 4 consecutive rows k0..k0+3 of one column
  bf16x4 v;
  v[0] = f2bf(base[(k0 + 0) * stride + col]);
  v[1] = f2bf(base[(k0 + 1) * stride + col]);
  v[2] = f2bf(base[(k0 + 2) * stride + col]);
  v[3] = f2bf(base[(k0 + 3) * stride + col]);
  return v;
}

__device__ __forceinline__ bf16x4 wfrag(const short* W, int stride, int row, int k0) {
  const short* p = W + row * stride + k0;
  bf16x4 v;
  v[0] = p[0];
  v[1] = p[1];
  v[2] = p[2];
  v[3] = p[3];
  return v;
}

__device__ __forceinline__ bf16x4 wfrag_cols(const short* W, int stride, int k0, int col) {
  bf16x4 v;
  v[0] = W[(k0 + 0) * stride + col];
  v[1] = W[(k0 + 1) * stride + col];
  v[2] = W[(k0 + 2) * stride + col];
  v[3] = W[(k0 + 3) * stride + col];
  return v;
}

// Forward pass of one 16-sample tile held in T.X; leaves H1, H2 in LDS and
// returns Y in the D layout (lane: Y[4q+j][r]).
__device__ __forceinline__ f32x4 tile_forward(const MlpWeights& w, WaveTiles& T, int r, int q) {
  const f32x4 zero = {0.f, 0.f, 0.f, 0.f};
  const bf16x4 ax = frag_rows(&T.X[0][0], kKin, r, 4 * q);
#pragma unroll
  for (int nt = 0; nt < 4; ++nt) {
    f32x4 acc = mfma(ax, wfrag(w.W1, kKin, nt * 16 + r, 4 * q), zero);
    const float b = w.p[oB1 + nt * 16 + r];
#pragma unroll
    for (int j = 0; j < 4; ++j) T.H1[4 * q + j][nt * 16 + r] = tanh_fast(acc[j] + b);
  }
  wave_sync();
#pragma unroll
  for (int nt = 0; nt < 4; ++nt) {
    f32x4 acc = zero;
#pragma unroll
    for (int kt = 0; kt < 4; ++kt)
      acc = mfma(frag_rows(&T.H1[0][0], kNh, r, kt * 16 + 4 * q), wfrag(w.W2, kNh, nt * 16 + r, kt * 16 + 4 * q),
                 acc);
    const float b = w.p[oB2 + nt * 16 + r];
#pragma unroll
    for (int j = 0; j < 4; ++j) T.H2[4 * q + j][nt * 16 + r] = tanh_fast(acc[j] + b);
  }
  wave_sync();
  f32x4 y = zero;
#pragma unroll
  for (int kt = 0; kt < 4; ++kt)
    y = mfma(frag_rows(&T.H2[0][0], kNh, r, kt * 16 + 4 * q), wfrag(w.W3, kNh, r, kt * 16 + 4 * q), y);
  const float b3 = w.p[oB3 + r];
#pragma unroll
  for (int j = 0; j < 4; ++j) y[j] += b3;
  return y;
}

__device__ __forceinline__ void load_x_tile(WaveTiles& T, const float* __restrict__ x, const int* __restrict__ idx,
                                            int n, int s0, int lane) {
  // 16 samples x 16 features (7 used): 4 entries per lane
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    const int flat = lane * 4 + e;
    const int s = flat >> 4, f = flat & 15;
    const int g = s0 + s;
    float v = 0.f;
    if (g < n && f < kNin) {
      const long long src = idx ? (long long)idx[g] : (long long)g;
      v = x[src * kNin + f];
    }
    T.X[s][f] = v;
  }
}

// ---- inference path: weights held in registers, 64 samples per pass ----
// Products are taken transposed, Z^T = W H^T: the weight tile is the A
// operand and the activations the B operand.  The MFMA result then holds,
// in lane (r, q), features 16mt+4q..+3 of sample r -- exactly the B-operand
// fragment of k-tile mt for the next layer -- so the activations stay in
// registers from layer to layer.  Every lane keeps its A fragments of W1
// (4), W2 (16) and W3 (4) and the biases of the features it produces; a
// pass runs 64 samples as four 16-sample tiles, with LDS used only to
// hand features in and outputs out (each lane builds and post-processes
// one sample).
struct WRegs {
  bf16x4 w1[4], w2[4][4], w3[4];
  float b1[4][4], b2[4][4], b3[4];
};

__device__ __forceinline__ void load_wregs(const MlpWeights& w, int r, int q, WRegs& R) {
#pragma unroll
  for (int mt = 0; mt < 4; ++mt) {
    R.w1[mt] = wfrag(w.W1, kKin, mt * 16 + r, 4 * q);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      R.b1[mt][j] = w.p[oB1 + mt * 16 + 4 * q + j];
      R.b2[mt][j] = w.p[oB2 + mt * 16 + 4 * q + j];
    }
#pragma unroll
    for (int kt = 0; kt < 4; ++kt) R.w2[mt][kt] = wfrag(w.W2, kNh, mt * 16 + r, kt * 16 + 4 * q);
    R.w3[mt] = wfrag(w.W3, kNh, r, mt * 16 + 4 * q);
  }
#pragma unroll
  for (int j = 0; j < 4; ++j) R.b3[j] = w.p[oB3 + 4 * q + j];
}

struct InferTiles {
  alignas(16) short X[64][kKin];  // bf16 features of the pass's 64 samples
  alignas(16) float Y[64][kKout];
};

// X (LDS, filled by the caller) -> Y (LDS) for the 64 samples of a pass
__device__ __forceinline__ void infer64(const WRegs& R, InferTiles& T, int r, int q) {
  const f32x4 zero = {0.f, 0.f, 0.f, 0.f};
  for (int tile = 0; tile < 4; ++tile) {
    // B operand of layer 1: features 4q..4q+3 of sample r (one 8-byte read)
    const bf16x4 xb = *reinterpret_cast<const bf16x4*>(&T.X[tile * 16 + r][4 * q]);
    bf16x4 h1[4], h2[4];
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
      const f32x4 acc = mfma(R.w1[mt], xb, zero);
#pragma unroll
      for (int j = 0; j < 4; ++j) h1[mt][j] = f2bf(tanh_fast(acc[j] + R.b1[mt][j]));
    }
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
      f32x4 acc = zero;
#pragma unroll
      for (int kt = 0; kt < 4; ++kt) acc = mfma(R.w2[mt][kt], h1[kt], acc);
#pragma unroll
      for (int j = 0; j < 4; ++j) h2[mt][j] = f2bf(tanh_fast(acc[j] + R.b2[mt][j]));
    }
    f32x4 y = zero;
#pragma unroll
    for (int kt = 0; kt < 4; ++kt) y = mfma(R.w3[kt], h2[kt], y);
#pragma unroll
    for (int j = 0; j < 4; ++j) T.Y[tile * 16 + r][4 * q + j] = y[j] + R.b3[j];
  }
  wave_sync();
}

__global__ void __launch_bounds__(kMlpBlock) k_mlp_forward(MlpWeights w, int n, const float* __restrict__ x,
                                                          float* __restrict__ y) {
  __shared__ InferTiles tiles[kWaves];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int r = lane & 15, q = lane >> 4;
  InferTiles& T = tiles[wv];
  WRegs R;
  load_wregs(w, r, q, R);
  const int npass = (n + 63) / 64;
  for (int pass = blockIdx.x * kWaves + wv; pass < npass; pass += gridDim.x * kWaves) {
    const int s = pass * 64 + lane;
#pragma unroll
    for (int f = 0; f < kKin; ++f) T.X[lane][f] = f2bf((s < n && f < kNin) ? x[(long long)s * kNin + f] : 0.f);
    wave_sync();
    infer64(R, T, r, q);
    if (s < n) {
#pragma unroll
      for (int o = 0; o < kNout; ++o) y[(long long)s * kNout + o] = T.Y[lane][o];
    }
    wave_sync();
  }
}

// Loss and gradient partials of a batch: per wavefront, sum over its tiles.
// partial[(block*kWaves + wave)*kNparam + i], lossp[block*kWaves + wave].
__global__ void __launch_bounds__(kMlpBlock) k_mlp_grad(MlpWeights w, int n, const float* __restrict__ x,
                                                       const float* __restrict__ t, const int* __restrict__ idx,
                                                       float inv_count, float* __restrict__ partial,
                                                       float* __restrict__ lossp) {
  __shared__ WaveTiles tiles[kWaves];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int r = lane & 15, q = lane >> 4;
  WaveTiles& T = tiles[wv];
  const f32x4 zero = {0.f, 0.f, 0.f, 0.f};
  f32x4 gW1[4], gW2[4][4], gW3[4];
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    gW1[a] = zero;
    gW3[a] = zero;
#pragma unroll
    for (int b = 0; b < 4; ++b) gW2[a][b] = zero;
  }
  float gb1 = 0.f, gb2 = 0.f, gb3 = 0.f, loss = 0.f;
  const int ntile = (n + 15) / 16;
  for (int tile = blockIdx.x * kWaves + wv; tile < ntile; tile += gridDim.x * kWaves) {
    const int s0 = tile * 16;
    load_x_tile(T, x, idx, n, s0, lane);
    wave_sync();
    const f32x4 yy = tile_forward(w, T, r, q);
    // dY = 2 (y - t) / (count) on the 5 outputs of valid samples
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int s = s0 + 4 * q + j;
      float d = 0.f;
      if (r < kNout && s < n) {
        const long long src = idx ? (long long)idx[s] : (long long)s;
        const float e = yy[j] - t[src * kNout + r];
        loss += e * e;
        d = 2.f * e * inv_count;
      }
      T.DY[4 * q + j][r] = d;
    }
    wave_sync();
    // dW3 += dY^T H2 ; db3
#pragma unroll
    for (int kt = 0; kt < 4; ++kt)
      gW3[kt] = mfma(frag_cols(&T.DY[0][0], kKout, 4 * q, r), frag_cols(&T.H2[0][0], kNh, 4 * q, kt * 16 + r),
                     gW3[kt]);
    if (lane < 16) {
#pragma unroll
      for (int s = 0; s < 16; ++s) gb3 += T.DY[s][lane];
    }
    // dH2 = dY W3 ; dZ2 = dH2 (1 - H2^2)
#pragma unroll
    for (int kt = 0; kt < 4; ++kt) {
      f32x4 d = mfma(frag_rows(&T.DY[0][0], kKout, r, 4 * q), wfrag_cols(w.W3, kNh, 4 * q, kt * 16 + r), zero);
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const float h = T.H2[4 * q + j][kt * 16 + r];
        T.D2[4 * q + j][kt * 16 + r] = d[j] * (1.f - h * h);
      }
    }
    wave_sync();
    // dW2 += dZ2^T H1 ; db2
#pragma unroll
    for (int ht = 0; ht < 4; ++ht) {
      const bf16x4 a = frag_cols(&T.D2[0][0], kNh, 4 * q, ht * 16 + r);
#pragma unroll
      for (int kt = 0; kt < 4; ++kt)
        gW2[ht][kt] = mfma(a, frag_cols(&T.H1[0][0], kNh, 4 * q, kt * 16 + r), gW2[ht][kt]);
    }
#pragma unroll
    for (int s = 0; s < 16; ++s) gb2 += T.D2[s][lane];
    // dH1 = dZ2 W2 ; dZ1 = dH1 (1 - H1^2)
#pragma unroll
    for (int kt = 0; kt < 4; ++kt) {
      f32x4 d = zero;
#pragma unroll
      for (int hk = 0; hk < 4; ++hk)
        d = mfma(frag_rows(&T.D2[0][0], kNh, r, hk * 16 + 4 * q), wfrag_cols(w.W2, kNh, hk * 16 + 4 * q, kt * 16 + r),
                 d);
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const float h = T.H1[4 * q + j][kt * 16 + r];
        T.D1[4 * q + j][kt * 16 + r] = d[j] * (1.f - h * h);
      }
    }
    wave_sync();
    // dW1 += dZ1^T X ; db1
#pragma unroll
    for (int ht = 0; ht < 4; ++ht)
      gW1[ht] = mfma(frag_cols(&T.D1[0][0], kNh, 4 * q, ht * 16 + r), frag_cols(&T.X[0][0], kKin, 4 * q, r), gW1[ht]);
#pragma unroll
    for (int s = 0; s < 16; ++s) gb1 += T.D1[s][lane];
    wave_sync();
  }
  // write this wavefront's partial gradient
  float* P = partial + (long long)(blockIdx.x * kWaves + wv) * kNparam;
#pragma unroll
  for (int ht = 0; ht < 4; ++ht)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int h = ht * 16 + 4 * q + j;
      P[oW1 + h * kKin + r] = r < kNin ? gW1[ht][j] : 0.f;
#pragma unroll
      for (int kt = 0; kt < 4; ++kt) P[oW2 + h * kNh + kt * 16 + r] = gW2[ht][kt][j];
    }
#pragma unroll
  for (int kt = 0; kt < 4; ++kt)
#pragma unroll
    for (int j = 0; j < 4; ++j) P[oW3 + (4 * q + j) * kNh + kt * 16 + r] = gW3[kt][j];
  P[oB1 + lane] = gb1;
  P[oB2 + lane] = gb2;
  if (lane < 16) P[oB3 + lane] = lane < kNout ? gb3 : 0.f;
  // loss partial: fixed butterfly over the wavefront
  for (int off = 32; off > 0; off >>= 1) loss += __shfl_xor(loss, off);
  if (lane == 0) lossp[blockIdx.x * kWaves + wv] = loss;
}

// Sum of the per-wavefront partials in a fixed order, in two stages so the
// work spreads over the chip: stage 1 sums groups of kRedGroup consecutive
// partials (one block row per group), stage 2 sums the group totals.
constexpr int kRedGroup = 32;

__global__ void __launch_bounds__(kMlpBlock) k_mlp_reduce1(int nparts, const float* __restrict__ partial,
                                                          const float* __restrict__ lossp, float* __restrict__ part2,
                                                          float* __restrict__ loss2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int p0 = blockIdx.y * kRedGroup, p1 = min(nparts, p0 + kRedGroup);
  if (i < kNparam) {
    float s = 0.f;
    for (int p = p0; p < p1; ++p) s += partial[(long long)p * kNparam + i];
    part2[(long long)blockIdx.y * kNparam + i] = s;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    float s = 0.f;
    for (int p = p0; p < p1; ++p) s += lossp[p];
    loss2[blockIdx.y] = s;
  }
}

__global__ void __launch_bounds__(kMlpBlock) k_mlp_reduce2(int ngroup, const float* __restrict__ part2,
                                                          const float* __restrict__ loss2, float* __restrict__ grad,
                                                          float* __restrict__ loss) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < kNparam) {
    float s = 0.f;
    for (int b = 0; b < ngroup; ++b) s += part2[(long long)b * kNparam + i];
    grad[i] = s;
  }
  if (i == 0) {
    float s = 0.f;
    for (int b = 0; b < ngroup; ++b) s += loss2[b];
    *loss = s;
  }
}

// padded entries of the layout stay exactly zero
__device__ __forceinline__ bool live_param(int i) {
  if (i >= oW1 && i < oB1) return ((i - oW1) % kKin) < kNin;
  if (i >= oW3 && i < oB3) return ((i - oW3) / kNh) < kNout;
  if (i >= oB3) return (i - oB3) < kNout;
  return true;
}

__global__ void __launch_bounds__(kMlpBlock) k_mlp_adam(float* __restrict__ p, float* __restrict__ m,
                                                       float* __restrict__ v, const float* __restrict__ g, float lr,
                                                       float b1, float b2, float eps, float c1, float c2,
                                                       short* __restrict__ W1, short* __restrict__ W2,
                                                       short* __restrict__ W3) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= kNparam) return;
  float pi = p[i];
  if (live_param(i)) {
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    pi -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  } else {
    pi = 0.f;
  }
  p[i] = pi;
  if (i < oB1) W1[i - oW1] = f2bf(pi);
  else if (i >= oW2 && i < oB2) W2[i - oW2] = f2bf(pi);
  else if (i >= oW3 && i < oB3) W3[i - oW3] = f2bf(pi);
}

__global__ void k_mlp_to_bf16(const float* __restrict__ p, short* __restrict__ W1, short* __restrict__ W2,
                              short* __restrict__ W3) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= kNparam) return;
  if (i < oB1) W1[i - oW1] = f2bf(p[i]);
  else if (i >= oW2 && i < oB2) W2[i - oW2] = f2bf(p[i]);
  else if (i >= oW3 && i < oB3) W3[i - oW3] = f2bf(p[i]);
}

}  // namespace pcfd

using namespace pcfd;

struct pcfd_mlp {
  hipStream_t stream = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  float* p = nullptr;  // FP32 parameters
  float* m = nullptr;  // Adam moments
  float* v = nullptr;
  float* g = nullptr;  // gradient
  float* loss = nullptr;
  short *W1 = nullptr, *W2 = nullptr, *W3 = nullptr;
  float* partial = nullptr;
  float* lossp = nullptr;
  float* part2 = nullptr;  // group totals of the partials
  float* loss2 = nullptr;
  int grid = 0;  // blocks of the gradient kernel
  int* scratch = nullptr;  // device address of the no-allocation scratch on this mlp's device
  pcfd_mlp* partner = nullptr;  // error network of the surrogate (borrowed)
  long long adam_t = 0;
  float last_ms = 0.f;
  std::string err;
};

namespace {
thread_local std::string g_mlp_err;

void mcheck(hipError_t e, const char* what) {
  if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

template <typename F>
int guarded_mlp(pcfd_mlp* c, F&& f) {
  if (!c) return PCFD_ERR_ARG;
  try {
    return f(*c);
  } catch (const std::exception& e) {
    c->err = e.what();
    return PCFD_ERR_HIP;
  }
}

MlpWeights weights(const pcfd_mlp& c) { return MlpWeights{c.W1, c.W2, c.W3, c.p}; }

void refresh_bf16(pcfd_mlp& c) {
  hipLaunchKernelGGL(k_mlp_to_bf16, dim3((kNparam + 255) / 256), dim3(256), 0, c.stream, c.p, c.W1, c.W2, c.W3);
  mcheck(hipGetLastError(), "k_mlp_to_bf16");
}

// loss/gradient of n samples (device arrays, optional index list)
void grad_device(pcfd_mlp& c, int n, const float* x, const float* t, const int* idx) {
  const int ntile = (n + 15) / 16;
  const int grid = std::max(1, std::min(c.grid, (ntile + kWaves - 1) / kWaves));
  hipLaunchKernelGGL(k_mlp_grad, dim3(grid), dim3(kMlpBlock), 0, c.stream, weights(c), n, x, t, idx,
                     1.0f / (float)((long long)n * kNout), c.partial, c.lossp);
  mcheck(hipGetLastError(), "k_mlp_grad");
  const int nparts = grid * kWaves, ngroup = (nparts + kRedGroup - 1) / kRedGroup;
  hipLaunchKernelGGL(k_mlp_reduce1, dim3((kNparam + kMlpBlock - 1) / kMlpBlock, ngroup), dim3(kMlpBlock), 0,
                     c.stream, nparts, c.partial, c.lossp, c.part2, c.loss2);
  mcheck(hipGetLastError(), "k_mlp_reduce1");
  hipLaunchKernelGGL(k_mlp_reduce2, dim3((kNparam + kMlpBlock - 1) / kMlpBlock), dim3(kMlpBlock), 0, c.stream,
                     ngroup, c.part2, c.loss2, c.g, c.loss);
  mcheck(hipGetLastError(), "k_mlp_reduce2");
}

struct DBuf {
  void* p = nullptr;
  explicit DBuf(size_t bytes) {
    if (bytes) mcheck(hipMalloc(&p, bytes), "hipMalloc");
  }
  ~DBuf() {
    if (p) (void)hipFree(p);
  }
  template <typename T>
  T* as() const {
    return (T*)p;
  }
};

// splitmix64: the batch index stream of training
uint64_t mix(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
}  // namespace

namespace pcfd {
int* scratch_on_current_device();
}

extern "C" {

int pcfd_mlp_create(unsigned long long seed, pcfd_mlp** out) {
  if (!out) {
    g_mlp_err = "pcfd_mlp_create: null argument";
    return PCFD_ERR_ARG;
  }
  *out = nullptr;
  pcfd_mlp* c = new pcfd_mlp();
  try {
    mcheck(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking), "hipStreamCreate");
    mcheck(hipEventCreate(&c->e0), "hipEventCreate");
    mcheck(hipEventCreate(&c->e1), "hipEventCreate");
    c->scratch = pcfd::scratch_on_current_device();
    c->grid = 256;
    const size_t P = sizeof(float) * kNparam;
    mcheck(hipMalloc((void**)&c->p, P), "hipMalloc");
    mcheck(hipMalloc((void**)&c->m, P), "hipMalloc");
    mcheck(hipMalloc((void**)&c->v, P), "hipMalloc");
    mcheck(hipMalloc((void**)&c->g, P), "hipMalloc");
    mcheck(hipMalloc((void**)&c->loss, sizeof(float)), "hipMalloc");
    mcheck(hipMalloc((void**)&c->W1, sizeof(short) * kNh * kKin), "hipMalloc");
    mcheck(hipMalloc((void**)&c->W2, sizeof(short) * kNh * kNh), "hipMalloc");
    mcheck(hipMalloc((void**)&c->W3, sizeof(short) * kKout * kNh), "hipMalloc");
    mcheck(hipMalloc((void**)&c->partial, sizeof(float) * (size_t)c->grid * kWaves * kNparam), "hipMalloc");
    mcheck(hipMalloc((void**)&c->lossp, sizeof(float) * (size_t)c->grid * kWaves), "hipMalloc");
    const size_t ngroup = ((size_t)c->grid * kWaves + kRedGroup - 1) / kRedGroup;
    mcheck(hipMalloc((void**)&c->part2, sizeof(float) * ngroup * kNparam), "hipMalloc");
    mcheck(hipMalloc((void**)&c->loss2, sizeof(float) * ngroup), "hipMalloc");
    // Glorot-uniform weights, zero biases (host-side, reproducible)
    std::vector<float> h(kNparam, 0.f);
    uint64_t s = seed;
    auto uni = [&](float a) {
      s = mix(s);
      return a * (2.f * (float)((s >> 11) * (1.0 / 9007199254740992.0)) - 1.f);
    };
    const float a1 = std::sqrt(6.f / (kNin + kNh)), a2 = std::sqrt(6.f / (kNh + kNh)),
                a3 = std::sqrt(6.f / (kNh + kNout));
    for (int hh = 0; hh < kNh; ++hh)
      for (int k = 0; k < kNin; ++k) h[oW1 + hh * kKin + k] = uni(a1);
    for (int i = 0; i < kNh * kNh; ++i) h[oW2 + i] = uni(a2);
    for (int o = 0; o < kNout; ++o)
      for (int k = 0; k < kNh; ++k) h[oW3 + o * kNh + k] = uni(a3);
    mcheck(hipMemcpy(c->p, h.data(), P, hipMemcpyHostToDevice), "hipMemcpy");
    mcheck(hipMemset(c->m, 0, P), "hipMemset");
    mcheck(hipMemset(c->v, 0, P), "hipMemset");
    refresh_bf16(*c);
    mcheck(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
  } catch (const std::exception& e) {
    g_mlp_err = e.what();
    pcfd_mlp_destroy(c);
    return PCFD_ERR_HIP;
  }
  *out = c;
  return PCFD_OK;
}

int pcfd_mlp_destroy(pcfd_mlp* c) {
  if (!c) return PCFD_ERR_ARG;
  float* fs[] = {c->p, c->m, c->v, c->g, c->loss, c->partial, c->lossp, c->part2, c->loss2};
  for (float* f : fs)
    if (f) (void)hipFree(f);
  short* ss[] = {c->W1, c->W2, c->W3};
  for (short* s : ss)
    if (s) (void)hipFree(s);
  if (c->e0) (void)hipEventDestroy(c->e0);
  if (c->e1) (void)hipEventDestroy(c->e1);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return PCFD_OK;
}

const char* pcfd_mlp_last_error(const pcfd_mlp* c) { return c ? c->err.c_str() : g_mlp_err.c_str(); }

double pcfd_mlp_last_kernel_ms(const pcfd_mlp* c) { return c ? (double)c->last_ms : 0.0; }

int pcfd_mlp_set_partner(pcfd_mlp* c, pcfd_mlp* partner) {
  return guarded_mlp(c, [&](pcfd_mlp& m) -> int {
    if (partner == c) {
      m.err = "pcfd_mlp_set_partner: a network cannot be its own error network";
      return PCFD_ERR_ARG;
    }
    m.partner = partner;
    return PCFD_OK;
  });
}

int pcfd_mlp_set_params(pcfd_mlp* c, const float* params) {
  return guarded_mlp(c, [&](pcfd_mlp& m) -> int {
    if (!params) {
      m.err = "pcfd_mlp_set_params: null params";
      return PCFD_ERR_ARG;
    }
    mcheck(hipMemcpyAsync(m.p, params, sizeof(float) * kNparam, hipMemcpyHostToDevice, m.stream), "copy");
    mcheck(hipMemsetAsync(m.m, 0, sizeof(float) * kNparam, m.stream), "hipMemset");
    mcheck(hipMemsetAsync(m.v, 0, sizeof(float) * kNparam, m.stream), "hipMemset");
    m.adam_t = 0;
    refresh_bf16(m);
    mcheck(hipStreamSynchronize(m.stream), "hipStreamSynchronize");
    return PCFD_OK;
  });
}

int pcfd_mlp_get_params(pcfd_mlp* c, float* params) {
  return guarded_mlp(c, [&](pcfd_mlp& m) -> int {
    if (!params) {
      m.err = "pcfd_mlp_get_params: null params";
      return PCFD_ERR_ARG;
    }
    mcheck(hipMemcpyAsync(params, m.p, sizeof(float) * kNparam, hipMemcpyDeviceToHost, m.stream), "copy");
    mcheck(hipStreamSynchronize(m.stream), "hipStreamSynchronize");
    return PCFD_OK;
  });
}

int pcfd_mlp_forward_device(pcfd_mlp* c, int n, const float* x, float* y) {
  return guarded_mlp(c, [&](pcfd_mlp& m) -> int {
    if (n < 0 || (n > 0 && (!x || !y))) {
      m.err = "pcfd_mlp_forward: invalid arguments";
      return PCFD_ERR_ARG;
    }
    if (n == 0) return PCFD_OK;
    // each wavefront loads its weight fragments once and loops over 64-sample
    // passes: about four blocks per CU
    const int npass = (n + 63) / 64;
    const int grid = std::max(1, std::min(1024, (npass + kWaves - 1) / kWaves));
    mcheck(hipEventRecord(m.e0, m.stream), "hipEventRecord");
    hipLaunchKernelGGL(k_mlp_forward, dim3(grid), dim3(kMlpBlock), 0, m.stream, weights(m), n, x, y);
    mcheck(hipGetLastError(), "k_mlp_forward");
    mcheck(hipEventRecord(m.e1, m.stream), "hipEventRecord");
    mcheck(hipEventSynchronize(m.e1), "hipEventSynchronize");
    mcheck(hipEventElapsedTime(&m.last_ms, m.e0, m.e1), "hipEventElapsedTime");
    return PCFD_OK;
  });
}

int pcfd_mlp_forward(pcfd_mlp* c, int n, const float* x, float* y) {
  return guarded_mlp(c, [&](pcfd_mlp& m) -> int {
    if (n < 0 || (n > 0 && (!x || !y))) {
      m.err = "pcfd_mlp_forward: invalid arguments";
      return PCFD_ERR_ARG;
    }
    if (n == 0) return PCFD_OK;
    DBuf dx(sizeof(float) * (size_t)n * kNin), dy(sizeof(float) * (size_t)n * kNout);
    mcheck(hipMemcpyAsync(dx.p, x, sizeof(float) * (size_t)n * kNin, hipMemcpyHostToDevice, m.stream), "copy");
    int rc = pcfd_mlp_forward_device(&m, n, dx.as<float>(), dy.as<float>());
    if (rc != PCFD_OK) return rc;
    mcheck(hipMemcpyAsync(y, dy.p, sizeof(float) * (size_t)n * kNout, hipMemcpyDeviceToHost, m.stream), "copy");
    mcheck(hipStreamSynchronize(m.stream), "hipStreamSynchronize");
    return PCFD_OK;
  });
}

int pcfd_mlp_loss_grad(pcfd_mlp* c, int n, const float* x, const float* t, float* loss, float* grad) {
  return guarded_mlp(c, [&](pcfd_mlp& m) -> int {
    if (n <= 0 || !x || !t) {
      m.err = "pcfd_mlp_loss_grad: invalid arguments";
      return PCFD_ERR_ARG;
    }
    DBuf dx(sizeof(float) * (size_t)n * kNin), dt(sizeof(float) * (size_t)n * kNout);
    mcheck(hipMemcpyAsync(dx.p, x, sizeof(float) * (size_t)n * kNin, hipMemcpyHostToDevice, m.stream), "copy");
    mcheck(hipMemcpyAsync(dt.p, t, sizeof(float) * (size_t)n * kNout, hipMemcpyHostToDevice, m.stream), "copy");
    grad_device(m, n, dx.as<float>(), dt.as<float>(), nullptr);
    float l = 0.f;
    mcheck(hipMemcpyAsync(&l, m.loss, sizeof(float), hipMemcpyDeviceToHost, m.stream), "copy");
    if (grad) mcheck(hipMemcpyAsync(grad, m.g, sizeof(float) * kNparam, hipMemcpyDeviceToHost, m.stream), "copy");
    mcheck(hipStreamSynchronize(m.stream), "hipStreamSynchronize");
    if (loss) *loss = l / (float)((long long)n * kNout);
    return PCFD_OK;
  });
}

int pcfd_mlp_train(pcfd_mlp* c, int n, const float* x, const float* t, int steps, int batch, float lr,
                   unsigned long long seed, float* loss_history) {
  return guarded_mlp(c, [&](pcfd_mlp& m) -> int {
    if (n <= 0 || !x || !t || steps < 0 || batch <= 0 || batch > n || !(lr > 0.f)) {
      m.err = "pcfd_mlp_train: invalid arguments";
      return PCFD_ERR_ARG;
    }
    DBuf dx(sizeof(float) * (size_t)n * kNin), dt(sizeof(float) * (size_t)n * kNout),
        didx(sizeof(int) * (size_t)batch);
    mcheck(hipMemcpyAsync(dx.p, x, sizeof(float) * (size_t)n * kNin, hipMemcpyHostToDevice, m.stream), "copy");
    mcheck(hipMemcpyAsync(dt.p, t, sizeof(float) * (size_t)n * kNout, hipMemcpyHostToDevice, m.stream), "copy");
    std::vector<int> hidx((size_t)batch);
    std::vector<float> hist;
    uint64_t s = seed;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    mcheck(hipEventRecord(m.e0, m.stream), "hipEventRecord");
    for (int it = 0; it < steps; ++it) {
      for (int k = 0; k < batch; ++k) {
        s = mix(s);
        hidx[k] = (int)(s % (uint64_t)n);
      }
      mcheck(hipMemcpyAsync(didx.p, hidx.data(), sizeof(int) * (size_t)batch, hipMemcpyHostToDevice, m.stream),
             "copy");
      grad_device(m, batch, dx.as<float>(), dt.as<float>(), didx.as<int>());
      ++m.adam_t;
      // cosine decay of the step size over the run
      const float lr_t = 0.5f * lr * (1.f + std::cos(3.14159265f * (float)it / (float)std::max(1, steps)));
      const float c1 = 1.f - std::pow(b1, (float)m.adam_t), c2 = 1.f - std::pow(b2, (float)m.adam_t);
      hipLaunchKernelGGL(k_mlp_adam, dim3((kNparam + kMlpBlock - 1) / kMlpBlock), dim3(kMlpBlock), 0, m.stream,
                         m.p, m.m, m.v, m.g, lr_t, b1, b2, eps, c1, c2, m.W1, m.W2, m.W3);
      mcheck(hipGetLastError(), "k_mlp_adam");
      if (loss_history) {
        float l = 0.f;
        mcheck(hipMemcpyAsync(&l, m.loss, sizeof(float), hipMemcpyDeviceToHost, m.stream), "copy");
        mcheck(hipStreamSynchronize(m.stream), "hipStreamSynchronize");
        loss_history[it] = l / (float)((long long)batch * kNout);
      } else {
        // the host-side index buffer is reused: wait for the copy
        mcheck(hipStreamSynchronize(m.stream), "hipStreamSynchronize");
      }
    }
    mcheck(hipEventRecord(m.e1, m.stream), "hipEventRecord");
    mcheck(hipEventSynchronize(m.e1), "hipEventSynchronize");
    mcheck(hipEventElapsedTime(&m.last_ms, m.e0, m.e1), "hipEventElapsedTime");
    return PCFD_OK;
  });
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Surrogate production rates with exact fallback.
//
// Features: f0 = (ln T - lnT_mid)/lnT_half, f1 = (ln rho - lnrho_mid)/
// lnrho_half, f2..f6 = (Y_i - y_mid)/y_scale.  Targets: asinh(wdot_i /
// (rho s0)) / sigma_i, inverted here.  A cell falls back to the exact rates
// when it lies outside the training box or when the predicted rates break
// element conservation -- the exact rates conserve every element's atoms
// exactly (every reaction does), so the molar imbalance of the two elements,
//   sum_e |sum_i atoms[e][i] w_i / M_i|   (for air: N and O atoms),
// over (sum_i |w_i/M_i| + rho * floor) estimates the surrogate's error.
// With an error network (same shape, trained on the first network's own
// errors: g_i = ln(|sigma_i (y_i - target_i)| + 1e-3), normalised by
// err_mid/err_scale) the predicted error e_i = exp(g_i) of output i in
// asinh units becomes a rate error rho s0 (sinh(|z_i| + e_i) - sinh(|z_i|)),
// z_i = sigma_i y_i, and a cell also falls back when the largest of these
// exceeds err_tol * (max_i |w_i| + rho * rate_floor).
namespace pcfd {

struct SurrogateDev {
  float lnT_mid, ilnT_half, lnrho_mid, ilnrho_half, y_mid, iy_scale, y_eps;  // i*: reciprocals
  double s0;
  double sigma[5];
  double T_min, T_max, rho_min, rho_max, tol, floor;
  double imw[5];  // 1 / molecular weight
  double atoms[2][5];
  double err_tol, rate_floor;
  double err_mid[5], err_scale[5];
};

// rates of one member from its transformed outputs y (FP32): rho s0 sinh(sigma y)
__device__ __forceinline__ void member_rates(const SurrogateDev& sd, double rho, const float* y, double* w) {
#pragma unroll
  for (int k = 0; k < 5; ++k) {
    const float z = (float)sd.sigma[k] * y[k];
    const float e = expf(z);
    w[k] = rho * sd.s0 * (double)(0.5f * (e - 1.f / e));
  }
}

template <bool ENS>
__global__ void __launch_bounds__(kMlpBlock) k_surrogate(MlpWeights w, MlpWeights w2, SurrogateDev sd, int n,
                                                        const double* __restrict__ rhoi, const double* __restrict__ T,
                                                        double* __restrict__ wdot, int* __restrict__ lst,
                                                        int* __restrict__ cnt) {
  __shared__ InferTiles tiles[kWaves];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int r = lane & 15, q = lane >> 4;
  InferTiles& Tl = tiles[wv];
  WRegs R, R2;
  load_wregs(w, r, q, R);
  if constexpr (ENS) load_wregs(w2, r, q, R2);
  const int npass = (n + 63) / 64;
  for (int pass = blockIdx.x * kWaves + wv; pass < npass; pass += gridDim.x * kWaves) {
    const int c = pass * 64 + lane;
    const bool valid = c < n;
    double ri[5], rho = 0.0, Tc = 1.0;
    if (valid) {
#pragma unroll
      for (int k = 0; k < 5; ++k) {
        ri[k] = rhoi[(long long)c * 5 + k];
        rho += ri[k];
      }
      Tc = T[c];
    } else {
#pragma unroll
      for (int k = 0; k < 5; ++k) ri[k] = 0.2;
      rho = 1.0;
    }
    // features (FP32): lane builds its own sample
    const float irho = (float)(1.0 / rho);
    Tl.X[lane][0] = f2bf((logf((float)Tc) - sd.lnT_mid) * sd.ilnT_half);
    Tl.X[lane][1] = f2bf((logf((float)rho) - sd.lnrho_mid) * sd.ilnrho_half);
#pragma unroll
    for (int k = 0; k < 5; ++k) {
      const float Y = (float)ri[k] * irho;
      Tl.X[lane][2 + k] = f2bf(((sd.y_eps > 0.f ? logf(Y + sd.y_eps) : Y) - sd.y_mid) * sd.iy_scale);
    }
#pragma unroll
    for (int f = kNin; f < kKin; ++f) Tl.X[lane][f] = 0;
    wave_sync();
    infer64(R, Tl, r, q);
    float ya[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) ya[k] = Tl.Y[lane][k];
    double wv5[5], werr = 0.0;
    member_rates(sd, rho, ya, wv5);
    if constexpr (ENS) {
      // every lane has read its outputs before the error network rewrites Y
      wave_sync();
      infer64(R2, Tl, r, q);
#pragma unroll
      for (int k = 0; k < 5; ++k) {
        const double e = exp((double)Tl.Y[lane][k] * sd.err_scale[k] + sd.err_mid[k]);
        const double z = fabs(sd.sigma[k] * (double)ya[k]);
        werr = fmax(werr, rho * sd.s0 * (sinh(z + e) - sinh(z)));
      }
    }
    bool fb = false;
    if (valid) {
      double m[5], mag = 0.0, wmax = 0.0;
#pragma unroll
      for (int k = 0; k < 5; ++k) {
        m[k] = wv5[k] * sd.imw[k];
        mag += fabs(m[k]);
        wmax = fmax(wmax, fabs(wv5[k]));
      }
      double eN = 0.0, eO = 0.0;
#pragma unroll
      for (int k = 0; k < 5; ++k) {
        eN += sd.atoms[0][k] * m[k];
        eO += sd.atoms[1][k] * m[k];
      }
      eN = fabs(eN);
      eO = fabs(eO);
      const bool out = !(Tc >= sd.T_min && Tc <= sd.T_max && rho >= sd.rho_min && rho <= sd.rho_max);
      bool bad = (eN + eO) > sd.tol * (mag + rho * sd.floor);
      if constexpr (ENS) bad = bad || !(werr <= sd.err_tol * (wmax + rho * sd.rate_floor));
#pragma unroll
      for (int k = 0; k < 5; ++k) wdot[(long long)c * 5 + k] = wv5[k];
      fb = out || bad;
    }
    // the pass's fallback cells, compacted in lane order into its own 64
    // slots (no atomics): cnt[pass] of them at lst[pass*64 ...]
    const unsigned long long mask = __ballot(fb);
    if (fb) lst[(long long)pass * 64 + __popcll(mask & ((1ull << lane) - 1ull))] = c;
    if (lane == 0) cnt[pass] = __popcll(mask);
    wave_sync();
  }
}

// Exact rates of the fallback cells: entry t of the compacted list lives in
// the pass p with off[p] <= t < off[p+1] (binary search).
__global__ void __launch_bounds__(kChemBlock) k_wdot_list(const ChemDev* __restrict__ md, int npass,
                                                         const int* __restrict__ lst, const int* __restrict__ off,
                                                         const double* __restrict__ rhoi,
                                                         const double* __restrict__ T, double* __restrict__ wdot) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= off[npass]) return;
  int lo = 0, hi = npass;  // off[lo] <= t < off[hi]
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (off[mid] <= t) lo = mid;
    else hi = mid;
  }
  const int c = lst[(long long)lo * 64 + (t - off[lo])];
  const ChemDev& m = *md;
  double r[PCFD_CHEM_MAXS], w[PCFD_CHEM_MAXS];
  for (int k = 0; k < m.ns; ++k) r[k] = rhoi[(long long)c * m.ns + k];
  wdot_of(m, r, T[c], w);
  for (int k = 0; k < m.ns; ++k) wdot[(long long)c * m.ns + k] = w[k];
}

const ChemDev* chem_device_model(const pcfd_chem* c);
int chem_nspecies(const pcfd_chem* c);
const double* chem_mw(const pcfd_chem* c);

// Scratch of pcfd_chem_surrogate_wdot_device_noalloc: the fallback list (64
// slots per 64-cell pass), the per-pass counts and their offsets, for at most
// PCFD_SURROGATE_NOALLOC_MAX_CELLS cells (17.3 MB per device, held until the
// process ends).  Every entry a call reads is written earlier in the same
// call (each pass writes its count, the scan writes every offset), so
// nothing needs clearing.  Host code never passes the array's name to a
// kernel: it is a host-side stand-in; the device address comes from
// hipGetSymbolAddress, once per device.
constexpr int kNoAllocPasses = PCFD_SURROGATE_NOALLOC_MAX_CELLS / 64;
constexpr size_t kNoAllocInts = (size_t)kNoAllocPasses * 64 + kNoAllocPasses + (kNoAllocPasses + 1);
static __device__ __align__(256) int scratch[kNoAllocInts];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
int* g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device (looked up once per device).
int* scratch_on_current_device() {
  int dev = 0;
  mcheck(hipGetDevice(&dev), "hipGetDevice");
  if (dev < 0 || dev >= kMaxDevices) throw std::runtime_error("device id out of range for the scratch table");
  std::lock_guard<std::mutex> g(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    int* ptr = nullptr;
    mcheck(hipGetSymbolAddress((void**)&ptr, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
    g_scratch_ptr[dev] = ptr;
  }
  return g_scratch_ptr[dev];
}

// Surrogate pass + fixed-order scan of the per-pass fallback counts + exact
// rates of the fallback list, on the mlp's stream; lst/cnt/off hold at least
// npass*64, npass and npass+1 ints.
int surrogate_launch(pcfd_chem* chem, pcfd_mlp& m, const pcfd_surrogate_desc* d, int n, const double* rhoi,
                     const double* T, double* wdot, int* lst, int* cnt, int* off, int* nfallback, double* ms);

}  // namespace pcfd

extern "C" {

int pcfd_chem_surrogate_wdot_device(pcfd_chem* chem, pcfd_mlp* mlp, const pcfd_surrogate_desc* d, int n,
                                    const double* rhoi, const double* T, double* wdot, int* nfallback,
                                    double* ms) {
  return guarded_mlp(mlp, [&](pcfd_mlp& m) -> int {
    if (!chem || !d || n < 0 || (n > 0 && (!rhoi || !T || !wdot))) {
      m.err = "pcfd_chem_surrogate_wdot: invalid arguments";
      return PCFD_ERR_ARG;
    }
    if (chem_nspecies(chem) != kNout) {
      m.err = "pcfd_chem_surrogate_wdot: the surrogate is built for 5 species (N, N2, O2, NO, O)";
      return PCFD_ERR_UNSUPPORTED;
    }
    if (nfallback) *nfallback = 0;
    if (n == 0) return PCFD_OK;
    const int npass = (n + 63) / 64;
    // fallback list: 64 slots per pass, per-pass counts, their offsets
    DBuf dlst(sizeof(int) * (size_t)npass * 64), dcnt(sizeof(int) * (size_t)npass),
        doff(sizeof(int) * ((size_t)npass + 1));
    return surrogate_launch(chem, m, d, n, rhoi, T, wdot, dlst.as<int>(), dcnt.as<int>(), doff.as<int>(), nfallback,
                            ms);
  });
}

int pcfd_chem_surrogate_wdot_device_noalloc(pcfd_chem* chem, pcfd_mlp* mlp, const pcfd_surrogate_desc* d, int n,
                                            const double* rhoi, const double* T, double* wdot, int* nfallback,
                                            double* ms) {
  return guarded_mlp(mlp, [&](pcfd_mlp& m) -> int {
    if (!chem || !d || n < 0 || (n > 0 && (!rhoi || !T || !wdot))) {
      m.err = "pcfd_chem_surrogate_wdot_device_noalloc: invalid arguments";
      return PCFD_ERR_ARG;
    }
    if (n > PCFD_SURROGATE_NOALLOC_MAX_CELLS) {
      m.err = "pcfd_chem_surrogate_wdot_device_noalloc: " + std::to_string(n) +
              " cells exceed the fixed scratch of PCFD_SURROGATE_NOALLOC_MAX_CELLS = " +
              std::to_string(PCFD_SURROGATE_NOALLOC_MAX_CELLS) + " (split the call)";
      return PCFD_ERR_ARG;
    }
    if (chem_nspecies(chem) != kNout) {
      m.err = "pcfd_chem_surrogate_wdot_device_noalloc: the surrogate is built for 5 species (N, N2, O2, NO, O)";
      return PCFD_ERR_UNSUPPORTED;
    }
    if (nfallback) *nfallback = 0;
    if (n == 0) return PCFD_OK;
    int* const base = m.scratch;
    int* const lst = base;
    int* const cnt = base + (size_t)kNoAllocPasses * 64;
    int* const off = cnt + kNoAllocPasses;
    return surrogate_launch(chem, m, d, n, rhoi, T, wdot, lst, cnt, off, nfallback, ms);
  });
}

}  // extern "C"

namespace pcfd {

int surrogate_launch(pcfd_chem* chem, pcfd_mlp& m, const pcfd_surrogate_desc* d, int n, const double* rhoi,
                     const double* T, double* wdot, int* lst, int* cnt, int* off, int* nfallback, double* ms) {
  SurrogateDev sd{};
  sd.lnT_mid = (float)d->lnT_mid;
  sd.ilnT_half = (float)(1.0 / d->lnT_half);
  sd.lnrho_mid = (float)d->lnrho_mid;
  sd.ilnrho_half = (float)(1.0 / d->lnrho_half);
  sd.y_mid = (float)d->y_mid;
  sd.iy_scale = (float)(1.0 / d->y_scale);
  sd.y_eps = (float)d->y_eps;
  sd.s0 = d->rate_scale;
  for (int k = 0; k < 5; ++k) {
    sd.sigma[k] = d->sigma[k];
    sd.imw[k] = 1.0 / chem_mw(chem)[k];
    sd.atoms[0][k] = d->atoms[0][k];
    sd.atoms[1][k] = d->atoms[1][k];
  }
  sd.T_min = d->T_min;
  sd.T_max = d->T_max;
  sd.rho_min = d->rho_min;
  sd.rho_max = d->rho_max;
  sd.tol = d->tol;
  sd.floor = d->floor;
  sd.err_tol = d->err_tol;
  sd.rate_floor = d->rate_floor;
  const bool ens = m.partner != nullptr;
  for (int k = 0; k < 5; ++k) {
    sd.err_mid[k] = d->err_mid[k];
    sd.err_scale[k] = d->err_scale[k];
    if (ens && !(d->err_scale[k] > 0.0)) {
      m.err = "surrogate: an error network is set; err_scale must be > 0";
      return PCFD_ERR_ARG;
    }
  }
  if (ens && !(d->err_tol > 0.0 && d->rate_floor >= 0.0)) {
    m.err = "surrogate: an error network is set; err_tol must be > 0 and rate_floor >= 0";
    return PCFD_ERR_ARG;
  }
  // the error network's BF16 weights are written on its own stream
  if (ens) mcheck(hipStreamSynchronize(m.partner->stream), "hipStreamSynchronize");
  // each wavefront loads its weight fragments once and loops over 64-sample
  // passes: about four blocks per CU
  const int npass = (n + 63) / 64;
  const int grid = std::max(1, std::min(1024, (npass + kWaves - 1) / kWaves));
  mcheck(hipEventRecord(m.e0, m.stream), "hipEventRecord");
  if (ens)
    hipLaunchKernelGGL(k_surrogate<true>, dim3(grid), dim3(kMlpBlock), 0, m.stream, weights(m), weights(*m.partner),
                       sd, n, rhoi, T, wdot, lst, cnt);
  else
    hipLaunchKernelGGL(k_surrogate<false>, dim3(grid), dim3(kMlpBlock), 0, m.stream, weights(m), weights(m), sd, n,
                       rhoi, T, wdot, lst, cnt);
  mcheck(hipGetLastError(), "k_surrogate");
  launch_scan_counts(m.stream, npass, cnt, off);
  hipLaunchKernelGGL(k_wdot_list, dim3((n + kChemBlock - 1) / kChemBlock), dim3(kChemBlock), 0, m.stream,
                     chem_device_model(chem), npass, lst, off, rhoi, T, wdot);
  mcheck(hipGetLastError(), "k_wdot_list");
  mcheck(hipEventRecord(m.e1, m.stream), "hipEventRecord");
  int nf = 0;
  mcheck(hipMemcpyAsync(&nf, off + npass, sizeof(int), hipMemcpyDeviceToHost, m.stream), "copy");
  mcheck(hipStreamSynchronize(m.stream), "hipStreamSynchronize");
  mcheck(hipEventElapsedTime(&m.last_ms, m.e0, m.e1), "hipEventElapsedTime");
  if (nfallback) *nfallback = nf;
  if (ms) *ms = m.last_ms;
  return PCFD_OK;
}

}  // namespace pcfd

extern "C" {

int pcfd_chem_surrogate_wdot(pcfd_chem* chem, pcfd_mlp* mlp, const pcfd_surrogate_desc* d, int n,
                             const double* rhoi, const double* T, double* wdot, int* nfallback, double* ms) {
  return guarded_mlp(mlp, [&](pcfd_mlp& m) -> int {
    if (n < 0 || (n > 0 && (!rhoi || !T || !wdot))) {
      m.err = "pcfd_chem_surrogate_wdot: invalid arguments";
      return PCFD_ERR_ARG;
    }
    if (n == 0) return PCFD_OK;
    DBuf dr(sizeof(double) * (size_t)n * 5), dt(sizeof(double) * (size_t)n), dw(sizeof(double) * (size_t)n * 5);
    mcheck(hipMemcpyAsync(dr.p, rhoi, sizeof(double) * (size_t)n * 5, hipMemcpyHostToDevice, m.stream), "copy");
    mcheck(hipMemcpyAsync(dt.p, T, sizeof(double) * (size_t)n, hipMemcpyHostToDevice, m.stream), "copy");
    mcheck(hipStreamSynchronize(m.stream), "hipStreamSynchronize");
    int rc = pcfd_chem_surrogate_wdot_device(chem, &m, d, n, dr.as<double>(), dt.as<double>(), dw.as<double>(),
                                             nfallback, ms);
    if (rc != PCFD_OK) return rc;
    mcheck(hipMemcpyAsync(wdot, dw.p, sizeof(double) * (size_t)n * 5, hipMemcpyDeviceToHost, m.stream), "copy");
    mcheck(hipStreamSynchronize(m.stream), "hipStreamSynchronize");
    return PCFD_OK;
  });
}

}  // extern "C"
