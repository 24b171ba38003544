// Internal context shared by the HIP translation units of libpcfd.so.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/pcfd.h"

namespace pcfd {

constexpr int kBlock = 256;  // threads per workgroup for node kernels (4 wavefronts)
constexpr int kW = 64;       // wavefront width = interleave width of the ELL/SELL layouts

// Device arrays.  Node state is structure-of-arrays: component k of node i
// lives at base[k*stride + i].
//
// gELL (node-gather lists): nodes are cut into chunks of 64 consecutive ids;
// chunk c holds gwidth[c] slots, slot s of lane l (node 64c+l) at
// index (gbase[c]+s)*64 + l.  Slots list the node's neighbours in edge order
// (the order the reference's edge loop adds their fluxes to the node: lower
// neighbours ascending, then the node's own edges in the point-to-point
// map's order), -1 after the last one.  Consecutive lanes read consecutive
// addresses.
//
// SELL (block matrix): rows grouped by colour, within a colour sorted by
// degree (descending) then id, padded to whole chunks of 64 rows (srow = -1
// for padding).  Chunk k stores its off-diagonal blocks as
// [slot][entry 0..24][lane] from sbase[k]*25*64, the diagonal block as
// [entry][lane] at k*25*64; slot s of a row is its s-th gELL neighbour,
// scol = -1 for padding.
struct Dev {
  // geometry
  double* xyz = nullptr;    // 3 x nnode
  double* vol = nullptr;    // nnode
  double* vol23 = nullptr;  // nnode, vol^(2/3) (the von Neumann time-step limit's volume factor)
  double* sw = nullptr;     // 6 x nnode, weighted LSQ moments
  // gELL gather lists
  int* gwidth = nullptr;    // nchunk_g
  int* gbase = nullptr;     // nchunk_g
  int* gcol = nullptr;      // nslot_g*64
  double* gnrm = nullptr;   // 4 x nslot_g*64: canonical (n0->n1) unit normal + |a|
  // half-edges
  int* he_ptr = nullptr;    // nnode+1: half-edges of node i, in half-edge order
  int* he_id = nullptr;     // nbnode
  int* he_tag = nullptr;    // nbnode
  double* he_a = nullptr;   // 4 x nbnode
  int* bnd_nodes = nullptr; // nodes with >= 1 half-edge
  int* bidx = nullptr;      // nnode: index into bnd_nodes, -1 for nodes without half-edges
  double* jdb = nullptr;    // 25 x nbnd boundary part of the Jacobian diagonal
  double* jqb = nullptr;    // 5 x nbnd boundary-node states after the Jacobian's BC re-evaluation
  // boundary-condition tables per face tag
  int* bc_type = nullptr;   // ntag
  double* qref = nullptr;   // 5 x ntag (conservative), aux pressure at [5*ntag + tag]
  double* twall = nullptr;  // ntag non-dimensional noSlip wall temperature (<0 adiabatic)
  int* he_nrm = nullptr;    // nbnode: wall-normal neighbour of a noSlip half-edge's node, else -1
  int* spos = nullptr;      // nnode: SELL row of each node
  // state
  double* q = nullptr;      // 5 x ntot (interior then phantoms)
  double* grad = nullptr;   // 27 x nnode (term*3+dir)
  double* lim = nullptr;    // 5 x nnode
  double* lim2 = nullptr;   // 5 x nnode (pre-clip copy)
  // pressure clip in edge order
  int* edges = nullptr;            // 2 x nedge, the edge list (n0, n1) in the reference's numbering
  unsigned char* clip_flag = nullptr;  // nnode
  int* clip_cnt = nullptr;         // per 64-edge pass
  int* clip_off = nullptr;         // passes + 1
  int* clip_list = nullptr;        // nedge (the flagged edges, in order)
  double* b = nullptr;      // 5 x nnode
  double* dt = nullptr;     // nnode
  double* dt_next = nullptr;  // nnode: the time step being computed while the Jacobian rows read dt
  double* qjac = nullptr;   // 5 x ntot: states the Jacobian rows read when hard-set nodes are committed early
  double* qrk = nullptr;    // 5 x ntot: the state a four-stage Runge-Kutta step starts from (allocated on first use)
  // SELL block matrix
  int* srow = nullptr;      // nrow_pad: node of each padded row (-1 = padding)
  int* swidth = nullptr;    // nchunk_s
  int* sbase = nullptr;     // nchunk_s
  int* scol = nullptr;      // nslot_s*64
  double* Md = nullptr;     // nslot_s*64*25 (FP64 mode)
  float* Mf = nullptr;      // nslot_s*64*25 (FP32 mode)
  double* Dd = nullptr;     // nchunk_s*64*25
  float* Df = nullptr;      // nchunk_s*64*25
  int* piv = nullptr;       // nrow_pad, 3-bit packed row permutation
  double* bsd = nullptr;    // 5 x nrow_pad right-hand side in SELL row order
  float* bsf = nullptr;
  double* xd = nullptr;     // 5 x nnode AoS correction, FP64 mode
  float* xf = nullptr;      // 5 x nnode AoS correction, FP32 mode
  // reductions
  double* red = nullptr;    // partial sums
  double* red_out = nullptr;
  int* icount = nullptr;
};

struct Ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  hipStream_t stream2 = nullptr;  // Jacobian rows, overlapped with the edge passes (default priority)
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  bool overlap = true;            // PCFD_JACOBIAN_STREAM=0 turns the overlap off
  bool jac_pending = false;
  bool has_hardset = false;       // a Dirichlet / sonic-inflow / noSlip tag exists
  std::string err;
  int nnode = 0, nbnode = 0, nedge = 0, ntot = 0, ntag = 0, nnz = 0, nbnd = 0;
  int nchunk_g = 0, nchunk_s = 0, nrow_pad = 0;
  long long nslot_g = 0, nslot_s = 0;
  long long nblocks = 0;  // nnode + 2*nedge (reference count)
  int ncolor = 0;
  bool has_noslip = false;
  std::vector<int> color_chunk;  // [ncolor+1] chunk range of each colour
  std::vector<int> h_bc_type;
  // host copies used by downloads
  std::vector<int> h_row_ptr, h_col, h_srow, h_spos, h_sbase;
  pcfd_params prm{};
  bool have_params = false;
  int iter = 1;
  double cfl = 0.0;
  double res = 0.0, res_nm1 = 0.0;  // residGlobal of the last two iterations (dynamic CFL)
  bool lu_done = false;
  bool jac_valid = false;
  int nred = 0;  // number of partial-sum blocks
  Dev d;
  // timing
  bool timing = false;
  double phase_ms[PCFD_NPHASE] = {0};
  hipEvent_t ev0 = nullptr, ev1 = nullptr;  // phase timer
  hipEvent_t ev2 = nullptr, ev3 = nullptr;  // kernel sub-timer inside a phase
  hipEvent_t ev4 = nullptr, ev5 = nullptr;  // Jacobian rows on the second stream
  long long device_bytes = 0;
};

#define PCFD_HIP(call)                                                                                   \
  do {                                                                                                   \
    hipError_t e_ = (call);                                                                              \
    if (e_ != hipSuccess) throw std::runtime_error(std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

inline int nblk(long long n, int bs = kBlock) { return (int)((n + bs - 1) / bs); }

// Workgroups are handed to the 8 XCDs round-robin (workgroup b runs on XCD
// b%8).  xcd_block() renumbers them so XCD x runs one contiguous range of
// logical blocks: neighbouring nodes (which share gathered neighbour data)
// then sit in one XCD's L2.  A bijection on [0, gridDim.x).
__device__ __forceinline__ int xcd_block() {
  const int nb = gridDim.x, b = blockIdx.x;
  const int q = nb >> 3, r = nb & 7, x = b & 7;
  return x * q + (x < r ? x : r) + (b >> 3);
}

// gELL neighbour lists as kernel arguments
struct Gell {
  const int* width;
  const int* base;
  const int* col;
  const double* nrm;
  long long stride;  // nslot_g*64, component stride of nrm
};

__device__ __forceinline__ long long gbase_of(const Gell& g, int i) {
  return (long long)g.base[i >> 6] * kW + (i & 63);
}

// Per-block partial sums of b^2 for the residual norms: total and per
// equation, reduced with a fixed-shape tree (deterministic), written to
// red[k*gridDim.x + blk].  Every thread of the block must call it.
__device__ __forceinline__ void block_norm_partials(const double* bi, double* red, int blk) {
  __shared__ double sh[6][kBlock];
  double tot = 0.0;
#pragma unroll
  for (int k = 0; k < 5; ++k) {
    double v = bi[k] * bi[k];
    sh[1 + k][threadIdx.x] = v;
    tot += v;
  }
  sh[0][threadIdx.x] = tot;
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
#pragma unroll
      for (int k = 0; k < 6; ++k) sh[k][threadIdx.x] += sh[k][threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < 6; ++k) red[k * gridDim.x + blk] = sh[k][0];
  }
}

// gELL traversal helpers: slot s of node i
__host__ __device__ inline long long gell_index(const int* gbase, int i, int s) {
  return ((long long)gbase[i >> 6] + s) * kW + (i & 63);
}

// timers: phase k (events ev0/ev1) and kernel-only sub-timers (ev2/ev3)
void phase_begin(Ctx& c);
void phase_end(Ctx& c, int k);
void sub_begin(Ctx& c);
void sub_end(Ctx& c, int k);

// launchers (each enqueues on ctx.stream; throw std::runtime_error on failure)
void launch_lsq_moments(Ctx& c);
void launch_update_bcs(Ctx& c);
void launch_gradient(Ctx& c, int nterms);
void launch_limiter(Ctx& c);
void launch_gradient_limiter(Ctx& c, int nterms);  // LSQ gradient + limiter in one pass, then the clip
void launch_residual(Ctx& c, int higher_order);  // inviscid, then viscous / noSlip edits
void launch_viscous(Ctx& c, int higher_order);
void launch_norms(Ctx& c, double* host_norms6);
void launch_timestep(Ctx& c, double cfl);
double launch_dtmin(Ctx& c);
void launch_jacobian(Ctx& c);
void launch_jacobian_overlapped(Ctx& c);  // rows on stream2; jacobian_join() before the matrix is used
void jacobian_join(Ctx& c);
void launch_prepare_sgs(Ctx& c);
void launch_sgs(Ctx& c, int nsgs, double* delta_xnorm);
void launch_apply_dq(Ctx& c, int explicit_mode, int* nzeroed);
void launch_zero_count(Ctx& c);                          // the zeroed-row counter
void launch_rk_stage(Ctx& c, double a, const double* q0);  // q = q0 + a dt/vol b (ApplyDQ)

}  // namespace pcfd
