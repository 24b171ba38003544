// Internal context shared by the HIP translation units of libpcfd.so.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/pcfd.h"

namespace pcfd {

constexpr int kBlock = 256;  // threads per workgroup for node kernels (4 wavefronts)

// Device arrays.  State is structure-of-arrays: component k of node i lives
// at base[k*stride + i].
struct Dev {
  // geometry
  double* xyz = nullptr;    // 3 x nnode
  double* vol = nullptr;    // nnode
  double* sw = nullptr;     // 6 x nnode, weighted LSQ moments
  // node-gather CSR over interior edges, neighbours ascending
  int* row_ptr = nullptr;   // nnode+1
  int* col = nullptr;       // nnz (= 2*nedge)
  double* nrm = nullptr;    // 4 x nnz: canonical (n0->n1) unit normal + |a| of the entry's edge
  int* blk_t = nullptr;     // nnz: block index of the transposed entry (col row, this node col)
  // half-edges
  int* he_ptr = nullptr;    // nnode+1: half-edges of node i, in half-edge order
  int* he_id = nullptr;     // nbnode
  int* he_tag = nullptr;    // nbnode
  double* he_a = nullptr;   // 4 x nbnode
  int* bnd_nodes = nullptr; // nodes with >= 1 half-edge
  // boundary-condition tables per face tag
  int* bc_type = nullptr;   // ntag
  double* qref = nullptr;   // 5 x ntag (conservative), + 1 x ntag aux pressure at [5*ntag]
  // state
  double* q = nullptr;      // 5 x ntot (interior then phantoms)
  double* grad = nullptr;   // 27 x nnode (term*3+dir)
  double* lim = nullptr;    // 5 x nnode
  double* lim2 = nullptr;   // 5 x nnode (pre-clip copy)
  double* b = nullptr;      // 5 x nnode
  double* dt = nullptr;     // nnode
  double* xd = nullptr;     // 5 x nnode (x in FP64 mode)
  float* xf = nullptr;      // 5 x nnode (x in FP32 mode)
  // block CRS: row i blocks at [blk_ptr[i], blk_ptr[i+1]), diagonal first
  int* blk_ptr = nullptr;   // nnode+1
  double* Md = nullptr;     // nblocks*25 (FP64 mode)
  float* Mf = nullptr;      // nblocks*25 (FP32 mode)
  int* piv = nullptr;       // 5 x nnode
  // colouring
  int* color_nodes = nullptr;
  // reductions
  double* red = nullptr;    // partial sums
  double* red_out = nullptr;
  int* icount = nullptr;
};

struct Ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  int nnode = 0, nbnode = 0, nedge = 0, ntot = 0, ntag = 0, nnz = 0, nbnd = 0;
  long long nblocks = 0;
  int ncolor = 0;
  std::vector<int> color_ptr;  // host copy
  std::vector<int> h_bc_type;
  pcfd_params prm{};
  bool have_params = false;
  int iter = 1;
  double cfl = 0.0;
  bool lu_done = false;
  bool jac_valid = false;
  int nred = 0;  // number of partial-sum blocks
  Dev d;
  // timing
  bool timing = false;
  double phase_ms[9] = {0};
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  long long device_bytes = 0;
};

#define PCFD_HIP(call)                                                                   \
  do {                                                                                   \
    hipError_t e_ = (call);                                                              \
    if (e_ != hipSuccess) throw std::runtime_error(std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

inline int nblk(long long n, int bs = kBlock) { return (int)((n + bs - 1) / bs); }

// launchers (each enqueues on ctx.stream; throw std::runtime_error on failure)
void launch_lsq_moments(Ctx& c);
void launch_update_bcs(Ctx& c);
void launch_gradient(Ctx& c, int nterms);
void launch_limiter(Ctx& c);
void launch_residual(Ctx& c, int higher_order);
void launch_norms(Ctx& c, double* host_norms6);
void launch_timestep(Ctx& c, double cfl);
double launch_dtmin(Ctx& c);
void launch_jacobian(Ctx& c);
void launch_prepare_sgs(Ctx& c);
void launch_sgs(Ctx& c, int nsgs, double* delta_xnorm);
void launch_apply_dq(Ctx& c, int explicit_mode, int* nzeroed);

}  // namespace pcfd
