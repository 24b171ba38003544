/* C++ host layer over the pcfd C ABI, named after the reference classes it
 * stands in for, so a ucs.x call site changes its receiver and little else.
 *
 *   pcfd::ucs::SolutionSpace  the hot-path part of SolutionSpace<Real>
 *                             (ucs/solutionSpace.h; NewtonIterate at
 *                             ucs/solutionSpace.tcc:572-904)
 *   pcfd::ucs::CRS            CRS<Real> (ucs/crs.h:13-81) over the device
 *                             block matrix (CRSMatrix, ucs/crsmatrix.h:10-107)
 *   pcfd::ucs::Mesh           a library-built mesh: Mesh<Real>::Read* readers
 *                             (ucs/mesh.tcc) or the synthetic generator
 *   pcfd::ucs::Error          thrown where the reference would Abort
 *                             (ucs/exceptions.h:30-49); carries the ABI code
 *
 * Header-only, C++17, depends on pcfd.h and pcfd_mesh.h only.  Arrays cross
 * in the reference layouts: q[node*nvars+k], qgrad[node*27+t*3+d],
 * l/b/x[node*5+k], M[blk*25+r*5+c] in ia/ja order.
 */
#ifndef PCFD_UCS_HPP
#define PCFD_UCS_HPP

#include <stdexcept>
#include <string>
#include <vector>

#include "pcfd.h"
#include "pcfd_mesh.h"

namespace pcfd {
namespace ucs {

class Error : public std::runtime_error {
 public:
  Error(int code, const std::string& what) : std::runtime_error(what), code_(code) {}
  int code() const { return code_; }

 private:
  int code_;
};

inline void Check(const pcfd_ctx* ctx, int rc, const char* what) {
  if (rc != PCFD_OK) throw Error(rc, std::string(what) + ": " + pcfd_last_error(ctx));
}

/* Boundary conditions per face tag: BCTypes id (ucs/bc_defines.h:4-30) and
 * the reference state Qref (5 conservative values, ucs/bc.tcc:45-68). */
struct BoundaryConditions {
  std::vector<int> type;     /* [nfactag] */
  std::vector<double> qref;  /* [nfactag*5] */
  std::vector<double> twall; /* [nfactag] noSlip wall temperature / T_ref, <0 adiabatic; empty: adiabatic */
};

/* Owner of a mesh built by the library: read from a file with the reader
 * the reference's Mesh would use (Mesh<Type>::Read*, ucs/mesh.tcc), or
 * generated.  data() is what Describe() takes. */
class Mesh {
 public:
  static Mesh Read(const std::string& path, bool rcm = true, int format = PCFD_MESH_FMT_AUTO) {
    pcfd_mesh* m = nullptr;
    if (pcfd_mesh_read(path.c_str(), format, rcm ? 1 : 0, &m) != 0) throw Error(PCFD_ERR_ARG, pcfd_mesh_error());
    return Mesh(m);
  }
  static Mesh ReadGMSH_Master(const std::string& path) { return Read(path, true, PCFD_MESH_FMT_GMSH); }
  static Mesh ReadUGRID_Ascii(const std::string& path) { return Read(path, true, PCFD_MESH_FMT_UGRID); }
  static Mesh ReadCRUNCH_Ascii(const std::string& path) { return Read(path, true, PCFD_MESH_FMT_FIELDVIEW); }
  static Mesh ReadSU2_Ascii(const std::string& path) { return Read(path, true, PCFD_MESH_FMT_SU2); }
  static Mesh Build(const pcfd_mesh_spec& spec) {
    pcfd_mesh* m = nullptr;
    if (pcfd_mesh_build(&spec, &m) != 0) throw Error(PCFD_ERR_ARG, pcfd_mesh_error());
    return Mesh(m);
  }
  Mesh(Mesh&& o) noexcept : m_(o.m_) { o.m_ = nullptr; }
  Mesh& operator=(Mesh&& o) noexcept {
    if (this != &o) {
      if (m_) pcfd_mesh_free(m_);
      m_ = o.m_;
      o.m_ = nullptr;
    }
    return *this;
  }
  Mesh(const Mesh&) = delete;
  Mesh& operator=(const Mesh&) = delete;
  ~Mesh() {
    if (m_) pcfd_mesh_free(m_);
  }
  const pcfd_mesh_data& data() const { return *pcfd_mesh_get(m_); }

 private:
  explicit Mesh(pcfd_mesh* m) : m_(m) {}
  pcfd_mesh* m_ = nullptr;
};

/* Mesh arrays as CalcMetrics leaves them (ucs/mesh.h:199-254): edges with
 * n0<n1, half-edges with phantom ids nnode+k.  The pointers stay owned by
 * `mesh`; pcfd_create copies them. */
inline pcfd_mesh_desc Describe(const pcfd_mesh_data& mesh, const BoundaryConditions& bc) {
  if ((int)bc.type.size() != mesh.nfactag || (int)bc.qref.size() != 5 * mesh.nfactag)
    throw Error(PCFD_ERR_ARG, "BoundaryConditions: need one type and 5 Qref values per face tag");
  pcfd_mesh_desc d{};
  d.nnode = mesh.nnode;
  d.nbnode = mesh.nbedge;
  d.nedge = mesh.nedge;
  d.edges = mesh.edges;
  d.edge_a = mesh.edge_a;
  d.bedges = mesh.bedges;
  d.bedge_a = mesh.bedge_a;
  d.bfactag = mesh.bfactag;
  d.xyz = mesh.xyz;
  d.vol = mesh.vol;
  d.nfactag = mesh.nfactag;
  d.bc_type = bc.type.data();
  d.qref = bc.qref.data();
  d.twall = bc.twall.empty() ? nullptr : bc.twall.data();
  if (!bc.twall.empty() && (int)bc.twall.size() != mesh.nfactag)
    throw Error(PCFD_ERR_ARG, "BoundaryConditions: twall needs one value per face tag");
  return d;
}

/* Device-resident solution space: q, qgrad, limiter, residual, dt and the
 * block system live on the GPU; the host sees them only through Get/Set. */
class SolutionSpace {
 public:
  explicit SolutionSpace(const pcfd_mesh_desc& mesh) : nnode_(mesh.nnode), nbnode_(mesh.nbnode) {
    Check(nullptr, pcfd_create(&mesh, &ctx_), "pcfd_create");
  }
  ~SolutionSpace() {
    if (ctx_) pcfd_destroy(ctx_);
  }
  SolutionSpace(const SolutionSpace&) = delete;
  SolutionSpace& operator=(const SolutionSpace&) = delete;

  pcfd_ctx* ctx() const { return ctx_; }
  int nnode() const { return nnode_; }
  int nbnode() const { return nbnode_; }

  /* Param keywords (ucs/param.tcc) */
  void SetParam(const pcfd_params& p) { Check(ctx_, pcfd_set_params(ctx_, &p), "pcfd_set_params"); }

  /* q of interior nodes then phantoms, nvars per node (5 or 10) */
  void SetQ(const double* q, int nvars) { Check(ctx_, pcfd_upload_q(ctx_, q, nvars), "pcfd_upload_q"); }
  void GetQ(double* q, int nvars) const { Check(ctx_, pcfd_download_q(ctx_, q, nvars), "pcfd_download_q"); }

  /* UpdateBCs (ucs/bc.tcc:1399-1457) */
  void UpdateBCs() { Check(ctx_, pcfd_update_bcs(ctx_), "UpdateBCs"); }
  /* Gradient::Compute (ucs/gradient.tcc:56-112); qgrad may be null */
  void ComputeGradients(double* qgrad = nullptr) { Check(ctx_, pcfd_gradient(ctx_, qgrad), "Gradient::Compute"); }
  /* Limiter::Compute (ucs/limiters.tcc:52-132); l may be null */
  void ComputeLimiters(double* l = nullptr) { Check(ctx_, pcfd_limiter(ctx_, l), "Limiter::Compute"); }
  /* SpatialResidual (ucs/residual.tcc:12-122): returns residGlobal,
   * components[0..4] receive the per-equation norms when given */
  double SpatialResidual(int iter, double* b = nullptr, double* components = nullptr) {
    double n[6];
    Check(ctx_, pcfd_residual(ctx_, iter, b, n), "SpatialResidual");
    if (components)
      for (int k = 0; k < 5; ++k) components[k] = n[k + 1];
    return n[0];
  }
  /* ComputeTimesteps (ucs/timestep.tcc:6-77): returns dt_min */
  double ComputeTimesteps(double cfl, double* dt = nullptr) {
    double dtmin = 0.0;
    Check(ctx_, pcfd_timestep(ctx_, cfl, dt, &dtmin), "ComputeTimesteps");
    return dtmin;
  }
  /* ComputeJacobians (ucs/jacobian.tcc:12-250) into the device CRS */
  void ComputeJacobians() { Check(ctx_, pcfd_jacobian(ctx_), "ComputeJacobians"); }
  /* NaN/Inf rows of x zeroed, then ApplyDQ (ucs/solutionSpace.tcc:772-804);
   * returns the number of rows zeroed */
  int ApplyDQ() {
    int nz = 0;
    Check(ctx_, pcfd_apply_dq(ctx_, &nz), "ApplyDQ");
    return nz;
  }
  /* ExplicitSolve (ucs/solve.tcc:70-140) */
  int ExplicitSolve() {
    int nz = 0;
    Check(ctx_, pcfd_explicit_update(ctx_, &nz), "ExplicitSolve");
    return nz;
  }
  /* One NewtonIterate (ucs/solutionSpace.tcc:572-904) in the reference phase
   * order; returns residGlobal of this iteration */
  double NewtonIterate(double* components = nullptr) {
    double r[6];
    Check(ctx_, pcfd_iterate(ctx_, r), "NewtonIterate");
    if (components)
      for (int k = 0; k < 5; ++k) components[k] = r[k + 1];
    return r[0];
  }
  /* PreIterate: iteration counter 1, limiter 1, first time step */
  void PreIterate() { Check(ctx_, pcfd_reset_iteration(ctx_), "PreIterate"); }
  int iter() const { return pcfd_get_iteration(ctx_); }
  /* Solve loop until residGlobal <= drop * residGlobal(first) or max_iter */
  int Solve(int max_iter, double drop, std::vector<double>* history = nullptr, double* secs = nullptr) {
    std::vector<double> h((size_t)max_iter);
    int n = 0;
    double s = 0.0;
    Check(ctx_, pcfd_solve(ctx_, max_iter, drop, h.data(), &n, &s), "Solve");
    if (history) history->assign(h.begin(), h.begin() + n);
    if (secs) *secs = s;
    return n;
  }
  void Synchronize() { Check(ctx_, pcfd_synchronize(ctx_), "pcfd_synchronize"); }

 private:
  pcfd_ctx* ctx_ = nullptr;
  int nnode_ = 0, nbnode_ = 0;
};

/* CRS<Real> over the device block system of a SolutionSpace.  The matrix
 * is filled by SolutionSpace::ComputeJacobians and b by SpatialResidual, as
 * in the reference where both write crs->A and crs->b. */
class CRS {
 public:
  explicit CRS(SolutionSpace& space) : space_(space) {}
  /* CRSMatrix::PrepareSGS (ucs/crsmatrix.tcc:839-876): LU of the diagonal */
  void PrepareSGS() { Check(space_.ctx(), pcfd_prepare_sgs(space_.ctx()), "PrepareSGS"); }
  /* BlankX + CRS::SGS (ucs/crs.tcc:61-173): x0 = 0, nSgs symmetric sweeps;
   * returns |xOld - xNorm| of the last two sweeps like the reference; x (AoS
   * node*5+k) receives the solution when given */
  double SGS(int nSgs, double* x = nullptr) {
    double d = 0.0;
    Check(space_.ctx(), pcfd_sgs(space_.ctx(), nSgs, x, &d), "CRS::SGS");
    return d;
  }
  /* CRSMatrix::M copied out in the caller's ia/ja order (diagonal first per
   * row, then neighbours ascending, ucs/crsmatrix.tcc:47-97) */
  void GetMatrix(const int* ia, const int* ja, double* M) const {
    Check(space_.ctx(), pcfd_download_jacobian(space_.ctx(), ia, ja, M), "CRSMatrix::M");
  }
  /* Row order of the multicolour sweep (Gauss-Seidel order) */
  std::vector<int> SweepOrder() const {
    std::vector<int> o((size_t)space_.nnode());
    Check(space_.ctx(), pcfd_sgs_order(space_.ctx(), o.data()), "pcfd_sgs_order");
    return o;
  }

 private:
  SolutionSpace& space_;
};

}  // namespace ucs
}  // namespace pcfd

#endif
