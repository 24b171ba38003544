// ucs_binding.hpp -- the binding a ucs.x maintainer adds on the reference side
// to hand a SolutionSpace to libpcfd.so.  It includes the reference's own
// headers (ucs/solutionSpace.h and what it pulls in) and include/pcfd.h, and
// fills the C ABI's descriptors from the reference's objects:
//   DescribeMesh    Mesh<Real> edges/half-edges/metrics + BoundaryConditions
//                   (ucs/uns_base.h:12-37, ucs/mesh.h:199-254, ucs/bc.h:31-33)
//   DescribeParams  Param<Real> after PostCompute (ucs/param.h:50-178)
// The reference's build compiles it as part of ucs.x; here it is compiled
// against the reference's headers by oracle/build_ref.py (its harness checks
// the descriptors it produces against this repository's own mesh and
// parameters).  It calls no pcfd_* function itself, so it links without
// libpcfd.so; a caller then passes the descriptors to pcfd_create /
// pcfd_set_params (or pcfd::ucs::SolutionSpace in include/pcfd_ucs.hpp).
#pragma once

#include <vector>

#include "pcfd.h"

namespace pcfd_binding {

// Owns the arrays a pcfd_mesh_desc points at.
struct MeshArrays {
  std::vector<int> edges, bedges, bfactag, bc_type;
  std::vector<double> edge_a, bedge_a, xyz, vol, qref, twall;
  pcfd_mesh_desc desc{};
};

// The mesh of one SolutionSpace after BuildMaps + CalcMetrics.  One GPU holds
// one partition with no parallel ghosts (gnode == 0): half-edge k's phantom
// is node nnode + k.
template <class SpaceT>
bool DescribeMesh(SpaceT* s, MeshArrays& a) {
  auto* m = s->m;
  auto* bc = s->bc;
  auto* param = s->param;
  const int nnode = m->GetNumNodes(), nedge = m->GetNumEdges(), nbedge = m->GetNumBoundaryEdges();
  if (m->GetNumParallelNodes() != 0) return false;
  a.edges.resize(2 * (size_t)nedge);
  a.edge_a.resize(4 * (size_t)nedge);
  for (int e = 0; e < nedge; ++e) {  // Edges<Real>, ucs/uns_base.h:12-21
    a.edges[2 * e] = m->edges[e].n[0];
    a.edges[2 * e + 1] = m->edges[e].n[1];
    for (int k = 0; k < 4; ++k) a.edge_a[4 * e + k] = m->edges[e].a[k];
  }
  a.bedges.resize(2 * (size_t)nbedge);
  a.bedge_a.resize(4 * (size_t)nbedge);
  a.bfactag.resize(nbedge);
  for (int e = 0; e < nbedge; ++e) {  // HalfEdges<Real>, ucs/uns_base.h:27-36
    a.bedges[2 * e] = m->bedges[e].n[0];
    a.bedges[2 * e + 1] = m->bedges[e].n[1];
    for (int k = 0; k < 4; ++k) a.bedge_a[4 * e + k] = m->bedges[e].a[k];
    a.bfactag[e] = m->bedges[e].factag;
  }
  a.xyz.assign(m->xyz, m->xyz + 3 * (size_t)nnode);
  a.vol.assign(m->vol, m->vol + (size_t)nnode);
  // per face tag: BCTypes id, Qref (conservative) and the non-dimensional
  // wall temperature of noSlip tags (ucs/bc.tcc:45-68, 1284-1290)
  const int ntag = bc->largest_bc_id + 1;
  a.bc_type.assign(ntag, 0);
  a.qref.assign(5 * (size_t)ntag, 0.0);
  a.twall.assign(ntag, -1.0);
  for (int t = 1; t < ntag; ++t) {
    auto* obj = bc->GetBCObj(t);
    a.bc_type[t] = obj->GetBCType();
    const double* q = obj->Qref ? obj->Qref : s->eqnset->Qinf;
    for (int k = 0; k < 5; ++k) a.qref[5 * t + k] = q[k];
    if (a.bc_type[t] == 4 /* noSlip */) a.twall[t] = obj->twall / param->ref_temperature;
  }
  a.desc = pcfd_mesh_desc{nnode, nbedge, nedge, a.edges.data(), a.edge_a.data(), a.bedges.data(),
                          a.bedge_a.data(), a.bfactag.data(), a.xyz.data(), a.vol.data(), ntag,
                          a.bc_type.data(), a.qref.data(), a.twall.data()};
  return true;
}

// Param<Real> -> pcfd_params (ucs/param.h; the keywords of ucs/param.tcc).
template <class SpaceT>
pcfd_params DescribeParams(SpaceT* s, int precision) {
  auto* param = s->param;
  pcfd_params p{};
  p.gamma = param->gamma;
  p.sorder = param->sorder;
  p.nfirst = param->nFirstOrderSteps;
  p.chi = param->chi;
  p.limiter = param->limiter;
  p.venkat_K = 1.0;  // the modified Venkatakrishnan limiter's K is fixed at 1 (ucs/limiters.tcc:566-567)
  p.limiter_refresh = param->limiterRefresh;
  p.limiter_freeze = param->limiterFreeze;
  p.cfl = param->cfl;
  p.cfl_start = param->cflStart;
  p.cfl_ramp = param->cflRampingSteps;
  p.nsgs = param->nSgs;
  p.symmetry2d = param->symmetry2D;
  p.jac_freq = param->jacobianFreq;
  p.precision = precision;  // PCFD_PLAIN_FP64, or PCFD_CHOSEN_FP32
  p.gradient_terms = 9;     // {rho, rho u, rho v, rho w, E, T, u, v, w} (ucs/compressible.tcc:1008-1027)
  p.viscous = param->viscous;
  // Re / Mach with the freestream velocity as the Mach number (ucs/compressible.tcc:752-755)
  p.re_tilde = param->viscous ? param->Re / param->GetVelocity(s->iter) : 1.0;
  p.prandtl = param->Pr;
  p.prandtl_t = param->PrT;
  p.sutherland_s = 110.4 / param->ref_temperature;  // ucs/eqnset.h:251-268
  p.gradient_type = param->gradType;
  p.global_dt = !param->useLocalTimeStepping;
  p.fixed_dt = param->dt;
  p.enable_vnn = param->enableVNN;
  p.vnn = param->VNN;
  p.dynamic_cfl = param->dynamicCFL;
  p.jac_field_type = param->fieldJacType;
  p.flux_type = param->flux_id == AlgebraicFluxType ? PCFD_FLUX_ALGEBRAIC : PCFD_FLUX_ROE;
  p.jac_boundary_type = param->boundaryJacType;
  return p;
}

}  // namespace pcfd_binding
