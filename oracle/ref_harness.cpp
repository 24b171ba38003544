// Driver over the reference solver (ProteusCFD ucs, compiled from its own
// sources by oracle/build_ref.py) that writes golden vectors for the oracle's
// tests.  TEST INFRASTRUCTURE ONLY: it runs in the build container, never on
// the GPU box, and nothing in the product links it.
//
// Usage:  ref_harness <dir/> <casename> <mode> [args]
//   mode "phases" <iter> <nsgs>: reads <dir>/q_in.bin ((nnode + gnode +
//        nbnode) x nvars doubles: interior rows, then the boundary phantoms;
//        only the conservative values are used), then runs one implicit
//        step phase by phase as SolutionSpace::NewtonIterate orders them
//        (ucs/solutionSpace.tcc:639-804) and writes every intermediate array.
//   mode "pointwise": reads <dir>/pw_in.bin records and evaluates the
//        eqnset's boundary states and ApplyDQ on each (see below).
//   mode "describe": the C-ABI descriptors the reference-side binding
//        (proteuscfd_amd/examples/ucs_binding.hpp) builds from this space.
//   mode "solve" <maxiter> <drop>: the reference's own steady loop
//        (ucs/solve.tcc:24-66: PreIterate, then PreTimeAdvance /
//        NewtonIterate / PostTimeAdvance) from freestream until residGlobal
//        falls below drop * its first value; writes the history and the
//        final state.
//   mode "trace" <n>: the same loop for n iterations, writing the last
//        iteration's Jacobian, time step, residual, correction and state.
//   mode "time" <n>: n iterations of that loop, each timed on the wall clock
//        (ref_iter_seconds.bin): the CPU baseline of bench.py.
//   mode "chemdb": writes the species database file <dir>/<casename> from
//        <dir>/chemdb.txt (see mode_chemdb), for the finite-rate eqnset.
// Arrays are written raw (little-endian float64 or int32) to
// <dir>/ref_<name>.bin; scalars go to <dir>/ref_meta.txt.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <limits>
#include <string>
#include <vector>

#include "general.h"
#include "exceptions.h"
#include "bc.h"
#include "mesh.h"
#include "param.h"
#include "eqnset.h"
#include "create_functions.h"
#include "parallel.h"
#include "solutionSpaceBase.h"
#include "solutionSpace.h"
#include "solutionOrdering.h"
#include "temporalControl.h"
#include "gradient.h"
#include "limiters.h"
#include "residual.h"
#include "timestep.h"
#include "jacobian.h"
#include "crs.h"
#include "h5layer.h"
#include <mpi.h>

#include "../proteuscfd_amd/examples/ucs_binding.hpp"

namespace {

std::string g_dir;
std::ofstream g_meta;

template <class T>
void dump(const std::string& name, const T* a, size_t n) {
  std::string path = g_dir + "ref_" + name + ".bin";
  FILE* f = std::fopen(path.c_str(), "wb");
  if (!f) {
    std::cerr << "cannot write " << path << std::endl;
    std::exit(3);
  }
  if (n) std::fwrite(a, sizeof(T), n, f);
  std::fclose(f);
}

std::vector<double> load(const std::string& name) {
  std::string path = g_dir + name;
  FILE* f = std::fopen(path.c_str(), "rb");
  if (!f) {
    std::cerr << "cannot read " << path << std::endl;
    std::exit(3);
  }
  std::fseek(f, 0, SEEK_END);
  long bytes = std::ftell(f);
  std::fseek(f, 0, SEEK_SET);
  std::vector<double> v(bytes / sizeof(double));
  if (std::fread(v.data(), sizeof(double), v.size(), f) != v.size()) std::exit(3);
  std::fclose(f);
  return v;
}

template <class T>
void meta(const char* key, T v) {
  g_meta.precision(17);
  g_meta << key << " " << v << "\n";
}

// Mesh arrays in the reference's own order (edges, half-edges, volumes).
void dump_mesh(SolutionSpace<Real>* s) {
  Mesh<Real>* m = s->m;
  const Int nnode = m->GetNumNodes(), nbnode = m->GetNumBoundaryNodes(), gnode = m->GetNumParallelNodes();
  const Int nedge = m->GetNumEdges(), nbedge = m->GetNumBoundaryEdges();
  meta("nnode", nnode);
  meta("gnode", gnode);
  meta("nbnode", nbnode);
  meta("nedge", nedge);
  meta("nbedge", nbedge);
  std::vector<Int> en(2 * (size_t)nedge), bn(2 * (size_t)nbedge), bt(nbedge);
  std::vector<Real> ea(4 * (size_t)nedge), ba(4 * (size_t)nbedge);
  for (Int e = 0; e < nedge; ++e) {
    for (int k = 0; k < 2; ++k) en[2 * e + k] = m->edges[e].n[k];
    for (int k = 0; k < 4; ++k) ea[4 * e + k] = m->edges[e].a[k];
  }
  for (Int e = 0; e < nbedge; ++e) {
    for (int k = 0; k < 2; ++k) bn[2 * e + k] = m->bedges[e].n[k];
    for (int k = 0; k < 4; ++k) ba[4 * e + k] = m->bedges[e].a[k];
    bt[e] = m->bedges[e].factag;
  }
  dump("edges", en.data(), en.size());
  dump("edge_a", ea.data(), ea.size());
  dump("bedges", bn.data(), bn.size());
  dump("bedge_a", ba.data(), ba.size());
  dump("bfactag", bt.data(), bt.size());
  dump("xyz", m->xyz, 3 * (size_t)nnode);
  dump("vol", m->vol, (size_t)nnode);
}

void dump_q(SolutionSpace<Real>* s, const char* name) {
  Mesh<Real>* m = s->m;
  const Int nvars = s->eqnset->neqn + s->eqnset->nauxvars;
  const size_t rows = (size_t)m->GetNumNodes() + m->GetNumParallelNodes() + m->GetNumBoundaryNodes();
  dump(name, s->q, rows * nvars);
}

int mode_phases(SolutionSpace<Real>* s, Int iter, Int nsgs) {
  Mesh<Real>* m = s->m;
  EqnSet<Real>* eq = s->eqnset;
  Param<Real>* p = s->param;
  const Int nnode = m->GetNumNodes();
  const Int neqn = eq->neqn, nvars = neqn + eq->nauxvars;
  const Int nrows = nnode + m->GetNumParallelNodes() + m->GetNumBoundaryNodes();
  std::vector<double> qin = load("q_in.bin");
  if (qin.size() != (size_t)nrows * nvars) {
    std::cerr << "q_in.bin holds " << qin.size() << " values, expected " << (size_t)nrows * nvars << std::endl;
    return 2;
  }
  s->PreIterate();
  s->iter = iter;
  for (Int i = 0; i < nrows; ++i) {
    for (Int k = 0; k < neqn; ++k) s->q[i * nvars + k] = qin[(size_t)i * nvars + k];
    eq->ComputeAuxiliaryVariables(&s->q[i * nvars]);
  }
  // NewtonIterate order (ucs/solutionSpace.tcc:661-720)
  UpdateBCs(s);
  dump_q(s, "q_bc");
  s->grad->Compute();
  dump("grad", s->qgrad, (size_t)nnode * s->grad->GetNterms() * 3);
  if (p->limiter) {
    s->limiter->Compute(s);
    dump("limiter", s->limiter->l, (size_t)nnode * neqn);
  }
  std::vector<Real> comps = ComputeResiduals(s);
  dump("residual", s->crs->b, (size_t)nnode * neqn);
  dump("residual_norms", comps.data(), comps.size());
  // dt of this iteration's CFL (PreTimeAdvance, :623-634), then the
  // Jacobian on it (the loop itself builds the Jacobian before the dt update)
  p->UpdateCFL(s->iter, 9999.0);
  meta("cfl", p->GetCFL());
  Real dtmin = ComputeTimesteps(s);
  meta("dtmin", dtmin);
  dump("dt", s->GetFieldData("timestep", FIELDS::STATE_NONE), (size_t)nnode);
  ComputeJacobians(s);
  CRSMatrix<Real>* A = s->crs->A;
  dump("ia", A->ia, (size_t)nnode + 1);
  dump("ja", A->ja, (size_t)A->nblocks);
  dump("jacobian", A->M, (size_t)A->nblocks * neqn * neqn);
  meta("nblocks", A->nblocks);
  A->PrepareSGS();
  s->crs->BlankX();
  Real dd = s->crs->SGS(nsgs, NULL, NULL, NULL);
  meta("sgs_delta", dd);
  dump("dq", s->crs->x, (size_t)nnode * neqn);
  // NaN/Inf rows of the correction zeroed and counted, then ApplyDQ
  // (ucs/solutionSpace.tcc:772-804)
  Int zerocount = 0;
  for (Int j = 0; j < nnode; ++j) {
    bool zero = false;
    for (Int k = 0; k < neqn; ++k)
      if (std::isnan(s->crs->x[j * neqn + k]) || std::isinf(s->crs->x[j * neqn + k])) zero = true;
    if (zero) {
      ++zerocount;
      for (Int k = 0; k < neqn; ++k) s->crs->x[j * neqn + k] = 0.0;
    }
  }
  meta("zerocount", zerocount);
  for (Int j = 0; j < nnode; ++j) eq->ApplyDQ(&s->crs->x[j * neqn], &s->q[j * nvars], &m->xyz[j * 3]);
  dump_q(s, "q_new");
  return 0;
}

// pw_in.bin: records of 32 doubles
//   [0] kind: 0 farfield state, 1 inviscid-wall state, 2 ApplyDQ
//   kinds 0/1: [1..10] QL (conservative + aux), [11..20] QR on entry (the
//              reference iterates on QR, so its entry value matters),
//              [21..24] a (unit normal, |a|), [25] vdotn; output QR (10)
//   kind 2:    [1..5] dQ, [6..15] Q; output Q after the update (10)
int mode_pointwise(SolutionSpace<Real>* s) {
  EqnSet<Real>* eq = s->eqnset;
  std::vector<double> in = load("pw_in.bin");
  const size_t nrec = in.size() / 32;
  std::vector<double> out(nrec * 10, 0.0);
  const Int nvars = eq->neqn + eq->nauxvars;
  std::vector<Real> QL(nvars), QR(nvars), Qinf(eq->Qinf, eq->Qinf + nvars), dQ(eq->neqn);
  Real xyz[3] = {0.0, 0.0, 0.0};
  for (size_t r = 0; r < nrec; ++r) {
    const double* rec = &in[r * 32];
    const int kind = (int)rec[0];
    if (kind == 0 || kind == 1) {
      for (Int k = 0; k < nvars; ++k) {
        QL[k] = rec[1 + k];
        QR[k] = rec[11 + k];
      }
      Real a[4] = {rec[21], rec[22], rec[23], rec[24]};
      if (kind == 0) eq->GetFarfieldBoundaryVariables(QL.data(), QR.data(), Qinf.data(), a, rec[25], 1.0);
      else eq->GetInviscidWallBoundaryVariables(QL.data(), QR.data(), Qinf.data(), a, rec[25], 1.0);
      for (Int k = 0; k < nvars; ++k) out[r * 10 + k] = QR[k];
    } else {
      for (Int k = 0; k < eq->neqn; ++k) dQ[k] = rec[1 + k];
      for (Int k = 0; k < nvars; ++k) QL[k] = rec[6 + k];
      eq->ApplyDQ(dQ.data(), QL.data(), xyz);
      for (Int k = 0; k < nvars; ++k) out[r * 10 + k] = QL[k];
    }
  }
  dump("pointwise", out.data(), out.size());
  return 0;
}

int mode_describe(SolutionSpace<Real>* s) {
  pcfd_binding::MeshArrays a;
  if (!pcfd_binding::DescribeMesh(s, a)) return 2;
  const pcfd_mesh_desc& d = a.desc;
  dump("desc_edges", d.edges, 2 * (size_t)d.nedge);
  dump("desc_edge_a", d.edge_a, 4 * (size_t)d.nedge);
  dump("desc_bedges", d.bedges, 2 * (size_t)d.nbnode);
  dump("desc_bedge_a", d.bedge_a, 4 * (size_t)d.nbnode);
  dump("desc_bfactag", d.bfactag, (size_t)d.nbnode);
  dump("desc_bc_type", d.bc_type, (size_t)d.nfactag);
  dump("desc_qref", d.qref, 5 * (size_t)d.nfactag);
  dump("desc_twall", d.twall, (size_t)d.nfactag);
  const pcfd_params p = pcfd_binding::DescribeParams(s, PCFD_CHOSEN_FP32);
  const double v[] = {p.gamma, (double)p.sorder, (double)p.nfirst, p.chi, (double)p.limiter, p.venkat_K,
                      (double)p.limiter_refresh, (double)p.limiter_freeze, p.cfl, p.cfl_start, (double)p.cfl_ramp,
                      (double)p.nsgs, (double)p.symmetry2d, (double)p.jac_freq, (double)p.precision,
                      (double)p.gradient_terms, (double)p.viscous, p.re_tilde, p.prandtl, p.prandtl_t,
                      p.sutherland_s, (double)p.gradient_type, (double)p.global_dt, p.fixed_dt,
                      (double)p.enable_vnn, p.vnn, (double)p.dynamic_cfl, (double)p.jac_field_type,
                      (double)p.flux_type, (double)p.jac_boundary_type};
  dump("desc_params", v, sizeof(v) / sizeof(v[0]));
  return 0;
}

int mode_solve(SolutionSpace<Real>* s, Int maxiter, Real drop) {
  std::vector<Real> hist;
  s->PreIterate();
  for (Int it = 1; it <= maxiter; ++it) {
    s->PreTimeAdvance();
    s->NewtonIterate();
    // residGlobal of this iteration (ucs/residual.tcc:47, private in
    // SolutionSpace): the norm of b, which the linear solve leaves alone
    const Real res = ParallelL2Norm(s->p, s->crs->b, s->m->GetNumNodes() * s->eqnset->neqn);
    s->PostTimeAdvance();
    hist.push_back(res);
    if (res <= drop * hist[0]) break;
  }
  meta("iterations", hist.size());
  dump("history", hist.data(), hist.size());
  dump_q(s, "q_final");
  return 0;
}

// time <n>: n iterations of the reference's own steady loop, each timed on
// the wall clock (PreTimeAdvance + NewtonIterate + PostTimeAdvance): the CPU
// baseline's per-iteration cost of the reference solver itself.
int mode_time(SolutionSpace<Real>* s, Int n) {
  std::vector<Real> secs;
  s->PreIterate();
  for (Int it = 1; it <= n; ++it) {
    const auto t0 = std::chrono::steady_clock::now();
    s->PreTimeAdvance();
    s->NewtonIterate();
    s->PostTimeAdvance();
    secs.push_back(std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count());
  }
  meta("iterations", secs.size());
  dump("iter_seconds", secs.data(), secs.size());
  return 0;
}

// trace <n>: the steady loop for n iterations; the last one's Jacobian and
// time step (after PreTimeAdvance) and residual, correction and state (after
// NewtonIterate) are written.
int mode_trace(SolutionSpace<Real>* s, Int n) {
  const Int nnode = s->m->GetNumNodes(), neqn = s->eqnset->neqn;
  s->PreIterate();
  for (Int it = 1; it <= n; ++it) {
    dump_q(s, "q_pre");
    s->PreTimeAdvance();
    if (it == n) {
      CRSMatrix<Real>* A = s->crs->A;
      dump("ia", A->ia, (size_t)nnode + 1);
      dump("ja", A->ja, (size_t)A->nblocks);
      dump("jacobian", A->M, (size_t)A->nblocks * neqn * neqn);
      dump("dt", s->GetFieldData("timestep", FIELDS::STATE_NONE), (size_t)nnode);
      dump_q(s, "q_jac");
    }
    s->NewtonIterate();
    if (it == n) {
      dump("residual", s->crs->b, (size_t)nnode * neqn);
      dump("dq", s->crs->x, (size_t)nnode * neqn);
      dump_q(s, "q_new");
    }
    s->PostTimeAdvance();
  }
  return 0;
}

// chemdb: <dir>/chemdb.txt lines "symbol MW coeff15 c1[7] c2[7]" (MW in g/mol,
// c1 the T <= 1000 K NASA-7 set, c2 the T > 1000 K set) -> the species
// database file <dir>/<casename> in the layout Species::GetDBInfo reads
// (ucs/species.tcc:140-320).  The transport curve fits it also reads are
// written as one row each, valid from 0 to 1e6 K with zero coefficients (the
// inviscid path never uses their values; the model's set-up prints them).
int mode_chemdb(const std::string& out) {
  std::ifstream in(g_dir + "chemdb.txt");
  if (!in) {
    std::cerr << "cannot read chemdb.txt" << std::endl;
    return 3;
  }
  hid_t fid = HDF_CreateFile(g_dir + out, 1);
  if (fid < 0) return 3;
  std::string sym;
  while (in >> sym) {
    Real mw, c15, c1[7], c2[7], fit[6] = {0.0, 1.0e6, 0.0, 0.0, 0.0, 0.0};
    in >> mw >> c15;
    for (int k = 0; k < 7; ++k) in >> c1[k];
    for (int k = 0; k < 7; ++k) in >> c2[k];
    const std::string dir = "/species/" + sym;
    HDF_WriteScalar(fid, dir, "MW", &mw);
    HDF_WriteScalar(fid, dir, "NASA7_burcat_coeff15", &c15);
    HDF_WriteArray(fid, dir, "NASA7_burcat1", c1, 7);
    HDF_WriteArray(fid, dir, "NASA7_burcat2", c2, 7);
    HDF_WriteArray(fid, dir, "k", fit, 1, 6);
    HDF_WriteArray(fid, dir, "mu", fit, 1, 6);
  }
  HDF_CloseFile(fid);
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc < 4) {
    std::cerr << "usage: ref_harness <dir/> <casename> phases|pointwise|describe|solve|chemdb [args]" << std::endl;
    return 1;
  }
  HDF_TurnOffErrorHandling();
  g_dir = argv[1];
  const std::string casename = argv[2], mode = argv[3];
  if (mode == "chemdb") return mode_chemdb(casename);
  MPI_Init(&argc, &argv);
  PObj<Real> pobj;
  std::vector<Param<Real>*> paramList;
  SolutionOrdering<Real> operations;
  TemporalControl<Real> temporalControl;
  ReadParamFile(paramList, casename, g_dir);
  operations.Read(casename, g_dir);
  temporalControl.Read(casename, g_dir);
  std::vector<SolutionSpaceBase<Real>*> spaces = CreateSolutionSpaces(paramList, pobj, temporalControl);
  SolutionSpace<Real>* s = static_cast<SolutionSpace<Real>*>(spaces[0]);
  g_meta.open(g_dir + "ref_meta.txt");
  meta("gamma", s->param->gamma);
  meta("Re", s->param->Re);
  meta("ref_temperature", s->param->ref_temperature);
  meta("ref_velocity", s->param->ref_velocity);
  meta("ref_density", s->param->ref_density);
  meta("ref_pressure", s->param->ref_pressure);
  meta("ref_time", s->param->ref_time);
  meta("ref_length", s->param->ref_length);
  meta("beta", s->GetFieldData("beta", FIELDS::STATE_NONE)[0]);
  if (s->param->eqnset_id == CompressibleEulerFR || s->param->eqnset_id == CompressibleNSFR)
    meta("Pref", static_cast<CompressibleFREqnSet<Real>*>(s->eqnset)->Pref);
  meta("neqn", s->eqnset->neqn);
  meta("nvars", s->eqnset->neqn + s->eqnset->nauxvars);
  meta("nterms", s->grad->GetNterms());
  {
    const Int nvars = s->eqnset->neqn + s->eqnset->nauxvars;
    dump("qinf", s->eqnset->Qinf, (size_t)nvars);
  }
  dump_mesh(s);
  int rc = 1;
  if (mode == "phases" && argc >= 6) rc = mode_phases(s, std::atoi(argv[4]), std::atoi(argv[5]));
  else if (mode == "pointwise") rc = mode_pointwise(s);
  else if (mode == "describe") rc = mode_describe(s);
  else if (mode == "solve" && argc >= 6) rc = mode_solve(s, std::atoi(argv[4]), std::atof(argv[5]));
  else if (mode == "trace" && argc >= 5) rc = mode_trace(s, std::atoi(argv[4]));
  else if (mode == "time" && argc >= 5) rc = mode_time(s, std::atoi(argv[4]));
  else std::cerr << "unknown mode or missing arguments" << std::endl;
  g_meta.close();
  MPI_Finalize();
  return rc;
}
