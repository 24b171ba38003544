// A ucs.x-style host program on the pcfd C++ layer: builds the sin^2-bump
// mesh, runs the phases of one residual evaluation one call at a time (as
// SolutionSpace::NewtonIterate's call sites do, ucs/solutionSpace.tcc:
// 572-904), then `iters` Newton iterations, and writes what it computed to a
// binary file for comparison:
//
//   int32  nnode, nbnode, iters
//   f64    b[nnode*5]           residual of the freestream-plus-wave state
//   f64    qgrad[nnode*27]      its gradient
//   f64    resid[iters]         residGlobal per Newton iteration
//   f64    q[(nnode+nbnode)*5]  state after the iterations
//   int32  order[nnode]         SGS row order
//
// usage: ucs_newton N|MESHFILE ITERS OUT [precision]
//   N: the sin^2-bump N^3 mesh, slip floor; MESHFILE: a GMSH / UGRID /
//   FieldView / SU2 file read with RCM numbering, farfield on every face tag
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "pcfd_ucs.hpp"

using pcfd::ucs::BoundaryConditions;
using pcfd::ucs::CRS;
using pcfd::ucs::SolutionSpace;

int main(int argc, char** argv) {
  if (argc < 4) {
    std::fprintf(stderr, "usage: %s N ITERS OUT [precision]\n", argv[0]);
    return 2;
  }
  const bool from_file = std::strspn(argv[1], "0123456789") != std::strlen(argv[1]);
  const int n = from_file ? 0 : std::atoi(argv[1]), iters = std::atoi(argv[2]);
  const int precision = argc > 4 ? std::atoi(argv[4]) : PCFD_PLAIN_FP64;
  const double gamma = 1.4, mach = 0.5;

  int rc = 0;
  try {
    pcfd_mesh_spec spec;
    pcfd_mesh_default_spec(PCFD_MESH_BUMP, &spec);
    spec.nx = spec.ny = spec.nz = n;
    const pcfd::ucs::Mesh mesh = from_file ? pcfd::ucs::Mesh::Read(argv[1]) : pcfd::ucs::Mesh::Build(spec);
    const pcfd_mesh_data& m = mesh.data();

    // freestream (ucs/compressible.tcc:823-867): rho 1, |u| = M, p = 1/gamma,
    // along x for the bump; for a file at 5 deg incidence and 3 deg sideslip, so
    // that no box face is exactly tangential (the farfield BC switches branch
    // on the sign of the normal velocity)
    const double pi = 3.141592653589793;
    const double al = from_file ? 5.0 * pi / 180.0 : 0.0, be = from_file ? 3.0 * pi / 180.0 : 0.0;
    const double dir[3] = {std::cos(al) * std::cos(be), std::sin(be), std::sin(al) * std::cos(be)};
    const double p = 1.0 / gamma;
    const double qinf[5] = {1.0, mach * dir[0], mach * dir[1], mach * dir[2], p / (gamma - 1.0) + 0.5 * mach * mach};
    // bump: farfield on all sides, slip floor (face tag 5 = zmin); file: farfield everywhere
    BoundaryConditions bc;
    bc.type.assign(m.nfactag, 6);
    bc.type[0] = 0;
    if (!from_file) bc.type[5] = 3;
    for (int t = 0; t < m.nfactag; ++t)
      for (int k = 0; k < 5; ++k) bc.qref.push_back(qinf[k]);

    SolutionSpace space(pcfd::ucs::Describe(m, bc));
    CRS crs(space);
    pcfd_params prm{};
    prm.gamma = gamma;
    prm.sorder = 2;
    prm.nfirst = 0;
    prm.chi = 0.0;
    prm.limiter = 3;
    prm.venkat_K = 1.0;
    prm.limiter_refresh = 1;
    prm.limiter_freeze = 999999;
    prm.cfl = 10.0;
    prm.cfl_start = 1.0;
    prm.cfl_ramp = 5;
    prm.nsgs = 4;
    prm.symmetry2d = -1;
    prm.jac_freq = 1;
    prm.precision = precision;
    prm.gradient_terms = 9;
    space.SetParam(prm);

    // freestream with a smooth wave on the interior nodes
    const int ntot = m.nnode + m.nbedge;
    std::vector<double> q((size_t)ntot * 5);
    for (int i = 0; i < ntot; ++i) {
      double w = 0.0;
      if (i < m.nnode) w = 0.02 * std::sin(3.0 * m.xyz[3 * i]) * std::cos(2.0 * m.xyz[3 * i + 1]);
      const double r = 1.0 + w, u = mach * (1.0 - w), pp = p * (1.0 + 0.5 * w);
      const double vel[3] = {u * dir[0], u * dir[1] + 0.1 * w, u * dir[2]};
      q[(size_t)i * 5 + 0] = r;
      q[(size_t)i * 5 + 1] = r * vel[0];
      q[(size_t)i * 5 + 2] = r * vel[1];
      q[(size_t)i * 5 + 3] = r * vel[2];
      q[(size_t)i * 5 + 4] = pp / (gamma - 1.0) + 0.5 * r * (vel[0] * vel[0] + vel[1] * vel[1] + vel[2] * vel[2]);
    }

    // one residual evaluation, phase by phase
    space.SetQ(q.data(), 5);
    space.PreIterate();
    space.UpdateBCs();
    std::vector<double> qgrad((size_t)m.nnode * 27), b((size_t)m.nnode * 5);
    space.ComputeGradients(qgrad.data());
    space.ComputeLimiters();
    space.SpatialResidual(2, b.data());

    // Newton iterations from the same state
    space.SetQ(q.data(), 5);
    space.PreIterate();
    std::vector<int> order = crs.SweepOrder();
    std::vector<double> resid;
    for (int it = 0; it < iters; ++it) resid.push_back(space.NewtonIterate());
    std::vector<double> qout((size_t)ntot * 5);
    space.GetQ(qout.data(), 5);

    FILE* f = std::fopen(argv[3], "wb");
    if (!f) throw pcfd::ucs::Error(PCFD_ERR_ARG, "cannot open output");
    const int hdr[3] = {m.nnode, m.nbedge, iters};
    std::fwrite(hdr, sizeof(int), 3, f);
    std::fwrite(b.data(), sizeof(double), b.size(), f);
    std::fwrite(qgrad.data(), sizeof(double), qgrad.size(), f);
    std::fwrite(resid.data(), sizeof(double), resid.size(), f);
    std::fwrite(qout.data(), sizeof(double), qout.size(), f);
    std::fwrite(order.data(), sizeof(int), order.size(), f);
    std::fclose(f);
    std::printf("ucs_newton: %s, %d nodes, %d iterations, residGlobal %.6e -> %.6e\n", argv[1], m.nnode, iters,
                resid.empty() ? 0.0 : resid.front(), resid.empty() ? 0.0 : resid.back());
  } catch (const pcfd::ucs::Error& e) {
    std::fprintf(stderr, "pcfd error %d: %s\n", e.code(), e.what());
    rc = 1;
  }
  return rc;
}
