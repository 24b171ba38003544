// Finite-rate chemistry: mass production rates, source term and its
// finite-difference Jacobian, one thread per cell, FP64.
//
// Reference loop (ucs/chem.tcc:575-583, ucs/reaction.tcc:762-856): for each
// species i, for each reaction j, a rate computed from scratch.  The
// per-reaction quantities (concentrations X, Kf, Kb = Kf/Kc, the reactant
// and product products, the third-body concentration) do not depend on i,
// so each reaction is evaluated once and its contribution
// (nu_ij * M_j * net_j) * MW_i added to species i in reaction order, which
// is the reference's arithmetic in the reference's summation order.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/pcfd_chem.h"
#include "chem_model.h"

namespace pcfd {

__global__ void __launch_bounds__(kChemBlock) k_wdot(const ChemDev* __restrict__ md, int n,
                                                    const double* __restrict__ rhoi, const double* __restrict__ T,
                                                    double* __restrict__ wdot) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  const ChemDev& m = *md;
  double r[PCFD_CHEM_MAXS], w[PCFD_CHEM_MAXS];
  for (int k = 0; k < m.ns; ++k) r[k] = rhoi[(long long)c * m.ns + k];
  wdot_of(m, r, T[c], w);
  for (int k = 0; k < m.ns; ++k) wdot[(long long)c * m.ns + k] = w[k];
}

// SourceTerm + SourceTermJacobian (ucs/eqnset.tcc:162-187): one-sided
// differences with h = 1e-8 on every FR variable.  The velocity columns are
// exactly zero (the source does not read u, v, w: (S - S)/h = 0) and the
// momentum and energy rows are zero, so only the species and T columns
// are evaluated.
__global__ void __launch_bounds__(kChemBlock) k_source(const ChemDev* __restrict__ md, int n,
                                                      const double* __restrict__ q, const double* __restrict__ vol,
                                                      double* __restrict__ src, double* __restrict__ jac) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  const ChemDev& m = *md;
  const int ns = m.ns, neqn = ns + 4;
  double Q[PCFD_CHEM_MAXS + 4], s[PCFD_CHEM_MAXS], sp[PCFD_CHEM_MAXS];
  for (int k = 0; k < neqn; ++k) Q[k] = q[(long long)c * neqn + k];
  const double v = vol[c];
  source_of(m, Q, v, s);
  double* so = src + (long long)c * neqn;
  for (int k = 0; k < neqn; ++k) so[k] = k < ns ? s[k] : 0.0;
  if (!jac) return;
  double* J = jac + (long long)c * neqn * neqn;
  for (int k = 0; k < neqn * neqn; ++k) J[k] = 0.0;
  const double h = 1.0e-8;
  for (int i = 0; i < neqn; ++i) {
    if (i >= ns && i < ns + 3) continue;
    const double keep = Q[i];
    Q[i] += h;
    source_of(m, Q, v, sp);
    Q[i] = keep;
    for (int j = 0; j < ns; ++j) J[j * neqn + i] = (sp[j] - s[j]) / h;
  }
}

}  // namespace pcfd

using namespace pcfd;

struct pcfd_chem {
  ChemDev host{};
  ChemDev* dev = nullptr;
  hipStream_t stream = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  float last_ms = 0.f;
  std::string err;
};

namespace {
thread_local std::string g_chem_err;

void check(hipError_t e, const char* what) {
  if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

template <typename F>
int guarded_chem(pcfd_chem* c, F&& f) {
  if (!c) return PCFD_ERR_ARG;
  try {
    return f(*c);
  } catch (const std::exception& e) {
    c->err = e.what();
    return PCFD_ERR_HIP;
  }
}

struct DevBuf {
  void* p = nullptr;
  explicit DevBuf(size_t bytes) {
    if (bytes) check(hipMalloc(&p, bytes), "hipMalloc");
  }
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  template <typename T>
  T* as() const {
    return (T*)p;
  }
};

int launch_source(pcfd_chem& c, int n, const double* q, const double* vol, double* src, double* jac) {
  if (n == 0) return PCFD_OK;
  check(hipEventRecord(c.e0, c.stream), "hipEventRecord");
  hipLaunchKernelGGL(k_source, dim3((n + kChemBlock - 1) / kChemBlock), dim3(kChemBlock), 0, c.stream, c.dev, n, q,
                     vol, src, jac);
  check(hipGetLastError(), "k_source");
  check(hipEventRecord(c.e1, c.stream), "hipEventRecord");
  check(hipEventSynchronize(c.e1), "hipEventSynchronize");
  check(hipEventElapsedTime(&c.last_ms, c.e0, c.e1), "hipEventElapsedTime");
  return PCFD_OK;
}
}  // namespace

namespace pcfd {
// accessors for the surrogate (kernels_mlp.hip)
const ChemDev* chem_device_model(const pcfd_chem* c) { return c->dev; }
int chem_nspecies(const pcfd_chem* c) { return c->host.ns; }
const double* chem_mw(const pcfd_chem* c) { return c->host.mw; }
}  // namespace pcfd

extern "C" {

int pcfd_chem_create(const pcfd_chem_desc* d, pcfd_chem** out) {
  if (!d || !out) {
    g_chem_err = "pcfd_chem_create: null argument";
    return PCFD_ERR_ARG;
  }
  *out = nullptr;
  if (d->nspecies < 1 || d->nspecies > PCFD_CHEM_MAXS || d->nreactions < 0 || d->nreactions > PCFD_CHEM_MAXR ||
      !d->mw || !d->nasa_lo || !d->nasa_hi || (d->nreactions > 0 && (!d->rxn_type || !d->third_body || !d->A ||
                                                                    !d->EA || !d->n || !d->nloc || !d->loc ||
                                                                    !d->nup || !d->nupp || !d->tb_eff)) ||
      !(d->ref_density > 0.0) || !(d->ref_time > 0.0) || !(d->ref_temperature > 0.0)) {
    g_chem_err = "pcfd_chem_create: invalid model description";
    return PCFD_ERR_ARG;
  }
  ChemDev m{};
  m.ns = d->nspecies;
  m.nr = d->nreactions;
  for (int s = 0; s < m.ns; ++s) {
    if (!(d->mw[s] > 0.0)) {
      g_chem_err = "pcfd_chem_create: molecular weights must be positive (kg/mol)";
      return PCFD_ERR_ARG;
    }
    m.mw[s] = d->mw[s];
    for (int k = 0; k < 7; ++k) {
      m.lo[s][k] = d->nasa_lo[7 * s + k];
      m.hi[s][k] = d->nasa_hi[7 * s + k];
    }
  }
  for (int r = 0; r < m.nr; ++r) {
    if (d->rxn_type[r] < 0 || d->rxn_type[r] > 3 || d->nloc[r] < 1 || d->nloc[r] > PCFD_CHEM_MAXS) {
      g_chem_err = "pcfd_chem_create: reaction " + std::to_string(r) + " invalid (type or species count)";
      return PCFD_ERR_ARG;
    }
    m.type[r] = d->rxn_type[r];
    m.tb[r] = d->third_body[r];
    m.A[r] = d->A[r];
    m.EA[r] = d->EA[r];
    m.n[r] = d->n[r];
    m.nloc[r] = d->nloc[r];
    for (int k = 0; k < m.nloc[r]; ++k) {
      const int g = d->loc[PCFD_CHEM_MAXS * r + k];
      if (g < 0 || g >= m.ns) {
        g_chem_err = "pcfd_chem_create: reaction " + std::to_string(r) + " names a species out of range";
        return PCFD_ERR_ARG;
      }
      m.loc[r][k] = g;
      m.nup[r][k] = d->nup[PCFD_CHEM_MAXS * r + k];
      m.nupp[r][k] = d->nupp[PCFD_CHEM_MAXS * r + k];
      m.eff[r][k] = d->tb_eff[PCFD_CHEM_MAXS * r + k];
    }
  }
  m.ref_density = d->ref_density;
  m.ref_time = d->ref_time;
  m.ref_temperature = d->ref_temperature;
  pcfd_chem* c = new pcfd_chem();
  c->host = m;
  try {
    check(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking), "hipStreamCreate");
    check(hipEventCreate(&c->e0), "hipEventCreate");
    check(hipEventCreate(&c->e1), "hipEventCreate");
    check(hipMalloc((void**)&c->dev, sizeof(ChemDev)), "hipMalloc");
    check(hipMemcpy(c->dev, &c->host, sizeof(ChemDev), hipMemcpyHostToDevice), "hipMemcpy");
  } catch (const std::exception& e) {
    g_chem_err = e.what();
    pcfd_chem_destroy(c);
    return PCFD_ERR_HIP;
  }
  *out = c;
  return PCFD_OK;
}

int pcfd_chem_destroy(pcfd_chem* c) {
  if (!c) return PCFD_ERR_ARG;
  if (c->dev) (void)hipFree(c->dev);
  if (c->e0) (void)hipEventDestroy(c->e0);
  if (c->e1) (void)hipEventDestroy(c->e1);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return PCFD_OK;
}

const char* pcfd_chem_last_error(const pcfd_chem* c) { return c ? c->err.c_str() : g_chem_err.c_str(); }

double pcfd_chem_last_kernel_ms(const pcfd_chem* c) { return c ? (double)c->last_ms : 0.0; }

int pcfd_chem_wdot(pcfd_chem* c, int n, const double* rhoi, const double* T, double* wdot) {
  return guarded_chem(c, [&](pcfd_chem& cc) -> int {
    if (n < 0 || (n > 0 && (!rhoi || !T || !wdot))) {
      cc.err = "pcfd_chem_wdot: invalid arguments";
      return PCFD_ERR_ARG;
    }
    if (n == 0) return PCFD_OK;
    const int ns = cc.host.ns;
    DevBuf dr(sizeof(double) * (size_t)n * ns), dt(sizeof(double) * (size_t)n), dw(sizeof(double) * (size_t)n * ns);
    check(hipMemcpyAsync(dr.p, rhoi, sizeof(double) * (size_t)n * ns, hipMemcpyHostToDevice, cc.stream), "copy");
    check(hipMemcpyAsync(dt.p, T, sizeof(double) * (size_t)n, hipMemcpyHostToDevice, cc.stream), "copy");
    check(hipEventRecord(cc.e0, cc.stream), "hipEventRecord");
    hipLaunchKernelGGL(k_wdot, dim3((n + kChemBlock - 1) / kChemBlock), dim3(kChemBlock), 0, cc.stream, cc.dev, n,
                       dr.as<double>(), dt.as<double>(), dw.as<double>());
    check(hipGetLastError(), "k_wdot");
    check(hipEventRecord(cc.e1, cc.stream), "hipEventRecord");
    check(hipMemcpyAsync(wdot, dw.p, sizeof(double) * (size_t)n * ns, hipMemcpyDeviceToHost, cc.stream), "copy");
    check(hipStreamSynchronize(cc.stream), "hipStreamSynchronize");
    check(hipEventElapsedTime(&cc.last_ms, cc.e0, cc.e1), "hipEventElapsedTime");
    return PCFD_OK;
  });
}

int pcfd_chem_wdot_device(pcfd_chem* c, int n, const double* rhoi, const double* T, double* wdot) {
  return guarded_chem(c, [&](pcfd_chem& cc) -> int {
    if (n < 0 || (n > 0 && (!rhoi || !T || !wdot))) {
      cc.err = "pcfd_chem_wdot_device: invalid arguments";
      return PCFD_ERR_ARG;
    }
    if (n == 0) return PCFD_OK;
    check(hipEventRecord(cc.e0, cc.stream), "hipEventRecord");
    hipLaunchKernelGGL(k_wdot, dim3((n + kChemBlock - 1) / kChemBlock), dim3(kChemBlock), 0, cc.stream, cc.dev, n,
                       rhoi, T, wdot);
    check(hipGetLastError(), "k_wdot");
    check(hipEventRecord(cc.e1, cc.stream), "hipEventRecord");
    check(hipEventSynchronize(cc.e1), "hipEventSynchronize");
    check(hipEventElapsedTime(&cc.last_ms, cc.e0, cc.e1), "hipEventElapsedTime");
    return PCFD_OK;
  });
}

int pcfd_chem_source(pcfd_chem* c, int n, const double* q, const double* vol, double* src, double* jac) {
  return guarded_chem(c, [&](pcfd_chem& cc) -> int {
    if (n < 0 || (n > 0 && (!q || !vol || !src))) {
      cc.err = "pcfd_chem_source: invalid arguments";
      return PCFD_ERR_ARG;
    }
    if (n == 0) return PCFD_OK;
    const size_t neqn = (size_t)cc.host.ns + 4;
    DevBuf dq(sizeof(double) * n * neqn), dv(sizeof(double) * n), ds(sizeof(double) * n * neqn),
        dj(jac ? sizeof(double) * n * neqn * neqn : 0);
    check(hipMemcpyAsync(dq.p, q, sizeof(double) * n * neqn, hipMemcpyHostToDevice, cc.stream), "copy");
    check(hipMemcpyAsync(dv.p, vol, sizeof(double) * n, hipMemcpyHostToDevice, cc.stream), "copy");
    launch_source(cc, n, dq.as<double>(), dv.as<double>(), ds.as<double>(), jac ? dj.as<double>() : nullptr);
    check(hipMemcpyAsync(src, ds.p, sizeof(double) * n * neqn, hipMemcpyDeviceToHost, cc.stream), "copy");
    if (jac)
      check(hipMemcpyAsync(jac, dj.p, sizeof(double) * n * neqn * neqn, hipMemcpyDeviceToHost, cc.stream), "copy");
    check(hipStreamSynchronize(cc.stream), "hipStreamSynchronize");
    return PCFD_OK;
  });
}

int pcfd_chem_source_device(pcfd_chem* c, int n, const double* q, const double* vol, double* src, double* jac) {
  return guarded_chem(c, [&](pcfd_chem& cc) -> int {
    if (n < 0 || (n > 0 && (!q || !vol || !src))) {
      cc.err = "pcfd_chem_source_device: invalid arguments";
      return PCFD_ERR_ARG;
    }
    return launch_source(cc, n, q, vol, src, jac);
  });
}

}  // extern "C"
