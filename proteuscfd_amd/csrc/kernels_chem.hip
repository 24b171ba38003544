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

// Species::GetCp (ucs/species.tcc:43-53) with the coefficient set of
// GetThermoCoeff (T > 1000 K: high), J/(kg K)
__device__ __forceinline__ double cp_of(const ChemDev& m, int s, double T) {
  const double* a = T > 1000.0 ? m.hi[s] : m.lo[s];
  const double cp_R = a[0] + T * (a[1] + T * (a[2] + T * (a[3] + T * a[4])));
  return cp_R * (kUnivR / m.mw[s]);
}

// ChemModel::GetP (ucs/chem.tcc:988-997): Dalton's law over the species'
// ideal-gas EOS, P += (rho_i R_i) T in species order, SI units
__device__ __forceinline__ double p_of(const ChemDev& m, const double* rhoiDim, double TDim) {
  double P = 0.0;
  for (int i = 0; i < m.ns; ++i) P += rhoiDim[i] * (kUnivR / m.mw[i]) * TDim;
  return P;
}

// ComputeAuxiliaryVariables (ucs/compressibleFR.tcc:754-814), one thread per
// state.  The states are AoS (ns + 4 doubles each, as the reference stores
// them): a block stages its slice of q and of aux through LDS so that the
// HBM reads and writes are contiguous across the block.
__global__ void __launch_bounds__(kChemBlock) k_fr_aux(const ChemDev* __restrict__ mp, int n,
                                                      const double* __restrict__ q, double pref, double s_ref,
                                                      double* __restrict__ aux) {
  __shared__ double sh[kChemBlock * (2 + 2 * PCFD_CHEM_MAXS)];
  const ChemDev& m = *mp;
  const int ns = m.ns, neqn = ns + 4, naux = 2 + 2 * ns;
  const long long b0 = (long long)blockIdx.x * blockDim.x;
  const int nb = (int)(n - b0 < (long long)blockDim.x ? n - b0 : (long long)blockDim.x);
  for (int k = threadIdx.x; k < nb * neqn; k += blockDim.x) sh[k] = q[b0 * neqn + k];
  __syncthreads();
  const int t = threadIdx.x;
  double X[2 + 2 * PCFD_CHEM_MAXS];
  if (t < nb) {
    const double* Q = sh + t * neqn;
    double rho = 0.0, rd[PCFD_CHEM_MAXS];
    for (int i = 0; i < ns; ++i) {
      rho += Q[i];
      rd[i] = Q[i] * m.ref_density;
    }
    const double TDim = Q[ns + 3] * m.ref_temperature;
    X[0] = p_of(m, rd, TDim) / pref;
    X[1] = rho;
    for (int i = 0; i < ns; ++i) {
      X[2 + i] = (cp_of(m, i, TDim) - kUnivR / m.mw[i]) / s_ref;
      X[2 + ns + i] = Q[i] / m.mw[i] / 1000.0;
    }
  }
  __syncthreads();
  if (t < nb)
    for (int k = 0; k < naux; ++k) sh[t * naux + k] = X[k];
  __syncthreads();
  for (int k = threadIdx.x; k < nb * naux; k += blockDim.x) aux[b0 * naux + k] = sh[k];
}

// NewtonFindTGivenP (ucs/compressibleFR.tcc:2342-2378), one thread per state
__global__ void __launch_bounds__(kChemBlock) k_fr_t_from_p(const ChemDev* __restrict__ mp, int n, const double* __restrict__ rhoi,
                                                           const double* __restrict__ p,
                                                           const double* __restrict__ Tinit, double pref,
                                                           double* __restrict__ T, int* __restrict__ iters) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const ChemDev& m = *mp;
  const int ns = m.ns, maxit = 28;
  const double tol = 1.0e-15, Tref = m.ref_temperature;
  double rd[PCFD_CHEM_MAXS];
  for (int i = 0; i < ns; ++i) rd[i] = rhoi[(long long)t * ns + i] * m.ref_density;
  double TDim = Tinit[t] * Tref;
  const double PgoalDim = p[t] * pref;
  int j = 0;
  for (j = 0; j < maxit; ++j) {
    const double TpDim = TDim + 1.0e-8;
    const double z = PgoalDim - p_of(m, rd, TDim);
    const double zp = PgoalDim - p_of(m, rd, TpDim);
    const double dzdT = (zp - z) / (TpDim - TDim);
    const double dT = -z / dzdT;
    if (fabs(dT / Tref) < tol) break;
    TDim += dT;
  }
  T[t] = TDim / Tref;
  if (iters) iters[t] = j;
}

int pcfd_chem_fr_aux(pcfd_chem* c, int n, const double* q, double ref_pressure, double ref_velocity, double* aux) {
  return guarded_chem(c, [&](pcfd_chem& cc) -> int {
    if (n < 0 || (n > 0 && (!q || !aux)) || !(ref_pressure > 0.0) || !(ref_velocity > 0.0)) {
      cc.err = "pcfd_chem_fr_aux: invalid arguments (n >= 0, buffers, ref_pressure > 0, ref_velocity > 0)";
      return PCFD_ERR_ARG;
    }
    if (n == 0) return PCFD_OK;
    const int ns = cc.host.ns;
    const size_t nq = (size_t)n * (ns + 4), na = (size_t)n * (2 + 2 * ns);
    DevBuf dq(sizeof(double) * nq), da(sizeof(double) * na);
    check(hipMemcpyAsync(dq.p, q, sizeof(double) * nq, hipMemcpyHostToDevice, cc.stream), "copy");
    check(hipEventRecord(cc.e0, cc.stream), "hipEventRecord");
    const double s_ref = ref_velocity * ref_velocity / cc.host.ref_temperature;
    hipLaunchKernelGGL(k_fr_aux, dim3((n + kChemBlock - 1) / kChemBlock), dim3(kChemBlock), 0, cc.stream, cc.dev, n,
                       dq.as<double>(), ref_pressure, s_ref, da.as<double>());
    check(hipGetLastError(), "k_fr_aux");
    check(hipEventRecord(cc.e1, cc.stream), "hipEventRecord");
    check(hipMemcpyAsync(aux, da.p, sizeof(double) * na, hipMemcpyDeviceToHost, cc.stream), "copy");
    check(hipStreamSynchronize(cc.stream), "hipStreamSynchronize");
    check(hipEventElapsedTime(&cc.last_ms, cc.e0, cc.e1), "hipEventElapsedTime");
    return PCFD_OK;
  });
}

int pcfd_chem_t_from_p(pcfd_chem* c, int n, const double* rhoi, const double* p, const double* Tinit,
                       double ref_pressure, double* T, int* iters) {
  return guarded_chem(c, [&](pcfd_chem& cc) -> int {
    if (n < 0 || (n > 0 && (!rhoi || !p || !Tinit || !T)) || !(ref_pressure > 0.0)) {
      cc.err = "pcfd_chem_t_from_p: invalid arguments (n >= 0, buffers, ref_pressure > 0)";
      return PCFD_ERR_ARG;
    }
    if (n == 0) return PCFD_OK;
    const int ns = cc.host.ns;
    DevBuf dr(sizeof(double) * (size_t)n * ns), dp(sizeof(double) * (size_t)n), dt0(sizeof(double) * (size_t)n),
        dT(sizeof(double) * (size_t)n), di(sizeof(int) * (size_t)n);
    check(hipMemcpyAsync(dr.p, rhoi, sizeof(double) * (size_t)n * ns, hipMemcpyHostToDevice, cc.stream), "copy");
    check(hipMemcpyAsync(dp.p, p, sizeof(double) * (size_t)n, hipMemcpyHostToDevice, cc.stream), "copy");
    check(hipMemcpyAsync(dt0.p, Tinit, sizeof(double) * (size_t)n, hipMemcpyHostToDevice, cc.stream), "copy");
    check(hipEventRecord(cc.e0, cc.stream), "hipEventRecord");
    hipLaunchKernelGGL(k_fr_t_from_p, dim3((n + kChemBlock - 1) / kChemBlock), dim3(kChemBlock), 0, cc.stream,
                       cc.dev, n, dr.as<double>(), dp.as<double>(), dt0.as<double>(), ref_pressure, dT.as<double>(),
                       di.as<int>());
    check(hipGetLastError(), "k_fr_t_from_p");
    check(hipEventRecord(cc.e1, cc.stream), "hipEventRecord");
    check(hipMemcpyAsync(T, dT.p, sizeof(double) * (size_t)n, hipMemcpyDeviceToHost, cc.stream), "copy");
    if (iters) check(hipMemcpyAsync(iters, di.p, sizeof(int) * (size_t)n, hipMemcpyDeviceToHost, cc.stream), "copy");
    check(hipStreamSynchronize(cc.stream), "hipStreamSynchronize");
    check(hipEventElapsedTime(&cc.last_ms, cc.e0, cc.e1), "hipEventElapsedTime");
    return PCFD_OK;
  });
}

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
