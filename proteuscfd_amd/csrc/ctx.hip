// C ABI of libpcfd.so: context lifetime, transfers and single phases.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <stdexcept>

#include "../../include/pcfd_mesh.h"
#include "common.h"

using namespace pcfd;

struct pcfd_ctx {
  Ctx c;
};

namespace {

thread_local std::string g_create_err;

template <typename T>
void dalloc(Ctx& c, T** p, size_t n) {
  if (n == 0) n = 1;
  PCFD_HIP(hipMalloc((void**)p, n * sizeof(T)));
  c.device_bytes += (long long)(n * sizeof(T));
}

template <typename T>
void dput(Ctx& c, T** p, const std::vector<T>& h) {
  dalloc(c, p, h.size());
  if (!h.empty()) PCFD_HIP(hipMemcpy(*p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
}

template <typename T>
void dfree(T*& p) {
  if (p) (void)hipFree(p);
  p = nullptr;
}

void free_all(Ctx& c) {
  Dev& d = c.d;
  dfree(d.xyz); dfree(d.vol); dfree(d.sw); dfree(d.row_ptr); dfree(d.col); dfree(d.nrm); dfree(d.blk_t);
  dfree(d.he_ptr); dfree(d.he_id); dfree(d.he_tag); dfree(d.he_a); dfree(d.bnd_nodes); dfree(d.bc_type);
  dfree(d.qref); dfree(d.q); dfree(d.grad); dfree(d.lim); dfree(d.lim2); dfree(d.b); dfree(d.dt); dfree(d.xd);
  dfree(d.xf); dfree(d.blk_ptr); dfree(d.Md); dfree(d.Mf); dfree(d.piv); dfree(d.color_nodes); dfree(d.red);
  dfree(d.red_out); dfree(d.icount);
  if (c.ev0) (void)hipEventDestroy(c.ev0);
  if (c.ev1) (void)hipEventDestroy(c.ev1);
  if (c.stream) (void)hipStreamDestroy(c.stream);
  c.ev0 = c.ev1 = nullptr;
  c.stream = nullptr;
}

double aux_pressure_host(const double* Q, double gm1) {
  double u = Q[1] / Q[0], v = Q[2] / Q[0], w = Q[3] / Q[0];
  double V2 = u * u + v * v + w * w;
  return gm1 * (Q[4] - 0.5 * Q[0] * V2);
}

template <typename F>
int guarded(pcfd_ctx* ctx, F&& f) {
  if (!ctx) return PCFD_ERR_ARG;
  try {
    return f(ctx->c);
  } catch (const std::exception& e) {
    ctx->c.err = e.what();
    return PCFD_ERR_HIP;
  }
}

int need_params(Ctx& c) {
  if (!c.have_params) {
    c.err = "pcfd_set_params has not been called";
    return PCFD_ERR_STATE;
  }
  return PCFD_OK;
}

void ensure_matrix(Ctx& c) {
  if (c.prm.precision == PCFD_CHOSEN_FP32) {
    if (!c.d.Mf) dalloc(c, &c.d.Mf, (size_t)c.nblocks * 25);
    if (!c.d.xf) dalloc(c, &c.d.xf, 5 * (size_t)c.nnode);
  } else {
    if (!c.d.Md) dalloc(c, &c.d.Md, (size_t)c.nblocks * 25);
  }
}

}  // namespace

// phase timing helpers (shared with solver.hip)
namespace pcfd {
void phase_begin(Ctx& c) {
  if (c.timing) PCFD_HIP(hipEventRecord(c.ev0, c.stream));
}
void phase_end(Ctx& c, int k) {
  if (!c.timing) return;
  PCFD_HIP(hipEventRecord(c.ev1, c.stream));
  PCFD_HIP(hipEventSynchronize(c.ev1));
  float ms = 0.f;
  PCFD_HIP(hipEventElapsedTime(&ms, c.ev0, c.ev1));
  c.phase_ms[k] += ms;
}
void ensure_matrix_alloc(Ctx& c) { ensure_matrix(c); }
}  // namespace pcfd

extern "C" {

const char* pcfd_version(void) { return "pcfd-gfx950 0.1"; }

const char* pcfd_last_error(const pcfd_ctx* ctx) {
  if (!ctx) return g_create_err.c_str();
  return ctx->c.err.c_str();
}

int pcfd_create(const pcfd_mesh_desc* m, pcfd_ctx** out) {
  if (!m || !out) {
    g_create_err = "pcfd_create: null argument";
    return PCFD_ERR_ARG;
  }
  *out = nullptr;
  if (m->nnode <= 0 || m->nbnode < 0 || m->nedge < 0 || m->nfactag <= 0 || !m->xyz || !m->vol || !m->bc_type ||
      !m->qref || (m->nedge > 0 && (!m->edges || !m->edge_a)) ||
      (m->nbnode > 0 && (!m->bedges || !m->bedge_a || !m->bfactag))) {
    g_create_err = "pcfd_create: invalid mesh description";
    return PCFD_ERR_ARG;
  }
  const int nn = m->nnode, nb = m->nbnode, ne = m->nedge, nt = m->nfactag;
  // validate indices on the host before anything reaches the device
  for (long long e = 0; e < ne; ++e) {
    int a = m->edges[2 * e], b = m->edges[2 * e + 1];
    if (a < 0 || b < 0 || a >= nn || b >= nn || a == b) {
      g_create_err = "pcfd_create: edge " + std::to_string(e) + " has invalid nodes";
      return PCFD_ERR_ARG;
    }
  }
  for (int h = 0; h < nb; ++h) {
    int a = m->bedges[2 * h], ph = m->bedges[2 * h + 1], tag = m->bfactag[h];
    if (a < 0 || a >= nn || ph != nn + h || tag < 0 || tag >= nt) {
      g_create_err = "pcfd_create: half-edge " + std::to_string(h) +
                     " invalid (node range, phantom id must be nnode+k, factag range)";
      return PCFD_ERR_ARG;
    }
  }
  pcfd_ctx* ctx = new pcfd_ctx();
  Ctx& c = ctx->c;
  try {
    PCFD_HIP(hipGetDevice(&c.device));
    PCFD_HIP(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking));
    PCFD_HIP(hipEventCreate(&c.ev0));
    PCFD_HIP(hipEventCreate(&c.ev1));
    c.nnode = nn;
    c.nbnode = nb;
    c.nedge = ne;
    c.ntot = nn + nb;
    c.ntag = nt;
    c.nnz = 2 * ne;
    // node-gather CSR, neighbours ascending, with the canonical edge normal
    std::vector<int> deg(nn + 1, 0);
    for (int e = 0; e < ne; ++e) {
      deg[m->edges[2 * e] + 1]++;
      deg[m->edges[2 * e + 1] + 1]++;
    }
    std::vector<int> row_ptr(nn + 1, 0);
    for (int i = 0; i < nn; ++i) row_ptr[i + 1] = row_ptr[i] + deg[i + 1];
    std::vector<std::pair<int, int>> ent(c.nnz);
    std::vector<int> fill(row_ptr.begin(), row_ptr.end() - 1);
    for (int e = 0; e < ne; ++e) {
      int a = m->edges[2 * e], b = m->edges[2 * e + 1];
      ent[fill[a]++] = {b, e};
      ent[fill[b]++] = {a, e};
    }
    for (int i = 0; i < nn; ++i) std::sort(ent.begin() + row_ptr[i], ent.begin() + row_ptr[i + 1]);
    std::vector<int> col(c.nnz);
    std::vector<double> nrm(4 * (size_t)c.nnz);
    for (int p = 0; p < c.nnz; ++p) {
      col[p] = ent[p].first;
      int e = ent[p].second;
      for (int k = 0; k < 4; ++k) nrm[(size_t)k * c.nnz + p] = m->edge_a[4 * (size_t)e + k];
    }
    for (int i = 0; i < nn; ++i)
      for (int p = row_ptr[i] + 1; p < row_ptr[i + 1]; ++p)
        if (col[p] == col[p - 1]) throw std::runtime_error("pcfd_create: duplicate edge between nodes");
    ent.clear();
    ent.shrink_to_fit();
    // half-edges per node in half-edge order
    std::vector<int> hdeg(nn + 1, 0);
    for (int h = 0; h < nb; ++h) hdeg[m->bedges[2 * h] + 1]++;
    std::vector<int> he_ptr(nn + 1, 0);
    for (int i = 0; i < nn; ++i) he_ptr[i + 1] = he_ptr[i] + hdeg[i + 1];
    std::vector<int> he_id(nb), hfill(he_ptr.begin(), he_ptr.end() - 1);
    for (int h = 0; h < nb; ++h) he_id[hfill[m->bedges[2 * h]]++] = h;
    std::vector<int> bnd;
    for (int i = 0; i < nn; ++i)
      if (he_ptr[i + 1] > he_ptr[i]) bnd.push_back(i);
    c.nbnd = (int)bnd.size();
    std::vector<int> he_tag(m->bfactag, m->bfactag + nb);
    std::vector<double> he_a(4 * (size_t)nb);
    for (int h = 0; h < nb; ++h)
      for (int k = 0; k < 4; ++k) he_a[(size_t)k * nb + h] = m->bedge_a[4 * (size_t)h + k];
    // SoA geometry
    std::vector<double> xyz(3 * (size_t)nn);
    for (int i = 0; i < nn; ++i)
      for (int k = 0; k < 3; ++k) xyz[(size_t)k * nn + i] = m->xyz[3 * (size_t)i + k];
    std::vector<double> vol(m->vol, m->vol + nn);
    c.h_bc_type.assign(m->bc_type, m->bc_type + nt);
    std::vector<double> qref(6 * (size_t)nt, 0.0);
    for (int t = 0; t < nt; ++t)
      for (int k = 0; k < 5; ++k) qref[(size_t)k * nt + t] = m->qref[5 * (size_t)t + k];
    // block-CRS pattern: row i = [diag, neighbours ascending]
    std::vector<int> blk_ptr(nn + 1);
    for (int i = 0; i <= nn; ++i) blk_ptr[i] = row_ptr[i] + i;
    c.nblocks = (long long)nn + c.nnz;
    // colouring for the SGS sweeps
    std::vector<int> colour(nn);
    int ncol = pcfd_color_nodes(nn, ne, m->edges, colour.data());
    if (ncol < 0) throw std::runtime_error(pcfd_mesh_error());
    c.ncolor = ncol;
    c.color_ptr.assign(ncol + 1, 0);
    for (int i = 0; i < nn; ++i) c.color_ptr[colour[i] + 1]++;
    for (int k = 0; k < ncol; ++k) c.color_ptr[k + 1] += c.color_ptr[k];
    std::vector<int> cnodes(nn), cfill(c.color_ptr.begin(), c.color_ptr.end() - 1);
    for (int i = 0; i < nn; ++i) cnodes[cfill[colour[i]]++] = i;

    Dev& d = c.d;
    dput(c, &d.xyz, xyz);
    dput(c, &d.vol, vol);
    dalloc(c, &d.sw, 6 * (size_t)nn);
    dput(c, &d.row_ptr, row_ptr);
    dput(c, &d.col, col);
    dput(c, &d.nrm, nrm);
    dput(c, &d.he_ptr, he_ptr);
    dput(c, &d.he_id, he_id);
    dput(c, &d.he_tag, he_tag);
    dput(c, &d.he_a, he_a);
    dput(c, &d.bnd_nodes, bnd);
    dput(c, &d.bc_type, c.h_bc_type);
    dput(c, &d.qref, qref);
    dalloc(c, &d.q, 5 * (size_t)c.ntot);
    dalloc(c, &d.grad, 27 * (size_t)nn);
    dalloc(c, &d.lim, 5 * (size_t)nn);
    dalloc(c, &d.lim2, 5 * (size_t)nn);
    dalloc(c, &d.b, 5 * (size_t)nn);
    dalloc(c, &d.dt, (size_t)nn);
    dalloc(c, &d.xd, 5 * (size_t)nn);
    dput(c, &d.blk_ptr, blk_ptr);
    dalloc(c, &d.piv, (size_t)nn);
    dput(c, &d.color_nodes, cnodes);
    dalloc(c, &d.red, 6 * (size_t)std::max(nblk(nn), 1024));
    dalloc(c, &d.red_out, 8);
    dalloc(c, &d.icount, 4);
    PCFD_HIP(hipMemset(d.grad, 0, 27 * sizeof(double) * (size_t)nn));
    launch_lsq_moments(c);
    PCFD_HIP(hipStreamSynchronize(c.stream));
  } catch (const std::exception& e) {
    g_create_err = e.what();
    free_all(c);
    delete ctx;
    return PCFD_ERR_HIP;
  }
  *out = ctx;
  return PCFD_OK;
}

int pcfd_destroy(pcfd_ctx* ctx) {
  if (!ctx) return PCFD_ERR_ARG;
  (void)hipStreamSynchronize(ctx->c.stream);
  free_all(ctx->c);
  delete ctx;
  return PCFD_OK;
}

int pcfd_set_params(pcfd_ctx* ctx, const pcfd_params* p) {
  return guarded(ctx, [&](Ctx& c) -> int {
    if (!p) {
      c.err = "pcfd_set_params: null params";
      return PCFD_ERR_ARG;
    }
    if (p->gamma <= 1.0 || p->sorder < 1 || p->sorder > 3 || p->limiter < 0 || p->limiter > 3 || p->nsgs < 0 ||
        (p->precision != PCFD_PLAIN_FP64 && p->precision != PCFD_CHOSEN_FP32) ||
        (p->gradient_terms != 5 && p->gradient_terms != 9)) {
      c.err = "pcfd_set_params: parameter out of range";
      return PCFD_ERR_ARG;
    }
    c.prm = *p;
    if (c.prm.limiter_refresh <= 0) c.prm.limiter_refresh = 1;
    if (c.prm.jac_freq <= 0) c.prm.jac_freq = 1;
    c.have_params = true;
    // aux pressure of each reference state (ComputeAuxiliaryVariables form)
    std::vector<double> qref(6 * (size_t)c.ntag);
    PCFD_HIP(hipMemcpy(qref.data(), c.d.qref, qref.size() * sizeof(double), hipMemcpyDeviceToHost));
    for (int t = 0; t < c.ntag; ++t) {
      double Q[5];
      for (int k = 0; k < 5; ++k) Q[k] = qref[(size_t)k * c.ntag + t];
      qref[5 * (size_t)c.ntag + t] = aux_pressure_host(Q, c.prm.gamma - 1.0);
    }
    PCFD_HIP(hipMemcpy(c.d.qref, qref.data(), qref.size() * sizeof(double), hipMemcpyHostToDevice));
    if (c.prm.nsgs > 0) ensure_matrix(c);
    c.jac_valid = false;
    c.lu_done = false;
    return PCFD_OK;
  });
}

int pcfd_info(const pcfd_ctx* ctx, int* ncolor, long long* nblocks, long long* device_bytes) {
  if (!ctx) return PCFD_ERR_ARG;
  if (ncolor) *ncolor = ctx->c.ncolor;
  if (nblocks) *nblocks = ctx->c.nblocks;
  if (device_bytes) *device_bytes = ctx->c.device_bytes;
  return PCFD_OK;
}

int pcfd_upload_q(pcfd_ctx* ctx, const double* q, int nvars) {
  return guarded(ctx, [&](Ctx& c) -> int {
    if (!q || nvars < 5) {
      c.err = "pcfd_upload_q: need q and nvars >= 5";
      return PCFD_ERR_ARG;
    }
    std::vector<double> s(5 * (size_t)c.ntot);
    for (int i = 0; i < c.ntot; ++i)
      for (int k = 0; k < 5; ++k) s[(size_t)k * c.ntot + i] = q[(size_t)i * nvars + k];
    PCFD_HIP(hipMemcpyAsync(c.d.q, s.data(), s.size() * sizeof(double), hipMemcpyHostToDevice, c.stream));
    PCFD_HIP(hipStreamSynchronize(c.stream));
    return PCFD_OK;
  });
}

int pcfd_download_q(pcfd_ctx* ctx, double* q, int nvars) {
  return guarded(ctx, [&](Ctx& c) -> int {
    if (!q || nvars < 5) {
      c.err = "pcfd_download_q: need q and nvars >= 5";
      return PCFD_ERR_ARG;
    }
    std::vector<double> s(5 * (size_t)c.ntot);
    PCFD_HIP(hipMemcpyAsync(s.data(), c.d.q, s.size() * sizeof(double), hipMemcpyDeviceToHost, c.stream));
    PCFD_HIP(hipStreamSynchronize(c.stream));
    const double gamma = c.have_params ? c.prm.gamma : 1.4;
    for (int i = 0; i < c.ntot; ++i) {
      double* Q = q + (size_t)i * nvars;
      for (int k = 0; k < 5; ++k) Q[k] = s[(size_t)k * c.ntot + i];
      if (nvars >= 10) {
        double gm1 = gamma - 1.0;
        double u = Q[1] / Q[0], v = Q[2] / Q[0], w = Q[3] / Q[0];
        double P = gm1 * (Q[4] - 0.5 * Q[0] * (u * u + v * v + w * w));
        double v2h = 0.5 * (u * u + v * v + w * w);
        Q[5] = gamma * (gm1 * (Q[4] - Q[0] * v2h)) / Q[0];
        Q[6] = P;
        Q[7] = u;
        Q[8] = v;
        Q[9] = w;
      }
    }
    return PCFD_OK;
  });
}

int pcfd_update_bcs(pcfd_ctx* ctx) {
  return guarded(ctx, [&](Ctx& c) -> int {
    if (int r = need_params(c)) return r;
    launch_update_bcs(c);
    PCFD_HIP(hipStreamSynchronize(c.stream));
    return PCFD_OK;
  });
}

static int download_soa(Ctx& c, const double* dsrc, int ncomp, double* host) {
  std::vector<double> s((size_t)ncomp * c.nnode);
  PCFD_HIP(hipMemcpyAsync(s.data(), dsrc, s.size() * sizeof(double), hipMemcpyDeviceToHost, c.stream));
  PCFD_HIP(hipStreamSynchronize(c.stream));
  for (int i = 0; i < c.nnode; ++i)
    for (int k = 0; k < ncomp; ++k) host[(size_t)i * ncomp + k] = s[(size_t)k * c.nnode + i];
  return PCFD_OK;
}

static int upload_soa(Ctx& c, double* ddst, int ncomp, const double* host) {
  std::vector<double> s((size_t)ncomp * c.nnode);
  for (int i = 0; i < c.nnode; ++i)
    for (int k = 0; k < ncomp; ++k) s[(size_t)k * c.nnode + i] = host[(size_t)i * ncomp + k];
  PCFD_HIP(hipMemcpyAsync(ddst, s.data(), s.size() * sizeof(double), hipMemcpyHostToDevice, c.stream));
  PCFD_HIP(hipStreamSynchronize(c.stream));
  return PCFD_OK;
}

int pcfd_gradient(pcfd_ctx* ctx, double* qgrad_out) {
  return guarded(ctx, [&](Ctx& c) -> int {
    if (int r = need_params(c)) return r;
    int nt = qgrad_out ? 9 : c.prm.gradient_terms;
    launch_gradient(c, nt);
    if (qgrad_out) return download_soa(c, c.d.grad, 27, qgrad_out);
    PCFD_HIP(hipStreamSynchronize(c.stream));
    return PCFD_OK;
  });
}

int pcfd_limiter(pcfd_ctx* ctx, double* l_out) {
  return guarded(ctx, [&](Ctx& c) -> int {
    if (int r = need_params(c)) return r;
    if (c.prm.limiter == 0) {
      c.err = "pcfd_limiter: limiter is off (limiter == 0)";
      return PCFD_ERR_STATE;
    }
    launch_limiter(c);
    if (l_out) return download_soa(c, c.d.lim, 5, l_out);
    PCFD_HIP(hipStreamSynchronize(c.stream));
    return PCFD_OK;
  });
}

int pcfd_set_limiter(pcfd_ctx* ctx, const double* l) {
  return guarded(ctx, [&](Ctx& c) -> int {
    if (!l) {
      c.err = "pcfd_set_limiter: null";
      return PCFD_ERR_ARG;
    }
    return upload_soa(c, c.d.lim, 5, l);
  });
}

int pcfd_residual(pcfd_ctx* ctx, int iter, double* b_out, double norms[6]) {
  return guarded(ctx, [&](Ctx& c) -> int {
    if (int r = need_params(c)) return r;
    int ho = (c.prm.sorder > 1 && iter > c.prm.nfirst) ? 1 : 0;
    launch_residual(c, ho);
    launch_norms(c, norms);
    if (b_out) return download_soa(c, c.d.b, 5, b_out);
    PCFD_HIP(hipStreamSynchronize(c.stream));
    return PCFD_OK;
  });
}

int pcfd_timestep(pcfd_ctx* ctx, double cfl, double* dt_out, double* dtmin) {
  return guarded(ctx, [&](Ctx& c) -> int {
    if (int r = need_params(c)) return r;
    launch_timestep(c, cfl);
    if (dtmin) *dtmin = launch_dtmin(c);
    if (dt_out) return download_soa(c, c.d.dt, 1, dt_out);
    PCFD_HIP(hipStreamSynchronize(c.stream));
    return PCFD_OK;
  });
}

int pcfd_set_timestep(pcfd_ctx* ctx, const double* dt) {
  return guarded(ctx, [&](Ctx& c) -> int {
    if (!dt) {
      c.err = "pcfd_set_timestep: null";
      return PCFD_ERR_ARG;
    }
    return upload_soa(c, c.d.dt, 1, dt);
  });
}

int pcfd_jacobian(pcfd_ctx* ctx) {
  return guarded(ctx, [&](Ctx& c) -> int {
    if (int r = need_params(c)) return r;
    ensure_matrix(c);
    launch_jacobian(c);
    PCFD_HIP(hipStreamSynchronize(c.stream));
    c.jac_valid = true;
    c.lu_done = false;
    return PCFD_OK;
  });
}

int pcfd_download_jacobian(pcfd_ctx* ctx, const int* ia, const int* ja, double* M) {
  return guarded(ctx, [&](Ctx& c) -> int {
    if (!ia || !ja || !M) {
      c.err = "pcfd_download_jacobian: null argument";
      return PCFD_ERR_ARG;
    }
    if (!c.jac_valid || c.lu_done) {
      c.err = "pcfd_download_jacobian: no unfactored Jacobian on the device";
      return PCFD_ERR_STATE;
    }
    std::vector<double> h((size_t)c.nblocks * 25);
    if (c.prm.precision == PCFD_CHOSEN_FP32) {
      std::vector<float> f(h.size());
      PCFD_HIP(hipMemcpy(f.data(), c.d.Mf, f.size() * sizeof(float), hipMemcpyDeviceToHost));
      for (size_t k = 0; k < f.size(); ++k) h[k] = f[k];
    } else {
      PCFD_HIP(hipMemcpy(h.data(), c.d.Md, h.size() * sizeof(double), hipMemcpyDeviceToHost));
    }
    std::vector<int> row_ptr(c.nnode + 1), col(c.nnz);
    PCFD_HIP(hipMemcpy(row_ptr.data(), c.d.row_ptr, row_ptr.size() * sizeof(int), hipMemcpyDeviceToHost));
    PCFD_HIP(hipMemcpy(col.data(), c.d.col, col.size() * sizeof(int), hipMemcpyDeviceToHost));
    for (int i = 0; i < c.nnode; ++i) {
      for (int k = ia[i]; k < ia[i + 1]; ++k) {
        int j = ja[k];
        long long src = -1;
        if (j == i) {
          src = row_ptr[i] + i;
        } else {
          const int* b = col.data() + row_ptr[i];
          const int* e = col.data() + row_ptr[i + 1];
          const int* f = std::lower_bound(b, e, j);
          if (f != e && *f == j) src = (long long)row_ptr[i] + i + 1 + (f - b);
        }
        if (src < 0) {
          c.err = "pcfd_download_jacobian: block (" + std::to_string(i) + "," + std::to_string(j) +
                  ") is not in the mesh graph";
          return PCFD_ERR_ARG;
        }
        std::memcpy(M + (size_t)k * 25, h.data() + (size_t)src * 25, 25 * sizeof(double));
      }
    }
    return PCFD_OK;
  });
}

int pcfd_prepare_sgs(pcfd_ctx* ctx) {
  return guarded(ctx, [&](Ctx& c) -> int {
    if (!c.jac_valid) {
      c.err = "pcfd_prepare_sgs: no Jacobian";
      return PCFD_ERR_STATE;
    }
    if (!c.lu_done) launch_prepare_sgs(c);
    c.lu_done = true;
    PCFD_HIP(hipStreamSynchronize(c.stream));
    return PCFD_OK;
  });
}

int pcfd_sgs(pcfd_ctx* ctx, int nsgs, double* x_out, double* delta) {
  return guarded(ctx, [&](Ctx& c) -> int {
    if (!c.lu_done) {
      c.err = "pcfd_sgs: call pcfd_prepare_sgs first";
      return PCFD_ERR_STATE;
    }
    if (nsgs < 0) {
      c.err = "pcfd_sgs: nsgs < 0";
      return PCFD_ERR_ARG;
    }
    launch_sgs(c, nsgs, delta);
    if (x_out) {
      if (c.prm.precision == PCFD_CHOSEN_FP32) {
        std::vector<float> f(5 * (size_t)c.nnode);
        PCFD_HIP(hipMemcpyAsync(f.data(), c.d.xf, f.size() * sizeof(float), hipMemcpyDeviceToHost, c.stream));
        PCFD_HIP(hipStreamSynchronize(c.stream));
        for (int i = 0; i < c.nnode; ++i)
          for (int k = 0; k < 5; ++k) x_out[(size_t)i * 5 + k] = f[(size_t)k * c.nnode + i];
        return PCFD_OK;
      }
      return download_soa(c, c.d.xd, 5, x_out);
    }
    PCFD_HIP(hipStreamSynchronize(c.stream));
    return PCFD_OK;
  });
}

int pcfd_sgs_order(pcfd_ctx* ctx, int* order) {
  return guarded(ctx, [&](Ctx& c) -> int {
    if (!order) return PCFD_ERR_ARG;
    PCFD_HIP(hipMemcpy(order, c.d.color_nodes, sizeof(int) * c.nnode, hipMemcpyDeviceToHost));
    return PCFD_OK;
  });
}

int pcfd_apply_dq(pcfd_ctx* ctx, int* nzeroed) {
  return guarded(ctx, [&](Ctx& c) -> int {
    if (int r = need_params(c)) return r;
    launch_apply_dq(c, 0, nzeroed);
    PCFD_HIP(hipStreamSynchronize(c.stream));
    return PCFD_OK;
  });
}

int pcfd_explicit_update(pcfd_ctx* ctx, int* nzeroed) {
  return guarded(ctx, [&](Ctx& c) -> int {
    if (int r = need_params(c)) return r;
    launch_apply_dq(c, 1, nzeroed);
    PCFD_HIP(hipStreamSynchronize(c.stream));
    return PCFD_OK;
  });
}

int pcfd_phase_times(pcfd_ctx* ctx, double ms[9], int reset) {
  if (!ctx) return PCFD_ERR_ARG;
  if (ms)
    for (int k = 0; k < 9; ++k) ms[k] = ctx->c.phase_ms[k];
  if (reset)
    for (int k = 0; k < 9; ++k) ctx->c.phase_ms[k] = 0.0;
  return PCFD_OK;
}

int pcfd_enable_phase_timing(pcfd_ctx* ctx, int on) {
  if (!ctx) return PCFD_ERR_ARG;
  ctx->c.timing = on != 0;
  return PCFD_OK;
}

int pcfd_synchronize(pcfd_ctx* ctx) {
  return guarded(ctx, [&](Ctx& c) -> int {
    PCFD_HIP(hipStreamSynchronize(c.stream));
    return PCFD_OK;
  });
}

}  // extern "C"
