// C ABI of libpcfd.so: context lifetime, layouts, transfers and single phases.
#include <algorithm>
#include <cmath>
#include <cstdlib>
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
  dfree(d.xyz); dfree(d.vol); dfree(d.vol23); dfree(d.sw); dfree(d.gwidth); dfree(d.gbase); dfree(d.gcol); dfree(d.gnrm); dfree(d.qrk);
  dfree(d.he_ptr); dfree(d.he_id); dfree(d.he_tag); dfree(d.he_a); dfree(d.bnd_nodes); dfree(d.bidx);
  dfree(d.jdb); dfree(d.jqb); dfree(d.bc_type); dfree(d.twall); dfree(d.he_nrm); dfree(d.spos);
  dfree(d.qref); dfree(d.q); dfree(d.grad); dfree(d.lim); dfree(d.lim2); dfree(d.edges); dfree(d.clip_flag); dfree(d.clip_cnt);
  dfree(d.clip_off); dfree(d.clip_list); dfree(d.b); dfree(d.dt);
  dfree(d.srow); dfree(d.swidth); dfree(d.sbase); dfree(d.scol); dfree(d.Md); dfree(d.Mf); dfree(d.Dd);
  dfree(d.Df); dfree(d.piv); dfree(d.bsd); dfree(d.bsf); dfree(d.xd); dfree(d.xf); dfree(d.red);
  dfree(d.red_out); dfree(d.icount); dfree(d.dt_next); dfree(d.qjac);
  for (hipEvent_t* e : {&c.ev0, &c.ev1, &c.ev2, &c.ev3, &c.ev4, &c.ev5, &c.ev_fork, &c.ev_join}) {
    if (*e) (void)hipEventDestroy(*e);
    *e = nullptr;
  }
  if (c.stream2) (void)hipStreamDestroy(c.stream2);
  if (c.stream) (void)hipStreamDestroy(c.stream);
  c.stream = c.stream2 = nullptr;
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

template <typename T>
void fetch(std::vector<double>& h, const T* src, size_t n) {
  std::vector<T> t(n);
  PCFD_HIP(hipMemcpy(t.data(), src, n * sizeof(T), hipMemcpyDeviceToHost));
  h.resize(n);
  for (size_t k = 0; k < n; ++k) h[k] = (double)t[k];
}

int need_params(Ctx& c) {
  if (!c.have_params) {
    c.err = "pcfd_set_params has not been called";
    return PCFD_ERR_STATE;
  }
  return PCFD_OK;
}

void ensure_matrix(Ctx& c) {
  const size_t off = (size_t)c.nslot_s * kW * 25, dg = (size_t)c.nchunk_s * kW * 25;
  if (c.prm.precision == PCFD_CHOSEN_FP32) {
    if (!c.d.Mf) dalloc(c, &c.d.Mf, off);
    if (!c.d.Df) dalloc(c, &c.d.Df, dg);
    if (!c.d.bsf) dalloc(c, &c.d.bsf, 5 * (size_t)c.nrow_pad);
    if (!c.d.xf) dalloc(c, &c.d.xf, 5 * (size_t)c.nnode);
  } else {
    if (!c.d.Md) dalloc(c, &c.d.Md, off);
    if (!c.d.Dd) dalloc(c, &c.d.Dd, dg);
    if (!c.d.bsd) dalloc(c, &c.d.bsd, 5 * (size_t)c.nrow_pad);
    if (!c.d.xd) dalloc(c, &c.d.xd, 5 * (size_t)c.nnode);
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
void sub_begin(Ctx& c) {
  if (c.timing) PCFD_HIP(hipEventRecord(c.ev2, c.stream));
}
void sub_end(Ctx& c, int k) {
  if (!c.timing) return;
  PCFD_HIP(hipEventRecord(c.ev3, c.stream));
  PCFD_HIP(hipEventSynchronize(c.ev3));
  float ms = 0.f;
  PCFD_HIP(hipEventElapsedTime(&ms, c.ev2, c.ev3));
  c.phase_ms[k] += ms;
}
void ensure_matrix_alloc(Ctx& c) { ensure_matrix(c); }
void ensure_overlap_buffers(Ctx& c) {
  if (!c.d.dt_next) dalloc(c, &c.d.dt_next, (size_t)c.nnode);
  if (c.has_hardset && !c.d.qjac) dalloc(c, &c.d.qjac, 5 * (size_t)c.ntot);
}
}  // namespace pcfd

extern "C" {

const char* pcfd_version(void) { return "pcfd-gfx950 0.2"; }

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
    PCFD_HIP(hipEventCreate(&c.ev2));
    PCFD_HIP(hipEventCreate(&c.ev3));
    PCFD_HIP(hipEventCreate(&c.ev4));
    PCFD_HIP(hipEventCreate(&c.ev5));
    // second stream for the Jacobian rows (default priority, like the first)
    PCFD_HIP(hipStreamCreateWithFlags(&c.stream2, hipStreamNonBlocking));
    PCFD_HIP(hipEventCreateWithFlags(&c.ev_fork, hipEventDisableTiming));
    PCFD_HIP(hipEventCreateWithFlags(&c.ev_join, hipEventDisableTiming));
    if (const char* e = std::getenv("PCFD_JACOBIAN_STREAM")) c.overlap = std::strcmp(e, "0") != 0;
    c.nnode = nn;
    c.nbnode = nb;
    c.nedge = ne;
    c.ntot = nn + nb;
    c.ntag = nt;
    c.nnz = 2 * ne;
    // ---- node adjacency in edge order, with the edge id: a node's slots
    //      hold its neighbours in the order the reference's edge loop adds
    //      their fluxes to it (Driver, ucs/driver.tcc:5-67)
    std::vector<int> row_ptr(nn + 1, 0);
    for (int e = 0; e < ne; ++e) {
      row_ptr[m->edges[2 * e] + 1]++;
      row_ptr[m->edges[2 * e + 1] + 1]++;
    }
    for (int i = 0; i < nn; ++i) row_ptr[i + 1] += row_ptr[i];
    std::vector<std::pair<int, int>> ent(c.nnz);
    {
      std::vector<int> fill(row_ptr.begin(), row_ptr.end() - 1);
      for (int e = 0; e < ne; ++e) {
        int a = m->edges[2 * e], b = m->edges[2 * e + 1];
        ent[fill[a]++] = {b, e};
        ent[fill[b]++] = {a, e};
      }
    }
    {
      std::vector<int> stamp(nn, -1);
      for (int i = 0; i < nn; ++i)
        for (int p = row_ptr[i]; p < row_ptr[i + 1]; ++p) {
          if (stamp[ent[p].first] == i) throw std::runtime_error("pcfd_create: duplicate edge between nodes");
          stamp[ent[p].first] = i;
        }
    }
    auto deg = [&](int i) { return row_ptr[i + 1] - row_ptr[i]; };
    // ---- gELL: chunks of 64 consecutive nodes
    c.nchunk_g = (nn + kW - 1) / kW;
    std::vector<int> gwidth(c.nchunk_g, 0), gbase(c.nchunk_g, 0);
    long long slots = 0;
    for (int k = 0; k < c.nchunk_g; ++k) {
      int w = 0;
      for (int i = k * kW; i < std::min(nn, (k + 1) * kW); ++i) w = std::max(w, deg(i));
      gwidth[k] = w;
      gbase[k] = (int)slots;
      slots += w;
    }
    if (slots * kW > (1LL << 31)) throw std::runtime_error("pcfd_create: mesh too large for 32-bit slot offsets");
    c.nslot_g = slots;
    std::vector<int> gcol((size_t)slots * kW, -1);
    std::vector<double> gnrm(4 * (size_t)slots * kW, 0.0);
    const size_t gstride = (size_t)slots * kW;
    for (int i = 0; i < nn; ++i) {
      for (int p = row_ptr[i]; p < row_ptr[i + 1]; ++p) {
        size_t idx = ((size_t)gbase[i / kW] + (p - row_ptr[i])) * kW + (i % kW);
        gcol[idx] = ent[p].first;
        int e = ent[p].second;
        for (int k = 0; k < 4; ++k) gnrm[k * gstride + idx] = m->edge_a[4 * (size_t)e + k];
      }
    }
    c.h_row_ptr = row_ptr;
    c.h_col.resize(c.nnz);
    for (int p = 0; p < c.nnz; ++p) c.h_col[p] = ent[p].first;
    ent.clear();
    ent.shrink_to_fit();
    // ---- half-edges per node in half-edge order
    std::vector<int> he_ptr(nn + 1, 0);
    for (int h = 0; h < nb; ++h) he_ptr[m->bedges[2 * h] + 1]++;
    for (int i = 0; i < nn; ++i) he_ptr[i + 1] += he_ptr[i];
    std::vector<int> he_id(nb);
    {
      std::vector<int> hfill(he_ptr.begin(), he_ptr.end() - 1);
      for (int h = 0; h < nb; ++h) he_id[hfill[m->bedges[2 * h]]++] = h;
    }
    std::vector<int> bnd, bidx(nn, -1);
    for (int i = 0; i < nn; ++i)
      if (he_ptr[i + 1] > he_ptr[i]) {
        bidx[i] = (int)bnd.size();
        bnd.push_back(i);
      }
    c.nbnd = (int)bnd.size();
    std::vector<int> he_tag(m->bfactag, m->bfactag + nb);
    std::vector<double> he_a(4 * (size_t)nb);
    for (int h = 0; h < nb; ++h)
      for (int k = 0; k < 4; ++k) he_a[(size_t)k * nb + h] = m->bedge_a[4 * (size_t)h + k];
    // ---- SoA geometry and BC tables
    std::vector<double> xyz(3 * (size_t)nn);
    for (int i = 0; i < nn; ++i)
      for (int k = 0; k < 3; ++k) xyz[(size_t)k * nn + i] = m->xyz[3 * (size_t)i + k];
    std::vector<double> vol(m->vol, m->vol + nn);
    c.h_bc_type.assign(m->bc_type, m->bc_type + nt);
    // hard-set BC types overwrite boundary node states inside the Jacobian
    for (int h = 0; h < nb; ++h) {
      const int ty = c.h_bc_type[m->bfactag[h]];
      if (ty == 1 /* dirichlet */ || ty == 7 /* sonicInflow */ || ty == 4 /* noSlip */) c.has_hardset = true;
    }
    std::vector<double> qref(6 * (size_t)nt, 0.0);
    for (int t = 0; t < nt; ++t)
      for (int k = 0; k < 5; ++k) qref[(size_t)k * nt + t] = m->qref[5 * (size_t)t + k];
    std::vector<double> twall(nt, -1.0);
    if (m->twall) twall.assign(m->twall, m->twall + nt);
    // wall-normal neighbour of each noSlip half-edge's node: over the
    // neighbours in ascending id, the last one maximising -unit(x_pt - x_wall).a
    // from a start of 0 (ucs/bc.tcc:1182-1210)
    std::vector<int> he_nrm(nb, -1);
    for (int h = 0; h < nb; ++h) {
      if (c.h_bc_type[m->bfactag[h]] != 4 /* noSlip */) continue;
      c.has_noslip = true;
      const int L = m->bedges[2 * h];
      const double* wx = &m->xyz[3 * (size_t)L];
      const double* a = &m->bedge_a[4 * (size_t)h];
      double dotmax = 0.0;
      for (int p = row_ptr[L]; p < row_ptr[L + 1]; ++p) {
        const int pt = c.h_col[p];
        double dx[3];
        for (int k = 0; k < 3; ++k) dx[k] = m->xyz[3 * (size_t)pt + k] - wx[k];
        const double mag = std::sqrt(dx[0] * dx[0] + dx[1] * dx[1] + dx[2] * dx[2]);
        for (int k = 0; k < 3; ++k) dx[k] = dx[k] / mag;
        double dot = 0.0;
        for (int k = 0; k < 3; ++k) dot -= dx[k] * a[k];
        if (dot >= dotmax) {
          he_nrm[h] = pt;
          dotmax = dot;
        }
      }
    }
    c.nblocks = (long long)nn + c.nnz;
    // ---- colouring and the SELL row order
    std::vector<int> colour(nn);
    int ncol = pcfd_color_nodes(nn, ne, m->edges, colour.data());
    if (ncol < 0) throw std::runtime_error(pcfd_mesh_error());
    c.ncolor = ncol;
    std::vector<std::vector<int>> members(ncol);
    for (int i = 0; i < nn; ++i) members[colour[i]].push_back(i);
    std::vector<int> srow;
    c.color_chunk.assign(ncol + 1, 0);
    for (int k = 0; k < ncol; ++k) {
      auto& v = members[k];
      std::stable_sort(v.begin(), v.end(), [&](int a, int b) { return deg(a) > deg(b); });
      srow.insert(srow.end(), v.begin(), v.end());
      while (srow.size() % kW) srow.push_back(-1);
      c.color_chunk[k + 1] = (int)(srow.size() / kW);
    }
    c.nrow_pad = (int)srow.size();
    c.nchunk_s = c.nrow_pad / kW;
    std::vector<int> swidth(c.nchunk_s, 0), sbase(c.nchunk_s, 0), spos(nn, -1);
    slots = 0;
    for (int k = 0; k < c.nchunk_s; ++k) {
      int w = 0;
      for (int l = 0; l < kW; ++l) {
        int i = srow[(size_t)k * kW + l];
        if (i >= 0) {
          w = std::max(w, deg(i));
          spos[i] = k * kW + l;
        }
      }
      swidth[k] = w;
      sbase[k] = (int)slots;
      slots += w;
    }
    if (slots * kW * 25 > (1LL << 40)) throw std::runtime_error("pcfd_create: matrix too large");
    c.nslot_s = slots;
    std::vector<int> scol((size_t)slots * kW, -1);
    for (int k = 0; k < c.nchunk_s; ++k)
      for (int l = 0; l < kW; ++l) {
        int i = srow[(size_t)k * kW + l];
        if (i < 0) continue;
        for (int p = row_ptr[i]; p < row_ptr[i + 1]; ++p)
          scol[((size_t)sbase[k] + (p - row_ptr[i])) * kW + l] = c.h_col[p];
      }
    c.h_srow = srow;
    c.h_spos = spos;
    c.h_sbase = sbase;

    Dev& d = c.d;
    dput(c, &d.xyz, xyz);
    dput(c, &d.vol, vol);
    {
      // host pow, so the limit has the same bits as a host evaluation
      std::vector<double> v23(vol.size());
      for (size_t i = 0; i < vol.size(); ++i) v23[i] = std::pow(vol[i], 2.0 / 3.0);
      dput(c, &d.vol23, v23);
    }
    dalloc(c, &d.sw, 6 * (size_t)nn);
    dput(c, &d.gwidth, gwidth);
    dput(c, &d.gbase, gbase);
    dput(c, &d.gcol, gcol);
    dput(c, &d.gnrm, gnrm);
    dput(c, &d.he_ptr, he_ptr);
    dput(c, &d.he_id, he_id);
    dput(c, &d.he_tag, he_tag);
    dput(c, &d.he_a, he_a);
    dput(c, &d.bnd_nodes, bnd);
    dput(c, &d.bidx, bidx);
    dalloc(c, &d.jdb, 25 * (size_t)std::max(c.nbnd, 1));
    dalloc(c, &d.jqb, 5 * (size_t)std::max(c.nbnd, 1));
    dput(c, &d.bc_type, c.h_bc_type);
    dput(c, &d.qref, qref);
    dput(c, &d.twall, twall);
    dput(c, &d.he_nrm, he_nrm);
    dput(c, &d.spos, spos);
    dalloc(c, &d.q, 5 * (size_t)c.ntot);
    dalloc(c, &d.grad, 27 * (size_t)nn);
    dalloc(c, &d.lim, 5 * (size_t)nn);
    dalloc(c, &d.lim2, 5 * (size_t)nn);
    {
      std::vector<int> ev(m->edges, m->edges + 2 * (size_t)ne);
      dput(c, &d.edges, ev);
      const size_t npass = ((size_t)ne + 63) / 64;
      dalloc(c, &d.clip_flag, (size_t)nn);
      dalloc(c, &d.clip_cnt, std::max<size_t>(npass, 1));
      dalloc(c, &d.clip_off, npass + 1);
      dalloc(c, &d.clip_list, std::max<size_t>((size_t)ne, 1));
    }
    dalloc(c, &d.b, 5 * (size_t)nn);
    dalloc(c, &d.dt, (size_t)nn);
    dput(c, &d.srow, srow);
    dput(c, &d.swidth, swidth);
    dput(c, &d.sbase, sbase);
    dput(c, &d.scol, scol);
    dalloc(c, &d.piv, (size_t)c.nrow_pad);
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
  if (ctx->c.stream2) (void)hipStreamSynchronize(ctx->c.stream2);
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
        (p->gradient_terms != 5 && p->gradient_terms != 9) ||
        (p->viscous && (p->gradient_terms != 9 || !(p->re_tilde > 0.0) || !(p->prandtl > 0.0) ||
                        !(p->prandtl_t > 0.0) || !(p->sutherland_s >= 0.0)))) {
      c.err = "pcfd_set_params: parameter out of range (viscous runs need gradient_terms 9, Re/M, Pr, PrT > 0)";
      return PCFD_ERR_ARG;
    }
    if (p->flux_type == 2 || p->jac_field_type == 2 || p->jac_boundary_type == 2) {
      c.err = "pcfd_set_params: the HLLC flux (fluxType 2) and the complex-step Jacobians (jacobianFieldType 2, "
              "jacobianBoundaryType 2) are not built";
      return PCFD_ERR_UNSUPPORTED;
    }
    if ((p->gradient_type != 0 && p->gradient_type != 1) || (p->global_dt != 0 && p->global_dt != 1) ||
        (p->enable_vnn != 0 && p->enable_vnn != 1) || (p->enable_vnn && !(p->vnn > 0.0)) ||
        (p->dynamic_cfl != 0 && p->dynamic_cfl != 1) || (p->jac_field_type != 0 && p->jac_field_type != 1) ||
        (p->jac_boundary_type != 0 && p->jac_boundary_type != 1) ||
        (p->flux_type != PCFD_FLUX_ROE && p->flux_type != PCFD_FLUX_ALGEBRAIC) ||
        (p->rk_stages != 0 && p->rk_stages != 1 && p->rk_stages != 4)) {
      c.err = "pcfd_set_params: option out of range (gradient_type, jac_field_type, jac_boundary_type, global_dt, "
              "enable_vnn, dynamic_cfl 0/1; flux_type PCFD_FLUX_ROE/PCFD_FLUX_ALGEBRAIC; vnn > 0 when enabled; "
              "rk_stages 0, 1 or 4)";
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
    std::vector<double> off, dg;
    const size_t noff = (size_t)c.nslot_s * kW * 25, ndg = (size_t)c.nchunk_s * kW * 25;
    if (c.prm.precision == PCFD_CHOSEN_FP32) {
      fetch(off, c.d.Mf, noff);
      fetch(dg, c.d.Df, ndg);
    } else {
      fetch(off, c.d.Md, noff);
      fetch(dg, c.d.Dd, ndg);
    }
    for (int i = 0; i < c.nnode; ++i) {
      const int r = c.h_spos[i], k = r / kW, l = r % kW;
      for (int b = ia[i]; b < ia[i + 1]; ++b) {
        const int j = ja[b];
        double* out = M + (size_t)b * 25;
        if (j == i) {
          for (int e = 0; e < 25; ++e) out[e] = dg[((size_t)k * 25 + e) * kW + l];
          continue;
        }
        const int* cb = c.h_col.data() + c.h_row_ptr[i];
        const int* ce = c.h_col.data() + c.h_row_ptr[i + 1];
        const int* f = std::find(cb, ce, j);  // a row's slots follow the edge order
        if (f == ce) {
          c.err = "pcfd_download_jacobian: block (" + std::to_string(i) + "," + std::to_string(j) +
                  ") is not in the mesh graph";
          return PCFD_ERR_ARG;
        }
        const size_t s = (size_t)c.h_sbase[k] + (size_t)(f - cb);
        for (int e = 0; e < 25; ++e) out[e] = off[(s * 25 + e) * kW + l];
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
      std::vector<double> h;
      if (c.prm.precision == PCFD_CHOSEN_FP32) fetch(h, c.d.xf, 5 * (size_t)c.nnode);
      else fetch(h, c.d.xd, 5 * (size_t)c.nnode);
      for (size_t k = 0; k < 5 * (size_t)c.nnode; ++k) x_out[k] = h[k];
      return PCFD_OK;
    }
    PCFD_HIP(hipStreamSynchronize(c.stream));
    return PCFD_OK;
  });
}

int pcfd_sgs_order(pcfd_ctx* ctx, int* order) {
  return guarded(ctx, [&](Ctx& c) -> int {
    if (!order) return PCFD_ERR_ARG;
    int k = 0;
    for (int r : c.h_srow)
      if (r >= 0) order[k++] = r;
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
    for (int k = 0; k < PCFD_NPHASE; ++k) ms[k] = ctx->c.phase_ms[k];
  if (reset)
    for (int k = 0; k < PCFD_NPHASE; ++k) ctx->c.phase_ms[k] = 0.0;
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
