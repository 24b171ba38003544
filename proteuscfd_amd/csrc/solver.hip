// Host driver of the steady implicit (or explicit) iteration, in the
// reference's phase order:
//   PreTimeAdvance   (ucs/solutionSpace.tcc:572-636): Jacobian with the
//                    previous dt, UpdateCFL, ComputeTimesteps
//   NewtonIterate    (ucs/solutionSpace.tcc:639-904): UpdateBCs, gradient,
//                    limiter, residual + norms, PrepareSGS + SGS (or the
//                    explicit update), NaN rows, ApplyDQ
//   PostTimeAdvance  (ucs/solutionSpace.tcc:907-916): iter++
#include <chrono>
#include <cmath>
#include <stdexcept>
#include <utility>

#include "common.h"

namespace pcfd {
void ensure_matrix_alloc(Ctx& c);
void ensure_overlap_buffers(Ctx& c);
}  // namespace pcfd

using namespace pcfd;

struct pcfd_ctx {
  Ctx c;
};

namespace {

// Param::UpdateCFL, prescribed ramp (ucs/param.tcc:718-725)
double cfl_at(const pcfd_params& p, int iter) {
  if (p.cfl_ramp > 0 && iter < p.cfl_ramp) return (p.cfl - p.cfl_start) / (double)p.cfl_ramp * (double)iter + p.cfl_start;
  return p.cfl;
}

// Param::UpdateCFL (ucs/param.tcc:695-726): with dynamicCFL the CFL moves by
// 10% per iteration in the direction of the residual trend that
// PreTimeAdvance computes from the last two residuals
// (ucs/solutionSpace.tcc:629-633), capped at cfl; cflStart on iteration 1.
double cfl_update(const pcfd_params& p, int iter, double res, double res_nm1, double cfl_prev) {
  if (!p.dynamic_cfl) return cfl_at(p, iter);
  double delta = 9999.0;
  if (res_nm1 > 1.0e-21) delta = (res_nm1 - res) / res_nm1;
  if (iter > 1) {
    const double sgn = delta > 0.0 ? 1.0 : -1.0;
    const double A = 0.0 > cfl_prev + 0.1 * sgn * cfl_prev ? 0.0 : cfl_prev + 0.1 * sgn * cfl_prev;
    return A < p.cfl ? A : p.cfl;
  }
  return p.cfl_start;
}

void iterate(Ctx& c, double* resid) {
  const pcfd_params& p = c.prm;
  // PreTimeAdvance
  const bool do_jac = p.nsgs > 0 && (c.iter % p.jac_freq == 0 || c.iter == 1);
  // with phase timing on, every phase runs alone so that its time is its own
  const bool overlap = do_jac && c.overlap && !c.timing;
  if (do_jac) {
    ensure_matrix_alloc(c);
    phase_begin(c);
    if (overlap) {
      ensure_overlap_buffers(c);
      launch_jacobian_overlapped(c);  // rows read c.d.dt on the second stream...
      std::swap(c.d.dt, c.d.dt_next);  // ...while the new time step goes to the other buffer
    } else {
      launch_jacobian(c);
    }
    phase_end(c, 5);
    c.jac_valid = true;
    c.lu_done = false;
  }
  c.cfl = cfl_update(p, c.iter, c.res, c.res_nm1, c.cfl);
  phase_begin(c);
  launch_timestep(c, c.cfl);
  phase_end(c, 4);
  // NewtonIterate
  phase_begin(c);
  launch_update_bcs(c);
  phase_end(c, 0);
  const bool ho = p.sorder > 1 && c.iter > p.nfirst;
  const bool grad_now = ho || p.viscous;  // ucs/solutionSpace.tcc:671-677
  const bool lim_now = (c.iter % p.limiter_refresh == 0) && p.sorder > 1 && p.limiter && c.iter > p.nfirst &&
                       c.iter < p.limiter_freeze;
  if (grad_now && lim_now && p.gradient_type == 0 && !c.timing) {
    // one pass for the weighted-LSQ gradient and the limiter (phase timing
    // keeps them apart so that each has its own time)
    launch_gradient_limiter(c, p.gradient_terms);
  } else {
    if (grad_now) {
      phase_begin(c);
      launch_gradient(c, p.gradient_terms);
      phase_end(c, 1);
    }
    if (lim_now) {
      phase_begin(c);
      launch_limiter(c);
      phase_end(c, 2);
    }
  }
  phase_begin(c);
  launch_residual(c, ho ? 1 : 0);
  phase_end(c, 3);
  launch_norms(c, resid);
  c.res_nm1 = c.res;  // ucs/solutionSpace.tcc:705-706
  c.res = resid[0];
  if (p.nsgs > 0) {
    jacobian_join(c);
    if (!c.lu_done) {
      phase_begin(c);
      launch_prepare_sgs(c);
      phase_end(c, 6);
      c.lu_done = true;
    }
    phase_begin(c);
    launch_sgs(c, p.nsgs, nullptr);
    phase_end(c, 7);
    phase_begin(c);
    launch_apply_dq(c, 0, nullptr);
    phase_end(c, 8);
  } else if (p.rk_stages == 4) {
    // four-stage Runge-Kutta (Jameson's coefficients): the first stage uses
    // the residual above; the state it starts from is kept in qrk
    phase_begin(c);
    if (!c.d.qrk) PCFD_HIP(hipMalloc((void**)&c.d.qrk, sizeof(double) * 5 * (size_t)c.ntot));
    PCFD_HIP(hipMemcpyAsync(c.d.qrk, c.d.q, sizeof(double) * 5 * (size_t)c.ntot, hipMemcpyDeviceToDevice,
                            c.stream));
    launch_zero_count(c);
    static const double a[4] = {0.25, 1.0 / 3.0, 0.5, 1.0};
    for (int k = 0; k < 4; ++k) {
      if (k > 0) {
        launch_update_bcs(c);
        if (grad_now && lim_now && p.gradient_type == 0) {
          launch_gradient_limiter(c, p.gradient_terms);
        } else {
          if (grad_now) launch_gradient(c, p.gradient_terms);
          if (lim_now) launch_limiter(c);
        }
        launch_residual(c, ho ? 1 : 0);
      }
      launch_rk_stage(c, a[k], c.d.qrk);
    }
    phase_end(c, 8);
  } else {
    phase_begin(c);
    launch_apply_dq(c, 1, nullptr);
    phase_end(c, 8);
  }
  c.iter++;
}

}  // namespace

extern "C" {

int pcfd_reset_iteration(pcfd_ctx* ctx) {
  if (!ctx) return PCFD_ERR_ARG;
  Ctx& c = ctx->c;
  if (!c.have_params) {
    c.err = "pcfd_set_params has not been called";
    return PCFD_ERR_STATE;
  }
  try {
    jacobian_join(c);
    // the counter starts at 0 (ucs/solutionSpaceBase.h:45): the first
    // iteration uses cflStart and, as iter > firstOrderSteps fails, first order
    c.iter = 0;
    c.jac_valid = false;
    c.lu_done = false;
    // Limiter ctor fills 1.0 (ucs/limiters.tcc:39-43); PreIterate computes dt
    std::vector<double> ones(5 * (size_t)c.nnode, 1.0);
    PCFD_HIP(hipMemcpy(c.d.lim, ones.data(), ones.size() * sizeof(double), hipMemcpyHostToDevice));
    c.res = c.res_nm1 = 0.0;
    c.cfl = cfl_update(c.prm, c.iter, 0.0, 0.0, c.prm.cfl_start);
    launch_timestep(c, c.cfl);
    PCFD_HIP(hipStreamSynchronize(c.stream));
  } catch (const std::exception& e) {
    c.err = e.what();
    return PCFD_ERR_HIP;
  }
  return PCFD_OK;
}

int pcfd_get_iteration(const pcfd_ctx* ctx) { return ctx ? ctx->c.iter : -1; }

int pcfd_iterate(pcfd_ctx* ctx, double resid[6]) {
  if (!ctx) return PCFD_ERR_ARG;
  Ctx& c = ctx->c;
  if (!c.have_params) {
    c.err = "pcfd_set_params has not been called";
    return PCFD_ERR_STATE;
  }
  try {
    double r[6];
    iterate(c, r);
    if (resid)
      for (int k = 0; k < 6; ++k) resid[k] = r[k];
  } catch (const std::exception& e) {
    c.err = e.what();
    return PCFD_ERR_HIP;
  }
  return PCFD_OK;
}

int pcfd_solve(pcfd_ctx* ctx, int max_iter, double drop, double* history, int* niter, double* secs) {
  if (!ctx) return PCFD_ERR_ARG;
  Ctx& c = ctx->c;
  if (!c.have_params) {
    c.err = "pcfd_set_params has not been called";
    return PCFD_ERR_STATE;
  }
  if (max_iter < 1 || !(drop > 0.0)) {
    c.err = "pcfd_solve: max_iter >= 1 and drop > 0 required";
    return PCFD_ERR_ARG;
  }
  try {
    auto t0 = std::chrono::steady_clock::now();
    double first = -1.0;
    int n = 0;
    for (; n < max_iter; ++n) {
      double r[6];
      iterate(c, r);
      if (history) history[n] = r[0];
      if (!std::isfinite(r[0])) {
        c.err = "pcfd_solve: residual diverged";
        if (niter) *niter = n + 1;
        return PCFD_ERR_STATE;
      }
      if (n == 0) first = r[0];
      if (r[0] <= drop * first) {
        ++n;
        break;
      }
    }
    PCFD_HIP(hipStreamSynchronize(c.stream));
    auto t1 = std::chrono::steady_clock::now();
    if (niter) *niter = n;
    if (secs) *secs = std::chrono::duration<double>(t1 - t0).count();
  } catch (const std::exception& e) {
    c.err = e.what();
    return PCFD_ERR_HIP;
  }
  return PCFD_OK;
}

}  // extern "C"
