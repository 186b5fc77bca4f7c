// Ground-state solver loop (becgpp/solvers.py:48-226) driving the device
// kernels.  Fields stay in HBM; each iteration synchronises once for the
// inner products of the CG step (batched with the previous step's residual)
// and once per Armijo trial.  Scalar logic is the reference's, in the same
// order, on the host.
#include <math.h>
#include <stdio.h>
#include <string.h>
#include <algorithm>
#include <chrono>

#include "../../include/becgpp_hip.h"
#include "bec_ops.h"

enum { S_P0 = 0, S_P1, S_G0, S_G1, S_D, S_Z, S_BEST, S_COUNT };

void solver_release(Plan& p) {
  if (!p.st_alloc) return;
  for (int i = 0; i < 8; ++i)
    if (p.st[i]) (void)hipFree(p.st[i]);
  for (int i = 0; i < 2; ++i)
    if (p.stphi[i]) (void)hipFree(p.stphi[i]);
  memset(p.st, 0, sizeof(p.st));
  memset(p.stphi, 0, sizeof(p.stphi));
  p.st_alloc = false;
}

static void solver_alloc(Plan& p) {
  if (p.st_alloc) return;
  for (int i = 0; i < S_COUNT; ++i) HIPCHK(hipMalloc(&p.st[i], sizeof(double2) * p.n));
  for (int i = 0; i < 2; ++i) HIPCHK(hipMalloc(&p.stphi[i], sizeof(double) * p.n));
  p.st_alloc = true;
}

static double now_s() {
  using namespace std::chrono;
  return duration<double>(steady_clock::now().time_since_epoch()).count();
}

struct C2 {
  double re, im;
};
static inline C2 cm(double a, double b) { return C2{a, b}; }
static inline C2 operator+(C2 a, C2 b) { return cm(a.re + b.re, a.im + b.im); }
static inline C2 operator-(C2 a, C2 b) { return cm(a.re - b.re, a.im - b.im); }
static inline C2 operator*(double s, C2 a) { return cm(s * a.re, s * a.im); }
static inline C2 operator*(C2 a, C2 b) { return cm(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re); }
static inline C2 conj(C2 a) { return cm(a.re, -a.im); }

extern "C" int bec_ground_state(bec_plan* h, const bec_phys* php, const bec_solver_cfg* cf,
                                void* psi_io, bec_solve_out* out, double* trace, int trace_cap) {
  try {
    if (!h || !php || !cf || !psi_io || !out) throw BecError(1, "null argument");
    Plan& p = h->p;
    const Phys ph = {php->beta2, php->beta3, php->Omega, php->G_C};
    if (cf->cg_beta < 0 || cf->cg_beta > 2) throw BecError(1, "cg_beta must be 'pr', 'pr_precond' or 'none'");
    solver_alloc(p);
    const bool lr = plan_long_range(p, ph);
    const double dv = p.dv;
    const long long n = p.n;
    double2* P[2] = {p.st[S_P0], p.st[S_P1]};
    double2* Gr[2] = {p.st[S_G0], p.st[S_G1]};
    double2* D = p.st[S_D];
    double2* Z = p.st[S_Z];
    double2* BEST = p.st[S_BEST];
    double* phi = p.stphi[0];
    double* phit = p.stphi[1];
    hipStream_t s = p.stream;

    const int maxit = cf->maxit;
    const double res_tol = cf->res_tol, etol = cf->energy_tol;
    const int check = std::max(1, cf->check);
    double alpha = cf->step;
    const double alpha_min = 1e-8, alpha_max = cf->step_max;
    const int ls_max = cf->linesearch_max;
    const int stall_max = std::max(1, cf->linesearch_stall);
    const double c1 = 1e-4, shrink = 0.5, growth = 1.1;
    const double bb_min = 0.05, bb_max = alpha_max, bb_mix = 0.5;
    const int restart_period = cf->cg_restart;
    const int conv_window = cf->conv_window;
    const double floor_ = cf->precond_shift_min;
    int n_tr = 0;

    // psi = psi0 / norm_of(psi0)                                   (solvers.py:56-58)
    add_job(p, C_NORM, launch_norm2((const double2*)psi_io, n, hpart_of(p, C_NORM), s), 1, 0, R_NORM);
    plan_sync(p);
    const double nrm0 = sqrt(p.hred[R_NORM] * dv);
    int c = 0;
    launch_scale((const double2*)psi_io, P[c], nrm0, n, s);
    HIPCHK(hipMemsetAsync(D, 0, sizeof(double2) * n, s));

    // E, phi = energy_phi(psi)                                       (solvers.py:106)
    long long n_energy = 1;
    run_energy_sums(p, P[c], nullptr, 0.0, lr ? p.rho : nullptr);
    if (lr) run_conv(p, 0, p.rho, -ph.G_C, phi, p.rho);
    HIPCHK(hipGetLastError());
    plan_sync(p);
    Energy e0 = energy_from_sums(p, ph, lr, false);
    double E = e0.E;
    double mu = e0.E + e0.Econtact + 2.0 * e0.Ethree + e0.Egrav;
    const double t0 = now_s();
    if (trace && trace_cap > 0) {
      trace[0] = 0.0; trace[1] = 0.0; trace[2] = E; trace[3] = NAN;
      n_tr = 1;
    }
    // grad, mu, resid = residual(psi, phi)                           (solvers.py:112)
    run_H(p, P[c], nullptr, 0.0, 1.0, nullptr, ph, true, lr ? phi : nullptr, 1.0, nullptr, mu, 0.5,
          -ph.Omega, Gr[c], true);
    HIPCHK(hipGetLastError());
    plan_sync(p);
    double resid_rel = sqrt(p.hred[R_HLAST] * dv) / std::max(1.0, fabs(mu));
    double best_resid = resid_rel;
    HIPCHK(hipMemcpyAsync(BEST, P[c], sizeof(double2) * n, hipMemcpyDeviceToDevice, s));

    bool have_prev = false;          // psi_prev / grad_prev differ from psi / grad
    bool converged = false;
    int stop = BEC_STOP_MAXIT;
    double E_ref = E;
    int it_ref = 0;
    int n_fail = 0, n_iter = 0;
    double sigma = floor_;
    double zg_prev = 0.0, gg_prev = 0.0;
    bool pending = false;            // an accepted step whose residual is still on the device
    int it_acc = 0;
    double trial_acc = 0.0;

    int it = 0;
    while (true) {
      const bool run_it = it < maxit;
      if (run_it) {
        // sigma, z = _preconditioner(grad, psi, phi)                 (solvers.py:129)
        // the mask's max rho comes from the per-block maxima of the last H pass
        launch_sigma(p, P[c], lr ? phi : nullptr, 1.0, part_of(p, C_HLAST), p.hlast_blocks, 4, 2, ph.beta2,
                     ph.beta3, floor_, s);
        run_precond(p, Gr[c], Z);
        const double2* pp = have_prev ? P[1 - c] : P[c];
        const double2* gp = have_prev ? Gr[1 - c] : Gr[c];
        add_job(p, C_DOTS, launch_dots(Z, P[c], D, Gr[c], gp, pp, n, hpart_of(p, C_DOTS), s), 16, 0, R_DOTS);
        HIPCHK(hipGetLastError());
      }
      plan_sync(p);
      if (pending) {
        // residual bookkeeping of the step accepted at iteration it_acc (solvers.py:196-221)
        pending = false;
        resid_rel = sqrt(p.hred[R_HLAST] * dv) / std::max(1.0, fabs(mu));
        if (trace && n_tr < trace_cap) {
          double* r = trace + 4 * n_tr;
          r[0] = now_s() - t0; r[1] = it_acc + 1; r[2] = E; r[3] = resid_rel;
          ++n_tr;
        }
        if (resid_rel < best_resid) {
          best_resid = resid_rel;
          HIPCHK(hipMemcpyAsync(BEST, P[c], sizeof(double2) * n, hipMemcpyDeviceToDevice, s));
        }
        if (cf->verbose && it_acc % check == 0)
          printf("  [gs] it=%d E=%.8f res=%.2e best=%.2e step=%.2e\n", it_acc, E, resid_rel, best_resid,
                 trial_acc);
        if (resid_rel < res_tol) {
          converged = true;
          stop = BEC_STOP_CONVERGED;
          if (cf->verbose) printf("[gs] converged it=%d: E=%.10g res=%.3e\n", it_acc, E, resid_rel);
          break;
        }
        if (it_acc - it_ref >= conv_window) {
          const double dE_win = fabs(E - E_ref) / std::max(1.0, fabs(E));
          if (dE_win < etol) {
            converged = true;
            stop = BEC_STOP_ENERGY_CONVERGED;
            if (cf->verbose)
              printf("[gs] energy converged it=%d: dE_window=%.2e res=%.3e best_res=%.3e\n", it_acc, dE_win,
                     resid_rel, best_resid);
            break;
          }
          E_ref = E;
          it_ref = it_acc;
        }
      }
      if (!run_it) break;
      n_iter = it + 1;
      const double* r = p.hred + R_DOTS;
      sigma = p.hred[R_SIGMA];
      const C2 cz = cm(r[0] * dv, r[1] * dv);        // <psi, z>
      const double pp_ = r[2] * dv;                  // <psi, psi>
      const C2 pd = cm(r[3] * dv, r[4] * dv);        // <psi, d>
      const C2 pg = cm(r[5] * dv, r[6] * dv);        // <psi, g>
      const double zg_raw = r[7] * dv, zy_raw = r[8] * dv;
      const C2 py = cm(r[9] * dv, r[10] * dv);       // <psi, y>
      const double gdr = r[11] * dv, gg = r[12] * dv, gy = r[13] * dv;
      const double sy = r[14] * dv, ss = r[15] * dv;
      // projected z_p = z - psi c
      const double zg = zg_raw - (conj(cz) * pg).re;   // Re <z_p, g>
      const double zy = zy_raw - (conj(cz) * py).re;   // Re <z_p, y>
      if (it == 0) { zg_prev = zg; gg_prev = gg; }     // z_prev = z, grad_prev = grad (solvers.py:115,121)
      double beta;
      if (it == 0 || cf->cg_beta == BEC_CG_NONE || (restart_period > 0 && it % restart_period == 0)) {
        beta = 0.0;
      } else {
        if (cf->cg_beta == BEC_CG_PR_PRECOND) {
          const double den = std::max(zg_prev, 1e-300);
          beta = std::max(0.0, zy / den);
        } else {
          const double den = std::max(gg_prev, 1e-300);
          beta = std::max(0.0, gy / den);
        }
        if (!std::isfinite(beta)) beta = 0.0;
      }
      // dirc = proj_tangent(z_p + beta d)                             (solvers.py:142-149)
      const C2 cw = cz - pp_ * cz + beta * pd;
      double az = 1.0, ad = beta, ag = 0.0;
      C2 gam = cm(0.0, 0.0) - (cz + cw);
      double gd = zg + beta * gdr - (conj(pg) * (cz + cw)).re;
      if (gd <= 0) {
        az = 1.0; ad = 0.0; ag = 0.0; gam = cm(-cz.re, -cz.im);
        gd = zg_raw - (conj(pg) * cz).re;
      }
      if (gd <= 0) {
        az = 0.0; ad = 0.0; ag = 1.0; gam = cm(-pg.re, -pg.im);
        gd = gg - (conj(pg) * pg).re;
      }
      if (it > 0) {                                    // Barzilai-Borwein (solvers.py:150-158)
        if (sy > 1e-30 && std::isfinite(sy)) {
          double bb = ss / sy;
          if (std::isfinite(bb)) {
            bb = std::min(std::max(bb, bb_min), bb_max);
            alpha = (1.0 - bb_mix) * alpha + bb_mix * bb;
          }
        }
      }
      alpha = std::min(std::max(alpha, alpha_min), alpha_max);
      launch_combine(D, Z, Gr[c], P[c], az, ad, ag, gam.re, gam.im, n, s);
      // Armijo backtracking                                          (solvers.py:160-174)
      bool accepted = false;
      double trial = alpha;
      Energy et{};
      for (int ls = 0; ls < ls_max; ++ls) {
        run_energy_sums(p, P[c], D, trial, lr ? p.rho : nullptr);
        if (lr) run_conv(p, 0, p.rho, -ph.G_C, phit, p.rho);
        HIPCHK(hipGetLastError());
        plan_sync(p);
        ++n_energy;
        et = energy_from_sums(p, ph, lr, true);
        if (std::isfinite(et.E) && et.E <= E - c1 * trial * gd) {
          accepted = true;
          break;
        }
        trial *= shrink;
        if (trial < alpha_min) break;
      }
      if (!accepted) {                                 // (solvers.py:175-185)
        n_fail += 1;
        alpha = std::max(alpha * 0.5, alpha_min);
        launch_combine(D, Z, nullptr, P[c], 1.0, 0.0, 0.0, -cz.re, -cz.im, n, s);   // direction = z
        if (alpha <= alpha_min || n_fail >= stall_max) {
          stop = BEC_STOP_LINE_SEARCH_STALLED;
          if (cf->verbose)
            printf("[gs] line search stalled it=%d: E=%.10g res=%.3e best_res=%.3e\n", it, E, resid_rel,
                   best_resid);
          break;
        }
        ++it;
        continue;
      }
      // accept                                                        (solvers.py:186-196)
      n_fail = 0;
      zg_prev = zg;
      gg_prev = gg;
      E = et.E;
      mu = et.E + et.Econtact + 2.0 * et.Ethree + et.Egrav;
      alpha = std::min(alpha_max, trial * growth);
      const double nrm = sqrt(et.nrm2);
      const double inv2 = 1.0 / et.nrm2;
      // psi_new = (psi - trial d) / nrm materialised into the other buffer;
      // grad_new = H psi_new - mu psi_new into the other grad buffer
      run_H(p, P[c], D, trial, nrm, P[1 - c], ph, true, lr ? phit : nullptr, inv2, lr ? phi : nullptr, mu,
            0.5, -ph.Omega, Gr[1 - c], true);
      HIPCHK(hipGetLastError());
      c = 1 - c;
      have_prev = true;
      pending = true;
      it_acc = it;
      trial_acc = trial;
      ++it;
    }
    if (best_resid < resid_rel) {
      HIPCHK(hipMemcpyAsync(psi_io, BEST, sizeof(double2) * n, hipMemcpyDeviceToDevice, s));
    } else {
      HIPCHK(hipMemcpyAsync(psi_io, P[c], sizeof(double2) * n, hipMemcpyDeviceToDevice, s));
    }
    plan_sync(p);
    out->iters = n_iter;
    out->converged = converged ? 1 : 0;
    out->stop_reason = stop;
    out->n_trace = n_tr;
    out->E = E;
    out->resid_rel = resid_rel;
    out->best_resid = best_resid;
    out->sigma = sigma;
    out->walltime = now_s() - t0;
    out->n_energy = n_energy;
    return 0;
  } catch (const BecError& e) {
    bec_set_error(e.what());
    return e.code;
  } catch (const std::exception& e) {
    bec_set_error(e.what());
    return 2;
  }
}
