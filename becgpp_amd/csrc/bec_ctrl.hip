// Controller kernels of the device-driven solver loop (becgpp/solvers.py:48-226).
//
// One step of the loop is a fixed launch sequence (shift selection,
// preconditioner, inner products, k_ctrl_dir, best copy, direction, trial
// energy, k_ctrl_armijo, failure direction, H of the accepted state).  Every
// launch covers the instances whose flag word carries its bit, and the two
// controllers rewrite the flags and the per-instance scalar table between
// them, so the host never waits inside a step.  An instance advances by one
// CG iteration per step when its first Armijo trial is accepted and retries
// its trial in the following steps otherwise, exactly as the reference's
// inner loop does.
//
// Each controller runs one workgroup per instance: it folds that instance's
// partial rows in a fixed order (deterministic), then thread 0 applies the
// reference's scalar logic to the instance's state.
#include <math.h>

#include "../../include/becgpp_hip.h"
#include "bec_internal.h"

#define CTRL_WG 1024

// Sums (q < K) or maxima (q >= K) of the R rows of KP values at rows[],
// folded as: lane j of value q takes rows j, j + J, ... in order (J =
// CTRL_WG / KP), a fixed pairwise tree combines the lanes of each wavefront,
// and the wavefronts' results are added in order; out[q] for q < KP.
// A fold runs in two phases so that a controller issues the loads of all its
// folds (and of its instance record) before it waits for any: issue() loads
// the first U rows of the lane, finish() adds them in order, continues over
// any further rows U at a time, and reduces over the lanes (every thread
// calls both).  The rows were written by other workgroups on other XCDs, so
// each wait is a trip to memory: one such trip per controller in the usual
// case (at most U J rows per fold).
template <int KP, int U>
struct RowFold {
  static constexpr int J = CTRL_WG / KP;
  double x[U];
  __device__ __forceinline__ void issue(const double* __restrict__ rows, int R) {
    const int q = threadIdx.x % KP, j = threadIdx.x / KP;
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int r = j + u * J;
      x[u] = r < R ? rows[(long long)r * KP + q] : 0.0;
    }
  }
  __device__ __forceinline__ void finish(const double* __restrict__ rows, int R, int K, double* out, double* sh) {
    const int q = threadIdx.x % KP, j = threadIdx.x / KP;
    const bool is_sum = q < K;
    double acc = is_sum ? 0.0 : -INFINITY;
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (j + u * J < R) acc = is_sum ? acc + x[u] : fmax(acc, x[u]);
    for (int r0 = j + U * J; r0 < R; r0 += U * J) {
      double y[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int r = r0 + u * J;
        y[u] = r < R ? rows[(long long)r * KP + q] : 0.0;
      }
#pragma unroll
      for (int u = 0; u < U; ++u)
        if (r0 + u * J < R) acc = is_sum ? acc + y[u] : fmax(acc, y[u]);
    }
    // the lanes' values: a pairwise tree over the W lanes of a wavefront
    // (shuffles), then the wavefronts' results in wavefront order
    constexpr int W = 64 / KP, NW = CTRL_WG / 64;
#pragma unroll
    for (int h = W / 2; h >= 1; h >>= 1) {
      const double v = __shfl_down(acc, h * KP, 64);
      acc = is_sum ? acc + v : fmax(acc, v);
    }
    if ((threadIdx.x & 63) < KP) sh[(threadIdx.x >> 6) * KP + q] = acc;
    __syncthreads();
    if (threadIdx.x < KP) {
      double t = sh[q];
      for (int w = 1; w < NW; ++w) t = is_sum ? t + sh[w * KP + q] : fmax(t, sh[w * KP + q]);
      out[q] = t;
    }
    __syncthreads();
  }
};

__device__ __forceinline__ double clk_now() { return (double)wall_clock64(); }

// one progress record into the instance's ring in host-mapped memory; the
// host prints the records after each poll (count after the records)
__device__ __forceinline__ void vlog_put(const CtrlArgs& c, int b, DevInst* q, int kind, int it, double E,
                                         double res, double best, double step) {
  if (!c.vlog) return;
  double* r = c.vlog + ((long long)b * VLOG_CAP + (q->n_log % VLOG_CAP)) * VLOG_REC;
  r[0] = kind; r[1] = it; r[2] = E; r[3] = res; r[4] = best; r[5] = step;
  q->n_log += 1;
  c.vlog[(long long)BEC_BMAX * VLOG_CAP * VLOG_REC + b] = (double)q->n_log;
}

// the instance record is worked on in LDS: one coalesced load and store
// instead of a chain of dependent global accesses by thread 0
constexpr int DEVINST_WORDS = (int)(sizeof(DevInst) / sizeof(unsigned long long));
static_assert(sizeof(DevInst) % sizeof(unsigned long long) == 0, "DevInst is a whole number of 8-byte words");
__device__ __forceinline__ void inst_load(DevInst* dst, const DevInst* src) {
  const unsigned long long* s = (const unsigned long long*)src;
  unsigned long long* d = (unsigned long long*)dst;
  for (int i = threadIdx.x; i < DEVINST_WORDS; i += blockDim.x) d[i] = s[i];
}
__device__ __forceinline__ void inst_store(DevInst* dst, const DevInst* src) {
  const unsigned long long* s = (const unsigned long long*)src;
  unsigned long long* d = (unsigned long long*)dst;
  for (int i = threadIdx.x; i < DEVINST_WORDS; i += blockDim.x) d[i] = s[i];
}

// ---------------------------------------------------------------------------
// k_ctrl_dir: residual bookkeeping of the step accepted last (solvers.py:
// 187-221), stop rules, then the CG direction, Barzilai-Borwein trial step
// and Armijo start (solvers.py:128-159) from this step's inner products.
// ---------------------------------------------------------------------------
// thread 0 of k_ctrl_dir: the scalar logic on the LDS copy of the record
__device__ __forceinline__ void ctrl_dir_scalar(const CtrlArgs& c, int b, unsigned f, bool dir, DevInst* q,
                                             const double* r, const double* hl, double sigma) {
  if (!dir) {                         // retrying a trial, or finished
    c.fl[b] = (f & FL_SEARCH) ? (FL_SEARCH | (q->lr ? FL_SLR : 0u)) : 0u;
    return;
  }
  const double dv = c.dv;
  unsigned nf = 0u;
  bool done = false;
  if (q->clk0 == 0.0) q->clk0 = clk_now();
  if (q->pending) {
    q->pending = 0;
    q->resid_rel = sqrt(hl[0] * dv) / fmax(1.0, fabs(q->mu));
    if (c.trace && b == 0 && q->n_tr < c.trace_cap) {
      double* t = c.trace + 4 * q->n_tr;
      t[0] = (clk_now() - q->clk0) / c.clk_hz;
      t[1] = q->it_acc + 1;
      t[2] = q->E;
      t[3] = q->resid_rel;
      q->n_tr += 1;
    }
    if (q->resid_rel < q->best_resid) {
      q->best_resid = q->resid_rel;
      nf |= FL_BEST;
    }
    if (c.verbose && q->it_acc % c.check == 0)
      vlog_put(c, b, q, VLOG_PROGRESS, q->it_acc, q->E, q->resid_rel, q->best_resid, q->trial_acc);
    if (q->resid_rel < c.res_tol) {
      q->converged = 1;
      q->stop = BEC_STOP_CONVERGED;
      done = true;
    } else if (q->it_acc - q->it_ref >= c.conv_window) {
      const double dE_win = fabs(q->E - q->E_ref) / fmax(1.0, fabs(q->E));
      if (dE_win < c.etol) {
        q->converged = 1;
        q->stop = BEC_STOP_ENERGY_CONVERGED;
        done = true;
      } else {
        q->E_ref = q->E;
        q->it_ref = q->it_acc;
      }
    }
  }
  if (!done && q->it >= c.maxit) done = true;
  if (done) {
    q->done = 1;
    c.hdone[b] = 1;
    c.fl[b] = nf;                     // the best copy still happens this step
    return;
  }
  const int it = q->it;
  q->n_iter = it + 1;
  q->sigma = sigma;
  // <psi,z>, <psi,psi>, <psi,d>, <psi,g>, Re<z,g>, Re<z,y>, <psi,y>, Re<g,d>, <g,g>, Re<g,y>, Re<s,y>, <s,s>
  const double czr = r[0] * dv, czi = r[1] * dv;
  const double pp = r[2] * dv;
  const double pdr = r[3] * dv, pdi = r[4] * dv;
  const double pgr = r[5] * dv, pgi = r[6] * dv;
  const double zg_raw = r[7] * dv, zy_raw = r[8] * dv;
  const double pyr = r[9] * dv, pyi = r[10] * dv;
  const double gdr = r[11] * dv, gg = r[12] * dv, gy = r[13] * dv;
  const double sy = r[14] * dv, ss = r[15] * dv;
  // Re(conj(a) b)
  const double cz_pg = czr * pgr + czi * pgi;
  const double cz_py = czr * pyr + czi * pyi;
  const double zg = zg_raw - cz_pg;                  // Re <z_p, g>, z_p = z - psi <psi, z>
  const double zy = zy_raw - cz_py;                  // Re <z_p, y>
  if (it == 0) { q->zg_prev = zg; q->gg_prev = gg; }
  double beta;
  if (it == 0 || c.cg_beta == BEC_CG_NONE || (c.restart_period > 0 && it % c.restart_period == 0)) {
    beta = 0.0;
  } else {
    if (c.cg_beta == BEC_CG_PR_PRECOND) beta = fmax(0.0, zy / fmax(q->zg_prev, 1e-300));
    else beta = fmax(0.0, gy / fmax(q->gg_prev, 1e-300));
    if (!isfinite(beta)) beta = 0.0;
  }
  // dirc = proj_tangent(z_p + beta d) = z + beta d + gam psi   (solvers.py:142-149)
  const double cwr = czr - pp * czr + beta * pdr, cwi = czi - pp * czi + beta * pdi;
  double az = 1.0, ad = beta, ag = 0.0;
  double gre = -(czr + cwr), gim = -(czi + cwi);
  double gd = zg_raw + beta * gdr - (pgr * (czr + cwr) + pgi * (czi + cwi));
  if (gd <= 0) {
    az = 1.0; ad = 0.0; ag = 0.0; gre = -czr; gim = -czi;
    gd = zg_raw - cz_pg;
  }
  if (gd <= 0) {
    az = 0.0; ad = 0.0; ag = 1.0; gre = -pgr; gim = -pgi;
    gd = gg - (pgr * pgr + pgi * pgi);
  }
  if (it > 0 && sy > 1e-30 && isfinite(sy)) {      // Barzilai-Borwein (solvers.py:150-158)
    double bb = ss / sy;
    if (isfinite(bb)) {
      bb = fmin(fmax(bb, c.bb_min), c.bb_max);
      q->alpha = (1.0 - c.bb_mix) * q->alpha + c.bb_mix * bb;
    }
  }
  q->alpha = fmin(fmax(q->alpha, c.alpha_min), c.alpha_max);
  q->gd = gd;
  q->trial = q->alpha;
  q->ls = 0;
  q->cz_re = czr;
  q->cz_im = czi;
  q->zg_cur = zg;                    // committed to zg_prev / gg_prev on acceptance
  q->gg_cur = gg;
  double* tb = c.tab + (long long)b * DS_N;
  tb[DS_AZ] = az; tb[DS_AD] = ad; tb[DS_AG] = ag; tb[DS_GRE] = gre; tb[DS_GIM] = gim;
  tb[DS_T] = q->trial;
  c.fl[b] = nf | FL_COMB | FL_SEARCH | (q->lr ? FL_SLR : 0u);
}

// ---------------------------------------------------------------------------
// k_ctrl_armijo: energy of the trial psi - t d (normalised in the sums,
// operators.py:32-33 via energy_components), the Armijo test, and the
// accept / shrink / fail transitions (solvers.py:160-196).
// ---------------------------------------------------------------------------
// thread 0 of k_ctrl_armijo: the Armijo test and transitions on the LDS copy
__device__ __forceinline__ void ctrl_armijo_scalar(const CtrlArgs& c, int b, unsigned f, DevInst* q, const double* e0,
                                                const double* e1, const double* e2, const double* gr) {
  const double dv = c.dv;
  const InstPhys ph = c.iphys[b];
  double kin = e0[0] + e1[0], lz = e0[1] + e1[1];
  if (c.ndim == 3) { kin += e2[0]; lz += e2[1]; }
  const double nrm2 = e0[2] * dv;
  const double i1 = 1.0 / nrm2, i2 = i1 * i1, i3 = i2 * i1;
  const double Ekin = kin * dv * i1, Etrap = e0[3] * dv * i1;
  const double Econtact = 0.5 * ph.beta2 * (e0[4] * dv) * i2;
  const double Ethree = (ph.beta3 / 3.0) * (e0[5] * dv) * i3;
  const double Erot = -ph.Omega * (lz * dv * i1);
  const double Egrav = (f & FL_SLR) ? 0.5 * (-ph.G_C) * (gr[0] * dv) * i2 : 0.0;
  const double Et = Ekin + Etrap + Econtact + Ethree + Erot + Egrav;
  q->n_energy += 1;
  double* tb = c.tab + (long long)b * DS_N;
  if (isfinite(Et) && Et <= q->E - c.c1 * q->trial * q->gd) {   // accepted (solvers.py:186-196)
    q->n_fail = 0;
    q->zg_prev = q->zg_cur;
    q->gg_prev = q->gg_cur;
    q->E = Et;
    q->mu = Et + Econtact + 2.0 * Ethree + Egrav;
    q->alpha = fmin(c.alpha_max, q->trial * c.growth);
    tb[DS_T] = q->trial;
    tb[DS_SC] = sqrt(nrm2);
    tb[DS_MU] = q->mu;
    tb[DS_PHISC] = q->lr ? -ph.G_C * i1 : 0.0;
    q->pending = 1;
    q->it_acc = q->it;
    q->trial_acc = q->trial;
    q->it += 1;
    c.fl[b] = (f & FL_BEST) | FL_ACC | FL_DIR;
    return;
  }
  q->trial *= c.shrink;
  q->ls += 1;
  if (!(q->trial < c.alpha_min || q->ls >= c.ls_max)) {          // next trial in the following step
    tb[DS_T] = q->trial;
    c.fl[b] = (f & FL_BEST) | FL_SEARCH | (q->lr ? FL_SLR : 0u);
    return;
  }
  // line search failed: direction = z_p, psi unchanged (solvers.py:175-185)
  q->n_fail += 1;
  q->alpha = fmax(q->alpha * 0.5, c.alpha_min);
  tb[DS_AZ] = 1.0; tb[DS_AD] = 0.0; tb[DS_AG] = 0.0; tb[DS_GRE] = -q->cz_re; tb[DS_GIM] = -q->cz_im;
  q->it += 1;
  if (q->alpha <= c.alpha_min || q->n_fail >= c.stall_max) {
    q->stop = BEC_STOP_LINE_SEARCH_STALLED;
    q->done = 1;
    c.hdone[b] = 1;
    if (c.verbose) vlog_put(c, b, q, VLOG_STALL, q->it - 1, q->E, q->resid_rel, q->best_resid, q->trial);
    c.fl[b] = f & FL_BEST;
    return;
  }
  c.fl[b] = (f & FL_BEST) | FL_FAIL | FL_DIR;
}

// ---------------------------------------------------------------------------
// k_ctrl_line: the whole Armijo backtracking of one iteration (solvers.py:
// 160-196) from the line coefficients of this step.  With p = |psi|^2,
// q = Re(conj(psi) d), r = |d|^2 the trial u = psi - t d has the density
// p - 2 t q + t^2 r and the norm n(t) = 1 - 2 t <q> + t^2 <r> (psi has unit
// norm), and its energy after normalisation is
//   E(t) = QL(t)/n + (beta2/2) Q4(t)/n^2 + (beta3/3) Q6(t)/n^3 - (G_C/2) P(t)/n^2
// with QL = <u,(T + V - Omega L_z) u>, Q4 = sum rho^2, Q6 = sum rho^3,
// P = sum rho (K*rho) (all times dV).  Their values at t = 0 are the state's,
// carried from the trial that produced it (L00, Spp, Sppp, Kpp); the cross
// term of QL uses <psi, L d> = <H psi, d> - sum (beta2 p + beta3 p^2 + phi) q
// with H psi = g + mu psi.  The reference's sequence of trials alpha,
// alpha/2, ... and its acceptance test then run here, in the same order, on
// these energies: every trial of the line costs no further pass over the
// fields.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void ctrl_line_scalar(const CtrlArgs& c, int b, unsigned f, DevInst* q, const double* e0,
                                                 const double* e1, const double* e2, const double* kq,
                                                 const double* kr, const InstPhys& ph) {
  const double dv = c.dv;
  double kin = e0[0] + e1[0], lz = e0[1] + e1[1];
  if (c.ndim == 3) { kin += e2[0]; lz += e2[1]; }
  const double a1 = e0[2] * dv, a2 = e0[3] * dv;
  const double L11 = (kin - ph.Omega * lz + e0[4]) * dv;
  const double L01 = q->gd + q->mu * a1 - (ph.beta2 * e0[5] + ph.beta3 * e0[10] + e0[19]) * dv;
  // sum rho^2: Spp - 4t pq + t^2 (4 qq + 2 pr) - 4t^3 qr + t^4 rr
  const double c41 = -4.0 * e0[5] * dv, c42 = (4.0 * e0[6] + 2.0 * e0[7]) * dv, c43 = -4.0 * e0[8] * dv,
               c44 = e0[9] * dv;
  // sum rho^3 with rho = p + t B + t^2 r, B = -2 q
  const double c61 = -6.0 * e0[10] * dv, c62 = (3.0 * e0[11] + 12.0 * e0[12]) * dv,
               c63 = (-8.0 * e0[13] - 12.0 * e0[14]) * dv, c64 = (12.0 * e0[15] + 3.0 * e0[16]) * dv,
               c65 = -6.0 * e0[17] * dv, c66 = e0[18] * dv;
  // sum rho (K*rho): Kpp - 4t Kpq + t^2 (4 Kqq + 2 Kpr) - 4t^3 Kqr + t^4 Krr; the
  // stored potential is phi = -G_C K*p, so K(p,q) = sum phi q / (-G_C)
  const bool lr = (f & FL_SLR) != 0u;
  double cp1 = 0.0, cp2 = 0.0, cp3 = 0.0, cp4 = 0.0;
  if (lr) {
    const double ig = 1.0 / (-ph.G_C);
    cp1 = -4.0 * e0[19] * dv * ig;
    cp2 = 4.0 * kq[0] * dv + 2.0 * e0[20] * dv * ig;
    cp3 = -4.0 * kq[1] * dv;
    cp4 = kr[0] * dv;
  }
  double t = q->trial;
  bool accepted = false;
  double Et = 0.0, Ec = 0.0, E3 = 0.0, Eg = 0.0, nt = 1.0, QL = 0.0, Q4 = 0.0, Q6 = 0.0, P = 0.0;
  for (int ls = 0; ls < c.ls_max; ++ls) {
    nt = 1.0 + t * (-2.0 * a1 + t * a2);
    const double i1 = 1.0 / nt, i2 = i1 * i1, i3 = i2 * i1;
    QL = q->L00 + t * (-2.0 * L01 + t * L11);
    Q4 = q->Spp + t * (c41 + t * (c42 + t * (c43 + t * c44)));
    Q6 = q->Sppp + t * (c61 + t * (c62 + t * (c63 + t * (c64 + t * (c65 + t * c66)))));
    P = lr ? q->Kpp + t * (cp1 + t * (cp2 + t * (cp3 + t * cp4))) : 0.0;
    Ec = 0.5 * ph.beta2 * Q4 * i2;
    E3 = (ph.beta3 / 3.0) * Q6 * i3;
    Eg = lr ? 0.5 * (-ph.G_C) * P * i2 : 0.0;
    Et = QL * i1 + Ec + E3 + Eg;
    q->n_energy += 1;
    if (isfinite(Et) && Et <= q->E - c.c1 * t * q->gd) {
      accepted = true;
      break;
    }
    t *= c.shrink;
    if (t < c.alpha_min) break;
  }
  q->trial = t;
  double* tb = c.tab + (long long)b * DS_N;
  if (accepted) {                                                 // (solvers.py:186-196)
    const double i1 = 1.0 / nt, i2 = i1 * i1;
    q->n_fail = 0;
    q->zg_prev = q->zg_cur;
    q->gg_prev = q->gg_cur;
    q->E = Et;
    q->mu = Et + Ec + 2.0 * E3 + Eg;
    q->alpha = fmin(c.alpha_max, t * c.growth);
    q->L00 = QL * i1;
    q->Spp = Q4 * i2;
    q->Sppp = Q6 * i2 * i1;
    q->Kpp = P * i2;
    tb[DS_T] = t;
    tb[DS_SC] = sqrt(nt);
    tb[DS_MU] = q->mu;
    tb[DS_PHISC] = 1.0;
    // phi of the new state: -G_C K*(p - 2t q + t^2 r) / n = (phi + 2t G_C K*q - t^2 G_C K*r) / n
    tb[DS_CQ] = 2.0 * t * ph.G_C;
    tb[DS_CR] = -t * t * ph.G_C;
    tb[DS_CS] = i1;
    q->pending = 1;
    q->it_acc = q->it;
    q->trial_acc = t;
    q->it += 1;
    c.fl[b] = (f & FL_BEST) | FL_ACC | FL_DIR | (lr ? FL_PHIC : 0u);
    return;
  }
  // line search failed: direction = z_p, psi unchanged (solvers.py:175-185)
  q->n_fail += 1;
  q->alpha = fmax(q->alpha * 0.5, c.alpha_min);
  tb[DS_AZ] = 1.0; tb[DS_AD] = 0.0; tb[DS_AG] = 0.0; tb[DS_GRE] = -q->cz_re; tb[DS_GIM] = -q->cz_im;
  q->it += 1;
  if (q->alpha <= c.alpha_min || q->n_fail >= c.stall_max) {
    q->stop = BEC_STOP_LINE_SEARCH_STALLED;
    q->done = 1;
    c.hdone[b] = 1;
    if (c.verbose) vlog_put(c, b, q, VLOG_STALL, q->it - 1, q->E, q->resid_rel, q->best_resid, q->trial);
    c.fl[b] = f & FL_BEST;
    return;
  }
  c.fl[b] = (f & FL_BEST) | FL_FAIL | FL_DIR;
}

__global__ __launch_bounds__(CTRL_WG) void k_ctrl_line(CtrlArgs c) {
  __shared__ double sh[CTRL_WG];
  __shared__ double e0[32], e1[2], e2[2], kq[2], kr[1];
  __shared__ DevInst qs;
  const int b = blockIdx.x;
  // every load is issued before any is waited for: the flag word, the rows
  // (the long-range ones too, used only when the flag asks for them), the
  // instance record and its physics
  const unsigned f = c.fl[b];
  const double* r0 = c.r_e0 + (long long)b * c.n_e0 * 32;
  const double* r1 = c.r_e1 + (long long)b * c.n_e1 * 2;
  const double* r2 = c.ndim == 3 ? c.r_e2 + (long long)b * c.n_e2 * 2 : nullptr;
  const double* rq = c.r_kq + (long long)b * c.n_kq * 2;
  const double* rr = c.r_kr + (long long)b * c.n_kr;
  RowFold<32, 32> f0;
  RowFold<2, 2> f1, f2, fq;
  RowFold<1, 2> fr;
  f0.issue(r0, c.n_e0);
  f1.issue(r1, c.n_e1);
  if (r2) f2.issue(r2, c.n_e2);
  fq.issue(rq, c.n_kq);
  fr.issue(rr, c.n_kr);
  const InstPhys ph = c.iphys[b];
  if (!(f & FL_SEARCH)) {             // finished this step: only a pending best copy remains
    if (threadIdx.x == 0) c.fl[b] = f & FL_BEST;
    return;
  }
  const bool lr = (f & FL_SLR) != 0u;
  inst_load(&qs, c.st + b);
  f0.finish(r0, c.n_e0, 21, e0, sh);
  f1.finish(r1, c.n_e1, 2, e1, sh);
  if (r2) f2.finish(r2, c.n_e2, 2, e2, sh);
  if (lr) {
    fq.finish(rq, c.n_kq, 2, kq, sh);
    fr.finish(rr, c.n_kr, 1, kr, sh);
  }
  if (threadIdx.x == 0) ctrl_line_scalar(c, b, f, &qs, e0, e1, e2, kq, kr, ph);
  __syncthreads();
  inst_store(c.st + b, &qs);
}

__global__ __launch_bounds__(CTRL_WG) void k_ctrl_dir(CtrlArgs c) {
  __shared__ double sh[CTRL_WG];
  __shared__ double r[16], hl[4];
  __shared__ DevInst qs;
  const int b = blockIdx.x;
  DevInst* q = &qs;
  const unsigned f = c.fl[b];
  // the rows of the accepted state's H pass are loaded with the inner
  // products and used when the record says that a step is pending; all
  // loads (flags, rows, record, shift) are issued before any is waited for
  const double* rd = c.r_dots + (long long)b * c.n_dots * 16;
  const double* rh = c.r_hl + (long long)b * c.n_hl * 4;
  RowFold<16, 16> fd;
  RowFold<4, 4> fh;
  fd.issue(rd, c.n_dots);
  fh.issue(rh, c.n_hl);
  const double sig = c.sigma[b];
  inst_load(q, c.st + b);
  const bool dir = (f & FL_DIR) != 0u;
  __syncthreads();
  if (dir) {
    fd.finish(rd, c.n_dots, 16, r, sh);
    if (q->pending) fh.finish(rh, c.n_hl, 2, hl, sh);
  }
  if (threadIdx.x == 0) ctrl_dir_scalar(c, b, f, dir, q, r, hl, sig);
  __syncthreads();
  inst_store(c.st + b, q);
}

__global__ __launch_bounds__(CTRL_WG) void k_ctrl_armijo(CtrlArgs c) {
  __shared__ double sh[CTRL_WG];
  __shared__ double e0[8], e1[2], e2[2], gr[1];
  __shared__ DevInst qs;
  const int b = blockIdx.x;
  const unsigned f = c.fl[b];
  const bool search = (f & FL_SEARCH) != 0u;
  if (!search) {                      // finished this step: only a pending best copy remains
    if (threadIdx.x == 0) c.fl[b] = f & FL_BEST;
    return;
  }
  const bool lr = (f & FL_SLR) != 0u;
  const double* r0 = c.r_e0 + (long long)b * c.n_e0 * 8;
  const double* r1 = c.r_e1 + (long long)b * c.n_e1 * 2;
  const double* r2 = c.ndim == 3 ? c.r_e2 + (long long)b * c.n_e2 * 2 : nullptr;
  const double* rg = lr ? c.r_gr + (long long)b * c.n_gr : nullptr;
  RowFold<8, 8> f0;
  RowFold<2, 2> f1, f2;
  RowFold<1, 2> fg;
  f0.issue(r0, c.n_e0);
  f1.issue(r1, c.n_e1);
  if (r2) f2.issue(r2, c.n_e2);
  if (lr) fg.issue(rg, c.n_gr);
  inst_load(&qs, c.st + b);
  f0.finish(r0, c.n_e0, 6, e0, sh);
  f1.finish(r1, c.n_e1, 2, e1, sh);
  if (r2) f2.finish(r2, c.n_e2, 2, e2, sh);
  if (lr) fg.finish(rg, c.n_gr, 1, gr, sh);
  if (threadIdx.x == 0) ctrl_armijo_scalar(c, b, f, &qs, e0, e1, e2, gr);
  __syncthreads();
  inst_store(c.st + b, &qs);
}

void launch_ctrl_dir(const CtrlArgs& c, hipStream_t s) {
  prof_pre(SITE_CTRL_DIR, s);
  hipLaunchKernelGGL(k_ctrl_dir, dim3(c.nb), dim3(CTRL_WG), 0, s, c);
  prof_post(SITE_CTRL_DIR, s);
}
void launch_ctrl_armijo(const CtrlArgs& c, hipStream_t s) {
  prof_pre(SITE_CTRL_ARMIJO, s);
  hipLaunchKernelGGL(k_ctrl_armijo, dim3(c.nb), dim3(CTRL_WG), 0, s, c);
  prof_post(SITE_CTRL_ARMIJO, s);
}
void launch_ctrl_line(const CtrlArgs& c, hipStream_t s) {
  prof_pre(SITE_CTRL_LINE, s);
  hipLaunchKernelGGL(k_ctrl_line, dim3(c.nb), dim3(CTRL_WG), 0, s, c);
  prof_post(SITE_CTRL_LINE, s);
}
