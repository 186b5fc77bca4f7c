// Ground-state solver loop (becgpp/solvers.py:48-226) over a batch of
// instances of one grid, or over the batches of several grids at once.
// Fields stay in HBM, laid out [instance][point]; every kernel launch covers
// all instances of its grid that take part in the current phase.  The scalar
// logic of each instance is the reference's, in the same order, run by
// controller kernels on the device (bec_ctrl.hip) between graph-replayed
// steps; the same logic on the host (host_loop) is kept for comparison.
//
// psi_prev and grad_prev are kept as the differences s = psi - psi_prev and
// y = grad - grad_prev, written by the H passes when a step is accepted, so
// every instance updates psi and grad in place.
#include <math.h>
#include <stdio.h>
#include <string.h>
#include <algorithm>
#include <chrono>
#include <functional>
#include <vector>

#include "../../include/becgpp_hip.h"
#include "bec_ops.h"

enum { S_PSI = 0, S_SD, S_G, S_Y, S_D, S_Z, S_BEST, S_COUNT };

static void dev_release(Plan& p);

void solver_release(Plan& p) {
  dev_release(p);
  for (int i = 0; i < 8; ++i)
    if (p.st[i]) (void)hipFree(p.st[i]);
  for (int i = 0; i < 6; ++i)
    if (p.stphi[i]) (void)hipFree(p.stphi[i]);
  for (int i = 0; i < 8; ++i)
    if (p.st32[i]) (void)hipFree(p.st32[i]);
  memset(p.st, 0, sizeof(p.st));
  memset(p.stphi, 0, sizeof(p.stphi));
  memset(p.st32, 0, sizeof(p.st32));
  p.scap = 0;
}

static void solver_alloc(Plan& p, int nb) {
  if (nb <= p.scap) return;
  HIPCHK(hipStreamSynchronize(p.stream));
  solver_release(p);
  for (int i = 0; i < S_COUNT; ++i) HIPCHK(hipMalloc(&p.st[i], sizeof(double2) * p.n * nb));
  for (int i = 0; i < 6; ++i) HIPCHK(hipMalloc(&p.stphi[i], sizeof(double) * p.n * nb));
  p.scap = nb;
}

// complex64 copies of the solver state for an FP32 phase (sized like st)
static void solver_alloc32(Plan& p) {
  if (p.st32[0]) return;
  HIPCHK(hipStreamSynchronize(p.stream));
  for (int i = 0; i < S_COUNT; ++i) HIPCHK(hipMalloc(&p.st32[i], sizeof(float2) * p.n * p.scap));
}

static void dev_release(Plan& p) {
  void* d[] = {p.dinst, p.flags, p.dtab, p.dtrace};
  for (void* q : d)
    if (q) (void)hipFree(q);
  if (p.hdone) (void)hipHostFree(p.hdone);
  if (p.hvlog) (void)hipHostFree(p.hvlog);
  p.hvlog = nullptr;
  p.hvlog_dev = nullptr;
  if (p.cstream) (void)hipStreamDestroy(p.cstream);
  p.cstream = nullptr;

  p.dinst = nullptr; p.flags = nullptr; p.dtab = nullptr; p.dtrace = nullptr; p.dtrace_cap = 0;
  p.hdone = nullptr; p.hdone_dev = nullptr;
}

// device-loop state for nb instances (sized with the solver state) and a
// trace buffer of trace_cap rows
static void dev_alloc(Plan& p, int nb, int trace_cap) {
  if (!p.dinst) {
    HIPCHK(hipMalloc(&p.dinst, sizeof(DevInst) * BEC_BMAX));
    HIPCHK(hipMalloc(&p.flags, sizeof(unsigned) * BEC_BMAX));
    HIPCHK(hipMalloc(&p.dtab, sizeof(double) * DS_N * BEC_BMAX));
    HIPCHK(hipHostMalloc(&p.hdone, sizeof(int) * BEC_BMAX, hipHostMallocMapped | hipHostMallocCoherent));
    HIPCHK(hipHostGetDevicePointer((void**)&p.hdone_dev, p.hdone, 0));
    const size_t nv = (size_t)BEC_BMAX * VLOG_CAP * VLOG_REC + BEC_BMAX;
    HIPCHK(hipHostMalloc(&p.hvlog, sizeof(double) * nv, hipHostMallocMapped | hipHostMallocCoherent));
    HIPCHK(hipHostGetDevicePointer((void**)&p.hvlog_dev, p.hvlog, 0));
  }
  if (!p.cstream) HIPCHK(hipStreamCreateWithFlags(&p.cstream, hipStreamNonBlocking));

  if (trace_cap > p.dtrace_cap) {
    if (p.dtrace) (void)hipFree(p.dtrace);
    p.dtrace = nullptr;
    p.dtrace_cap = 0;
    HIPCHK(hipMalloc(&p.dtrace, sizeof(double) * 4 * (size_t)trace_cap));
    p.dtrace_cap = trace_cap;
  }
  (void)nb;
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

// per-instance scalar state of the loop (the reference's local variables)
struct Inst {
  double E, mu, resid_rel, best_resid, alpha, sigma, zg_prev, gg_prev, zg_cur, gg_cur, E_ref, trial, gd,
      trial_acc;
  double cz_re, cz_im;
  double L00, Spp, Sppp, Kpp;   // polynomial search: the state's quadratic form and pair sums (normalised)
  int it_ref, n_fail, n_iter, stop, it_acc, ls;
  bool converged, done, pending, lr;
  long long n_energy;
  Energy et;
};

static Batch mask_batch(int nb, unsigned long long m) { return Batch{nb, m}; }
static unsigned long long low_bits(int nb) { return nb >= 64 ? ~0ull : ((1ull << nb) - 1ull); }

// Prints the progress records the controllers wrote since the last call.
static void vlog_flush(const Plan& p, int nb, std::vector<int>& printed) {
  const double* cnt = p.hvlog + (size_t)BEC_BMAX * VLOG_CAP * VLOG_REC;
  for (int b = 0; b < nb; ++b) {
    const int n = (int)((volatile const double*)cnt)[b];
    for (int k = std::max(printed[b], n - VLOG_CAP); k < n; ++k) {
      const double* r = p.hvlog + ((size_t)b * VLOG_CAP + (size_t)(k % VLOG_CAP)) * VLOG_REC;
      if ((int)r[0] == VLOG_PROGRESS)
        printf("  [gs%s%d] it=%d E=%.8f res=%.2e best=%.2e step=%.2e\n", nb > 1 ? " " : "", nb > 1 ? b : 0,
               (int)r[1], r[2], r[3], r[4], r[5]);
      else
        printf("[gs] line search stalled it=%d: E=%.10g res=%.3e best_res=%.3e\n", (int)r[1], r[2], r[3], r[4]);
    }
    printed[b] = n;
  }
  fflush(stdout);
}

// One plan's share of the device-driven loop (bec_ctrl.hip): its instances'
// state in device memory, the controllers' arguments and the launches of one
// step.  run_loops() drives one or several of them, each plan's steps on its
// own stream.
struct DevLoop {
  Plan* pp = nullptr;
  int nb = 0;
  const bec_solver_cfg* cf = nullptr;
  std::vector<Inst>* I = nullptr;
  bool lr_any = false, lp = false;
  double* trace = nullptr;
  int trace_cap = 0;
  int* n_tr = nullptr;
  hipStream_t s = nullptr;             // the caller's stream (p.stream on entry)
  CtrlArgs c;
  std::vector<DevInst> di;
  std::vector<int> printed;
  double floor_ = 0.0;
  long long max_steps = 0;
  // a queue of more solves than slots: whether stopped slots should be
  // refilled now, the refill itself (called with every stream drained), and
  // whether solves are still waiting for a slot
  std::function<bool()> want_refill;
  std::function<void()> refill;
  std::function<bool()> pending;
  double2 *PSI = nullptr, *SD = nullptr, *G = nullptr, *Y = nullptr, *D = nullptr, *Z = nullptr, *BEST = nullptr;
  double *phi = nullptr, *phit = nullptr, *LQ = nullptr, *LR = nullptr, *KQ = nullptr, *KR = nullptr;

  // Uploads the instances' state and fills the controllers' arguments.
  void begin(Plan& p, int nb_, const bec_solver_cfg* cf_, std::vector<Inst>& I_, bool lr_any_, double* trace_,
             int trace_cap_, int& n_tr_, bool lp_) {
    pp = &p;
    nb = nb_;
    cf = cf_;
    I = &I_;
    lr_any = lr_any_;
    lp = lp_;
    trace = trace_;
    trace_cap = trace_cap_;
    n_tr = &n_tr_;
    s = p.stream;
    // an FP32 phase runs the same passes on the complex64 fields (pointers
    // reinterpreted by the launches, which all see p.f32)
    double2* const* F = p.f32 ? (double2* const*)p.st32 : p.st;
    PSI = F[S_PSI];
    SD = F[S_SD];
    G = F[S_G];
    Y = F[S_Y];
    D = F[S_D];
    Z = F[S_Z];
    BEST = F[S_BEST];
    phi = p.stphi[0];
    phit = p.stphi[1];
    LQ = p.stphi[2];                   // q = Re(conj(psi) d), r = |d|^2 and their convolutions
    LR = p.stphi[3];
    KQ = p.stphi[4];
    KR = p.stphi[5];
    dev_alloc(p, nb, trace ? trace_cap : 0);
    di.assign(nb, DevInst{});
    printed.assign(nb, 0);
    upload(low_bits(nb), s);

    memset(&c, 0, sizeof(c));
    c.nb = nb;
    c.ndim = p.ndim;
    c.dv = p.dv;
    c.maxit = cf->maxit;
    c.check = std::max(1, cf->check);
    c.ls_max = cf->linesearch_max;
    c.stall_max = std::max(1, cf->linesearch_stall);
    c.restart_period = cf->cg_restart;
    c.conv_window = cf->conv_window;
    c.cg_beta = cf->cg_beta;
    c.verbose = cf->verbose;
    c.res_tol = cf->res_tol;
    c.etol = cf->energy_tol;
    c.alpha_min = 1e-8;
    c.alpha_max = cf->step_max;
    c.c1 = 1e-4;
    c.shrink = 0.5;
    c.growth = 1.1;
    c.bb_min = 0.05;
    c.bb_max = cf->step_max;
    c.bb_mix = 0.5;
    c.r_dots = part_of(p, C_DOTS);
    c.r_hl = part_of(p, C_HLAST);
    c.n_hl = p.hlast_rows;
    c.r_e0 = part_of(p, C_EFIRST);
    c.r_e1 = part_of(p, C_EREST0);
    c.r_e2 = part_of(p, C_EREST1);
    c.r_gr = part_of(p, C_EGRAV);
    if (lp) c.r_e0 = part_of(p, C_LINE);
    c.r_kq = part_of(p, C_EGRAV);
    c.r_kr = part_of(p, C_EGRAV2);
    c.sigma = p.sigma;
    c.iphys = p.iphys;
    c.st = (DevInst*)p.dinst;
    c.fl = p.flags;
    c.tab = p.dtab;
    c.trace = trace ? p.dtrace : nullptr;
    c.trace_cap = trace ? trace_cap : 0;
    int dev = 0, khz = 0;
    HIPCHK(hipGetDevice(&dev));
    HIPCHK(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev));
    c.clk_hz = 1e3 * (double)std::max(khz, 1);
    c.hdone = p.hdone_dev;
    c.vlog = cf->verbose ? p.hvlog_dev : nullptr;
    floor_ = cf->precond_shift_min;
    // every step resolves a trial of each running instance: maxit iterations of
    // at most linesearch_max trials bound the number of steps
    max_steps = (long long)(cf->maxit + 2) * (long long)(std::max(1, cf->linesearch_max) + 1) + 64;
  }

  // One step on p.stream (the loop's stream while it runs).  The polynomial
  // search resolves a whole iteration of every running instance (an accepted
  // trial or a failed line search) per step; the trial search one Armijo trial.
  void step() {
    Plan& p = *pp;
    hipStream_t st = p.stream;
    const Batch BD = dev_batch(p, nb, FL_DIR);
    launch_sigma(p, nb, act_of(BD), PSI, lr_any ? phi : nullptr, part_of(p, C_HLAST), p.hlast_rows, 4, 2, floor_,
                 st, p.f32);
    // z = P^-1 g, with the CG inner products accumulated by its last pass
    const DotFields df{PSI, D, G, Y, SD};
    c.n_dots = run_precond(p, BD, G, Z, &df);
    launch_ctrl_dir(c, st);
    HArgs a;
    memset(&a, 0, sizeof(a));
    a.psi = PSI;
    a.d = D;
    a.psi_w = PSI;
    a.sdiff = SD;
    a.local_on = true;
    a.kf = 0.5;
    a.rf_omega = 1;
    a.out = G;
    a.ydiff = Y;
    a.reduce = true;
    if (lp) {
      const EnergyRows er = run_line_sums(p, dev_batch(p, nb, FL_SEARCH), PSI, D, Z, G, lr_any ? phi : nullptr,
                                          lr_any ? LQ : nullptr, lr_any ? LR : nullptr);
      c.n_e0 = er.r0;
      c.n_e1 = er.r1;
      c.n_e2 = er.r2;
      if (lr_any) {
        Batch BL = dev_batch(p, nb, FL_SLR);
        BL.f32 = 0;                      // the convolutions are FP64 in every phase
        // K*q and K*r as one complex convolution of q + i r
        c.n_kq = c.n_kr = run_conv_pair(p, BL, LQ, LR, KQ, KR, C_EGRAV, C_EGRAV2);
      }
      launch_ctrl_line(c, st);
      launch_post(PSI, BEST, D, Z, p.dtab, p.flags, p.n, nb, p.f32, st);
      // the last H pass first forms the accepted state's potential
      // phi <- (phi + 2t G_C K*q - t^2 G_C K*r) / n (instances flagged FL_PHIC)
      a.phi = lr_any ? phi : nullptr;
      a.phi_out = lr_any ? phi : nullptr;
      a.kq = lr_any ? KQ : nullptr;
      a.kr = lr_any ? KR : nullptr;
      c.n_hl = run_H(p, dev_batch(p, nb, FL_ACC), a);
      return;
    }
    // trial energy of psi - t d; instances with a new direction form d in the
    // same pass, and the pass keeps its share of H for the accepting step
    EnergyFuse fu;
    fu.hy_out = p.w1;
    fu.z = Z;
    fu.g = G;
    fu.d_out = D;
    fu.comb_bit = FL_COMB;
    const EnergyRows er = run_energy_sums(p, dev_batch(p, nb, FL_SEARCH), PSI, D, nullptr,
                                          lr_any ? p.rho : nullptr, &fu);
    c.n_e0 = er.r0;
    c.n_e1 = er.r1;
    c.n_e2 = er.r2;
    if (lr_any) c.n_gr = run_conv(p, 0, dev_batch(p, nb, FL_SLR), p.rho, 1.0, phit, p.rho);
    launch_ctrl_armijo(c, st);
    launch_post(PSI, BEST, D, Z, p.dtab, p.flags, p.n, nb, p.f32, st);
    a.phi = lr_any ? phit : nullptr;
    a.phi_out = lr_any ? phi : nullptr;
    c.n_hl = run_H_accept_dev(p, dev_batch(p, nb, FL_ACC), a);
  }

  // The slots of mask start their instances: each record from I[] (the
  // scalar state slots_begin set up), flags FL_DIR, done flag cleared.  On
  // stream st, synchronised.
  void upload(unsigned long long mask, hipStream_t st) {
    Plan& p = *pp;
    const unsigned fdir = FL_DIR;
    for (int b = 0; b < nb; ++b) {
      if (!((mask >> b) & 1ull)) continue;
      const Inst& q = (*I)[b];
      memset(&di[b], 0, sizeof(DevInst));
      di[b].E = q.E;
      di[b].mu = q.mu;
      di[b].E_ref = q.E_ref;
      di[b].resid_rel = q.resid_rel;
      di[b].best_resid = q.best_resid;
      di[b].alpha = q.alpha;
      di[b].sigma = q.sigma;
      di[b].stop = BEC_STOP_MAXIT;
      di[b].lr = q.lr ? 1 : 0;
      di[b].n_energy = q.n_energy;
      di[b].L00 = q.L00;
      di[b].Spp = q.Spp;
      di[b].Sppp = q.Sppp;
      di[b].Kpp = q.Kpp;
      di[b].n_tr = *n_tr;
      p.hdone[b] = 0;
      printed[b] = 0;
      if (cf->verbose) p.hvlog[(size_t)BEC_BMAX * VLOG_CAP * VLOG_REC + b] = 0.0;
      // copies on the given stream, never the legacy one
      HIPCHK(hipMemcpyAsync((DevInst*)p.dinst + b, &di[b], sizeof(DevInst), hipMemcpyHostToDevice, st));
      HIPCHK(hipMemcpyAsync(p.flags + b, &fdir, sizeof(unsigned), hipMemcpyHostToDevice, st));
    }
    HIPCHK(hipStreamSynchronize(st));
  }

  // the slots whose instances have stopped (the controllers' host-mapped flags)
  unsigned long long done_mask() const {
    unsigned long long m = 0;
    for (int b = 0; b < nb; ++b)
      if (((volatile int*)pp->hdone)[b] != 0) m |= 1ull << b;
    return m;
  }

  // every instance has stopped and no solve waits for a slot
  bool finished() const {
    return done_mask() == low_bits(nb) && !(pending && pending());
  }

  // The final state of the slots of mask read back into I[] (on stream st).
  void harvest(unsigned long long mask, hipStream_t st) {
    Plan& p = *pp;
    for (int b = 0; b < nb; ++b)
      if ((mask >> b) & 1ull)
        HIPCHK(hipMemcpyAsync(&di[b], (DevInst*)p.dinst + b, sizeof(DevInst), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    for (int b = 0; b < nb; ++b) {
      if (!((mask >> b) & 1ull)) continue;
      Inst& q = (*I)[b];
      q.E = di[b].E;
      q.resid_rel = di[b].resid_rel;
      q.best_resid = di[b].best_resid;
      q.n_iter = di[b].n_iter;
      q.converged = di[b].converged != 0;
      q.stop = di[b].stop;
      q.sigma = di[b].sigma;
      q.n_energy = di[b].n_energy;
      q.done = true;
    }
  }

  void log() {
    if (cf->verbose) vlog_flush(*pp, nb, printed);
  }

  // Reads the instances' final state back into I[] and the trace rows.
  void end() {
    Plan& p = *pp;
    log();
    HIPCHK(hipMemcpyAsync(di.data(), p.dinst, sizeof(DevInst) * nb, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    for (int b = 0; b < nb; ++b) {
      Inst& q = (*I)[b];
      q.E = di[b].E;
      q.resid_rel = di[b].resid_rel;
      q.best_resid = di[b].best_resid;
      q.n_iter = di[b].n_iter;
      q.converged = di[b].converged != 0;
      q.stop = di[b].stop;
      q.sigma = di[b].sigma;
      q.n_energy = di[b].n_energy;
      q.done = true;
    }
    if (trace && di[0].n_tr > *n_tr) {
      std::vector<double> rows(4 * (size_t)di[0].n_tr);
      HIPCHK(hipMemcpyAsync(rows.data(), p.dtrace, sizeof(double) * rows.size(), hipMemcpyDeviceToHost, s));
      HIPCHK(hipStreamSynchronize(s));
      for (int r = *n_tr; r < di[0].n_tr; ++r)
        for (int k = 0; k < 4; ++k) trace[4 * r + k] = rows[4 * r + k];
      *n_tr = di[0].n_tr;
    }
  }
};

// Drives the loops L[] (distinct plans) to the end.  Each chunk of K steps of
// every loop still running is captured once into ONE hipGraph: the first
// loop's stream is the capture's origin, the other loops' streams fork from
// it and join back, so each plan's launches form a branch of their own and
// one graph launch advances every instance of every grid.  The host polls the
// instances' done flags one chunk behind, so the device always has work
// queued; when a loop's instances have all stopped, the graph is captured
// again without its branch.  While a kernel is being timed the steps are
// launched directly, so that its events bracket each launch.
static void run_loops(const std::vector<DevLoop*>& L) {
  const int nl = (int)L.size();
  std::vector<hipStream_t> cs(nl);
  std::vector<hipEvent_t> evs;
  hipGraph_t graph[2] = {nullptr, nullptr};
  hipGraphExec_t gexec[2] = {nullptr, nullptr};
  auto drop_graphs = [&]() {
    for (int g = 0; g < 2; ++g) {
      if (gexec[g]) (void)hipGraphExecDestroy(gexec[g]);
      if (graph[g]) (void)hipGraphDestroy(graph[g]);
      gexec[g] = nullptr;
      graph[g] = nullptr;
    }
  };
  auto cleanup = [&]() {
    for (int i = 0; i < nl; ++i) {
      if (cs[i]) (void)hipStreamSynchronize(cs[i]);
      L[i]->pp->stream = L[i]->s;
    }
    drop_graphs();
    for (hipEvent_t e : evs) (void)hipEventDestroy(e);
  };
  try {
    auto mkev = [&]() {
      hipEvent_t e = nullptr;
      HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
      evs.push_back(e);
      return e;
    };
    long long max_steps = 0;
    for (int i = 0; i < nl; ++i) {
      cs[i] = L[i]->pp->cstream;
      max_steps = std::max(max_steps, L[i]->max_steps);
      // the loop's stream follows the work queued so far on its caller's stream
      const hipEvent_t e = mkev();
      HIPCHK(hipEventRecord(e, L[i]->s));
      HIPCHK(hipStreamWaitEvent(cs[i], e, 0));
    }
    const hipEvent_t ev[2] = {mkev(), mkev()};
    const hipEvent_t fork = mkev();
    std::vector<hipEvent_t> join(nl);
    for (int i = 0; i < nl; ++i) join[i] = mkev();
    for (int i = 0; i < nl; ++i) L[i]->pp->stream = cs[i];
    const hipStream_t c0 = cs[0];      // graphs launch here and the polls' events are recorded here
    std::vector<int> live(nl);
    for (int i = 0; i < nl; ++i) live[i] = i;

    // k steps of every live loop, forked from c0 and joined back to it
    auto issue = [&](int k_steps) {
      if (nl > 1) {
        HIPCHK(hipEventRecord(fork, c0));
        for (int i : live)
          if (i != 0) HIPCHK(hipStreamWaitEvent(cs[i], fork, 0));
      }
      for (int k = 0; k < k_steps; ++k)
        for (int i : live) L[i]->step();
      if (nl > 1)
        for (int i : live)
          if (i != 0) {
            HIPCHK(hipEventRecord(join[i], cs[i]));
            HIPCHK(hipStreamWaitEvent(c0, join[i], 0));
          }
    };
    auto capture = [&](int k_steps, int g) {
      HIPCHK(hipStreamBeginCapture(c0, hipStreamCaptureModeThreadLocal));
      try {
        issue(k_steps);
      } catch (...) {
        hipGraph_t g2 = nullptr;
        (void)hipStreamEndCapture(c0, &g2);
        if (g2) (void)hipGraphDestroy(g2);
        throw;
      }
      HIPCHK(hipStreamEndCapture(c0, &graph[g]));
      HIPCHK(hipGraphInstantiate(&gexec[g], graph[g], nullptr, nullptr, 0));
    };
    auto drain = [&]() {
      for (int i = 0; i < nl; ++i) HIPCHK(hipStreamSynchronize(cs[i]));
    };
    // steps per chunk: 4, and 16 once the loop has run K_LONG_AFTER steps (a
    // long solve then launches a quarter of the graphs; a short one wastes at
    // most two chunks of no-op steps after its last instance stops)
    const int K = 4, K_LONG = 16;
    const long long K_LONG_AFTER = 256;
    const bool use_graph = !prof_active();
    long long steps = 0;
    for (long long chunk = 0;; ++chunk) {
      const int g = (steps >= K_LONG_AFTER) ? 1 : 0;
      const int k_steps = g ? K_LONG : K;
      if (use_graph) {
        if (!gexec[g]) {                 // captured on idle streams
          drain();
          capture(k_steps, g);
        }
        HIPCHK(hipGraphLaunch(gexec[g], c0));
      } else {
        issue(k_steps);
      }
      steps += k_steps;
      HIPCHK(hipGetLastError());
      HIPCHK(hipEventRecord(ev[chunk & 1], c0));
      if (chunk > 0) {
        HIPCHK(hipEventSynchronize(ev[(chunk - 1) & 1]));
        std::vector<int> still;
        for (int i : live) {
          L[i]->log();
          if (L[i]->want_refill && L[i]->want_refill()) {
            // stopped slots take waiting solves: every stream drained first
            // (the graph stays valid: the same launches over the same slots)
            drain();
            L[i]->refill();
          }
          if (!L[i]->finished()) still.push_back(i);
        }
        if (still.empty()) break;
        if (still.size() != live.size()) {
          // the chunk just launched may still run: its executable graph is
          // destroyed only once every stream has drained
          live.swap(still);
          drain();
          drop_graphs();
        }
      }
      if (steps > max_steps) throw BecError(2, "internal: the device loop did not finish");
    }
    drain();
  } catch (...) {
    cleanup();
    throw;
  }
  cleanup();
}

// One plan's share of a solve phase: its slots (instances in the loop) and
// their scalar state.  Slot b solves from the state at src[b] with physics
// ph[b] and returns its state to dst[b] and its results to out[b];
// slots_begin sets slots up (normalised psi, initial energy and residual),
// slots_end writes them out.  A plain batch fills every slot once; a queue
// (more solves than slots) refills the slots whose solves have stopped.
struct Phase {
  Plan* pp = nullptr;
  int nb = 0;
  const bec_phys* php = nullptr;       // plain batch: physics of slot b at php[b]
  const bec_solver_cfg* cf = nullptr;
  double2* psi_io = nullptr;           // plain batch: states [nb][n], in and out
  bec_solve_out* out = nullptr;        // plain batch: results [nb]
  double* trace = nullptr;
  int trace_cap = 0;
  bool f32 = false;
  std::vector<Phys> ph;
  std::vector<Inst> I;
  std::vector<const double2*> src;
  std::vector<double2*> dst;
  std::vector<bec_solve_out*> outp;
  std::vector<double> t0;
  unsigned long long all = 0, lrmask = 0;
  bool lr_any = false, lp = false;
  int n_tr = 0;
};

// Checks and sizes a phase of nb slots; lr_any: some solve of the phase has a
// long-range kernel (the loop's steps then carry the convolutions).
static void phase_init(Phase& f, bool lr_any) {
  Plan& p = *f.pp;
  const int nb = f.nb;
  const bec_solver_cfg* cf = f.cf;
  if (nb < 1 || nb > BEC_BMAX) throw BecError(1, "batch size must be in [1, " + std::to_string(BEC_BMAX) + "]");
  if (cf->cg_beta < 0 || cf->cg_beta > 2) throw BecError(1, "cg_beta must be 'pr', 'pr_precond' or 'none'");
  if (nb > 1 && !plan_batchable(p))
    throw BecError(1, "batched solves need each instance's lines to fill whole workgroups, which this grid (N = " +
                          std::to_string(p.N) + ") does not: solve its instances one at a time");
  if (lr_any && !p.khat[0]) throw BecError(1, "kernel spectrum not built for this plan");
  plan_reserve(p, nb);
  solver_alloc(p, nb);
  f.all = low_bits(nb);
  f.ph.assign(nb, Phys{0.0, 0.0, 0.0, 0.0});
  f.I.assign(nb, Inst{});
  f.src.assign(nb, nullptr);
  f.dst.assign(nb, nullptr);
  f.outp.assign(nb, nullptr);
  f.t0.assign(nb, 0.0);
  f.lrmask = 0;
  f.lr_any = lr_any;
  f.n_tr = 0;
  // the polynomial line search (one step per iteration) unless the trial-per-step
  // search is asked for or the loop runs on the host
  f.lp = cf->line_poly != 0 && !cf->host_loop;
  if (f.f32) {
    if (lr_any && !f.lp) throw BecError(1, "internal: FP32 phase with a long-range kernel needs the polynomial search");
    solver_alloc32(p);
  }
}

// Sets up the slots of mask (ph, src, dst, outp filled by the caller):
// psi = psi0 / norm_of(psi0), the initial energy and residual
// (solvers.py:56-112); f32: their states are narrowed to complex64 for an
// FP32 phase.  The other slots' fields are untouched.
static void slots_begin(Phase& f, unsigned long long mask) {
  Plan& p = *f.pp;
  const int nb = f.nb;
  const bec_solver_cfg* cf = f.cf;
  const double dv = p.dv;
  const long long n = p.n;
  double2* PSI = p.st[S_PSI];
  double2* SD = p.st[S_SD];
  double2* G = p.st[S_G];
  double2* Y = p.st[S_Y];
  double2* D = p.st[S_D];
  double2* BEST = p.st[S_BEST];
  double* phi = p.stphi[0];     // long-range potential of psi (scaled)
  double* phit = p.stphi[1];    // unscaled convolution of the latest trial
  hipStream_t s = p.stream;
  const double floor_ = cf->precond_shift_min;
  const bool whole = mask == f.all;
  auto slots = [&](auto fn) {
    for (int b = 0; b < nb; ++b)
      if ((mask >> b) & 1ull) fn(b);
  };
  plan_set_phys(p, f.ph.data(), nb);
  slots([&](int b) {
    Inst& q = f.I[b];
    memset(&q, 0, sizeof(Inst));
    q.lr = plan_long_range(p, f.ph[b]);
    if (q.lr) f.lrmask |= 1ull << b;
    else f.lrmask &= ~(1ull << b);
    q.alpha = cf->step;
    q.sigma = floor_;
    q.stop = BEC_STOP_MAXIT;
  });
  if ((f.lrmask & mask) && !f.lr_any) throw BecError(2, "internal: a long-range slot in a phase without long range");
  const bool lr_any = f.lr_any;

  // psi = psi0 / norm_of(psi0)                                        (solvers.py:56-58)
  // the initial states in place [nb][n] when they are, else staged in SD
  bool in_place = true;
  slots([&](int b) { in_place = in_place && f.src[b] == f.psi_io + (long long)b * n; });
  const double2* in = f.psi_io;
  if (!in_place) {
    slots([&](int b) {
      HIPCHK(hipMemcpyAsync(SD + (long long)b * n, f.src[b], sizeof(double2) * n, hipMemcpyDeviceToDevice, s));
    });
    in = SD;
  }
  add_job(p, C_NORM, launch_norm2(in, n, nb, mask, hpart_of(p, C_NORM), s), 1, 0, R_NORM, Batch{nb, mask});
  plan_sync(p);
  std::vector<double> nrm0(nb, 1.0);
  slots([&](int b) { nrm0[b] = sqrt(hres(p, b)[R_NORM] * dv); });
  launch_scale(in, PSI, nrm0.data(), n, nb, s, mask);
  auto zero = [&](void* base, size_t el) {
    if (whole) {
      HIPCHK(hipMemsetAsync(base, 0, el * n * nb, s));
    } else {
      slots([&](int b) { HIPCHK(hipMemsetAsync((char*)base + el * n * b, 0, el * n, s)); });
    }
  };
  zero(SD, sizeof(double2));
  zero(Y, sizeof(double2));
  zero(D, sizeof(double2));
  zero(phi, sizeof(double));
  zero(phit, sizeof(double));

  // E, phi = energy_phi(psi)                                            (solvers.py:106)
  const Batch SL{nb, mask};
  run_energy_sums(p, SL, PSI, nullptr, nullptr, lr_any ? p.rho : nullptr);
  if (f.lrmask & mask) run_conv(p, 0, mask_batch(nb, f.lrmask & mask), p.rho, 1.0, phi, p.rho);
  HIPCHK(hipGetLastError());
  plan_sync(p);
  std::vector<double> mu_arr(nb, 0.0), phisc(nb, 0.0);
  const double now = now_s();
  slots([&](int b) {
    Inst& q = f.I[b];
    const Energy e0 = energy_from_sums(p, b, f.ph[b], q.lr, false, -f.ph[b].G_C);
    const double* rs = hres(p, b);
    q.L00 = e0.Ekin + e0.Etrap + e0.Erot;
    q.Spp = rs[R_EFIRST + 4] * dv;
    q.Sppp = rs[R_EFIRST + 5] * dv;
    q.Kpp = q.lr ? rs[R_EGRAV] * dv : 0.0;
    q.E = e0.E;
    q.mu = e0.E + e0.Econtact + 2.0 * e0.Ethree + e0.Egrav;
    q.E_ref = e0.E;
    q.n_energy = 1;
    mu_arr[b] = q.mu;
    phisc[b] = q.lr ? -f.ph[b].G_C : 0.0;
    f.t0[b] = now;
  });
  if (f.trace && f.trace_cap > 0 && nb == 1) {
    f.trace[0] = 0.0; f.trace[1] = 0.0; f.trace[2] = f.I[0].E; f.trace[3] = NAN;
    f.n_tr = 1;
  }
  // grad, mu, resid = residual(psi, phi); phi becomes the scaled potential  (solvers.py:112)
  {
    HArgs a;
    memset(&a, 0, sizeof(a));
    a.psi = PSI;
    a.local_on = true;
    a.phi = lr_any ? phi : nullptr;
    a.phisc = phisc.data();
    a.phi_out = lr_any ? phi : nullptr;
    a.mu = mu_arr.data();
    a.kf = 0.5;
    a.rf_omega = 1;
    a.out = G;
    a.reduce = true;
    run_H(p, SL, a);
  }
  HIPCHK(hipGetLastError());
  plan_sync(p);
  slots([&](int b) {
    f.I[b].resid_rel = sqrt(hres(p, b)[R_HLAST] * dv) / std::max(1.0, fabs(f.I[b].mu));
    f.I[b].best_resid = f.I[b].resid_rel;
  });
  launch_copy(PSI, BEST, n, nb, mask_act(mask), s);

  if (f.f32) {
    auto narrow = [&](int i) {
      if (whole) {
        launch_narrow(p.st[i], p.st32[i], n * nb, s);
      } else {
        slots([&](int b) { launch_narrow(p.st[i] + (long long)b * n, p.st32[i] + (long long)b * n, n, s); });
      }
    };
    narrow(S_PSI);
    narrow(S_G);
    narrow(S_BEST);
    auto zero32 = [&](int i) {
      if (whole) {
        HIPCHK(hipMemsetAsync(p.st32[i], 0, sizeof(float2) * n * nb, s));
      } else {
        slots([&](int b) { HIPCHK(hipMemsetAsync(p.st32[i] + (long long)b * n, 0, sizeof(float2) * n, s)); });
      }
    };
    zero32(S_SD);
    zero32(S_Y);
    zero32(S_D);
    HIPCHK(hipGetLastError());
  }
}

// A plain batch: every slot once, states in place at psi_io ([nb][n]).
static void phase_begin(Phase& f) {
  const long long n = f.pp->n;
  bool lr = false;
  for (int b = 0; b < f.nb; ++b)
    lr = lr || plan_long_range(*f.pp, Phys{f.php[b].beta2, f.php[b].beta3, f.php[b].Omega, f.php[b].G_C});
  phase_init(f, lr);
  for (int b = 0; b < f.nb; ++b) {
    f.ph[b] = Phys{f.php[b].beta2, f.php[b].beta3, f.php[b].Omega, f.php[b].G_C};
    f.src[b] = f.psi_io + (long long)b * n;
    f.dst[b] = f.psi_io + (long long)b * n;
    f.outp[b] = f.out + b;
  }
  slots_begin(f, f.all);
}

// The same loop with its scalar logic on the host (cfg host_loop): one
// synchronisation per iteration for the inner products and one per Armijo
// round.  Kept to check the device-driven loop's trajectory.
static void host_loop(Phase& f) {
  Plan& p = *f.pp;
  const int nb = f.nb;
  const bec_solver_cfg* cf = f.cf;
  std::vector<Inst>& I = f.I;
  const std::vector<Phys>& ph = f.ph;
  const double dv = p.dv;
  const long long n = p.n;
  double2* PSI = p.st[S_PSI];
  double2* SD = p.st[S_SD];
  double2* G = p.st[S_G];
  double2* Y = p.st[S_Y];
  double2* D = p.st[S_D];
  double2* Z = p.st[S_Z];
  double2* BEST = p.st[S_BEST];
  double* phi = p.stphi[0];
  double* phit = p.stphi[1];
  hipStream_t s = p.stream;
  const int maxit = cf->maxit;
  const double res_tol = cf->res_tol, etol = cf->energy_tol;
  const int check = std::max(1, cf->check);
  const double alpha_min = 1e-8, alpha_max = cf->step_max;
  const int ls_max = cf->linesearch_max;
  const int stall_max = std::max(1, cf->linesearch_stall);
  const double c1 = 1e-4, shrink = 0.5, growth = 1.1;
  const double bb_min = 0.05, bb_max = alpha_max, bb_mix = 0.5;
  const int restart_period = cf->cg_restart;
  const int conv_window = cf->conv_window;
  const double floor_ = cf->precond_shift_min;
  const unsigned long long lrmask = f.lrmask;
  const bool lr_any = f.lr_any;
  double* trace = f.trace;
  const int trace_cap = f.trace_cap;
  int& n_tr = f.n_tr;
  const double t0 = f.t0[0];
  std::vector<double> mu_arr(nb), phisc(nb), tarr(nb), divarr(nb);

int it = 0;
while (true) {
  unsigned long long run = 0;          // instances that execute iteration `it`
  for (int b = 0; b < nb; ++b)
    if (!I[b].done && it < maxit) run |= 1ull << b;
  if (run) {
    // sigma, z = _preconditioner(grad, psi, phi)                      (solvers.py:129)
    launch_sigma(p, nb, mask_act(run), PSI, lr_any ? phi : nullptr, part_of(p, C_HLAST), p.hlast_rows, 4, 2,
                 floor_, s);
    run_precond(p, mask_batch(nb, run), G, Z);
    add_job(p, C_DOTS, launch_dots(Z, PSI, D, G, Y, SD, n, nb, mask_act(run), nullptr, hpart_of(p, C_DOTS), s),
            16, 0, R_DOTS, mask_batch(nb, run));
    HIPCHK(hipGetLastError());
  }
  plan_sync(p);
  unsigned long long bestmask = 0;
  for (int b = 0; b < nb; ++b) {
    Inst& q = I[b];
    if (!q.pending) continue;
    // residual bookkeeping of the step accepted at iteration it_acc     (solvers.py:196-221)
    q.pending = false;
    q.resid_rel = sqrt(hres(p, b)[R_HLAST] * dv) / std::max(1.0, fabs(q.mu));
    if (trace && nb == 1 && n_tr < trace_cap) {
      double* r = trace + 4 * n_tr;
      r[0] = now_s() - t0; r[1] = q.it_acc + 1; r[2] = q.E; r[3] = q.resid_rel;
      ++n_tr;
    }
    if (q.resid_rel < q.best_resid) {
      q.best_resid = q.resid_rel;
      bestmask |= 1ull << b;
    }
    if (cf->verbose && q.it_acc % check == 0)
      printf("  [gs%s%d] it=%d E=%.8f res=%.2e best=%.2e step=%.2e\n", nb > 1 ? " " : "", nb > 1 ? b : 0,
             q.it_acc, q.E, q.resid_rel, q.best_resid, q.trial_acc);
    if (q.resid_rel < res_tol) {
      q.converged = true;
      q.stop = BEC_STOP_CONVERGED;
      q.done = true;
      continue;
    }
    if (q.it_acc - q.it_ref >= conv_window) {
      const double dE_win = fabs(q.E - q.E_ref) / std::max(1.0, fabs(q.E));
      if (dE_win < etol) {
        q.converged = true;
        q.stop = BEC_STOP_ENERGY_CONVERGED;
        q.done = true;
        continue;
      }
      q.E_ref = q.E;
      q.it_ref = q.it_acc;
    }
  }
  if (bestmask) launch_copy(PSI, BEST, n, nb, mask_act(bestmask), s);
  unsigned long long act = 0;
  for (int b = 0; b < nb; ++b) {
    if (I[b].done) continue;
    if (it >= maxit) { I[b].done = true; continue; }
    act |= 1ull << b;
  }
  if (!act) break;
  // per-instance CG direction and trial step                          (solvers.py:128-159)
  CombArgs cb;
  memset(&cb, 0, sizeof(cb));
  for (int b = 0; b < nb; ++b) {
    if (!((act >> b) & 1ull)) continue;
    Inst& q = I[b];
    q.n_iter = it + 1;
    const double* r = hres(p, b) + R_DOTS;
    q.sigma = hres(p, b)[R_SIGMA];
    const C2 cz = cm(r[0] * dv, r[1] * dv);        // <psi, z>
    const double pp_ = r[2] * dv;                  // <psi, psi>
    const C2 pd = cm(r[3] * dv, r[4] * dv);        // <psi, d>
    const C2 pg = cm(r[5] * dv, r[6] * dv);        // <psi, g>
    const double zg_raw = r[7] * dv, zy_raw = r[8] * dv;
    const C2 py = cm(r[9] * dv, r[10] * dv);       // <psi, y>
    const double gdr = r[11] * dv, gg = r[12] * dv, gy = r[13] * dv;
    const double sy = r[14] * dv, ss = r[15] * dv;
    const double zg = zg_raw - (conj(cz) * pg).re;   // Re <z_p, g>, z_p = z - psi <psi, z>
    const double zy = zy_raw - (conj(cz) * py).re;   // Re <z_p, y>
    if (it == 0) { q.zg_prev = zg; q.gg_prev = gg; }  // z_prev = z, grad_prev = grad (solvers.py:115,121)
    double beta;
    if (it == 0 || cf->cg_beta == BEC_CG_NONE || (restart_period > 0 && it % restart_period == 0)) {
      beta = 0.0;
    } else {
      if (cf->cg_beta == BEC_CG_PR_PRECOND) beta = std::max(0.0, zy / std::max(q.zg_prev, 1e-300));
      else beta = std::max(0.0, gy / std::max(q.gg_prev, 1e-300));
      if (!std::isfinite(beta)) beta = 0.0;
    }
    // dirc = proj_tangent(z_p + beta d), with the two descent fallbacks   (solvers.py:142-149)
    const C2 cw = cz - pp_ * cz + beta * pd;
    double az = 1.0, ad = beta, ag = 0.0;
    C2 gam = cm(0.0, 0.0) - (cz + cw);
    double gd = zg_raw + beta * gdr - (conj(pg) * (cz + cw)).re;
    if (gd <= 0) {
      az = 1.0; ad = 0.0; ag = 0.0; gam = cm(-cz.re, -cz.im);
      gd = zg_raw - (conj(pg) * cz).re;
    }
    if (gd <= 0) {
      az = 0.0; ad = 0.0; ag = 1.0; gam = cm(-pg.re, -pg.im);
      gd = gg - (conj(pg) * pg).re;
    }
    if (it > 0 && sy > 1e-30 && std::isfinite(sy)) {   // Barzilai-Borwein (solvers.py:150-158)
      double bb = ss / sy;
      if (std::isfinite(bb)) {
        bb = std::min(std::max(bb, bb_min), bb_max);
        q.alpha = (1.0 - bb_mix) * q.alpha + bb_mix * bb;
      }
    }
    q.alpha = std::min(std::max(q.alpha, alpha_min), alpha_max);
    if (cf->verbose >= 2)
      printf("  [gs%d] it=%d E=%.15e gd=%.3e beta=%.3e alpha=%.3e zg=%.3e zy=%.3e gg=%.3e sy=%.3e ss=%.3e pg=%.2e\n",
             b, it, q.E, gd, beta, q.alpha, zg, zy, gg, sy, ss, pg.re);
    q.gd = gd;
    q.trial = q.alpha;
    q.ls = 0;
    q.cz_re = cz.re;
    q.cz_im = cz.im;
    q.zg_cur = zg;           // committed to zg_prev / gg_prev on acceptance
    q.gg_cur = gg;
    cb.az[b] = az; cb.ad[b] = ad; cb.ag[b] = ag; cb.gre[b] = gam.re; cb.gim[b] = gam.im;
  }
  launch_combine(D, Z, G, PSI, cb, nullptr, n, nb, mask_act(act), s);
  // Armijo backtracking, one round per trial for all searching instances (solvers.py:160-174)
  unsigned long long T = act, acc = 0, fail = 0;
  while (T) {
    for (int b = 0; b < nb; ++b) tarr[b] = I[b].trial;
    const Batch BT = mask_batch(nb, T);
    run_energy_sums(p, BT, PSI, D, tarr.data(), lr_any ? p.rho : nullptr);
    if (T & lrmask) run_conv(p, 0, mask_batch(nb, T & lrmask), p.rho, 1.0, phit, p.rho);
    HIPCHK(hipGetLastError());
    plan_sync(p);
    for (int b = 0; b < nb; ++b) {
      if (!((T >> b) & 1ull)) continue;
      Inst& q = I[b];
      ++q.n_energy;
      q.et = energy_from_sums(p, b, ph[b], q.lr, true, -ph[b].G_C);
      if (std::isfinite(q.et.E) && q.et.E <= q.E - c1 * q.trial * q.gd) {
        acc |= 1ull << b;
        T &= ~(1ull << b);
        continue;
      }
      if (cf->verbose >= 2)
        printf("  [gs%d] it=%d reject trial=%.3e Et-E=%.3e need=%.3e\n", b, it, q.trial, q.et.E - q.E,
               -c1 * q.trial * q.gd);
      q.trial *= shrink;
      q.ls += 1;
      if (q.trial < alpha_min || q.ls >= ls_max) {
        fail |= 1ull << b;
        T &= ~(1ull << b);
      }
    }
  }
  if (fail) {                                                    // (solvers.py:175-185)
    CombArgs cf2;
    memset(&cf2, 0, sizeof(cf2));
    for (int b = 0; b < nb; ++b) {
      if (!((fail >> b) & 1ull)) continue;
      Inst& q = I[b];
      q.n_fail += 1;
      q.alpha = std::max(q.alpha * 0.5, alpha_min);
      cf2.az[b] = 1.0;                                           // direction = z_p
      cf2.gre[b] = -q.cz_re;
      cf2.gim[b] = -q.cz_im;
      if (q.alpha <= alpha_min || q.n_fail >= stall_max) {
        q.stop = BEC_STOP_LINE_SEARCH_STALLED;
        q.done = true;
        if (cf->verbose)
          printf("[gs] line search stalled it=%d: E=%.10g res=%.3e best_res=%.3e\n", it, q.E, q.resid_rel,
                 q.best_resid);
      }
    }
    launch_combine(D, Z, nullptr, PSI, cf2, nullptr, n, nb, mask_act(fail), s);
  }
  if (acc) {                                                     // (solvers.py:186-196)
    for (int b = 0; b < nb; ++b) {
      if (!((acc >> b) & 1ull)) continue;
      Inst& q = I[b];
      q.n_fail = 0;
      q.zg_prev = q.zg_cur;
      q.gg_prev = q.gg_cur;
      q.E = q.et.E;
      q.mu = q.et.E + q.et.Econtact + 2.0 * q.et.Ethree + q.et.Egrav;
      q.alpha = std::min(alpha_max, q.trial * growth);
      const double inv2 = 1.0 / q.et.nrm2;
      tarr[b] = q.trial;
      divarr[b] = sqrt(q.et.nrm2);
      mu_arr[b] = q.mu;
      phisc[b] = q.lr ? -ph[b].G_C * inv2 : 0.0;
      q.pending = true;
      q.it_acc = it;
      q.trial_acc = q.trial;
    }
    // psi <- (psi - trial d)/nrm in place (s = change), grad <- H psi - mu psi (y = change)
    HArgs a;
    memset(&a, 0, sizeof(a));
    a.psi = PSI;
    a.d = D;
    a.t = tarr.data();
    a.div = divarr.data();
    a.psi_w = PSI;
    a.sdiff = SD;
    a.local_on = true;
    a.phi = lr_any ? phit : nullptr;
    a.phisc = phisc.data();
    a.phi_out = lr_any ? phi : nullptr;
    a.mu = mu_arr.data();
    a.kf = 0.5;
    a.rf_omega = 1;
    a.out = G;
    a.ydiff = Y;
    a.reduce = true;
    run_H(p, mask_batch(nb, acc), a);
    HIPCHK(hipGetLastError());
  }
  ++it;
}
}

// The slots of mask written out: the returned iterate (solvers.py:225-226)
// to dst[b] and the results to outp[b]; an FP32 phase's states are widened
// back first.
static void slots_end(Phase& f, unsigned long long mask) {
  Plan& p = *f.pp;
  const int nb = f.nb;
  const long long n = p.n;
  hipStream_t s = p.stream;
  double2* PSI = p.st[S_PSI];
  double2* BEST = p.st[S_BEST];
  const bool whole = mask == f.all;
  auto slots = [&](auto fn) {
    for (int b = 0; b < nb; ++b)
      if ((mask >> b) & 1ull) fn(b);
  };
  if (f.f32) {
    for (int i : {S_PSI, S_BEST}) {
      if (whole) launch_widen(p.st32[i], p.st[i], n * nb, s);
      else slots([&](int b) { launch_widen(p.st32[i] + (long long)b * n, p.st[i] + (long long)b * n, n, s); });
    }
    HIPCHK(hipGetLastError());
  }
  // return the lowest-residual iterate seen                             (solvers.py:225-226)
  unsigned long long from_best = 0;
  slots([&](int b) {
    if (f.I[b].best_resid < f.I[b].resid_rel) from_best |= 1ull << b;
  });
  bool in_place = true;
  slots([&](int b) { in_place = in_place && f.dst[b] == f.psi_io + (long long)b * n; });
  if (in_place) {
    launch_copy(BEST, f.psi_io, n, nb, mask_act(from_best), s);
    launch_copy(PSI, f.psi_io, n, nb, mask_act(mask & ~from_best), s);
  } else {
    slots([&](int b) {
      const double2* from = (((from_best >> b) & 1ull) ? BEST : PSI) + (long long)b * n;
      HIPCHK(hipMemcpyAsync(f.dst[b], from, sizeof(double2) * n, hipMemcpyDeviceToDevice, s));
    });
  }
  plan_sync(p);
  const double now = now_s();
  slots([&](int b) {
    bec_solve_out& o = *f.outp[b];
    o.iters = f.I[b].n_iter;
    o.converged = f.I[b].converged ? 1 : 0;
    o.stop_reason = f.I[b].stop;
    o.n_trace = (nb == 1) ? f.n_tr : 0;
    o.E = f.I[b].E;
    o.resid_rel = f.I[b].resid_rel;
    o.best_resid = f.I[b].best_resid;
    o.sigma = f.I[b].sigma;
    o.walltime = now - f.t0[b];
    o.n_energy = f.I[b].n_energy;
  });
}

static void phase_end(Phase& f) { slots_end(f, f.all); }

// One phase of the solve over the plans of F (distinct): each plan's set-up,
// then their loops together (run_loops; the host-driven loop for a single
// FP64 phase that asks for it), then each plan's outputs.
static void run_phases(const std::vector<Phase*>& F) {
  for (Phase* f : F) phase_begin(*f);
  if (F.size() == 1 && !F[0]->f32 && F[0]->cf->host_loop) {
    host_loop(*F[0]);
  } else {
    std::vector<DevLoop> L(F.size());
    std::vector<DevLoop*> lp(F.size());
    try {
      for (size_t i = 0; i < F.size(); ++i) {
        Phase& f = *F[i];
        if (f.f32) f.pp->f32 = 1;
        L[i].begin(*f.pp, f.nb, f.cf, f.I, f.lr_any, (!f.f32 && f.trace && f.nb == 1) ? f.trace : nullptr,
                   f.trace_cap, f.n_tr, f.lp);
        lp[i] = &L[i];
      }
      run_loops(lp);
      for (DevLoop* l : lp) l->end();
    } catch (...) {
      for (Phase* f : F) f->pp->f32 = 0;
      throw;
    }
    for (Phase* f : F) f->pp->f32 = 0;
  }
  for (Phase* f : F) phase_end(*f);
}

// Whether a plan's instances take the FP32 warm phase: warm_res_tol above
// res_tol, transforms without the radices 7, 11 and 13 (built for FP64 only),
// and with long range the polynomial search (its convolutions, sums and test
// stay FP64).
static bool warm_applies(const Plan& p, int nb, const bec_phys* php, const bec_solver_cfg* cf) {
  bool lr = false;
  for (int b = 0; b < nb && b < BEC_BMAX; ++b)
    lr = lr || plan_long_range(p, Phys{php[b].beta2, php[b].beta3, php[b].Omega, php[b].G_C});
  const bool lp = cf->line_poly != 0 && !cf->host_loop;
  return cf->warm_res_tol > cf->res_tol && !(lr && !lp) && !p.fpN.odd;
}

// The solves of ng plans (distinct grids; nb[g] instances each, physics
// php and states psi[g] laid out group after group): the FP32 warm phase of
// the plans it applies to, then the FP64 loop of all of them, which restarts
// from the warm phase's returned states.  trace only for one instance.
static int solve_groups(int ng, Plan* const* P, const int* nbs, const bec_phys* php,
                        const bec_solver_cfg* const* cfs, double2* const* psi, bec_solve_out* out, double* trace,
                        int trace_cap) {
  std::vector<int> off(ng + 1, 0);
  for (int g = 0; g < ng; ++g) off[g + 1] = off[g] + nbs[g];
  std::vector<bec_solver_cfg> wcf(ng);
  std::vector<bec_solve_out> ow(off[ng]);
  std::vector<Phase> warm(ng), main(ng);
  std::vector<Phase*> fw, fm;
  for (int g = 0; g < ng; ++g) {
    Phase& m = main[g];
    m.pp = P[g];
    m.nb = nbs[g];
    m.php = php + off[g];
    m.cf = cfs[g];
    m.psi_io = psi[g];
    m.out = out + off[g];
    m.trace = ng == 1 ? trace : nullptr;
    m.trace_cap = ng == 1 ? trace_cap : 0;
    fm.push_back(&m);
    if (!warm_applies(*P[g], nbs[g], m.php, cfs[g])) continue;
    wcf[g] = *cfs[g];
    wcf[g].res_tol = cfs[g]->warm_res_tol;
    Phase& w = warm[g];
    w = m;
    w.cf = &wcf[g];
    w.out = ow.data() + off[g];
    w.trace = nullptr;
    w.trace_cap = 0;
    w.f32 = true;
    fw.push_back(&w);
  }
  if (!fw.empty()) run_phases(fw);
  run_phases(fm);
  for (int g = 0; g < ng; ++g) {
    const bool warmed = !warm[g].I.empty();
    for (int b = off[g]; b < off[g + 1]; ++b) {
      out[b].warm_iters = warmed ? ow[b].iters : 0;
      if (!warmed) continue;
      out[b].iters += ow[b].iters;
      out[b].walltime += ow[b].walltime;
      out[b].n_energy += ow[b].n_energy;
    }
  }
  return 0;
}

// One phase of a queue of nq solves (states psi_io [nq][n], physics php[nq],
// results out[nq]) on nb slots of plan p: the first nb solves fill the slots;
// whenever at least an eighth of the slots (or all occupied ones) have
// stopped and solves are waiting, the loop drains, writes the stopped slots
// out and sets the waiting solves up in them, and goes on with the same
// graph.  Every solve follows the ground_state loop exactly, as in a batch.
static void queue_phase(Plan& p, int nb, int nq, const bec_phys* php, const bec_solver_cfg* cf, double2* psi_io,
                        bec_solve_out* out, bool f32) {
  bool lr = false;
  for (int k = 0; k < nq; ++k)
    lr = lr || plan_long_range(p, Phys{php[k].beta2, php[k].beta3, php[k].Omega, php[k].G_C});
  Phase f;
  f.pp = &p;
  f.nb = nb;
  f.cf = cf;
  f.f32 = f32;
  phase_init(f, lr);
  const long long n = p.n;
  int next = 0;
  auto assign = [&](int b) {
    const int k = next++;
    f.ph[b] = Phys{php[k].beta2, php[k].beta3, php[k].Omega, php[k].G_C};
    f.src[b] = psi_io + (long long)k * n;
    f.dst[b] = psi_io + (long long)k * n;
    f.outp[b] = out + k;
  };
  unsigned long long occ = 0;          // slots holding a solve not yet written out
  for (int b = 0; b < nb && next < nq; ++b) {
    assign(b);
    occ |= 1ull << b;
  }
  slots_begin(f, occ);
  DevLoop L;
  try {
    if (f32) p.f32 = 1;
    L.begin(p, nb, cf, f.I, f.lr_any, nullptr, 0, f.n_tr, f.lp);
    L.pending = [&]() { return next < nq; };
    L.want_refill = [&]() {
      if (next >= nq) return false;
      const unsigned long long dn = L.done_mask() & occ;
      const int ready = __builtin_popcountll(dn), busy = __builtin_popcountll(occ);
      return ready > 0 && (ready >= std::max(1, nb / 8) || ready == busy);
    };
    L.refill = [&]() {
      const unsigned long long dn = L.done_mask() & occ;
      L.harvest(dn, p.stream);
      slots_end(f, dn);                // (FP64 set-up and write-out: the launches' own
      occ &= ~dn;                      //  batches carry the precision, not p.f32)
      unsigned long long fresh = 0;
      for (int b = 0; b < nb && next < nq; ++b)
        if ((dn >> b) & 1ull) {
          assign(b);
          fresh |= 1ull << b;
        }
      if (fresh) {
        slots_begin(f, fresh);
        L.upload(fresh, p.stream);
        occ |= fresh;
      }
    };
    run_loops({&L});
    L.harvest(occ, p.stream);
  } catch (...) {
    p.f32 = 0;
    throw;
  }
  p.f32 = 0;
  slots_end(f, occ);
}

// The queue of nq solves on nb slots: the FP32 warm phase over the whole
// queue when it applies (as for a batch), then the FP64 phase over it.
static int solve_queue(Plan& p, int nb, int nq, const bec_phys* php, const bec_solver_cfg* cf, double2* psi_io,
                       bec_solve_out* out) {
  bool lr = false;
  for (int k = 0; k < nq; ++k)
    lr = lr || plan_long_range(p, Phys{php[k].beta2, php[k].beta3, php[k].Omega, php[k].G_C});
  const bool lp = cf->line_poly != 0 && !cf->host_loop;
  const bool warm = cf->warm_res_tol > cf->res_tol && !(lr && !lp) && !p.fpN.odd;
  std::vector<bec_solve_out> ow(nq);
  if (warm) {
    bec_solver_cfg w = *cf;
    w.res_tol = cf->warm_res_tol;
    queue_phase(p, nb, nq, php, &w, psi_io, ow.data(), true);
  }
  queue_phase(p, nb, nq, php, cf, psi_io, out, false);
  for (int k = 0; k < nq; ++k) {
    out[k].warm_iters = warm ? ow[k].iters : 0;
    if (!warm) continue;
    out[k].iters += ow[k].iters;
    out[k].walltime += ow[k].walltime;
    out[k].n_energy += ow[k].n_energy;
  }
  return 0;
}

static int solve_batch(Plan& p, int nb, const bec_phys* php, const bec_solver_cfg* cf, double2* psi_io,
                       bec_solve_out* out, double* trace, int trace_cap) {
  Plan* P = &p;
  return solve_groups(1, &P, &nb, php, &cf, &psi_io, out, trace, trace_cap);
}


extern "C" int bec_ground_state(bec_plan* h, const bec_phys* php, const bec_solver_cfg* cf, void* psi_io,
                                bec_solve_out* out, double* trace, int trace_cap) {
  try {
    if (!h || !php || !cf || !psi_io || !out) throw BecError(1, "null argument");
    return solve_batch(h->p, 1, php, cf, (double2*)psi_io, out, trace, trace_cap);
  } catch (const BecError& e) {
    bec_set_error(e.what());
    return e.code;
  } catch (const std::exception& e) {
    bec_set_error(e.what());
    return 2;
  }
}

// Reducing line passes need whole workgroups per instance: one instance's
// lines (n / N) must be a multiple of the smallest workgroup of both lengths.
bool plan_batchable(const Plan& p) {
  const int nlk = (int)(p.n / p.N);
  return nlk % lpw_min(p.lcN.tpl) == 0 && nlk % lpw_min(p.lcM.tpl) == 0;
}

extern "C" int bec_plan_max_batch(bec_plan* h, int* nb_max) {
  try {
    if (!h || !nb_max) throw BecError(1, "null argument");
    *nb_max = plan_batchable(h->p) ? BEC_BMAX : 1;
    return 0;
  } catch (const BecError& e) {
    bec_set_error(e.what());
    return e.code;
  }
}

extern "C" int bec_plan_reserve(bec_plan* h, int nb) {
  try {
    if (!h) throw BecError(1, "null plan");
    Plan& p = h->p;
    plan_reserve(p, nb);
    solver_alloc(p, nb);
    solver_alloc32(p);
    dev_alloc(p, nb, 0);
    HIPCHK(hipStreamSynchronize(p.stream));
    return 0;
  } catch (const BecError& e) {
    bec_set_error(e.what());
    return e.code;
  } catch (const std::exception& e) {
    bec_set_error(e.what());
    return 2;
  }
}

// Autotuner: times the line passes of one loop step -- H psi on every axis,
// the energies, the preconditioner and (with a kernel spectrum loaded) the
// convolution -- over nb instances for each candidate schedule, and keeps the
// fastest: the stage order of each transform length and a power-of-two factor
// on the lines per workgroup of contiguous and of strided passes.
extern "C" int bec_plan_autotune(bec_plan* h, int nb, double* ms_best, double* ms_default, char* report,
                                 int report_cap) {
  try {
    if (!h) throw BecError(1, "null plan");
    Plan& p = h->p;
    plan_reserve(p, nb);
    solver_alloc(p, nb);
    hipStream_t s = p.stream;
    const long long n = p.n;
    double2* PSI = p.st[S_PSI];
    double2* G = p.st[S_G];
    double2* Z = p.st[S_Z];
    // a smooth, finite state for every instance: psi = 1, rho = 1
    std::vector<double2> one(n, make_double2(1.0, 0.0));
    for (int b = 0; b < nb; ++b)
      HIPCHK(hipMemcpyAsync(PSI + (size_t)b * n, one.data(), sizeof(double2) * n, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemsetAsync(p.rho, 0, sizeof(double) * n * nb, s));
    HIPCHK(hipStreamSynchronize(s));
    std::vector<Phys> ph(nb, Phys{1.0, 0.0, 0.5, 0.0});
    plan_set_phys(p, ph.data(), nb);
    std::vector<double> mu(nb, 0.0);
    const unsigned long long all = (nb >= 64) ? ~0ull : ((1ull << nb) - 1ull);
    const Batch ALL{nb, all};
    const bool conv = p.khat[0] != nullptr;
    auto work = [&]() {
      HArgs a;
      memset(&a, 0, sizeof(a));
      a.psi = PSI;
      a.local_on = true;
      a.mu = mu.data();
      a.kf = 0.5;
      a.rf_omega = 1;
      a.out = G;
      run_H(p, ALL, a);
      run_precond(p, ALL, G, Z, nullptr);
      run_energy_sums(p, ALL, PSI, nullptr, nullptr, conv ? p.rho : nullptr);
      if (conv) run_conv(p, 0, ALL, p.rho, 1.0, p.loc, nullptr);
    };
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    auto timed = [&]() {
      work();                                  // warm-up
      plan_sync(p);                            // fold (and drop) the queued host reductions
      double tot = 0.0;
      for (int r = 0; r < 3; ++r) {
        HIPCHK(hipEventRecord(e0, s));
        work();
        HIPCHK(hipEventRecord(e1, s));
        HIPCHK(hipEventSynchronize(e1));
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, e0, e1));
        tot += ms;
        plan_sync(p);
      }
      return tot / 3.0;
    };
    const auto oN = stage_orders(p.fpN, 4);
    const auto oM = conv ? stage_orders(p.fpM, 3) : std::vector<std::vector<int>>{{}};
    double best = 1e300, base = -1.0;
    std::vector<int> bN, bM;
    int bl[2] = {0, 0};
    for (const auto& rn : oN)
      for (const auto& rm : oM)
        for (int lc = -1; lc <= 1; ++lc)
          for (int ls = -1; ls <= 1; ++ls) {
            set_stage_order(p.fpN, rn);
            if (conv) set_stage_order(p.fpM, rm);
            p.lsh[0] = lc;
            p.lsh[1] = ls;
            const double t = timed();
            if (rn == oN[0] && (!conv || rm == oM[0]) && lc == 0 && ls == 0) base = t;
            if (t < best) {
              best = t;
              bN = rn;
              bM = rm;
              bl[0] = lc;
              bl[1] = ls;
            }
          }
    set_stage_order(p.fpN, bN);
    if (conv) set_stage_order(p.fpM, bM);
    p.lsh[0] = bl[0];
    p.lsh[1] = bl[1];
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (ms_best) *ms_best = best;
    if (ms_default) *ms_default = base;
    if (report && report_cap > 0) {
      std::string r = "N=" + std::to_string(p.N) + " stages [";
      for (size_t i = 0; i < bN.size(); ++i) r += (i ? "," : "") + std::to_string(bN[i]);
      r += "]";
      if (conv) {
        r += " M=" + std::to_string(p.M) + " stages [";
        for (size_t i = 0; i < bM.size(); ++i) r += (i ? "," : "") + std::to_string(bM[i]);
        r += "]";
      }
      r += " lines/workgroup x2^" + std::to_string(bl[0]) + " contiguous, x2^" + std::to_string(bl[1]) + " strided";
      snprintf(report, (size_t)report_cap, "%s", r.c_str());
    }
    return 0;
  } catch (const BecError& e) {
    bec_set_error(e.what());
    return e.code;
  } catch (const std::exception& e) {
    bec_set_error(e.what());
    return 2;
  }
}

extern "C" int bec_ground_state_batch(bec_plan* h, int nb, const bec_phys* php, const bec_solver_cfg* cf,
                                      void* psi_io, bec_solve_out* out) {
  try {
    if (!h || !php || !cf || !psi_io || !out) throw BecError(1, "null argument");
    return solve_batch(h->p, nb, php, cf, (double2*)psi_io, out, nullptr, 0);
  } catch (const BecError& e) {
    bec_set_error(e.what());
    return e.code;
  } catch (const std::exception& e) {
    bec_set_error(e.what());
    return 2;
  }
}

extern "C" int bec_ground_state_uneven(int ng, bec_plan* const* plans, const int* nb, const bec_phys* php,
                                       const bec_solver_cfg* const* cfg, void* const* psi_io, bec_solve_out* out) {
  try {
    if (!plans || !nb || !php || !cfg || !psi_io || !out) throw BecError(1, "null argument");
    if (ng < 1 || ng > BEC_GMAX) throw BecError(1, "number of grids must be in [1, " + std::to_string(BEC_GMAX) + "]");
    std::vector<Plan*> P(ng);
    std::vector<double2*> psi(ng);
    for (int g = 0; g < ng; ++g) {
      if (!plans[g] || !cfg[g] || !psi_io[g]) throw BecError(1, "null argument");
      if (nb[g] < 1 || nb[g] > BEC_BMAX)
        throw BecError(1, "batch size must be in [1, " + std::to_string(BEC_BMAX) + "]");
      if (cfg[g]->host_loop) throw BecError(1, "solves on several grids at once run the device-driven loop only");
      P[g] = &plans[g]->p;
      psi[g] = (double2*)psi_io[g];
      for (int h = 0; h < g; ++h)
        if (P[h] == P[g]) throw BecError(1, "each grid of one call needs a plan of its own");
    }
    return solve_groups(ng, P.data(), nb, php, cfg, psi.data(), out, nullptr, 0);
  } catch (const BecError& e) {
    bec_set_error(e.what());
    return e.code;
  } catch (const std::exception& e) {
    bec_set_error(e.what());
    return 2;
  }
}

extern "C" int bec_ground_state_queue(bec_plan* h, int nb, int nq, const bec_phys* php, const bec_solver_cfg* cf,
                                      void* psi_io, bec_solve_out* out) {
  try {
    if (!h || !php || !cf || !psi_io || !out) throw BecError(1, "null argument");
    if (nb < 1 || nb > BEC_BMAX) throw BecError(1, "slots must be in [1, " + std::to_string(BEC_BMAX) + "]");
    if (nq < 1) throw BecError(1, "the queue needs at least one solve");
    if (cf->host_loop) throw BecError(1, "a queue of solves runs the device-driven loop only");
    return solve_queue(h->p, std::min(nb, nq), nq, php, cf, (double2*)psi_io, out);
  } catch (const BecError& e) {
    bec_set_error(e.what());
    return e.code;
  } catch (const std::exception& e) {
    bec_set_error(e.what());
    return 2;
  }
}
