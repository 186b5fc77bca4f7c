/* C-ABI of libbecgpp_hip.so: the MI355X (gfx950) FP64 ground-state path of
 * becGPP.  Plain pointers and sizes only.
 *
 * Conventions
 *  - Every function returns 0 on success, 1 for an invalid argument
 *    (the Python layer raises ValueError), 2 for a HIP/runtime failure
 *    (RuntimeError); bec_last_error() gives the message of the last failure
 *    on the calling thread.
 *  - Field pointers are DEVICE pointers to C-contiguous arrays in the
 *    reference's layout (indexing "ij", axis 0 = x): complex fields are
 *    interleaved complex128 (double re, double im), real fields float64, with
 *    N^ndim points.  The caller owns them; the plan owns its workspace.
 *  - Work is ordered on the plan's stream (bec_plan_set_stream, default 0).
 *    Functions that return host scalars synchronise that stream.  A plan is
 *    not re-entrant: one thread per plan, or external locking.
 *
 * The reference (becGPP 1.2.1) is pure Python; each entry point names the
 * Python function whose arithmetic it replaces (file:line in becgpp/).
 */
#ifndef BECGPP_HIP_H
#define BECGPP_HIP_H
#ifdef __cplusplus
extern "C" {
#endif

typedef struct bec_plan bec_plan;

/* kernel kinds (grid.py:23-55 resolve_kernel) */
#define BEC_KERNEL_NONE 0
#define BEC_KERNEL_NEWTON 1
#define BEC_KERNEL_LOG 2
#define BEC_KERNEL_TABLE 3   /* real-space table supplied by the caller (q2d, W) */

/* which spectrum: the interaction kernel K or the q2d virial kernel W
 * (interactions.py:112 _build_kernel_fft(G, which)) */
#define BEC_WHICH_K 0
#define BEC_WHICH_W 1

/* cg_beta variants (solvers.py:75-82) */
#define BEC_CG_PR_PRECOND 0
#define BEC_CG_PR 1
#define BEC_CG_NONE 2

/* stop reasons (solvers.py:117,180,207,215) */
#define BEC_STOP_MAXIT 0
#define BEC_STOP_CONVERGED 1
#define BEC_STOP_ENERGY_CONVERGED 2
#define BEC_STOP_LINE_SEARCH_STALLED 3

typedef struct bec_phys {
  double beta2, beta3, Omega, G_C;
} bec_phys;

/* solver knobs, one field per CFG key read by solvers.py:59-88 */
typedef struct bec_solver_cfg {
  int maxit;
  double res_tol;
  double energy_tol;
  int conv_window;
  double step;
  double step_max;
  int linesearch_max;
  int linesearch_stall;
  double precond_shift_min;
  int cg_beta;
  int cg_restart;
  int verbose;
  int check;
  int host_loop;   /* 0: the loop's scalar logic runs in controller kernels on the
                      device (default); 1: on the host, one synchronisation per
                      iteration and Armijo trial (no reference counterpart) */
  double warm_res_tol;  /* > res_tol: an FP32 warm phase first runs the loop on
                           complex64 fields to this relative residual, then the
                           FP64 loop polishes from its best state; with long
                           range this needs line_poly (the convolutions, sums
                           and the Armijo test stay FP64); 0 = FP64 only (no
                           reference counterpart) */
  int line_poly;        /* 1: the device loop prices every Armijo trial of an
                           iteration from one pass of line coefficients (the
                           energy along psi - t d is a polynomial in t over a
                           polynomial norm), so each step completes an
                           iteration of every instance; 0: one trial per step
                           (no reference counterpart; the trials, their order
                           and the acceptance test are the reference's) */
} bec_solver_cfg;

typedef struct bec_solve_out {
  int iters;            /* iterations performed (solvers.py:128) */
  int converged;        /* loop flag before the final residual (solvers.py:206,214) */
  int stop_reason;      /* BEC_STOP_* */
  int n_trace;          /* trace rows written */
  double E;             /* energy of the last accepted iterate */
  double resid_rel;     /* relative residual of the last accepted iterate */
  double best_resid;    /* lowest relative residual seen (solvers.py:199-201) */
  double sigma;         /* last Sobolev shift (solvers.py:232) */
  double walltime;      /* seconds, from after the initial energy (solvers.py:107) */
  long long n_energy;   /* energy evaluations (line-search trials + initial) */
  int warm_iters;       /* of iters, those of the FP32 warm phase */
} bec_solve_out;

const char* bec_last_error(void);

/* Plan: grid geometry, FFT twiddles, workspaces, kernel spectra.
 * Replaces make_grid (grid.py:69-101) for the device side: x_i = (i-N/2) dx,
 * dx = 2L/N, k = 2 pi fftfreq(N, dx), V = trap_coeff r^s (0 if s = 0),
 * padded length M = pad N.  N <= 1024 must factor into 2^a 3^b 5^c 7^d
 * 11^e 13^f; the long-range kernels also need pad N <= 1024 (a plan with a
 * larger pad N solves without long range, and bec_plan_build_kernel /
 * bec_plan_set_kernel_table refuse it); ndim is 2 or 3. */
int bec_plan_create(int ndim, int N, double L, int pad, double s, double trap_coeff, bec_plan** out);
int bec_plan_destroy(bec_plan* p);
int bec_plan_set_stream(bec_plan* p, void* hip_stream);

/* Kernel spectrum of the free-space convolution (interactions.py:112-151):
 * cell-averaged kernel on the wrap-around displacement grid, exact self cell,
 * FFT, stored as the real-even half spectrum [(M/2+1)^ndim].
 * bec_plan_build_kernel builds newton (2D/3D) or log (2D) on the device;
 * bec_plan_set_kernel_table takes a caller-built real-space table [M^ndim]
 * (device pointer; q2d K_eff and its virial kernel W, interactions.py:64-94). */
int bec_plan_build_kernel(bec_plan* p, int which, int kind);
int bec_plan_set_kernel_table(bec_plan* p, int which, const double* ker_real);
/* clear_kernel_cache (interactions.py:174-176) */
int bec_plan_clear_kernels(bec_plan* p);

/* sum |psi|^2 over the grid (norm_of, grid.py:108-109, before dV and sqrt) */
int bec_norm2(bec_plan* p, const void* psi, double* sum_out);

/* out = scale * dV * (K * rho)_crop  (kernel_convolution, interactions.py:154-164,
 * with scale = 1; long_range_phi, interactions.py:167-171, with scale = -G_C). */
int bec_conv(bec_plan* p, int which, const double* rho, double scale, double* out);

/* energy_components (operators.py:10-29): comps[8] = E, Ekin, Etrap, Econtact,
 * Ethree, Erot, Egrav, Lz (host); phi_out (device, nullable) receives the
 * long-range potential (zeros when the kernel is off). */
int bec_energy(bec_plan* p, const void* psi, const bec_phys* ph, double* comps, double* phi_out);

/* apply_H (operators.py:36-45); phi (nullable) reuses a long-range potential,
 * otherwise it is computed when the kernel is on. */
int bec_apply_H(bec_plan* p, const void* psi, const bec_phys* ph, const double* phi, void* out);

/* Lz_op (grid.py:112-116): out = -i (x d_y - y d_x) psi */
int bec_Lz(bec_plan* p, const void* psi, void* out);

/* compute_residual (operators.py:48-52): out3 = resid_abs, mu, resid_rel */
int bec_residual(bec_plan* p, const void* psi, const bec_phys* ph, double* out3);

/* _preconditioner (solvers.py:28-45): out = proj_tangent(IFFT(FFT(g)/(k^2/2+sigma)));
 * sigma_out (host) = max(lower median of the bulk local potential, shift_min). */
int bec_precondition(bec_plan* p, const void* g, const void* psi, const bec_phys* ph,
                     const double* phi, double shift_min, void* out, double* sigma_out);

/* moments for observables (operators.py:55-103): out5 = sum rho, sum r^2 rho,
 * sum (x^2+y^2) rho, sum z^2 rho, max rho (host, before dV). */
int bec_moments(bec_plan* p, const void* psi, double* out5);

/* Weight of a 2D state in the lowest Landau level (vortex.py:61-84, lll_weight):
 * sum over m = 0..Mcap of |<q_m,psi>|^2 / (<q_m,q_m><psi,psi>) with
 * q_m = r^m exp(-r^2/2) exp(i m theta) / sqrt(pi m!), theta = sign(Omega)
 * atan2(y, x), Mcap = min(max(Mmax, ceil(<r^2> + 10 sqrt(max(<r^2>,1)) + 30)),
 * max(Mmax, floor((0.8 L)^2))); clamped to [0, 1].  mcap_out (nullable). */
int bec_lll_weight(bec_plan* p, const void* psi, double Omega, int Mmax, double* w_out, int* mcap_out);

/* Resampling of a state between grids (grid.py:140-161, resample_state):
 * multilinear interpolation of the n0^ndim grid with spacing dx0 (nodes
 * (j - n0/2) dx0) at the nodes of the n1^ndim grid with spacing dx1, zero
 * outside the old grid's span (torch grid_sample, align_corners=True, zero
 * padding), then divided by its norm sqrt(sum |psi|^2 dx1^ndim).  norm_out
 * receives that norm before the division; the state is left undivided (and
 * the caller reports no state) when the norm is not finite or <= 1e-14, or
 * when n0 = 1 (norm 0).  Needs no plan; runs on `stream` (a hipStream_t,
 * null for the default stream) and synchronises it. */
int bec_resample(int ndim, int n0, double dx0, const void* psi_old, int n1, double dx1, void* psi_new,
                 double* norm_out, void* stream);

/* Batched diagnostics (the diagnostic bundle of a sweep's states in one set
 * of launches; no reference counterpart beyond the per-state functions they
 * batch).  psi holds nb <= 64 states of the plan's grid one after the other.
 *  - bec_observe_batch (operators.py:55-103 and observables' inputs):
 *    out[16 b + 0..7] = the energy components as bec_energy with ph[b],
 *    [8..12] = the moments of bec_moments, [13] = sum rho (W*rho) dV with the
 *    q2d virial kernel W (when the plan holds it and G_C != 0, else NaN);
 *  - bec_lll_weight_batch (vortex.py:61-84): w_out[b] as bec_lll_weight with
 *    Omega[b];
 *  - bec_vortex_count_batch (vortex.py:10-48): out4[4 b..] as
 *    bec_vortex_count with xcm, ycm, R, ell = prm4[4 b .. 4 b + 3]. */
int bec_observe_batch(bec_plan* p, int nb, const bec_phys* ph, const void* psi, double* out);
int bec_lll_weight_batch(bec_plan* p, int nb, const void* psi, const double* Omega, int Mmax, double* w_out);
int bec_vortex_count_batch(bec_plan* p, int nb, const void* psi, const double* prm4, double frac, int* out4);

/* bec_resample without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same arguments, kernels, results and errors as bec_resample; its partial
 * sums live in a fixed 1024-row device array of the library (one per GPU,
 * held until the process ends) instead of a buffer allocated per call, and a
 * grid needing more rows is refused (status 1).  Calls that overlap on one
 * GPU share that array and race. */
int bec_resample_noalloc(int ndim, int n0, double dx0, const void* psi_old, int n1, double dx1, void* psi_new,
                         double* norm_out, void* stream);

/* Vortex count of a 2D state (vortex.py:10-48, _count_vortices_2d): the phase
 * winding of every grid plaquette, counted where the plaquette's mean density
 * smoothed by exp(-ell^2 k^2 / 2) exceeds frac times its maximum and its first
 * corner lies within R of (xcm, ycm).  out4 = {Nv, Nplus, Nminus, Nnet}. */
int bec_vortex_count(bec_plan* p, const void* psi, double xcm, double ycm, double R, double ell, double frac,
                     int* out4);

/* Thomas-Fermi chemical potential of unit norm (thomasfermi.py:48-62,
 * _bisect_mu_for_norm): rho(mu) is the upper root of b3 rho^2 + b2 rho =
 * mu - base (0 where there is none) on the plan's grid; hi is widened as
 * lo + 2 (hi - lo) + 1e-9 while sum rho(hi) dV < 1 (at most 200 times, else
 * *mu_out = NaN), then the bracket is narrowed by 2^steps (32 points priced
 * per pass over base) and *mu_out is its midpoint: the crossing of the norm
 * through 1, as the reference's `steps` halvings give it.  The search runs on
 * the device with one synchronisation at the end.  base: device float64
 * field (V_eff [+ phi]). */
int bec_tf_mu(bec_plan* p, const double* base, double b2, double b3, double lo, double hi, int steps,
              double* mu_out);

/* The compressed Thomas-Fermi fixed point with long range
 * (thomasfermi.py:87-119): from n_io (device, the normalised initial density)
 * each iteration forms phi = -G_C K*n (the plan's kernel spectrum), base =
 * veff + phi, mu of unit norm on base (as bec_tf_mu with the bracket
 * [min base, max base + |b2| + |b3| + 1] and 90 halvings), n_new = rho(mu)
 * normalised, its L1 change dn, n <- (1 - relax) n + relax n_new
 * normalised; it stops when dn < tol or after iters iterations, or early when
 * mu cannot be found (out4[0] = NaN) or n_new vanishes.  The iteration state
 * stays on the device (no per-iteration host synchronisation).
 * out4 (host) = mu of the last iteration, dn, iterations, converged (1/0);
 * n_io holds the final density. */
int bec_tf_nonlocal(bec_plan* p, const double* veff, double b2, double b3, double G_C, int iters, double relax,
                    double tol, double* n_io, double* out4);

/* Seed fields (seeds.py:11-103): psi = A core (1 + amp_noise xi_a)
 * exp(i (phase_noise xi_p + winding theta + vortex_sign sum_k theta_k)),
 * normalised to unit norm (norm_of), with A = exp(-r^2/2) (envelope 0) or
 * sqrt(max(env, 0)) of the device field env (envelope 1), core = prod_k
 * tanh(|r - r_k| / core) over the n_vortices in-plane positions
 * vortices[2k], vortices[2k+1] (device; lines along z in 3D), theta =
 * atan2(y, x), theta_k = atan2(y - y_k, x - x_k).  xi_a and xi_p are the
 * normals of bec_noise(rng, 0, i) and bec_noise(rng, 1, i) at point i
 * (C order).  The reference draws its noise from torch's generator; the
 * envelopes, vortex positions and the order of the factors are its. */
typedef struct bec_seed_spec {
  int envelope;
  double amp_noise;
  double phase_noise;
  int winding;
  double vortex_sign;
  double core;
  unsigned long long rng;
} bec_seed_spec;
int bec_seed(bec_plan* p, const bec_seed_spec* spec, const double* env, const double* vortices, int n_vortices,
             void* psi_out);

/* The seeds' noise on the host (no device work): out[k] = the standard normal
 * of point first + k in stream `stream` of seed rng -- Box-Muller on outputs
 * 2i and 2i+1 of the splitmix64 sequence seeded at key = output 0 of the
 * sequence seeded at (rng mod 2^32) + 2^32 stream.  No reference counterpart. */
int bec_noise(unsigned long long rng, unsigned long long stream, unsigned long long first, long long count,
              double* out);

/* ground_state loop (solvers.py:48-226): Riemannian Sobolev-preconditioned
 * CG with Polak-Ribiere beta, Barzilai-Borwein trial step and Armijo line
 * search.  psi_io holds psi0 on entry (normalised internally) and the
 * returned iterate (best-residual rule, solvers.py:225-226) on exit.
 * trace (host, nullable) receives up to trace_cap rows of
 * (walltime, iteration, E, resid_rel) (solvers.py:110-111,197-198). */
int bec_ground_state(bec_plan* p, const bec_phys* ph, const bec_solver_cfg* cfg, void* psi_io,
                     bec_solve_out* out, double* trace, int trace_cap);

/* Batched ground_state: nb (<= 64) independent solves on one grid, e.g. the
 * seeds of ground_state_multiseed (solvers.py:238-263) or the points of a
 * sweep without continuation (modes.py:295-387), advanced together so that
 * every kernel launch covers all instances still iterating.  ph[nb] gives
 * each instance's physics (the grid, kernel and solver knobs are shared);
 * psi_io holds the nb states one after the other; out[nb] as above (no
 * trace).  Each instance follows the ground_state loop exactly. */
int bec_ground_state_batch(bec_plan* p, int nb, const bec_phys* ph, const bec_solver_cfg* cfg, void* psi_io,
                           bec_solve_out* out);

/* A queue of nq solves on one grid, more than one batch holds: nb (<= 64)
 * slots advance together as in bec_ground_state_batch, and whenever some of
 * them have stopped while solves are waiting, those slots take the next
 * solves (their set-up between two graph launches of the same loop), so the
 * batch stays full until the queue drains.  ph[nq], psi_io ([nq][n]) and
 * out[nq] as in bec_ground_state_batch; every solve follows the ground_state
 * loop exactly.  The reference solves a sweep's points and seeds one after
 * another (modes.py:295-387, solvers.py:238-263).  Device-driven loop only. */
int bec_ground_state_queue(bec_plan* p, int nb, int nq, const bec_phys* ph, const bec_solver_cfg* cfg,
                           void* psi_io, bec_solve_out* out);

/* Batched ground_state over ng (<= 8) grids at once, e.g. the seeds and
 * points of a sweep whose autobox grids differ in size (modes.py:295-387,
 * which solves them one after another).  Grid g has its own plan plans[g]
 * (distinct plans), nb[g] (<= 64) instances with physics ph[off_g ..] and
 * solver settings cfg[g], and its states at psi_io[g] ([nb[g]][n_g]); ph and
 * out[] run group after group (off_g = nb[0] + .. + nb[g-1]).  Every grid's
 * instances follow the ground_state loop exactly, as in
 * bec_ground_state_batch; the steps of all grids are captured into one
 * hipGraph (one branch per grid), so one graph launch advances every
 * instance of every grid.  The device-driven loop only (cfg host_loop = 0). */
int bec_ground_state_uneven(int ng, bec_plan* const* plans, const int* nb, const bec_phys* ph,
                            const bec_solver_cfg* const* cfg, void* const* psi_io, bec_solve_out* out);

/* Largest batch the plan's grid supports in bec_ground_state_batch: 64, or 1
 * when one instance's lines do not fill whole workgroups of its line passes
 * (small or awkward N, e.g. 2D N = 20 or 50).  No reference counterpart: the
 * reference solves seeds one after the other (solvers.py:238-263). */
int bec_plan_max_batch(bec_plan* p, int* nb_max);

/* Sizes the plan's workspace, solver state and device-loop buffers for nb
 * instances ahead of bec_ground_state_batch (which otherwise sizes them on
 * first use).  No counterpart in the reference, which allocates per call:
 * after it, a batched solve on this plan allocates nothing. */
int bec_plan_reserve(bec_plan* p, int nb);

/* Autotuner (no counterpart in the reference): times the line passes of one
 * solver step over nb instances for every candidate schedule -- the stage
 * (radix) order of each transform length and a power-of-two factor on the
 * lines per workgroup of contiguous and of strided passes -- and keeps the
 * fastest in the plan for its later calls.  ms_best / ms_default (nullable)
 * receive the step-pass time of the chosen and of the default schedule;
 * report (nullable) a one-line description. */
int bec_plan_autotune(bec_plan* p, int nb, double* ms_best, double* ms_default, char* report, int report_cap);

/* Measurement hooks (no counterpart in the reference, which has no profiler
 * hooks): time every launch of one kernel with HIP events on its stream;
 * bytes sums the algorithmic bytes of the timed launches (each array element
 * the launch must read or write counted once, per instance it covered).
 * id 0 switches timing off.  Ids: kinetic passes 100 + 10*axis + kind, line
 * FFT passes 2000 + 1000*contig + 100*load + 10*dir + mul, inner products 10,
 * shift selection 20-23 (becgpp_amd/csrc/bec_internal.h). */
int bec_prof_watch(int id);
int bec_prof_read(double* total_ms, long long* count, double* bytes);
/* id -1 in bec_prof_watch times every launch site at once; then
 * bec_prof_sites lists the sites that launched (n_out = their number, at most
 * cap ids written) and bec_prof_read_site gives one site's totals (id -1:
 * all sites). */
int bec_prof_sites(int* ids, int cap, int* n_out);
int bec_prof_read_site(int id, double* total_ms, long long* count, double* bytes);

#ifdef __cplusplus
}
#endif
#endif
