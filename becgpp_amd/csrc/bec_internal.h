// Host-side plan structure and kernel-launch wrappers of the becGPP library.
//
// Every kernel works on a batch of nb instances of one grid geometry: fields
// are stored instance after instance ([b][n], n = N^ndim points), lines are
// numbered across instances (line l belongs to instance l / nl), and the
// per-instance scalars of a launch travel by value in its arguments.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>
#include "bec_common.h"

#define BEC_BMAX 64          // instances per launch (per-instance scalars by value; 64-bit masks)
#define BEC_GMAX 8           // grids advanced together by bec_ground_state_uneven

// line-pass axis modes (which axis a pass transforms, and how lines map to
// memory).  2D: axis 1 = y (contiguous), axis 0 = x (stride N).
// 3D: axis 2 = z (contiguous), axis 1 = y (stride N), axis 0 = x (stride N^2).
enum AxisMode { A2Y = 0, A2X = 1, A3Z = 2, A3Y = 3, A3X = 4 };

// kinetic / rotation / energy pass kinds
enum KinKind { KE_FIRST = 0, KE_REST = 1, KH_FIRST = 2, KH_MID = 3, KH_LAST = 4,
               KL_FIRST = 5 };   // KL_FIRST: line coefficients of the polynomial Armijo search (first axis)

// generic line-FFT pass options
enum LoadKind { LD_C = 0, LD_R = 1, LD_HERM = 2, LD_R2 = 3 };   // LD_R2: two real fields as re + i im
enum DirKind { DIR_FWD = 0, DIR_INV = 1, DIR_FMI = 2 };
enum MulKind { MUL_NONE = 0, MUL_KHAT2 = 1, MUL_KHAT3 = 2, MUL_PRE2 = 3, MUL_PRE3 = 4, MUL_GAUSS = 5 };
// ST_CDOT: complex store + the 16 CG inner products; ST_RE2: the real part to
// out and the imaginary part to out2 (a pair of real convolutions), with
// sum(re dot), sum(re dot2) in partials and sum(im dot2) in partials2
enum StoreKind { ST_C = 0, ST_RE = 1, ST_CDOT = 2, ST_RE2 = 3 };

struct LineCfg {       // threads per line / lines per workgroup for a length
  int tpl, lpw;
};

// per-instance physics table entry (device memory, constant during a solve)
struct InstPhys {
  double beta2, beta3, Omega, G_C;
};

// Which instances of a batch a launch covers: a host bit mask, or (in the
// device-driven solver loop) the per-instance flag words fl[b] tested
// against bit.
struct Act {
  unsigned long long m;
  const unsigned* fl;
  unsigned bit;
};

inline Act mask_act(unsigned long long m) { return Act{m, nullptr, 0u}; }

// per-instance flags of the device-driven loop
enum InstFlag : unsigned {
  FL_DIR = 1,       // preconditioner, shift and inner products this step
  FL_COMB = 2,      // a new search direction is formed this step
  FL_SEARCH = 4,    // an Armijo trial is evaluated this step
  FL_SLR = 8,       // ... with its long-range convolution
  FL_FAIL = 16,     // the line search failed: direction <- preconditioned gradient
  FL_ACC = 32,      // the trial was accepted: psi and the gradient are updated
  FL_BEST = 64,     // the state is the lowest-residual one so far: keep a copy
  FL_PHIC = 128,    // polynomial search accepted with long range: combine the new potential
};

// per-instance scalars of the device-driven loop, DS_N doubles per instance
// (DS_CQ, DS_CR, DS_CS: the accepted state's potential (phi + CQ K*q + CR K*r) CS
// in the polynomial search, q = Re(conj(psi) d), r = |d|^2)
enum DevScal { DS_T = 0, DS_SC = 1, DS_MU = 2, DS_PHISC = 3, DS_AZ = 4, DS_AD = 5, DS_AG = 6, DS_GRE = 7,
               DS_GIM = 8, DS_CQ = 9, DS_CR = 10, DS_CS = 11, DS_N = 16 };

// Arguments of the generic line-FFT pass (pad-on-load / crop-on-store).
struct PassArgs {
  FftPlan fp;
  int tpl, lpw;
  int nb, nl;                          // instances, lines per instance
  Act act;                             // active instances
  int ni;                              // line l' = o*ni + i within an instance
  long long in_so, in_si, in_s, in_ib; // input: b*in_ib + o*in_so + i*in_si, element stride in_s
  int n_load, off_load;                // loaded window [off_load, off_load+n_load) of the line
  long long out_so, out_si, out_s, out_ib;
  int n_store, off_store;              // stored window
  double out_scale;
  const void* in;
  void* out;
  const void* in2;                     // LD_R2: the field loaded as the imaginary part
  void* out2;                          // ST_RE2: the imaginary part's output
  const double* khat; int H, M;        // real-even kernel spectrum, half table [H]^d
  int kfull;                           // the lines cover all M frequencies of the transformed real axis
                                       //   (a pair convolution), not its H = M/2 + 1
  const double* sigma; int Nk; double invNdx;   // per-instance Sobolev shifts and grid frequencies
  double gl2;                          // MUL_GAUSS: the line's factor exp(-gl2 k^2)
  const double* gl2s;                  // MUL_GAUSS: per-instance gl2 (device; overrides gl2)
  const double* dot;                   // optional: rows of sum(out_re * dot) (dot laid out like out)
  const double* dot2;                  // optional with dot: a second sum(out_re * dot2) in the same rows
  double* partials;                    // per-block partial rows, device (nullable)
  double* hpartials;                   // the same rows in pinned host memory (nullable)
  double* partials2;                   // ST_RE2: rows of sum(im dot2), device
  unsigned long long* ucount;          // measurement: algorithmic bytes of the instances covered
  unsigned long long ubytes;           //   (ubytes per instance, added once per instance; nullable)
  unsigned long long lbytes;           //   and lbytes once per launch that covers an instance (the
                                       //   kernel spectrum, shared by the instances)
  const double2* dpsi; const double2* dd; const double2* dg; const double2* dy; const double2* ds;  // ST_CDOT
  int f32;                             // complex fields (in/out, ST_CDOT) are float2: an FP32 phase
  int lsh[2];                          // tuned lines-per-workgroup shift: contiguous, strided passes
};

// Arguments of the kinetic/rotation/energy/H passes (line length N, no pad).
struct KinArgs {
  FftPlan fp;
  int tpl, lpw, N, nl, nb;
  Act act;
  long long istride;                                     // points per instance
  double dx, invNdx;
  const double2* psi; const double2* d;                  // loaded value (psi - t[b] d) / sc[b]
  double2* psi_w;                                        // KH_FIRST: loaded value written back
  double2* sdiff;                                        // KH_FIRST: loaded value - psi
  double* rho_out;                                       // KE_FIRST: |loaded|^2; KL_FIRST: Re(conj(psi) d)
  double* r_out;                                         // KL_FIRST: |d|^2
  const double2* acc; double2* out;
  double2* ydiff;                                        // KH_LAST: new out - previous out
  double kf, rf; int rf_omega;                           // multiplier kf k^2 + rf c k (rf = -Omega[b])
  const InstPhys* iphys; int local_on;                   // local terms in KH_LAST / KE_FIRST
  double tc, s_half; int s_mode;
  const double* phi;                                     // long-range potential rows x phisc[b]
  double* phi_out;
  // KH_LAST (polynomial search): instances flagged FL_PHIC first take
  // phi <- (phi + CQ kq + CR kr) CS (table scalars), written to phi_out
  const double* kq; const double* kr;
  double t[BEC_BMAX], sc[BEC_BMAX], mu[BEC_BMAX], phisc[BEC_BMAX];
  const double* tab; unsigned tabm;                      // device scalars: bit k of tabm reads DS slot k
  double* partials; double* hpartials;
  unsigned long long* ucount;                            // measurement: algorithmic bytes of the instances
  unsigned long long ubytes;                             //   covered (ubytes each; nullable)
  // device-driven loop fusions
  double2* hy_out;                                       // KE_FIRST: also IFFT(m FFT(loaded)) of this axis
  const double2* z; const double2* g; double2* d_out;    // KE_FIRST: instances with comb_bit form
  unsigned comb_bit;                                     //   d = az z + ad d + ag g + gam psi first
  int acc_sc;                                            // KH_MID/KH_LAST: acc is divided by sc[b]
  int f32;                                               // complex fields are float2: an FP32 phase
  int lsh[2];                                            // tuned lines-per-workgroup shift (as PassArgs)
};

// per-instance combination d <- az z + ad d + ag g + gam psi
struct CombArgs {
  double az[BEC_BMAX], ad[BEC_BMAX], ag[BEC_BMAX], gre[BEC_BMAX], gim[BEC_BMAX];
};

struct SigmaStats {
  unsigned long long kmin, kmax;       // order keys of the bulk's min / max (this evaluation)
  unsigned long long blo, bhi;         // value range of the histogram bins (the previous evaluation's)
  unsigned int cnt, bucket, rank_in, ncand;
};

// a pending reduction: per-block rows in pinned host memory -> hred
struct RedJob {
  int site, nblk, K, KM, slot;        // nblk rows per instance
  int nb;                             // instances of the launch
  unsigned long long act;             // instances whose rows were written
};
#define BEC_MAX_JOBS 16
struct RedJobs {
  RedJob j[BEC_MAX_JOBS];
  int n;
};

#define SIGMA_BINS 2048
#define SIGMA_LDS_CAND 4096
#define SIGMA_MAX_BLOCKS 256
#define SIGMA_U 4                    // elements per thread and round of the shift's element passes

struct Plan {
  int ndim, N, pad, M, H;
  long long n, nM, nS1, nS2;         // points per instance field, padded points, conv spectra sizes
  double Lbox, dx, dv, s, tc, s_half; int s_mode;
  double invNdx;
  hipStream_t stream;
  FftPlan fpN, fpM;
  double2* twN; double2* twM;
  LineCfg lcN, lcM;
  int lsh[2];                        // autotuned log2 factor on the lines per workgroup (contiguous, strided)
  double* tfbuf;                     // Thomas-Fermi search: its state and block rows
  int* tfflag; int* tfflag_dev;      // pinned host-mapped: the TF fixed point finished
  double* dgbuf;                     // diagnostics: 1024 block rows of 16 values
  double* khat[2];                   // [0] = K, [1] = W (virial) real-even spectra
  int kkind[2];                      // 0 none, 1 newton, 2 log, 3 table
  // batch workspace, sized for cap instances
  int cap;
  double2* w1;                       // n complex per instance (H accumulator, preconditioner)
  double2* s1; double2* s2;          // conv spectra per instance
  double* rho;                       // n real per instance
  double* loc; double* cand; unsigned* hist; SigmaStats* stats; double* sigma; void* sig_part;
  InstPhys* iphys;                   // per-instance physics of the current call
  double* partials; long long part_stride; int max_rows;   // one region per PartSite
  double* hpart; double* hpart_dev;  // pinned host copy of the regions (host view / kernel view)
  RedJobs pend;                      // reductions whose rows the host folds at the next sync
  int hlast_rows;                    // rows per instance of the last H pass (rho max)
  double* hred; double* hred_dev;    // pinned results [instance][RSTRIDE]
  // solver state, sized for scap instances
  int scap;
  double2* st[8];
  double* stphi[6];                  // phi (scaled), raw trial convolution; polynomial search: q, r, K*q, K*r
  float2* st32[8];                   // the same fields in complex64 (FP32 warm phase)
  int f32;                           // 1 while an FP32 phase runs: device-loop passes use st32
  // device-driven loop state, sized for scap instances
  void* dinst; unsigned* flags; double* dtab; double* dtrace; int dtrace_cap;
  int* hdone; int* hdone_dev;        // pinned: instance finished (written by the controller)
  double* hvlog; double* hvlog_dev;  // pinned: progress records + per-instance record counts
  hipStream_t cstream;               // plan-owned stream the loop's steps are captured on
};

// reduction slots, RSTRIDE doubles per instance in Plan::hred
#define RSTRIDE 64
enum RedSlot {
  R_EFIRST = 0,    // 6: kin, lz, srho, sV, s2, s3
  R_EREST = 8,     // 2 per remaining axis: 8..11
  R_EGRAV = 12,    // 1: sum(phi_raw * rho)
  R_WVIR = 14,     // 1: sum(rho * W*rho) (q2d virial kernel)
  R_HLAST = 16,    // 3: |h|^2, Re(psi* h), max rho
  R_NORM = 20,     // 1
  R_DOTS = 24,     // 16 inner products
  R_SIGMA = 44,    // 1 sigma value
  R_MOM = 48,      // 5: mass, R2 rho, rperp2 rho, z2 rho, max rho
  R_RES2 = 56,     // 1: |h - mu psi|^2
};

// partial-row regions, one per reducing launch site
// C_LINE (rows of up to 32 values: the polynomial search's first pass) is the
// last site and spans two regions; the device buffer holds C_COUNT + 1.
enum PartSite { C_EFIRST = 0, C_EREST0, C_EREST1, C_EGRAV, C_HLAST, C_NORM, C_DOTS, C_MOM,
                C_RES2, C_EGRAV2, C_LINE, C_COUNT };

// ---------------------------------------------------------------------------
// device-driven solver loop (bec_ctrl.hip): per-instance state of the loop of
// solvers.py:48-226 kept in device memory, advanced by two controller kernels
// per step.  Doubles first, then counters.
// ---------------------------------------------------------------------------
struct DevInst {
  double E, mu, resid_rel, best_resid, alpha, sigma, zg_prev, gg_prev, zg_cur, gg_cur, E_ref;
  double trial, gd, trial_acc, cz_re, cz_im;
  // polynomial search: sums of the current (normalised) state, carried from
  // the accepted trial: <psi,(T + V - Omega L_z) psi> dv, sum rho^2 dv,
  // sum rho^3 dv, sum rho (K*rho) dv
  double L00, Spp, Sppp, Kpp;
  double clk0;
  long long n_energy;
  int it, it_ref, n_fail, n_iter, stop, it_acc, ls, converged, done, pending, lr, n_tr, n_log, pad_;
};

struct CtrlArgs {
  int nb, ndim;
  double dv;
  // solver knobs (solvers.py:59-88)
  int maxit, check, ls_max, stall_max, restart_period, conv_window, cg_beta, verbose;
  double res_tol, etol, alpha_min, alpha_max, c1, shrink, growth, bb_min, bb_max, bb_mix;
  // device partial rows of this step's reducing launches, rows per instance
  const double* r_dots; int n_dots;      // 16 sums per row
  const double* r_hl; int n_hl;          // |h|^2, Re<psi,h>, max rho (4 per row)
  const double* r_e0; int n_e0;          // first energy pass: 6 sums (8 per row)
  const double* r_e1; int n_e1;          // second axis: 2 sums
  const double* r_e2; int n_e2;          // third axis (3D): 2 sums
  const double* r_gr; int n_gr;          // sum phi_raw rho (1 per row)
  const double* r_kq; int n_kq;          // polynomial search: sum q K*q, sum r K*q (2 per row)
  const double* r_kr; int n_kr;          //   sum r K*r (1 per row)
  const double* sigma;                   // Sobolev shift per instance
  const InstPhys* iphys;
  DevInst* st; unsigned* fl; double* tab;
  double* trace; int trace_cap; double clk_hz;   // convergence history of instance 0
  int* hdone;                            // host-mapped: instance finished
  double* vlog;                          // host-mapped progress records, VLOG_CAP per instance
};
// progress record of the device-driven loop (printed by the host when verbose)
#define VLOG_CAP 64
#define VLOG_REC 6                       // kind, it, E, resid, best, step
enum { VLOG_PROGRESS = 1, VLOG_STALL = 2 };
void launch_ctrl_dir(const CtrlArgs& c, hipStream_t s);
void launch_ctrl_armijo(const CtrlArgs& c, hipStream_t s);
// polynomial Armijo search: every trial of the line from one pass of sums
void launch_ctrl_line(const CtrlArgs& c, hipStream_t s);
// accepted polynomial trials with long range: phi <- (phi + CQ kq + CR kr) CS
void launch_phi_comb(double* phi, const double* kq, const double* kr, const double* tab, const unsigned* fl,
                     long long n, int nb, hipStream_t s);

// ---------------------------------------------------------------------------
// launch wrappers (bec_kernels.hip); reducing launches return rows/instance
// ---------------------------------------------------------------------------
LineCfg line_cfg(const FftPlan& fp);
int lpw_min(int tpl);                  // fewest lines per workgroup: whole wavefronts
int launch_pass(int contig, int ld, int dir, int mul, int st, const PassArgs& a, hipStream_t s);
int launch_kin(int ax, int kind, const KinArgs& a, hipStream_t s);
int elem_blocks(long long n);
int launch_norm2(const double2* psi, long long n, int nb, unsigned long long act, double* hpart, hipStream_t s);
int reduce_blocks(long long n);      // workgroups per instance of the reducing elementwise kernels
void launch_scale(const double2* in, double2* out, const double* divs, long long n, int nb, hipStream_t s,
                  unsigned long long act = ~0ull);   // the instances of act only
void launch_combine(double2* d, const double2* z, const double2* g, const double2* psi, const CombArgs& c,
                    const double* tab, long long n, int nb, Act act, hipStream_t s);
int launch_dots(const double2* z, const double2* psi, const double2* d, const double2* g, const double2* y,
                const double2* sv, long long n, int nb, Act act, double* partials, double* hpart, hipStream_t s);
int launch_res2(const double2* h, const double2* psi, double mu, double2* g_out, long long n, double* hpart,
                hipStream_t s);
void launch_rho(const double2* psi, double* rho, long long n, int nb, hipStream_t s);
// per instance of nb (fields [b][n]): rows b * blocks + block
int launch_moments(const double2* psi, int ndim, int N, double dx, double* partials, double* hpart,
                   hipStream_t s, int nb = 1);
void launch_copy(const double2* src, double2* dst, long long n, int nb, Act mask, hipStream_t s);
// device-driven loop, after the Armijo controller: BEST <- psi for FL_BEST,
// D <- az Z + gam psi (table coefficients) for FL_FAIL
// f32: the complex fields hold float2 (FP32 phase; pointers reinterpreted)
void launch_post(const double2* psi, double2* best, double2* d, const double2* z, const double* tab,
                 const unsigned* fl, long long n, int nb, int f32, hipStream_t s);
void launch_narrow(const double2* in, float2* out, long long n, hipStream_t s);
// diagnostics: LLL projections of a chunk of m (rows of 3 sums per m), phases
// and plaquette windings of the vortex count
// (batched over nb instances, fields [b][n]; per-instance parameters in the
// device array prm; rows of instance b at b * blocks + block)
int launch_lll(const double2* psi, int N, double dx, const double* prm, int m0, int nb, double* rows, hipStream_t s);
int launch_vortex(const double2* psi, const double* nbar, int N, double dx, const double* prm, int nb, double* rows,
                  hipStream_t s);
int launch_nbar(const double2* f, long long n, int nb, double* nbar, double* rows, hipStream_t s);
void launch_abs2c(const double2* psi, long long n, double2* out, hipStream_t s);
// multilinear resampling from an n0-point grid (first node x0, span (n0-1) dx0)
// onto the n1-point grid of spacing dx1; per-block rows of sum |out|^2
int launch_resample(const double2* in, int ndim, int n0, double x0, double span, int n1, double dx1, double2* out,
                    double* rows, hipStream_t s);
void launch_widen(const float2* in, double2* out, long long n, hipStream_t s);
void launch_sigma(const Plan& p, int nb, Act act, const double2* psi, const double* phi,
                  const double* mrows, int nrows, int mkp, int mq, double floor_, hipStream_t s, int f32 = 0);
void launch_kernel_real(int ndim, int M, double dx, int kind, double2* out, hipStream_t s);
void launch_extract_khat(int ndim, int M, int H, const double2* kf, double* khat, hipStream_t s);
void launch_init_sigma(const Plan& p, int nb, hipStream_t s);

// live timing of one watched launch site (bec_prof.cpp); site ids:
//   kinetic passes 100 + 10*axis + kind;  line-FFT passes
//   2000 + 1000*contig + 100*load + 10*dir + mul;  elementwise kernels < 100
void prof_pre(int id, hipStream_t s, double bytes = 0.0);
void prof_post(int id, hipStream_t s);
// bytes: algorithmic bytes of the launch (summed per watched launch)
bool prof_events(int id, hipEvent_t* start, hipEvent_t* stop, double bytes);   // events for hipExtLaunchKernelGGL
// device counter to which watched launches of the device-driven loop add their
// algorithmic bytes per instance covered (their instances are not known on the host)
unsigned long long* prof_counter(int id);
bool prof_active();                  // some launch site is being timed
enum ElemSite { SITE_DOTS = 10, SITE_COMBINE = 11, SITE_NORM = 12, SITE_SIGMA_LOCAL = 20,
                SITE_SIGMA_GATHER = 22, SITE_SIGMA_SELECT = 23, SITE_CTRL_DIR = 30, SITE_CTRL_ARMIJO = 31,
                SITE_CTRL_LINE = 32, SITE_POST = 33, SITE_PHI_COMB = 34 };
