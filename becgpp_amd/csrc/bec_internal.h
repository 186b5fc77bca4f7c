// Host-side plan structure and kernel-launch wrappers of the becGPP library.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>
#include "bec_common.h"

// line-pass axis modes (which axis a pass transforms, and how lines map to
// memory).  2D: axis 1 = y (contiguous), axis 0 = x (stride N).
// 3D: axis 2 = z (contiguous), axis 1 = y (stride N), axis 0 = x (stride N^2).
enum AxisMode { A2Y = 0, A2X = 1, A3Z = 2, A3Y = 3, A3X = 4 };

// kinetic / rotation / energy pass kinds
enum KinKind { KE_FIRST = 0, KE_REST = 1, KH_FIRST = 2, KH_MID = 3, KH_LAST = 4 };

// generic line-FFT pass options
enum LoadKind { LD_C = 0, LD_R = 1, LD_HERM = 2 };
enum DirKind { DIR_FWD = 0, DIR_INV = 1, DIR_FMI = 2 };
enum MulKind { MUL_NONE = 0, MUL_KHAT2 = 1, MUL_KHAT3 = 2, MUL_PRE2 = 3, MUL_PRE3 = 4 };
enum StoreKind { ST_C = 0, ST_RE = 1 };

struct LineCfg {       // threads per line / lines per workgroup for a length
  int tpl, lpw;
};

// Arguments of the generic line-FFT pass (pad-on-load / crop-on-store).
struct PassArgs {
  FftPlan fp;
  int tpl, lpw;
  int nlines, ni;                      // line l = o*ni + i
  long long in_so, in_si, in_s;        // input: base = o*in_so + i*in_si, element stride in_s
  int n_load, off_load;                // loaded window [off_load, off_load+n_load) of the line
  long long out_so, out_si, out_s;
  int n_store, off_store;              // stored window
  double out_scale;
  const void* in;
  void* out;
  const double* khat; int H, M;        // real-even kernel spectrum, half table [H]^d
  const double* sigma; int Nk; double invNdx;   // Sobolev shift and grid frequencies
  const double* dot;                   // optional: accumulate sum(out_re * dot)
  double* partials;                    // per-block partial rows, device (nullable)
  double* hpartials;                   // the same rows in pinned host memory (nullable)
};

// Arguments of the kinetic/rotation/energy/H passes (line length N, no pad).
struct KinArgs {
  FftPlan fp;
  int tpl, lpw, N, nlines;
  double dx, invNdx;
  const double2* psi; const double2* d; double t, sc;    // loaded value (psi - t d) sc
  double2* psi_out;                                      // KH_FIRST: materialised loaded value
  double* rho_out;                                       // KE_FIRST: |loaded|^2
  const double2* acc; double2* out;
  double kf, rf;                                         // multiplier kf k^2 + rf lzc k
  double beta2, beta3, tc, s_half; int s_mode;           // local terms (KE_FIRST, KH_LAST)
  const double* phi; double phisc; double* phi_out;
  double mu;
  double* partials;                    // per-block partial rows, device (nullable)
  double* hpartials;                   // the same rows in pinned host memory (nullable)
};

struct SigmaStats {
  unsigned long long kmin, kmax;
  unsigned int cnt, bucket, rank_in, ncand;
};

// a pending reduction: per-block rows in pinned host memory -> hred[slot ..]
struct RedJob {
  int site, nblk, K, KM, slot;
};
#define BEC_MAX_JOBS 16
struct RedJobs {
  RedJob j[BEC_MAX_JOBS];
  int n;
};

#define SIGMA_BINS 2048
#define SIGMA_LDS_CAND 4096
#define SIGMA_MAX_BLOCKS 256

struct Plan {
  int ndim, N, pad, M, H;
  long long n, nM, nS1, nS2;         // field points, padded points, conv spectra sizes
  double Lbox, dx, dv, s, tc, s_half; int s_mode;
  double invNdx;
  hipStream_t stream;
  FftPlan fpN, fpM;
  double2* twN; double2* twM;
  LineCfg lcN, lcM;
  double* khat[2];                   // [0] = K, [1] = W (virial) real-even spectra
  int kkind[2];                      // 0 none, 1 newton, 2 log, 3 q2d / external
  double2* w1; double2* w2;          // n-point complex scratch
  double2* s1; double2* s2;          // conv spectra
  double* rho;                       // n-point real scratch
  double* partials; long long part_stride; int max_blocks;   // one region per PartSite
  double* hpart; double* hpart_dev;  // pinned host copy of the regions (host view / kernel view)
  RedJobs pend;                      // reductions whose rows the host folds at the next sync
  int hlast_blocks;                  // rows of the last H pass (rho max for the shift mask)
  double* red;                       // device reduction outputs (256 doubles)
  double* hred;                      // pinned, device-mapped host mirror (host view)
  double* hred_dev;                  // the same memory as seen by kernels
  // Sobolev shift selection
  double* loc; double* cand; unsigned* hist; SigmaStats* stats; double* sigma; void* sig_part;
  // solver state (allocated on first solve)
  double2* st[8]; double* stphi[2];
  bool st_alloc;
};

// reduction slot layout in Plan::red (device) / Plan::hred (host)
enum RedSlot {
  R_EFIRST = 0,    // 6: kin, lz, srho, sV, s2, s3
  R_EREST = 8,     // 2 per remaining axis (up to 2 axes): 8..11
  R_EGRAV = 12,    // 1: sum(phi_raw * rho)
  R_HLAST = 16,    // 3: |h|^2, Re(psi* h), max rho
  R_NORM = 20,     // 1
  R_DOTS = 24,     // 17 inner products
  R_SIGMA = 48,    // 1 sigma value
  R_MOM = 52,      // 5: mass, R2 rho, rperp2 rho, z2 rho, max rho
  R_RES2 = 60,     // 1: |h - mu psi|^2
  R_TMP = 64
};

// partial-row regions, one per reducing launch site
enum PartSite { C_EFIRST = 0, C_EREST0, C_EREST1, C_EGRAV, C_HLAST, C_NORM, C_DOTS, C_MOM,
                C_RES2, C_COUNT };

// ---------------------------------------------------------------------------
// launch wrappers (bec_kernels.hip)
// ---------------------------------------------------------------------------
LineCfg line_cfg(int L);
int launch_pass(int contig, int ld, int dir, int mul, int st, const PassArgs& a, hipStream_t s);
int launch_kin(int ax, int kind, const KinArgs& a, hipStream_t s);
int launch_norm2(const double2* psi, long long n, double* hpart, hipStream_t s);
void launch_scale(const double2* in, double2* out, double sc, long long n, hipStream_t s);
void launch_combine(double2* d, const double2* z, const double2* g, const double2* psi, double az,
                    double ad, double ag, double gre, double gim, long long n, hipStream_t s);
int launch_dots(const double2* z, const double2* psi, const double2* d, const double2* g,
                const double2* gp, const double2* pp, long long n, double* hpart, hipStream_t s);
int launch_res2(const double2* h, const double2* psi, double mu, double2* g_out, long long n,
                double* hpart, hipStream_t s);
void launch_rho(const double2* psi, double* rho, long long n, hipStream_t s);
int launch_moments(const double2* psi, int ndim, int N, double dx, double* partials, double* hpart,
                   hipStream_t s);
void launch_sigma(const Plan& p, const double2* psi, const double* phi, double phisc, const double* mrows,
                  int nrows, int mkp, int mq, double beta2, double beta3, double floor_, hipStream_t s);
void launch_kernel_real(int ndim, int M, double dx, int kind, double2* out, hipStream_t s);
void launch_extract_khat(int ndim, int M, int H, const double2* kf, double* khat, hipStream_t s);
void launch_init_sigma(const Plan& p, hipStream_t s);

// live timing of one watched launch site (bec_prof.cpp); site ids:
//   kinetic passes 100 + 10*axis + kind;  line-FFT passes
//   2000 + 1000*contig + 100*load + 10*dir + mul;  elementwise kernels < 100
void prof_pre(int id, hipStream_t s);
bool prof_events(int id, hipEvent_t* start, hipEvent_t* stop);   // events for hipExtLaunchKernelGGL
void prof_post(int id, hipStream_t s);
enum ElemSite { SITE_DOTS = 10, SITE_COMBINE = 11, SITE_NORM = 12, SITE_SIGMA_LOCAL = 20,
                SITE_SIGMA_HIST = 21, SITE_SIGMA_GATHER = 22, SITE_SIGMA_SELECT = 23 };
