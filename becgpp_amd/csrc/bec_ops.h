// Multi-pass operator drivers shared by the C-ABI entry points and the solver.
// Every driver covers a batch of instances (fields laid out [b][n]); the
// single-instance entry points use a batch of one.
#pragma once
#include "bec_internal.h"
#include <stdexcept>
#include <string>
#include <vector>

struct BecError : std::runtime_error {
  int code;
  BecError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define HIPCHK(x)                                                                        \
  do {                                                                                   \
    hipError_t e_ = (x);                                                                 \
    if (e_ != hipSuccess)                                                                \
      throw BecError(2, std::string("HIP error '") + hipGetErrorString(e_) + "' at " +  \
                            __FILE__ + ":" + std::to_string(__LINE__));                  \
  } while (0)

typedef InstPhys Phys;

// the instances a driver call covers.  Host-driven calls give a bit mask;
// the device-driven solver loop gives the flag words and the flag bit
// instead, reads its per-instance scalars from the device table tab, and
// keeps every reduction as device rows (dev) for its controller kernels.
struct Batch {
  int nb;
  unsigned long long act;     // bit b set: instance b takes part
  const unsigned* fl = nullptr;
  unsigned bit = 0;
  const double* tab = nullptr;
  bool dev = false;
  int f32 = 0;                // complex fields are float2 (FP32 phase)
};
inline Batch one_instance() { return Batch{1, 1ull}; }
inline Batch dev_batch(const Plan& p, int nb, unsigned bit) {
  Batch B{nb, 0ull};
  B.fl = p.flags;
  B.bit = bit;
  B.tab = p.dtab;
  B.dev = true;
  B.f32 = p.f32;
  return B;
}
inline Act act_of(const Batch& B) { return Act{B.act, B.fl, B.bit}; }

// rows per instance of the reducing launches of one driver call
struct EnergyRows {
  int r0, r1, r2;             // first pass (6 sums), second and third axis (2 sums each)
};

// energy components of (psi - t d) [/ its norm when normalise], one instance
struct Energy {
  double E, Ekin, Etrap, Econtact, Ethree, Erot, Egrav, Lz;
  double nrm2;    // sum |psi - t d|^2 dv of the unnormalised state
};

bool plan_long_range(const Plan& p, const Phys& ph);
void plan_reserve(Plan& p, int nb);            // grow the batch workspace to nb instances
void plan_set_phys(Plan& p, const Phys* ph, int nb);   // per-instance physics table (synchronous)
void plan_sync(Plan& p);                       // synchronise, then fold the queued reductions
double* part_of(Plan& p, int site);            // device partial-row region of a reducing launch site
double* hpart_of(Plan& p, int site);           // pinned host region, as seen by kernels
void add_job(Plan& p, int site, int nblk, int K, int KM, int slot, const Batch& B);   // queue a host fold
inline const double* hres(const Plan& p, int b) { return p.hred + (long long)b * RSTRIDE; }

// out = scale * dV * (K * rho)_crop per instance; with dot, the sums
// sum(out_unscaled... ) -> hred[b][R_EGRAV] use out as computed (scale included)
// dot2 (with dot): a second sum(out * dot2) in the same rows; site: the
// partial-row region of the sums (device loop) or their host fold
int run_conv(Plan& p, int which, const Batch& B, const double* rho, double scale, double* out, const double* dot,
             const double* dot2 = nullptr, int site = C_EGRAV, int slot = R_EGRAV);
// The convolutions of two real fields at once (device-driven loop only): kq =
// dV (K * q)_crop and kr = dV (K * r)_crop from one complex convolution of
// q + i r (K is real and even); rows of sum(kq q), sum(kq r) in part_of(site_qr)
// and of sum(kr r) in part_of(site_rr), the same count per instance (returned).
int run_conv_pair(Plan& p, const Batch& B, const double* q, const double* r, double* kq, double* kr, int site_qr,
                  int site_rr);
// kinetic/L_z/local sums of psi - t[b] d into hred[b][R_EFIRST..], [R_EREST..];
// rho_out (optional) receives |psi - t d|^2
// Fusions of the device-driven loop into the first energy pass: instances
// flagged comb_bit first form d = az z + ad d + ag g + gam psi (coefficients
// from the table); hy_out receives that pass's share of H (psi - t d).
struct EnergyFuse {
  double2* hy_out;
  const double2* z; const double2* g; double2* d_out; unsigned comb_bit;
};
EnergyRows run_energy_sums(Plan& p, const Batch& B, const double2* psi, const double2* d, const double* t,
                           double* rho_out, const EnergyFuse* fuse = nullptr);
// Line coefficients of the polynomial Armijo search (device-driven loop, rows
// only): the first axis' pass (KL_FIRST: new direction for instances with
// FL_COMB, local sums, q = Re(conj(psi) d) and r = |d|^2 to q_out / r_out when
// given, kinetic and rotation forms of d) into part_of(C_LINE); the other
// axes' forms of d into C_EREST0 / C_EREST1.
EnergyRows run_line_sums(Plan& p, const Batch& B, const double2* psi, double2* d, const double2* z,
                         const double2* g, const double* phi, double* q_out, double* r_out);
// assemble instance b's Energy from hred (after plan_sync); sgrav = sum(phi rho)
// with phi the potential as stored (its scale is passed as gscale)
Energy energy_from_sums(const Plan& p, int b, const Phys& ph, bool lr, bool normalise, double gscale);
// H passes: out = (kf k^2 + rf c k) psi + [local] psi - mu psi with psi =
// (psi_in - t d) / div; with psi_w the transformed state is written back
// (psi_w may equal psi_in) and sdiff receives its change; ydiff receives the
// change of out; phi * phisc[b] is the long-range potential (stored to
// phi_out).  Reductions -> hred[b][R_HLAST..] and device rows when reduce.
struct HArgs {
  const double2* psi; const double2* d; const double* t; const double* div;
  double2* psi_w; double2* sdiff;
  bool local_on; const double* phi; const double* phisc; double* phi_out;
  const double* kq; const double* kr;   // last pass: the accepted state's phi combined first (KinArgs::kq)
  const double* mu;
  double kf, rf; int rf_omega;
  double2* out; double2* ydiff;
  bool reduce;
};
int run_H(Plan& p, const Batch& B, const HArgs& h);   // returns rows per instance of the last pass
// H of the accepted trial (psi - t d) / sc in the device-driven loop: the first
// axis' share is in p.w1 from the trial's energy pass (EnergyFuse::hy_out),
// the remaining axes are transformed here; psi is updated in place (change to
// sdiff), out receives H psi - mu psi (change to ydiff).
int run_H_accept_dev(Plan& p, const Batch& B, const HArgs& h);
// z = IFFT(FFT(g) / (k^2/2 + sigma[b])) with sigma read from p.sigma (device).
// With dots, the last pass also accumulates the 16 CG inner products of z
// with (psi, d, g, y, s) into device rows (part_of(C_DOTS)); returns their
// rows per instance (0 without dots).
struct DotFields {
  const double2* psi; const double2* d; const double2* g; const double2* y; const double2* s;
};
int run_precond(Plan& p, const Batch& B, const double2* g, double2* out, const DotFields* dots = nullptr);

void solver_release(Plan& p);
bool plan_batchable(const Plan& p);            // batched solves possible on this grid
// FFT stage schedules (the autotuner's search space)
void set_stage_order(FftPlan& fp, const std::vector<int>& rad);
std::vector<std::vector<int>> stage_orders(const FftPlan& fp, int cap);
void bec_set_error(const char* msg);

struct bec_plan {
  Plan p;
};
