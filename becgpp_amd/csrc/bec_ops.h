// Multi-pass operator drivers shared by the C-ABI entry points and the solver.
#pragma once
#include "bec_internal.h"
#include <stdexcept>
#include <string>

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

struct Phys { double beta2, beta3, Omega, G_C; };

// energy components of (psi - t d) [/ its norm when normalise]
struct Energy {
  double E, Ekin, Etrap, Econtact, Ethree, Erot, Egrav, Lz;
  double nrm2;    // sum |psi - t d|^2 dv of the unnormalised state
};

bool plan_long_range(const Plan& p, const Phys& ph);
void plan_sync(Plan& p);                       // synchronise, then fold the queued reductions
double* part_of(Plan& p, int site);            // device partial-row region of a reducing launch site
double* hpart_of(Plan& p, int site);           // pinned host region, as seen by kernels
void add_job(Plan& p, int site, int nblk, int K, int KM, int slot);   // queue a host fold

// out = scale * dV * (K * rho)_crop, optionally sum(out * dot) -> red[R_EGRAV]
void run_conv(Plan& p, int which, const double* rho, double scale, double* out, const double* dot);
// kinetic/L_z/local sums of psi - t d into red[R_EFIRST..], red[R_EREST..];
// rho_out (optional) receives |psi - t d|^2
void run_energy_sums(Plan& p, const double2* psi, const double2* d, double t, double* rho_out);
// assemble Energy from red (host copy must be current); sgrav = sum(phi_raw rho)
Energy energy_from_sums(const Plan& p, const Phys& ph, bool lr, bool normalise);
// H passes: out = (kf k^2 + rf c k) psi + [local] psi - mu psi with psi = (psi_in - t d)/div
// materialised into psi_out when given; phi*phisc is the long-range potential
// (copied to phi_out when given).  Reductions -> red[R_HLAST..] when reduce.
void run_H(Plan& p, const double2* psi_in, const double2* d, double t, double div, double2* psi_out,
           const Phys& ph, bool local_on, const double* phi, double phisc, double* phi_out, double mu,
           double kf, double rf, double2* out, bool reduce);
// z = IFFT(FFT(g) / (k^2/2 + sigma)) with sigma read from p.sigma (device)
void run_precond(Plan& p, const double2* g, double2* out);

void solver_release(Plan& p);
void bec_set_error(const char* msg);

struct bec_plan {
  Plan p;
};
