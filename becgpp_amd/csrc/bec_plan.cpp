// Plan management, multi-pass operator drivers and the single-instance C-ABI
// entry points (the solver loop lives in bec_solver.cpp).
#include <math.h>
#include <stdio.h>
#include <string.h>
#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "../../include/becgpp_hip.h"
#include "bec_ops.h"

static thread_local std::string g_err;

void bec_set_error(const char* msg) { g_err = msg; }

extern "C" const char* bec_last_error(void) { return g_err.c_str(); }

#define BEC_TRY(...)                                \
  try {                                             \
    __VA_ARGS__;                                    \
    return 0;                                       \
  } catch (const BecError& e_) {                    \
    g_err = e_.what();                              \
    return e_.code;                                 \
  } catch (const std::exception& e_) {              \
    g_err = e_.what();                              \
    return 2;                                       \
  }

static void bad(const std::string& m) { throw BecError(1, m); }

// ---------------------------------------------------------------------------
// FFT plans
// ---------------------------------------------------------------------------
// Stage radices of L = 2^a 3^b 5^c 7^d 11^e 13^f.  When 8 divides L the order
// starts and ends with a power-of-two radix (8, 4 or 2): with tpl = L/8
// threads per line those two stages can run from and to registers (fft_reg).
// reg reports that they do, which is chosen for L = 2^a only: a line of
// L = 3 2^a on tpl = 3 2^(a-3) threads needs four or more lines per
// workgroup for whole wavefronts, and the same line through LDS on a power
// of two threads (tpl = 64 at L = 384, one wavefront per line, four times the
// workgroups) measured 9% more CG iterations per second on the 2D lattice at
// N = 384 (3 instances) while L = 2^a lines lost 7-11% through LDS.
static bool factor(int L, std::vector<int>& rad, bool& reg) {
  rad.clear();
  int n = L, a = 0;
  while (n % 2 == 0) { ++a; n /= 2; }
  std::vector<int> p2, odd;
  for (int i = 0; i < a / 3; ++i) p2.push_back(8);
  if (a % 3 == 1) p2.push_back(2);
  if (a % 3 == 2) p2.push_back(4);
  while (n % 3 == 0) { odd.push_back(3); n /= 3; }
  while (n % 5 == 0) { odd.push_back(5); n /= 5; }
  for (int q : {7, 11, 13})
    while (n % q == 0) { odd.push_back(q); n /= q; }
  if (n != 1) return false;
  reg = false;
  if (a >= 3 && (L & (L - 1)) == 0)    // L = 2^a: tpl = L/8 is a power of two
    for (int l = 1; l * (L / 8) <= BEC_WG; l *= 2)
      if ((l * (L / 8)) % 64 == 0) reg = true;
  if (L == 8) reg = true;              // one thread per line, one stage
  if (a >= 3 && p2.size() == 1 && !odd.empty()) { p2 = {4, 2}; }
  // a spare radix 2 between first and last pairs with a 3: one stage fewer
  if (a >= 3 && p2.size() >= 3 && p2.back() == 2 && !odd.empty() && odd[0] == 3) {
    p2.pop_back();
    odd[0] = 6;
  }
  if (a >= 3) {
    rad.push_back(p2[0]);
    for (int r : odd) rad.push_back(r);
    for (size_t i = 1; i < p2.size(); ++i) rad.push_back(p2[i]);
  } else {
    for (int r : p2) rad.push_back(r);
    for (int r : odd) rad.push_back(r);
  }
  return (int)rad.size() <= BEC_MAX_STAGES;
}

// Stage radices (in order) of fp, with each stage's Ns, twiddle stride and
// j / Ns multiplier; any order of the factors is a valid Stockham transform.
void set_stage_order(FftPlan& fp, const std::vector<int>& rad) {
  const int L = fp.L;
  fp.nst = (int)rad.size();
  fp.odd = 0;
  for (int r : rad)
    if (r == 7 || r == 11 || r == 13) fp.odd = 1;
  int ns = 1;
  for (int i = 0; i < BEC_MAX_STAGES; ++i) {
    fp.rad[i] = i < fp.nst ? rad[i] : 1;
    fp.ns[i] = ns;
    fp.step[i] = i < fp.nst ? L / (ns * rad[i]) : 1;
    // j / ns == umulhi(j, mag) for 0 <= j < 2^12: mag = floor(2^32/ns) + 1 exceeds
    // 2^32/ns by at most 1/ns, so the error j (mag ns - 2^32) / 2^32 < j ns / 2^32 < 1/ns
    fp.mag[i] = ns > 1 ? (unsigned)((1ull << 32) / (unsigned long long)ns) + 1u : 0u;
    if (i < fp.nst) ns *= rad[i];
  }
}

// Distinct stage orders of fp's radices that keep its kind of transform (a
// register-resident one needs a first and a last radix dividing 8), at most
// cap of them, the current order first.
std::vector<std::vector<int>> stage_orders(const FftPlan& fp, int cap) {
  std::vector<int> r(fp.rad, fp.rad + fp.nst);
  std::vector<std::vector<int>> out{r};
  std::vector<int> q = r;
  std::sort(q.begin(), q.end());
  do {
    if ((int)out.size() >= cap) break;
    if (fp.reg && fp.nst > 1 && (8 % q.front() != 0 || 8 % q.back() != 0)) continue;
    if (q != r) out.push_back(q);
  } while (std::next_permutation(q.begin(), q.end()));
  return out;
}

static void make_fft(int L, FftPlan& fp, double2** tw_dev, hipStream_t s) {
  std::vector<int> rad;
  bool reg = false;
  if (L < 2 || L > BEC_MAX_L || !factor(L, rad, reg))
    bad("transform length " + std::to_string(L) + " unsupported: need 2^a 3^b 5^c 7^d 11^e 13^f <= " +
        std::to_string(BEC_MAX_L));
  fp.L = L;
  fp.reg = reg ? 1 : 0;
  fp.wide = 0;
  set_stage_order(fp, rad);
  // twiddles exp(-2 pi i m / L) evaluated in extended precision
  std::vector<double2> tw(L);
  const long double pi = 3.141592653589793238462643383279502884L;
  for (int m = 0; m < L; ++m) {
    const long double ang = -2.0L * pi * (long double)m / (long double)L;
    tw[m].x = (double)cosl(ang);
    tw[m].y = (double)sinl(ang);
  }
  for (int m = 0; m < L; ++m) {      // exact values on the axes
    if ((4 * m) % L == 0) {
      const int q = (4 * m) / L;
      const double c[4] = {1.0, 0.0, -1.0, 0.0}, sn[4] = {0.0, -1.0, 0.0, 1.0};
      tw[m].x = c[q];
      tw[m].y = sn[q];
    }
  }
  HIPCHK(hipMalloc(tw_dev, sizeof(double2) * L));
  HIPCHK(hipMemcpy(*tw_dev, tw.data(), sizeof(double2) * L, hipMemcpyHostToDevice));
  fp.tw = *tw_dev;
}

// ---------------------------------------------------------------------------
// workspace, reductions, physics table
// ---------------------------------------------------------------------------
bool plan_long_range(const Plan& p, const Phys& ph) { return p.kkind[0] != 0 && fabs(ph.G_C) > 1e-15; }

double* part_of(Plan& p, int site) { return p.partials + (size_t)site * p.part_stride; }
double* hpart_of(Plan& p, int site) { return p.hpart_dev + (size_t)site * p.part_stride; }

static void free_batch(Plan& p) {
  void* ptrs[] = {p.w1, p.s1, p.s2, p.rho, p.loc, p.cand, p.hist, p.stats, p.sigma, p.sig_part, p.iphys,
                  p.partials};
  for (void* q : ptrs)
    if (q) (void)hipFree(q);
  if (p.hpart) (void)hipHostFree(p.hpart);
  if (p.hred) (void)hipHostFree(p.hred);
  p.w1 = nullptr; p.s1 = nullptr; p.s2 = nullptr; p.rho = nullptr; p.loc = nullptr; p.cand = nullptr;
  p.hist = nullptr; p.stats = nullptr; p.sigma = nullptr; p.sig_part = nullptr; p.iphys = nullptr;
  p.partials = nullptr; p.hpart = nullptr; p.hred = nullptr; p.hpart_dev = nullptr; p.hred_dev = nullptr;
  p.cap = 0;
}

void plan_reserve(Plan& p, int nb) {
  if (nb < 1 || nb > BEC_BMAX) bad("batch size must be in [1, " + std::to_string(BEC_BMAX) + "]");
  if (nb <= p.cap) return;
  HIPCHK(hipStreamSynchronize(p.stream));
  free_batch(p);
  const long long B = nb;
  HIPCHK(hipMalloc(&p.w1, sizeof(double2) * p.n * B));
  HIPCHK(hipMalloc(&p.s1, sizeof(double2) * p.nS1 * B));
  if (p.nS2) HIPCHK(hipMalloc(&p.s2, sizeof(double2) * p.nS2 * B));
  HIPCHK(hipMalloc(&p.rho, sizeof(double) * p.n * B));
  HIPCHK(hipMalloc(&p.loc, sizeof(double) * p.n * B));
  HIPCHK(hipMalloc(&p.cand, sizeof(double) * p.n * B));
  HIPCHK(hipMalloc(&p.hist, sizeof(unsigned) * SIGMA_BINS * B));
  HIPCHK(hipMalloc(&p.stats, sizeof(SigmaStats) * B));
  HIPCHK(hipMalloc(&p.sigma, sizeof(double) * B));
  HIPCHK(hipMalloc(&p.sig_part, sizeof(unsigned long long) * 4 * SIGMA_MAX_BLOCKS * B));
  HIPCHK(hipMalloc(&p.iphys, sizeof(InstPhys) * B));
  // partial rows: at most one per line of an instance, or one per elementwise block
  const long long rows = std::max<long long>(std::max<long long>(p.n / p.N, 1024), (long long)p.N + 8);
  p.max_rows = (int)rows;
  p.part_stride = 16LL * rows * B + 16 * BEC_XCDS;
  HIPCHK(hipMalloc(&p.partials, sizeof(double) * p.part_stride * (C_COUNT + 1)));
  HIPCHK(hipHostMalloc(&p.hpart, sizeof(double) * p.part_stride * C_COUNT,
                       hipHostMallocMapped | hipHostMallocCoherent));
  HIPCHK(hipHostGetDevicePointer((void**)&p.hpart_dev, p.hpart, 0));
  HIPCHK(hipHostMalloc(&p.hred, sizeof(double) * RSTRIDE * B, hipHostMallocMapped | hipHostMallocCoherent));
  memset(p.hred, 0, sizeof(double) * RSTRIDE * B);
  HIPCHK(hipHostGetDevicePointer((void**)&p.hred_dev, p.hred, 0));
  launch_init_sigma(p, nb, p.stream);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(p.stream));
  p.cap = nb;
}

void plan_set_phys(Plan& p, const Phys* ph, int nb) {
  HIPCHK(hipMemcpyAsync(p.iphys, ph, sizeof(InstPhys) * nb, hipMemcpyHostToDevice, p.stream));
  HIPCHK(hipStreamSynchronize(p.stream));
}

// Fold the pinned per-block rows of every queued reduction into hred, per
// active instance, rows in order (deterministic): sums for the first K
// values, maxima for the next KM.
static void fold_jobs(Plan& p) {
  for (int i = 0; i < p.pend.n; ++i) {
    const RedJob& j = p.pend.j[i];
    const int KT = j.K + j.KM;
    const int KP = KT <= 1 ? 1 : (KT <= 2 ? 2 : (KT <= 4 ? 4 : (KT <= 8 ? 8 : 16)));
    const double* rows = p.hpart + (size_t)j.site * p.part_stride;
    for (int b = 0; b < j.nb; ++b) {
      if (!((j.act >> b) & 1ull)) continue;
      const double* rb = rows + (size_t)b * j.nblk * KP;
      for (int q = 0; q < KT; ++q) {
        double s = (q < j.K) ? 0.0 : -INFINITY;
        for (int r = 0; r < j.nblk; ++r) {
          const double x = rb[(size_t)r * KP + q];
          s = (q < j.K) ? s + x : std::max(s, x);
        }
        p.hred[(size_t)b * RSTRIDE + j.slot + q] = s;
      }
    }
  }
  p.pend.n = 0;
}

void add_job(Plan& p, int site, int nblk, int K, int KM, int slot, const Batch& B) {
  for (int i = 0; i < p.pend.n; ++i)
    if (p.pend.j[i].site == site) throw BecError(2, "internal: reduction site reused before a sync");
  if (p.pend.n == BEC_MAX_JOBS) throw BecError(2, "internal: too many pending reductions");
  RedJob& j = p.pend.j[p.pend.n++];
  j.site = site;
  j.nblk = nblk;
  j.K = K;
  j.KM = KM;
  j.slot = slot;
  j.nb = B.nb;
  j.act = B.act;
}

void plan_sync(Plan& p) {
  HIPCHK(hipStreamSynchronize(p.stream));
  fold_jobs(p);
}

// ---------------------------------------------------------------------------
// drivers
// ---------------------------------------------------------------------------
static PassArgs pass_base(const Plan& p, const FftPlan& fp, const LineCfg& lc, const Batch& B) {
  PassArgs a;
  memset(&a, 0, sizeof(a));
  a.fp = fp;
  a.tpl = lc.tpl;
  a.lpw = lc.lpw;
  a.nb = B.nb;
  a.act = act_of(B);
  a.ni = 1;
  a.out_scale = 1.0;
  a.H = p.H;
  a.M = p.M;
  a.Nk = p.N;
  a.invNdx = p.invNdx;
  a.sigma = p.sigma;
  a.f32 = B.f32;
  a.lsh[0] = p.lsh[0];
  a.lsh[1] = p.lsh[1];
  return a;
}

int run_conv(Plan& p, int which, const Batch& B, const double* rho, double scale, double* out,
             const double* dot, const double* dot2, int site, int slot) {
  if (!p.khat[which]) throw BecError(1, "kernel spectrum not built for this plan");
  const int N = p.N, M = p.M, H = p.H, i0 = (p.M - p.N) / 2;
  const double fac = scale * p.dv / (double)p.nM;
  const long long n = p.n, S1 = p.nS1, S2 = p.nS2;
  int nb_rows = 0;
  if (p.ndim == 2) {
    PassArgs c1 = pass_base(p, p.fpM, p.lcM, B);
    c1.nl = N; c1.in_so = N; c1.in_s = 1; c1.in_ib = n; c1.n_load = N; c1.off_load = i0;
    c1.out_so = H; c1.out_s = 1; c1.out_ib = S1; c1.n_store = H; c1.off_store = 0;
    c1.in = rho; c1.out = p.s1;
    launch_pass(1, LD_R, DIR_FWD, MUL_NONE, ST_C, c1, p.stream);
    PassArgs c2 = pass_base(p, p.fpM, p.lcM, B);
    c2.nl = H; c2.ni = H; c2.in_si = 1; c2.in_s = H; c2.in_ib = S1; c2.n_load = N; c2.off_load = i0;
    c2.out_si = 1; c2.out_s = H; c2.out_ib = S1; c2.n_store = N; c2.off_store = i0;
    c2.in = p.s1; c2.out = p.s1; c2.khat = p.khat[which];
    launch_pass(0, LD_C, DIR_FMI, MUL_KHAT2, ST_C, c2, p.stream);
    PassArgs c3 = pass_base(p, p.fpM, p.lcM, B);
    c3.nl = N; c3.in_so = H; c3.in_s = 1; c3.in_ib = S1;
    c3.out_so = N; c3.out_s = 1; c3.out_ib = n; c3.n_store = N; c3.off_store = i0; c3.out_scale = fac;
    c3.in = p.s1; c3.out = out; c3.dot = dot; c3.dot2 = dot ? dot2 : nullptr;
    if (dot) (B.dev ? c3.partials : c3.hpartials) = B.dev ? part_of(p, site) : hpart_of(p, site);
    nb_rows = launch_pass(1, LD_HERM, DIR_INV, MUL_NONE, ST_RE, c3, p.stream);
  } else {
    const long long NH = (long long)N * H, MH = (long long)M * H;
    PassArgs c1 = pass_base(p, p.fpM, p.lcM, B);
    c1.nl = N * N; c1.in_so = N; c1.in_s = 1; c1.in_ib = n; c1.n_load = N; c1.off_load = i0;
    c1.out_so = H; c1.out_s = 1; c1.out_ib = S1; c1.n_store = H; c1.off_store = 0;
    c1.in = rho; c1.out = p.s1;
    launch_pass(1, LD_R, DIR_FWD, MUL_NONE, ST_C, c1, p.stream);
    PassArgs c2 = pass_base(p, p.fpM, p.lcM, B);
    c2.nl = (int)NH; c2.ni = H; c2.in_so = NH; c2.in_si = 1; c2.in_s = H; c2.in_ib = S1; c2.n_load = N;
    c2.off_load = i0;
    c2.out_so = MH; c2.out_si = 1; c2.out_s = H; c2.out_ib = S2; c2.n_store = M; c2.off_store = 0;
    c2.in = p.s1; c2.out = p.s2;
    launch_pass(0, LD_C, DIR_FWD, MUL_NONE, ST_C, c2, p.stream);
    PassArgs c3 = pass_base(p, p.fpM, p.lcM, B);
    c3.nl = (int)MH; c3.ni = (int)MH; c3.in_si = 1; c3.in_s = MH; c3.in_ib = S2; c3.n_load = N; c3.off_load = i0;
    c3.out_si = 1; c3.out_s = MH; c3.out_ib = S2; c3.n_store = N; c3.off_store = i0;
    c3.in = p.s2; c3.out = p.s2; c3.khat = p.khat[which];
    launch_pass(0, LD_C, DIR_FMI, MUL_KHAT3, ST_C, c3, p.stream);
    PassArgs c4 = pass_base(p, p.fpM, p.lcM, B);
    c4.nl = (int)NH; c4.ni = H; c4.in_so = MH; c4.in_si = 1; c4.in_s = H; c4.in_ib = S2; c4.n_load = M;
    c4.off_load = 0;
    c4.out_so = NH; c4.out_si = 1; c4.out_s = H; c4.out_ib = S1; c4.n_store = N; c4.off_store = i0;
    c4.in = p.s2; c4.out = p.s1;
    launch_pass(0, LD_C, DIR_INV, MUL_NONE, ST_C, c4, p.stream);
    PassArgs c5 = pass_base(p, p.fpM, p.lcM, B);
    c5.nl = N * N; c5.in_so = H; c5.in_s = 1; c5.in_ib = S1;
    c5.out_so = N; c5.out_s = 1; c5.out_ib = n; c5.n_store = N; c5.off_store = i0; c5.out_scale = fac;
    c5.in = p.s1; c5.out = out; c5.dot = dot; c5.dot2 = dot ? dot2 : nullptr;
    if (dot) (B.dev ? c5.partials : c5.hpartials) = B.dev ? part_of(p, site) : hpart_of(p, site);
    nb_rows = launch_pass(1, LD_HERM, DIR_INV, MUL_NONE, ST_RE, c5, p.stream);
  }
  if (dot && !B.dev) add_job(p, site, nb_rows, dot2 ? 2 : 1, 0, slot, B);
  return nb_rows;
}

// K*q and K*r as one complex convolution of q + i r: K is real and even, so
// the result's real and imaginary parts are the two convolutions.  The first
// pass loads the pair and keeps all M frequencies of its axis (no half
// spectrum), the middle passes run over those M columns (kfull), and the last
// one stores the real part to kq and the imaginary part to kr with the sums
// the polynomial line search needs.  Half the launches of two real
// convolutions, one read of the kernel spectrum, and no Hermitian handling.
int run_conv_pair(Plan& p, const Batch& B, const double* q, const double* r, double* kq, double* kr, int site_qr,
                  int site_rr) {
  if (!p.khat[0]) throw BecError(1, "kernel spectrum not built for this plan");
  if (!B.dev) throw BecError(2, "internal: pair convolutions run in the device-driven loop only");
  const int N = p.N, M = p.M, i0 = (p.M - p.N) / 2;
  const double fac = p.dv / (double)p.nM;
  const long long n = p.n, S1 = p.nS1, S2 = p.nS2;
  int nb_rows = 0;
  PassArgs last;
  if (p.ndim == 2) {
    // rows x: y lines (contiguous), the pair padded -> s1 [N][M]
    PassArgs c1 = pass_base(p, p.fpM, p.lcM, B);
    c1.nl = N; c1.in_so = N; c1.in_s = 1; c1.in_ib = n; c1.n_load = N; c1.off_load = i0;
    c1.out_so = M; c1.out_s = 1; c1.out_ib = S1; c1.n_store = M; c1.off_store = 0;
    c1.in = q; c1.in2 = r; c1.out = p.s1;
    launch_pass(1, LD_R2, DIR_FWD, MUL_NONE, ST_C, c1, p.stream);
    // the M columns ky: x lines (stride M), padded, x K^, cropped
    PassArgs c2 = pass_base(p, p.fpM, p.lcM, B);
    c2.nl = M; c2.ni = M; c2.in_si = 1; c2.in_s = M; c2.in_ib = S1; c2.n_load = N; c2.off_load = i0;
    c2.out_si = 1; c2.out_s = M; c2.out_ib = S1; c2.n_store = N; c2.off_store = i0;
    c2.in = p.s1; c2.out = p.s1; c2.khat = p.khat[0]; c2.kfull = 1;
    launch_pass(0, LD_C, DIR_FMI, MUL_KHAT2, ST_C, c2, p.stream);
    last = pass_base(p, p.fpM, p.lcM, B);
    last.nl = N; last.in_so = M; last.in_s = 1; last.in_ib = S1; last.n_load = M; last.off_load = 0;
  } else {
    const long long NM = (long long)N * M, MM = (long long)M * M;
    // z lines (contiguous), the pair padded -> s1 [x][y][kz], kz in [0, M)
    PassArgs c1 = pass_base(p, p.fpM, p.lcM, B);
    c1.nl = N * N; c1.in_so = N; c1.in_s = 1; c1.in_ib = n; c1.n_load = N; c1.off_load = i0;
    c1.out_so = M; c1.out_s = 1; c1.out_ib = S1; c1.n_store = M; c1.off_store = 0;
    c1.in = q; c1.in2 = r; c1.out = p.s1;
    launch_pass(1, LD_R2, DIR_FWD, MUL_NONE, ST_C, c1, p.stream);
    // y lines (x, kz), padded -> s2 [x][ky][kz]
    PassArgs c2 = pass_base(p, p.fpM, p.lcM, B);
    c2.nl = (int)NM; c2.ni = M; c2.in_so = NM; c2.in_si = 1; c2.in_s = M; c2.in_ib = S1; c2.n_load = N;
    c2.off_load = i0;
    c2.out_so = MM; c2.out_si = 1; c2.out_s = M; c2.out_ib = S2; c2.n_store = M; c2.off_store = 0;
    c2.in = p.s1; c2.out = p.s2; c2.kfull = 1;
    launch_pass(0, LD_C, DIR_FWD, MUL_NONE, ST_C, c2, p.stream);
    // x lines (ky, kz): padded, x K^, cropped
    PassArgs c3 = pass_base(p, p.fpM, p.lcM, B);
    c3.nl = (int)MM; c3.ni = (int)MM; c3.in_si = 1; c3.in_s = MM; c3.in_ib = S2; c3.n_load = N; c3.off_load = i0;
    c3.out_si = 1; c3.out_s = MM; c3.out_ib = S2; c3.n_store = N; c3.off_store = i0;
    c3.in = p.s2; c3.out = p.s2; c3.khat = p.khat[0]; c3.kfull = 1;
    launch_pass(0, LD_C, DIR_FMI, MUL_KHAT3, ST_C, c3, p.stream);
    // y lines back (x, kz), cropped -> s1
    PassArgs c4 = pass_base(p, p.fpM, p.lcM, B);
    c4.nl = (int)NM; c4.ni = M; c4.in_so = MM; c4.in_si = 1; c4.in_s = M; c4.in_ib = S2; c4.n_load = M;
    c4.off_load = 0;
    c4.out_so = NM; c4.out_si = 1; c4.out_s = M; c4.out_ib = S1; c4.n_store = N; c4.off_store = i0;
    c4.in = p.s2; c4.out = p.s1; c4.kfull = 1;
    launch_pass(0, LD_C, DIR_INV, MUL_NONE, ST_C, c4, p.stream);
    last = pass_base(p, p.fpM, p.lcM, B);
    last.nl = N * N; last.in_so = M; last.in_s = 1; last.in_ib = S1; last.n_load = M; last.off_load = 0;
  }
  // the contiguous axis back, cropped: K*q = re, K*r = im, with their sums
  last.out_so = N; last.out_s = 1; last.out_ib = n; last.n_store = N; last.off_store = i0; last.out_scale = fac;
  last.in = p.s1; last.out = kq; last.out2 = kr; last.dot = q; last.dot2 = r;
  last.partials = part_of(p, site_qr);
  last.partials2 = part_of(p, site_rr);
  nb_rows = launch_pass(1, LD_C, DIR_INV, MUL_NONE, ST_RE2, last, p.stream);
  return nb_rows;
}

static KinArgs kin_base(const Plan& p, const Batch& B) {
  KinArgs a;
  memset(&a, 0, sizeof(a));
  a.fp = p.fpN;
  a.tpl = p.lcN.tpl;
  a.lpw = p.lcN.lpw;
  a.N = p.N;
  a.f32 = B.f32;
  a.lsh[0] = p.lsh[0];
  a.lsh[1] = p.lsh[1];
  a.nl = (int)(p.n / p.N);
  a.nb = B.nb;
  a.act = act_of(B);
  a.istride = p.n;
  a.dx = p.dx;
  a.invNdx = p.invNdx;
  a.tc = p.tc;
  a.s_mode = p.s_mode;
  a.s_half = p.s_half;
  a.iphys = p.iphys;
  for (int b = 0; b < BEC_BMAX; ++b) { a.t[b] = 0.0; a.sc[b] = 1.0; a.mu[b] = 0.0; a.phisc[b] = 1.0; }
  return a;
}

static const int kAxes2[2] = {A2Y, A2X}, kAxes3[3] = {A3Z, A3Y, A3X};

EnergyRows run_energy_sums(Plan& p, const Batch& B, const double2* psi, const double2* d, const double* t,
                           double* rho_out, const EnergyFuse* fuse) {
  const int* ax = p.ndim == 2 ? kAxes2 : kAxes3;
  EnergyRows er{0, 0, 0};
  for (int k = 0; k < p.ndim; ++k) {
    KinArgs a = kin_base(p, B);
    a.psi = psi;
    a.d = d;
    if (t)
      for (int b = 0; b < B.nb; ++b) a.t[b] = t[b];
    if (B.dev && d) {
      a.tab = B.tab;
      a.tabm = 1u << DS_T;
    }
    const int site = (k == 0) ? C_EFIRST : (k == 1 ? C_EREST0 : C_EREST1);
    if (B.dev) a.partials = part_of(p, site);
    else a.hpartials = hpart_of(p, site);
    if (k == 0) a.rho_out = rho_out;
    if (k == 0 && fuse) {
      a.hy_out = fuse->hy_out;
      a.kf = 0.5;
      a.rf_omega = 1;
      a.z = fuse->z;
      a.g = fuse->g;
      a.d_out = fuse->d_out;
      a.comb_bit = fuse->comb_bit;
    }
    const int rows = launch_kin(ax[k], k == 0 ? KE_FIRST : KE_REST, a, p.stream);
    (k == 0 ? er.r0 : (k == 1 ? er.r1 : er.r2)) = rows;
    if (B.dev) continue;
    if (k == 0) add_job(p, C_EFIRST, rows, 6, 0, R_EFIRST, B);
    else add_job(p, site, rows, 2, 0, R_EREST + 2 * (k - 1), B);
  }
  return er;
}

EnergyRows run_line_sums(Plan& p, const Batch& B, const double2* psi, double2* d, const double2* z,
                         const double2* g, const double* phi, double* q_out, double* r_out) {
  const int* ax = p.ndim == 2 ? kAxes2 : kAxes3;
  EnergyRows er{0, 0, 0};
  for (int k = 0; k < p.ndim; ++k) {
    KinArgs a = kin_base(p, B);
    if (k == 0) {
      a.psi = psi;
      a.d = d;
      a.z = z;
      a.g = g;
      a.d_out = d;
      a.comb_bit = FL_COMB;
      a.tab = B.tab;
      a.phi = phi;
      a.rho_out = q_out;
      a.r_out = r_out;
      a.partials = part_of(p, C_LINE);
      er.r0 = launch_kin(ax[0], KL_FIRST, a, p.stream);
    } else {
      a.psi = d;                    // the direction alone: its kinetic and rotation forms
      a.partials = part_of(p, k == 1 ? C_EREST0 : C_EREST1);
      (k == 1 ? er.r1 : er.r2) = launch_kin(ax[k], KE_REST, a, p.stream);
    }
  }
  return er;
}

Energy energy_from_sums(const Plan& p, int b, const Phys& ph, bool lr, bool normalise, double gscale) {
  const double* r = hres(p, b);
  double kin = r[R_EFIRST + 0], lz = r[R_EFIRST + 1];
  for (int k = 1; k < p.ndim; ++k) {
    kin += r[R_EREST + 2 * (k - 1)];
    lz += r[R_EREST + 2 * (k - 1) + 1];
  }
  const double dv = p.dv;
  Energy e;
  e.nrm2 = r[R_EFIRST + 2] * dv;
  const double i1 = normalise ? 1.0 / e.nrm2 : 1.0;
  const double i2 = i1 * i1, i3 = i2 * i1;
  e.Ekin = kin * dv * i1;
  e.Etrap = r[R_EFIRST + 3] * dv * i1;
  e.Econtact = 0.5 * ph.beta2 * (r[R_EFIRST + 4] * dv) * i2;
  e.Ethree = (ph.beta3 / 3.0) * (r[R_EFIRST + 5] * dv) * i3;
  e.Lz = lz * dv * i1;
  e.Erot = -ph.Omega * e.Lz;
  e.Egrav = lr ? 0.5 * gscale * (r[R_EGRAV] * dv) * i2 : 0.0;
  e.E = e.Ekin + e.Etrap + e.Econtact + e.Ethree + e.Erot + e.Egrav;
  return e;
}

int run_H(Plan& p, const Batch& B, const HArgs& h) {
  int last_rows = 0;
  const int* ax = p.ndim == 2 ? kAxes2 : kAxes3;
  const bool transformed = h.d || h.div;   // (device-driven: the step and norm come from the table)
  if (transformed && !h.psi_w) throw BecError(2, "internal: a transformed state must be written back");
  const double2* psi_rest = h.psi_w ? h.psi_w : h.psi;
  for (int k = 0; k < p.ndim; ++k) {
    KinArgs a = kin_base(p, B);
    a.kf = h.kf;
    a.rf = h.rf;
    a.rf_omega = h.rf_omega;
    const bool last = (k == p.ndim - 1);
    if (k == 0) {
      a.psi = h.psi;
      a.d = h.d;
      for (int b = 0; b < B.nb; ++b) {
        if (h.t) a.t[b] = h.t[b];
        if (h.div) a.sc[b] = h.div[b];
      }
      a.psi_w = h.psi_w;
      a.sdiff = h.sdiff;
      a.out = last ? h.out : p.w1;
      if (B.dev && h.d) {              // accepting pass: (psi - t d) / ||psi - t d|| from the table
        a.tab = B.tab;
        a.tabm = (1u << DS_T) | (1u << DS_SC);
      }
    } else {
      a.psi = psi_rest;
      a.acc = p.w1;
      a.out = last ? h.out : p.w1;
    }
    if (last) {
      a.local_on = h.local_on ? 1 : 0;
      if (!h.local_on) a.s_mode = 0;
      a.phi = h.local_on ? h.phi : nullptr;
      a.phi_out = h.phi_out;
      a.kq = h.kq;
      a.kr = h.kr;
      for (int b = 0; b < B.nb; ++b) {
        if (h.phisc) a.phisc[b] = h.phisc[b];
        if (h.mu) a.mu[b] = h.mu[b];
      }
      a.ydiff = h.ydiff;
      a.partials = h.reduce ? part_of(p, C_HLAST) : nullptr;
      a.hpartials = (h.reduce && !B.dev) ? hpart_of(p, C_HLAST) : nullptr;
      if (B.dev) {
        a.tab = B.tab;
        a.tabm |= (h.mu ? 0u : 1u << DS_MU) | (h.phisc ? 0u : 1u << DS_PHISC);
      }
    }
    const int kind = (k == 0) ? KH_FIRST : (last ? KH_LAST : KH_MID);
    const int rows = launch_kin(ax[k], kind, a, p.stream);
    if (last) last_rows = rows;
    if (last && h.reduce) {
      if (!B.dev) add_job(p, C_HLAST, rows, 2, 1, R_HLAST, B);
      p.hlast_rows = rows;
    }
  }
  return last_rows;
}

int run_H_accept_dev(Plan& p, const Batch& B, const HArgs& h) {
  const int* ax = p.ndim == 2 ? kAxes2 : kAxes3;
  int rows = 0;
  for (int k = 1; k < p.ndim; ++k) {
    const bool last = (k == p.ndim - 1);
    KinArgs a = kin_base(p, B);
    a.kf = 0.5;
    a.rf_omega = 1;
    a.psi = h.psi;
    a.d = h.d;
    a.tab = B.tab;
    a.tabm = (1u << DS_T) | (1u << DS_SC);
    a.acc = p.w1;
    a.acc_sc = (k == 1) ? 1 : 0;     // the trial's share is of the unnormalised state
    if (last) {
      a.out = h.out;
      a.psi_w = h.psi_w;
      a.sdiff = h.sdiff;
      a.local_on = 1;
      a.phi = h.phi;
      a.phi_out = h.phi_out;
      a.tabm |= (1u << DS_MU) | (1u << DS_PHISC);
      a.ydiff = h.ydiff;
      a.partials = part_of(p, C_HLAST);
    } else {
      a.out = p.w1;
    }
    rows = launch_kin(ax[k], last ? KH_LAST : KH_MID, a, p.stream);
  }
  p.hlast_rows = rows;
  return rows;
}

static int last_precond_pass(Plan& p, PassArgs c, const DotFields* dots) {
  if (!dots) return (launch_pass(1, LD_C, DIR_INV, MUL_NONE, ST_C, c, p.stream), 0);
  c.dpsi = dots->psi; c.dd = dots->d; c.dg = dots->g; c.dy = dots->y; c.ds = dots->s;
  c.partials = part_of(p, C_DOTS);
  return launch_pass(1, LD_C, DIR_INV, MUL_NONE, ST_CDOT, c, p.stream);
}

int run_precond(Plan& p, const Batch& B, const double2* g, double2* out, const DotFields* dots) {
  const int N = p.N;
  const long long n = p.n;
  const double inv_n = 1.0 / (double)p.n;
  if (p.ndim == 2) {
    PassArgs a = pass_base(p, p.fpN, p.lcN, B);
    a.nl = N; a.in_so = N; a.in_s = 1; a.in_ib = n; a.n_load = N; a.out_so = N; a.out_s = 1; a.out_ib = n;
    a.n_store = N; a.in = g; a.out = p.w1;
    launch_pass(1, LD_C, DIR_FWD, MUL_NONE, ST_C, a, p.stream);
    PassArgs b = pass_base(p, p.fpN, p.lcN, B);
    b.nl = N; b.ni = N; b.in_si = 1; b.in_s = N; b.in_ib = n; b.n_load = N; b.out_si = 1; b.out_s = N;
    b.out_ib = n; b.n_store = N; b.in = p.w1; b.out = p.w1;
    launch_pass(0, LD_C, DIR_FMI, MUL_PRE2, ST_C, b, p.stream);
    PassArgs c = a;
    c.in = p.w1; c.out = out; c.out_scale = inv_n;
    return last_precond_pass(p, c, dots);
  } else {
    const long long N2 = (long long)N * N;
    PassArgs z = pass_base(p, p.fpN, p.lcN, B);
    z.nl = (int)N2; z.in_so = N; z.in_s = 1; z.in_ib = n; z.n_load = N; z.out_so = N; z.out_s = 1; z.out_ib = n;
    z.n_store = N; z.in = g; z.out = p.w1;
    launch_pass(1, LD_C, DIR_FWD, MUL_NONE, ST_C, z, p.stream);
    PassArgs y = pass_base(p, p.fpN, p.lcN, B);
    y.nl = (int)N2; y.ni = N; y.in_so = N2; y.in_si = 1; y.in_s = N; y.in_ib = n; y.n_load = N;
    y.out_so = N2; y.out_si = 1; y.out_s = N; y.out_ib = n; y.n_store = N; y.in = p.w1; y.out = p.w1;
    launch_pass(0, LD_C, DIR_FWD, MUL_NONE, ST_C, y, p.stream);
    PassArgs x = pass_base(p, p.fpN, p.lcN, B);
    x.nl = (int)N2; x.ni = (int)N2; x.in_si = 1; x.in_s = N2; x.in_ib = n; x.n_load = N;
    x.out_si = 1; x.out_s = N2; x.out_ib = n; x.n_store = N; x.in = p.w1; x.out = p.w1;
    launch_pass(0, LD_C, DIR_FMI, MUL_PRE3, ST_C, x, p.stream);
    launch_pass(0, LD_C, DIR_INV, MUL_NONE, ST_C, y, p.stream);
    PassArgs zi = z;
    zi.in = p.w1; zi.out = out; zi.out_scale = inv_n;
    return last_precond_pass(p, zi, dots);
  }
}

// ---------------------------------------------------------------------------
// plan lifecycle
// ---------------------------------------------------------------------------
static void plan_free(Plan& p) {
  solver_release(p);
  free_batch(p);
  if (p.tfflag) (void)hipHostFree(p.tfflag);
  p.tfflag = nullptr;
  p.tfflag_dev = nullptr;
  void* ptrs[] = {p.twN, p.twM, p.khat[0], p.khat[1], p.tfbuf, p.dgbuf};
  for (void* q : ptrs)
    if (q) (void)hipFree(q);
}

extern "C" int bec_plan_create(int ndim, int N, double L, int pad, double s, double trap_coeff,
                               bec_plan** out) {
  bec_plan* h = nullptr;
  try {
    if (!out) bad("null output pointer");
    *out = nullptr;
    if (ndim != 2 && ndim != 3) bad("ndim must be 2 or 3");
    if (!(L > 0.0)) bad("L must be positive");
    if (pad < 2) pad = 2;
    h = new bec_plan();
    Plan& p = h->p;
    memset(&p, 0, sizeof(Plan));
    p.ndim = ndim;
    p.N = N;
    p.pad = pad;
    p.M = pad * N;
    p.H = p.M / 2 + 1;
    if (p.M % 2) bad("padded length must be even");
    p.n = (ndim == 2) ? (long long)N * N : (long long)N * N * N;
    p.nM = (ndim == 2) ? (long long)p.M * p.M : (long long)p.M * p.M * p.M;
    // convolution spectra per instance, sized for a pair convolution (all M
    // frequencies of the first transformed axis; a single real field's
    // convolution uses H = M/2 + 1 of them)
    p.nS1 = (ndim == 2) ? (long long)N * p.M : (long long)N * N * p.M;
    p.nS2 = (ndim == 2) ? 0 : (long long)N * p.M * p.M;
    p.Lbox = L;
    p.dx = 2.0 * L / N;
    p.dv = (ndim == 2) ? p.dx * p.dx : p.dx * p.dx * p.dx;
    p.s = s;
    p.tc = trap_coeff;
    p.s_half = s / 2.0;
    p.s_mode = (s > 0) ? (p.s_half == 1.0 ? 1 : (p.s_half == 2.0 ? 2 : 3)) : 0;
    p.invNdx = 1.0 / (N * p.dx);
    p.stream = 0;
    make_fft(N, p.fpN, &p.twN, p.stream);
    // the padded transforms exist only for M <= BEC_MAX_L; a larger grid
    // solves without long range and refuses a kernel spectrum (conv_check)
    if (p.M <= BEC_MAX_L) make_fft(p.M, p.fpM, &p.twM, p.stream);
    // batched reducing passes need one instance's lines (n / N of them) to fill
    // whole workgroups; keep the power-of-two thread count where L/8 would not
    const int nlk = (int)(p.n / N);
    if (p.fpN.reg && nlk % lpw_min(N / 8) != 0) p.fpN.reg = 0;
    if (p.fpM.reg && nlk % lpw_min(p.M / 8) != 0) p.fpM.reg = 0;
    // a line of L/8 threads that leaves one instance's lines short of a
    // whole-wavefront workgroup (2D N = 16: 16 lines of 2 threads) runs
    // through LDS on about L/4 threads, which fills one (the register and
    // compile-time plans measured faster wherever they batch: 2D N = 128
    // ran 18% more CG iterations/s on them than in this form)
    if (ndim == 2 && !p.fpN.reg && nlk % lpw_min(line_cfg(p.fpN).tpl) != 0) {
      p.fpN.wide = 1;
      if (nlk % lpw_min(line_cfg(p.fpN).tpl) != 0) p.fpN.wide = 0;
    }
    p.lcN = line_cfg(p.fpN);
    p.lcM = p.twM ? line_cfg(p.fpM) : p.lcN;
    plan_reserve(p, 1);
    *out = h;
    return 0;
  } catch (const BecError& e) {
    g_err = e.what();
    if (h) { plan_free(h->p); delete h; }
    return e.code;
  } catch (const std::exception& e) {
    g_err = e.what();
    if (h) { plan_free(h->p); delete h; }
    return 2;
  }
}

extern "C" int bec_plan_destroy(bec_plan* h) {
  BEC_TRY({
    if (!h) return 0;
    (void)hipStreamSynchronize(h->p.stream);
    plan_free(h->p);
    delete h;
  });
}

extern "C" int bec_plan_set_stream(bec_plan* h, void* stream) {
  BEC_TRY({
    if (!h) bad("null plan");
    h->p.stream = (hipStream_t)stream;
  });
}

// forward FFT of a full M^d complex array in place (kernel spectra)
static void fft_full_M(Plan& p, double2* a) {
  const int M = p.M;
  const Batch B = one_instance();
  if (p.ndim == 2) {
    PassArgs r = pass_base(p, p.fpM, p.lcM, B);
    r.nl = M; r.in_so = M; r.in_s = 1; r.n_load = M; r.out_so = M; r.out_s = 1; r.n_store = M;
    r.in = a; r.out = a;
    launch_pass(1, LD_C, DIR_FWD, MUL_NONE, ST_C, r, p.stream);
    PassArgs c = pass_base(p, p.fpM, p.lcM, B);
    c.nl = M; c.ni = M; c.in_si = 1; c.in_s = M; c.n_load = M; c.out_si = 1; c.out_s = M;
    c.n_store = M; c.in = a; c.out = a;
    launch_pass(0, LD_C, DIR_FWD, MUL_NONE, ST_C, c, p.stream);
  } else {
    const long long M2 = (long long)M * M;
    PassArgs z = pass_base(p, p.fpM, p.lcM, B);
    z.nl = (int)M2; z.in_so = M; z.in_s = 1; z.n_load = M; z.out_so = M; z.out_s = 1; z.n_store = M;
    z.in = a; z.out = a;
    launch_pass(1, LD_C, DIR_FWD, MUL_NONE, ST_C, z, p.stream);
    PassArgs y = pass_base(p, p.fpM, p.lcM, B);
    y.nl = (int)M2; y.ni = M; y.in_so = M2; y.in_si = 1; y.in_s = M; y.n_load = M;
    y.out_so = M2; y.out_si = 1; y.out_s = M; y.n_store = M; y.in = a; y.out = a;
    launch_pass(0, LD_C, DIR_FWD, MUL_NONE, ST_C, y, p.stream);
    PassArgs x = pass_base(p, p.fpM, p.lcM, B);
    x.nl = (int)M2; x.ni = (int)M2; x.in_si = 1; x.in_s = M2; x.n_load = M;
    x.out_si = 1; x.out_s = M2; x.n_store = M; x.in = a; x.out = a;
    launch_pass(0, LD_C, DIR_FWD, MUL_NONE, ST_C, x, p.stream);
  }
}

static void finish_kernel(Plan& p, int which, double2* full) {
  fft_full_M(p, full);
  const long long nh = (p.ndim == 2) ? (long long)p.H * p.H : (long long)p.H * p.H * p.H;
  if (!p.khat[which]) HIPCHK(hipMalloc(&p.khat[which], sizeof(double) * nh));
  launch_extract_khat(p.ndim, p.M, p.H, full, p.khat[which], p.stream);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(p.stream));
}

// exact self-cell averages (interactions.py:37-43, 130-133, 146)
static const double SELF_NEWTON_3D = 2.3800773640;

// long-range kernels need the padded transforms (pad N <= BEC_MAX_L)
static void conv_check(const Plan& p) {
  if (!p.twM)
    bad("the padded length pad*N = " + std::to_string(p.M) + " exceeds " + std::to_string(BEC_MAX_L) +
        ": a long-range kernel needs pad*N <= " + std::to_string(BEC_MAX_L));
}

extern "C" int bec_plan_build_kernel(bec_plan* h, int which, int kind) {
  BEC_TRY({
    if (!h) bad("null plan");
    Plan& p = h->p;
    if (which != BEC_WHICH_K && which != BEC_WHICH_W) bad("which must be K or W");
    if (kind != BEC_KERNEL_NEWTON && kind != BEC_KERNEL_LOG) bad("device-built kernels: newton or log");
    if (kind == BEC_KERNEL_LOG && p.ndim != 2) bad("kernel='log' is two-dimensional only");
    conv_check(p);
    double2* full = nullptr;
    HIPCHK(hipMalloc(&full, sizeof(double2) * p.nM));
    try {
      launch_kernel_real(p.ndim, p.M, p.dx, kind, full, p.stream);
      double self;
      if (p.ndim == 3) self = SELF_NEWTON_3D / p.dx;
      else if (kind == BEC_KERNEL_NEWTON) self = 4.0 * asinh(1.0) / p.dx;
      else self = -log(p.dx) + (1.5 + 0.5 * log(2.0) - 0.25 * M_PI);
      HIPCHK(hipStreamSynchronize(p.stream));
      const double2 sc = make_double2(self, 0.0);
      HIPCHK(hipMemcpy(full, &sc, sizeof(double2), hipMemcpyHostToDevice));
      finish_kernel(p, which, full);
    } catch (...) {
      (void)hipFree(full);
      throw;
    }
    HIPCHK(hipFree(full));
    p.kkind[which] = kind;
  });
}

__global__ void k_real_to_complex(const double* in, double2* out, long long n) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    out[i] = make_double2(in[i], 0.0);
}

extern "C" int bec_plan_set_kernel_table(bec_plan* h, int which, const double* ker_real) {
  BEC_TRY({
    if (!h || !ker_real) bad("null argument");
    Plan& p = h->p;
    if (which != BEC_WHICH_K && which != BEC_WHICH_W) bad("which must be K or W");
    conv_check(p);
    double2* full = nullptr;
    HIPCHK(hipMalloc(&full, sizeof(double2) * p.nM));
    try {
      hipLaunchKernelGGL(k_real_to_complex, dim3(2048), dim3(256), 0, p.stream, ker_real, full, p.nM);
      finish_kernel(p, which, full);
    } catch (...) {
      (void)hipFree(full);
      throw;
    }
    HIPCHK(hipFree(full));
    p.kkind[which] = BEC_KERNEL_TABLE;
  });
}

extern "C" int bec_plan_clear_kernels(bec_plan* h) {
  BEC_TRY({
    if (!h) bad("null plan");
    Plan& p = h->p;
    HIPCHK(hipStreamSynchronize(p.stream));
    for (int w = 0; w < 2; ++w) {
      if (p.khat[w]) HIPCHK(hipFree(p.khat[w]));
      p.khat[w] = nullptr;
      p.kkind[w] = 0;
    }
  });
}

// ---------------------------------------------------------------------------
// single-instance operator entry points
// ---------------------------------------------------------------------------
static Phys phys_of(const bec_phys* ph) {
  if (!ph) bad("null physics parameters");
  Phys q;
  q.beta2 = ph->beta2;
  q.beta3 = ph->beta3;
  q.Omega = ph->Omega;
  q.G_C = ph->G_C;
  return q;
}

extern "C" int bec_norm2(bec_plan* h, const void* psi, double* sum_out) {
  BEC_TRY({
    if (!h || !psi || !sum_out) bad("null argument");
    Plan& p = h->p;
    const Batch B = one_instance();
    add_job(p, C_NORM, launch_norm2((const double2*)psi, p.n, 1, 1ull, hpart_of(p, C_NORM), p.stream), 1, 0,
            R_NORM, B);
    HIPCHK(hipGetLastError());
    plan_sync(p);
    *sum_out = hres(p, 0)[R_NORM];
  });
}

extern "C" int bec_conv(bec_plan* h, int which, const double* rho, double scale, double* out) {
  BEC_TRY({
    if (!h || !rho || !out) bad("null argument");
    if (which != BEC_WHICH_K && which != BEC_WHICH_W) bad("which must be K or W");
    Plan& p = h->p;
    run_conv(p, which, one_instance(), rho, scale, out, nullptr);
    HIPCHK(hipGetLastError());
  });
}

extern "C" int bec_energy(bec_plan* h, const void* psi, const bec_phys* php, double* comps,
                          double* phi_out) {
  BEC_TRY({
    if (!h || !psi || !comps) bad("null argument");
    Plan& p = h->p;
    const Phys ph = phys_of(php);
    const bool lr = plan_long_range(p, ph);
    const Batch B = one_instance();
    run_energy_sums(p, B, (const double2*)psi, nullptr, nullptr, lr ? p.rho : nullptr);
    if (lr) {
      double* phi = phi_out ? phi_out : p.loc;
      run_conv(p, 0, B, p.rho, -ph.G_C, phi, p.rho);
    } else if (phi_out) {
      HIPCHK(hipMemsetAsync(phi_out, 0, sizeof(double) * p.n, p.stream));
    }
    HIPCHK(hipGetLastError());
    plan_sync(p);
    const Energy e = energy_from_sums(p, 0, ph, lr, false, 1.0);
    const double v[8] = {e.E, e.Ekin, e.Etrap, e.Econtact, e.Ethree, e.Erot, e.Egrav, e.Lz};
    memcpy(comps, v, sizeof(v));
  });
}

// H psi for one instance: phi (nullable) is the potential as given; when
// absent and the kernel is on it is computed into scratch
static void apply_H_one(Plan& p, const Phys& ph, const double2* psi, const double* phi_in, double2* out,
                        bool reduce) {
  const Batch B = one_instance();
  plan_set_phys(p, &ph, 1);
  const bool lr = plan_long_range(p, ph);
  const double* phi = phi_in;
  if (!phi && lr) {
    launch_rho(psi, p.rho, p.n, 1, p.stream);
    run_conv(p, 0, B, p.rho, -ph.G_C, p.loc, nullptr);
    phi = p.loc;
  }
  HArgs a;
  memset(&a, 0, sizeof(a));
  a.psi = psi;
  a.local_on = true;
  a.phi = phi;
  a.kf = 0.5;
  a.rf_omega = 1;
  a.out = out;
  a.reduce = reduce;
  run_H(p, B, a);
}

extern "C" int bec_apply_H(bec_plan* h, const void* psi, const bec_phys* php, const double* phi,
                           void* out) {
  BEC_TRY({
    if (!h || !psi || !out) bad("null argument");
    Plan& p = h->p;
    apply_H_one(p, phys_of(php), (const double2*)psi, phi, (double2*)out, false);
    HIPCHK(hipGetLastError());
  });
}

extern "C" int bec_Lz(bec_plan* h, const void* psi, void* out) {
  BEC_TRY({
    if (!h || !psi || !out) bad("null argument");
    Plan& p = h->p;
    HArgs a;
    memset(&a, 0, sizeof(a));
    a.psi = (const double2*)psi;
    a.local_on = false;
    a.kf = 0.0;
    a.rf = 1.0;
    a.out = (double2*)out;
    run_H(p, one_instance(), a);
    HIPCHK(hipGetLastError());
  });
}

extern "C" int bec_residual(bec_plan* h, const void* psi, const bec_phys* php, double* out3) {
  BEC_TRY({
    if (!h || !psi || !out3) bad("null argument");
    Plan& p = h->p;
    const Phys ph = phys_of(php);
    // h = H psi (mu = 0) into the conv scratch, Re <psi, H psi> from the fused sums
    double2* hbuf = p.s1;
    if (p.nS1 < p.n) throw BecError(2, "internal: scratch too small");
    apply_H_one(p, ph, (const double2*)psi, nullptr, hbuf, true);
    plan_sync(p);
    const double mu = hres(p, 0)[R_HLAST + 1] * p.dv;
    add_job(p, C_RES2, launch_res2(hbuf, (const double2*)psi, mu, nullptr, p.n, hpart_of(p, C_RES2), p.stream),
            1, 0, R_RES2, one_instance());
    HIPCHK(hipGetLastError());
    plan_sync(p);
    const double resid = sqrt(hres(p, 0)[R_RES2] * p.dv);
    out3[0] = resid;
    out3[1] = mu;
    out3[2] = resid / std::max(1.0, fabs(mu));
  });
}

extern "C" int bec_precondition(bec_plan* h, const void* g, const void* psi, const bec_phys* php,
                                const double* phi, double shift_min, void* out, double* sigma_out) {
  BEC_TRY({
    if (!h || !g || !psi || !out) bad("null argument");
    Plan& p = h->p;
    const Phys ph = phys_of(php);
    const Batch B = one_instance();
    plan_set_phys(p, &ph, 1);
    const bool lr = plan_long_range(p, ph);
    const double* phiu = nullptr;
    if (lr) {
      if (phi) phiu = phi;
      else {   // the potential goes to rho: the shift selection uses loc and cand
        launch_rho((const double2*)psi, p.rho, p.n, 1, p.stream);
        run_conv(p, 0, B, p.rho, -ph.G_C, p.rho, nullptr);
        phiu = p.rho;
      }
    }
    // max rho: the sigma pass folds the per-block maxima of the moments pass
    const int nbm = launch_moments((const double2*)psi, p.ndim, p.N, p.dx, part_of(p, C_MOM), hpart_of(p, C_MOM),
                                   p.stream);
    add_job(p, C_MOM, nbm, 4, 1, R_MOM, B);
    launch_sigma(p, 1, mask_act(1ull), (const double2*)psi, phiu, part_of(p, C_MOM), nbm, 8, 4, shift_min,
                 p.stream);
    double2* z = (double2*)out;
    run_precond(p, B, (const double2*)g, z);
    // projection onto the tangent space: z - psi <psi, z> dv
    add_job(p, C_DOTS, launch_dots(z, (const double2*)psi, z, z, z, z, p.n, 1, mask_act(1ull), nullptr,
                                   hpart_of(p, C_DOTS), p.stream),
            16, 0, R_DOTS, B);
    plan_sync(p);
    const double* r = hres(p, 0);
    CombArgs cb;
    memset(&cb, 0, sizeof(cb));
    cb.az[0] = 1.0;
    cb.gre[0] = -r[R_DOTS + 0] * p.dv;
    cb.gim[0] = -r[R_DOTS + 1] * p.dv;
    launch_combine(z, z, nullptr, (const double2*)psi, cb, nullptr, p.n, 1, mask_act(1ull), p.stream);
    HIPCHK(hipGetLastError());
    plan_sync(p);
    if (sigma_out) *sigma_out = hres(p, 0)[R_SIGMA];
  });
}

// Thomas-Fermi chemical potential with unit norm (thomasfermi.py:48-62): hi
// is widened while the norm at hi is below 1 (at most 200 times: NaN), then
// `steps` bisection steps run on the device without a host round trip.
// Diagnostics over a batch of nb states of one grid (fields [b][n]): every
// kernel launch covers all the states, their rows are copied to the host once
// per launch and folded per state in row order.
#define DIAG_ROWS (1024 * 16)
static double* diag_rows_dev(Plan& p) { return p.dgbuf; }
static double* diag_prm_dev(Plan& p) { return p.dgbuf + (size_t)DIAG_ROWS * BEC_BMAX; }
static void diag_alloc(Plan& p) {
  if (!p.dgbuf) HIPCHK(hipMalloc(&p.dgbuf, sizeof(double) * ((size_t)DIAG_ROWS * BEC_BMAX + 4 * BEC_BMAX)));
}
// rows of a diagnostics launch over nb states (g rows of kp values each)
static std::vector<double> diag_rows(Plan& p, int g, int kp, int nb) {
  std::vector<double> r((size_t)g * kp * nb);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(r.data(), p.dgbuf, sizeof(double) * r.size(), hipMemcpyDeviceToHost, p.stream));
  HIPCHK(hipStreamSynchronize(p.stream));
  return r;
}
static void diag_params(Plan& p, const std::vector<double>& prm) {
  HIPCHK(hipMemcpyAsync(diag_prm_dev(p), prm.data(), sizeof(double) * prm.size(), hipMemcpyHostToDevice,
                        p.stream));
}
static void check_batch(int nb) {
  if (nb < 1 || nb > BEC_BMAX) bad("batch size must be in [1, " + std::to_string(BEC_BMAX) + "]");
}

// per state: sum rho, r^2 rho, (x^2 + y^2) rho, z^2 rho, max rho (before dV)
static void moments_batch(Plan& p, int nb, const double2* psi, double* out5) {
  plan_reserve(p, nb);
  const Batch B{nb, nb >= 64 ? ~0ull : ((1ull << nb) - 1ull)};
  add_job(p, C_MOM, launch_moments(psi, p.ndim, p.N, p.dx, nullptr, hpart_of(p, C_MOM), p.stream, nb), 4, 1, R_MOM,
          B);
  HIPCHK(hipGetLastError());
  plan_sync(p);
  for (int b = 0; b < nb; ++b) memcpy(out5 + 5 * b, hres(p, b) + R_MOM, sizeof(double) * 5);
}

// Weight in the lowest Landau level (vortex.py:61-84) of each state: the
// orbitals m = 0 .. Mcap (Mcap from the state's <r^2> and L), five per launch
// for every state at once.
static void lll_batch(Plan& p, int nb, const double2* psi, const double* Omega, int Mmax, double* w_out,
                      int* mcap_out) {
  if (p.ndim != 2) bad("the LLL weight is two-dimensional only");
  diag_alloc(p);
  std::vector<double> mom(5 * (size_t)nb);
  moments_batch(p, nb, psi, mom.data());
  std::vector<int> mcap(nb);
  int mmax = 0;
  for (int b = 0; b < nb; ++b) {
    const double nrm2 = mom[5 * b];
    const double r2mean = nrm2 > 0.0 ? mom[5 * b + 1] / nrm2 : 0.0;
    const int est = (int)ceil(r2mean + 10.0 * sqrt(std::max(r2mean, 1.0)) + 30.0);
    const int cap2 = std::max(Mmax, (int)((0.80 * p.Lbox) * (0.80 * p.Lbox)));
    mcap[b] = std::min(std::max(Mmax, est), cap2);
    mmax = std::max(mmax, mcap[b]);
  }
  std::vector<double> w(nb, 0.0), prm(2 * (size_t)nb);
  for (int m0 = 0; m0 <= mmax; m0 += 5) {
    for (int b = 0; b < nb; ++b) {
      prm[2 * b] = Omega[b] >= 0.0 ? 1.0 : -1.0;
      prm[2 * b + 1] = std::max(0, std::min(5, mcap[b] + 1 - m0));
    }
    diag_params(p, prm);
    const int g = launch_lll(psi, p.N, p.dx, diag_prm_dev(p), m0, nb, diag_rows_dev(p), p.stream);
    const std::vector<double> r = diag_rows(p, g, 16, nb);
    for (int b = 0; b < nb; ++b) {
      const int mcnt = (int)prm[2 * b + 1];
      const double nrm2 = mom[5 * b];
      for (int q = 0; q < mcnt; ++q) {
        double nq = 0.0, re = 0.0, im = 0.0;
        for (int k = 0; k < g; ++k) {
          const double* row = r.data() + ((size_t)b * g + k) * 16;
          nq += row[3 * q];
          re += row[3 * q + 1];
          im += row[3 * q + 2];
        }
        if (nq > 0.0 && nrm2 > 0.0) w[b] += (re * re + im * im) / (nq * nrm2);
      }
    }
  }
  for (int b = 0; b < nb; ++b) {
    w_out[b] = std::min(std::max(w[b], 0.0), 1.0);
    if (mcap_out) mcap_out[b] = mcap[b];
  }
}

// Vortex count of each 2D state (vortex.py:10-48): the density smoothed with
// exp(-ell^2 k^2 / 2) (two line passes, ell per state), then the plaquettes'
// phase windings where the corner-averaged smoothed density exceeds frac
// times its maximum and the corner lies within R of (xcm, ycm).
// prm4[4b..4b+3] = xcm, ycm, R, ell; out4[4b..] = {Nv, N+, N-, Nnet}.
static void vortex_batch(Plan& p, int nb, const double2* psi, const double* prm4, double frac, int* out4) {
  if (p.ndim != 2) bad("the vortex count takes a two-dimensional state (a 3D state's midplane)");
  diag_alloc(p);
  plan_reserve(p, nb);
  const int N = p.N;
  const long long n = p.n;
  const Batch B{nb, nb >= 64 ? ~0ull : ((1ull << nb) - 1ull)};
  std::vector<double> prm(4 * (size_t)nb);
  for (int b = 0; b < nb; ++b) prm[b] = 0.5 * prm4[4 * b + 3] * prm4[4 * b + 3];
  diag_params(p, prm);                                  // gl2 per state
  launch_abs2c(psi, n * nb, p.w1, p.stream);
  PassArgs a = pass_base(p, p.fpN, p.lcN, B);          // x: lines along ix (stride N)
  a.nl = N; a.ni = N; a.in_si = 1; a.in_s = N; a.in_ib = n; a.n_load = N; a.out_si = 1; a.out_s = N;
  a.out_ib = n; a.n_store = N;
  a.in = p.w1; a.out = p.w1; a.gl2s = diag_prm_dev(p); a.out_scale = 1.0 / N;
  launch_pass(0, LD_C, DIR_FMI, MUL_GAUSS, ST_C, a, p.stream);
  PassArgs c = pass_base(p, p.fpN, p.lcN, B);          // y: contiguous lines
  c.nl = N; c.in_so = N; c.in_s = 1; c.in_ib = n; c.n_load = N; c.out_so = N; c.out_s = 1; c.out_ib = n;
  c.n_store = N;
  c.in = p.w1; c.out = p.w1; c.gl2s = diag_prm_dev(p); c.out_scale = 1.0 / N;
  launch_pass(1, LD_C, DIR_FMI, MUL_GAUSS, ST_C, c, p.stream);
  const int g = launch_nbar(p.w1, n, nb, p.rho, diag_rows_dev(p), p.stream);
  std::vector<double> r = diag_rows(p, g, 1, nb);
  for (int b = 0; b < nb; ++b) {
    double mx = 0.0;
    for (int k = 0; k < g; ++k) mx = std::max(mx, r[(size_t)b * g + k]);
    prm[4 * b] = prm4[4 * b];
    prm[4 * b + 1] = prm4[4 * b + 1];
    prm[4 * b + 2] = prm4[4 * b + 2];
    prm[4 * b + 3] = frac * mx;
  }
  diag_params(p, prm);
  const int gv = launch_vortex(psi, p.rho, N, p.dx, diag_prm_dev(p), nb, diag_rows_dev(p), p.stream);
  r = diag_rows(p, gv, 2, nb);
  for (int b = 0; b < nb; ++b) {
    double np_ = 0.0, nm = 0.0;
    for (int k = 0; k < gv; ++k) {
      np_ += r[((size_t)b * gv + k) * 2];
      nm += r[((size_t)b * gv + k) * 2 + 1];
    }
    const int ip = (int)llround(np_), im = (int)llround(nm);
    out4[4 * b] = ip + im;
    out4[4 * b + 1] = ip;
    out4[4 * b + 2] = im;
    out4[4 * b + 3] = ip - im;
  }
}



// Resampling between grids (grid.py:140-161): see include/becgpp_hip.h.
extern "C" int bec_resample(int ndim, int n0, double dx0, const void* psi_old, int n1, double dx1, void* psi_new,
                            double* norm_out, void* stream) {
  BEC_TRY({
    if (!psi_old || !psi_new || !norm_out) bad("null argument");
    if (ndim != 2 && ndim != 3) bad("ndim must be 2 or 3");
    if (n0 < 1 || n1 < 1 || !(dx0 > 0.0) || !(dx1 > 0.0)) bad("grid sizes and spacings must be positive");
    *norm_out = 0.0;
    const double x0 = -(double)(n0 / 2) * dx0;             // first and last node of the old grid
    const double x1 = x0 + (double)(n0 - 1) * dx0;
    if (!(x1 > x0)) return 0;                                // degenerate: no interpolation
    hipStream_t s = (hipStream_t)stream;
    const long long n = ndim == 2 ? (long long)n1 * n1 : (long long)n1 * n1 * n1;
    double* rows = nullptr;
    HIPCHK(hipMalloc(&rows, 1024 * sizeof(double)));
    std::vector<double> r(1024);
    int g = 0;
    try {
      g = launch_resample((const double2*)psi_old, ndim, n0, x0, x1 - x0, n1, dx1, (double2*)psi_new, rows, s);
      HIPCHK(hipGetLastError());
      HIPCHK(hipMemcpyAsync(r.data(), rows, (size_t)g * sizeof(double), hipMemcpyDeviceToHost, s));
      HIPCHK(hipStreamSynchronize(s));
    } catch (...) {
      (void)hipFree(rows);
      throw;
    }
    HIPCHK(hipFree(rows));
    double sum = 0.0;
    for (int b = 0; b < g; ++b) sum += r[b];
    const double nn = sqrt(sum * (ndim == 2 ? dx1 * dx1 : dx1 * dx1 * dx1));
    *norm_out = nn;
    if (std::isfinite(nn) && nn > 1e-14) {
      launch_scale((const double2*)psi_new, (double2*)psi_new, &nn, n, 1, s);
      HIPCHK(hipGetLastError());
    }
  });
}

extern "C" int bec_lll_weight(bec_plan* h, const void* psi, double Omega, int Mmax, double* w_out, int* mcap_out) {
  BEC_TRY({
    if (!h || !psi || !w_out) bad("null argument");
    lll_batch(h->p, 1, (const double2*)psi, &Omega, Mmax, w_out, mcap_out);
  });
}

extern "C" int bec_lll_weight_batch(bec_plan* h, int nb, const void* psi, const double* Omega, int Mmax,
                                    double* w_out) {
  BEC_TRY({
    if (!h || !psi || !Omega || !w_out) bad("null argument");
    check_batch(nb);
    lll_batch(h->p, nb, (const double2*)psi, Omega, Mmax, w_out, nullptr);
  });
}

extern "C" int bec_vortex_count(bec_plan* h, const void* psi, double xcm, double ycm, double R, double ell,
                                double frac, int* out4) {
  BEC_TRY({
    if (!h || !psi || !out4) bad("null argument");
    const double prm[4] = {xcm, ycm, R, ell};
    vortex_batch(h->p, 1, (const double2*)psi, prm, frac, out4);
  });
}

extern "C" int bec_vortex_count_batch(bec_plan* h, int nb, const void* psi, const double* prm4, double frac,
                                      int* out4) {
  BEC_TRY({
    if (!h || !psi || !prm4 || !out4) bad("null argument");
    check_batch(nb);
    vortex_batch(h->p, nb, (const double2*)psi, prm4, frac, out4);
  });
}

extern "C" int bec_moments(bec_plan* h, const void* psi, double* out5) {
  BEC_TRY({
    if (!h || !psi || !out5) bad("null argument");
    moments_batch(h->p, 1, (const double2*)psi, out5);
  });
}

// The quantities observables() needs, for nb states of one grid in one set of
// launches: out[16 b ..] = E, Ekin, Etrap, Econtact, Ethree, Erot, Egrav, Lz
// (as bec_energy with ph[b]), sum rho, sum r^2 rho, sum (x^2+y^2) rho,
// sum z^2 rho, max rho (before dV), and sum rho (W*rho) dV (the q2d virial
// kernel's term, when the plan holds that spectrum and G_C != 0; else NaN).
extern "C" int bec_observe_batch(bec_plan* h, int nb, const bec_phys* php, const void* psi, double* out) {
  BEC_TRY({
    if (!h || !php || !psi || !out) bad("null argument");
    check_batch(nb);
    Plan& p = h->p;
    plan_reserve(p, nb);
    std::vector<Phys> ph(nb);
    unsigned long long all = nb >= 64 ? ~0ull : ((1ull << nb) - 1ull), lr = 0, wm = 0;
    for (int b = 0; b < nb; ++b) {
      ph[b] = phys_of(php + b);
      if (plan_long_range(p, ph[b])) {
        lr |= 1ull << b;
        if (p.khat[1]) wm |= 1ull << b;
      }
    }
    const double2* f = (const double2*)psi;
    run_energy_sums(p, Batch{nb, all}, f, nullptr, nullptr, p.rho);
    if (lr) run_conv(p, 0, Batch{nb, lr}, p.rho, 1.0, p.loc, p.rho);
    if (wm) run_conv(p, 1, Batch{nb, wm}, p.rho, 1.0, p.cand, p.rho, nullptr, C_EGRAV2, R_WVIR);
    add_job(p, C_MOM, launch_moments(f, p.ndim, p.N, p.dx, nullptr, hpart_of(p, C_MOM), p.stream, nb), 4, 1, R_MOM,
            Batch{nb, all});
    HIPCHK(hipGetLastError());
    plan_sync(p);
    for (int b = 0; b < nb; ++b) {
      const bool l = (lr >> b) & 1ull;
      const Energy e = energy_from_sums(p, b, ph[b], l, false, -ph[b].G_C);
      double* o = out + 16 * (size_t)b;
      const double v[8] = {e.E, e.Ekin, e.Etrap, e.Econtact, e.Ethree, e.Erot, e.Egrav, e.Lz};
      memcpy(o, v, sizeof(v));
      memcpy(o + 8, hres(p, b) + R_MOM, sizeof(double) * 5);
      o[13] = ((wm >> b) & 1ull) ? hres(p, b)[R_WVIR] * p.dv : NAN;
      o[14] = o[15] = 0.0;
    }
  });
}
