// Initial states on the device: the Thomas-Fermi reference densities
// (thomasfermi.py:33-156) and the seed fields built on them (seeds.py:11-103).
//
// Chemical potential.  rho(mu) = the upper root of b3 rho^2 + b2 rho =
// mu - base (0 where there is none) is non-decreasing in mu, and so is the
// norm N(mu) = sum rho dV; mu is the crossing N(mu) = 1.  Each round of the
// search evaluates N at 32 points in one pass over base (one partial row of
// 32 sums per workgroup) and a one-workgroup step folds the rows in order and
// narrows the bracket.  The upper end is first widened as the reference does
// (hi <- lo + 2 (hi - lo) + 1e-9, at most 200 times: 32 consecutive widenings
// are priced per round), then the bracket is cut into 32 per round; 18 rounds
// narrow it by 2^90, the reference's 90 halvings, and mu is its midpoint.
// The rounds are a fixed launch sequence (rounds after the crossing is found
// return at once), so nothing waits on the host.
//
// Nonlocal fixed point (compressed TF with the long-range potential): one
// iteration is the convolution, base = V_eff + phi, the search, the new
// density normalised, its L1 change, the relaxed mix and its normalisation.
// The state (bracket, sums, iteration count, convergence) stays in device
// memory; the host queues iterations in chunks and reads one host-mapped flag
// per chunk.
//
// Seeds.  psi = A core (1 + a xi_a) exp(i (p xi_p + m theta + w sum_k theta_k)),
// normalised: A the Gaussian exp(-r^2/2) or sqrt of a TF density, core the
// product of tanh(|r - r_k| / xi) over the vortices r_k (in-plane distances),
// theta = atan2(y, x), theta_k = atan2(y - y_k, x - x_k).  xi_a, xi_p are
// standard normals from a counter-based generator: the splitmix64 stream
// seeded at key(rng, stream) (outputs 2i and 2i + 1 for point i) through the
// Box-Muller map, so every point's noise is a function of (rng, stream,
// index) alone and a CPU restatement draws the same numbers.
#include <math.h>
#include <string.h>
#include <vector>

#include "../../include/becgpp_hip.h"
#include "bec_ops.h"

#define TF_K 32             // points priced per search round
#define TF_ROUNDS_EXPAND 7  // rounds covering the reference's 200 widenings (7 x 32 candidates)
#define TF_ROWS 1024

enum TfPhase { TF_EXPAND = 0, TF_SECTION = 1, TF_FOUND = 2, TF_FAILED = 3 };

struct TfDev {
  double lo, hi, mu;
  double ssum, dn, nsum;
  int phase, nexp, rounds_left;
  int skip, done, conv, iters, failed;
};

__device__ __forceinline__ double tf_root(double rhs, double b2, double b3) {
  if (b3 > 1e-30) {
    const double disc = b2 * b2 + 4.0 * b3 * rhs;
    if (!(disc > 0.0)) return 0.0;
    const double root = (-b2 + sqrt(disc)) / (2.0 * b3);
    return root > 0.0 ? root : 0.0;
  }
  if (fabs(b2) > 1e-30) {
    const double r = rhs / b2;
    return r > 0.0 ? r : 0.0;
  }
  return 0.0;
}

#define TF_LOOP(n) \
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < (n); i += (long long)gridDim.x * blockDim.x)

// K block-level sums (and KM maxima after them) into row `row` of width W
template <int K, int KM, int W>
__device__ __forceinline__ void tf_row(const double* v, double* rows, int row) {
  __shared__ double red[(K + KM) * 4];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < K + KM; ++q) {
    const double s = q < K ? wave_sum(v[q]) : wave_max(v[q]);
    if (lane == 0) red[q * 4 + wid] = s;
  }
  __syncthreads();
  if (threadIdx.x < K + KM) {
    const int q = threadIdx.x;
    double s = red[q * 4];
    for (int w = 1; w < (int)(blockDim.x >> 6); ++w) s = q < K ? s + red[q * 4 + w] : fmax(s, red[q * 4 + w]);
    rows[(long long)row * W + q] = s;
  }
}

// the 32 chemical potentials a round prices (identical in the pricing and the
// step kernels): widening candidates h_0 = hi, h_{j+1} = lo + 2 (h_j - lo) + 1e-9,
// or the interior points lo + (hi - lo) (j + 1) / 32
__device__ __forceinline__ void tf_points(const TfDev* s, double* mu) {
  if (s->phase == TF_EXPAND) {
    double h = s->hi;
    for (int j = 0; j < TF_K; ++j) {
      mu[j] = h;
      h = s->lo + 2.0 * (h - s->lo) + 1e-9;
    }
  } else {
    for (int j = 0; j < TF_K; ++j) mu[j] = s->lo + (s->hi - s->lo) * (double)(j + 1) / (double)TF_K;
  }
}

__global__ __launch_bounds__(256) void k_tf_price(const double* __restrict__ base, long long n, const TfDev* st,
                                                  double b2, double b3, double* rows) {
  __shared__ double mu[TF_K];
  const TfDev s = *st;
  if (s.skip || (s.phase != TF_EXPAND && s.phase != TF_SECTION)) return;
  if (threadIdx.x == 0) tf_points(&s, mu);
  __syncthreads();
  double v[TF_K];
#pragma unroll
  for (int j = 0; j < TF_K; ++j) v[j] = 0.0;
  TF_LOOP(n) {
    const double b = base[i];
#pragma unroll
    for (int j = 0; j < TF_K; ++j) v[j] += tf_root(mu[j] - b, b2, b3);
  }
  tf_row<TF_K, 0, TF_K>(v, rows, blockIdx.x);
}

// one workgroup: fold the rows in order, then widen or narrow the bracket
__global__ __launch_bounds__(256) void k_tf_narrow(const double* __restrict__ rows, int nrows, TfDev* st,
                                                   double dv) {
  __shared__ double N[TF_K];
  __shared__ double mu[TF_K];
  if (st->skip || (st->phase != TF_EXPAND && st->phase != TF_SECTION)) return;
  if (threadIdx.x < TF_K) {
    double s = 0.0;
    for (int r = 0; r < nrows; ++r) s += rows[(long long)r * TF_K + threadIdx.x];
    N[threadIdx.x] = s * dv;
  }
  if (threadIdx.x == 0) tf_points(st, mu);
  __syncthreads();
  if (threadIdx.x != 0) return;
  TfDev& s = *st;
  if (s.phase == TF_EXPAND) {
    for (int j = 0; j < TF_K; ++j) {
      if (s.nexp + j > 200) {              // the 201st widening gives up (thomasfermi.py:51-55)
        s.phase = TF_FAILED;
        s.mu = NAN;
        return;
      }
      if (!(N[j] < 1.0)) {
        s.hi = mu[j];
        s.phase = TF_SECTION;
        if (s.rounds_left <= 0) {          // no narrowing asked for: the bracket's midpoint
          s.mu = 0.5 * (s.lo + s.hi);
          s.phase = TF_FOUND;
        }
        return;
      }
    }
    s.nexp += TF_K;
    s.hi = s.lo + 2.0 * (mu[TF_K - 1] - s.lo) + 1e-9;
    return;
  }
  int j = 0;
  while (j < TF_K - 1 && N[j] < 1.0) ++j;
  if (j < TF_K - 1) {
    const double lo = j == 0 ? s.lo : mu[j - 1];
    s.hi = mu[j];
    s.lo = lo;
  } else {
    s.lo = mu[TF_K - 2];
  }
  if (--s.rounds_left <= 0) {
    s.mu = 0.5 * (s.lo + s.hi);
    s.phase = TF_FOUND;
  }
}

// base = veff + phi with its per-block min and max
__global__ __launch_bounds__(256) void k_tf_base(const double* __restrict__ veff, const double* __restrict__ phi,
                                                 double* base, long long n, double* rows) {
  double v[2] = {INFINITY, -INFINITY};
  TF_LOOP(n) {
    const double b = veff[i] + phi[i];
    base[i] = b;
    v[0] = fmin(v[0], b);
    v[1] = fmax(v[1], b);
  }
  v[0] = -v[0];                    // the minimum as a maximum of negatives
  tf_row<0, 2, 2>(v, rows, blockIdx.x);
}

// start of a search: the bracket [min base, max base + |b2| + |b3| + 1]
// (thomasfermi.py:99-100) from the base rows, or the given lo, hi
__global__ __launch_bounds__(64) void k_tf_begin(const double* __restrict__ rows, int nrows, TfDev* st, double b2,
                                                 double b3, int rounds) {
  if (threadIdx.x != 0) return;
  TfDev& s = *st;
  s.skip = s.done;
  if (s.skip) return;
  if (rows) {
    double mn = INFINITY, mx = -INFINITY;
    for (int r = 0; r < nrows; ++r) {
      mn = fmin(mn, -rows[2 * r]);
      mx = fmax(mx, rows[2 * r + 1]);
    }
    s.lo = mn;
    s.hi = mx + fabs(b2) + fabs(b3) + 1.0;
  }
  s.phase = TF_EXPAND;
  s.nexp = 0;
  s.rounds_left = rounds;
}

// the new density rho(mu - base) and its sum
__global__ __launch_bounds__(256) void k_tf_density(const double* __restrict__ base, long long n, const TfDev* st,
                                                    double b2, double b3, double* nnew, double* rows) {
  const TfDev s = *st;
  if (s.skip || s.phase != TF_FOUND) return;
  double v[1] = {0.0};
  TF_LOOP(n) {
    const double r = tf_root(s.mu - base[i], b2, b3);
    nnew[i] = r;
    v[0] += r;
  }
  tf_row<1, 0, 1>(v, rows, blockIdx.x);
}

__global__ __launch_bounds__(64) void k_tf_fold_density(const double* __restrict__ rows, int nrows, TfDev* st,
                                                        double dv, int* hdone) {
  if (threadIdx.x != 0) return;
  TfDev& s = *st;
  if (s.skip) return;
  if (s.phase != TF_FOUND) {             // no chemical potential: the fixed point gives up
    s.failed = 1;
    s.done = 1;
    s.skip = 1;
    if (hdone) *hdone = 1;
    return;
  }
  double sum = 0.0;
  for (int r = 0; r < nrows; ++r) sum += rows[r];
  s.ssum = sum * dv;
  if (!(s.ssum > 0.0)) {                 // an empty density: stop with the current n (thomasfermi.py:105-106)
    s.done = 1;
    s.skip = 1;
    if (hdone) *hdone = 1;
  }
}

// n_new / ssum, its L1 change from n, and n <- (1 - relax) n + relax n_new
__global__ __launch_bounds__(256) void k_tf_mix(double* nvec, const double* __restrict__ nnew, long long n,
                                                const TfDev* st, double relax, double* rows) {
  const TfDev s = *st;
  if (s.skip) return;
  double v[2] = {0.0, 0.0};
  TF_LOOP(n) {
    const double nn = nnew[i] / s.ssum;
    const double o = nvec[i];
    v[0] += fabs(nn - o);
    const double m = (1.0 - relax) * o + relax * nn;
    nvec[i] = m;
    v[1] += m;
  }
  tf_row<2, 0, 2>(v, rows, blockIdx.x);
}

__global__ __launch_bounds__(64) void k_tf_fold_mix(const double* __restrict__ rows, int nrows, TfDev* st, double dv,
                                                    double tol, int maxit, int* hdone) {
  if (threadIdx.x != 0) return;
  TfDev& s = *st;
  if (s.skip) return;
  double d = 0.0, m = 0.0;
  for (int r = 0; r < nrows; ++r) {
    d += rows[2 * r];
    m += rows[2 * r + 1];
  }
  s.dn = d * dv;
  s.nsum = m * dv;
  s.iters += 1;
  if (s.dn < tol) s.conv = 1;
  if (s.conv || s.iters >= maxit) {
    s.done = 1;
    if (hdone) *hdone = 1;
  }
}

__global__ __launch_bounds__(256) void k_tf_scale(double* nvec, long long n, const TfDev* st) {
  const TfDev s = *st;
  if (s.skip) return;
  TF_LOOP(n) nvec[i] = nvec[i] / s.nsum;
}

static int tf_blocks(long long n) {
  long long g = (n + 255) / 256;
  if (g > TF_ROWS) g = TF_ROWS;
  return (int)(g < 1 ? 1 : g);
}

// plan buffers of the TF search: the state, partial rows and a host-mapped flag
struct TfBufs {
  TfDev* st;
  double* rows;
  int* hdone;
  int* hdone_dev;
};
static TfBufs tf_bufs(Plan& p) {
  if (!p.tfbuf) HIPCHK(hipMalloc(&p.tfbuf, sizeof(TfDev) + sizeof(double) * TF_K * TF_ROWS + 64));
  if (!p.tfflag) {
    HIPCHK(hipHostMalloc(&p.tfflag, sizeof(int), hipHostMallocMapped | hipHostMallocCoherent));
    HIPCHK(hipHostGetDevicePointer((void**)&p.tfflag_dev, p.tfflag, 0));
  }
  TfBufs b;
  b.st = (TfDev*)p.tfbuf;
  b.rows = (double*)((char*)p.tfbuf + ((sizeof(TfDev) + 63) / 64) * 64);
  b.hdone = p.tfflag;
  b.hdone_dev = p.tfflag_dev;
  return b;
}

// the rounds of one chemical-potential search on `base`, queued on s
static void tf_search(const TfBufs& b, const double* base, long long n, double b2, double b3, double dv, int rounds,
                      hipStream_t s) {
  const int g = tf_blocks(n);
  for (int r = 0; r < TF_ROUNDS_EXPAND + rounds; ++r) {
    hipLaunchKernelGGL(k_tf_price, dim3(g), dim3(256), 0, s, base, n, b.st, b2, b3, b.rows);
    hipLaunchKernelGGL(k_tf_narrow, dim3(1), dim3(256), 0, s, b.rows, g, b.st, dv);
  }
}

static int section_rounds(int steps) { return (steps + 4) / 5; }   // 5 halvings per round of 32

extern "C" int bec_tf_mu(bec_plan* h, const double* base, double b2, double b3, double lo, double hi, int steps,
                         double* mu_out) {
  try {
    if (!h || !base || !mu_out) throw BecError(1, "null argument");
    if (steps < 0) throw BecError(1, "steps must be >= 0");
    Plan& p = h->p;
    const TfBufs b = tf_bufs(p);
    TfDev s0;
    memset(&s0, 0, sizeof(s0));
    s0.lo = lo;
    s0.hi = hi;
    HIPCHK(hipMemcpyAsync(b.st, &s0, sizeof(s0), hipMemcpyHostToDevice, p.stream));
    const int R = section_rounds(steps);
    hipLaunchKernelGGL(k_tf_begin, dim3(1), dim3(64), 0, p.stream, nullptr, 0, b.st, b2, b3, R);
    tf_search(b, base, p.n, b2, b3, p.dv, R, p.stream);
    HIPCHK(hipGetLastError());
    TfDev s1;
    HIPCHK(hipMemcpyAsync(&s1, b.st, sizeof(s1), hipMemcpyDeviceToHost, p.stream));
    HIPCHK(hipStreamSynchronize(p.stream));
    if (s1.phase == TF_FAILED) *mu_out = NAN;
    else if (s1.phase == TF_FOUND) *mu_out = s1.mu;
    else *mu_out = 0.5 * (s1.lo + s1.hi);
    return 0;
  } catch (const BecError& e) {
    bec_set_error(e.what());
    return e.code;
  } catch (const std::exception& e) {
    bec_set_error(e.what());
    return 2;
  }
}

extern "C" int bec_tf_nonlocal(bec_plan* h, const double* veff, double b2, double b3, double G_C, int iters,
                               double relax, double tol, double* n_io, double* out4) {
  try {
    if (!h || !veff || !n_io || !out4) throw BecError(1, "null argument");
    if (iters < 0) throw BecError(1, "iters must be >= 0");
    Plan& p = h->p;
    if (!p.khat[0]) throw BecError(1, "kernel spectrum not built for this plan");
    plan_reserve(p, 1);
    const TfBufs b = tf_bufs(p);
    const long long n = p.n;
    double* phi = p.loc;               // workspace of the plan: potential, base, new density
    double* base = p.cand;
    double* nnew = p.rho;
    TfDev s0;
    memset(&s0, 0, sizeof(s0));
    s0.dn = INFINITY;
    s0.mu = NAN;
    s0.done = iters == 0 ? 1 : 0;
    *b.hdone = s0.done;
    HIPCHK(hipMemcpyAsync(b.st, &s0, sizeof(s0), hipMemcpyHostToDevice, p.stream));
    const int g = tf_blocks(n), R = section_rounds(90);
    auto iteration = [&]() {
      run_conv(p, 0, one_instance(), n_io, -G_C, phi, nullptr);
      hipLaunchKernelGGL(k_tf_base, dim3(g), dim3(256), 0, p.stream, veff, phi, base, n, b.rows);
      hipLaunchKernelGGL(k_tf_begin, dim3(1), dim3(64), 0, p.stream, b.rows, g, b.st, b2, b3, R);
      tf_search(b, base, n, b2, b3, p.dv, R, p.stream);
      hipLaunchKernelGGL(k_tf_density, dim3(g), dim3(256), 0, p.stream, base, n, b.st, b2, b3, nnew, b.rows);
      hipLaunchKernelGGL(k_tf_fold_density, dim3(1), dim3(64), 0, p.stream, b.rows, g, b.st, p.dv, b.hdone_dev);
      hipLaunchKernelGGL(k_tf_mix, dim3(g), dim3(256), 0, p.stream, n_io, nnew, n, b.st, relax, b.rows);
      hipLaunchKernelGGL(k_tf_fold_mix, dim3(1), dim3(64), 0, p.stream, b.rows, g, b.st, p.dv, tol, iters,
                         b.hdone_dev);
      hipLaunchKernelGGL(k_tf_scale, dim3(g), dim3(256), 0, p.stream, n_io, n, b.st);
    };
    // iterations queued in chunks; the device's done flag is read one chunk behind
    const int CH = 8;
    hipEvent_t ev[2];
    HIPCHK(hipEventCreateWithFlags(&ev[0], hipEventDisableTiming));
    HIPCHK(hipEventCreateWithFlags(&ev[1], hipEventDisableTiming));
    try {
      int queued = 0;
      for (int chunk = 0; queued < iters; ++chunk) {
        for (int k = 0; k < CH && queued < iters; ++k, ++queued) iteration();
        HIPCHK(hipGetLastError());
        HIPCHK(hipEventRecord(ev[chunk & 1], p.stream));
        if (chunk > 0) {
          HIPCHK(hipEventSynchronize(ev[(chunk - 1) & 1]));
          if (*(volatile int*)b.hdone) break;
        }
      }
      HIPCHK(hipStreamSynchronize(p.stream));
    } catch (...) {
      (void)hipStreamSynchronize(p.stream);
      (void)hipEventDestroy(ev[0]);
      (void)hipEventDestroy(ev[1]);
      throw;
    }
    (void)hipEventDestroy(ev[0]);
    (void)hipEventDestroy(ev[1]);
    TfDev s1;
    HIPCHK(hipMemcpy(&s1, b.st, sizeof(s1), hipMemcpyDeviceToHost));
    out4[0] = s1.failed ? NAN : s1.mu;
    out4[1] = s1.dn;
    out4[2] = s1.iters;
    out4[3] = s1.conv;
    return 0;
  } catch (const BecError& e) {
    bec_set_error(e.what());
    return e.code;
  } catch (const std::exception& e) {
    bec_set_error(e.what());
    return 2;
  }
}

// ---------------------------------------------------------------------------
// seeds
// ---------------------------------------------------------------------------
__host__ __device__ __forceinline__ unsigned long long sm_mix(unsigned long long z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
#define SM_GAMMA 0x9E3779B97F4A7C15ull
// output k (0-based) of the splitmix64 stream seeded at `state`
__host__ __device__ __forceinline__ unsigned long long sm_out(unsigned long long state, unsigned long long k) {
  return sm_mix(state + (k + 1ull) * SM_GAMMA);
}
__host__ __device__ __forceinline__ unsigned long long noise_key(unsigned long long rng, unsigned long long stream) {
  return sm_out((rng & 0xFFFFFFFFull) | (stream << 32), 0);
}
// standard normal of point i: Box-Muller on outputs 2i, 2i + 1 of the key's stream
__device__ __forceinline__ double noise_normal(unsigned long long key, unsigned long long i) {
  const unsigned long long a = sm_out(key, 2ull * i), c = sm_out(key, 2ull * i + 1ull);
  const double u1 = (double)((a >> 11) + 1ull) * 0x1.0p-53;
  const double u2 = (double)(c >> 11) * 0x1.0p-53;
  return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

struct SeedArgs {
  int ndim, N;
  double dx;
  int envelope;
  double amp_noise, phase_noise;
  int winding;
  double vortex_sign, core;
  unsigned long long ka, kp;
};

__global__ __launch_bounds__(256) void k_seed(SeedArgs a, const double* __restrict__ env,
                                              const double* __restrict__ vort, int nv, double2* out, long long n,
                                              double* rows) {
  double v[1] = {0.0};
  const int N = a.N;
  TF_LOOP(n) {
    int ix, iy, iz = 0;
    if (a.ndim == 2) {
      ix = (int)(i / N);
      iy = (int)(i % N);
    } else {
      ix = (int)(i / ((long long)N * N));
      iy = (int)((i / N) % N);
      iz = (int)(i % N);
    }
    const double x = (double)(ix - N / 2) * a.dx, y = (double)(iy - N / 2) * a.dx;
    double A;
    if (a.envelope) {
      A = sqrt(fmax(env[i], 0.0));
    } else {
      const double z = (double)(iz - N / 2) * a.dx;
      double r2 = x * x + y * y;
      if (a.ndim == 3) r2 += z * z;
      A = exp(-0.5 * r2);
    }
    double core = 1.0, ph = 0.0;
    for (int k = 0; k < nv; ++k) {
      const double dxk = x - vort[2 * k], dyk = y - vort[2 * k + 1];
      ph += a.vortex_sign * atan2(dyk, dxk);
      core *= tanh(sqrt(dxk * dxk + dyk * dyk) / a.core);
    }
    const double amp = A * core * (1.0 + a.amp_noise * noise_normal(a.ka, (unsigned long long)i));
    if (a.phase_noise != 0.0) ph += a.phase_noise * noise_normal(a.kp, (unsigned long long)i);
    if (a.winding != 0) ph += (double)a.winding * atan2(y, x);
    double sn, cs;
    sincos(ph, &sn, &cs);
    const double2 w = make_double2(amp * cs, amp * sn);
    out[i] = w;
    v[0] += w.x * w.x + w.y * w.y;
  }
  tf_row<1, 0, 1>(v, rows, blockIdx.x);
}

extern "C" int bec_seed(bec_plan* h, const bec_seed_spec* spec, const double* env, const double* vortices,
                        int n_vortices, void* psi_out) {
  try {
    if (!h || !spec || !psi_out) throw BecError(1, "null argument");
    if (spec->envelope && !env) throw BecError(1, "a density envelope needs its field");
    if (n_vortices < 0 || (n_vortices > 0 && !vortices)) throw BecError(1, "bad vortex list");
    if (n_vortices > 0 && !(spec->core > 0.0)) throw BecError(1, "the vortex core radius must be positive");
    Plan& p = h->p;
    const TfBufs b = tf_bufs(p);
    SeedArgs a;
    a.ndim = p.ndim;
    a.N = p.N;
    a.dx = p.dx;
    a.envelope = spec->envelope;
    a.amp_noise = spec->amp_noise;
    a.phase_noise = spec->phase_noise;
    a.winding = spec->winding;
    a.vortex_sign = spec->vortex_sign;
    a.core = spec->core;
    a.ka = noise_key(spec->rng, 0);
    a.kp = noise_key(spec->rng, 1);
    const int g = tf_blocks(p.n);
    hipLaunchKernelGGL(k_seed, dim3(g), dim3(256), 0, p.stream, a, env, vortices, n_vortices, (double2*)psi_out,
                       p.n, b.rows);
    HIPCHK(hipGetLastError());
    std::vector<double> r(g);
    HIPCHK(hipMemcpyAsync(r.data(), b.rows, sizeof(double) * g, hipMemcpyDeviceToHost, p.stream));
    HIPCHK(hipStreamSynchronize(p.stream));
    double sum = 0.0;
    for (int k = 0; k < g; ++k) sum += r[k];
    const double nrm = sqrt(sum * p.dv);
    if (!(nrm > 0.0) || !std::isfinite(nrm)) throw BecError(1, "the seed has no norm");
    launch_scale((const double2*)psi_out, (double2*)psi_out, &nrm, p.n, 1, p.stream);
    HIPCHK(hipGetLastError());
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

// host draw of the generator (bec_noise): the same numbers the seed kernel uses
extern "C" int bec_noise(unsigned long long rng, unsigned long long stream, unsigned long long first, long long count,
                         double* out) {
  if (!out || count < 0) {
    bec_set_error("bad arguments");
    return 1;
  }
  const unsigned long long key = noise_key(rng, stream);
  for (long long k = 0; k < count; ++k) {
    const unsigned long long i = first + (unsigned long long)k;
    const unsigned long long a = sm_out(key, 2ull * i), c = sm_out(key, 2ull * i + 1ull);
    const double u1 = (double)((a >> 11) + 1ull) * 0x1.0p-53;
    const double u2 = (double)(c >> 11) * 0x1.0p-53;
    out[k] = sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
  }
  return 0;
}
