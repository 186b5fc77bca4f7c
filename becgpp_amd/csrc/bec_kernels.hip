// HIP kernels of the becGPP ground-state path for CDNA4 (gfx950), FP64.
//
// All transforms are line transforms: a workgroup loads lpw lines of one axis
// into LDS (coalesced: whole rows for the contiguous axis, tiles of lpw
// adjacent columns for strided axes), runs a mixed-radix Stockham FFT in LDS
// with tpl threads per line, applies a per-frequency multiplier, optionally
// transforms back, and writes the result through a fused epilogue.  Every
// launch covers a batch of instances of one grid (line l of instance l / nl).
// The work is HBM/LDS-bound FP64 vector arithmetic: there is no MFMA here.
#include "bec_internal.h"
#include <hip/hip_ext.h>
#include <algorithm>
#include <type_traits>

#define TWO_PI 6.283185307179586

// The file is compiled once per translation unit BEC_TU = 0..7 (the build
// runs them in parallel): 0 holds the elementwise kernels and the launch
// dispatch, 1..5 the kinetic passes of one axis mode each, 6 and 7 the
// generic line passes.  Without BEC_TU every part is compiled.
#ifndef BEC_TU
#define BEC_TU -1
#endif
#define BEC_TU_IS(k) (BEC_TU < 0 || BEC_TU == (k))

__device__ __forceinline__ double kfreq(int j, int N, double invNdx) {
  const int f = (j < (N + 1) / 2) ? j : j - N;
  return TWO_PI * ((double)f * invNdx);
}
__device__ __forceinline__ double xcoord(int i, int N, double dx) { return (double)(i - N / 2) * dx; }

// XCD-aware block order: the dispatcher deals workgroups round-robin over the
// BEC_XCDS chiplets, each with its own L2.  Logical block (b % XCDS) * (nb /
// XCDS) + b / XCDS gives every XCD one contiguous range of lines, so adjacent
// strided columns (which share cache lines) are read through one L2.  nb is
// a multiple of BEC_XCDS.
__device__ __forceinline__ int xcd_block(int b, int nb) {
  return (b % BEC_XCDS) * (nb / BEC_XCDS) + b / BEC_XCDS;
}

__device__ __forceinline__ bool inst_on(unsigned long long act, int b) { return (act >> b) & 1ull; }
__device__ __forceinline__ bool act_on(const Act& a, int b) {
  return a.fl ? (a.fl[b] & a.bit) != 0u : ((a.m >> b) & 1ull) != 0ull;
}
// does any line of [l0, l0 + lpw) belong to an active instance?  (block-uniform)
__device__ __forceinline__ bool block_on(const Act& a, int l0, int lpw, int nl, int total) {
  if (l0 >= total) return false;
  const int b0 = l0 / nl, b1 = min(l0 + lpw - 1, total - 1) / nl;
  for (int b = b0; b <= b1; ++b)
    if (act_on(a, b)) return true;
  return false;
}

__device__ __forceinline__ double trap_value(double r2, double tc, int s_mode, double s_half) {
  if (s_mode == 0) return 0.0;
  const double p = (s_mode == 1) ? r2 : (s_mode == 2 ? r2 * r2 : pow(r2, s_half));
  return tc * p;
}

// ---------------------------------------------------------------------------
// Deterministic reductions.  Each workgroup folds its threads' K sums and KM
// maxima in a fixed tree order and stores one row of KP values (KP = K+KM
// rounded up to a power of two) at row index `row`, in device memory and/or
// pinned host memory.  The host folds the host rows in row order after the
// stream synchronisation it performs anyway; device consumers fold the device
// rows in a later launch.  No workgroup issues a device-wide fence (on a
// multi-XCD device an agent-scope fence writes back and invalidates the L2).
// ---------------------------------------------------------------------------
template <int K, int KM>
__device__ __forceinline__ void block_partials(const double* v, const double* m, double* partials,
                                               double* hpart, long long row) {
  constexpr int KT = K + KM;
  constexpr int KP = KT <= 1 ? 1 : (KT <= 2 ? 2 : (KT <= 4 ? 4 : (KT <= 8 ? 8 : (KT <= 16 ? 16 : 32))));
  static_assert(KT <= 32, "at most 32 reduced values");
  __shared__ double red[(KT <= 16 ? 16 : 32) * 16];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
#pragma unroll
  for (int q = 0; q < K; ++q) {
    const double s = wave_sum(v[q]);
    if (lane == 0) red[q * 16 + wid] = s;
  }
#pragma unroll
  for (int q = 0; q < KM; ++q) {
    const double s = wave_max(m[q]);
    if (lane == 0) red[(K + q) * 16 + wid] = s;
  }
  __syncthreads();
  if (threadIdx.x < KT) {
    const int q = threadIdx.x;
    double s = red[q * 16];
    for (int w = 1; w < nw; ++w) s = (q < K) ? s + red[q * 16 + w] : fmax(s, red[q * 16 + w]);
    if (partials) partials[row * KP + q] = s;
    if (hpart) hpart[row * KP + q] = s;
  }
}

// ---------------------------------------------------------------------------
// axis geometry of the kinetic passes (within one instance)
// ---------------------------------------------------------------------------
template <int AX> struct Ax;
template <> struct Ax<A2Y> {   // rows: line l = ix, element e = iy
  static constexpr bool contig = true; static constexpr int ndim = 2; static constexpr int lsign = 1;
  __device__ static long long base(int l, int N) { return (long long)l * N; }
  __device__ static long long stride(int) { return 1; }
  __device__ static void coords(int l, int e, int, int& ix, int& iy, int& iz) { ix = l; iy = e; iz = 0; }
  __device__ static int lidx(int l, int) { return l; }
};
template <> struct Ax<A2X> {   // columns: line l = iy, element e = ix
  static constexpr bool contig = false; static constexpr int ndim = 2; static constexpr int lsign = -1;
  __device__ static long long base(int l, int) { return (long long)l; }
  __device__ static long long stride(int N) { return N; }
  __device__ static void coords(int l, int e, int, int& ix, int& iy, int& iz) { ix = e; iy = l; iz = 0; }
  __device__ static int lidx(int l, int) { return l; }
};
template <> struct Ax<A3Z> {   // line l = ix*N + iy, element e = iz
  static constexpr bool contig = true; static constexpr int ndim = 3; static constexpr int lsign = 0;
  __device__ static long long base(int l, int N) { return (long long)l * N; }
  __device__ static long long stride(int) { return 1; }
  __device__ static void coords(int l, int e, int N, int& ix, int& iy, int& iz) { ix = l / N; iy = l % N; iz = e; }
  __device__ static int lidx(int l, int) { return 0; }
};
template <> struct Ax<A3Y> {   // line l = ix*N + iz, element e = iy
  static constexpr bool contig = false; static constexpr int ndim = 3; static constexpr int lsign = 1;
  __device__ static long long base(int l, int N) { return (long long)(l / N) * N * N + (l % N); }
  __device__ static long long stride(int N) { return N; }
  __device__ static void coords(int l, int e, int N, int& ix, int& iy, int& iz) { ix = l / N; iy = e; iz = l % N; }
  __device__ static int lidx(int l, int N) { return l / N; }
};
template <> struct Ax<A3X> {   // line l = iy*N + iz, element e = ix
  static constexpr bool contig = false; static constexpr int ndim = 3; static constexpr int lsign = -1;
  __device__ static long long base(int l, int) { return (long long)l; }
  __device__ static long long stride(int N) { return (long long)N * N; }
  __device__ static void coords(int l, int e, int N, int& ix, int& iy, int& iz) { ix = e; iy = l / N; iz = l % N; }
  __device__ static int lidx(int l, int N) { return l / N; }
};

// per-instance scalar of a kinetic pass: from the device table when its bit
// is set in tabm, otherwise the launch argument
__device__ __forceinline__ double kin_scal(const KinArgs& a, int slot, int b, const double* byval) {
  return (a.tab && ((a.tabm >> slot) & 1u)) ? a.tab[(long long)b * DS_N + slot] : byval[b];
}

// the per-instance scalars a kinetic pass uses, read once per workgroup
struct KinScal {
  double t, sc, mu, phisc, beta2, beta3, rf;
  bool pc;                             // phi <- (phi + cq kq + cr kr) cs first (KH_LAST)
  double cq, cr, cs;
};
template <bool PC = false>
__device__ __forceinline__ KinScal kin_scal_all(const KinArgs& a, int b) {
  KinScal k;
  k.t = kin_scal(a, DS_T, b, a.t);
  k.sc = kin_scal(a, DS_SC, b, a.sc);
  k.mu = kin_scal(a, DS_MU, b, a.mu);
  k.phisc = kin_scal(a, DS_PHISC, b, a.phisc);
  const bool need_phys = a.local_on || a.rf_omega;
  const InstPhys ip = need_phys ? a.iphys[b] : InstPhys{0.0, 0.0, 0.0, 0.0};
  k.beta2 = ip.beta2;
  k.beta3 = ip.beta3;
  k.rf = a.rf_omega ? -ip.Omega : a.rf;
  k.pc = PC && a.kq && a.act.fl && (a.act.fl[b] & FL_PHIC);
  k.cq = k.pc ? a.tab[(long long)b * DS_N + DS_CQ] : 0.0;
  k.cr = k.pc ? a.tab[(long long)b * DS_N + DS_CR] : 0.0;
  k.cs = k.pc ? a.tab[(long long)b * DS_N + DS_CS] : 1.0;
  return k;
}
// The new-direction coefficients of instance b for the first energy pass of
// the device-driven loop (on when its flag word carries comb_bit), read at
// the kernel's start so that z and g load together with psi and d.
struct CombScal {
  bool on;
  double az, ad, ag, gre, gim;
};
__device__ __forceinline__ CombScal comb_scal(const KinArgs& a, int b, bool inst_on) {
  CombScal c{false, 0.0, 0.0, 0.0, 0.0, 0.0};
  if (a.d_out && inst_on && (a.act.fl[b] & a.comb_bit)) {
    const double* t = a.tab + (long long)b * DS_N;
    c.on = true;
    c.az = t[DS_AZ];
    c.ad = t[DS_AD];
    c.ag = t[DS_AG];
    c.gre = t[DS_GRE];
    c.gim = t[DS_GIM];
  }
  return c;
}


// Kinetic + rotation line pass.  Along an axis with wavenumber k the operator
// T - Omega L_z acts as the multiplier kf k^2 + rf c k, where c = +x for the y
// axis, -y for the x axis and 0 for the z axis (L_z = -i(x d_y - y d_x)).
// The loaded value is (psi - t[b] d) / sc[b].
//   KE_*: forward transform only; accumulate sum k^2/2 |t_k|^2 / N (kinetic)
//         and c sum k |t_k|^2 / N (L_z) per line (Parseval), and on the first
//         pass the local sums rho, V rho, rho^2, rho^3.  With hy_out the first
//         pass also stores IFFT(m FFT(loaded)) of its axis (the first term of
//         H of the trial state, reused when the trial is accepted).
//   KH_*: forward, multiply, inverse; FIRST writes, MID accumulates, LAST adds
//         the local potential (V + b2 rho + b3 rho^2 + phi) psi - mu psi.  The
//         pass that loads a transformed state writes it back (psi_w, and its
//         change to sdiff).
// Loads and stores: thread t owns one line of the block -- its FFT line
// t / tpl on a contiguous axis, column t mod lpw on a strided one (adjacent
// threads read adjacent columns) -- and its elements e0 + j tpl, j < BEC_EPT
// (N <= BEC_EPT tpl).  The line's addressing is computed once per thread, and
// all of a phase's loads are issued before any is used.
#define BEC_EPT 8
#ifndef BEC_KH_EARLY2D
#define BEC_KH_EARLY2D 1       // 2D H passes load their epilogue before the transforms
#endif
#ifndef BEC_KH_CHUNK
#define BEC_KH_CHUNK 4         // elements per chunk of the H passes' epilogue (3D axes)
#endif
#ifndef BEC_CDOT_CHUNK
#define BEC_CDOT_CHUNK 2       // elements per chunk of the inner-product epilogue
#endif
struct LineOwn {
  int cc, e0;                 // block line and first element of the thread
};
// LDS pitch of a block's lines.  On a strided axis adjacent threads hold the
// same element of adjacent lines; one extra element per line puts those 16-byte
// accesses in different banks.  A contiguous line is stored with one skipped
// slot per eight elements (lx<true>), L + L/8 slots.
template <bool CONTIG>
__host__ __device__ __forceinline__ int lpitch(int L) { return CONTIG ? L + (L >> 3) : L + 1; }
template <bool CONTIG>
__device__ __forceinline__ LineOwn line_owner(int tpl, int lpw) {
  LineOwn o;
  if (CONTIG) {
    o.cc = threadIdx.x / tpl;
    o.e0 = threadIdx.x - o.cc * tpl;
  } else {
    o.cc = threadIdx.x & (lpw - 1);
    o.e0 = threadIdx.x >> (__ffs(lpw) - 1);
  }
  return o;
}

// squared coordinates of a line's fixed axes; the element axis is filled per element
template <int AX>
__device__ __forceinline__ void line_sq(int li, int N, double dx, double* c2) {
  using X = Ax<AX>;
  int ix, iy, iz;
  X::coords(li, N / 2, N, ix, iy, iz);
  const double xx = xcoord(ix, N, dx), yy = xcoord(iy, N, dx), zz = xcoord(iz, N, dx);
  c2[0] = xx * xx;
  c2[1] = yy * yy;
  c2[2] = (X::ndim == 3) ? zz * zz : 0.0;
}
// element axis of each mode: 0 = x, 1 = y, 2 = z
template <int AX> struct AxE;
template <> struct AxE<A2Y> { static constexpr int k = 1; };
template <> struct AxE<A2X> { static constexpr int k = 0; };
template <> struct AxE<A3Z> { static constexpr int k = 2; };
template <> struct AxE<A3Y> { static constexpr int k = 1; };
template <> struct AxE<A3X> { static constexpr int k = 0; };
// trap potential at element e of a line, r^2 summed as x^2 + y^2 [+ z^2]
template <int AX>
__device__ __forceinline__ double line_V(const double* c2, int e, const KinArgs& a) {
  constexpr int K = AxE<AX>::k;
  const double ee = xcoord(e, a.N, a.dx);
  const double x2 = (K == 0) ? ee * ee : c2[0];
  const double y2 = (K == 1) ? ee * ee : c2[1];
  double r2 = x2 + y2;
  if (Ax<AX>::ndim == 3) r2 += (K == 2) ? ee * ee : c2[2];
  return trap_value(r2, a.tc, a.s_mode, a.s_half);
}

// C is the storage of the complex fields: double2, or float2 in an FP32 phase
// (values are widened to FP64 in registers; the transforms run in C; every
// sum accumulates in FP64).
// PC: the KH_LAST variant that first combines the accepted state's potential
// (KinArgs::kq; long-range grids of the polynomial search)
template <int AX, int KIND, class C, bool ODD = false, int LC = 0, bool PC = false>
__global__ __launch_bounds__(BEC_WG) void k_kin(KinArgs a) {
  extern __shared__ double2 smem_d[];
  C* smem = (C*)smem_d;
  C* twl = smem;                       // twiddles [N]
  using X = Ax<AX>;
  C* lds = smem + a.N;                 // lines [lpw][lpitch<X::contig>(N)]
  const C* fpsi = (const C*)a.psi;
  const C* fd = (const C*)a.d;
  C* fout = (C*)a.out;
  constexpr bool KE = (KIND == KE_FIRST || KIND == KE_REST);
  const int N = LC ? LC : a.N, tpl = LC ? lc_tpl(LC) : a.tpl, lpw = a.lpw, nl = a.nl;
  const int lb = xcd_block(blockIdx.x, gridDim.x);
  const int l0 = lb * lpw;
  const int total = a.nb * nl;
  const long long strd = X::stride(N);
  if (l0 >= total) return;
  // instances of the block's lines: b0 .. b1 (one when lpw divides nl)
  const int b0 = l0 / nl, b1 = min(l0 + lpw - 1, total - 1) / nl;
  const bool on0 = act_on(a.act, b0), on1 = (b1 == b0) ? on0 : act_on(a.act, b1);
  if (!on0 && !on1) return;            // lpw <= nl: no instance between b0 and b1
  // a workgroup's lines span instances b0 and b1 only
  const KinScal S0 = kin_scal_all<PC>(a, b0);
  const KinScal S1 = (b1 == b0) ? S0 : kin_scal_all<PC>(a, b1);
  CombScal C0{false, 0.0, 0.0, 0.0, 0.0, 0.0}, C1 = C0;
  if (KIND == KE_FIRST) {
    C0 = comb_scal(a, b0, on0);
    C1 = (b1 == b0) ? C0 : comb_scal(a, b1, on1);
  }
  const bool brow = on0;               // this block stores a partial row
  if (a.ucount && brow && threadIdx.x == 0 && l0 % nl == 0) atomicAdd(a.ucount, a.ubytes);
  stage_twiddles(twl, a.fp);
  // the line this thread loads, transforms and stores, and its elements lane + j tpl
  const LineOwn ow = line_owner<X::contig>(tpl, lpw);
  const int lane = ow.e0;
  const int ol = l0 + ow.cc;
  const int ob = ol / nl;
  const bool oon = ol < total && (ob == b0 ? on0 : on1);
  const int oli = ol - ob * nl;
  const long long obase = (long long)ob * a.istride + X::base(oli, N);
  const KinScal& OS = (ob == b0) ? S0 : S1;
  double c2[3] = {0.0, 0.0, 0.0};
  if (KIND == KE_FIRST || KIND == KH_LAST) line_sq<AX>(oli, N, a.dx, c2);
  C* s = lds + ow.cc * lpitch<X::contig>(N);
  double loc[4] = {0.0, 0.0, 0.0, 0.0};
  C x[BEC_EPT];
  // loads of the H passes' epilogue; on 2D axes they are issued before the
  // transforms (in 3D the extra registers would halve the waves per SIMD)
  constexpr bool EARLY = BEC_KH_EARLY2D && (KIND == KH_MID || KIND == KH_LAST) && X::ndim == 2;
  double2 ps[BEC_EPT], acc[BEC_EPT], old[BEC_EPT];
  double phv[BEC_EPT];
  {
    // phase 1: issue every load of this thread's elements (with the new
    // direction's z and g when this instance forms one)
    const CombScal& OC = (ob == b0) ? C0 : C1;
    const bool comb = KIND == KE_FIRST && oon && OC.on;
    double2 p0[BEC_EPT], dd[BEC_EPT], zz[BEC_EPT], gg[BEC_EPT];
#pragma unroll
    for (int j = 0; j < BEC_EPT; ++j) {
      const int e = ow.e0 + j * tpl;
      if (oon && e < N) {
        const long long adr = obase + e * strd;
        p0[j] = to_d2(fpsi[adr]);
        if (a.d) dd[j] = to_d2(fd[adr]);
        if (KIND == KE_FIRST && comb) {
          if (OC.az != 0.0) zz[j] = to_d2(((const C*)a.z)[adr]);
          if (OC.ag != 0.0) gg[j] = to_d2(((const C*)a.g)[adr]);
        }
      }
    }
    // new search directions (device-driven loop): d = az z + ad d + ag g + gam psi
    if (KIND == KE_FIRST && comb) {
      const double az = OC.az, ad = OC.ad, ag = OC.ag, gre = OC.gre, gim = OC.gim;
#pragma unroll
      for (int j = 0; j < BEC_EPT; ++j) {
        const int e = ow.e0 + j * tpl;
        if (e >= N) continue;
        double2 r = make_double2(gre * p0[j].x - gim * p0[j].y, gre * p0[j].y + gim * p0[j].x);
        if (az != 0.0) { r.x += az * zz[j].x; r.y += az * zz[j].y; }
        if (ad != 0.0) { r.x += ad * dd[j].x; r.y += ad * dd[j].y; }
        if (ag != 0.0) { r.x += ag * gg[j].x; r.y += ag * gg[j].y; }
        dd[j] = r;
        ((C*)a.d_out)[obase + e * strd] = from_d2<C>(r);
      }
    }
    // phase 2: the loaded values
#pragma unroll
    for (int j = 0; j < BEC_EPT; ++j) {
      const int e = ow.e0 + j * tpl;
      x[j] = cmk<C>(0, 0);
      if (e >= N) continue;
      double2 v = make_double2(0.0, 0.0);
      if (oon) {
        const long long adr = obase + e * strd;
        v = p0[j];
        if (a.d) {
          v.x = v.x - OS.t * dd[j].x;
          v.y = v.y - OS.t * dd[j].y;
        }
        if (OS.sc != 1.0) { v.x = v.x / OS.sc; v.y = v.y / OS.sc; }
        if (KIND == KH_FIRST || KIND == KH_LAST) {
          if (a.psi_w) ((C*)a.psi_w)[adr] = from_d2<C>(v);
          if (a.sdiff) ((C*)a.sdiff)[adr] = from_d2<C>(csub(v, p0[j]));
        }
        if (KIND == KE_FIRST) {
          const double rho = cabs2(v);
          const double V = line_V<AX>(c2, e, a);
          loc[0] += rho;
          loc[1] += V * rho;
          loc[2] += rho * rho;
          loc[3] += rho * rho * rho;
          if (a.rho_out) a.rho_out[adr] = rho;
        }
      }
      x[j] = from_d2<C>(v);
      if (EARLY && KIND == KH_LAST) ps[j] = to_d2(x[j]);    // the state as written to psi_w
    }
  }
  if (EARLY && oon) {
#pragma unroll
    for (int j = 0; j < BEC_EPT; ++j) {
      const int e = ow.e0 + j * tpl;
      if (e >= N) continue;
      const long long adr = obase + e * strd;
      acc[j] = a.acc ? to_d2(((const C*)a.acc)[adr]) : make_double2(0.0, 0.0);
      if (KIND == KH_LAST) {
        if (!a.psi_w) ps[j] = to_d2(fpsi[adr]);
        phv[j] = (a.local_on && a.phi) ? a.phi[adr] : 0.0;
        if (OS.pc) phv[j] = (phv[j] + OS.cq * a.kq[adr] + OS.cr * a.kr[adr]) * OS.cs;
        if (a.ydiff) old[j] = to_d2(fout[adr]);
      }
    }
  }
  fft_any<-1, ODD, X::contig, LC>(x, s, a.fp, twl, lane, tpl, oon);
  const double invN = 1.0 / (double)N;
  const double lzc = (X::lsign == 0 || !oon) ? 0.0 : (double)X::lsign * xcoord(X::lidx(oli, N), N, a.dx);
  const bool mult = !KE || (KIND == KE_FIRST && a.hy_out);
  const double rfl = oon ? OS.rf : 0.0;
  double kin = 0.0, lz = 0.0;
  if (oon) {
#pragma unroll
    for (int j = 0; j < BEC_EPT; ++j) {
      const int k = lane + j * tpl;
      if (k >= N) continue;
      const double kv = kfreq(k, N, a.invNdx);
      if (KE) {
        const double a2 = cabs2(x[j]);
        kin += 0.5 * kv * kv * a2;
        lz += kv * a2;
      }
      if (mult) x[j] = cscale(x[j], a.kf * kv * kv + rfl * lzc * kv);
    }
  }
  if (KE) {
    kin *= invN;
    lz *= lzc * invN;
    if (brow) {
      if (KIND == KE_FIRST) {
        double v6[6] = {kin, lz, loc[0], loc[1], loc[2], loc[3]};
        block_partials<6, 0>(v6, nullptr, a.partials, a.hpartials, lb);
      } else {
        double v2[2] = {kin, lz};
        block_partials<2, 0>(v2, nullptr, a.partials, a.hpartials, lb);
      }
    }
    if (!mult) return;
  }
  fft_any<+1, ODD, X::contig, LC>(x, s, a.fp, twl, lane, tpl, oon);
  double hsum[2] = {0.0, 0.0};
  double hmax[1] = {0.0};
  if (oon) {
    // the epilogue in chunks of EC elements: loads of a chunk (issued before
    // the transforms on 2D axes, all in one chunk there), then its results,
    // in element order
    constexpr int EC = EARLY ? BEC_EPT : BEC_KH_CHUNK;
#pragma unroll
    for (int j0 = 0; j0 < BEC_EPT; j0 += EC) {
#pragma unroll
      for (int j = j0; j < j0 + EC; ++j) {
        const int e = ow.e0 + j * tpl;
        if (EARLY || e >= N) continue;
        const long long adr = obase + e * strd;
        if (KIND == KH_MID || KIND == KH_LAST) acc[j] = a.acc ? to_d2(((const C*)a.acc)[adr]) : make_double2(0.0, 0.0);
        if (KIND == KH_LAST) {
          ps[j] = to_d2(fpsi[adr]);
          phv[j] = (a.local_on && a.phi) ? a.phi[adr] : 0.0;
          if (OS.pc) phv[j] = (phv[j] + OS.cq * a.kq[adr] + OS.cr * a.kr[adr]) * OS.cs;
          if (a.ydiff) old[j] = to_d2(fout[adr]);
        }
      }
#pragma unroll
      for (int j = j0; j < j0 + EC; ++j) {
        const int e = ow.e0 + j * tpl;
        if (e >= N) continue;
        const long long adr = obase + e * strd;
        double2 u = to_d2(x[j]);
        u.x *= invN;
        u.y *= invN;
        if (KE) {
          ((C*)a.hy_out)[adr] = from_d2<C>(u);
          continue;
        }
        if (KIND == KH_FIRST) {
          fout[adr] = from_d2<C>(u);
          continue;
        }
        double2 p = acc[j];
        if (a.acc_sc) {
          p.x /= OS.sc;
          p.y /= OS.sc;
        }
        if (KIND == KH_MID) {
          fout[adr] = from_d2<C>(cadd(p, u));
          continue;
        }
        const double2 psv = ps[j];
        const double rho = cabs2(psv);
        double W = 0.0;
        if (a.local_on) {
          double ph = 0.0;
          if (a.phi) {
            ph = phv[j] * OS.phisc;
            if (OS.pc) a.phi_out[adr] = phv[j];          // the accepted state's potential
            else if (a.phi_out) a.phi_out[adr] = ph;
          }
          W = line_V<AX>(c2, e, a) + OS.beta2 * rho + OS.beta3 * rho * rho + ph;
        }
        const double mu = OS.mu;
        double2 h;
        h.x = p.x + u.x + W * psv.x - mu * psv.x;
        h.y = p.y + u.y + W * psv.y - mu * psv.y;
        if (a.ydiff) ((C*)a.ydiff)[adr] = from_d2<C>(csub(h, old[j]));
        fout[adr] = from_d2<C>(h);
        hsum[0] += cabs2(h);
        hsum[1] += psv.x * h.x + psv.y * h.y;
        hmax[0] = fmax(hmax[0], rho);
      }
    }
  }
  if (KIND == KH_LAST && brow && (a.partials || a.hpartials))
    block_partials<2, 1>(hsum, hmax, a.partials, a.hpartials, lb);
}

// Line coefficients of the polynomial Armijo search (device-driven loop).
// Along the search line u(t) = psi - t d every term of the energy is a
// polynomial in t: with p = |psi|^2, q = Re(conj(psi) d), r = |d|^2 the
// density is rho(t) = p - 2 t q + t^2 r, so sum rho^2 has degree 4, sum rho^3
// degree 6, the norm and the quadratic forms of T + V - Omega L_z degree 2 and
// the long-range pair sum degree 4 (with K*q and K*r).  One pass over psi and
// d therefore prices every trial of the line search.  This is the pass of the
// first axis: it forms the new direction d = az z + ad d + ag g + gam psi
// (instances with the comb bit), accumulates the local sums, writes q and r
// for the long-range convolutions (rho_out, r_out), and transforms d along
// its axis for the kinetic and rotation forms of d (Parseval).  The other
// axes run KE_REST on d alone.  Sums of a row (LC_K):
//   0 kin(d) of this axis   1 L_z(d) of this axis   2 sum q   3 sum r
//   4 sum V r   5 p q   6 q^2   7 p r   8 q r   9 r^2
//   10 p^2 q   11 p^2 r   12 p q^2   13 q^3   14 p q r   15 q^2 r   16 p r^2
//   17 q r^2   18 r^3   19 phi q   20 phi r     (phi: the state's potential)
#define LC_K 21
template <int AX, class C, bool ODD = false, int LC = 0>
__global__ __launch_bounds__(BEC_WG) void k_lcoef(KinArgs a) {
  extern __shared__ double2 smem_d[];
  C* smem = (C*)smem_d;
  C* twl = smem;                       // twiddles [N]
  using X = Ax<AX>;
  C* lds = smem + a.N;                 // lines [lpw][lpitch<X::contig>(N)]
  const C* fpsi = (const C*)a.psi;
  const C* fd = (const C*)a.d;
  const int N = LC ? LC : a.N, tpl = LC ? lc_tpl(LC) : a.tpl, lpw = a.lpw, nl = a.nl;
  const int lb = xcd_block(blockIdx.x, gridDim.x);
  const int l0 = lb * lpw;
  const int total = a.nb * nl;
  const long long strd = X::stride(N);
  if (l0 >= total) return;
  const int b0 = l0 / nl, b1 = min(l0 + lpw - 1, total - 1) / nl;
  const bool on0 = act_on(a.act, b0), on1 = (b1 == b0) ? on0 : act_on(a.act, b1);
  if (!on0 && !on1) return;
  const CombScal C0 = comb_scal(a, b0, on0);
  const CombScal C1 = (b1 == b0) ? C0 : comb_scal(a, b1, on1);
  const bool brow = on0;
  if (a.ucount && brow && threadIdx.x == 0 && l0 % nl == 0) atomicAdd(a.ucount, a.ubytes);
  stage_twiddles(twl, a.fp);
  const LineOwn ow = line_owner<X::contig>(tpl, lpw);
  const int lane = ow.e0;
  const int ol = l0 + ow.cc;
  const int ob = ol / nl;
  const bool oon = ol < total && (ob == b0 ? on0 : on1);
  const int oli = ol - ob * nl;
  const long long obase = (long long)ob * a.istride + X::base(oli, N);
  double c2[3];
  line_sq<AX>(oli, N, a.dx, c2);
  C* s = lds + ow.cc * lpitch<X::contig>(N);
  double v[LC_K];
#pragma unroll
  for (int q = 0; q < LC_K; ++q) v[q] = 0.0;
  C x[BEC_EPT];
  {
    const CombScal& OC = (ob == b0) ? C0 : C1;
    const bool comb = oon && OC.on;
    double2 p0[BEC_EPT], dd[BEC_EPT], zz[BEC_EPT], gg[BEC_EPT];
    double fv[BEC_EPT];
    // phase 1: every load of the thread's elements
#pragma unroll
    for (int j = 0; j < BEC_EPT; ++j) {
      const int e = ow.e0 + j * tpl;
      fv[j] = 0.0;
      if (oon && e < N) {
        const long long adr = obase + e * strd;
        p0[j] = to_d2(fpsi[adr]);
        dd[j] = to_d2(fd[adr]);
        if (comb) {
          if (OC.az != 0.0) zz[j] = to_d2(((const C*)a.z)[adr]);
          if (OC.ag != 0.0) gg[j] = to_d2(((const C*)a.g)[adr]);
        }
        if (a.phi) fv[j] = a.phi[adr];
      }
    }
    // phase 2: the new direction, stored as the field's type holds it
    if (comb) {
      const double az = OC.az, ad = OC.ad, ag = OC.ag, gre = OC.gre, gim = OC.gim;
#pragma unroll
      for (int j = 0; j < BEC_EPT; ++j) {
        const int e = ow.e0 + j * tpl;
        if (e >= N) continue;
        double2 r = make_double2(gre * p0[j].x - gim * p0[j].y, gre * p0[j].y + gim * p0[j].x);
        if (az != 0.0) { r.x += az * zz[j].x; r.y += az * zz[j].y; }
        if (ad != 0.0) { r.x += ad * dd[j].x; r.y += ad * dd[j].y; }
        if (ag != 0.0) { r.x += ag * gg[j].x; r.y += ag * gg[j].y; }
        const C rc = from_d2<C>(r);
        ((C*)a.d_out)[obase + e * strd] = rc;
        dd[j] = to_d2(rc);
      }
    }
    // phase 3: the local sums and the fields of the convolutions
#pragma unroll
    for (int j = 0; j < BEC_EPT; ++j) {
      const int e = ow.e0 + j * tpl;
      x[j] = cmk<C>(0, 0);
      if (!oon || e >= N) continue;
      const double P = cabs2(p0[j]);
      const double Q = p0[j].x * dd[j].x + p0[j].y * dd[j].y;
      const double R = cabs2(dd[j]);
      const double V = line_V<AX>(c2, e, a);
      const double PQ = P * Q, PR = P * R, QQ = Q * Q, QR = Q * R, RR = R * R;
      v[2] += Q;        v[3] += R;        v[4] += V * R;
      v[5] += PQ;       v[6] += QQ;       v[7] += PR;       v[8] += QR;       v[9] += RR;
      v[10] += P * PQ;  v[11] += P * PR;  v[12] += P * QQ;  v[13] += Q * QQ;  v[14] += P * QR;
      v[15] += Q * QR;  v[16] += P * RR;  v[17] += Q * RR;  v[18] += R * RR;
      v[19] += fv[j] * Q;
      v[20] += fv[j] * R;
      if (a.rho_out) {
        const long long adr = obase + e * strd;
        a.rho_out[adr] = Q;
        a.r_out[adr] = R;
      }
      x[j] = from_d2<C>(dd[j]);
    }
  }
  fft_any<-1, ODD, X::contig, LC>(x, s, a.fp, twl, lane, tpl, oon);
  const double invN = 1.0 / (double)N;
  const double lzc = (X::lsign == 0 || !oon) ? 0.0 : (double)X::lsign * xcoord(X::lidx(oli, N), N, a.dx);
  double kin = 0.0, lz = 0.0;
  if (oon) {
#pragma unroll
    for (int j = 0; j < BEC_EPT; ++j) {
      const int k = lane + j * tpl;
      if (k >= N) continue;
      const double kv = kfreq(k, N, a.invNdx);
      const double a2 = cabs2(x[j]);
      kin += 0.5 * kv * kv * a2;
      lz += kv * a2;
    }
  }
  v[0] = kin * invN;
  v[1] = lz * lzc * invN;
  if (brow) block_partials<LC_K, 0>(v, nullptr, a.partials, a.hpartials, lb);
}

// The 16 inner products of one CG step (sums, scaled by dv on the host):
//  0,1 <psi,z>   2 <psi,psi>   3,4 <psi,d>   5,6 <psi,g>   7 Re<z,g>
//  8 Re<z,y>   9,10 <psi,y>   11 Re<g,d>   12 <g,g>   13 Re<g,y>
//  14 Re<s,y>  15 <s,s>    (y = g - g_prev, s = psi - psi_prev, <u,v> = sum conj(u) v)
__device__ __forceinline__ void dot_accum(double* v, double2 zi, double2 pi, double2 di, double2 gi, double2 y,
                                          double2 s) {
  v[0] += pi.x * zi.x + pi.y * zi.y;  v[1] += pi.x * zi.y - pi.y * zi.x;
  v[2] += cabs2(pi);
  v[3] += pi.x * di.x + pi.y * di.y;  v[4] += pi.x * di.y - pi.y * di.x;
  v[5] += pi.x * gi.x + pi.y * gi.y;  v[6] += pi.x * gi.y - pi.y * gi.x;
  v[7] += zi.x * gi.x + zi.y * gi.y;
  v[8] += zi.x * y.x + zi.y * y.y;
  v[9] += pi.x * y.x + pi.y * y.y;    v[10] += pi.x * y.y - pi.y * y.x;
  v[11] += gi.x * di.x + gi.y * di.y;
  v[12] += cabs2(gi);
  v[13] += gi.x * y.x + gi.y * y.y;
  v[14] += s.x * y.x + s.y * y.y;
  v[15] += cabs2(s);
}

// C: storage (and transform type) of the complex fields, as in k_kin; the
// real fields (LD_R input, ST_RE output, khat, dot) are FP64.  LC != 0: built
// for lines of LC points on the register plan of that length (fft_reg_c).
template <bool CONTIG, int LD, int DIR, int MUL, int ST, class C, bool ODD = false, int LC = 0>
__global__ __launch_bounds__(BEC_WG) void k_pass(PassArgs a) {
  using F = typename CT<C>::F;
  extern __shared__ double2 smem_d[];
  C* smem = (C*)smem_d;
  const int L = LC ? LC : a.fp.L, tpl = LC ? lc_tpl(LC) : a.tpl, lpw = a.lpw, nl = a.nl;
  C* twl = smem;                       // twiddles [L]
  C* lds = smem + L;                   // lines [lpw][lpitch<CONTIG>(L)]
  const int lb = xcd_block(blockIdx.x, gridDim.x);
  const int l0 = lb * lpw;
  const int total = a.nb * nl;
  if (l0 >= total) return;
  const int b0 = l0 / nl, b1 = min(l0 + lpw - 1, total - 1) / nl;
  const bool on0 = act_on(a.act, b0), on1 = (b1 == b0) ? on0 : act_on(a.act, b1);
  bool any = on0 || on1;
  for (int b = b0 + 1; b < b1 && !any; ++b) any = act_on(a.act, b);
  if (!any) return;
  auto inst_active = [&](int b) { return b == b0 ? on0 : (b == b1 ? on1 : act_on(a.act, b)); };
  const bool brow = on0;               // this block stores a partial row
  // measurement: the bytes of each active instance whose first line is here,
  // and the launch's shared bytes with its first active instance
  if (a.ucount && threadIdx.x == 0)
    for (int b = b0; b <= b1; ++b)
      if ((long long)b * nl >= l0 && inst_active(b)) {
        bool first = true;
        for (int c = 0; c < b && first && a.lbytes; ++c) first = !act_on(a.act, c);
        atomicAdd(a.ucount, a.ubytes + (first ? a.lbytes : 0ull));
      }
  stage_twiddles(twl, a.fp);
  // the line this thread loads, transforms and stores, and its elements lane + j tpl
  const LineOwn ow = line_owner<CONTIG>(tpl, lpw);
  const int lane = ow.e0;
  const int ol = l0 + ow.cc;
  const int ob = ol / nl;
  const bool oon = ol < total && inst_active(ob);
  const int oli = ol - ob * nl;
  const int oo = oli / a.ni, oi = oli - oo * a.ni;
  C* s = lds + ow.cc * lpitch<CONTIG>(L);
  C v[BEC_EPT];
  // kernel spectrum of the line's elements, loaded with the line itself
  double kh[(MUL == MUL_KHAT2 || MUL == MUL_KHAT3) ? BEC_EPT : 1];
  {
    const long long ibase = ob * a.in_ib + oo * a.in_so + oi * a.in_si;
#pragma unroll
    for (int j = 0; j < BEC_EPT; ++j) {
      const int e = lane + j * tpl;
      v[j] = cmk<C>(0, 0);
      if (!oon || e >= L) continue;
      if (LD == LD_HERM) {
        const int half = L / 2;
        const int src = (e <= half) ? e : L - e;
        const C w = ((const C*)a.in)[ibase + src * a.in_s];
        v[j] = (e <= half) ? w : cmk<C>(w.x, -w.y);
      } else {
        const int r = e - a.off_load;
        if (r >= 0 && r < a.n_load) {
          if (LD == LD_C) v[j] = ((const C*)a.in)[ibase + r * a.in_s];
          else if (LD == LD_R2)
            v[j] = cmk<C>((F)((const double*)a.in)[ibase + r * a.in_s], (F)((const double*)a.in2)[ibase + r * a.in_s]);
          else v[j] = cmk<C>((F)((const double*)a.in)[ibase + r * a.in_s], 0);
        }
      }
    }
    if (MUL == MUL_KHAT2 || MUL == MUL_KHAT3) {
      // 2D spectrum [kx][ky] (element stride H); 3D spectrum [ky][kx][kz], the
      // line's kz column contiguous.  The line's other frequencies run over
      // [0, H) (half spectrum) or [0, M) (kfull: a pair convolution), both
      // folded to [0, H) by the spectrum's evenness.
      long long kb;
      const int nz = a.kfull ? a.M : a.H;
      if (MUL == MUL_KHAT2) {
        kb = min(oli, a.M - oli);                      // column index ky
      } else {
        const int ky = oli / nz, kx = oli - ky * nz;
        kb = ((long long)min(ky, a.M - ky) * a.H + min(kx, a.M - kx)) * a.H;
      }
      const long long kst = (MUL == MUL_KHAT2) ? (long long)a.H : 1;
#pragma unroll
      for (int j = 0; j < BEC_EPT; ++j) {
        const int k = lane + j * tpl;
        kh[j] = (oon && k < L) ? a.khat[kb + min(k, L - k) * kst] : 0.0;
      }
    }
  }
  const double sig = ((MUL == MUL_PRE2 || MUL == MUL_PRE3) && oon) ? a.sigma[ob] : 0.0;
  if (DIR == DIR_INV) fft_any<+1, ODD, CONTIG, LC>(v, s, a.fp, twl, lane, tpl, oon);
  else fft_any<-1, ODD, CONTIG, LC>(v, s, a.fp, twl, lane, tpl, oon);
  if (DIR == DIR_FMI) {
    if (oon) {
      if (MUL == MUL_KHAT2 || MUL == MUL_KHAT3) {
#pragma unroll
        for (int j = 0; j < BEC_EPT; ++j) v[j] = cscale(v[j], kh[j]);
      } else if (MUL == MUL_GAUSS) {
#pragma unroll
        for (int j = 0; j < BEC_EPT; ++j) {
          const int k = lane + j * tpl;
          if (k >= L) continue;
          const double kx = kfreq(k, a.Nk, a.invNdx);
          v[j] = cscale(v[j], exp(-(a.gl2s ? a.gl2s[ob] : a.gl2) * kx * kx));
        }
      } else {
        // denominator 1/2 (kx^2 + ky^2 [+ kz^2]) + sigma, summed in that order
        const double ky = kfreq(MUL == MUL_PRE2 ? oli : oli / a.Nk, a.Nk, a.invNdx);
        const double kz = (MUL == MUL_PRE3) ? kfreq(oli % a.Nk, a.Nk, a.invNdx) : 0.0;
#pragma unroll
        for (int j = 0; j < BEC_EPT; ++j) {
          const int k = lane + j * tpl;
          if (k >= L) continue;
          const double kx = kfreq(k, a.Nk, a.invNdx);
          double k2 = kx * kx + ky * ky;
          if (MUL == MUL_PRE3) k2 = k2 + kz * kz;
          const double den = 0.5 * k2 + sig;
          v[j] = cmk<C>(v[j].x / (F)den, v[j].y / (F)den);
        }
      }
    }
    fft_any<+1, ODD, CONTIG, LC>(v, s, a.fp, twl, lane, tpl, oon);
  }
  // stores: line position q = lane + j tpl goes to output element q - off_store
  constexpr bool RE = (ST == ST_RE || ST == ST_RE2);
  double dsum[3] = {0.0, 0.0, 0.0};
  double zv[ST == ST_CDOT ? 16 : 1];
#pragma unroll
  for (int q = 0; q < (ST == ST_CDOT ? 16 : 1); ++q) zv[q] = 0.0;
  if (oon) {
    const long long obase = ob * a.out_ib + oo * a.out_so + oi * a.out_si;
    // elements in chunks of EC: the five fields of the inner products are
    // loaded for EC elements at a time (fewer registers), summed in the same
    // element order as one chunk would
    constexpr int EC = (ST == ST_CDOT) ? BEC_CDOT_CHUNK : BEC_EPT;
#pragma unroll
    for (int j0 = 0; j0 < BEC_EPT; j0 += EC) {
      double dv[EC], dv2[EC];
      C fp[ST == ST_CDOT ? EC : 1], fd[ST == ST_CDOT ? EC : 1], fg[ST == ST_CDOT ? EC : 1];
      C fy[ST == ST_CDOT ? EC : 1], fs[ST == ST_CDOT ? EC : 1];
#pragma unroll
      for (int i = 0; i < EC; ++i) {
        const int j = j0 + i;
        const int q = lane + j * tpl, e = q - a.off_store;
        if (q >= L || e < 0 || e >= a.n_store) continue;
        const long long adr = obase + e * a.out_s;
        if (RE && a.dot) dv[i] = a.dot[adr];
        if (RE && a.dot2) dv2[i] = a.dot2[adr];
        if (ST == ST_CDOT) {
          fp[i] = ((const C*)a.dpsi)[adr]; fd[i] = ((const C*)a.dd)[adr]; fg[i] = ((const C*)a.dg)[adr];
          fy[i] = ((const C*)a.dy)[adr]; fs[i] = ((const C*)a.ds)[adr];
        }
      }
#pragma unroll
      for (int i = 0; i < EC; ++i) {
        const int j = j0 + i;
        const int q = lane + j * tpl, e = q - a.off_store;
        if (q >= L || e < 0 || e >= a.n_store) continue;
        const long long adr = obase + e * a.out_s;
        const C w = v[j];
        if (ST == ST_CDOT) {
          const C zi = cscale(w, a.out_scale);
          ((C*)a.out)[adr] = zi;
          dot_accum(zv, to_d2(zi), to_d2(fp[i]), to_d2(fd[i]), to_d2(fg[i]), to_d2(fy[i]), to_d2(fs[i]));
        } else if (ST == ST_C) {
          ((C*)a.out)[adr] = cscale(w, a.out_scale);
        } else if (ST == ST_RE2) {
          const double r = (double)w.x * a.out_scale, im = (double)w.y * a.out_scale;
          ((double*)a.out)[adr] = r;
          ((double*)a.out2)[adr] = im;
          if (a.dot) {
            dsum[0] += r * dv[i];
            dsum[1] += r * dv2[i];
            dsum[2] += im * dv2[i];
          }
        } else {
          const double r = (double)w.x * a.out_scale;
          ((double*)a.out)[adr] = r;
          if (a.dot) dsum[0] += r * dv[i];
          if (a.dot2) dsum[1] += r * dv2[i];
        }
      }
    }
  }
  if (ST == ST_RE && a.dot && brow) {
    if (a.dot2) block_partials<2, 0>(dsum, nullptr, a.partials, a.hpartials, lb);
    else block_partials<1, 0>(dsum, nullptr, a.partials, a.hpartials, lb);
  }
  if (ST == ST_CDOT && brow) block_partials<16, 0>(zv, nullptr, a.partials, a.hpartials, lb);
  if (ST == ST_RE2 && a.dot && brow) {
    block_partials<2, 0>(dsum, nullptr, a.partials, nullptr, lb);
    block_partials<1, 0>(dsum + 2, nullptr, a.partials2, nullptr, lb);
  }
}

#if BEC_TU_IS(0)
// ---------------------------------------------------------------------------
// elementwise kernels: grid (chunks, instances); reducing ones store row
// blockIdx.y * gridDim.x + blockIdx.x
// ---------------------------------------------------------------------------
#define ELEM_LOOP(n) \
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < (n); i += (long long)gridDim.x * blockDim.x)

struct DblArr {
  double v[BEC_BMAX];
};

__global__ __launch_bounds__(BEC_WG) void k_norm2(const double2* __restrict__ psi, long long n,
                                                  unsigned long long act, double* hpart) {
  const int b = blockIdx.y;
  if (!inst_on(act, b)) return;
  const double2* p = psi + (long long)b * n;
  double v[1] = {0.0};
  ELEM_LOOP(n) v[0] += cabs2(p[i]);
  block_partials<1, 0>(v, nullptr, nullptr, hpart, (long long)b * gridDim.x + blockIdx.x);
}

// out = in / div[b] for the instances of act (in place when in == out)
__global__ __launch_bounds__(BEC_WG) void k_scale(const double2* in, double2* out, DblArr div, long long n,
                                                  unsigned long long act) {
  const int b = blockIdx.y;
  if (!((act >> b) & 1ull)) return;
  const double sc = div.v[b];
  ELEM_LOOP(n) {
    const double2 v = in[(long long)b * n + i];
    out[(long long)b * n + i] = make_double2(v.x / sc, v.y / sc);
  }
}

__global__ __launch_bounds__(BEC_WG) void k_rho(const double2* __restrict__ in, double* out, long long n) {
  const long long off = (long long)blockIdx.y * n;
  ELEM_LOOP(n) out[off + i] = cabs2(in[off + i]);
}

__global__ __launch_bounds__(BEC_WG) void k_copy(const double2* __restrict__ src, double2* dst, long long n,
                                                 Act mask) {
  const int b = blockIdx.y;
  if (!act_on(mask, b)) return;
  const long long off = (long long)b * n;
  ELEM_LOOP(n) dst[off + i] = src[off + i];
}

// d <- az z + ad d + ag g + gam psi   (gam complex), per instance; the
// coefficients come from the launch arguments or the device table tab
__global__ __launch_bounds__(BEC_WG) void k_combine(double2* d, const double2* __restrict__ z,
                                                    const double2* __restrict__ g,
                                                    const double2* __restrict__ psi, CombArgs c,
                                                    const double* __restrict__ tab, long long n, Act act) {
  const int b = blockIdx.y;
  if (!act_on(act, b)) return;
  const long long off = (long long)b * n;
  double az = c.az[b], ad = c.ad[b], ag = c.ag[b], gre = c.gre[b], gim = c.gim[b];
  if (tab) {
    const double* t = tab + (long long)b * DS_N;
    az = t[DS_AZ]; ad = t[DS_AD]; ag = t[DS_AG]; gre = t[DS_GRE]; gim = t[DS_GIM];
  }
  if (!g) ag = 0.0;
  ELEM_LOOP(n) {
    const long long j = off + i;
    const double2 p = psi[j];
    double2 r = make_double2(gre * p.x - gim * p.y, gre * p.y + gim * p.x);
    if (az != 0.0) { const double2 v = z[j]; r.x += az * v.x; r.y += az * v.y; }
    if (ad != 0.0) { const double2 v = d[j]; r.x += ad * v.x; r.y += ad * v.y; }
    if (ag != 0.0) { const double2 v = g[j]; r.x += ag * v.x; r.y += ag * v.y; }
    d[j] = r;
  }
}

// the CG inner products (dot_accum) of whole fields
__global__ __launch_bounds__(BEC_WG) void k_dots(const double2* __restrict__ z,
                                                 const double2* __restrict__ psi,
                                                 const double2* __restrict__ d,
                                                 const double2* __restrict__ g,
                                                 const double2* __restrict__ yv,
                                                 const double2* __restrict__ sv, long long n,
                                                 Act act, double* partials, double* hpart) {
  const int b = blockIdx.y;
  if (!act_on(act, b)) return;
  const long long off = (long long)b * n;
  double v[16];
#pragma unroll
  for (int q = 0; q < 16; ++q) v[q] = 0.0;
  ELEM_LOOP(n) {
    const long long j = off + i;
    dot_accum(v, z[j], psi[j], d[j], g[j], yv[j], sv[j]);
  }
  block_partials<16, 0>(v, nullptr, partials, hpart, (long long)b * gridDim.x + blockIdx.x);
}

// g = h - mu psi, and sum |g|^2 (one instance)
// after the Armijo controller: keep the lowest-residual state, and after a
// failed line search restart the direction from the preconditioned gradient
template <class C>
__global__ __launch_bounds__(BEC_WG) void k_post(const C* __restrict__ psi, C* best, C* d,
                                                 const C* __restrict__ z, const double* __restrict__ tab,
                                                 const unsigned* __restrict__ fl, long long n) {
  const int b = blockIdx.y;
  const unsigned f = fl[b];
  const bool cp = (f & FL_BEST) != 0u, fail = (f & FL_FAIL) != 0u;
  if (!cp && !fail) return;
  const long long off = (long long)b * n;
  const double* t = tab + (long long)b * DS_N;
  const double az = t[DS_AZ], gre = t[DS_GRE], gim = t[DS_GIM];
  ELEM_LOOP(n) {
    const long long j = off + i;
    const C pc = psi[j];
    if (cp) best[j] = pc;
    if (fail) {
      const double2 zz = to_d2(z[j]), p = to_d2(pc);
      d[j] = from_d2<C>(make_double2(az * zz.x + gre * p.x - gim * p.y, az * zz.y + gre * p.y + gim * p.x));
    }
  }
}

__global__ __launch_bounds__(BEC_WG) void k_res2(const double2* __restrict__ h,
                                                 const double2* __restrict__ psi, double mu,
                                                 double2* g, long long n, double* hpart) {
  double v[1] = {0.0};
  ELEM_LOOP(n) {
    const double2 hi = h[i], p = psi[i];
    const double2 r = make_double2(hi.x - mu * p.x, hi.y - mu * p.y);
    if (g) g[i] = r;
    v[0] += cabs2(r);
  }
  block_partials<1, 0>(v, nullptr, nullptr, hpart, blockIdx.x);
}

// moments (one instance): mass, sum R2 rho, sum (X^2+Y^2) rho, sum Z^2 rho; max rho
// per instance (blockIdx.y): sum rho, r^2 rho, (x^2 + y^2) rho, z^2 rho, max rho
__global__ __launch_bounds__(BEC_WG) void k_moments(const double2* __restrict__ psi0, int ndim, int N,
                                                    double dx, double* partials, double* hpart) {
  double v[4] = {0.0, 0.0, 0.0, 0.0};
  double m[1] = {0.0};
  const long long n = (ndim == 2) ? (long long)N * N : (long long)N * N * N;
  const double2* psi = psi0 + (long long)blockIdx.y * n;
  ELEM_LOOP(n) {
    const double rho = cabs2(psi[i]);
    int ix, iy, iz = 0;
    if (ndim == 2) { ix = (int)(i / N); iy = (int)(i % N); }
    else { ix = (int)(i / ((long long)N * N)); iy = (int)((i / N) % N); iz = (int)(i % N); }
    const double xx = xcoord(ix, N, dx), yy = xcoord(iy, N, dx);
    const double zz = (ndim == 3) ? xcoord(iz, N, dx) : 0.0;
    double r2 = xx * xx + yy * yy;
    if (ndim == 3) r2 += zz * zz;
    v[0] += rho;
    v[1] += r2 * rho;
    v[2] += (xx * xx + yy * yy) * rho;
    v[3] += zz * zz * rho;
    m[0] = fmax(m[0], rho);
  }
  block_partials<4, 1>(v, m, partials, hpart, (long long)blockIdx.y * gridDim.x + blockIdx.x);
}

// ---------------------------------------------------------------------------
// Sobolev shift: lower median of the local potential over the bulk
// {rho > 0.05 max rho}, per instance (grid y), in three launches:
//   1. the masked values, their min / max / count rows, and a histogram of
//      SIGMA_BINS bins over the previous evaluation's value range (values
//      outside it fall in the end bins: the bins stay ordered, which is all the
//      selection needs);
//   2. the rows folded, the bin holding rank (cnt-1)/2 located, its members
//      gathered;
//   3. exact selection among them in one workgroup; the range just found
//      becomes the next evaluation's bin range.
// No atomics across the grid on a single address except the candidate counter.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int sigma_bin(double v, double lo, double scale) {
  int b = (int)((v - lo) * scale);
  return b < 0 ? 0 : (b >= SIGMA_BINS ? SIGMA_BINS - 1 : b);
}
// the bins' value range: [lo, hi] with SIGMA_BINS / (hi - lo) bins per unit,
// or a single bin when the range is empty
__device__ __forceinline__ void sigma_range(const SigmaStats* st, double& lo, double& scale) {
  lo = key_to_double(st->blo);
  const double hi = key_to_double(st->bhi);
  scale = (hi > lo) ? (double)SIGMA_BINS / (hi - lo) : 0.0;
}
template <class C>
__global__ __launch_bounds__(BEC_WG) void k_sigma_local(const C* __restrict__ psi,
                                                        const double* __restrict__ phi,
                                                        const double* mrows, int nrows, int mkp, int mq,
                                                        const InstPhys* iphys, int ndim, int N, double dx,
                                                        double tc, int s_mode, double s_half, long long n,
                                                        Act act, double* loc, unsigned long long* part,
                                                        const SigmaStats* stats, unsigned* hists) {
  __shared__ unsigned long long smin[BEC_WG / 64], smax[BEC_WG / 64], scnt[BEC_WG / 64];
  __shared__ double smx[BEC_WG / 64];
  __shared__ unsigned h[SIGMA_BINS];
  const int b = blockIdx.y;
  if (!act_on(act, b)) return;
  const long long off = (long long)b * n;
  for (int k = threadIdx.x; k < SIGMA_BINS; k += blockDim.x) h[k] = 0u;
  double blo, bscale;
  sigma_range(stats + b, blo, bscale);
  // max rho of the state: fold the per-block maxima of the pass that computed it
  double rm = 0.0;
  const double* mr = mrows + (long long)b * nrows * mkp;
  for (int r = threadIdx.x; r < nrows; r += blockDim.x) rm = fmax(rm, mr[(long long)r * mkp + mq]);
  rm = wave_max(rm);
  if ((threadIdx.x & 63) == 0) smx[threadIdx.x >> 6] = rm;
  __syncthreads();
  rm = smx[0];
  for (int w = 1; w < (int)(blockDim.x >> 6); ++w) rm = fmax(rm, smx[w]);
  const double thr = 0.05 * rm;
  const double b2 = iphys[b].beta2, b3 = iphys[b].beta3;
  unsigned long long kmn = ~0ull, kmx = 0ull, cnt = 0ull;
  // SIGMA_U elements per thread and round, their loads issued first
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long i0 = blockIdx.x * (long long)blockDim.x + threadIdx.x; i0 < n; i0 += SIGMA_U * stride) {
    C pv[SIGMA_U];
    double fv[SIGMA_U];
#pragma unroll
    for (int u = 0; u < SIGMA_U; ++u) {
      const long long i = i0 + u * stride;
      if (i < n) {
        pv[u] = psi[off + i];
        fv[u] = phi ? phi[off + i] : 0.0;
      }
    }
#pragma unroll
    for (int u = 0; u < SIGMA_U; ++u) {
      const long long i = i0 + u * stride;
      if (i >= n) continue;
      const double rho = cabs2(pv[u]);
      double val = __longlong_as_double(0x7ff8000000000000ll);    // NaN = outside the bulk
      if (rho > thr) {
        int ix, iy, iz = 0;
        if (ndim == 2) { ix = (int)(i / N); iy = (int)(i % N); }
        else { ix = (int)(i / ((long long)N * N)); iy = (int)((i / N) % N); iz = (int)(i % N); }
        const double xx = xcoord(ix, N, dx), yy = xcoord(iy, N, dx);
        double r2 = xx * xx + yy * yy;
        if (ndim == 3) { const double zz = xcoord(iz, N, dx); r2 += zz * zz; }
        double lv = trap_value(r2, tc, s_mode, s_half) + 2.0 * b2 * rho + 3.0 * b3 * (rho * rho);
        if (phi) lv = lv + fv[u];
        val = lv;
        const unsigned long long k = order_key(lv);
        kmn = k < kmn ? k : kmn;
        kmx = k > kmx ? k : kmx;
        cnt++;
        atomicAdd(&h[sigma_bin(lv, blo, bscale)], 1u);
      }
      loc[off + i] = val;
    }
  }
  for (int o2 = 32; o2 > 0; o2 >>= 1) {
    const unsigned long long x = __shfl_down(kmn, o2, 64), y = __shfl_down(kmx, o2, 64);
    kmn = x < kmn ? x : kmn;
    kmx = y > kmx ? y : kmx;
    cnt += __shfl_down(cnt, o2, 64);
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) { smin[wid] = kmn; smax[wid] = kmx; scnt[wid] = cnt; }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < (int)(blockDim.x >> 6); ++w) {
      kmn = smin[w] < kmn ? smin[w] : kmn;
      kmx = smax[w] > kmx ? smax[w] : kmx;
      cnt += scnt[w];
    }
    unsigned long long* pr = part + ((long long)b * gridDim.x + blockIdx.x) * 4;
    pr[0] = kmn;
    pr[1] = kmx;
    pr[2] = cnt;
  }
  __syncthreads();
  unsigned* hist = hists + (long long)b * SIGMA_BINS;
  for (int k = threadIdx.x; k < SIGMA_BINS; k += blockDim.x)
    if (h[k]) atomicAdd(&hist[k], h[k]);
}

// fold the local pass's per-block (min, max, count) rows of one instance;
// every block does it (at most SIGMA_MAX_BLOCKS rows), block 0 records it
__device__ __forceinline__ void sigma_fold(const unsigned long long* part, int nb1, SigmaStats* st,
                                           unsigned long long& kmn, unsigned long long& kmx,
                                           unsigned long long& cnt) {
  __shared__ unsigned long long a[BEC_WG], c[BEC_WG], d[BEC_WG];
  unsigned long long mn = ~0ull, mx = 0ull, ct = 0ull;
  for (int r = threadIdx.x; r < nb1; r += blockDim.x) {
    const unsigned long long x = part[4 * r], y = part[4 * r + 1];
    mn = x < mn ? x : mn;
    mx = y > mx ? y : mx;
    ct += part[4 * r + 2];
  }
  a[threadIdx.x] = mn; c[threadIdx.x] = mx; d[threadIdx.x] = ct;
  __syncthreads();
  for (int h = blockDim.x / 2; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) {
      a[threadIdx.x] = a[threadIdx.x + h] < a[threadIdx.x] ? a[threadIdx.x + h] : a[threadIdx.x];
      c[threadIdx.x] = c[threadIdx.x + h] > c[threadIdx.x] ? c[threadIdx.x + h] : c[threadIdx.x];
      d[threadIdx.x] += d[threadIdx.x + h];
    }
    __syncthreads();
  }
  kmn = a[0]; kmx = c[0]; cnt = d[0];
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->kmin = kmn;
    st->kmax = kmx;
    st->cnt = (unsigned)cnt;
  }
  __syncthreads();
}

// every block locates the bin holding the rank-k element, k = (cnt-1)/2
// (parallel prefix over the histogram), then gathers that bin's members
__global__ __launch_bounds__(BEC_WG) void k_sigma_gather(const double* __restrict__ loc, long long n,
                                                         const unsigned long long* part, int nb1,
                                                         SigmaStats* stats, const unsigned* hists, double* cands,
                                                         Act act) {
  __shared__ unsigned pre[BEC_WG];
  __shared__ int sbin;
  __shared__ unsigned srank;
  const int b = blockIdx.y;
  if (!act_on(act, b)) return;
  SigmaStats* st = stats + b;
  const unsigned* hist = hists + (long long)b * SIGMA_BINS;
  double* cand = cands + (long long)b * n;
  const long long off = (long long)b * n;
  unsigned long long kmn, kmx, cnt64;
  sigma_fold(part + (long long)b * nb1 * 4, nb1, st, kmn, kmx, cnt64);
  const unsigned cnt = (unsigned)cnt64;
  if (!(cnt > 0 && key_to_double(kmx) > key_to_double(kmn))) return;
  double lo, scale;
  sigma_range(st, lo, scale);
  const unsigned k = (cnt - 1u) / 2u;
  constexpr int PER = SIGMA_BINS / BEC_WG;
  unsigned own = 0;
  for (int j = 0; j < PER; ++j) own += hist[threadIdx.x * PER + j];
  pre[threadIdx.x] = own;
  if (threadIdx.x == 0) {              // defined even if no bin holds rank k
    sbin = 0;
    srank = 0u;
  }
  __syncthreads();
  for (int o2 = 1; o2 < BEC_WG; o2 <<= 1) {          // inclusive Hillis-Steele scan
    const unsigned add = (int)threadIdx.x >= o2 ? pre[threadIdx.x - o2] : 0u;
    __syncthreads();
    pre[threadIdx.x] += add;
    __syncthreads();
  }
  const unsigned before = pre[threadIdx.x] - own;
  if (before <= k && k < pre[threadIdx.x]) {           // exactly one thread
    unsigned run = before;
    int bb = threadIdx.x * PER;
    for (int j = 0; j < PER - 1; ++j, ++bb) {
      const unsigned hb = hist[bb];
      if (run + hb > k) break;
      run += hb;
    }
    sbin = bb;
    srank = k - run;
  }
  __syncthreads();
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->bucket = (unsigned)sbin;
    st->rank_in = srank;
  }
  const int bin = sbin;
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long i0 = blockIdx.x * (long long)blockDim.x + threadIdx.x; i0 < n; i0 += SIGMA_U * stride) {
    double v[SIGMA_U];
#pragma unroll
    for (int u = 0; u < SIGMA_U; ++u) {
      const long long i = i0 + u * stride;
      v[u] = (i < n) ? loc[off + i] : __longlong_as_double(0x7ff8000000000000ll);
    }
#pragma unroll
    for (int u = 0; u < SIGMA_U; ++u)
      if (v[u] == v[u] && sigma_bin(v[u], lo, scale) == bin) {
        const unsigned at = atomicAdd(&st->ncand, 1u);
        if (at < (unsigned long long)n) cand[at] = v[u];     // an instance holds n candidates at most
      }
  }
}

// one workgroup per instance: exact selection (bitonic sort in LDS, or an
// 8-bit radix select over the order keys for a large bin), sigma =
// max(median, floor), then reset the histogram and statistics
__global__ __launch_bounds__(BEC_WG) void k_sigma_select(SigmaStats* stats, unsigned* hists,
                                                         const double* __restrict__ cands, long long n,
                                                         double floor_, double* sigmas, double* hred, Act act) {
  __shared__ double sv[SIGMA_LDS_CAND];
  __shared__ unsigned dh[256];
  __shared__ unsigned long long prefix;
  __shared__ unsigned rank_s;
  const int b = blockIdx.x;
  if (!act_on(act, b)) return;
  SigmaStats* st = stats + b;
  unsigned* hist = hists + (long long)b * SIGMA_BINS;
  const double* cand = cands + (long long)b * n;
  const unsigned cnt = st->cnt;
  const double lo = key_to_double(st->kmin), hi = key_to_double(st->kmax);
  double med = floor_;
  bool have = false;
  if (cnt > 0 && !(hi > lo)) {
    med = lo;
    have = true;
  } else if (cnt > 0) {
    // the candidates read stay within the instance's n slots, the rank within them
    const unsigned m = (unsigned)min((unsigned long long)st->ncand, (unsigned long long)n);
    const unsigned r = m ? min(st->rank_in, m - 1u) : 0u;
    if (m <= SIGMA_LDS_CAND) {
      unsigned P = 1;
      while (P < m) P <<= 1;
      for (unsigned i = threadIdx.x; i < P; i += blockDim.x) sv[i] = (i < m) ? cand[i] : INFINITY;
      __syncthreads();
      for (unsigned kk = 2; kk <= P; kk <<= 1) {
        for (unsigned j = kk >> 1; j > 0; j >>= 1) {
          for (unsigned i = threadIdx.x; i < P; i += blockDim.x) {
            const unsigned ixj = i ^ j;
            if (ixj > i) {
              const double x = sv[i], y = sv[ixj];
              const bool up = ((i & kk) == 0);
              if ((x > y) == up) { sv[i] = y; sv[ixj] = x; }
            }
          }
          __syncthreads();
        }
      }
      med = sv[r];
      have = true;
    } else {
      if (threadIdx.x == 0) { prefix = 0ull; rank_s = r; }
      __syncthreads();
      for (int shift = 56; shift >= 0; shift -= 8) {
        for (int q = threadIdx.x; q < 256; q += blockDim.x) dh[q] = 0u;
        __syncthreads();
        const unsigned long long pr = prefix;
        const unsigned long long hmask = (shift == 56) ? 0ull : (~0ull << (shift + 8));
        for (unsigned i = threadIdx.x; i < m; i += blockDim.x) {
          const unsigned long long kk = order_key(cand[i]);
          if ((kk & hmask) == pr) atomicAdd(&dh[(kk >> shift) & 0xffu], 1u);
        }
        __syncthreads();
        if (threadIdx.x == 0) {
          unsigned run = 0, rr = rank_s;
          int q = 0;
          for (; q < 255; ++q) {
            if (run + dh[q] > rr) break;
            run += dh[q];
          }
          prefix = pr | ((unsigned long long)q << shift);
          rank_s = rr - run;
        }
        __syncthreads();
      }
      med = key_to_double(prefix);
      have = true;
    }
  }
  __syncthreads();
  for (int q = threadIdx.x; q < SIGMA_BINS; q += blockDim.x) hist[q] = 0u;
  if (threadIdx.x == 0) {
    double sg = have ? med : floor_;
    sg = fmax(sg, floor_);
    sigmas[b] = sg;
    if (hred) hred[(long long)b * RSTRIDE + R_SIGMA] = sg;
    st->blo = st->kmin;                // the next evaluation bins over this range
    st->bhi = st->kmax;
    st->kmin = ~0ull;
    st->kmax = 0ull;
    st->cnt = 0u;
    st->bucket = 0u;
    st->rank_in = 0u;
    st->ncand = 0u;
  }
}

__global__ __launch_bounds__(BEC_WG) void k_sigma_reset(SigmaStats* stats, unsigned* hists, double* sigmas) {
  const int b = blockIdx.x;
  for (int q = threadIdx.x; q < SIGMA_BINS; q += blockDim.x) hists[(long long)b * SIGMA_BINS + q] = 0u;
  if (threadIdx.x == 0) {
    SigmaStats* st = stats + b;
    st->kmin = ~0ull;
    st->kmax = 0ull;
    st->blo = 0ull;                    // no range yet: one bin
    st->bhi = 0ull;
    st->cnt = 0u;
    st->bucket = 0u;
    st->rank_in = 0u;
    st->ncand = 0u;
    sigmas[b] = 0.0;
  }
}

// ---------------------------------------------------------------------------
// kernel spectrum construction (becgpp/interactions.py:117-148): Gauss-Legendre
// 4x4 (2D) / 4x4x4 (3D) cell average of the point kernel on the wrap-around
// displacement grid; the self cell is filled by the host afterwards.
// kind: 1 = newton (1/r), 2 = log (-ln r).
// ---------------------------------------------------------------------------
__constant__ double c_gl_nodes[4] = {-0.8611363115940526, -0.3399810435848563, 0.3399810435848563,
                                     0.8611363115940526};
__constant__ double c_gl_wts[4] = {0.3478548451374538, 0.6521451548625461, 0.6521451548625461,
                                   0.3478548451374538};

__global__ __launch_bounds__(BEC_WG) void k_kernel_real(int ndim, int M, double dx, int kind,
                                                        double2* out) {
  const long long n = (ndim == 2) ? (long long)M * M : (long long)M * M * M;
  ELEM_LOOP(n) {
    int jx, jy, jz = 0;
    if (ndim == 2) { jx = (int)(i / M); jy = (int)(i % M); }
    else { jx = (int)(i / ((long long)M * M)); jy = (int)((i / M) % M); jz = (int)(i % M); }
    const double DX = (double)(jx <= M / 2 ? jx : jx - M) * dx;
    const double DY = (double)(jy <= M / 2 ? jy : jy - M) * dx;
    const double DZ = (double)(jz <= M / 2 ? jz : jz - M) * dx;
    double ker = 0.0;
    if (ndim == 2) {
      for (int a = 0; a < 4; ++a) {
        for (int b = 0; b < 4; ++b) {
          const double ux = DX + 0.5 * dx * c_gl_nodes[a];
          const double uy = DY + 0.5 * dx * c_gl_nodes[b];
          const double r2 = ux * ux + uy * uy;
          const double val = (kind == 1) ? 1.0 / sqrt(r2) : -0.5 * log(r2);
          ker += 0.25 * c_gl_wts[a] * c_gl_wts[b] * val;
        }
      }
    } else {
      for (int a = 0; a < 4; ++a)
        for (int b = 0; b < 4; ++b)
          for (int c = 0; c < 4; ++c) {
            const double ux = DX + 0.5 * dx * c_gl_nodes[a];
            const double uy = DY + 0.5 * dx * c_gl_nodes[b];
            const double uz = DZ + 0.5 * dx * c_gl_nodes[c];
            ker += (c_gl_wts[a] * c_gl_wts[b] * c_gl_wts[c] / 8.0) / sqrt(ux * ux + uy * uy + uz * uz);
          }
    }
    out[i] = make_double2(ker, 0.0);
  }
}

__global__ __launch_bounds__(BEC_WG) void k_extract_khat(int ndim, int M, int H,
                                                         const double2* __restrict__ kf, double* khat) {
  const long long n = (ndim == 2) ? (long long)H * H : (long long)H * H * H;
  ELEM_LOOP(n) {
    long long src;
    if (ndim == 2) {
      src = (i / H) * M + (i % H);
    } else {
      // khat[(ky H + kx) H + kz] from the full [kz][ky][kx] spectrum
      const long long kz = i % H, kx = (i / H) % H, ky = i / ((long long)H * H);
      src = (kz * M + ky) * M + kx;
    }
    khat[i] = kf[src].x;
  }
}

// ---------------------------------------------------------------------------
// launch wrappers
// ---------------------------------------------------------------------------
int elem_blocks(long long n) {
  long long g = (n + BEC_WG - 1) / BEC_WG;
  if (g > 1024) g = 1024;
  if (g < 1) g = 1;
  return (int)g;
}

// Threads per line: L/8 for register-resident transforms, else the power of
// two covering L with at most 8 elements per thread.
LineCfg line_cfg(const FftPlan& fp) {
  const int L = fp.L;
  int tpl = 1;
  if (fp.reg) {
    tpl = L / 8;
  } else {
    const int need = fp.wide ? (L + 3) / 4 : (L + 7) / 8;
    while (tpl < need) tpl <<= 1;
    if (tpl > BEC_WG) tpl = BEC_WG;
  }
  LineCfg c;
  c.tpl = tpl;
  c.lpw = BEC_WG / tpl;
  return c;
}

// Lines per workgroup for a launch: whole wavefronts (at least one), at most
// BEC_WG threads, a power of two dividing the lines of one instance when
// possible (so that a reducing workgroup never spans two instances), aiming
// at ~1024 workgroups so that small problems still spread over all 256 CUs.
// A strided pass takes at least 4 adjacent columns per workgroup, so that
// every row segment it touches is 64 contiguous bytes.
int lpw_min(int tpl) {
  int l = 1;
  while ((l * tpl) % 64 != 0 && l * tpl < BEC_WG) l *= 2;
  return l;
}
#endif  // BEC_TU_IS(0)

// whole = false (passes that store no partial rows): a workgroup may span two
// instances, so l need not divide nl -- the 2D convolution's strided pass runs
// over the H = M/2 + 1 columns of the half spectrum, an odd count.
static int pick_lpw(int tpl, int nl, int total, bool strided, bool whole = true) {
  const int lmin = lpw_min(tpl);
  const int lmax = BEC_WG / tpl > lmin ? BEC_WG / tpl : lmin;
  const int want = total / 1024;
  auto fits = [&](int l) { return !whole || nl % l == 0; };
  int l = lmin;
  if (strided)
    while (l < 4 && l * 2 <= lmax && fits(l * 2)) l *= 2;
  while (l * 2 <= lmax && l * 2 <= want && fits(l * 2)) l *= 2;
  return l;
}
// The heuristic count scaled by 2^shift (the plan's autotuned schedule):
// halving keeps l a divisor of nl; doubling only while it still divides nl
// (whole) and the workgroup stays within BEC_WG threads.
static int shifted_lpw(int l, int tpl, int nl, int shift, bool whole = true) {
  const int lmin = lpw_min(tpl);
  while (shift > 0 && l * 2 * tpl <= BEC_WG && (!whole || nl % (l * 2) == 0)) { l *= 2; --shift; }
  while (shift < 0 && l / 2 >= lmin) { l /= 2; ++shift; }
  return l;
}
static unsigned long long lowmask(int nb) { return nb >= 64 ? ~0ull : ((1ull << nb) - 1ull); }
// Algorithmic bytes per instance of a line pass: every array element the pass
// must read or write, once (re-reads through the caches are not counted).
template <int LD, int ST>
static double pass_bytes(const PassArgs& a) {
  const double lines = a.nl;
  const double cel = a.f32 ? 8.0 : 16.0;              // complex element
  const double in_el = (LD == LD_R) ? 8.0 : (LD == LD_R2 ? 16.0 : cel);
  const double loaded = (LD == LD_HERM) ? (double)(a.fp.L / 2 + 1) : (double)a.n_load;
  double b = lines * loaded * in_el;
  const double st_el = (ST == ST_RE) ? 8.0 : (ST == ST_RE2 ? 16.0 : cel);
  b += lines * a.n_store * st_el;
  if ((ST == ST_RE || ST == ST_RE2) && a.dot) b += lines * a.n_store * 8.0;
  if ((ST == ST_RE || ST == ST_RE2) && a.dot2) b += lines * a.n_store * 8.0;
  if (ST == ST_CDOT) b += lines * a.n_store * 5.0 * cel;
  return b;
}
template <int KIND>
static double kin_bytes(const KinArgs& a) {
  const double n = (double)a.nl * a.N, c = (a.f32 ? 8.0 : 16.0) * n, r = 8.0 * n;
  double b = c;                                          // psi
  if (a.d) b += c;
  if (KIND == KE_FIRST) {
    if (a.rho_out) b += r;
    if (a.hy_out) b += c;
    if (a.d_out) b += 2.0 * c;                           // z read, new d written (direction steps)
  }
  if (KIND == KH_FIRST || KIND == KH_LAST) {
    if (a.psi_w) b += c;
    if (a.sdiff) b += c;
  }
  if (KIND == KH_FIRST || KIND == KH_MID || KIND == KH_LAST) b += c;   // out
  if ((KIND == KH_MID || KIND == KH_LAST) && a.acc) b += c;
  if (KIND == KH_LAST) {
    if (a.local_on && a.phi) b += r;
    if (a.phi_out) b += r;
    if (a.kq) b += 2.0 * r;                              // K*q, K*r of the accepted state's potential
    if (a.ydiff) b += 2.0 * c;                           // previous out read, difference written
  }
  return b;
}
// workgroup count rounded up to a whole number of XCD rounds (see xcd_block)
static int xcd_grid(int nlines, int lpw) {
  const int b = (nlines + lpw - 1) / lpw;
  return ((b + BEC_XCDS - 1) / BEC_XCDS) * BEC_XCDS;
}

// Calls f(std::integral_constant<int, LC>) with LC the compile-time length of
// the plan (reg_plan_len) when the kernels were built for it, else LC = 0
// (the generic transform).  Built: 128, 192, 256, 384 and, with P512, 512.
template <bool P512, class F>
static void with_lc(const FftPlan& fp, F&& f) {
  const int r = reg_plan_len(fp);
  switch (r ? r : lds_plan_len(fp)) {
    case 128: f(std::integral_constant<int, 128>{}); return;
    case 192: f(std::integral_constant<int, 192>{}); return;
    case 256: f(std::integral_constant<int, 256>{}); return;
    case 384: f(std::integral_constant<int, 384>{}); return;
    case 512:
      if constexpr (P512) {
        f(std::integral_constant<int, 512>{});
        return;
      }
      break;
    default: break;
  }
  f(std::integral_constant<int, 0>{});
}

template <bool CONTIG, int LD, int DIR, int MUL, int ST>
static int go_pass(const PassArgs& a0, hipStream_t s) {
  PassArgs a = a0;
  const int total = a.nb * a.nl;
  // only passes that store partial rows need whole workgroups per instance
  const bool whole = ST == ST_CDOT || ((ST == ST_RE || ST == ST_RE2) && a.dot);
  a.lpw = shifted_lpw(pick_lpw(a.tpl, a.nl, total, !CONTIG, whole), a.tpl, a.nl, a.lsh[CONTIG ? 0 : 1], whole);
  const int blocks = xcd_grid(total, a.lpw);
  const size_t csz = a.f32 ? sizeof(float2) : sizeof(double2);
  const size_t shm = ((size_t)a.lpw * lpitch<CONTIG>(a.fp.L) + a.fp.L) * csz;
  // site id: 2000 + 1000 contig + 100 load + 10 direction + multiplier (+5 for
  // the 16 inner products, +7 for a pair's two real outputs); the complex
  // passes of a pair convolution count load 4
  const int id = 2000 + 1000 * (CONTIG ? 1 : 0) + 100 * (LD == LD_C && a.kfull ? 4 : LD) + 10 * DIR + MUL +
                 (ST == ST_CDOT ? 5 : (ST == ST_RE2 ? 7 : 0));
  hipEvent_t e0 = nullptr, e1 = nullptr;
  const double bpi = pass_bytes<LD, ST>(a);
  // the kernel spectrum is read once per launch, whatever the instances
  const double bpl = (MUL == MUL_KHAT2 || MUL == MUL_KHAT3)
                         ? 8.0 * (MUL == MUL_KHAT2 ? (double)a.H * a.H : (double)a.H * a.H * a.H)
                         : 0.0;
  if (a.act.fl) {
    a.ucount = prof_counter(id);
    a.ubytes = (unsigned long long)bpi;
    a.lbytes = (unsigned long long)bpl;
  }
  const int nact = __builtin_popcountll(a.act.m & lowmask(a.nb));
  const bool ev = prof_events(id, &e0, &e1, a.act.fl ? 0.0 : bpi * nact + (nact ? bpl : 0.0));
  const dim3 grid(blocks), block(a.tpl * a.lpw);
  // FP32 phases run the plain transforms and the preconditioner passes only
  constexpr bool kF32 = (LD == LD_C) && ST != ST_RE2 && (MUL == MUL_NONE || MUL == MUL_PRE2 || MUL == MUL_PRE3);
  if (a.f32) {
    if constexpr (kF32) {
      if (ev) hipExtLaunchKernelGGL((k_pass<CONTIG, LD, DIR, MUL, ST, float2>), grid, block, shm, s, e0, e1, 0, a);
      else hipLaunchKernelGGL((k_pass<CONTIG, LD, DIR, MUL, ST, float2>), grid, block, shm, s, a);
    } else {
      abort();
    }
  } else {
    if (a.fp.odd) {
      if (ev) hipExtLaunchKernelGGL((k_pass<CONTIG, LD, DIR, MUL, ST, double2, true>), grid, block, shm, s, e0, e1, 0, a);
      else hipLaunchKernelGGL((k_pass<CONTIG, LD, DIR, MUL, ST, double2, true>), grid, block, shm, s, a);
    } else {
      with_lc<true>(a.fp, [&](auto lc) {
        constexpr int LC = decltype(lc)::value;
        if (ev) hipExtLaunchKernelGGL((k_pass<CONTIG, LD, DIR, MUL, ST, double2, false, LC>), grid, block, shm, s, e0, e1, 0, a);
        else hipLaunchKernelGGL((k_pass<CONTIG, LD, DIR, MUL, ST, double2, false, LC>), grid, block, shm, s, a);
      });
    }
  }
  return (a.nl + a.lpw - 1) / a.lpw;
}

// launch_pass: the cases are compiled in two translation units (BEC_TU 6:
// plain transforms and the preconditioner; 7: smoothing and convolution); each
// returns -1 for a case it does not hold.
#define BEC_P(C, LDV, D, MU, S)                                                 \
  if (contig == (C) && ld == (LDV) && dir == (D) && mul == (MU) && st == (S)) { \
    return go_pass<C, LDV, D, MU, S>(a, s);                                     \
  }
int launch_pass_plain(int contig, int ld, int dir, int mul, int st, const PassArgs& a, hipStream_t s);
int launch_pass_conv(int contig, int ld, int dir, int mul, int st, const PassArgs& a, hipStream_t s);
#if BEC_TU_IS(6)
int launch_pass_plain(int contig, int ld, int dir, int mul, int st, const PassArgs& a, hipStream_t s) {
  // plain transforms (kernel spectra, preconditioner)
  BEC_P(1, LD_C, DIR_FWD, MUL_NONE, ST_C)
  BEC_P(0, LD_C, DIR_FWD, MUL_NONE, ST_C)
  BEC_P(1, LD_C, DIR_INV, MUL_NONE, ST_C)
  BEC_P(1, LD_C, DIR_INV, MUL_NONE, ST_CDOT)
  BEC_P(0, LD_C, DIR_INV, MUL_NONE, ST_C)
  // preconditioner middle pass
  BEC_P(0, LD_C, DIR_FMI, MUL_PRE2, ST_C)
  BEC_P(0, LD_C, DIR_FMI, MUL_PRE3, ST_C)
  return -1;
}
#endif
#if BEC_TU_IS(7)
int launch_pass_conv(int contig, int ld, int dir, int mul, int st, const PassArgs& a, hipStream_t s) {
  // Gaussian smoothing (vortex diagnostic)
  BEC_P(1, LD_C, DIR_FMI, MUL_GAUSS, ST_C)
  BEC_P(0, LD_C, DIR_FMI, MUL_GAUSS, ST_C)
  // convolution
  BEC_P(1, LD_R, DIR_FWD, MUL_NONE, ST_C)
  BEC_P(0, LD_C, DIR_FMI, MUL_KHAT2, ST_C)
  BEC_P(0, LD_C, DIR_FMI, MUL_KHAT3, ST_C)
  BEC_P(1, LD_HERM, DIR_INV, MUL_NONE, ST_RE)
  // pair convolution (two real fields at once)
  BEC_P(1, LD_R2, DIR_FWD, MUL_NONE, ST_C)
  BEC_P(1, LD_C, DIR_INV, MUL_NONE, ST_RE2)
  return -1;
}
#endif
#undef BEC_P
#if BEC_TU_IS(0)
int launch_pass(int contig, int ld, int dir, int mul, int st, const PassArgs& a, hipStream_t s) {
  int r = launch_pass_plain(contig, ld, dir, mul, st, a, s);
  if (r < 0) r = launch_pass_conv(contig, ld, dir, mul, st, a, s);
  if (r < 0) abort();
  return r;
}
#endif

template <int AX, int KIND>
static int go_kin(const KinArgs& a0, hipStream_t s) {
  KinArgs a = a0;
  const int total = a.nb * a.nl;
  a.lpw = shifted_lpw(pick_lpw(a.tpl, a.nl, total, !Ax<AX>::contig), a.tpl, a.nl,
                      a.lsh[Ax<AX>::contig ? 0 : 1]);
  const int blocks = xcd_grid(total, a.lpw);
  const size_t csz = a.f32 ? sizeof(float2) : sizeof(double2);
  const size_t shm = ((size_t)a.lpw * lpitch<Ax<AX>::contig>(a.N) + a.N) * csz;
  const int id = 100 + 10 * AX + KIND;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  const double bpi = kin_bytes<KIND>(a);
  if (a.act.fl) {
    a.ucount = prof_counter(id);
    a.ubytes = (unsigned long long)bpi;
  }
  const bool ev = prof_events(id, &e0, &e1, a.act.fl ? 0.0 : bpi * __builtin_popcountll(a.act.m & lowmask(a.nb)));
  const dim3 grid(blocks), block(a.tpl * a.lpw);
  auto go = [&](auto pc) {
    constexpr bool PC = decltype(pc)::value;
    if (a.f32) {
      if (ev) hipExtLaunchKernelGGL((k_kin<AX, KIND, float2, false, 0, PC>), grid, block, shm, s, e0, e1, 0, a);
      else hipLaunchKernelGGL((k_kin<AX, KIND, float2, false, 0, PC>), grid, block, shm, s, a);
    } else if (a.fp.odd) {
      if (ev) hipExtLaunchKernelGGL((k_kin<AX, KIND, double2, true, 0, PC>), grid, block, shm, s, e0, e1, 0, a);
      else hipLaunchKernelGGL((k_kin<AX, KIND, double2, true, 0, PC>), grid, block, shm, s, a);
    } else {
      with_lc<false>(a.fp, [&](auto lc) {
        constexpr int LC = decltype(lc)::value;
        if (ev) hipExtLaunchKernelGGL((k_kin<AX, KIND, double2, false, LC, PC>), grid, block, shm, s, e0, e1, 0, a);
        else hipLaunchKernelGGL((k_kin<AX, KIND, double2, false, LC, PC>), grid, block, shm, s, a);
      });
    }
  };
  if constexpr (KIND == KH_LAST) {
    if (a.kq) go(std::true_type{});
    else go(std::false_type{});
  } else {
    go(std::false_type{});
  }
  return (a.nl + a.lpw - 1) / a.lpw;
}

// Line-coefficient pass of the polynomial search (first axis of a grid only).
template <int AX>
static int go_lcoef(const KinArgs& a0, hipStream_t s) {
  if constexpr (AX != A2Y && AX != A3Z) {
    abort();
    return 0;
  } else {
    KinArgs a = a0;
    const int total = a.nb * a.nl;
    a.lpw = shifted_lpw(pick_lpw(a.tpl, a.nl, total, !Ax<AX>::contig), a.tpl, a.nl, a.lsh[0]);
    const int blocks = xcd_grid(total, a.lpw);
    const size_t csz = a.f32 ? sizeof(float2) : sizeof(double2);
    const size_t shm = ((size_t)a.lpw * lpitch<Ax<AX>::contig>(a.N) + a.N) * csz;
    const int id = 100 + 10 * AX + KL_FIRST;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    // psi and d read, z read and d written (direction steps; g is read too
    // only on a step that falls back to the gradient direction), phi read,
    // q and r written (long range)
    const double n = (double)a.nl * a.N, c = (a.f32 ? 8.0 : 16.0) * n, r = 8.0 * n;
    const double bpi = 2.0 * c + (a.d_out ? 2.0 * c : 0.0) + (a.phi ? r : 0.0) + (a.rho_out ? 2.0 * r : 0.0);
    if (a.act.fl) {
      a.ucount = prof_counter(id);
      a.ubytes = (unsigned long long)bpi;
    }
    const bool ev = prof_events(id, &e0, &e1, a.act.fl ? 0.0 : bpi * __builtin_popcountll(a.act.m & lowmask(a.nb)));
    const dim3 grid(blocks), block(a.tpl * a.lpw);
    if (a.f32) {
      if (ev) hipExtLaunchKernelGGL((k_lcoef<AX, float2>), grid, block, shm, s, e0, e1, 0, a);
      else hipLaunchKernelGGL((k_lcoef<AX, float2>), grid, block, shm, s, a);
    } else if (a.fp.odd) {
      if (ev) hipExtLaunchKernelGGL((k_lcoef<AX, double2, true>), grid, block, shm, s, e0, e1, 0, a);
      else hipLaunchKernelGGL((k_lcoef<AX, double2, true>), grid, block, shm, s, a);
    } else {
      with_lc<false>(a.fp, [&](auto lc) {
        constexpr int LC = decltype(lc)::value;
        if (ev) hipExtLaunchKernelGGL((k_lcoef<AX, double2, false, LC>), grid, block, shm, s, e0, e1, 0, a);
        else hipLaunchKernelGGL((k_lcoef<AX, double2, false, LC>), grid, block, shm, s, a);
      });
    }
    return (a.nl + a.lpw - 1) / a.lpw;
  }
}

// launch_kin: one translation unit per axis mode (BEC_TU 1..5)
#define BEC_K(AXV)                                                  \
  switch (kind) {                                                   \
    case KE_FIRST: return go_kin<AXV, KE_FIRST>(a, s);              \
    case KE_REST: return go_kin<AXV, KE_REST>(a, s);                \
    case KH_FIRST: return go_kin<AXV, KH_FIRST>(a, s);              \
    case KH_MID: return go_kin<AXV, KH_MID>(a, s);                  \
    case KL_FIRST: return go_lcoef<AXV>(a, s);                      \
    default: return go_kin<AXV, KH_LAST>(a, s);                     \
  }
int launch_kin_a2y(int kind, const KinArgs& a, hipStream_t s);
int launch_kin_a2x(int kind, const KinArgs& a, hipStream_t s);
int launch_kin_a3z(int kind, const KinArgs& a, hipStream_t s);
int launch_kin_a3y(int kind, const KinArgs& a, hipStream_t s);
int launch_kin_a3x(int kind, const KinArgs& a, hipStream_t s);
#if BEC_TU_IS(1)
int launch_kin_a2y(int kind, const KinArgs& a, hipStream_t s) { BEC_K(A2Y) }
#endif
#if BEC_TU_IS(2)
int launch_kin_a2x(int kind, const KinArgs& a, hipStream_t s) { BEC_K(A2X) }
#endif
#if BEC_TU_IS(3)
int launch_kin_a3z(int kind, const KinArgs& a, hipStream_t s) { BEC_K(A3Z) }
#endif
#if BEC_TU_IS(4)
int launch_kin_a3y(int kind, const KinArgs& a, hipStream_t s) { BEC_K(A3Y) }
#endif
#if BEC_TU_IS(5)
int launch_kin_a3x(int kind, const KinArgs& a, hipStream_t s) { BEC_K(A3X) }
#endif
#undef BEC_K
#if BEC_TU_IS(0)
int launch_kin(int ax, int kind, const KinArgs& a, hipStream_t s) {
  switch (ax) {
    case A2Y: return launch_kin_a2y(kind, a, s);
    case A2X: return launch_kin_a2x(kind, a, s);
    case A3Z: return launch_kin_a3z(kind, a, s);
    case A3Y: return launch_kin_a3y(kind, a, s);
    case A3X: return launch_kin_a3x(kind, a, s);
  }
  abort();
  return 0;
}

int launch_norm2(const double2* psi, long long n, int nb, unsigned long long act, double* hpart, hipStream_t s) {
  const int g = elem_blocks(n);
  hipLaunchKernelGGL(k_norm2, dim3(g, nb), dim3(BEC_WG), 0, s, psi, n, act, hpart);
  return g;
}
void launch_scale(const double2* in, double2* out, const double* divs, long long n, int nb, hipStream_t s,
                  unsigned long long act) {
  DblArr d;
  for (int b = 0; b < BEC_BMAX; ++b) d.v[b] = b < nb ? divs[b] : 1.0;
  hipLaunchKernelGGL(k_scale, dim3(elem_blocks(n), nb), dim3(BEC_WG), 0, s, in, out, d, n, act);
}
void launch_rho(const double2* psi, double* rho, long long n, int nb, hipStream_t s) {
  hipLaunchKernelGGL(k_rho, dim3(elem_blocks(n), nb), dim3(BEC_WG), 0, s, psi, rho, n);
}
void launch_post(const double2* psi, double2* best, double2* d, const double2* z, const double* tab,
                 const unsigned* fl, long long n, int nb, int f32, hipStream_t s) {
  prof_pre(SITE_POST, s);
  if (f32)
    hipLaunchKernelGGL(k_post<float2>, dim3(elem_blocks(n), nb), dim3(BEC_WG), 0, s, (const float2*)psi,
                       (float2*)best, (float2*)d, (const float2*)z, tab, fl, n);
  else
    hipLaunchKernelGGL(k_post<double2>, dim3(elem_blocks(n), nb), dim3(BEC_WG), 0, s, psi, best, d, z, tab, fl, n);
  prof_post(SITE_POST, s);
}
// ---------------------------------------------------------------------------
// Diagnostics of a 2D state (vortex.py:10-84)
// ---------------------------------------------------------------------------
// Lowest-Landau-level projections for m = m0 .. m0+mcnt-1 (mcnt <= 5): per
// block the sums |q_m|^2 and conj(q_m) psi (re, im), q_m = r^m exp(-r^2/2)
// exp(i m theta) / sqrt(pi m!), theta = sgn atan2(y, x); log |q_m| advances
// by log r - log(m)/2 from one m to the next, the phase by e^{i theta}.
// per instance b (blockIdx.y): sgn = prm[2b], orbitals m0 .. m0 + mcnt - 1
// with mcnt = prm[2b + 1] (at most 5; 0 gives zero rows)
__global__ __launch_bounds__(BEC_WG) void k_lll(const double2* __restrict__ psi0, int N, double dx,
                                                const double* __restrict__ prm, int m0, double* rows) {
  double v[15];
#pragma unroll
  for (int q = 0; q < 15; ++q) v[q] = 0.0;
  const long long n = (long long)N * N;
  const int b = blockIdx.y;
  const double sgn = prm[2 * b];
  const int mcnt = (int)prm[2 * b + 1];
  const double2* psi = psi0 + (long long)b * n;
  const double lpi = 1.1447298858494002;           // log(pi)
  if (mcnt > 0) ELEM_LOOP(n) {
    const int ix = (int)(i / N), iy = (int)(i % N);
    const double x = xcoord(ix, N, dx), y = sgn * xcoord(iy, N, dx);
    const double r2 = x * x + y * y, r = sqrt(r2);
    const double lr = log(fmax(r, 1e-300));
    double lm = (double)m0 * lr - 0.5 * r2 - 0.5 * (lpi + lgamma((double)m0 + 1.0));
    const double th = atan2(y, x);
    double2 e = make_double2(cos((double)m0 * th), sin((double)m0 * th));
    const double2 u = make_double2(cos(th), sin(th));
    const double2 p = psi[i];
#pragma unroll
    for (int q = 0; q < 5; ++q) {
      if (q < mcnt) {
        const double mag = exp(lm);
        const double2 qm = make_double2(mag * e.x, mag * e.y);
        v[3 * q] += mag * mag;
        v[3 * q + 1] += qm.x * p.x + qm.y * p.y;     // Re conj(q) psi
        v[3 * q + 2] += qm.x * p.y - qm.y * p.x;     // Im conj(q) psi
        const int m = m0 + q + 1;
        lm += lr - 0.5 * log((double)m);
        e = cmul(e, u);
      }
    }
  }
  block_partials<15, 0>(v, nullptr, rows, nullptr, (long long)blockIdx.y * gridDim.x + blockIdx.x);
}
int launch_lll(const double2* psi, int N, double dx, const double* prm, int m0, int nb, double* rows,
               hipStream_t s) {
  const int g = elem_blocks((long long)N * N);
  hipLaunchKernelGGL(k_lll, dim3(g, nb), dim3(BEC_WG), 0, s, psi, N, dx, prm, m0, rows);
  return g;
}
// wrap an angle difference into [-pi, pi): (d + pi) mod 2 pi - pi, the modulus
// taking the sign of the divisor
__device__ __forceinline__ double wrap_pi(double d) {
  const double tp = 6.283185307179586;
  double r = fmod(d + 3.141592653589793, tp);
  if (r < 0.0) r += tp;
  return r - 3.141592653589793;
}
// plaquettes (ix, iy), ix, iy < N-1: the winding of the phase around corners
// (ix,iy) -> (ix+1,iy) -> (ix+1,iy+1) -> (ix,iy+1) -> back, counted when the
// corner-averaged smoothed density exceeds thr and corner (ix,iy) lies within
// R of (xcm, ycm); rows of (number of windings >= 1, number <= -1) per block
// per instance b (blockIdx.y): xcm, ycm, R, thr = prm[4b .. 4b + 3]
__global__ __launch_bounds__(BEC_WG) void k_vortex(const double2* __restrict__ psi0,
                                                   const double* __restrict__ nbar0, int N, double dx,
                                                   const double* __restrict__ prm, double* rows) {
  double v[2] = {0.0, 0.0};
  const long long np = (long long)(N - 1) * (N - 1);
  const int b = blockIdx.y;
  const double2* psi = psi0 + (long long)b * N * N;
  const double* nbar = nbar0 + (long long)b * N * N;
  const double xcm = prm[4 * b], ycm = prm[4 * b + 1], R = prm[4 * b + 2], thr = prm[4 * b + 3];
  ELEM_LOOP(np) {
    const int ix = (int)(i / (N - 1)), iy = (int)(i % (N - 1));
    const long long a0 = (long long)ix * N + iy, a1 = a0 + N, a2 = a1 + 1, a3 = a0 + 1;
    const double cell = 0.25 * (nbar[a0] + nbar[a1] + nbar[a2] + nbar[a3]);
    const double xx = xcoord(ix, N, dx) - xcm, yy = xcoord(iy, N, dx) - ycm;
    if (!(cell > thr) || !(sqrt(xx * xx + yy * yy) < R)) continue;
    const double t0 = atan2(psi[a0].y, psi[a0].x), t1 = atan2(psi[a1].y, psi[a1].x);
    const double t2 = atan2(psi[a2].y, psi[a2].x), t3 = atan2(psi[a3].y, psi[a3].x);
    const double tot = wrap_pi(t1 - t0) + wrap_pi(t2 - t1) + wrap_pi(t3 - t2) + wrap_pi(t0 - t3);
    const double w = rint(tot / 6.283185307179586);   // round half to even
    if (w >= 1.0) v[0] += 1.0;
    if (w <= -1.0) v[1] += 1.0;
  }
  block_partials<2, 0>(v, nullptr, rows, nullptr, (long long)blockIdx.y * gridDim.x + blockIdx.x);
}
int launch_vortex(const double2* psi, const double* nbar, int N, double dx, const double* prm, int nb, double* rows,
                  hipStream_t s) {
  const int g = elem_blocks((long long)(N - 1) * (N - 1));
  hipLaunchKernelGGL(k_vortex, dim3(g, nb), dim3(BEC_WG), 0, s, psi, nbar, N, dx, prm, rows);
  return g;
}
// nbar = max(Re f, 0) (the smoothed density) and its per-block maxima
__global__ __launch_bounds__(BEC_WG) void k_nbar(const double2* __restrict__ f, long long n, double* nbar,
                                                 double* rows) {
  double m[1] = {0.0};
  const long long off = (long long)blockIdx.y * n;
  ELEM_LOOP(n) {
    const double v = fmax(f[off + i].x, 0.0);
    nbar[off + i] = v;
    m[0] = fmax(m[0], v);
  }
  block_partials<0, 1>(nullptr, m, rows, nullptr, (long long)blockIdx.y * gridDim.x + blockIdx.x);
}
int launch_nbar(const double2* f, long long n, int nb, double* nbar, double* rows, hipStream_t s) {
  const int g = elem_blocks(n);
  hipLaunchKernelGGL(k_nbar, dim3(g, nb), dim3(BEC_WG), 0, s, f, n, nbar, rows);
  return g;
}
// |psi|^2 as a complex field (its smoothing input)
__global__ __launch_bounds__(BEC_WG) void k_abs2c(const double2* __restrict__ psi, long long n, double2* out) {
  ELEM_LOOP(n) out[i] = make_double2(cabs2(psi[i]), 0.0);
}
void launch_abs2c(const double2* psi, long long n, double2* out, hipStream_t s) {
  hipLaunchKernelGGL(k_abs2c, dim3(elem_blocks(n)), dim3(BEC_WG), 0, s, psi, n, out);
}

// ---------------------------------------------------------------------------
// Resampling between grids (grid.py:140-161): multilinear interpolation of the
// old grid's nodes x0 + j dx0 (j < n0) at the new grid's nodes, nodes outside
// the old grid's span taken as zero (torch grid_sample, align_corners=True,
// zero padding, whose arithmetic and summation order this follows: the
// normalised coordinate g = 2 (c - x0) / span - 1, the index ((g + 1) / 2)
// (n0 - 1), corner weights (hi - t) and (t - lo) multiplied last axis first).
// Per block: the sum of |out|^2.
// ---------------------------------------------------------------------------
struct Lin {
  int lo;
  double w0, w1;    // weights of lo and lo + 1
};
__device__ __forceinline__ Lin lin_axis(int i, int n1, double dx1, double x0, double span, int n0) {
  const double c = (double)(i - n1 / 2) * dx1;
  const double g = 2.0 * (c - x0) / span - 1.0;
  const double t = ((g + 1.0) / 2.0) * (double)(n0 - 1);
  const double lo = floor(t);
  Lin l;
  l.lo = (int)lo;
  l.w0 = (lo + 1.0) - t;
  l.w1 = t - lo;
  return l;
}
__global__ __launch_bounds__(BEC_WG) void k_resample(const double2* __restrict__ in, int ndim, int n0, double x0,
                                                     double span, int n1, double dx1, double2* out, double* rows) {
  double v[1] = {0.0};
  const long long n = ndim == 2 ? (long long)n1 * n1 : (long long)n1 * n1 * n1;
  ELEM_LOOP(n) {
    double2 acc = make_double2(0.0, 0.0);
    if (ndim == 2) {
      const int ia = (int)(i / n1), ib = (int)(i % n1);
      const Lin A = lin_axis(ia, n1, dx1, x0, span, n0);    // axis 0 (grid_sample H)
      const Lin Bw = lin_axis(ib, n1, dx1, x0, span, n0);   // axis 1 (grid_sample W)
#pragma unroll
      for (int q = 0; q < 4; ++q) {                          // nw, ne, sw, se
        const int da = q >> 1, db = q & 1;
        const int ja = A.lo + da, jb = Bw.lo + db;
        if (ja < 0 || ja >= n0 || jb < 0 || jb >= n0) continue;
        const double w = (db ? Bw.w1 : Bw.w0) * (da ? A.w1 : A.w0);
        const double2 s = in[(long long)ja * n0 + jb];
        acc.x += s.x * w;
        acc.y += s.y * w;
      }
    } else {
      const long long nn = (long long)n1 * n1;
      const int ia = (int)(i / nn), ib = (int)((i / n1) % n1), ic = (int)(i % n1);
      const Lin A = lin_axis(ia, n1, dx1, x0, span, n0);    // axis 0 (D)
      const Lin Bh = lin_axis(ib, n1, dx1, x0, span, n0);   // axis 1 (H)
      const Lin Cw = lin_axis(ic, n1, dx1, x0, span, n0);   // axis 2 (W)
#pragma unroll
      for (int q = 0; q < 8; ++q) {                          // tnw .. bse
        const int da = q >> 2, db = (q >> 1) & 1, dc = q & 1;
        const int ja = A.lo + da, jb = Bh.lo + db, jc = Cw.lo + dc;
        if (ja < 0 || ja >= n0 || jb < 0 || jb >= n0 || jc < 0 || jc >= n0) continue;
        const double w = (dc ? Cw.w1 : Cw.w0) * (db ? Bh.w1 : Bh.w0) * (da ? A.w1 : A.w0);
        const double2 s = in[((long long)ja * n0 + jb) * n0 + jc];
        acc.x += s.x * w;
        acc.y += s.y * w;
      }
    }
    out[i] = acc;
    v[0] += cabs2(acc);
  }
  block_partials<1, 0>(v, nullptr, rows, nullptr, blockIdx.x);
}
int launch_resample(const double2* in, int ndim, int n0, double x0, double span, int n1, double dx1, double2* out,
                    double* rows, hipStream_t s) {
  const long long n = ndim == 2 ? (long long)n1 * n1 : (long long)n1 * n1 * n1;
  const int g = elem_blocks(n);
  hipLaunchKernelGGL(k_resample, dim3(g), dim3(BEC_WG), 0, s, in, ndim, n0, x0, span, n1, dx1, out, rows);
  return g;
}

__global__ __launch_bounds__(BEC_WG) void k_narrow(const double2* __restrict__ in, float2* out, long long n) {
  ELEM_LOOP(n) out[i] = from_d2<float2>(in[i]);
}
__global__ __launch_bounds__(BEC_WG) void k_widen(const float2* __restrict__ in, double2* out, long long n) {
  ELEM_LOOP(n) out[i] = to_d2(in[i]);
}
void launch_narrow(const double2* in, float2* out, long long n, hipStream_t s) {
  hipLaunchKernelGGL(k_narrow, dim3(elem_blocks(n)), dim3(BEC_WG), 0, s, in, out, n);
}
void launch_widen(const float2* in, double2* out, long long n, hipStream_t s) {
  hipLaunchKernelGGL(k_widen, dim3(elem_blocks(n)), dim3(BEC_WG), 0, s, in, out, n);
}
// phi <- (phi + CQ kq + CR kr) CS for the instances flagged FL_PHIC: the
// potential of the accepted state of a polynomial line search, from the
// current potential and the convolutions of q and r (coefficients in the table)
__global__ __launch_bounds__(BEC_WG) void k_phi_comb(double* phi, const double* __restrict__ kq,
                                                     const double* __restrict__ kr, const double* __restrict__ tab,
                                                     const unsigned* __restrict__ fl, long long n) {
  const int b = blockIdx.y;
  if (!(fl[b] & FL_PHIC)) return;
  const double* t = tab + (long long)b * DS_N;
  const double cq = t[DS_CQ], cr = t[DS_CR], cs = t[DS_CS];
  const long long off = (long long)b * n;
  ELEM_LOOP(n) phi[off + i] = (phi[off + i] + cq * kq[off + i] + cr * kr[off + i]) * cs;
}
void launch_phi_comb(double* phi, const double* kq, const double* kr, const double* tab, const unsigned* fl,
                     long long n, int nb, hipStream_t s) {
  prof_pre(SITE_PHI_COMB, s);
  hipLaunchKernelGGL(k_phi_comb, dim3(elem_blocks(n), nb), dim3(BEC_WG), 0, s, phi, kq, kr, tab, fl, n);
  prof_post(SITE_PHI_COMB, s);
}
void launch_copy(const double2* src, double2* dst, long long n, int nb, Act mask, hipStream_t s) {
  hipLaunchKernelGGL(k_copy, dim3(elem_blocks(n), nb), dim3(BEC_WG), 0, s, src, dst, n, mask);
}
static double act_units(const Act& a, int nb) { return a.fl ? nb : __builtin_popcountll(a.m & lowmask(nb)); }
void launch_combine(double2* d, const double2* z, const double2* g, const double2* psi, const CombArgs& c,
                    const double* tab, long long n, int nb, Act act, hipStream_t s) {
  prof_pre(SITE_COMBINE, s, 48.0 * (double)n * act_units(act, nb));
  hipLaunchKernelGGL(k_combine, dim3(elem_blocks(n), nb), dim3(BEC_WG), 0, s, d, z, g, psi, c, tab, n, act);
  prof_post(SITE_COMBINE, s);
}
int reduce_blocks(long long n) {
  long long g = (n + BEC_WG - 1) / BEC_WG;
  if (g > 256) g = 256;
  if (g < 1) g = 1;
  return (int)g;
}
int launch_dots(const double2* z, const double2* psi, const double2* d, const double2* g, const double2* y,
                const double2* sv, long long n, int nb, Act act, double* partials, double* hpart, hipStream_t s) {
  const int gr = reduce_blocks(n);
  prof_pre(SITE_DOTS, s, 96.0 * (double)n * act_units(act, nb));
  hipLaunchKernelGGL(k_dots, dim3(gr, nb), dim3(BEC_WG), 0, s, z, psi, d, g, y, sv, n, act, partials, hpart);
  prof_post(SITE_DOTS, s);
  return gr;
}
int launch_res2(const double2* h, const double2* psi, double mu, double2* g_out, long long n, double* hpart,
                hipStream_t s) {
  const int gr = elem_blocks(n);
  hipLaunchKernelGGL(k_res2, dim3(gr), dim3(BEC_WG), 0, s, h, psi, mu, g_out, n, hpart);
  return gr;
}
int launch_moments(const double2* psi, int ndim, int N, double dx, double* partials, double* hpart,
                   hipStream_t s, int nb) {
  const long long n = (ndim == 2) ? (long long)N * N : (long long)N * N * N;
  const int gr = elem_blocks(n);
  hipLaunchKernelGGL(k_moments, dim3(gr, nb), dim3(BEC_WG), 0, s, psi, ndim, N, dx, partials, hpart);
  return gr;
}
// Workgroups per instance of the shift's element passes: one per 4 BEC_WG
// (local pass) or BEC_WG (gather) elements, capped so that a batch launches
// about 2048 (local) or 4096 (gather) in all -- each workgroup pays a fixed
// share (the histogram merge, the rank search over the bins) -- but never
// fewer than 8 / 16 per instance.
static int sigma_blocks(long long n, int per, int nb, int cap, int total, int floor_) {
  long long g = (n + per - 1) / per;
  const long long lim = std::max<long long>(floor_, total / std::max(nb, 1));
  g = std::min<long long>(g, std::min<long long>(cap, lim));
  return (int)std::max<long long>(g, 1);
}
void launch_sigma(const Plan& p, int nb, Act act, const double2* psi, const double* phi,
                  const double* mrows, int nrows, int mkp, int mq, double floor_, hipStream_t s, int f32) {
  const int g1 = sigma_blocks(p.n, 4 * BEC_WG, nb, SIGMA_MAX_BLOCKS, 2048, 8);
  const int g = sigma_blocks(p.n, BEC_WG, nb, 1024, 4096, 16);
  unsigned long long* part = (unsigned long long*)p.sig_part;
  prof_pre(SITE_SIGMA_LOCAL, s);
  if (f32)
    hipLaunchKernelGGL(k_sigma_local<float2>, dim3((int)g1, nb), dim3(BEC_WG), 0, s, (const float2*)psi, phi, mrows,
                       nrows, mkp, mq, p.iphys, p.ndim, p.N, p.dx, p.tc, p.s_mode, p.s_half, p.n, act, p.loc, part,
                       p.stats, p.hist);
  else
    hipLaunchKernelGGL(k_sigma_local<double2>, dim3((int)g1, nb), dim3(BEC_WG), 0, s, psi, phi, mrows, nrows, mkp,
                       mq, p.iphys, p.ndim, p.N, p.dx, p.tc, p.s_mode, p.s_half, p.n, act, p.loc, part, p.stats,
                       p.hist);
  prof_post(SITE_SIGMA_LOCAL, s);
  prof_pre(SITE_SIGMA_GATHER, s);
  hipLaunchKernelGGL(k_sigma_gather, dim3(g, nb), dim3(BEC_WG), 0, s, p.loc, p.n, (const unsigned long long*)part,
                     (int)g1, p.stats, p.hist, p.cand, act);
  prof_post(SITE_SIGMA_GATHER, s);
  prof_pre(SITE_SIGMA_SELECT, s);
  hipLaunchKernelGGL(k_sigma_select, dim3(nb), dim3(BEC_WG), 0, s, p.stats, p.hist, p.cand, p.n, floor_, p.sigma,
                     p.hred_dev, act);
  prof_post(SITE_SIGMA_SELECT, s);
}
void launch_init_sigma(const Plan& p, int nb, hipStream_t s) {
  hipLaunchKernelGGL(k_sigma_reset, dim3(nb), dim3(BEC_WG), 0, s, p.stats, p.hist, p.sigma);
}
void launch_kernel_real(int ndim, int M, double dx, int kind, double2* out, hipStream_t s) {
  const long long n = (ndim == 2) ? (long long)M * M : (long long)M * M * M;
  hipLaunchKernelGGL(k_kernel_real, dim3(elem_blocks(n)), dim3(BEC_WG), 0, s, ndim, M, dx, kind, out);
}
void launch_extract_khat(int ndim, int M, int H, const double2* kf, double* khat, hipStream_t s) {
  const long long n = (ndim == 2) ? (long long)H * H : (long long)H * H * H;
  hipLaunchKernelGGL(k_extract_khat, dim3(elem_blocks(n)), dim3(BEC_WG), 0, s, ndim, M, H, kf, khat);
}

#endif  // BEC_TU_IS(0)