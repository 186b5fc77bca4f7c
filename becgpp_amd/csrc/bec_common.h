// Shared device/host definitions for the becGPP MI355X ground-state library.
//
// Layout conventions (identical to the reference's torch tensors):
//   * fields are C-contiguous, indexing "ij": element (ix, iy[, iz]) sits at
//     ix*N + iy (2D) or (ix*N + iy)*N + iz (3D); axis 0 = x;
//   * complex fields are interleaved double2 (torch complex128);
//   * x_i = (i - N/2) dx, k_j = 2 pi f_j / (N dx) with f_j the FFT frequency
//     index (Nyquist negative), as becgpp/grid.py:73-75.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define BEC_WG 256          // threads per workgroup of the line passes (maximum)
#define BEC_XCDS 8          // accelerator complex dies of an MI355X (one L2 each)
#define BEC_MAX_STAGES 16
#define BEC_MAX_L 1024      // longest line transform held in LDS

struct FftPlan {
  int L;                      // transform length (2^a 3^b 5^c 7^d 11^e 13^f)
  int nst;                    // number of Stockham stages
  int rad[BEC_MAX_STAGES];    // radix per stage (8, 6, 5, 4, 3, 2; 7, 11, 13 outside fft_reg)
  int ns[BEC_MAX_STAGES];     // Ns of the stage: product of the earlier radices
  int step[BEC_MAX_STAGES];   // twiddle stride L / (Ns R)
  unsigned mag[BEC_MAX_STAGES];   // floor(2^32 / Ns) + 1: j / Ns = umulhi(j, mag) for j < 2^12
  int reg;                    // 1: L = 8 tpl and the first and last radix divide 8 (fft_reg)
  int odd;                    // 1: a radix 7, 11 or 13 (kernels built with ODD)
  int wide;                   // 1: through LDS on about L/4 threads per line (not reg)
  const double2* tw;          // tw[m] = exp(-2 pi i m / L), m in [0, L)
};

// Complex arithmetic on interleaved pairs: double2 (complex128 fields) and
// float2 (the complex64 fields of an FP32 phase).  CT<C>::F is the real type.
template <class C> struct CT;
template <> struct CT<double2> { using F = double; };
template <> struct CT<float2> { using F = float; };
template <class C> __device__ __forceinline__ C cmk(typename CT<C>::F x, typename CT<C>::F y) {
  C c;
  c.x = x;
  c.y = y;
  return c;
}
template <class C> __device__ __forceinline__ C cadd(C a, C b) { return cmk<C>(a.x + b.x, a.y + b.y); }
template <class C> __device__ __forceinline__ C csub(C a, C b) { return cmk<C>(a.x - b.x, a.y - b.y); }
template <class C> __device__ __forceinline__ C cmul(C a, C b) {
  return cmk<C>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <class C> __device__ __forceinline__ C cscale(C a, double s) {
  using F = typename CT<C>::F;
  return cmk<C>(a.x * (F)s, a.y * (F)s);
}
// multiply by +i (S=+1) or -i (S=-1)
template <int S, class C> __device__ __forceinline__ C cmuli(C a) {
  return S > 0 ? cmk<C>(-a.y, a.x) : cmk<C>(a.y, -a.x);
}
// |a|^2 in FP64 whatever the storage
template <class C> __device__ __forceinline__ double cabs2(C a) {
  return (double)a.x * (double)a.x + (double)a.y * (double)a.y;
}
// widen / narrow one element
__device__ __forceinline__ double2 to_d2(double2 a) { return a; }
__device__ __forceinline__ double2 to_d2(float2 a) { return make_double2((double)a.x, (double)a.y); }
template <class C> __device__ __forceinline__ C from_d2(double2 a) {
  return cmk<C>((typename CT<C>::F)a.x, (typename CT<C>::F)a.y);
}

// ----------------------------------------------------------------------------
// radix-R DFT butterflies, X_p = sum_q v_q exp(SIGN 2 pi i p q / R)
// ----------------------------------------------------------------------------
template <int SIGN, class C> __device__ __forceinline__ void bfly2(C* v) {
  C a = v[0], b = v[1];
  v[0] = cadd(a, b);
  v[1] = csub(a, b);
}
template <int SIGN, class C> __device__ __forceinline__ void bfly3(C* v) {
  using F = typename CT<C>::F;
  // w = exp(SIGN 2 pi i / 3) = -1/2 + i SIGN sqrt(3)/2
  const F s = SIGN * 0.86602540378443864676;   // sqrt(3)/2
  C a = v[0], b = v[1], c = v[2];
  C t1 = cadd(b, c);
  C t2 = cmk<C>(a.x - F(0.5) * t1.x, a.y - F(0.5) * t1.y);
  C d = csub(b, c);
  C t3 = cmk<C>(-s * d.y, s * d.x);       // i s (b - c)
  v[0] = cadd(a, t1);
  v[1] = cadd(t2, t3);
  v[2] = csub(t2, t3);
}
template <int SIGN, class C> __device__ __forceinline__ void bfly4(C* v) {
  C a = v[0], b = v[1], c = v[2], d = v[3];
  C t0 = cadd(a, c), t1 = csub(a, c);
  C t2 = cadd(b, d), t3 = cmuli<SIGN>(csub(b, d));
  v[0] = cadd(t0, t2);
  v[2] = csub(t0, t2);
  v[1] = cadd(t1, t3);
  v[3] = csub(t1, t3);
}
template <int SIGN, class C> __device__ __forceinline__ void bfly8(C* v) {
  using F = typename CT<C>::F;
  const F h = 0.70710678118654752440;           // sqrt(2)/2
  C e[4] = {v[0], v[2], v[4], v[6]};
  C o[4] = {v[1], v[3], v[5], v[7]};
  bfly4<SIGN>(e);
  bfly4<SIGN>(o);
  // twiddles w8^p, w8 = exp(SIGN i pi/4)
  C o1 = cmk<C>(h * (o[1].x - SIGN * o[1].y), h * (o[1].y + SIGN * o[1].x));
  C o2 = cmuli<SIGN>(o[2]);
  C o3 = cmk<C>(h * (-o[3].x - SIGN * o[3].y), h * (-o[3].y + SIGN * o[3].x));
  v[0] = cadd(e[0], o[0]); v[4] = csub(e[0], o[0]);
  v[1] = cadd(e[1], o1);   v[5] = csub(e[1], o1);
  v[2] = cadd(e[2], o2);   v[6] = csub(e[2], o2);
  v[3] = cadd(e[3], o3);   v[7] = csub(e[3], o3);
}
template <int SIGN, class C> __device__ __forceinline__ void bfly5(C* v) {
  using F = typename CT<C>::F;
  // w = exp(SIGN 2 pi i / 5): cos and sin of 2 pi / 5 and 4 pi / 5
  const F c1 = 0.30901699437494742410, c2 = -0.80901699437494742410;
  const F s1 = SIGN * 0.95105651629515357212, s2 = SIGN * 0.58778525229247312917;
  const C a = v[0];
  const C t1 = cadd(v[1], v[4]), t2 = cadd(v[2], v[3]);
  const C t3 = csub(v[1], v[4]), t4 = csub(v[2], v[3]);
  const C r1 = cmk<C>(a.x + c1 * t1.x + c2 * t2.x, a.y + c1 * t1.y + c2 * t2.y);
  const C r2 = cmk<C>(a.x + c2 * t1.x + c1 * t2.x, a.y + c2 * t1.y + c1 * t2.y);
  // i (s1 t3 + s2 t4) and i (s2 t3 - s1 t4)
  const C u1 = cmk<C>(-(s1 * t3.y + s2 * t4.y), s1 * t3.x + s2 * t4.x);
  const C u2 = cmk<C>(-(s2 * t3.y - s1 * t4.y), s2 * t3.x - s1 * t4.x);
  v[0] = cmk<C>(a.x + t1.x + t2.x, a.y + t1.y + t2.y);
  v[1] = cadd(r1, u1);
  v[4] = csub(r1, u1);
  v[2] = cadd(r2, u2);
  v[3] = csub(r2, u2);
}
// radix 6 as two radix-3 DFTs of the even and odd inputs, A and B:
// X_k = A_{k mod 3} + w^k B_{k mod 3}, w = exp(SIGN i pi/3), w^3 = -1
template <int SIGN, class C> __device__ __forceinline__ void bfly6(C* v) {
  using F = typename CT<C>::F;
  const F h = SIGN * 0.86602540378443864676;    // sqrt(3)/2
  C e[3] = {v[0], v[2], v[4]};
  C o[3] = {v[1], v[3], v[5]};
  bfly3<SIGN>(e);
  bfly3<SIGN>(o);
  // w o1 = (1/2 + i h) o1,  w^2 o2 = (-1/2 + i h) o2
  const C o1 = cmk<C>(F(0.5) * o[1].x - h * o[1].y, F(0.5) * o[1].y + h * o[1].x);
  const C o2 = cmk<C>(-F(0.5) * o[2].x - h * o[2].y, -F(0.5) * o[2].y + h * o[2].x);
  v[0] = cadd(e[0], o[0]); v[3] = csub(e[0], o[0]);
  v[1] = cadd(e[1], o1);   v[4] = csub(e[1], o1);
  v[2] = cadd(e[2], o2);   v[5] = csub(e[2], o2);
}
template <int R, int SIGN, class C> __device__ __forceinline__ void bfly(C* v) {
  if constexpr (R == 2) bfly2<SIGN>(v);
  else if constexpr (R == 3) bfly3<SIGN>(v);
  else if constexpr (R == 4) bfly4<SIGN>(v);
  else if constexpr (R == 5) bfly5<SIGN>(v);
  else if constexpr (R == 6) bfly6<SIGN>(v);
  else bfly8<SIGN>(v);
}

// LDS position of element i of a line.  With PD one slot is skipped after
// every eight elements: on a contiguous axis the tpl threads of a line are
// adjacent lanes, and the first stage's stores (element j R + r by lane j)
// would otherwise put the 16-byte values of neighbouring lanes R * 16 bytes
// apart, all in the same few banks; with the skip they fall in distinct banks.
template <bool PD> __host__ __device__ __forceinline__ int lx(int i) { return PD ? i + (i >> 3) : i; }

// One Stockham autosort stage of radix R on a line of length L held in LDS.
// tpl threads cooperate on the line (lane in [0, tpl)); L/tpl <= 8 is
// guaranteed by the host, so a thread owns at most ceil(8/R) butterflies.
// tw is the twiddle table staged in LDS by the calling kernel.
// Every thread of the workgroup must call it (it contains barriers).
template <int R, int SIGN, bool PD = false, class C>
__device__ __forceinline__ void fft_stage(C* s, int L, int Ns, unsigned mag, int step, int lane, int tpl,
                                          const C* __restrict__ tw, bool active) {
  constexpr int MAXB = (8 + R - 1) / R;
  const int nb = L / R;
  C v[MAXB][R];
  if (active) {
#pragma unroll
    for (int b = 0; b < MAXB; ++b) {
      const int j = lane + b * tpl;
      if (j < nb) {
#pragma unroll
        for (int r = 0; r < R; ++r) v[b][r] = s[lx<PD>(j + r * nb)];
      }
    }
  }
  __syncthreads();
  if (active) {
#pragma unroll
    for (int b = 0; b < MAXB; ++b) {
      const int j = lane + b * tpl;
      if (j < nb) {
        const int k = (Ns == 1) ? 0 : j - (int)__umulhi((unsigned)j, mag) * Ns;   // j mod Ns
        if (Ns > 1) {
#pragma unroll
          for (int r = 1; r < R; ++r) {
            C w = tw[r * k * step];
            if (SIGN > 0) w.y = -w.y;
            v[b][r] = cmul(v[b][r], w);
          }
        }
        bfly<R, SIGN>(v[b]);
        const int idxD = (j - k) * R + k;
#pragma unroll
        for (int r = 0; r < R; ++r) s[lx<PD>(idxD + r * Ns)] = v[b][r];
      }
    }
  }
  __syncthreads();
}

// cos and sin of 2 pi k / R, k = 0 .. (R-1)/2, for the odd prime radices
// (rounded from extended-precision values)
template <int R> struct OddRoot;
template <> struct OddRoot<7> {
  static constexpr double c[4] = {1.0, 0.6234898018587335, -0.2225209339563144, -0.9009688679024191};
  static constexpr double s[4] = {0.0, 0.7818314824680298, 0.9749279121818236, 0.4338837391175581};
};
template <> struct OddRoot<11> {
  static constexpr double c[6] = {1.0, 0.8412535328311812, 0.41541501300188644, -0.14231483827328514,
                                  -0.6548607339452851, -0.9594929736144974};
  static constexpr double s[6] = {0.0, 0.5406408174555976, 0.9096319953545183, 0.9898214418809327,
                                  0.7557495743542583, 0.28173255684142967};
};
template <> struct OddRoot<13> {
  static constexpr double c[7] = {1.0, 0.8854560256532099, 0.5680647467311558, 0.12053668025532305,
                                  -0.3546048870425356, -0.7485107481711011, -0.970941817426052};
  static constexpr double s[7] = {0.0, 0.46472317204376856, 0.8229838658936564, 0.992708874098054,
                                  0.9350162426854148, 0.6631226582407952, 0.23931566428755777};
};

// One Stockham stage of an odd prime radix R (7, 11 or 13).  With a_q =
// u_q + u_{R-q} and b_q = u_q - u_{R-q} (q = 1 .. H, H = (R-1)/2),
//   X_p, X_{R-p} = u_0 + sum_q cos(2 pi pq/R) a_q  +-  SIGN i sum_q sin(2 pi pq/R) b_q,
// and each output pair is stored to LDS as soon as it is formed, so a thread
// holds R values per butterfly.  Only lengths run through fft_lds take these
// radices (a line of L = 8 tpl with a factor >= 7 never fills whole wavefronts).
template <int R, int SIGN, bool PD = false, class C>
__device__ __forceinline__ void fft_stage_odd(C* s, int L, int Ns, unsigned mag, int step, int lane, int tpl,
                                              const C* __restrict__ tw, bool active) {
  using F = typename CT<C>::F;
  constexpr int H = (R - 1) / 2;
  constexpr int MAXB = (8 + R - 1) / R;
  const int nb = L / R;
  C u0[MAXB], a[MAXB][H], b[MAXB][H];
  if (active) {
#pragma unroll
    for (int bb = 0; bb < MAXB; ++bb) {
      const int j = lane + bb * tpl;
      if (j < nb) {
        const int k = (Ns == 1) ? 0 : j - (int)__umulhi((unsigned)j, mag) * Ns;   // j mod Ns
        u0[bb] = s[lx<PD>(j)];
#pragma unroll
        for (int q = 1; q <= H; ++q) {
          C x = s[lx<PD>(j + q * nb)], y = s[lx<PD>(j + (R - q) * nb)];
          if (Ns > 1) {
            C wx = tw[q * k * step], wy = tw[(R - q) * k * step];
            if (SIGN > 0) { wx.y = -wx.y; wy.y = -wy.y; }
            x = cmul(x, wx);
            y = cmul(y, wy);
          }
          a[bb][q - 1] = cadd(x, y);
          b[bb][q - 1] = csub(x, y);
        }
      }
    }
  }
  __syncthreads();
  if (active) {
#pragma unroll
    for (int bb = 0; bb < MAXB; ++bb) {
      const int j = lane + bb * tpl;
      if (j < nb) {
        const int k = (Ns == 1) ? 0 : j - (int)__umulhi((unsigned)j, mag) * Ns;
        const int idxD = (j - k) * R + k;
        C x0 = u0[bb];
#pragma unroll
        for (int q = 0; q < H; ++q) x0 = cadd(x0, a[bb][q]);
        s[lx<PD>(idxD)] = x0;
#pragma unroll
        for (int p = 1; p <= H; ++p) {
          C re = u0[bb], im = cmk<C>(0, 0);
#pragma unroll
          for (int q = 1; q <= H; ++q) {
            const int m = (p * q) % R;
            const int mm = m <= H ? m : R - m;
            const F cs = (F)OddRoot<R>::c[mm];
            const F sn = (F)(m <= H ? OddRoot<R>::s[mm] : -OddRoot<R>::s[mm]);
            re = cmk<C>(re.x + cs * a[bb][q - 1].x, re.y + cs * a[bb][q - 1].y);
            im = cmk<C>(im.x + sn * b[bb][q - 1].x, im.y + sn * b[bb][q - 1].y);
          }
          const C t = cmuli<SIGN>(im);
          s[lx<PD>(idxD + p * Ns)] = cadd(re, t);
          s[lx<PD>(idxD + (R - p) * Ns)] = csub(re, t);
        }
      }
    }
  }
  __syncthreads();
}

// In-place unnormalised DFT of the LDS line s (SIGN=-1 forward, +1 inverse);
// twl = the plan's twiddle table copied to LDS (stage_twiddles).
template <int SIGN, bool ODD = false, bool PD = false, class C>
__device__ __forceinline__ void fft_lds(C* s, const FftPlan& p, const C* twl, int lane, int tpl,
                                        bool active) {
  for (int st = 0; st < p.nst; ++st) {
    const int R = p.rad[st], Ns = p.ns[st], step = p.step[st];
    const unsigned mag = p.mag[st];
    if (R == 8) fft_stage<8, SIGN, PD>(s, p.L, Ns, mag, step, lane, tpl, twl, active);
    else if (R == 4) fft_stage<4, SIGN, PD>(s, p.L, Ns, mag, step, lane, tpl, twl, active);
    else if (R == 3) fft_stage<3, SIGN, PD>(s, p.L, Ns, mag, step, lane, tpl, twl, active);
    else if (R == 5) fft_stage<5, SIGN, PD>(s, p.L, Ns, mag, step, lane, tpl, twl, active);
    else if (R == 6) fft_stage<6, SIGN, PD>(s, p.L, Ns, mag, step, lane, tpl, twl, active);
    else if (!ODD || R == 2) fft_stage<2, SIGN, PD>(s, p.L, Ns, mag, step, lane, tpl, twl, active);
    else if constexpr (ODD) {
      if (R == 7) fft_stage_odd<7, SIGN, PD>(s, p.L, Ns, mag, step, lane, tpl, twl, active);
      else if (R == 11) fft_stage_odd<11, SIGN, PD>(s, p.L, Ns, mag, step, lane, tpl, twl, active);
      else fft_stage_odd<13, SIGN, PD>(s, p.L, Ns, mag, step, lane, tpl, twl, active);
    }
  }
}

// ----------------------------------------------------------------------------
// Register-resident line transform.  A thread holds the 8 elements lane + m tpl
// (m < 8) of its line, L = 8 tpl.  The first stage (Ns = 1) takes its radix-R
// butterflies j = lane + b tpl straight from those registers (inputs
// j + r L/R = lane + (b + r 8/R) tpl); the last stage (Ns = L/R) leaves its
// outputs j + r Ns = lane + (b + r 8/R) tpl in them.  Only the stages between
// go through the LDS line, so a forward transform, a pointwise multiply and an
// inverse transform keep the spectrum in registers across the multiply.
// ----------------------------------------------------------------------------
template <int R, int SIGN, bool PD = false, class C>
__device__ __forceinline__ void reg_first(C (&v)[8], C* s, int lane, int tpl, bool active) {
  constexpr int Q = 8 / R;
  __syncthreads();                     // earlier readers of s are done
  if (active) {
#pragma unroll
    for (int b = 0; b < Q; ++b) {
      C u[R];
#pragma unroll
      for (int r = 0; r < R; ++r) u[r] = v[b + r * Q];
      bfly<R, SIGN>(u);
      const int j = lane + b * tpl;
#pragma unroll
      for (int r = 0; r < R; ++r) s[lx<PD>(j * R + r)] = u[r];
    }
  }
  __syncthreads();
}
template <int R, int SIGN, bool PD = false, class C>
__device__ __forceinline__ void reg_last(C (&v)[8], const C* s, int L, int lane, int tpl,
                                         const C* __restrict__ tw, bool active) {
  constexpr int Q = 8 / R;
  const int ns = L / R;                // twiddle stride L / (Ns R) = 1
  if (!active) return;
#pragma unroll
  for (int b = 0; b < Q; ++b) {
    const int j = lane + b * tpl;
    C u[R];
#pragma unroll
    for (int r = 0; r < R; ++r) u[r] = s[lx<PD>(j + r * ns)];
#pragma unroll
    for (int r = 1; r < R; ++r) {
      C w = tw[r * j];
      if (SIGN > 0) w.y = -w.y;
      u[r] = cmul(u[r], w);
    }
    bfly<R, SIGN>(u);
#pragma unroll
    for (int r = 0; r < R; ++r) v[b + r * Q] = u[r];
  }
}
template <int SIGN, bool PD = false, class C>
__device__ __forceinline__ void fft_reg(C (&v)[8], C* s, const FftPlan& p, const C* twl,
                                        int lane, int tpl, bool active) {
  const int n = p.nst;
  if (n == 1) {                        // L = 8, one thread per line
    if (active) bfly<8, SIGN>(v);
    return;
  }
  const int r0 = p.rad[0], rl = p.rad[n - 1];
  if (r0 == 8) reg_first<8, SIGN, PD>(v, s, lane, tpl, active);
  else if (r0 == 4) reg_first<4, SIGN, PD>(v, s, lane, tpl, active);
  else reg_first<2, SIGN, PD>(v, s, lane, tpl, active);
  for (int st = 1; st < n - 1; ++st) {
    const int R = p.rad[st], Ns = p.ns[st], step = p.step[st];
    const unsigned mag = p.mag[st];
    if (R == 8) fft_stage<8, SIGN, PD>(s, p.L, Ns, mag, step, lane, tpl, twl, active);
    else if (R == 4) fft_stage<4, SIGN, PD>(s, p.L, Ns, mag, step, lane, tpl, twl, active);
    else if (R == 3) fft_stage<3, SIGN, PD>(s, p.L, Ns, mag, step, lane, tpl, twl, active);
    else if (R == 5) fft_stage<5, SIGN, PD>(s, p.L, Ns, mag, step, lane, tpl, twl, active);
    else if (R == 6) fft_stage<6, SIGN, PD>(s, p.L, Ns, mag, step, lane, tpl, twl, active);
    else fft_stage<2, SIGN, PD>(s, p.L, Ns, mag, step, lane, tpl, twl, active);
  }
  if (rl == 8) reg_last<8, SIGN, PD>(v, s, p.L, lane, tpl, twl, active);
  else if (rl == 4) reg_last<4, SIGN, PD>(v, s, p.L, lane, tpl, twl, active);
  else reg_last<2, SIGN, PD>(v, s, p.L, lane, tpl, twl, active);
}
// The register-resident transform of a power-of-two line of compile-time
// length LC in the stage order the plans use (radix 8 first, radix 8 in the
// middle stages, the remaining 8, 4 or 2 last; tpl = LC / 8).  It performs
// the same operations in the same order as fft_reg on such a plan, with the
// strides, Ns and j mod Ns folded at compile time.
template <int LC> struct RegPlan {
  static constexpr int lg() { int a = 0; for (int n = LC; n > 1; n >>= 1) ++a; return a; }
  static constexpr int A = lg();
  static constexpr int TPL = LC / 8;
  static constexpr int RL = (A % 3 == 1) ? 2 : (A % 3 == 2 ? 4 : 8);
  static constexpr int NMID = (A % 3 == 0 ? A / 3 : A / 3 + 1) - 2;   // stages between the first and the last
  static_assert(LC >= 64 && (LC & (LC - 1)) == 0 && NMID <= 2, "register plans of 64..1024 points");
};
template <int R, int SIGN, bool PD, int L, int NS, int TPL, class C>
__device__ __forceinline__ void fft_stage_c(C* s, int lane, const C* __restrict__ tw, bool active) {
  constexpr int MAXB = (8 + R - 1) / R;
  constexpr int nb = L / R;
  constexpr int step = L / (NS * R);
  C v[MAXB][R];
  if (active) {
#pragma unroll
    for (int b = 0; b < MAXB; ++b) {
      const int j = lane + b * TPL;
      if (j < nb) {
#pragma unroll
        for (int r = 0; r < R; ++r) v[b][r] = s[lx<PD>(j + r * nb)];
      }
    }
  }
  __syncthreads();
  if (active) {
#pragma unroll
    for (int b = 0; b < MAXB; ++b) {
      const int j = lane + b * TPL;
      if (j < nb) {
        const int k = (NS == 1) ? 0 : j % NS;
        if constexpr (NS > 1) {
#pragma unroll
          for (int r = 1; r < R; ++r) {
            C w = tw[r * k * step];
            if (SIGN > 0) w.y = -w.y;
            v[b][r] = cmul(v[b][r], w);
          }
        }
        bfly<R, SIGN>(v[b]);
        const int idxD = (j - k) * R + k;
#pragma unroll
        for (int r = 0; r < R; ++r) s[lx<PD>(idxD + r * NS)] = v[b][r];
      }
    }
  }
  __syncthreads();
}
template <int SIGN, bool PD, int LC, class C>
__device__ __forceinline__ void fft_reg_c(C (&v)[8], C* s, const C* twl, int lane, bool active) {
  using P = RegPlan<LC>;
  reg_first<8, SIGN, PD>(v, s, lane, P::TPL, active);
  if constexpr (P::NMID >= 1) fft_stage_c<8, SIGN, PD, LC, 8, P::TPL>(s, lane, twl, active);
  if constexpr (P::NMID >= 2) fft_stage_c<8, SIGN, PD, LC, 64, P::TPL>(s, lane, twl, active);
  reg_last<P::RL, SIGN, PD>(v, s, LC, lane, P::TPL, twl, active);
}
// The transform through LDS of a line of compile-time length LC = 3 2^a
// (192, 384) in the plans' stage order (8, 3 or 6, 8) on the power-of-two
// thread count covering LC/8, as fft_line does for such a plan.
template <int LC> struct LdsPlan;
template <> struct LdsPlan<192> {
  static constexpr int TPL = 32, R0 = 8, R1 = 3, R2 = 8;
};
template <> struct LdsPlan<384> {
  static constexpr int TPL = 64, R0 = 8, R1 = 6, R2 = 8;
};
template <int SIGN, bool PD, int LC, class C>
__device__ __forceinline__ void fft_lds_c(C (&v)[8], C* s, const C* twl, int lane, bool active) {
  using P = LdsPlan<LC>;
  constexpr int T = P::TPL;
  __syncthreads();
  if (active) {
#pragma unroll
    for (int m = 0; m < 8; ++m) {
      const int e = lane + m * T;
      if (m * T < LC && e < LC) s[lx<PD>(e)] = v[m];
    }
  }
  __syncthreads();
  fft_stage_c<P::R0, SIGN, PD, LC, 1, T>(s, lane, twl, active);
  fft_stage_c<P::R1, SIGN, PD, LC, P::R0, T>(s, lane, twl, active);
  fft_stage_c<P::R2, SIGN, PD, LC, P::R0 * P::R1, T>(s, lane, twl, active);
  if (active) {
#pragma unroll
    for (int m = 0; m < 8; ++m) {
      const int e = lane + m * T;
      v[m] = (m * T < LC && e < LC) ? s[lx<PD>(e)] : cmk<C>(0.0, 0.0);
    }
  }
}
// threads per line of the kernels built for length LC
__host__ __device__ constexpr int lc_tpl(int LC) {
  return (LC & (LC - 1)) == 0 ? LC / 8 : (LC == 192 ? 32 : 64);
}

// host side: the compile-time length whose transform equals plan p's, or 0
__host__ __device__ inline int lds_plan_len(const FftPlan& p) {
  if (p.reg || p.wide || p.odd || p.nst != 3 || p.rad[0] != 8 || p.rad[2] != 8) return 0;
  if (p.L == 192 && p.rad[1] == 3) return 192;
  if (p.L == 384 && p.rad[1] == 6) return 384;
  return 0;
}
__host__ __device__ inline int reg_plan_len(const FftPlan& p) {
  if (!p.reg || p.odd || p.L < 64 || p.L > 1024 || (p.L & (p.L - 1)) != 0) return 0;
  int a = 0;
  for (int n = p.L; n > 1; n >>= 1) ++a;
  const int rl = (a % 3 == 1) ? 2 : (a % 3 == 2 ? 4 : 8);
  const int nst = (a % 3 == 0) ? a / 3 : a / 3 + 1;
  if (p.nst != nst || nst > 4) return 0;
  for (int i = 0; i < nst - 1; ++i)
    if (p.rad[i] != 8) return 0;
  return p.rad[nst - 1] == rl ? p.L : 0;
}

// The DFT of a line whose elements lane + m tpl (m < 8, those below L) are in
// v: register-resident when the plan allows it, else through the LDS line s
// (fft_lds).  Every thread of the workgroup calls it; results return in v.
// ODD compiles in the radices 7, 11 and 13 (plans with FftPlan::odd): their
// stages hold more registers, so the kernels of other plans are built without.
template <int SIGN, bool ODD = false, bool PD = false, class C>
__device__ __forceinline__ void fft_line(C (&v)[8], C* s, const FftPlan& p, const C* twl,
                                         int lane, int tpl, bool active) {
  if (p.reg) {
    fft_reg<SIGN, PD>(v, s, p, twl, lane, tpl, active);
    return;
  }
  __syncthreads();
  if (active) {
#pragma unroll
    for (int m = 0; m < 8; ++m) {
      const int e = lane + m * tpl;
      if (e < p.L) s[lx<PD>(e)] = v[m];
    }
  }
  __syncthreads();
  fft_lds<SIGN, ODD, PD>(s, p, twl, lane, tpl, active);
  if (active) {
#pragma unroll
    for (int m = 0; m < 8; ++m) {
      const int e = lane + m * tpl;
      v[m] = (e < p.L) ? s[lx<PD>(e)] : cmk<C>(0.0, 0.0);
    }
  }
}

// fft_line, or fft_reg_c / fft_lds_c when the kernel was built for the plan's length LC
template <int SIGN, bool ODD, bool PD, int LC, class C>
__device__ __forceinline__ void fft_any(C (&v)[8], C* s, const FftPlan& p, const C* twl, int lane, int tpl,
                                        bool active) {
  if constexpr (LC != 0 && (LC & (LC - 1)) == 0) fft_reg_c<SIGN, PD, LC>(v, s, twl, lane, active);
  else if constexpr (LC != 0) fft_lds_c<SIGN, PD, LC>(v, s, twl, lane, active);
  else fft_line<SIGN, ODD, PD>(v, s, p, twl, lane, tpl, active);
}

// Copies the twiddle table (L entries) into LDS; the caller's next barrier
// publishes it.  Issued together with the line loads, so its latency overlaps
// theirs instead of sitting inside every FFT stage.
template <class C>
__device__ __forceinline__ void stage_twiddles(C* twl, const FftPlan& p) {
  for (int i = threadIdx.x; i < p.L; i += blockDim.x) twl[i] = from_d2<C>(p.tw[i]);
}

// ----------------------------------------------------------------------------
// deterministic workgroup reductions (fixed tree order)
// ----------------------------------------------------------------------------
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_down(v, off, 64));
  return v;
}

// Sums K per-thread values over the workgroup and writes them to out[0..K).
// red must hold K*(blockDim.x/64) doubles of LDS.  Called by every thread.
template <int K>
__device__ __forceinline__ void block_sum_store(double (&v)[K], double* red, double* out) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
#pragma unroll
  for (int q = 0; q < K; ++q) {
    double s = wave_sum(v[q]);
    if (lane == 0) red[q * nw + wid] = s;
  }
  __syncthreads();
  if (threadIdx.x < K) {
    double s = 0.0;
    for (int w = 0; w < nw; ++w) s += red[threadIdx.x * nw + w];
    out[threadIdx.x] = s;
  }
  __syncthreads();
}

__device__ __forceinline__ uint64_t order_key(double v) {
  uint64_t u = __double_as_longlong(v);
  return (u >> 63) ? ~u : (u | 0x8000000000000000ull);
}
__device__ __forceinline__ double key_to_double(uint64_t k) {
  uint64_t u = (k >> 63) ? (k & 0x7fffffffffffffffull) : ~k;
  return __longlong_as_double(u);
}
