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
  int L;                      // transform length (2^a 3^b 5^c)
  int nst;                    // number of Stockham stages
  int rad[BEC_MAX_STAGES];    // radix per stage (8, 6, 5, 4, 3 or 2)
  int ns[BEC_MAX_STAGES];     // Ns of the stage: product of the earlier radices
  int step[BEC_MAX_STAGES];   // twiddle stride L / (Ns R)
  unsigned mag[BEC_MAX_STAGES];   // floor(2^32 / Ns) + 1: j / Ns = umulhi(j, mag) for j < 2^12
  int reg;                    // 1: L = 8 tpl and the first and last radix divide 8 (fft_reg)
  const double2* tw;          // tw[m] = exp(-2 pi i m / L), m in [0, L)
};

__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cscale(double2 a, double s) { return make_double2(a.x * s, a.y * s); }
// multiply by +i (S=+1) or -i (S=-1)
template <int S> __device__ __forceinline__ double2 cmuli(double2 a) {
  return S > 0 ? make_double2(-a.y, a.x) : make_double2(a.y, -a.x);
}
__device__ __forceinline__ double cabs2(double2 a) { return a.x * a.x + a.y * a.y; }

// ----------------------------------------------------------------------------
// radix-R DFT butterflies, X_p = sum_q v_q exp(SIGN 2 pi i p q / R)
// ----------------------------------------------------------------------------
template <int SIGN> __device__ __forceinline__ void bfly2(double2* v) {
  double2 a = v[0], b = v[1];
  v[0] = cadd(a, b);
  v[1] = csub(a, b);
}
template <int SIGN> __device__ __forceinline__ void bfly3(double2* v) {
  // w = exp(SIGN 2 pi i / 3) = -1/2 + i SIGN sqrt(3)/2
  const double s = SIGN * 0.86602540378443864676;   // sqrt(3)/2
  double2 a = v[0], b = v[1], c = v[2];
  double2 t1 = cadd(b, c);
  double2 t2 = make_double2(a.x - 0.5 * t1.x, a.y - 0.5 * t1.y);
  double2 d = csub(b, c);
  double2 t3 = make_double2(-s * d.y, s * d.x);       // i s (b - c)
  v[0] = cadd(a, t1);
  v[1] = cadd(t2, t3);
  v[2] = csub(t2, t3);
}
template <int SIGN> __device__ __forceinline__ void bfly4(double2* v) {
  double2 a = v[0], b = v[1], c = v[2], d = v[3];
  double2 t0 = cadd(a, c), t1 = csub(a, c);
  double2 t2 = cadd(b, d), t3 = cmuli<SIGN>(csub(b, d));
  v[0] = cadd(t0, t2);
  v[2] = csub(t0, t2);
  v[1] = cadd(t1, t3);
  v[3] = csub(t1, t3);
}
template <int SIGN> __device__ __forceinline__ void bfly8(double2* v) {
  const double h = 0.70710678118654752440;           // sqrt(2)/2
  double2 e[4] = {v[0], v[2], v[4], v[6]};
  double2 o[4] = {v[1], v[3], v[5], v[7]};
  bfly4<SIGN>(e);
  bfly4<SIGN>(o);
  // twiddles w8^p, w8 = exp(SIGN i pi/4)
  double2 o1 = make_double2(h * (o[1].x - SIGN * o[1].y), h * (o[1].y + SIGN * o[1].x));
  double2 o2 = cmuli<SIGN>(o[2]);
  double2 o3 = make_double2(h * (-o[3].x - SIGN * o[3].y), h * (-o[3].y + SIGN * o[3].x));
  v[0] = cadd(e[0], o[0]); v[4] = csub(e[0], o[0]);
  v[1] = cadd(e[1], o1);   v[5] = csub(e[1], o1);
  v[2] = cadd(e[2], o2);   v[6] = csub(e[2], o2);
  v[3] = cadd(e[3], o3);   v[7] = csub(e[3], o3);
}
template <int SIGN> __device__ __forceinline__ void bfly5(double2* v) {
  // w = exp(SIGN 2 pi i / 5): cos and sin of 2 pi / 5 and 4 pi / 5
  const double c1 = 0.30901699437494742410, c2 = -0.80901699437494742410;
  const double s1 = SIGN * 0.95105651629515357212, s2 = SIGN * 0.58778525229247312917;
  const double2 a = v[0];
  const double2 t1 = cadd(v[1], v[4]), t2 = cadd(v[2], v[3]);
  const double2 t3 = csub(v[1], v[4]), t4 = csub(v[2], v[3]);
  const double2 r1 = make_double2(a.x + c1 * t1.x + c2 * t2.x, a.y + c1 * t1.y + c2 * t2.y);
  const double2 r2 = make_double2(a.x + c2 * t1.x + c1 * t2.x, a.y + c2 * t1.y + c1 * t2.y);
  // i (s1 t3 + s2 t4) and i (s2 t3 - s1 t4)
  const double2 u1 = make_double2(-(s1 * t3.y + s2 * t4.y), s1 * t3.x + s2 * t4.x);
  const double2 u2 = make_double2(-(s2 * t3.y - s1 * t4.y), s2 * t3.x - s1 * t4.x);
  v[0] = make_double2(a.x + t1.x + t2.x, a.y + t1.y + t2.y);
  v[1] = cadd(r1, u1);
  v[4] = csub(r1, u1);
  v[2] = cadd(r2, u2);
  v[3] = csub(r2, u2);
}
// radix 6 as two radix-3 DFTs of the even and odd inputs, A and B:
// X_k = A_{k mod 3} + w^k B_{k mod 3}, w = exp(SIGN i pi/3), w^3 = -1
template <int SIGN> __device__ __forceinline__ void bfly6(double2* v) {
  const double h = SIGN * 0.86602540378443864676;    // sqrt(3)/2
  double2 e[3] = {v[0], v[2], v[4]};
  double2 o[3] = {v[1], v[3], v[5]};
  bfly3<SIGN>(e);
  bfly3<SIGN>(o);
  // w o1 = (1/2 + i h) o1,  w^2 o2 = (-1/2 + i h) o2
  const double2 o1 = make_double2(0.5 * o[1].x - h * o[1].y, 0.5 * o[1].y + h * o[1].x);
  const double2 o2 = make_double2(-0.5 * o[2].x - h * o[2].y, -0.5 * o[2].y + h * o[2].x);
  v[0] = cadd(e[0], o[0]); v[3] = csub(e[0], o[0]);
  v[1] = cadd(e[1], o1);   v[4] = csub(e[1], o1);
  v[2] = cadd(e[2], o2);   v[5] = csub(e[2], o2);
}
template <int R, int SIGN> __device__ __forceinline__ void bfly(double2* v) {
  if constexpr (R == 2) bfly2<SIGN>(v);
  else if constexpr (R == 3) bfly3<SIGN>(v);
  else if constexpr (R == 4) bfly4<SIGN>(v);
  else if constexpr (R == 5) bfly5<SIGN>(v);
  else if constexpr (R == 6) bfly6<SIGN>(v);
  else bfly8<SIGN>(v);
}

// One Stockham autosort stage of radix R on a line of length L held in LDS.
// tpl threads cooperate on the line (lane in [0, tpl)); L/tpl <= 8 is
// guaranteed by the host, so a thread owns at most ceil(8/R) butterflies.
// tw is the twiddle table staged in LDS by the calling kernel.
// Every thread of the workgroup must call it (it contains barriers).
template <int R, int SIGN>
__device__ __forceinline__ void fft_stage(double2* s, int L, int Ns, unsigned mag, int step, int lane, int tpl,
                                          const double2* __restrict__ tw, bool active) {
  constexpr int MAXB = (8 + R - 1) / R;
  const int nb = L / R;
  double2 v[MAXB][R];
  if (active) {
#pragma unroll
    for (int b = 0; b < MAXB; ++b) {
      const int j = lane + b * tpl;
      if (j < nb) {
#pragma unroll
        for (int r = 0; r < R; ++r) v[b][r] = s[j + r * nb];
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
            double2 w = tw[r * k * step];
            if (SIGN > 0) w.y = -w.y;
            v[b][r] = cmul(v[b][r], w);
          }
        }
        bfly<R, SIGN>(v[b]);
        const int idxD = (j - k) * R + k;
#pragma unroll
        for (int r = 0; r < R; ++r) s[idxD + r * Ns] = v[b][r];
      }
    }
  }
  __syncthreads();
}

// In-place unnormalised DFT of the LDS line s (SIGN=-1 forward, +1 inverse);
// twl = the plan's twiddle table copied to LDS (stage_twiddles).
template <int SIGN>
__device__ __forceinline__ void fft_lds(double2* s, const FftPlan& p, const double2* twl, int lane, int tpl,
                                        bool active) {
  for (int st = 0; st < p.nst; ++st) {
    const int R = p.rad[st], Ns = p.ns[st], step = p.step[st];
    const unsigned mag = p.mag[st];
    if (R == 8) fft_stage<8, SIGN>(s, p.L, Ns, mag, step, lane, tpl, twl, active);
    else if (R == 4) fft_stage<4, SIGN>(s, p.L, Ns, mag, step, lane, tpl, twl, active);
    else if (R == 3) fft_stage<3, SIGN>(s, p.L, Ns, mag, step, lane, tpl, twl, active);
    else if (R == 5) fft_stage<5, SIGN>(s, p.L, Ns, mag, step, lane, tpl, twl, active);
    else if (R == 6) fft_stage<6, SIGN>(s, p.L, Ns, mag, step, lane, tpl, twl, active);
    else fft_stage<2, SIGN>(s, p.L, Ns, mag, step, lane, tpl, twl, active);
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
template <int R, int SIGN>
__device__ __forceinline__ void reg_first(double2 (&v)[8], double2* s, int lane, int tpl, bool active) {
  constexpr int Q = 8 / R;
  __syncthreads();                     // earlier readers of s are done
  if (active) {
#pragma unroll
    for (int b = 0; b < Q; ++b) {
      double2 u[R];
#pragma unroll
      for (int r = 0; r < R; ++r) u[r] = v[b + r * Q];
      bfly<R, SIGN>(u);
      const int j = lane + b * tpl;
#pragma unroll
      for (int r = 0; r < R; ++r) s[j * R + r] = u[r];
    }
  }
  __syncthreads();
}
template <int R, int SIGN>
__device__ __forceinline__ void reg_last(double2 (&v)[8], const double2* s, int L, int lane, int tpl,
                                         const double2* __restrict__ tw, bool active) {
  constexpr int Q = 8 / R;
  const int ns = L / R;                // twiddle stride L / (Ns R) = 1
  if (!active) return;
#pragma unroll
  for (int b = 0; b < Q; ++b) {
    const int j = lane + b * tpl;
    double2 u[R];
#pragma unroll
    for (int r = 0; r < R; ++r) u[r] = s[j + r * ns];
#pragma unroll
    for (int r = 1; r < R; ++r) {
      double2 w = tw[r * j];
      if (SIGN > 0) w.y = -w.y;
      u[r] = cmul(u[r], w);
    }
    bfly<R, SIGN>(u);
#pragma unroll
    for (int r = 0; r < R; ++r) v[b + r * Q] = u[r];
  }
}
template <int SIGN>
__device__ __forceinline__ void fft_reg(double2 (&v)[8], double2* s, const FftPlan& p, const double2* twl,
                                        int lane, int tpl, bool active) {
  const int n = p.nst;
  if (n == 1) {                        // L = 8, one thread per line
    if (active) bfly<8, SIGN>(v);
    return;
  }
  const int r0 = p.rad[0], rl = p.rad[n - 1];
  if (r0 == 8) reg_first<8, SIGN>(v, s, lane, tpl, active);
  else if (r0 == 4) reg_first<4, SIGN>(v, s, lane, tpl, active);
  else reg_first<2, SIGN>(v, s, lane, tpl, active);
  for (int st = 1; st < n - 1; ++st) {
    const int R = p.rad[st], Ns = p.ns[st], step = p.step[st];
    const unsigned mag = p.mag[st];
    if (R == 8) fft_stage<8, SIGN>(s, p.L, Ns, mag, step, lane, tpl, twl, active);
    else if (R == 4) fft_stage<4, SIGN>(s, p.L, Ns, mag, step, lane, tpl, twl, active);
    else if (R == 3) fft_stage<3, SIGN>(s, p.L, Ns, mag, step, lane, tpl, twl, active);
    else if (R == 5) fft_stage<5, SIGN>(s, p.L, Ns, mag, step, lane, tpl, twl, active);
    else if (R == 6) fft_stage<6, SIGN>(s, p.L, Ns, mag, step, lane, tpl, twl, active);
    else fft_stage<2, SIGN>(s, p.L, Ns, mag, step, lane, tpl, twl, active);
  }
  if (rl == 8) reg_last<8, SIGN>(v, s, p.L, lane, tpl, twl, active);
  else if (rl == 4) reg_last<4, SIGN>(v, s, p.L, lane, tpl, twl, active);
  else reg_last<2, SIGN>(v, s, p.L, lane, tpl, twl, active);
}
// The DFT of a line whose elements lane + m tpl (m < 8, those below L) are in
// v: register-resident when the plan allows it, else through the LDS line s
// (fft_lds).  Every thread of the workgroup calls it; results return in v.
template <int SIGN>
__device__ __forceinline__ void fft_line(double2 (&v)[8], double2* s, const FftPlan& p, const double2* twl,
                                         int lane, int tpl, bool active) {
  if (p.reg) {
    fft_reg<SIGN>(v, s, p, twl, lane, tpl, active);
    return;
  }
  __syncthreads();
  if (active) {
#pragma unroll
    for (int m = 0; m < 8; ++m) {
      const int e = lane + m * tpl;
      if (e < p.L) s[e] = v[m];
    }
  }
  __syncthreads();
  fft_lds<SIGN>(s, p, twl, lane, tpl, active);
  if (active) {
#pragma unroll
    for (int m = 0; m < 8; ++m) {
      const int e = lane + m * tpl;
      v[m] = (e < p.L) ? s[e] : make_double2(0.0, 0.0);
    }
  }
}

// Copies the twiddle table (L entries) into LDS; the caller's next barrier
// publishes it.  Issued together with the line loads, so its latency overlaps
// theirs instead of sitting inside every FFT stage.
__device__ __forceinline__ void stage_twiddles(double2* twl, const FftPlan& p) {
  for (int i = threadIdx.x; i < p.L; i += blockDim.x) twl[i] = p.tw[i];
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
