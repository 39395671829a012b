// Stockham auto-sort FFT for FP64 complex lines resident in LDS (gfx950).
//
// A workgroup of T threads transforms NL lines of length L (power of two) that
// sit in LDS at line stride LS (double2 elements).  Element j of a line is
// stored at padx(j) = j + j/8: one pad slot every 8 elements keeps the strided
// writes of the early passes (stride R elements) on distinct LDS banks.
//
// Each pass of radix R (8 while at least 8 points remain, then 4 or 2) reads R
// points per butterfly from LDS, applies the twiddles, does the R-point DFT in
// registers and writes the outputs to their auto-sorted positions, so the
// result lands in natural order with no bit-reversal pass.  Twiddles come from
// one table tw[m] = exp(-2*pi*i*m/NS2D_TW_LEN) built on the host in extended
// precision: a butterfly with span P and offset k loads w = tw[k*NS2D_TW_LEN/(P*R)]
// once and forms w^r by repeated multiplication.
//
// SIGN = -1 is the forward transform (e^{-i}), +1 the inverse (e^{+i});
// neither is normalised, matching FFTW's conventions.
#pragma once
#include <hip/hip_runtime.h>

#ifndef NS2D_TW_LEN
#define NS2D_TW_LEN 8192
#endif

namespace ns2d {

__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cconj(double2 a) { return make_double2(a.x, -a.y); }
__device__ __forceinline__ double2 cscale(double s, double2 a) { return make_double2(s * a.x, s * a.y); }

// 1/x for x >= 1 (integer |k|^2, the Crank-Nicolson denominators 2 + D):
// the hardware reciprocal refined by two Newton steps, r' = r + r (1 - x r),
// 5 VALU operations against about 10 and two scaling steps for the IEEE
// division; the result is within one unit in the last place of 1/x.
__device__ __forceinline__ double rcp_nr(double x) {
  double r = __builtin_amdgcn_rcp(x);
  double e = __builtin_fma(-x, r, 1.0);
  r = __builtin_fma(r, e, r);
  e = __builtin_fma(-x, r, 1.0);
  return __builtin_fma(r, e, r);
}
// multiply by SIGN*i
template <int SIGN>
__device__ __forceinline__ double2 cmul_si(double2 a) {
  return SIGN < 0 ? make_double2(a.y, -a.x) : make_double2(-a.y, a.x);
}

// padded LDS position of element j of a line: one slot of padding per 2^PS
// elements (PS = 3 everywhere except column-pass shapes tuned otherwise)
template <int PS = 3>
__host__ __device__ constexpr int padx_s(int j) { return j + (j >> PS); }
__host__ __device__ constexpr int padx(int j) { return padx_s<3>(j); }
// LDS line stride for lines of length L: padded length, odd so that
// neighbouring lines start on different banks
__host__ __device__ constexpr int line_stride(int L) { return padx(L) + 1; }

template <int SIGN>
__device__ __forceinline__ void dft2(double2* v) {
  double2 a = v[0], b = v[1];
  v[0] = cadd(a, b);
  v[1] = csub(a, b);
}

template <int SIGN>
__device__ __forceinline__ void dft4(double2* v) {
  double2 t0 = cadd(v[0], v[2]);
  double2 t1 = csub(v[0], v[2]);
  double2 t2 = cadd(v[1], v[3]);
  double2 t3 = cmul_si<SIGN>(csub(v[1], v[3]));
  v[0] = cadd(t0, t2);
  v[2] = csub(t0, t2);
  v[1] = cadd(t1, t3);
  v[3] = csub(t1, t3);
}

template <int SIGN>
__device__ __forceinline__ void dft8(double2* v) {
  constexpr double H = 0.70710678118654752440084436210485;
  double2 e[4] = {v[0], v[2], v[4], v[6]};
  double2 o[4] = {v[1], v[3], v[5], v[7]};
  dft4<SIGN>(e);
  dft4<SIGN>(o);
  // o[k] *= W8^k, W8 = exp(SIGN*2*pi*i/8)
  double2 o1 = o[1], o3 = o[3];
  if (SIGN < 0) {
    o[1] = make_double2(H * (o1.x + o1.y), H * (o1.y - o1.x));
    o[3] = make_double2(H * (o3.y - o3.x), -H * (o3.x + o3.y));
  } else {
    o[1] = make_double2(H * (o1.x - o1.y), H * (o1.y + o1.x));
    o[3] = make_double2(-H * (o3.x + o3.y), H * (o3.x - o3.y));
  }
  o[2] = cmul_si<SIGN>(o[2]);
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    v[k] = cadd(e[k], o[k]);
    v[k + 4] = csub(e[k], o[k]);
  }
}

template <int R, int SIGN>
__device__ __forceinline__ void dft(double2* v) {
  if constexpr (R == 2) dft2<SIGN>(v);
  else if constexpr (R == 4) dft4<SIGN>(v);
  else dft8<SIGN>(v);
}

template <int SIGN>
__device__ __forceinline__ double2 twiddle(const double2* __restrict__ tw, int m) {
  double2 w = tw[m];
  return SIGN < 0 ? w : make_double2(w.x, -w.y);
}

// One radix-R Stockham pass with span P over NL lines.
template <int L, int R, int P, int NL, int LS, int T, int SIGN>
__device__ __forceinline__ void stockham_pass(double2* lds, const double2* __restrict__ tw, int tid) {
  constexpr int NBF = L / R;
  constexpr int TOT = NL * NBF;
  constexpr int IT = (TOT + T - 1) / T;
  constexpr int TWS = NS2D_TW_LEN / (P * R);
  double2 v[IT][R];
#pragma unroll
  for (int it = 0; it < IT; ++it) {
    const int b = tid + it * T;
    if ((TOT % T == 0) || b < TOT) {
      const int line = b / NBF, i = b % NBF;
      const double2* src = lds + line * LS;
#pragma unroll
      for (int r = 0; r < R; ++r) v[it][r] = src[padx(i + r * NBF)];
    }
  }
  __syncthreads();
#pragma unroll
  for (int it = 0; it < IT; ++it) {
    const int b = tid + it * T;
    if ((TOT % T == 0) || b < TOT) {
      const int line = b / NBF, i = b % NBF;
      const int k = i & (P - 1);
      if constexpr (P > 1) {
        const double2 w = twiddle<SIGN>(tw, k * TWS);
        double2 wr = w;
#pragma unroll
        for (int r = 1; r < R; ++r) {
          v[it][r] = cmul(v[it][r], wr);
          if (r + 1 < R) wr = cmul(wr, w);
        }
      }
      dft<R, SIGN>(v[it]);
      double2* dst = lds + line * LS;
      const int j0 = (i - k) * R + k;
#pragma unroll
      for (int r = 0; r < R; ++r) dst[padx(j0 + r * P)] = v[it][r];
    }
  }
  __syncthreads();
}

template <int L, int P, int NL, int LS, int T, int SIGN>
__device__ __forceinline__ void fft_passes(double2* lds, const double2* __restrict__ tw, int tid) {
  constexpr int REM = L / P;
  if constexpr (REM > 1) {
    constexpr int R = REM >= 8 ? 8 : REM;
    stockham_pass<L, R, P, NL, LS, T, SIGN>(lds, tw, tid);
    fft_passes<L, P * R, NL, LS, T, SIGN>(lds, tw, tid);
  }
}

// Transform NL lines of length L in LDS in place (natural order in and out,
// padded positions).  Requires a barrier before the call if the data was just
// written by other threads; ends with a barrier.
template <int L, int NL, int LS, int T, int SIGN>
__device__ __forceinline__ void fft_lds(double2* lds, const double2* __restrict__ tw, int tid) {
  static_assert((L & (L - 1)) == 0 && L >= 2, "power-of-two length");
  static_assert(L <= NS2D_TW_LEN, "twiddle table too short");
  fft_passes<L, 1, NL, LS, T, SIGN>(lds, tw, tid);
}

}  // namespace ns2d
