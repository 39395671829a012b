// Register-resident Stockham FFT for FP64 complex lines (gfx950).
//
// A line of N points (power of two) is owned by TL threads; thread l holds the
// E = N/TL points x[l + TL*e], e = 0..E-1 (the "stride-TL distribution").
// Each pass of radix R = min(16, E, remaining) runs E/R butterflies per thread
// entirely in registers: butterfly i = l + TL*m reads x[i + r*N/R], which is
// register m + r*E/R of the same thread.  Between passes the outputs are
// written to their Stockham positions in an LDS line buffer and read back in
// the stride-TL distribution (one exchange per pass boundary).  The outputs of
// the last pass land in natural order in the same stride-TL distribution, so a
// transform starts and ends in registers: an N = 4096 line with E = 16 costs
// three radix-16 passes and two LDS exchanges.
//
// LDS positions are padded (padx_s<PS>: one slot every 2^PS elements) so that
// the stride-R writes of the first pass hit distinct banks.  Twiddles: one load
// w = tw[k*NS2D_TW_LEN/(P*R)] per butterfly from the host-built table
// exp(-2*pi*i*m/NS2D_TW_LEN); the powers w^r are formed by multiplication
// (apply_twiddles below).
//
// SIGN = -1: forward (e^{-i}); +1: inverse (e^{+i}); unnormalised, as FFTW.
// BLOCK selects the exchange barrier: __syncthreads() when a line's threads
// span several waves, a wavefront-scope fence otherwise.
#pragma once
#include "ns2d_fft.hpp"

namespace ns2d {

template <bool BLOCK>
__device__ __forceinline__ void xsync() {
  if constexpr (BLOCK) {
    __syncthreads();
  } else {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
}

// 16-point DFT as 4 x 4: n = n1 + 4 n2, k = k1 + 4 k2
template <int SIGN>
__device__ __forceinline__ void dft16(double2* v) {
  constexpr double C1 = 0.92387953251128675612818318939679;  // cos(pi/8)
  constexpr double S1 = 0.38268343236508977172845998403040;  // sin(pi/8)
  constexpr double H = 0.70710678118654752440084436210485;   // cos(pi/4)
  double2 y[4][4];
#pragma unroll
  for (int n1 = 0; n1 < 4; ++n1) {
    double2 t[4] = {v[n1], v[n1 + 4], v[n1 + 8], v[n1 + 12]};
    dft4<SIGN>(t);
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) y[n1][k1] = t[k1];
  }
  // y[n1][k1] *= W16^(n1*k1), W16 = exp(SIGN*2*pi*i/16)
  auto rot = [](double2 a, double c, double s) {  // a * (c + SIGN*i*s)
    return SIGN < 0 ? make_double2(a.x * c + a.y * s, a.y * c - a.x * s)
                    : make_double2(a.x * c - a.y * s, a.y * c + a.x * s);
  };
  y[1][1] = rot(y[1][1], C1, S1);    // W^1
  y[1][2] = rot(y[1][2], H, H);      // W^2
  y[1][3] = rot(y[1][3], S1, C1);    // W^3
  y[2][1] = rot(y[2][1], H, H);      // W^2
  y[2][2] = cmul_si<SIGN>(y[2][2]);  // W^4
  y[2][3] = rot(y[2][3], -H, H);     // W^6
  y[3][1] = rot(y[3][1], S1, C1);    // W^3
  y[3][2] = rot(y[3][2], -H, H);     // W^6
  y[3][3] = rot(y[3][3], -C1, -S1);  // W^9
#pragma unroll
  for (int k1 = 0; k1 < 4; ++k1) {
    double2 t[4] = {y[0][k1], y[1][k1], y[2][k1], y[3][k1]};
    dft4<SIGN>(t);
#pragma unroll
    for (int k2 = 0; k2 < 4; ++k2) v[k1 + 4 * k2] = t[k2];
  }
}

template <int R, int SIGN>
__device__ __forceinline__ void dftr(double2* v) {
  if constexpr (R == 16) dft16<SIGN>(v);
  else if constexpr (R == 1) {}
  else dft<R, SIGN>(v);
}

// v[r] *= w^r for r = 1..R-1, with w^(4a+b) = (w^4)^a w^b so that only
// w, w^2, w^3 and the running (w^4)^a are live (at most five roundings deep)
template <int R>
__device__ __forceinline__ void apply_twiddles(double2* v, double2 w) {
  if constexpr (R == 2) {
    v[1] = cmul(v[1], w);
  } else {
    const double2 w2 = cmul(w, w);
    const double2 w3 = cmul(w2, w);
    v[1] = cmul(v[1], w);
    v[2] = cmul(v[2], w2);
    v[3] = cmul(v[3], w3);
    if constexpr (R > 4) {
      const double2 w4 = cmul(w2, w2);
      double2 base = w4;
#pragma unroll
      for (int a = 1; a < R / 4; ++a) {
        v[4 * a] = cmul(v[4 * a], base);
        v[4 * a + 1] = cmul(v[4 * a + 1], cmul(base, w));
        v[4 * a + 2] = cmul(v[4 * a + 2], cmul(base, w2));
        v[4 * a + 3] = cmul(v[4 * a + 3], cmul(base, w3));
        if (a + 1 < R / 4) base = cmul(base, w4);
      }
    }
  }
}

// Twiddles of one thread's butterflies, loaded once per kernel: the passes
// with span P > 1 use E/R of them each (one per butterfly), the same in both
// directions (conjugated for the inverse).  Preloading them keeps the table
// loads out of the transforms' barrier-separated passes.
template <int N, int TL, int P>
constexpr int tw_count() {
  if constexpr (N / P <= 1) {
    return 0;
  } else {
    constexpr int E = N / TL, REM = N / P;
    constexpr int R0 = E < 16 ? E : 16;
    constexpr int R = R0 < REM ? R0 : REM;
    return (P > 1 ? E / R : 0) + tw_count<N, TL, P * R>();
  }
}

template <int N, int TL>
struct TwReg {
  static constexpr int n = tw_count<N, TL, 1>();
  double2 w[n > 0 ? n : 1];
};

template <int N, int TL, int P, int OFF>
__device__ __forceinline__ void load_twiddles_pass(TwReg<N, TL>& t, const double2* __restrict__ tw, int l) {
  constexpr int E = N / TL, REM = N / P;
  if constexpr (REM > 1) {
    constexpr int R0 = E < 16 ? E : 16;
    constexpr int R = R0 < REM ? R0 : REM;
    constexpr int NB = E / R;
    constexpr int TWS = NS2D_TW_LEN / (P * R);
    if constexpr (P > 1) {
#pragma unroll
      for (int m = 0; m < NB; ++m) t.w[OFF + m] = tw[((l + TL * m) & (P - 1)) * TWS];
    }
    load_twiddles_pass<N, TL, P * R, OFF + (P > 1 ? NB : 0)>(t, tw, l);
  }
}

// t.w = this thread's twiddles (lane l of a line owned by TL threads)
template <int N, int TL>
__device__ __forceinline__ void load_twiddles(TwReg<N, TL>& t, const double2* __restrict__ tw, int l) {
  load_twiddles_pass<N, TL, 1, 0>(t, tw, l);
}

// Every power w^r (r = 1..R-1) of each butterfly's twiddle, loaded from the
// table once per kernel (w^r = exp(-2 pi i k r/(P R)) = tw[k r NS2D_TW_LEN/(P R)],
// k r < P R): the passes then multiply by table values instead of forming the
// powers by products (about 28 complex products per radix-16 butterfly), at
// the cost of (R-1) E/R registers pairs per pass, for shapes whose occupancy
// does not depend on their register count (one wave per SIMD).
template <int N, int TL, int P>
constexpr int tw_full_count() {
  if constexpr (N / P <= 1) {
    return 0;
  } else {
    constexpr int E = N / TL, REM = N / P;
    constexpr int R0 = E < 16 ? E : 16;
    constexpr int R = R0 < REM ? R0 : REM;
    return (P > 1 ? (E / R) * (R - 1) : 0) + tw_full_count<N, TL, P * R>();
  }
}

template <int N, int TL>
struct TwFull {
  static constexpr int n = tw_full_count<N, TL, 1>();
  double2 w[n > 0 ? n : 1];
};

template <int N, int TL, int P, int OFF>
__device__ __forceinline__ void load_twiddles_pass(TwFull<N, TL>& t, const double2* __restrict__ tw, int l) {
  constexpr int E = N / TL, REM = N / P;
  if constexpr (REM > 1) {
    constexpr int R0 = E < 16 ? E : 16;
    constexpr int R = R0 < REM ? R0 : REM;
    constexpr int NB = E / R;
    constexpr int TWS = NS2D_TW_LEN / (P * R);
    if constexpr (P > 1) {
#pragma unroll
      for (int m = 0; m < NB; ++m) {
        const int k = (l + TL * m) & (P - 1);
#pragma unroll
        for (int r = 1; r < R; ++r) t.w[OFF + m * (R - 1) + r - 1] = tw[k * r * TWS];
      }
    }
    load_twiddles_pass<N, TL, P * R, OFF + (P > 1 ? NB * (R - 1) : 0)>(t, tw, l);
  }
}

template <int N, int TL>
__device__ __forceinline__ void load_twiddles(TwFull<N, TL>& t, const double2* __restrict__ tw, int l) {
  load_twiddles_pass<N, TL, 1, 0>(t, tw, l);
}

// twiddle slots a butterfly of radix R uses in a set of type TW
template <typename TW>
struct TwSlots {
  template <int R>
  static constexpr int per_butterfly() { return 1; }
};
template <int N, int TL>
struct TwSlots<TwFull<N, TL>> {
  template <int R>
  static constexpr int per_butterfly() { return R - 1; }
};

// twiddle source: the global table, or the thread's preloaded set
template <int SIGN, int N, int TL>
__device__ __forceinline__ double2 tw_get(const double2* __restrict__ tw, int idx, int) {
  return twiddle<SIGN>(tw, idx);
}
template <int SIGN, int N, int TL>
__device__ __forceinline__ double2 tw_get(const TwReg<N, TL>& t, int, int slot) {
  double2 w = t.w[slot];
  // opaque at the point of use: the powers w^r of the pass are formed here,
  // not hoisted to the kernel start where they would all stay live
  asm volatile("" : "+v"(w.x), "+v"(w.y));
  return SIGN < 0 ? w : make_double2(w.x, -w.y);
}

// LDS slot of line position p in the exchange after a pass of span P and
// radix R.  Padded (one slot every 2^PS), or, with NS2D_FFT_SWIZZLE, XOR-
// swizzled within rows of 16 slots (256 B, the 64 banks' width): sixteen
// consecutive lanes write the positions a*P*R + k + r*P (k < P, a < 16/P) of
// their 16/P blocks, which a pass with P < 16 puts on the same bank columns
// of different rows; XORing a row's columns with P * (its block number mod
// 16/P) gives every lane of the sixteen its own column, and the reads of 16
// consecutive positions stay in one row (a permutation of its columns).  The
// slot never leaves its row, so a line needs no more than N slots.
#ifndef NS2D_FFT_SWIZZLE
#define NS2D_FFT_SWIZZLE 0
#endif
template <int PS, int P, int R>
__device__ __forceinline__ int fft_slot(int p) {
  if constexpr (!NS2D_FFT_SWIZZLE) {
    return padx_s<PS>(p);
  } else if constexpr (P >= 16) {
    return p;   // sixteen lanes, sixteen consecutive positions
  } else {
    constexpr int PR = P * R;
    constexpr int RS = PR >= 16 ? PR / 16 : 1;   // rows per block
    constexpr int M = 16 / P;                    // blocks under sixteen lanes
    return p ^ ((((p >> 4) / RS) & (M - 1)) * P);
  }
}

// x[r] *= w^r for butterfly m of a pass whose twiddle slots start at OFF:
// powers formed from one twiddle (table or TwReg), or read from a TwFull set
template <int SIGN, int N, int TL, int R, int OFF, typename TW>
__device__ __forceinline__ void apply_butterfly_twiddles(double2* x, const TW& tw, int idx, int m) {
  apply_twiddles<R>(x, tw_get<SIGN, N, TL>(tw, idx, OFF + m));
}
template <int SIGN, int N, int TL, int R, int OFF>
__device__ __forceinline__ void apply_butterfly_twiddles(double2* x, const TwFull<N, TL>& tw, int, int m) {
#pragma unroll
  for (int r = 1; r < R; ++r) {
    const double2 w = tw.w[OFF + m * (R - 1) + r - 1];
    x[r] = cmul(x[r], SIGN < 0 ? w : make_double2(w.x, -w.y));
  }
}

template <int N, int TL, int SIGN, bool BLOCK, int P, int PS, int OFF, typename TW>
__device__ __forceinline__ void fft_reg_pass(double2* v, double2* buf, const TW& tw, int l) {
  constexpr int E = N / TL;
  constexpr int REM = N / P;
  if constexpr (REM > 1) {
    constexpr int R0 = E < 16 ? E : 16;
    constexpr int R = R0 < REM ? R0 : REM;
    constexpr int NB = E / R;          // butterflies per thread
    constexpr bool LAST = (P * R == N);
    constexpr int TWS = NS2D_TW_LEN / (P * R);
#pragma unroll
    for (int m = 0; m < NB; ++m) {
      const int i = l + TL * m;
      double2 x[R];
#pragma unroll
      for (int r = 0; r < R; ++r) x[r] = v[m + r * NB];
      const int k = i & (P - 1);
      if constexpr (P > 1) apply_butterfly_twiddles<SIGN, N, TL, R, OFF>(x, tw, k * TWS, m);
      dftr<R, SIGN>(x);
      if constexpr (LAST) {
#pragma unroll
        for (int r = 0; r < R; ++r) v[m + r * NB] = x[r];
      } else {
        const int j0 = (i - k) * R + k;
#pragma unroll
        for (int r = 0; r < R; ++r) buf[fft_slot<PS, P, R>(j0 + r * P)] = x[r];
      }
    }
    if constexpr (!LAST) {
      xsync<BLOCK>();
#pragma unroll
      for (int e = 0; e < E; ++e) v[e] = buf[fft_slot<PS, P, R>(l + TL * e)];
      xsync<BLOCK>();
      fft_reg_pass<N, TL, SIGN, BLOCK, P * R, PS,
                   OFF + (P > 1 ? NB * TwSlots<TW>::template per_butterfly<R>() : 0)>(v, buf, tw, l);
    }
  }
}

// In-register FFT of one line: v holds x[l + TL*e] on entry and X[l + TL*e]
// on exit.  buf: this line's LDS exchange buffer (padx_s<PS>(N) elements).
// All TL threads of the line (and, with BLOCK, every thread of the block)
// must call.
// tw: the global table (const double2*) or the thread's preloaded TwReg<N, TL>
// (loaded with the same l).
template <int N, int TL, int SIGN, bool BLOCK, int PS = 3, typename TW>
__device__ __forceinline__ void fft_reg(double2* v, double2* buf, const TW& tw, int l) {
  static_assert((N & (N - 1)) == 0 && N >= 2 && N % TL == 0, "bad FFT shape");
  static_assert(N <= NS2D_TW_LEN, "twiddle table too short");
  fft_reg_pass<N, TL, SIGN, BLOCK, 1, PS, 0>(v, buf, tw, l);
}

}  // namespace ns2d
