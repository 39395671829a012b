// Fused HIP kernels for the pseudospectral vorticity RHS and its RK4CN / AB4
// integrators (gfx950, FP64).
//
// Spectral arrays live in HBM in the reference's layout [Nx][ld] of double2,
// kx rows (0..Nx/2, -Nx/2+1..-1), ky columns 0..Ny/2 (ld >= Ny/2+1, padding
// columns are kept at zero).  One right-hand side evaluation
// (reference NonlinearRHSBatch, Solver/src/solver.c:884-1032) is two passes:
//
//   kpass  (column pass, one workgroup per group of C ky-columns)
//          [FWD] x-forward FFT of N(x,ky) -> K = (1/(NxNy))^2 * r2c(u.grad w),
//                (0,0)=0, dealias factor, Kolmogorov forcing  (solver.c:998-1027)
//          [EPI] integrator stage epilogue: RK4CN stage sums and update
//                (solver.c:690,729,768,853) or CN-AB4 update (solver.c:333),
//                with the ForceConjugacy global write (solver.c:2089-2101)
//          [INV] four x-inverse FFTs of the next stage input S:
//                G1 = X^-1[S/k^2], G2 = X^-1[-i kx S/k^2], G3 = X^-1[i kx S],
//                G4 = X^-1[S]   (solver.c:904-957 before the y transform)
//   rpass  (row pass, one workgroup per x row)
//          u = c2r_y(i ky G1), v = c2r_y(G2), w_x = c2r_y(G3), w_y = c2r_y(i ky G4)
//          via two complex inverse FFTs (u+iv, w_x+iw_y); n = u w_x + v w_y
//          (solver.c:981, note the + sign); N = r2c_y(n) via a half-length
//          complex FFT; optionally w = c2r_y(G4) for max|w| (solver.c:2219-2291)
//
// The c2r_y here ignores the imaginary parts of the ky=0 and ky=Ny/2 entries,
// which is FFTW's c2r behaviour on a non-Hermitian input column.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "ns2d_fft.hpp"
#include "ns2d_internal.h"

namespace ns2d {

__device__ __forceinline__ int kx_of_row(int i, int nx) { return i <= nx / 2 ? i : i - nx; }

// pow(x, e) with the exponents the solver uses evaluated exactly
__device__ __forceinline__ double pow_k2(double x, double e) {
  if (e == 1.0) return x;
  if (e == 2.0) return x * x;
  if (e == 0.0) return 1.0;
  return pow(x, e);
}

// D_fac / dt of the CN update for |k|^2 = k2 (k2 > 0), kind chosen on the host
// from the reference's flag logic (solver.c:835-850 / :315-330).
__device__ __forceinline__ double visc_coeff(const PhysArgs& ph, int kind, double k2) {
  switch (kind) {
    case VISC_HYPER_DRAG: return ph.nu * pow_k2(k2, ph.hyper_pow) + ph.alpha * pow_k2(k2, ph.drag_pow);
    case VISC_K2_DRAG: return ph.nu * k2 + ph.alpha * pow_k2(k2, ph.drag_pow);
    case VISC_HYPER: return ph.nu * pow_k2(k2, ph.hyper_pow);
    default: return ph.nu * k2;
  }
}

// ApplyDealiasing multiplier, solver.c:1039-1099
__device__ __forceinline__ double dealias_factor(const PhysArgs& ph, int kx, int ky) {
  const double k2 = (double)(kx * kx + ky * ky);
  switch (ph.filter) {
    case FILTER_23:
      if (k2 >= ph.kmax_sqr) return 0.0;
      if (ph.kolm_kill && kx == 0) return 0.0;
      return 1.0;
    case FILTER_HOULI_REF: {
      // integer division k/(N/2) (truncating): only the Nyquist row/column
      // differ from 1; the two possible values come precomputed from the host
      const int qx = kx / (ph.nx / 2), qy = ky / (ph.ny / 2);
      const int q = qx * qx + qy * qy;
      return q == 0 ? 1.0 : (q == 1 ? ph.houli_edge : ph.houli_corner);
    }
    case FILTER_HOULI: {
      const double a = (double)kx / (ph.nx / 2.0), b = (double)ky / (ph.ny / 2.0);
      const double r2 = a * a + b * b;
      const double r4 = r2 * r2, r8 = r4 * r4, r16 = r8 * r8, r32 = r16 * r16;
      return exp(-36.0 * (r32 * r4));  // (sqrt(r2))^36
    }
    default: return 1.0;
  }
}

// ---------------------------------------------------------------------------
// kpass: column pass in k-space
// ---------------------------------------------------------------------------
template <int L, int C, int T>
__global__ __launch_bounds__(T) void kpass_kernel(KPassArgs a) {
  constexpr int LS = (C > 1) ? L + 1 : L;   // padded line stride in LDS
  constexpr int NE = C * L;                 // elements handled by the workgroup
  constexpr int IT = (NE + T - 1) / T;
  __shared__ double2 lds[C * LS];

  const int tid = threadIdx.x;
  const int g = xcd_remap((int)blockIdx.x, (int)gridDim.x);
  const int ky0 = g * C;
  const int nyf = a.nyf, ld = a.ld, nx = L;
  const double2* __restrict__ tw = a.tw;
  const PhysArgs& ph = a.ph;

  double2 S[IT];     // next stage input (registers), element e = tid + it*T
  double2 K[IT];

  // ---------------- FWD: x-forward FFT of the row-pass output -------------
  if (a.mode & KP_FWD) {
#pragma unroll
    for (int it = 0; it < IT; ++it) {
      const int e = tid + it * T;
      if ((NE % T == 0) || e < NE) {
        const int c = e % C, x = e / C, ky = ky0 + c;
        lds[c * LS + x] = (ky < nyf) ? a.nin[(size_t)x * ld + ky] : make_double2(0.0, 0.0);
      }
    }
    __syncthreads();
    fft_lds<L, C, LS, T, -1>(lds, tw, tid);
    const double nrm = ph.norm2;
#pragma unroll
    for (int it = 0; it < IT; ++it) {
      const int e = tid + it * T;
      if ((NE % T == 0) || e < NE) {
        const int c = e % C, i = e / C, ky = ky0 + c;
        const int kx = kx_of_row(i, nx);
        double2 v = lds[c * LS + i];
        if (a.mode & KP_RAW) {
          // plain unnormalised forward transform (initial conditions)
        } else if (kx == 0 && ky == 0) {
          v = make_double2(0.0, 0.0);
        } else {
          v = cscale(nrm, v);                                  // solver.c:1005
          v = cscale(dealias_factor(ph, kx, ky), v);           // solver.c:1019
          if (ph.force_on && kx == 0 && ky == ph.force_k)      // solver.c:1022-1027
            v.x += ph.force_val;
        }
        K[it] = v;
      }
    }
  }

  // ---------------- EPI: integrator stage ----------------------------------
  const int epi = a.epi;
  const bool conj_fix = (a.mode & KP_CONJFIX) && ky0 == 0;
#pragma unroll
  for (int it = 0; it < IT; ++it) {
    const int e = tid + it * T;
    if (!((NE % T == 0) || e < NE)) continue;
    const int c = e % C, i = e / C, ky = ky0 + c;
    if (ky >= nyf) { S[it] = make_double2(0.0, 0.0); continue; }
    const size_t idx = (size_t)i * ld + ky;
    const int kx = kx_of_row(i, nx);
    if (epi == EPI_FROM_STATE) { S[it] = a.w_hat[idx]; continue; }
    if (epi == EPI_FROM_INPUT) { S[it] = a.sin[idx]; continue; }
    // the state, with the ForceConjugacy write of the RHS just finished:
    // w_hat[kx<0, 0] = conj(stage_input[|kx|, 0])  (solver.c:2096-2102)
    double2 w;
    const bool fixed = conj_fix && c == 0 && kx < 0;
    if (fixed) w = cconj(a.conj_src[-kx]);
    else if (epi != EPI_RHS) w = a.w_hat[idx];
    const double2 k = K[it];
    switch (epi) {
      case EPI_RHS:
        a.kout[idx] = k;
        if (fixed) a.w_hat[idx] = w;
        break;
      case EPI_RK1:
        a.acc[idx] = cscale(RK4_B1, k);
        if (a.hist_store) a.hist_store[idx] = k;
        S[it] = cadd(w, cscale(a.coef, k));                  // solver.c:690
        break;
      case EPI_RK2:
      case EPI_RK3: {
        const double b = (epi == EPI_RK2) ? RK4_B2 : RK4_B3;
        a.acc[idx] = cadd(a.acc[idx], cscale(b, k));
        S[it] = cadd(w, cscale(a.coef, k));                  // solver.c:729,768
        break;
      }
      case EPI_RK4CN: {
        double2 nw;
        if (kx == 0 && ky == 0) {
          nw = make_double2(0.0, 0.0);                         // solver.c:860
        } else {
          const double2 sum = cadd(a.acc[idx], cscale(RK4_B4, k));
          const double D = a.dt * visc_coeff(ph, a.visc_kind, (double)(kx * kx + ky * ky));
          nw = cadd(cscale((2.0 - D) / (2.0 + D), w), cscale(2.0 * a.dt / (2.0 + D), sum));  // :853
        }
        a.w_hat[idx] = nw;
        S[it] = nw;
        break;
      }
      case EPI_AB4: {
        double2 nw;
        if (kx == 0 && ky == 0) {
          nw = make_double2(0.0, 0.0);                         // solver.c:337
        } else {
          const double2 h2 = a.h_new[idx], h1 = a.h_mid[idx], h0 = a.h_old[idx];
          const double2 sum = cadd(cadd(cadd(cscale(AB4_1, k), cscale(AB4_2, h2)), cscale(AB4_3, h1)),
                                   cscale(AB4_4, h0));
          const double D = a.dt * visc_coeff(ph, a.visc_kind, (double)(kx * kx + ky * ky));
          nw = cadd(cscale((2.0 - D) / (2.0 + D), w), cscale(2.0 * a.dt / (2.0 + D), sum));  // :333
        }
        a.h_old[idx] = k;   // the oldest history slot becomes the newest (:346-348)
        a.w_hat[idx] = nw;
        S[it] = nw;
        break;
      }
      default: break;
    }
  }

  // blow-up bookkeeping for the reference's per-step SystemCheck
  // (solver.c:2301-2326): the stage-1 row pass left max|w| of the state that
  // this step started from in stat[0]; record it and reset the accumulator.
  if ((a.mode & KP_STAT) && blockIdx.x == 0 && tid == 0) {
    unsigned long long* st = a.stat;
    const unsigned long long m = st[STAT_WMAX_CUR];
    st[STAT_WMAX_LAST] = m;
    if (__longlong_as_double((long long)m) >= MAX_VORT_LIM && st[STAT_FIRST_BAD] == ~0ull)
      st[STAT_FIRST_BAD] = st[STAT_ITER];
    st[STAT_WMAX_CUR] = 0ull;
  }
  if ((a.mode & KP_ITER) && blockIdx.x == 0 && tid == 0) a.stat[STAT_ITER] += 1ull;

  if (!(a.mode & KP_INV)) return;

  // ky=0 column of the next stage input, for the next ForceConjugacy
  if (ky0 == 0 && a.conj_dst) {
#pragma unroll
    for (int it = 0; it < IT; ++it) {
      const int e = tid + it * T;
      if ((NE % T == 0) || e < NE) {
        if (e % C == 0) a.conj_dst[e / C] = S[it];
      }
    }
  }

  // ---------------- INV: four x-inverse FFTs --------------------------------
#pragma unroll 1
  for (int f = 0; f < 4; ++f) {
    __syncthreads();
#pragma unroll
    for (int it = 0; it < IT; ++it) {
      const int e = tid + it * T;
      if ((NE % T == 0) || e < NE) {
        const int c = e % C, i = e / C, ky = ky0 + c;
        const int kx = kx_of_row(i, nx);
        const int k2i = kx * kx + ky * ky;
        const double inv = k2i ? 1.0 / (double)k2i : 0.0;
        const double2 s = S[it];
        double2 v;
        if (f == 0) v = cscale(inv, s);                                      // psi_hat
        else if (f == 1) v = make_double2(kx * inv * s.y, -(kx * inv) * s.x);  // -i kx psi_hat
        else if (f == 2) v = make_double2(-kx * s.y, kx * s.x);              // i kx w_hat
        else v = s;                                                          // w_hat
        lds[c * LS + i] = v;
      }
    }
    __syncthreads();
    fft_lds<L, C, LS, T, +1>(lds, tw, tid);
    double2* __restrict__ gout = a.g[f];
#pragma unroll
    for (int it = 0; it < IT; ++it) {
      const int e = tid + it * T;
      if ((NE % T == 0) || e < NE) {
        const int c = e % C, x = e / C, ky = ky0 + c;
        if (ky < nyf) gout[(size_t)x * ld + ky] = lds[c * LS + x];
      }
    }
  }
}

// ---------------------------------------------------------------------------
// rpass: row pass in (x, ky) space
// ---------------------------------------------------------------------------
template <int NY, int T>
__global__ __launch_bounds__(T) void rpass_kernel(RPassArgs a) {
  constexpr int H = NY / 2;
  constexpr int ITY = (NY + T - 1) / T;     // y points per thread
  constexpr int ITK = (H + 1 + T - 1) / T;  // ky points per thread
  __shared__ double2 lds[NY];
  __shared__ double red[T / 64];

  const int tid = threadIdx.x;
  const int x = blockIdx.x;
  const int ld = a.ld;
  const double2* __restrict__ tw = a.tw;
  const size_t row = (size_t)x * ld;

  double n[ITY];
  double2 z1[ITY];

  // builds the full-length spectrum of (A + i B) from the half spectra of two
  // real fields; the ky=0 and ky=NY/2 entries keep only their real parts
  auto build_pair = [&](int which) {
#pragma unroll
    for (int it = 0; it < ITK; ++it) {
      const int ky = tid + it * T;
      if (ky <= H) {
        double2 A, B;
        const double fky = (double)ky;
        if (which == 0) {         // u = i ky G1, v = G2
          const double2 g1 = a.g[0][row + ky];
          A = make_double2(-fky * g1.y, fky * g1.x);
          B = a.g[1][row + ky];
        } else if (which == 1) {  // w_x = G3, w_y = i ky G4
          A = a.g[2][row + ky];
          const double2 g4 = a.g[3][row + ky];
          B = make_double2(-fky * g4.y, fky * g4.x);
        } else {                  // w = G4, paired with zero
          A = a.g[3][row + ky];
          B = make_double2(0.0, 0.0);
        }
        if (ky == 0 || ky == H) { A.y = 0.0; B.y = 0.0; }
        lds[ky] = make_double2(A.x - B.y, A.y + B.x);
        if (ky > 0 && ky < H) lds[NY - ky] = make_double2(A.x + B.y, B.x - A.y);
      }
    }
  };

  if (a.mode & RP_NL) {
    build_pair(0);
    __syncthreads();
    fft_lds<NY, 1, NY, T, +1>(lds, tw, tid);
#pragma unroll
    for (int it = 0; it < ITY; ++it) {
      const int y = tid + it * T;
      if (y < NY) z1[it] = lds[y];
    }
    __syncthreads();
    build_pair(1);
    __syncthreads();
    fft_lds<NY, 1, NY, T, +1>(lds, tw, tid);
#pragma unroll
    for (int it = 0; it < ITY; ++it) {
      const int y = tid + it * T;
      if (y < NY) {
        const double2 z2 = lds[y];
        n[it] = z1[it].x * z2.x + z1[it].y * z2.y;   // u*w_x + v*w_y
      }
    }
  }

  if (a.mode & RP_W) {
    __syncthreads();
    build_pair(2);
    __syncthreads();
    fft_lds<NY, 1, NY, T, +1>(lds, tw, tid);
    double m = 0.0;
#pragma unroll
    for (int it = 0; it < ITY; ++it) {
      const int y = tid + it * T;
      if (y < NY) {
        const double w = lds[y].x * a.wscale;
        if (a.wreal) a.wreal[(size_t)x * NY + y] = w;
        const double aw = fabs(w);
        m = (aw > m || aw != aw) ? aw : m;
      }
    }
    // wave then block max; a NaN propagates (its bit pattern compares largest)
    unsigned long long mb = (unsigned long long)__double_as_longlong(m);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      unsigned long long o = __shfl_xor(mb, off, 64);
      mb = o > mb ? o : mb;
    }
    if ((tid & 63) == 0) red[tid >> 6] = __longlong_as_double((long long)mb);
    __syncthreads();
    if (tid == 0) {
      unsigned long long bm = 0;
      for (int w = 0; w < T / 64; ++w) {
        unsigned long long v = (unsigned long long)__double_as_longlong(red[w]);
        bm = v > bm ? v : bm;
      }
      atomicMax(a.wmax, bm);
    }
  }

  if (a.mode & RP_NL) {
    __syncthreads();
    double* ldsr = reinterpret_cast<double*>(lds);
#pragma unroll
    for (int it = 0; it < ITY; ++it) {
      const int y = tid + it * T;
      if (y < NY) ldsr[y] = n[it];
    }
    __syncthreads();
    // packed c[m] = n[2m] + i n[2m+1]; C = FFT_H(c)
    fft_lds<H, 1, H, T, -1>(lds, tw, tid);
    constexpr int TWS = NS2D_TW_LEN / NY;
    double2* __restrict__ out = a.nout + row;
#pragma unroll
    for (int it = 0; it < ITK; ++it) {
      const int k = tid + it * T;
      if (k <= H) {
        double2 X;
        if (k == 0 || k == H) {
          const double2 c0 = lds[0];
          X = make_double2(k == 0 ? c0.x + c0.y : c0.x - c0.y, 0.0);
        } else {
          const double2 p = lds[k];
          const double2 q = cconj(lds[H - k]);
          const double2 E = make_double2(0.5 * (p.x + q.x), 0.5 * (p.y + q.y));
          const double2 O = make_double2(0.5 * (p.y - q.y), -0.5 * (p.x - q.x));  // (p-q)/(2i)
          X = cadd(E, cmul(tw[k * TWS], O));
        }
        out[k] = X;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// forward transform of a real field (ICs): r2c along y per row
// ---------------------------------------------------------------------------
template <int NY, int T>
__global__ __launch_bounds__(T) void r2c_rows_kernel(const double* __restrict__ in, double2* __restrict__ out,
                                                       const double2* __restrict__ tw, int ld) {
  constexpr int H = NY / 2;
  constexpr int ITY = (NY + T - 1) / T;
  constexpr int ITK = (H + 1 + T - 1) / T;
  __shared__ double2 lds[H];
  const int tid = threadIdx.x, x = blockIdx.x;
  double* ldsr = reinterpret_cast<double*>(lds);
#pragma unroll
  for (int it = 0; it < ITY; ++it) {
    const int y = tid + it * T;
    if (y < NY) ldsr[y] = in[(size_t)x * NY + y];
  }
  __syncthreads();
  fft_lds<H, 1, H, T, -1>(lds, tw, tid);
  constexpr int TWS = NS2D_TW_LEN / NY;
#pragma unroll
  for (int it = 0; it < ITK; ++it) {
    const int k = tid + it * T;
    if (k <= H) {
      double2 X;
      if (k == 0 || k == H) {
        const double2 c0 = lds[0];
        X = make_double2(k == 0 ? c0.x + c0.y : c0.x - c0.y, 0.0);
      } else {
        const double2 p = lds[k];
        const double2 q = cconj(lds[H - k]);
        const double2 E = make_double2(0.5 * (p.x + q.x), 0.5 * (p.y + q.y));
        const double2 O = make_double2(0.5 * (p.y - q.y), -0.5 * (p.x - q.x));
        X = cadd(E, cmul(tw[k * TWS], O));
      }
      out[(size_t)x * ld + k] = X;
    }
  }
}

// ---------------------------------------------------------------------------
// measures: one workgroup per spectrum bin (annulus round(|k|) = b)
// ---------------------------------------------------------------------------
__device__ __forceinline__ long isqrt_floor(long v) {
  if (v < 0) return -1;
  long r = (long)sqrt((double)v);
  while (r * r > v) --r;
  while ((r + 1) * (r + 1) <= v) ++r;
  return r;
}

template <int T>
__global__ __launch_bounds__(T) void measures_kernel(MeasArgs a) {
  __shared__ double red[NQ][T / 64];
  const int b = blockIdx.x;
  const int tid = threadIdx.x;
  const int nx = a.nx, ny = a.ny, H = ny / 2;
  double acc[NQ];
#pragma unroll
  for (int q = 0; q < NQ; ++q) acc[q] = 0.0;
  if (b > 0) {
    // b^2 - b < k^2 <= b^2 + b  <=>  round(|k|) == b
    const long lo = (long)b * b - b, hi = (long)b * b + b;
    const int kxmax = b + 1;
    for (int kxo = tid; kxo <= 2 * kxmax; kxo += T) {
      const int kx = kxo - kxmax;
      if (kx < -nx / 2 + 1 || kx > nx / 2) continue;
      const long kx2 = (long)kx * kx;
      long ylo = (lo - kx2 < 0) ? 0 : isqrt_floor(lo - kx2) + 1;
      long yhi = isqrt_floor(hi - kx2);
      if (yhi > H) yhi = H;
      const int i = kx >= 0 ? kx : kx + nx;
      for (long ky = ylo; ky <= yhi; ++ky) {
        const size_t idx = (size_t)i * a.ld + ky;
        const double2 w = a.w_hat[idx];
        const double k2 = (double)(kx2 + ky * ky);
        const double c = (ky == 0 || ky == H) ? 1.0 : 2.0;
        const double a2 = w.x * w.x + w.y * w.y;
        const double pre = visc_coeff(a.ph, a.visc_kind, k2);
        acc[Q_ENRG] += c * a2 / k2;
        acc[Q_ENST] += c * a2;
        acc[Q_PALIN] += c * k2 * a2;
        // |u|^2+|v|^2 with u = i ky w/k^2, v = -i kx w/k^2
        const double uv2 = (double)(ky * ky) * a2 / (k2 * k2) + (double)kx2 * a2 / (k2 * k2);
        acc[Q_ENRG_DISS] += c * pre * uv2;
        acc[Q_ENST_DISS] += c * ((a.quirk_enst_diss && c == 2.0) ? 1.0 : pre) * a2;
        {
          // |i (kx u + ky v)|^2, sys_msr.c:186-193 (zero up to rounding)
          const double2 u = make_double2(-(double)ky * w.y / k2, (double)ky * w.x / k2);
          const double2 v = make_double2((double)kx * w.y / k2, -(double)kx * w.x / k2);
          const double dx = (double)kx * u.x + (double)ky * v.x, dy = (double)kx * u.y + (double)ky * v.y;
          acc[Q_DIV] += c * (dx * dx + dy * dy);
        }
        if (a.rhs) {
          double2 r = a.rhs[idx];
          // the flux terms ignore the forcing (sys_msr.c:128-132)
          if (a.ph.force_on && kx == 0 && ky == a.ph.force_k) r.x -= a.ph.force_val;
          const double deriv = 2.0 * (w.x * r.x + w.y * r.y);
          const double diss = pre * a2;
          acc[Q_DERIV] += c * deriv;
          acc[Q_DISS] += c * diss;
          acc[Q_DERIV_K2] += c * deriv / k2;
          acc[Q_DISS_K2] += c * diss / k2;
          if (k2 < a.sbst_hi && k2 >= a.sbst_lo) {
            acc[Q_SB_DERIV] += c * deriv;
            acc[Q_SB_DISS] += c * diss;
            acc[Q_SB_DERIV_K2] += c * deriv / k2;
            acc[Q_SB_DISS_K2] += c * diss / k2;
          }
        }
      }
    }
  }
  // deterministic block reduction: wave shuffles, then a fixed-order sum
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    double v = acc[q];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if ((tid & 63) == 0) red[q][tid >> 6] = v;
  }
  __syncthreads();
  if (tid < NQ) {
    double s = 0.0;
    for (int w = 0; w < T / 64; ++w) s += red[tid][w];
    a.out[(size_t)b * NQ + tid] = s;
  }
}

// ---------------------------------------------------------------------------
// host-side launchers (template dispatch on the transform lengths)
// ---------------------------------------------------------------------------
template <int L>
struct KCfg {
  static constexpr int C = L <= 256 ? 16 : (4096 / L);
  static constexpr int T = (C * L / 4) < 256 ? ((C * L / 4) < 64 ? 64 : C * L / 4) : 256;
};

template <int L>
static hipError_t launch_kpass_t(const KPassArgs& a, hipStream_t s) {
  constexpr int C = KCfg<L>::C, T = KCfg<L>::T;
  const int groups = (a.nyf + C - 1) / C;
  hipLaunchKernelGGL((kpass_kernel<L, C, T>), dim3(groups), dim3(T), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_kpass(int nx, const KPassArgs& a, hipStream_t s) {
  switch (nx) {
    case 8: return launch_kpass_t<8>(a, s);
    case 16: return launch_kpass_t<16>(a, s);
    case 32: return launch_kpass_t<32>(a, s);
    case 64: return launch_kpass_t<64>(a, s);
    case 128: return launch_kpass_t<128>(a, s);
    case 256: return launch_kpass_t<256>(a, s);
    case 512: return launch_kpass_t<512>(a, s);
    case 1024: return launch_kpass_t<1024>(a, s);
    case 2048: return launch_kpass_t<2048>(a, s);
    case 4096: return launch_kpass_t<4096>(a, s);
    default: return hipErrorInvalidValue;
  }
}

template <int NY>
struct RCfg {
  static constexpr int T = NY >= 1024 ? 256 : (NY / 4 < 64 ? 64 : NY / 4);
};

template <int NY>
static hipError_t launch_rpass_t(int nx, const RPassArgs& a, hipStream_t s) {
  constexpr int T = RCfg<NY>::T;
  hipLaunchKernelGGL((rpass_kernel<NY, T>), dim3(nx), dim3(T), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_rpass(int nx, int ny, const RPassArgs& a, hipStream_t s) {
  switch (ny) {
    case 8: return launch_rpass_t<8>(nx, a, s);
    case 16: return launch_rpass_t<16>(nx, a, s);
    case 32: return launch_rpass_t<32>(nx, a, s);
    case 64: return launch_rpass_t<64>(nx, a, s);
    case 128: return launch_rpass_t<128>(nx, a, s);
    case 256: return launch_rpass_t<256>(nx, a, s);
    case 512: return launch_rpass_t<512>(nx, a, s);
    case 1024: return launch_rpass_t<1024>(nx, a, s);
    case 2048: return launch_rpass_t<2048>(nx, a, s);
    case 4096: return launch_rpass_t<4096>(nx, a, s);
    default: return hipErrorInvalidValue;
  }
}

template <int NY>
static hipError_t launch_r2c_rows_t(int nx, const double* in, double2* out, const double2* tw, int ld,
                                    hipStream_t s) {
  constexpr int T = RCfg<NY>::T;
  hipLaunchKernelGGL((r2c_rows_kernel<NY, T>), dim3(nx), dim3(T), 0, s, in, out, tw, ld);
  return hipGetLastError();
}

hipError_t launch_r2c_rows(int nx, int ny, const double* in, double2* out, const double2* tw, int ld,
                           hipStream_t s) {
  switch (ny) {
    case 8: return launch_r2c_rows_t<8>(nx, in, out, tw, ld, s);
    case 16: return launch_r2c_rows_t<16>(nx, in, out, tw, ld, s);
    case 32: return launch_r2c_rows_t<32>(nx, in, out, tw, ld, s);
    case 64: return launch_r2c_rows_t<64>(nx, in, out, tw, ld, s);
    case 128: return launch_r2c_rows_t<128>(nx, in, out, tw, ld, s);
    case 256: return launch_r2c_rows_t<256>(nx, in, out, tw, ld, s);
    case 512: return launch_r2c_rows_t<512>(nx, in, out, tw, ld, s);
    case 1024: return launch_r2c_rows_t<1024>(nx, in, out, tw, ld, s);
    case 2048: return launch_r2c_rows_t<2048>(nx, in, out, tw, ld, s);
    case 4096: return launch_r2c_rows_t<4096>(nx, in, out, tw, ld, s);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_measures(int nbins, const MeasArgs& a, hipStream_t s) {
  hipLaunchKernelGGL((measures_kernel<256>), dim3(nbins), dim3(256), 0, s, a);
  return hipGetLastError();
}

}  // namespace ns2d
