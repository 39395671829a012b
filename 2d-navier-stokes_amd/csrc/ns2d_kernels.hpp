// Device code of the fused HIP kernels (gfx950, FP64), shared by the
// translation units that instantiate the pass shapes (ns2d_pass_tu.hip, built
// once per transform-length range) and the launchers (ns2d_kernels.hip).
#pragma once
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
#include "ns2d_fft_reg.hpp"
#include "ns2d_internal.h"
#include <cstdio>
#include <cstdlib>
#include <vector>

// NS2D_TW_PRELOAD = 1: each thread loads its FFT twiddles once per kernel
// (TwReg) instead of from the table inside every transform pass.  Off by
// default: the registers the preloaded set holds cost occupancy (and spills
// in the 1024-thread shapes), and MI355X ran the passes slower with it
// (1024^2: column pass 36.2 against 33.3 us, 4096^2 RK4CN 357 against 384
// steps/s).
#ifndef NS2D_TW_PRELOAD
#define NS2D_TW_PRELOAD 0
#endif

namespace ns2d {

#if NS2D_TW_PRELOAD
#define NS2D_TWIDDLES(N, TLN, name, table, lane) \
  TwReg<N, TLN> name##_reg_;                     \
  load_twiddles<N, TLN>(name##_reg_, table, lane); \
  const TwReg<N, TLN>& name = name##_reg_;
#else
#define NS2D_TWIDDLES(N, TLN, name, table, lane) const double2* const name = table;
#endif

__device__ __forceinline__ int kx_of_row(int i, int nx) { return i <= nx / 2 ? i : i - nx; }

// Returns v unchanged but opaque to the optimiser.  Used at the top of the
// per-field / per-batch loops so that index and twiddle arithmetic is redone in
// each iteration instead of being hoisted and kept live across the loop (which
// otherwise exhausts the register file and spills to scratch).
__device__ __forceinline__ int opaque(int v) {
  asm volatile("" : "+v"(v));
  return v;
}

// Element i of a device array, addressed as the array's base plus a 32-bit
// unsigned byte offset: the global load/store form with the base in scalar
// registers and one 32-bit offset per lane, instead of a 64-bit address
// formed per lane.  Every array a pass addresses this way is one field of at
// most 4096 x 2056 complex values (135 MB; grids beyond 4096 have no shapes).
template <typename T>
__device__ __forceinline__ T& at(T* p, size_t i) {
  const unsigned off = (unsigned)i * (unsigned)sizeof(T);
  return *(T*)((const char*)p + off);
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
      // integer division k/(N/2) (truncating): with kx in (-Nx/2, Nx/2] and
      // ky in [0, Ny/2] the quotients are 1 on the Nyquist row/column and 0
      // elsewhere, so only those differ from 1; the two possible values come
      // precomputed from the host
      const int q = (kx == ph.nx / 2) + (ky == ph.ny / 2);
      return q == 0 ? 1.0 : (q == 1 ? ph.houli_edge : ph.houli_corner);
    }
    case FILTER_HOULI:   // exp(-36 (|k|/k_N)^36), tabulated on the host
      return ph.filt_tab[kidx(kx < 0 ? kx + ph.nx : kx, ky, ph.ld, ph.nx)];
    default: return 1.0;
  }
}

// ---------------------------------------------------------------------------
// kpass: column pass in k-space
//
// A work item is a group of C adjacent column slots.  Thread tid owns slot
// c = tid % C and the rows x = l + TL*e (l = tid / C, TL = T/C, e < E = L/TL),
// so each global access of a wave covers C contiguous complex values of a row,
// and the column's x-transform runs in registers (fft_reg) with the slot's own
// LDS exchange buffer.  KS workgroups share a column group: each redoes the
// forward transform and takes 4/KS of the inverse fields; only split 0 writes
// the integrator state (the new state goes to a separate buffer w_out, so no
// workgroup reads what another writes).
//
// PK (packed pair): slot s >= 1 is column ky = s and slot 0 carries both the
// ky = 0 and the ky = Ny/2 = H column, so Ny/2 slots cover the Ny/2+1 columns
// (an even count: no work item is left holding the odd column alone).  Two
// columns go through one complex transform in both directions:
//   forward  the row pass leaves N(x,0) and N(x,H) real (they are the
//            y-transforms' k = 0 and k = H terms of real rows), so
//            Z = X[N(.,0) + i N(.,H)] gives K0 = (Z + conj Z(-kx))/2 and
//            KH = (Z - conj Z(-kx))/(2i);
//   inverse  the row pass uses one real number per entry of these columns:
//            after its i ky factor it keeps the real part (FFTW's c2r
//            projection), i.e. Re G2, Re G3, Re G4 at ky = 0 and H, Im G1 and
//            Im G4 at H (-H Im G = Re(i H G)), nothing of G1 at ky = 0.  With
//            Re X^-1[P] = X^-1[(P + conj P(-kx))/2], one transform of
//            (P + conj P(-kx))/2 + i (Q + conj Q(-kx))/2 gives Re X^-1[P] and
//            Re X^-1[Q]; the four transforms are
//              G1 slot: P = -i psi_H, Q = -i S_H  -> Im G1(x,H), Im G4(x,H)
//              G2 slot: P = -i kx psi_0, Q = -i kx psi_H  -> Re G2(x,0), Re G2(x,H)
//              G3 slot: P = i kx S_0, Q = i kx S_H        -> Re G3(x,0), Re G3(x,H)
//              G4 slot: P = S_0, Q = S_H                  -> Re G4(x,0), Re G4(x,H)
//            (G4(x,H) is written as two 8-byte halves from two slots; the
//            unused parts are written as zeros.)
// With PK = 1 the columns the transforms need wait in pairbuf ([split][2][Nx],
// per workgroup split): KH between the forward transform and its epilogue,
// then the next stage inputs S_0 and S_H.  With PK = 2 they stay in registers
// (E more per thread), so slot 0 costs only the extra LDS exchanges.
// ---------------------------------------------------------------------------
// PS: LDS padding (one slot per 2^PS elements); OFF: extra 16-B slots between
// the slots' lines, which sets where the neighbouring slot's accesses of one
// wave fall among the LDS banks (adjacent lanes work on adjacent slots)
template <int L, int C, int T, int KS, int PS, int OFF, int PK, int PF>
__device__ __forceinline__ void kpass_body(const KPassArgs& a, const int wi) {
  constexpr int TL = T / C;        // threads per column
  constexpr int E = L / TL;        // rows per thread
  constexpr int LS = padx_s<PS>(L) + OFF;
  constexpr int FPS = 4 / KS;      // inverse fields per split
  constexpr bool BLOCK = T > 64;
  static_assert(T % C == 0 && L % TL == 0 && E >= 2 && 4 % KS == 0, "bad kpass shape");
  __shared__ double2 lds[C * LS];

  const int tid = threadIdx.x;
  // with last_split (KS = 1 shapes), work items ngroups-1 and ngroups are the
  // last column group's two halves: two inverse fields each, split 0 writes
  const int ngroups = ((PK ? a.nyf - 1 : a.nyf) + C - 1) / C;
  const bool tail = KS == 1 && a.last_split && wi >= ngroups - 1;
  const int g = tail ? ngroups - 1 : wi / KS;
  const int split = tail ? wi - (ngroups - 1) : wi % KS;
  const int nfs = tail ? 2 : FPS;   // inverse fields of this work item
  const int c = tid % C, l = tid / C;
  const int ky = g * C + c;        // the slot; with PK slot 0 is the (0, Ny/2) pair
  const int nyf = a.nyf, ld = a.ld, nx = L;
  const int H = nyf - 1;           // ky of the Nyquist column
  const bool colv = ky < (PK ? H : nyf);
  const bool pair = PK && ky == 0;
  const bool pblk = PK && g == 0;  // uniform per workgroup
  const bool writer = split == 0;
  const PhysArgs& ph = a.ph;
  double2* buf = lds + c * LS;
  NS2D_TWIDDLES(L, TL, tw, a.tw, l)
  // the pair's held columns (see above)
  double2* __restrict__ hold = a.pairbuf + (size_t)split * 2 * L;

  double2 R[E];  // forward result K, then the next stage input S
  // PK = 2: the pair's second column (KH, then S_H) beside R; S_0 moves here
  // when the second column's epilogue runs (see the column loop)
  constexpr bool PKR = PK == 2;
  double2 RH[PKR ? E : 1];
#pragma unroll
  for (int e = 0; e < (PKR ? E : 1); ++e) RH[e] = make_double2(0.0, 0.0);

  // PF: the RK epilogue's state, accumulator and coefficient loads are issued
  // before the forward transform, so that their latency hides behind it
  // (unpacked shapes whose register budget leaves room for 2.5 E values)
  static_assert(!(PF && PK), "prefetch is for unpacked shapes");
  constexpr int EP = PF ? E : 1;
  double2 wpre[EP], apre[EP];
  double lpre[EP];
  // (every epilogue that reads the state runs after a forward transform on a
  // valid column, exactly when `pre` holds; the arrays are filled
  // unconditionally so that they stay in registers)
  const bool pfw = PF != 0;
  const bool pre = pfw && (a.mode & KP_FWD) && colv;
  if constexpr (PF != 0) {
#pragma unroll
    for (int it = 0; it < E; ++it) {
      wpre[it] = apre[it] = make_double2(0.0, 0.0);
      lpre[it] = 0.0;
    }
    const bool epi_rk = a.epi == EPI_RK1 || a.epi == EPI_RK2 || a.epi == EPI_RK3 || a.epi == EPI_RK4CN ||
                        a.epi == EPI_AB4;
    if (pfw && epi_rk && (a.mode & KP_FWD)) {   // uniform per workgroup
      const bool cf = (a.mode & KP_CONJFIX) && ky == 0;
      const bool need_acc = a.epi == EPI_RK4CN || ((a.epi == EPI_RK2 || a.epi == EPI_RK3) && writer);
      const bool need_lam = a.epi == EPI_RK4CN;
      // every row's loads issued without a branch around them: a slot past the
      // last column reads column 0 (in bounds) and keeps its zeros
      const int kyl = colv ? ky : 0;
#pragma unroll
      for (int it = 0; it < E; ++it) {
        const int i = l + TL * it;
        const int kx = kx_of_row(i, nx);
        const size_t idx = kidx(i, kyl, ld, nx);
        const bool cj = cf && kx < 0;
        const double2* pw = cj ? a.conj_src + (cj ? -kx : 0) : a.w_in + idx;
        const double2 w = *pw;
        if (pre) wpre[it] = make_double2(w.x, cj ? -w.y : w.y);
        if (need_acc) {
          const double2 av = at(a.acc, idx);
          if (pre) apre[it] = av;
        }
        if (need_lam) {
          const double m = at(a.lam, idx);
          if (pre) lpre[it] = m;
        }
      }
    }
  }

  // ---------------- FWD: x-forward FFT of the row-pass output -------------
  if (a.mode & KP_FWD) {
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int x = l + TL * e;
      if (pair) R[e] = make_double2(at(a.nin, nidx(x, 0, ld)).x, at(a.nin, nidx(x, H, ld)).x);
      else R[e] = colv ? at(a.nin, nidx(x, ky, ld)) : make_double2(0.0, 0.0);
      if ((e & 3) == 3) asm volatile("" ::: "memory");
    }
    fft_reg<L, TL, -1, BLOCK, PS>(R, buf, tw, opaque(l));
    if (pblk) {
      if (pair) {
#pragma unroll
        for (int e = 0; e < E; ++e) buf[padx_s<PS>(l + TL * e)] = R[e];
      }
      xsync<BLOCK>();
      if (pair) {
#pragma unroll
        for (int e = 0; e < E; ++e) {
          const int x = l + TL * e;
          const double2 z = R[e], zm = buf[padx_s<PS>((L - x) & (L - 1))];
          R[e] = make_double2(0.5 * (z.x + zm.x), 0.5 * (z.y - zm.y));    // K0
          const double2 kh = make_double2(0.5 * (z.y + zm.y), -0.5 * (z.x - zm.x));  // KH
          if constexpr (PKR) RH[e] = kh;
          else at(hold, L + x) = kh;
        }
      }
      xsync<BLOCK>();
    }
  }

  // ---------------- EPI: integrator stage, per column of the slot ----------
  // The state w as the epilogue sees it, with the ForceConjugacy write of the
  // RHS just finished: w_hat[kx<0, 0] = conj(stage_input[|kx|, 0])
  // (solver.c:2096-2102).  Padding columns (ky >= Ny/2+1) carry zeros.
#pragma unroll 1
  for (int q = 0; q < (pair ? 2 : 1); ++q) {
    // opaque: keeps the row offsets from being hoisted out of the loop and
    // held live across it
    const int kq = opaque(q ? H : ky);
    const int lq = opaque(l);
    auto index_of = [&](int it) -> size_t { return kidx(lq + TL * it, kq, ld, nx); };
    if (PKR && q == 1) {
      // S_0 to RH, KH (forward modes) to R
#pragma unroll
      for (int e = 0; e < (PKR ? E : 1); ++e) {
        const double2 t = R[e];
        R[e] = RH[e];
        RH[e] = t;
      }
    }
    if (a.mode & KP_FWD) {
      if (!PKR && q == 1) {
#pragma unroll
        for (int e = 0; e < E; ++e) R[e] = at(hold, L + lq + TL * e);
      }
      const double nrm = ph.norm2;
#pragma unroll
      for (int e = 0; e < E; ++e) {
        const int kx = kx_of_row(lq + TL * e, nx);
        double2 v = R[e];
        if (a.mode & KP_RAW) {
          // plain unnormalised forward transform (initial conditions)
        } else if ((kx == 0 && kq == 0) || !colv) {
          // (0,0) = 0; a slot past the last column holds no mode, and its ky
          // may lie past the row's end (ky >= ld), where the per-mode filter
          // table must not be read
          v = make_double2(0.0, 0.0);
        } else {
          v = cscale(nrm, v);                                  // solver.c:1005
          v = cscale(dealias_factor(ph, kx, kq), v);           // solver.c:1019
          if (ph.force_on && kx == 0 && kq == ph.force_k)      // solver.c:1022-1027
            v.x += ph.force_val;
        }
        R[e] = v;
      }
    }
    const bool conj_fix = (a.mode & KP_CONJFIX) && kq == 0;
    // one load from a selected address (no branch around it, so that the
    // loads of a thread's rows are all in flight before the first use)
    auto state_at = [&](int it) -> double2 {
      if constexpr (PF != 0) return wpre[it];
      const int i = lq + TL * it;
      const int kx = kx_of_row(i, nx);
      const bool cj = conj_fix && kx < 0;
      const double2* p = cj ? a.conj_src + (cj ? -kx : 0) : a.w_in + kidx(i, kq, ld, nx);
      const double2 v = *p;
      return make_double2(v.x, cj ? -v.y : v.y);
    };
    auto acc_at = [&](int it) -> double2 {
      if constexpr (PF != 0) return apre[it];
      return at(a.acc, index_of(it));
    };
    auto lam_at = [&](int it) -> double {
      if constexpr (PF != 0) return lpre[it];
      return at(a.lam, index_of(it));
    };
    switch (a.epi) {
      case EPI_FROM_STATE:
#pragma unroll
        for (int it = 0; it < E; ++it) {
          R[it] = colv ? at(a.w_in, index_of(it)) : make_double2(0.0, 0.0);
          if ((it & 3) == 3) asm volatile("" ::: "memory");
        }
        break;
      case EPI_FROM_INPUT:
#pragma unroll
        for (int it = 0; it < E; ++it) {
          R[it] = colv ? at(a.sin, index_of(it)) : make_double2(0.0, 0.0);
          if ((it & 3) == 3) asm volatile("" ::: "memory");
        }
        break;
      case EPI_RHS:
        if (writer && colv) {
#pragma unroll
          for (int it = 0; it < E; ++it) {
            at(a.kout, index_of(it)) = R[it];
            const int kx = kx_of_row(lq + TL * it, nx);
            if (conj_fix && kx < 0) at(a.w_out, index_of(it)) = cconj(at(a.conj_src, -kx));
          }
        }
        break;
      case EPI_RK1:
        if (colv) {
#pragma unroll
          for (int it = 0; it < E; ++it) {
            const double2 k = R[it];
            if (writer) {
              at(a.acc, index_of(it)) = cscale(RK4_B1, k);
              if (a.hist_store) at(a.hist_store, index_of(it)) = k;
            }
            R[it] = cadd(state_at(it), cscale(a.coef, k));     // solver.c:690
            if ((it & 3) == 3) asm volatile("" ::: "memory");
          }
        }
        break;
      case EPI_RK2:
      case EPI_RK3:
        if (colv) {
          const double b = (a.epi == EPI_RK2) ? RK4_B2 : RK4_B3;
#pragma unroll
          for (int it = 0; it < E; ++it) {
            const double2 k = R[it];
            if (writer) at(a.acc, index_of(it)) = cadd(acc_at(it), cscale(b, k));
            R[it] = cadd(state_at(it), cscale(a.coef, k));     // solver.c:729,768
            if ((it & 3) == 3) asm volatile("" ::: "memory");
          }
        }
        break;
      case EPI_RK4CN:
        if (colv) {
#pragma unroll
          for (int it = 0; it < E; ++it) {
            const int kx = kx_of_row(lq + TL * it, nx);
            // every mode computes; the (0,0) mode then takes 0 (solver.c:860)
            const double2 sum = cadd(acc_at(it), cscale(RK4_B4, R[it]));
            const double D = a.dt * lam_at(it);
            const double q = rcp_nr(2.0 + D);
            double2 nw = cadd(cscale((2.0 - D) * q, state_at(it)), cscale(2.0 * a.dt * q, sum));  // :853
            if (kx == 0 && kq == 0) nw = make_double2(0.0, 0.0);
            if (writer) at(a.w_out, index_of(it)) = nw;
            R[it] = nw;
            if ((it & 3) == 3) asm volatile("" ::: "memory");
          }
        }
        break;
      case EPI_AB4:
        if (colv) {
#pragma unroll
          for (int it = 0; it < E; ++it) {
            const int kx = kx_of_row(lq + TL * it, nx);
            const size_t idx = index_of(it);
            const double2 k = R[it];
            // every mode computes; the (0,0) mode then takes 0 (solver.c:337)
            const double2 sum = cadd(cadd(cadd(cscale(AB4_1, k), cscale(AB4_2, at(a.h_new, idx))),
                                          cscale(AB4_3, at(a.h_mid, idx))), cscale(AB4_4, at(a.h_old, idx)));
            const double D = a.dt * at(a.lam, idx);
            const double q = rcp_nr(2.0 + D);
            double2 nw = cadd(cscale((2.0 - D) * q, state_at(it)), cscale(2.0 * a.dt * q, sum));  // :333
            if (kx == 0 && kq == 0) nw = make_double2(0.0, 0.0);
            if (writer) {
              at(a.h_store, idx) = k;   // newest history entry (solver.c:346-348)
              at(a.w_out, idx) = nw;
            }
            R[it] = nw;
            if ((it & 3) == 3) asm volatile("" ::: "memory");
          }
        }
        break;
      default: break;
    }
    if (!colv) {
#pragma unroll
      for (int it = 0; it < E; ++it) R[it] = make_double2(0.0, 0.0);
    }
    // ky=0 column of the next stage input, for the next ForceConjugacy
    if ((a.mode & KP_INV) && writer && kq == 0 && a.conj_dst) {
#pragma unroll
      for (int it = 0; it < E; ++it) at(a.conj_dst, lq + TL * it) = R[it];
    }
    if (!PKR && pair && (a.mode & KP_INV)) {
#pragma unroll
      for (int it = 0; it < E; ++it) at(hold, q * L + lq + TL * it) = R[it];
    }
  }

  // blow-up bookkeeping for the reference's per-step SystemCheck
  // (solver.c:2301-2326): the stage-1 row pass left max|w| of the state that
  // this step started from in stat[0]; record it and reset the accumulator.
  if ((a.mode & KP_STAT) && wi == 0 && tid == 0) {
    unsigned long long* st = a.stat;
    const unsigned long long m = st[STAT_WMAX_CUR];
    st[STAT_WMAX_LAST] = m;
    // a NaN maximum also counts (the reference's >= test would let it pass)
    if (!(__longlong_as_double((long long)m) < MAX_VORT_LIM) && st[STAT_FIRST_BAD] == ~0ull)
      st[STAT_FIRST_BAD] = st[STAT_ITER];
    st[STAT_WMAX_CUR] = 0ull;
  }
  if ((a.mode & KP_ITER) && wi == 0 && tid == 0) a.stat[STAT_ITER] += 1ull;

  if (!(a.mode & KP_INV)) return;

  // ---------------- INV: x-inverse FFTs of the derived fields ---------------
  // f = 0: psi_hat = S/k^2, 1: -i kx psi_hat, 2: i kx S, 3: S (f is uniform:
  // only the first two fields compute 1/k^2)
  auto derive = [](int f, double2 s, int kx, int kyv) -> double2 {
    if (f >= 2) return f == 2 ? make_double2(-kx * s.y, kx * s.x) : s;
    const int k2i = kx * kx + kyv * kyv;
    const double inv = k2i ? rcp_nr((double)k2i) : 0.0;
    if (f == 0) return cscale(inv, s);
    return make_double2(kx * inv * s.y, -(kx * inv) * s.x);
  };
#pragma unroll 1
  for (int fi = 0; fi < nfs; ++fi) {
    const int f = split * nfs + fi;
    const int l_o = opaque(l), ky_o = opaque(ky);
    double2 F[E];
    if (pblk) {
      if (pair) {
        // Y = P + i Q in F, V = conj P + i conj Q to the line buffer
#pragma unroll
        for (int e = 0; e < E; ++e) {
          const int x = l_o + TL * e;
          const int kx = kx_of_row(x, nx);
          // S_0 and S_H: after the column loop RH holds S_0 and R holds S_H (PK = 2)
          const double2 s0 = PKR ? RH[PKR ? e : 0] : at(hold, x);
          const double2 sh = PKR ? R[e] : at(hold, L + x);
          double2 P, Q;
          if (f == 0) {
            const double2 ps = derive(0, sh, kx, H);
            P = make_double2(ps.y, -ps.x);
            Q = make_double2(sh.y, -sh.x);
          } else {
            P = derive(f, s0, kx, 0);
            Q = derive(f, sh, kx, H);
          }
          F[e] = make_double2(P.x - Q.y, P.y + Q.x);
          buf[padx_s<PS>(x)] = make_double2(P.x + Q.y, Q.x - P.y);
        }
      }
      xsync<BLOCK>();
      if (pair) {
#pragma unroll
        for (int e = 0; e < E; ++e) {
          const double2 vm = buf[padx_s<PS>((L - (l_o + TL * e)) & (L - 1))];
          F[e] = make_double2(0.5 * (F[e].x + vm.x), 0.5 * (F[e].y + vm.y));
        }
      }
      xsync<BLOCK>();
    }
    if (!pair) {
#pragma unroll
      for (int e = 0; e < E; ++e) F[e] = derive(f, R[e], kx_of_row(l_o + TL * e, nx), ky_o);
    }
    fft_reg<L, TL, +1, BLOCK, PS>(F, buf, tw, l_o);
    double2* __restrict__ gout = a.g[f];
    if (pair) {
      double* __restrict__ g4 = reinterpret_cast<double*>(a.g[3]);
#pragma unroll
      for (int e = 0; e < E; ++e) {
        const size_t r0 = gidx(l_o + TL * e, 0, ld), rh = gidx(l_o + TL * e, H, ld);
        if (f == 0) {
          at(gout, r0) = make_double2(0.0, 0.0);
          at(gout, rh) = make_double2(0.0, F[e].x);
          at(g4, 2 * rh + 1) = F[e].y;
        } else if (f == 3) {
          at(gout, r0) = make_double2(F[e].x, 0.0);
          at(g4, 2 * rh) = F[e].y;
        } else {
          at(gout, r0) = make_double2(F[e].x, 0.0);
          at(gout, rh) = make_double2(F[e].y, 0.0);
        }
      }
    } else if (colv) {
#pragma unroll
      for (int e = 0; e < E; ++e) at(gout, gidx(l_o + TL * e, ky_o, ld)) = F[e];
    }
  }
}

// W: minimum waves per SIMD the register allocation must allow (1 = no limit)
template <int L, int C, int T, int KS, int PS, int OFF, int PK, int W, int PF>
__global__ __launch_bounds__(T, W) void kpass_kernel(KPassArgs a) {
  kpass_body<L, C, T, KS, PS, OFF, PK, PF>(a, xcd_remap((int)blockIdx.x, (int)gridDim.x));
}

// All members of an ensemble size class in one launch: work items
// [m * per_member, (m+1) * per_member) belong to member m, whose arguments
// are tab[m] (uniform per workgroup, read with scalar loads).
template <int L, int C, int T, int KS, int PS, int OFF, int PK, int W, int PF>
__global__ __launch_bounds__(T, W) void kpass_batched(const KPassArgs* __restrict__ tab, int per_member) {
  const int b = xcd_remap((int)blockIdx.x, (int)gridDim.x);
  const int m = b / per_member;
  kpass_body<L, C, T, KS, PS, OFF, PK, PF>(tab[m], b - m * per_member);
}

// LDS padding of the row passes' line buffers: one 16-B slot per 2^RPS
#ifndef NS2D_RPASS_PS
#define NS2D_RPASS_PS 3
#endif
constexpr int RPS = NS2D_RPASS_PS;

// ---------------------------------------------------------------------------
// rpass: row pass in (x, ky) space, one row pair (x0, x0+1) per TLR threads
//
// Thread l of a pair owns y (or ky) = l + TLR*e, e < E = NY/TLR.  For each row
// the full-length spectrum of (A + iB) is assembled from the stored half
// spectra of two real fields (ky >= 0 read directly, ky < 0 as conjugates;
// ky = 0 and NY/2 keep only their real parts), transformed in registers and
// multiplied pointwise.  The two rows' products are packed as n(x0) + i n(x0+1)
// and transformed forward once; N(x0,k) = (W[k] + conj W[NY-k])/2 and
// N(x0+1,k) = (W[k] - conj W[NY-k])/(2i).
// ---------------------------------------------------------------------------
// PFR = 1: the four G fields of both rows are loaded at the start (2 x 4 x E
// values per thread, for shapes with few points per thread), so that the
// loads of the second row and of each row's second field pair overlap the
// first transforms.  PFR = 2 (single row): TLR threads own one row x0 = blk
// instead of a pair; the product n(x0) is transformed forward on its own
// (imaginary part zero) and N(x0,k) = W[k] with the imaginary parts of k = 0
// and NY/2 set to zero, as a real-input transform gives them.  Twice the
// workgroups and three transforms per workgroup instead of five, for one
// more forward transform per row pair.  PFR = 3: a single row whose four G
// fields are all loaded before its first transform.  PFR = 4 and 5: PFR = 3
// and 2 with the real product's forward transform done at half length
// (rpass_forward_half).  PFR + 8: the same shape with every twiddle power of
// its transforms preloaded (TwFull) instead of formed by products.
template <int NY, int TLR, bool BLOCK, typename TWH>
__device__ __forceinline__ void rpass_forward_half(double2 (&w)[NY / TLR], double2* buf, const RPassArgs& a,
                                                   int x0, int l, const TWH& twh);

// lane label of lane l in the half-length line of rpass_forward_half
template <int TLR>
__device__ __forceinline__ int half_lane(int l) { return (l >> 1) + ((l & 1) ? TLR / 2 : 0); }

// tw: the row transforms' twiddle set, twh: the half-length line's.  CFL:
// the instantiation that can reduce the CFL rate (mode RP_CFL); the step's
// kernels are built without that code.
template <int NY, int TLR, int PAIRS, int PFR, bool CFL, typename TW, typename TWH>
__device__ __forceinline__ void rpass_body_tw(const RPassArgs& a, const int blk, const TW& tw, const TWH& twh) {
  constexpr int T = TLR * PAIRS;
  constexpr int H = NY / 2;
  constexpr int E = NY / TLR;
  constexpr int LS = padx_s<RPS>(NY) + 1;
  constexpr bool BLOCK = TLR > 64;
  constexpr bool SR = PFR >= 2;
  constexpr bool UPF = PFR == 3 || PFR == 4;   // single row, G fields loaded up front
  constexpr bool HLF = PFR >= 4;               // half-length forward transform
  static_assert(PFR >= 0 && PFR <= 5, "rpass_body_tw takes the load mode alone");
  static_assert(E >= 2 && (BLOCK ? PAIRS == 1 : 64 % TLR == 0), "bad rpass shape");
  static_assert(!SR || PAIRS == 1, "single-row shapes own one row per workgroup");
  __shared__ double2 lds[PAIRS * LS];
  __shared__ double red[(T + 63) / 64];
  __shared__ double cred[CFL ? (T + 63) / 64 : 1];

  const int tid = threadIdx.x;
  const int q = tid / TLR, l = tid % TLR;
  const int x0 = SR ? blk : 2 * (blk * PAIRS + q);
  const bool valid = x0 < a.nx;   // uniform per block when BLOCK
  const int ld = a.ld;
  double2* buf = lds + q * LS;

  // full-length spectrum of A + iB for row x (which: 0 -> u + iv, 1 -> w_x + i w_y,
  // 2 -> w(x) + i w(x+1), 3 -> w(x))
  // Element e of lane lo holds ky = lo + TLR*e: elements below E/2 lie in
  // [0, H), those above E/2 in (H, NY), and only lane 0's elements 0 and E/2
  // are the real-valued ky = 0 and ky = H, so the mirror and edge tests are
  // fixed per element once the loop is unrolled.
  auto build = [&](double2* z, int which, int x, int lo) {
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int ky = lo + TLR * e;
      const bool edge = (e == 0 || e == E / 2) && lo == 0;
      const bool mirror = e > E / 2 || (e == E / 2 && lo != 0);
      const int kk = mirror ? NY - ky : ky;
      double2 A, B;
      const double fk = (double)kk;
      if (which == 0) {         // u = i ky G1, v = G2
        const double2 g1 = at(a.g[0], gidx(x, kk, ld));
        A = make_double2(-fk * g1.y, fk * g1.x);
        B = at(a.g[1], gidx(x, kk, ld));
      } else if (which == 1) {  // w_x = G3, w_y = i ky G4
        A = at(a.g[2], gidx(x, kk, ld));
        const double2 g4 = at(a.g[3], gidx(x, kk, ld));
        B = make_double2(-fk * g4.y, fk * g4.x);
      } else if (which == 2) {  // w of rows x and x+1 = G4
        A = at(a.g[3], gidx(x, kk, ld));
        B = at(a.g[3], gidx(x + 1, kk, ld));
      } else {                  // w of row x alone
        A = at(a.g[3], gidx(x, kk, ld));
        B = make_double2(0.0, 0.0);
      }
      if (edge) { A.y = 0.0; B.y = 0.0; }
      z[e] = mirror ? make_double2(A.x + B.y, B.x - A.y) : make_double2(A.x - B.y, A.y + B.x);
      if ((e & 3) == 3) asm volatile("" ::: "memory");
    }
  };

  // RP_CFL: this thread's largest pi (|u|/dx + |v|/dy) over its points of
  // u + iv (z1 after its transform); a NaN propagates
  double cm = 0.0;
  auto cfl_acc = [&](const double2* z) {
    if (CFL && (a.mode & RP_CFL)) {
#pragma unroll
      for (int e = 0; e < E; ++e) {
        const double uu = z[e].x * a.cfl_s, vv = z[e].y * a.cfl_s;
        const double c = M_PI * (fabs(uu) / a.cfl_dx + fabs(vv) / a.cfl_dy);
        cm = (c > cm || c != c) ? c : cm;
      }
    }
  };

  double n0[E];
  double2 w[E];
  if constexpr (PFR == 1) {
    if ((a.mode & RP_NL) && (BLOCK || valid)) {
      double2 pre[2][4][E];
#pragma unroll
      for (int r = 0; r < 2; ++r)
#pragma unroll
        for (int f = 0; f < 4; ++f)
#pragma unroll
          for (int e = 0; e < E; ++e) {
            const int ky = l + TLR * e;
            const int kk = ky > H ? NY - ky : ky;
            pre[r][f][e] = at(a.g[f], gidx(x0 + r, kk, ld));
          }
      // the same assembly as build() from the loaded values
      auto build_pre = [&](double2* z, int which, const double2 (&g)[4][E]) {
#pragma unroll
        for (int e = 0; e < E; ++e) {
          const int ky = l + TLR * e;
          const bool mirror = ky > H;
          const int kk = mirror ? NY - ky : ky;
          const double fk = (double)kk;
          double2 A, B;
          if (which == 0) {
            A = make_double2(-fk * g[0][e].y, fk * g[0][e].x);
            B = g[1][e];
          } else {
            A = g[2][e];
            B = make_double2(-fk * g[3][e].y, fk * g[3][e].x);
          }
          if (kk == 0 || kk == H) { A.y = 0.0; B.y = 0.0; }
          z[e] = mirror ? make_double2(A.x + B.y, B.x - A.y) : make_double2(A.x - B.y, A.y + B.x);
        }
      };
#pragma unroll
      for (int r = 0; r < 2; ++r) {
        double2 z1[E], z2[E];
        build_pre(z1, 0, pre[r]);
        fft_reg<NY, TLR, +1, BLOCK, RPS>(z1, buf, tw, opaque(l));
        cfl_acc(z1);
        build_pre(z2, 1, pre[r]);
        fft_reg<NY, TLR, +1, BLOCK, RPS>(z2, buf, tw, opaque(l));
#pragma unroll
        for (int e = 0; e < E; ++e) {
          const double nl = z1[e].x * z2[e].x + z1[e].y * z2[e].y;   // u*w_x + v*w_y
          if (r == 0) n0[e] = nl;
          else w[e] = make_double2(n0[e], nl);
        }
      }
    }
  } else if (SR && (a.mode & RP_NL) && (BLOCK || valid)) {
    double2 z1[E], z2[E];
    build(z1, 0, x0, opaque(l));
    if constexpr (UPF) build(z2, 1, x0, opaque(l));   // both field pairs loaded up front
    fft_reg<NY, TLR, +1, BLOCK, RPS>(z1, buf, tw, opaque(l));
    cfl_acc(z1);
    if constexpr (!UPF) build(z2, 1, x0, opaque(l));
    fft_reg<NY, TLR, +1, BLOCK, RPS>(z2, buf, tw, opaque(l));
#pragma unroll
    for (int e = 0; e < E; ++e) w[e] = make_double2(z1[e].x * z2[e].x + z1[e].y * z2[e].y, 0.0);
  } else if ((a.mode & RP_NL) && (BLOCK || valid)) {
#pragma unroll 1
    for (int r = 0; r < 2; ++r) {
      const int x = opaque(x0 + r);
      double2 z1[E], z2[E];
      build(z1, 0, x, opaque(l));
      fft_reg<NY, TLR, +1, BLOCK, RPS>(z1, buf, tw, opaque(l));
      cfl_acc(z1);
      build(z2, 1, x, opaque(l));
      fft_reg<NY, TLR, +1, BLOCK, RPS>(z2, buf, tw, opaque(l));
#pragma unroll
      for (int e = 0; e < E; ++e) {
        const double nl = z1[e].x * z2[e].x + z1[e].y * z2[e].y;   // u*w_x + v*w_y
        if (r == 0) n0[e] = nl;
        else w[e] = make_double2(n0[e], nl);
      }
    }
  }

  if (CFL && (a.mode & RP_CFL)) {
    // wave then block max, as for max|w| below (its own LDS words)
    unsigned long long mb = (unsigned long long)__double_as_longlong(cm != cm ? __longlong_as_double(0x7ff8000000000000ll) : cm);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      const unsigned long long o = __shfl_xor(mb, off, 64);
      mb = o > mb ? o : mb;
    }
    if ((tid & 63) == 0) cred[tid >> 6] = __longlong_as_double((long long)mb);
    __syncthreads();
    if (tid == 0) {
      unsigned long long bm = 0;
      for (int i = 0; i < (T + 63) / 64; ++i) {
        const unsigned long long v = (unsigned long long)__double_as_longlong(cred[i]);
        bm = v > bm ? v : bm;
      }
      atomicMax(a.cmax, bm);
    }
  }

  if (a.mode & RP_W) {
    double m = 0.0;
    if (BLOCK || valid) {
      double2 z[E];
      build(z, SR ? 3 : 2, x0, opaque(l));
      fft_reg<NY, TLR, +1, BLOCK, RPS>(z, buf, tw, opaque(l));
#pragma unroll
      for (int e = 0; e < E; ++e) {
        const int y = l + TLR * e;
        const double w0 = z[e].x * a.wscale, w1 = SR ? 0.0 : z[e].y * a.wscale;
        if (a.wreal) {
          at(a.wreal, (size_t)x0 * NY + y) = w0;
          if (!SR) at(a.wreal, (size_t)(x0 + 1) * NY + y) = w1;
        }
        const double aw = fmax(fabs(w0), fabs(w1));
        m = (aw > m || aw != aw) ? aw : m;
        if (w0 != w0 || w1 != w1) m = __longlong_as_double(0x7ff8000000000000ll);
      }
    }
    // wave then block max; a NaN propagates (its bit pattern compares largest)
    unsigned long long mb = (unsigned long long)__double_as_longlong(m);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      const unsigned long long o = __shfl_xor(mb, off, 64);
      mb = o > mb ? o : mb;
    }
    if ((tid & 63) == 0) red[tid >> 6] = __longlong_as_double((long long)mb);
    __syncthreads();
    if (tid == 0) {
      unsigned long long bm = 0;
      for (int i = 0; i < (T + 63) / 64; ++i) {
        const unsigned long long v = (unsigned long long)__double_as_longlong(red[i]);
        bm = v > bm ? v : bm;
      }
      atomicMax(a.wmax, bm);
    }
  }

  if (SR && (a.mode & RP_NL) && (BLOCK || valid)) {
    if constexpr (HLF) {
      rpass_forward_half<NY, TLR, BLOCK>(w, buf, a, x0, l, twh);
    } else {
      fft_reg<NY, TLR, -1, BLOCK, RPS>(w, buf, tw, opaque(l));
      // k = l + TLR*e <= H: every element below E/2, lane 0's element E/2 (k = H)
#pragma unroll
      for (int e = 0; e <= E / 2; ++e) {
        const int k = l + TLR * e;
        if (e < E / 2) at(a.nout, nidx(x0, k, ld)) = (e == 0 && l == 0) ? make_double2(w[e].x, 0.0) : w[e];
        else if (l == 0) at(a.nout, nidx(x0, k, ld)) = make_double2(w[e].x, 0.0);
      }
    }
  } else if ((a.mode & RP_NL) && (BLOCK || valid)) {
    fft_reg<NY, TLR, -1, BLOCK, RPS>(w, buf, tw, opaque(l));
    // W[NY-k] for the thread's own k: exchange through the line buffer
#pragma unroll
    for (int e = 0; e < E; ++e) buf[padx_s<RPS>(l + TLR * e)] = w[e];
    xsync<BLOCK>();
    // k = l + TLR*e <= H: every element below E/2, lane 0's element E/2
#pragma unroll
    for (int e = 0; e <= E / 2; ++e) {
      const int k = l + TLR * e;
      if (e < E / 2 || l == 0) {
        const double2 p = w[e];
        const double2 qm = cconj(buf[padx_s<RPS>((NY - k) & (NY - 1))]);
        at(a.nout, nidx(x0, k, ld)) = make_double2(0.5 * (p.x + qm.x), 0.5 * (p.y + qm.y));
        at(a.nout, nidx(x0 + 1, k, ld)) = make_double2(0.5 * (p.y - qm.y), -0.5 * (p.x - qm.x));
      }
    }
  }
}

template <int NY, int TLR, int PAIRS, int PFR, bool CFL = false>
__device__ __forceinline__ void rpass_body(const RPassArgs& a, const int blk) {
  constexpr int PM = PFR & 7;
  const int l = (int)threadIdx.x % TLR;
  if constexpr ((PFR & 8) != 0) {
    static_assert(PAIRS == 1 && TLR <= 64, "preloaded twiddle powers are for one-wave rows");
    TwFull<NY, TLR> tw;
    load_twiddles<NY, TLR>(tw, a.tw, l);
    if constexpr (PM >= 4) {
      TwFull<NY / 2, TLR> twh;
      load_twiddles<NY / 2, TLR>(twh, a.tw, half_lane<TLR>(l));
      rpass_body_tw<NY, TLR, PAIRS, PM, CFL>(a, blk, tw, twh);
    } else {
      rpass_body_tw<NY, TLR, PAIRS, PM, CFL>(a, blk, tw, a.tw);
    }
  } else {
    NS2D_TWIDDLES(NY, TLR, tw, a.tw, l)
    if constexpr (PM >= 4) {
      NS2D_TWIDDLES(NY / 2, TLR, twh, a.tw, half_lane<TLR>(l))
      rpass_body_tw<NY, TLR, PAIRS, PM, CFL>(a, blk, tw, twh);
    } else {
      rpass_body_tw<NY, TLR, PAIRS, PM, CFL>(a, blk, tw, a.tw);
    }
  }
}

// The value of lane l ^ 1 (the other lane of an even/odd pair; both lanes of a
// pair always lie in one wave): DPP quad permutation [1, 0, 3, 2] of each
// 32-bit half.
__device__ __forceinline__ double swap_lane_pair(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_mov_dpp((int)(b & 0xffffffffll), 0xB1, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), 0xB1, 0xF, 0xF, false);
  return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}

// Forward transform of one real row, n(y) = w[e].x at y = l + TLR*e, as a
// complex transform of half the length.  With z(m) = n(2m) + i n(2m+1),
// m < M = NY/2, and Z its M-point transform, the even and odd samples'
// transforms are E(k) = (Z(k) + conj Z(M-k))/2 and O(k) = (Z(k) - conj Z(M-k))/(2i),
// and N(k) = E(k) + exp(-2 pi i k/NY) O(k) for k <= M (k = M: E(0) - O(0)).
// Lanes 2j and 2j+1 hold n(2j + TLR*e) and n(2j+1 + TLR*e); each passes the
// other half of its values across the pair, so that lane 2j holds
// z(j + TLR*h) (its even e = 2h) and lane 2j+1 z(j + TLR/2 + TLR*h) (odd
// e = 2h+1): the stride-TLR distribution of the M-point line over the labels
// lh = j and j + TLR/2.  The transform leaves Z(lh + TLR*h) in the same
// distribution; Z(M-k) comes through the line buffer.  Writes N(x0, k) for
// k = 0..M, each once.
template <int NY, int TLR, bool BLOCK, typename TWH>
__device__ __forceinline__ void rpass_forward_half(double2 (&w)[NY / TLR], double2* buf, const RPassArgs& a,
                                                   int x0, int l, const TWH& twh) {
  constexpr int E = NY / TLR, EH = E / 2, M = NY / 2;
  static_assert(E >= 4 && TLR % 2 == 0, "half-length forward transform needs E >= 4 and an even TLR");
  const int ld = a.ld;
  const bool odd = (l & 1) != 0;
  const int lh = half_lane<TLR>(l);
  double2 z[EH];
#pragma unroll
  for (int h = 0; h < EH; ++h) {
    const double give = odd ? w[2 * h].x : w[2 * h + 1].x;
    const double take = swap_lane_pair(give);
    z[h] = odd ? make_double2(take, w[2 * h + 1].x) : make_double2(w[2 * h].x, take);
  }
  // exp(-2 pi i k/NY) of the outputs k = lh + TLR*h, loaded ahead of the transform
  double2 pw[EH];
#pragma unroll
  for (int h = 0; h < EH; ++h) pw[h] = at(a.tw, (lh + TLR * h) * (NS2D_TW_LEN / NY));
  fft_reg<M, TLR, -1, BLOCK, RPS>(z, buf, twh, opaque(lh));
  // (the transform's last exchange ends with a barrier: the buffer is free)
#pragma unroll
  for (int h = 0; h < EH; ++h) buf[padx_s<RPS>(lh + TLR * h)] = z[h];
  xsync<BLOCK>();
#pragma unroll
  for (int h = 0; h < EH; ++h) {
    const int k = lh + TLR * h;
    const double2 A = z[h];
    const double2 Bm = buf[padx_s<RPS>((M - k) & (M - 1))];   // Z(M-k); conj taken below
    const double2 Ev = make_double2(0.5 * (A.x + Bm.x), 0.5 * (A.y - Bm.y));
    const double2 Dv = make_double2(0.5 * (A.x - Bm.x), 0.5 * (A.y + Bm.y));
    const double2 Ov = make_double2(Dv.y, -Dv.x);
    if (k == 0) {
      // real-input transform: N(0) and N(M) are real
      at(a.nout, nidx(x0, 0, ld)) = make_double2(Ev.x + Ov.x, 0.0);
      at(a.nout, nidx(x0, M, ld)) = make_double2(Ev.x - Ov.x, 0.0);
    } else {
      at(a.nout, nidx(x0, k, ld)) = cadd(Ev, cmul(pw[h], Ov));
    }
  }
}

// W: minimum waves per SIMD the register allocation must allow (1 = no limit)
// Consecutive rows (row pairs) go to one XCD, whose L2 then serves the
// x-tiles of the mixed-space fields that neighbouring rows share.
template <int NY, int TLR, int PAIRS, int W, int PFR>
__global__ __launch_bounds__(TLR * PAIRS, W) void rpass_kernel(RPassArgs a) {
  rpass_body<NY, TLR, PAIRS, PFR>(a, xcd_remap((int)blockIdx.x, (int)gridDim.x));
}

// the same shape with the CFL-rate reduction compiled in (ns2d_cfl_rate)
template <int NY, int TLR, int PAIRS, int W, int PFR>
__global__ __launch_bounds__(TLR * PAIRS, W) void rpass_cfl_kernel(RPassArgs a) {
  rpass_body<NY, TLR, PAIRS, PFR, true>(a, xcd_remap((int)blockIdx.x, (int)gridDim.x));
}

template <int NY, int TLR, int PAIRS, int W, int PFR>
__global__ __launch_bounds__(TLR * PAIRS, W) void rpass_batched(const RPassArgs* __restrict__ tab, int per_member) {
  const int b = xcd_remap((int)blockIdx.x, (int)gridDim.x);
  const int m = b / per_member;
  rpass_body<NY, TLR, PAIRS, PFR>(tab[m], b - m * per_member);
}

// ---------------------------------------------------------------------------
// Discrete adjoint (vector-Jacobian product) of the RK4CN step.
//
// With K = N(S) = M .* r2c(u w_x + v w_y), u = c2r(U S), v = c2r(V S),
// w_x = c2r(X S), w_y = c2r(Y S) (U = i ky/|k|^2, V = -i kx/|k|^2, X = i kx,
// Y = i ky, M = dealias * (1/(NxNy))^2, zero at k = 0) and gradients packed as
// g = dJ/dRe + i dJ/dIm, the transposes are
//   r2c^T(l) = c2r(Wh .* l),   Wh = 1/2 on interior ky columns, 1 on ky = 0, Ny/2
//   c2r^T(f) = W2 .* r2c(f),   W2 = 2 on interior ky columns, 1 on ky = 0, Ny/2
// so, with l_n = r2c^T(M .* l_K):
//   N'(S)^T l_K = conj(U) W2 r2c(l_n w_x) + conj(X) W2 r2c(l_n u)
//               + conj(V) W2 r2c(l_n w_y) + conj(Y) W2 r2c(l_n v).
// rpass_adj forms the four products and their y transforms (the ky parts of
// the multipliers and W2 applied); kpass_adj does the four x transforms, the
// kx parts, the RK4CN adjoint recursion and r2c^T of the next stage.
// SR = 1: TLR threads own one row (l_n of that row transformed alone, with a
// zero imaginary part) instead of a row pair: twice the workgroups, five
// transforms per workgroup instead of nine.
// ---------------------------------------------------------------------------
template <int NY, int TLR, int PAIRS, int SR>
__device__ __forceinline__ void rpass_adj_body(const RAdjArgs& a, const int blk) {
  constexpr int H = NY / 2;
  constexpr int E = NY / TLR;
  constexpr int LS = padx_s<RPS>(NY) + 1;
  constexpr bool BLOCK = TLR > 64;
  static_assert(E >= 2 && (BLOCK ? PAIRS == 1 : 64 % TLR == 0), "bad rpass_adj shape");
  static_assert(!SR || PAIRS == 1, "single-row shapes own one row per workgroup");
  __shared__ double2 lds[PAIRS * LS];

  const int tid = threadIdx.x;
  const int q = tid / TLR, l = tid % TLR;
  const int x0 = SR ? blk : 2 * (blk * PAIRS + q);
  const bool valid = x0 < a.nx;
  if (!BLOCK && !valid) return;
  const int ld = a.ld;
  double2* buf = lds + q * LS;
  NS2D_TWIDDLES(NY, TLR, tw, a.tw, l)

  // full-length spectrum of A + iB from the half spectra of two real fields,
  // A from row xa of fa, B from row xb of fb (fb == nullptr only where hasb
  // is 0: B = 0)
  // (element e of lane lo holds ky = lo + TLR*e; the mirror and edge tests
  // are fixed per element as in the forward row pass)
  auto build = [&](double2* z, const double2* fa, int xa, const double2* fb, int xb, int mula, int mulb, int lo,
                   int hasb = 1) {
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int ky = lo + TLR * e;
      const bool edge = (e == 0 || e == E / 2) && lo == 0;
      const bool mirror = e > E / 2 || (e == E / 2 && lo != 0);
      const int kk = mirror ? NY - ky : ky;
      const double fk = (double)kk;
      double2 A = fa[gidx(xa, kk, ld)], B = hasb ? fb[gidx(xb, kk, ld)] : make_double2(0.0, 0.0);
      if (mula) A = make_double2(-fk * A.y, fk * A.x);   // times i ky
      if (mulb) B = make_double2(-fk * B.y, fk * B.x);
      if (edge) { A.y = 0.0; B.y = 0.0; }
      z[e] = mirror ? make_double2(A.x + B.y, B.x - A.y) : make_double2(A.x - B.y, A.y + B.x);
      if ((e & 3) == 3) asm volatile("" ::: "memory");
    }
  };
  // r2c of the real and imaginary parts of a transformed packed row:
  // P[k] = (Z[k] + conj Z[-k])/2, Q[k] = (Z[k] - conj Z[-k])/(2i); writes
  // W2 (mul ? -i k : 1) P to pout and W2 (mulq ? -i k : 1) Q to qout
  auto split_store = [&](double2* z, double2* pout, double2* qout, int x, int mulp, int mulq, int lo) {
#pragma unroll
    for (int e = 0; e < E; ++e) buf[padx_s<RPS>(lo + TLR * e)] = z[e];
    xsync<BLOCK>();
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int k = lo + TLR * e;
      if (k <= H) {
        const double2 zp = z[e];
        const double2 zm = cconj(buf[padx_s<RPS>((NY - k) & (NY - 1))]);
        double2 P = make_double2(0.5 * (zp.x + zm.x), 0.5 * (zp.y + zm.y));
        double2 Q = make_double2(0.5 * (zp.y - zm.y), -0.5 * (zp.x - zm.x));
        const double w2 = (k == 0 || k == H) ? 1.0 : 2.0;
        const double fk = (double)k;
        if (mulp) P = make_double2(fk * P.y, -fk * P.x);   // times -i k
        if (mulq) Q = make_double2(fk * Q.y, -fk * Q.x);
        pout[nidx(x, k, ld)] = cscale(w2, P);
        qout[nidx(x, k, ld)] = cscale(w2, Q);
      }
    }
    xsync<BLOCK>();
  };

  // l_n of rows x0 and x0+1 (SR: of row x0) from the x-transformed r2c^T field L
  double2 zl[E];
  if constexpr (SR) build(zl, a.lin, x0, nullptr, x0, 0, 0, opaque(l), 0);
  else build(zl, a.lin, x0, a.lin, x0 + 1, 0, 0, opaque(l));
  fft_reg<NY, TLR, +1, BLOCK, RPS>(zl, buf, tw, opaque(l));
#pragma unroll 1
  for (int r = 0; r < (SR ? 1 : 2); ++r) {
    const int x = opaque(x0 + r);
    double2 z1[E], z2[E];
    build(z1, a.g[0], x, a.g[1], x, 1, 0, opaque(l));   // u + i v
    fft_reg<NY, TLR, +1, BLOCK, RPS>(z1, buf, tw, opaque(l));
    build(z2, a.g[2], x, a.g[3], x, 0, 1, opaque(l));   // w_x + i w_y
    fft_reg<NY, TLR, +1, BLOCK, RPS>(z2, buf, tw, opaque(l));
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const double ln = (r == 0) ? zl[e].x : zl[e].y;
      const double2 u = z1[e], wv = z2[e];
      z1[e] = make_double2(ln * wv.x, ln * u.x);   // l_n w_x + i l_n u
      z2[e] = make_double2(ln * wv.y, ln * u.y);   // l_n w_y + i l_n v
    }
    fft_reg<NY, TLR, -1, BLOCK, RPS>(z1, buf, tw, opaque(l));
    split_store(z1, a.bout[0], a.bout[1], x, 1, 0, opaque(l));
    fft_reg<NY, TLR, -1, BLOCK, RPS>(z2, buf, tw, opaque(l));
    split_store(z2, a.bout[2], a.bout[3], x, 0, 1, opaque(l));
  }
}

template <int NY, int TLR, int PAIRS, int SR>
__global__ __launch_bounds__(TLR * PAIRS) void rpass_adj_kernel(RAdjArgs a) {
  rpass_adj_body<NY, TLR, PAIRS, SR>(a, xcd_remap((int)blockIdx.x, (int)gridDim.x));
}

template <int L, int C, int T>
__global__ __launch_bounds__(T) void kpass_adj_kernel(KAdjArgs a) {
  constexpr int TL = T / C;
  constexpr int E = L / TL;
  constexpr int LS = line_stride(L);
  constexpr bool BLOCK = T > 64;
  static_assert(T % C == 0 && L % TL == 0 && E >= 2, "bad kpass_adj shape");
  __shared__ double2 lds[C * LS];

  const int tid = threadIdx.x;
  const int g = xcd_remap((int)blockIdx.x, (int)gridDim.x);
  const int ky0 = g * C;
  const int c = tid % C, l = tid / C;
  const int ky = ky0 + c;
  const int nyf = a.nyf, ld = a.ld, nx = L;
  const bool colv = ky < nyf;
  const PhysArgs& ph = a.ph;
  double2* buf = lds + c * LS;
  const double h = a.dt;
  NS2D_TWIDDLES(L, TL, tw, a.tw, l)

  double2 LSacc[E];   // N'(S)^T l_K of the stage just finished
#pragma unroll
  for (int e = 0; e < E; ++e) LSacc[e] = make_double2(0.0, 0.0);
  if (a.mode & KA_FWD) {
#pragma unroll 1
    for (int f = 0; f < 4; ++f) {
      const int l_o = opaque(l), ky_o = opaque(ky);
      double2 F[E];
#pragma unroll
      for (int e = 0; e < E; ++e) {
        F[e] = colv ? a.bin[f][nidx(l_o + TL * e, ky_o, ld)] : make_double2(0.0, 0.0);
        if ((e & 3) == 3) asm volatile("" ::: "memory");
      }
      fft_reg<L, TL, -1, BLOCK>(F, buf, tw, l_o);
#pragma unroll
      for (int e = 0; e < E; ++e) {
        const int kx = kx_of_row(l_o + TL * e, nx);
        const int k2i = kx * kx + ky_o * ky_o;
        const double inv = k2i ? rcp_nr((double)k2i) : 0.0;
        const double2 v = F[e];
        double2 t;
        if (f == 0) t = cscale(inv, v);                                   // conj(U): 1/|k|^2 (-i ky in B1)
        else if (f == 1) t = make_double2(kx * v.y, -kx * v.x);          // conj(X) = -i kx
        else if (f == 2) t = make_double2(-(kx * inv) * v.y, (kx * inv) * v.x);  // conj(V) = i kx/|k|^2
        else t = v;                                                       // conj(Y): -i ky in B4
        LSacc[e] = cadd(LSacc[e], t);
      }
    }
  }

  // RK4CN adjoint recursion (backwards through solver.c:690-853)
  double2 LK[E];
#pragma unroll
  for (int e = 0; e < E; ++e) {
    LK[e] = make_double2(0.0, 0.0);
    if (!colv) continue;
    const int kx = kx_of_row(l + TL * e, nx);
    const size_t idx = kidx(l + TL * e, ky, ld, nx);
    const bool zero = (kx == 0 && ky == 0);
    const double D = h * a.lam[idx];
    const double q = rcp_nr(2.0 + D);
    const double A = (2.0 - D) * q, B = 2.0 * h * q;
    double2 lp = a.lamp_in[idx];   // dJ/dw_{n+1}
    if (zero) lp = make_double2(0.0, 0.0);
    switch (a.epi) {
      case ADJ_START:
        a.lacc[idx] = cscale(A, lp);
        LK[e] = cscale(B * RK4_B1, lp);
        break;
      case ADJ_S4:
        a.lacc[idx] = cadd(a.lacc[idx], LSacc[e]);
        LK[e] = cadd(cscale(B * RK4_B2, lp), cscale(h, LSacc[e]));
        break;
      case ADJ_S3:
        a.lacc[idx] = cadd(a.lacc[idx], LSacc[e]);
        LK[e] = cadd(cscale(B * RK4_B2, lp), cscale(0.5 * h, LSacc[e]));
        break;
      case ADJ_S2:
        a.lacc[idx] = cadd(a.lacc[idx], LSacc[e]);
        LK[e] = cadd(cscale(B * RK4_B1, lp), cscale(0.5 * h, LSacc[e]));
        break;
      case ADJ_S1: {
        // dJ/dw_n; with KA_INV it starts the adjoint of the previous step
        const double2 lw = cadd(a.lacc[idx], LSacc[e]);
        a.lamp_out[idx] = lw;
        if (a.mode & KA_INV) {
          const double Dp = h * a.lam_prev[idx];
          const double qp = rcp_nr(2.0 + Dp);
          const double Ap = (2.0 - Dp) * qp, Bp = 2.0 * h * qp;
          const double2 lq = zero ? make_double2(0.0, 0.0) : lw;
          a.lacc[idx] = cscale(Ap, lq);
          LK[e] = cscale(Bp * RK4_B1, lq);
        }
        break;
      }
      default: break;
    }
    if ((e & 3) == 3) asm volatile("" ::: "memory");
  }
  if (!(a.mode & KA_INV)) return;

  // l_n transform of the next stage: x-inverse of Wh .* M .* l_K
  {
    const int l_o = opaque(l), ky_o = opaque(ky);
    const double wh = (ky_o == 0 || ky_o == ph.ny / 2) ? 1.0 : 0.5;
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int kx = kx_of_row(l_o + TL * e, nx);
      double m = 0.0;
      if (colv && !(kx == 0 && ky_o == 0)) m = wh * ph.norm2 * dealias_factor(ph, kx, ky_o);
      LK[e] = cscale(m, LK[e]);
    }
    fft_reg<L, TL, +1, BLOCK>(LK, buf, tw, l_o);
    if (colv) {
#pragma unroll
      for (int e = 0; e < E; ++e) a.lout[gidx(l_o + TL * e, ky_o, ld)] = LK[e];
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
  __shared__ double2 lds[line_stride(H)];
  const int tid = threadIdx.x, x = blockIdx.x;
  double* ldsr = reinterpret_cast<double*>(lds);
#pragma unroll
  for (int it = 0; it < ITY; ++it) {
    const int y = tid + it * T;
    if (y < NY) ldsr[2 * padx(y >> 1) + (y & 1)] = in[(size_t)x * NY + y];
  }
  __syncthreads();
  fft_lds<H, 1, line_stride(H), T, -1>(lds, tw, tid);
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
        const double2 p = lds[padx(k)];
        const double2 q = cconj(lds[padx(H - k)]);
        const double2 E = make_double2(0.5 * (p.x + q.x), 0.5 * (p.y + q.y));
        const double2 O = make_double2(0.5 * (p.y - q.y), -0.5 * (p.x - q.x));
        X = cadd(E, cmul(tw[k * TWS], O));
      }
      out[nidx(x, k, ld)] = X;
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
  const int b = meas_bin((int)blockIdx.x);
  if (b >= a.nbins) return;
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
      // four ky per round: the loads of a round are issued together, then the
      // modes are summed in ky order (the order of a one-by-one walk)
      for (long ky0 = ylo; ky0 <= yhi; ky0 += 4) {
        double2 wv[4], rv[4];
        double pv[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const long kyl = ky0 + u <= yhi ? ky0 + u : yhi;   // a valid index
          const size_t idx = kidx(i, (int)kyl, a.ld, nx);
          wv[u] = a.w_hat[idx];
          pv[u] = a.lam[idx];
          rv[u] = a.rhs ? a.rhs[idx] : make_double2(0.0, 0.0);
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const long ky = ky0 + u;
          if (ky > yhi) break;
          const double2 w = wv[u];
          const double k2 = (double)(kx2 + ky * ky);
          const double c = (ky == 0 || ky == H) ? 1.0 : 2.0;
          const double a2 = w.x * w.x + w.y * w.y;
          const double pre = pv[u];
          acc[Q_ENRG] += c * a2 / k2;
          acc[Q_ENST] += c * a2;
          acc[Q_PALIN] += c * k2 * a2;
          // |u|^2+|v|^2 with u = i ky w/k^2, v = -i kx w/k^2
          const double uv2 = (double)(ky * ky) * a2 / (k2 * k2) + (double)kx2 * a2 / (k2 * k2);
          acc[Q_ENRG_DISS] += c * pre * uv2;
          acc[Q_ENST_DISS] += c * ((a.quirk_enst_diss && c == 2.0) ? 1.0 : pre) * a2;
          {
            // |i (kx u + ky v)|^2, sys_msr.c:186-193 (zero up to rounding)
            const double2 uu = make_double2(-(double)ky * w.y / k2, (double)ky * w.x / k2);
            const double2 vv = make_double2((double)kx * w.y / k2, -(double)kx * w.x / k2);
            const double dx = (double)kx * uu.x + (double)ky * vv.x, dy = (double)kx * uu.y + (double)ky * vv.y;
            acc[Q_DIV] += c * (dx * dx + dy * dy);
          }
          if (a.rhs) {
            double2 r = rv[u];
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


// kpass variants {L, C, T, KS, PS, OFF, PK, W, PF}: L-point columns, C column
// slots per workgroup, T threads, KS workgroups per column group, LDS padding
// shift PS and line offset OFF, PK = 1 packs the ky = 0 and ky = Ny/2 columns
// into slot 0 (PK = 2: with the pair's second column in registers), W minimum
// waves per SIMD, PF = 1 issues the RK epilogue's loads before the forward
// transform; the first listed for each L is the default.  Rows per thread
// E = L*C/T.  Measured on square grids: the unpacked shape is the default at
// every length (the packed pair's extra exchanges and held columns make slot
// 0 the last workgroup to finish; at 4096, one workgroup per CU, it removes
// the odd column group of the last round, which paid off with the 8-slot
// padding but no longer with the 16-slot one).  Padding one slot per 16
// (PS = 4) with the two columns' lines 8 slots apart (OFF = 8: adjacent lanes
// work on the two columns, whose accesses then fall on opposite halves of
// the banks) beats one slot per 8 with adjacent lines by 6 % at 2048 and 4 %
// at 4096.  Up to 1024 every workgroup runs in the first round and the pass
// takes as long as its busiest CU: the Ny/2+1 columns make an odd column
// count, and with two columns per workgroup the last group lands as a third
// workgroup on two CUs (-9 % at 1024 when it is left out); one column per
// workgroup (KS splits of the inverse fields) spreads the work over four or
// more workgroups per CU, so the odd one costs a quarter instead of a half:
// 1024 one column per 256 threads, 2 splits, 16-slot padding (-7.5 %, and
// -0.3 % per step against 8-slot padding in interleaved timing); 512 one column per
// 128 threads, 4 splits (-5 %); 256 one column per wave, 4 splits (-38 %,
// no workgroup barriers).
#define NS2D_KPASS_VARIANTS(X) \
  X(8, 16, 64, 1, 3, 1, 0, 1, 0) \
  X(8, 16, 64, 1, 3, 1, 1, 1, 0) \
  X(16, 16, 64, 1, 3, 1, 0, 1, 0) \
  X(16, 16, 64, 1, 3, 1, 1, 1, 0) \
  X(32, 16, 64, 1, 3, 1, 0, 1, 0) \
  X(32, 16, 64, 1, 3, 1, 1, 1, 0) \
  X(64, 16, 64, 1, 3, 1, 0, 1, 0) \
  X(64, 16, 64, 1, 3, 1, 1, 1, 0) \
  X(128, 16, 128, 1, 3, 1, 0, 1, 0) \
  X(128, 16, 128, 1, 3, 1, 1, 1, 0) \
  X(128, 8, 64, 2, 3, 1, 1, 1, 0) \
  X(128, 8, 128, 2, 3, 1, 1, 1, 0) \
  X(256, 1, 64, 4, 3, 1, 0, 1, 0) \
  X(256, 1, 64, 4, 3, 1, 0, 1, 1) \
  X(256, 8, 512, 2, 3, 1, 0, 1, 0) \
  X(256, 8, 256, 2, 3, 1, 0, 1, 0) \
  X(256, 8, 256, 2, 3, 1, 0, 1, 1) \
  X(256, 8, 256, 2, 3, 1, 1, 4, 0) \
  X(256, 16, 256, 1, 3, 1, 1, 1, 0) \
  X(256, 8, 128, 2, 3, 1, 1, 1, 0) \
  X(256, 8, 128, 4, 3, 1, 1, 1, 0) \
  X(256, 16, 256, 2, 3, 1, 0, 1, 0) \
  X(256, 32, 512, 2, 3, 1, 0, 1, 0) \
  X(256, 8, 512, 2, 4, 2, 0, 1, 0) \
  X(256, 8, 512, 2, 4, 8, 0, 1, 0) \
  X(256, 8, 256, 2, 4, 2, 0, 1, 0) \
  X(256, 8, 256, 2, 4, 8, 0, 1, 0) \
  X(256, 1, 64, 2, 3, 1, 0, 1, 0) \
  X(256, 2, 128, 4, 3, 1, 0, 1, 0) \
  X(256, 1, 128, 4, 3, 1, 0, 1, 0) \
  X(512, 1, 128, 4, 3, 1, 0, 1, 0) \
  X(512, 1, 128, 4, 3, 1, 0, 1, 1) \
  X(512, 2, 256, 2, 3, 1, 0, 1, 0) \
  X(512, 4, 256, 2, 3, 1, 0, 1, 0) \
  X(512, 4, 256, 2, 3, 1, 0, 1, 1) \
  X(512, 4, 256, 2, 3, 1, 1, 4, 0) \
  X(512, 8, 256, 2, 3, 1, 1, 1, 0) \
  X(512, 8, 256, 1, 3, 1, 1, 1, 0) \
  X(512, 4, 128, 4, 3, 1, 1, 1, 0) \
  X(512, 8, 256, 4, 3, 1, 1, 1, 0) \
  X(512, 8, 512, 2, 3, 1, 1, 1, 0) \
  X(512, 8, 512, 2, 3, 1, 0, 1, 0) \
  X(512, 8, 512, 1, 3, 1, 0, 1, 0) \
  X(512, 16, 512, 2, 3, 1, 0, 1, 0) \
  X(512, 4, 256, 2, 3, 1, 2, 1, 0) \
  X(512, 4, 512, 2, 3, 1, 0, 1, 0) \
  X(512, 2, 256, 2, 3, 8, 0, 1, 0) \
  X(512, 2, 256, 2, 4, 8, 0, 1, 0) \
  X(512, 8, 512, 1, 4, 2, 0, 1, 0) \
  X(512, 8, 512, 1, 4, 8, 0, 1, 0) \
  X(512, 8, 512, 2, 4, 2, 0, 1, 0) \
  X(512, 1, 128, 2, 3, 1, 0, 1, 0) \
  X(512, 1, 64, 4, 3, 1, 0, 1, 0) \
  X(512, 1, 64, 2, 3, 1, 0, 1, 0) \
  X(1024, 1, 256, 2, 4, 1, 0, 1, 0) \
  X(1024, 1, 256, 2, 3, 1, 0, 1, 0) \
  X(1024, 1, 256, 2, 3, 1, 0, 1, 1) \
  X(1024, 2, 512, 2, 4, 8, 0, 1, 0) \
  X(1024, 2, 512, 2, 3, 1, 0, 1, 0) \
  X(1024, 2, 256, 2, 3, 1, 0, 1, 0) \
  X(1024, 2, 256, 2, 3, 1, 0, 1, 1) \
  X(1024, 2, 256, 2, 3, 1, 1, 1, 0) \
  X(1024, 2, 256, 2, 3, 1, 1, 4, 0) \
  X(1024, 2, 256, 2, 3, 4, 1, 1, 0) \
  X(1024, 2, 256, 2, 3, 8, 1, 1, 0) \
  X(1024, 4, 256, 2, 3, 1, 1, 1, 0) \
  X(1024, 4, 256, 1, 3, 1, 1, 1, 0) \
  X(1024, 4, 256, 4, 3, 1, 1, 1, 0) \
  X(1024, 2, 128, 2, 3, 1, 1, 1, 0) \
  X(1024, 2, 128, 4, 3, 1, 1, 1, 0) \
  X(1024, 1, 64, 4, 3, 1, 1, 1, 0) \
  X(1024, 4, 512, 2, 3, 1, 1, 1, 0) \
  X(1024, 4, 512, 1, 3, 1, 1, 1, 0) \
  X(1024, 8, 512, 4, 4, 4, 1, 1, 0) \
  X(1024, 8, 1024, 4, 3, 4, 1, 1, 0) \
  X(1024, 8, 1024, 2, 3, 4, 1, 1, 0) \
  X(1024, 4, 256, 1, 3, 1, 0, 1, 0) \
  X(1024, 8, 512, 2, 3, 1, 0, 1, 0) \
  X(1024, 8, 512, 1, 3, 1, 0, 1, 0) \
  X(1024, 8, 1024, 1, 3, 1, 0, 1, 0) \
  X(1024, 4, 512, 1, 3, 1, 0, 1, 0) \
  X(1024, 4, 512, 2, 3, 1, 0, 1, 0) \
  X(1024, 2, 256, 2, 3, 1, 2, 1, 0) \
  X(1024, 2, 256, 2, 3, 1, 2, 3, 0) \
  X(1024, 2, 256, 4, 3, 1, 0, 1, 0) \
  X(1024, 2, 512, 4, 3, 1, 0, 1, 0) \
  X(1024, 2, 512, 2, 3, 8, 0, 1, 0) \
  X(1024, 2, 512, 2, 2, 8, 0, 1, 0) \
  X(1024, 2, 512, 2, 4, 1, 0, 1, 0) \
  X(1024, 2, 256, 2, 3, 8, 0, 1, 0) \
  X(1024, 2, 256, 2, 4, 8, 0, 1, 0) \
  X(1024, 8, 512, 1, 4, 2, 0, 1, 0) \
  X(1024, 8, 512, 1, 4, 8, 0, 1, 0) \
  X(1024, 4, 512, 1, 4, 4, 0, 1, 0) \
  X(1024, 1, 256, 4, 4, 1, 0, 1, 0) \
  X(1024, 1, 128, 2, 4, 1, 0, 1, 0) \
  X(2048, 2, 512, 1, 4, 8, 0, 1, 0) \
  X(2048, 2, 512, 1, 3, 1, 0, 1, 0) \
  X(2048, 2, 512, 1, 3, 1, 0, 1, 1) \
  X(2048, 2, 512, 1, 3, 1, 1, 4, 0) \
  X(2048, 2, 256, 2, 3, 1, 1, 1, 0) \
  X(2048, 2, 256, 1, 3, 1, 1, 1, 0) \
  X(2048, 4, 512, 1, 3, 1, 1, 1, 0) \
  X(2048, 1, 128, 4, 3, 1, 1, 1, 0) \
  X(2048, 4, 1024, 1, 3, 1, 1, 1, 0) \
  X(2048, 2, 512, 1, 3, 1, 2, 1, 0) \
  X(2048, 2, 512, 1, 3, 1, 2, 4, 0) \
  X(2048, 2, 512, 1, 3, 8, 0, 1, 0) \
  X(4096, 2, 1024, 1, 4, 8, 0, 1, 0) \
  X(4096, 2, 512, 1, 4, 8, 0, 1, 0) \
  X(4096, 2, 1024, 1, 3, 4, 1, 1, 0) \
  X(4096, 2, 512, 1, 3, 1, 1, 1, 0) \
  X(4096, 2, 512, 1, 4, 4, 1, 1, 0) \
  X(4096, 2, 512, 1, 4, 8, 1, 1, 0) \
  X(4096, 2, 512, 1, 3, 4, 1, 1, 0) \
  X(4096, 1, 256, 1, 3, 1, 1, 1, 0) \
  X(4096, 1, 256, 2, 3, 1, 1, 1, 0) \
  X(4096, 2, 512, 2, 3, 1, 1, 1, 0) \
  X(4096, 2, 1024, 1, 3, 1, 1, 1, 0) \
  X(4096, 2, 1024, 1, 3, 8, 1, 1, 0) \
  X(4096, 1, 512, 1, 3, 1, 1, 1, 0) \
  X(4096, 1, 512, 2, 3, 1, 1, 1, 0) \
  X(4096, 2, 512, 1, 3, 1, 0, 1, 0) \
  X(4096, 2, 1024, 1, 4, 8, 1, 1, 0) \
  X(4096, 2, 512, 1, 3, 8, 0, 1, 0)

struct KVariant {
  int L, C, T, KS, PS, OFF, PK, W, PF;
  void (*kern)(KPassArgs);
  void (*batched)(const KPassArgs*, int);
  // work items per column group count: the slots cover nyf columns (PK: nyf-1)
  int groups(int nyf) const { return ((PK ? nyf - 1 : nyf) + C - 1) / C; }
};

// rpass variants {NY, TLR, PAIRS, W, PFR}: TLR threads per row pair (E = NY/TLR
// points per thread), PAIRS row pairs per workgroup, W minimum waves per SIMD,
// PFR = 1 loads both rows' G fields up front (shapes with E = 4), PFR = 2 runs
// one row per TLR threads instead of a pair, PFR = 3 one row with its four G
// fields loaded before the first transform, PFR = 4 and 5 are PFR = 3 and 2
// with the real product's forward transform at half length; the first listed
// for each NY is the default.  Single rows win at every size from 256 (the
// grids up to 1024 run in one round and are bound by a workgroup's chain of
// transforms; at 1024 one wave per row with all loads up front, no workgroup
// barriers: -0.6 % per step in interleaved timing against 256 threads) and at
// 4096, where the mixed-space fields' 2-row tiles make a row pair's sequential
// halves read each 32-B run twice (single rows 260 against 287 us per launch,
// the two rows of a tile on one XCD).  Interleaved on one solver (median us
// per step): 512, the half-length one-wave shape 88.6 against 96.1 for
// one-wave row pairs and 97.3 for the same shape at full length; 2048,
// 256-thread single rows 491.0 against 549.6 for 256-thread row pairs (491.9
// at half length); 1024 and 4096 keep the full length (166.3 against 167.2,
// 2302 against 2355), 256 ties (51.43 against 51.44).  PFR + 8 (every
// twiddle power preloaded, one-wave rows only): 256, 48.84 against 51.06 for
// the same shape forming the powers, so the 256 default; 512 and 1024 slower
// (89.4 against 88.3, 185.2 against 164.7: a likely reason, not measured,
// is the scattered table reads of every wave at the kernel's start).
#define NS2D_RPASS_VARIANTS(X) \
  X(8, 1, 64, 1, 0) \
  X(16, 1, 64, 1, 0) \
  X(16, 4, 1, 1, 4) \
  X(16, 4, 1, 1, 10) \
  X(16, 4, 1, 1, 12) \
  X(32, 2, 32, 1, 0) \
  X(32, 4, 16, 1, 0) \
  X(64, 4, 16, 1, 0) \
  X(64, 8, 8, 1, 0) \
  X(128, 8, 8, 1, 0) \
  X(128, 16, 4, 1, 0) \
  X(128, 32, 2, 1, 0) \
  X(256, 64, 1, 1, 10) \
  X(256, 64, 1, 1, 2) \
  X(256, 64, 1, 1, 0) \
  X(256, 64, 1, 1, 1) \
  X(256, 16, 4, 1, 0) \
  X(256, 32, 2, 1, 0) \
  X(256, 64, 1, 1, 3) \
  X(256, 64, 1, 1, 4) \
  X(512, 64, 1, 1, 4) \
  X(512, 64, 1, 1, 0) \
  X(512, 64, 1, 1, 2) \
  X(512, 32, 2, 1, 0) \
  X(512, 128, 1, 1, 0) \
  X(512, 128, 1, 1, 1) \
  X(512, 64, 2, 1, 0) \
  X(512, 64, 4, 1, 0) \
  X(512, 128, 1, 1, 2) \
  X(512, 64, 1, 1, 3) \
  X(1024, 64, 1, 1, 3) \
  X(1024, 256, 1, 1, 2) \
  X(1024, 256, 1, 1, 0) \
  X(1024, 256, 1, 1, 1) \
  X(1024, 64, 1, 1, 0) \
  X(1024, 128, 1, 1, 0) \
  X(1024, 64, 2, 1, 0) \
  X(1024, 64, 4, 1, 0) \
  X(1024, 128, 1, 1, 2) \
  X(1024, 256, 1, 1, 3) \
  X(1024, 64, 1, 1, 2) \
  X(1024, 128, 1, 2, 2) \
  X(1024, 64, 1, 1, 4) \
  X(1024, 64, 1, 1, 5) \
  X(1024, 256, 1, 1, 5) \
  X(2048, 256, 1, 1, 2) \
  X(2048, 256, 1, 1, 0) \
  X(2048, 128, 1, 2, 0) \
  X(2048, 128, 1, 1, 0) \
  X(2048, 512, 1, 1, 0) \
  X(2048, 512, 1, 1, 1) \
  X(2048, 512, 1, 1, 2) \
  X(2048, 512, 1, 1, 3) \
  X(2048, 512, 1, 1, 5) \
  X(2048, 256, 1, 1, 5) \
  X(4096, 512, 1, 4, 2) \
  X(4096, 512, 1, 4, 0) \
  X(4096, 256, 1, 1, 0) \
  X(4096, 256, 1, 2, 0) \
  X(4096, 512, 1, 1, 0) \
  X(4096, 1024, 1, 1, 0) \
  X(4096, 1024, 1, 1, 1) \
  X(4096, 1024, 1, 1, 2) \
  X(4096, 1024, 1, 1, 3) \
  X(4096, 512, 1, 4, 5) \
  X(4096, 1024, 1, 1, 5)

struct RVariant {
  int NY, TLR, PAIRS, W, PFR;
  void (*kern)(RPassArgs);
  void (*batched)(const RPassArgs*, int);
  void (*cfl)(RPassArgs);   // kern with the CFL-rate reduction (mode RP_CFL)
};

// The shape tables, assembled from the per-length translation units in
// list order (so a shape's index among those of its length is its position
// in the lists above).
const std::vector<KVariant>& kvariant_table();
const std::vector<RVariant>& rvariant_table();

}  // namespace ns2d
