// Internal definitions shared by the kernels and the C-ABI host code.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>

// Spectral arrays ([Nx][ld] complex, kx rows, ky columns) are stored in
// x-tiles: element (x, ky) at ((x/XT)*ld + ky)*XT + x%XT, so XT consecutive
// rows of a column are contiguous (XT = 1: plain row-major).  Two layouts:
// NS2D_XTILE_K for the k-space arrays only the column pass's epilogue (and
// the measures, the adjoint's k-space recursion and the host copies) touches
// -- state, accumulator, history, coefficient and filter tables, RHS output;
// 0 there means column-major, element (x, ky) at ky*Nx + x, so that a column
// group's columns are whole contiguous runs; NS2D_XTILE_M for the
// mixed-space fields the row and column passes hand to each other (G1..G4,
// N, the adjoint's tape and fields).
// Defaults (measured on MI355X against row-major arrays, per RK4CN step):
// column-major k-space and 2-row mixed-space tiles, -7.5 % at 1024^2, -6 % at
// 4096^2 (the column pass -22 %, its 2-column groups reading whole 32-B
// runs of a tile instead of 16 B per row; the row pass +25 %), -7 % for CN-AB4
// at 4096^2.
#ifndef NS2D_XTILE_K
#define NS2D_XTILE_K 0
#endif
// The mixed-space fields go two ways: G (the column pass's four inverse
// fields, and the adjoint's l_n transform) from the column pass to the row
// pass, N (the row pass's output, the adjoint's B fields, the initial
// conditions' row transforms) from the row pass to the column pass; each set
// has its own tile height.
#ifndef NS2D_XTILE_M
#define NS2D_XTILE_M 2
#endif
#ifndef NS2D_XTILE_G
#define NS2D_XTILE_G NS2D_XTILE_M
#endif
#ifndef NS2D_XTILE_N
#define NS2D_XTILE_N NS2D_XTILE_M
#endif

namespace ns2d {

template <int XT>
__host__ __device__ __forceinline__ size_t tile_idx(int x, int ky, int ld) {
  // at most 8: the tile height must divide every Nx the ABI accepts
  static_assert(XT == 1 || XT == 2 || XT == 4 || XT == 8, "tile height must be 1, 2, 4 or 8");
  constexpr int S = XT == 1 ? 0 : XT == 2 ? 1 : XT == 4 ? 2 : 3;
  if constexpr (XT == 1) return (size_t)x * ld + ky;
  else return (((size_t)(x >> S) * ld + ky) << S) + (size_t)(x & (XT - 1));
}
constexpr int XT_K = NS2D_XTILE_K, XT_G = NS2D_XTILE_G, XT_N = NS2D_XTILE_N;
// k-space array index (nx: the grid's Nx)
__host__ __device__ __forceinline__ size_t kidx(int x, int ky, int ld, int nx) {
  if constexpr (XT_K == 0) return (size_t)ky * nx + x;
  else return tile_idx<XT_K == 0 ? 1 : XT_K>(x, ky, ld);
}
// mixed-space index of a field the column pass writes and the row pass reads
__host__ __device__ __forceinline__ size_t gidx(int x, int ky, int ld) { return tile_idx<XT_G>(x, ky, ld); }
// ... of a field the row pass writes and the column pass reads
__host__ __device__ __forceinline__ size_t nidx(int x, int ky, int ld) { return tile_idx<XT_N>(x, ky, ld); }

// RK4 Butcher tableau and AB4 weights (reference Solver/src/solver.c:29-47)
constexpr double RK4_A21 = 0.5, RK4_A32 = 0.5, RK4_A43 = 1.0;
constexpr double RK4_B1 = 1.0 / 6.0, RK4_B2 = 1.0 / 3.0, RK4_B3 = 1.0 / 3.0, RK4_B4 = 1.0 / 6.0;
constexpr double AB4_1 = 55.0 / 24.0, AB4_2 = -59.0 / 24.0, AB4_3 = 37.0 / 24.0, AB4_4 = -3.0 / 8.0;
// data_types.h:171-173
constexpr double MAX_VORT_LIM = 1e100;

enum { FILTER_23 = 0, FILTER_HOULI_REF = 1, FILTER_HOULI = 2, FILTER_NONE = 3 };
enum { VISC_K2 = 0, VISC_HYPER = 1, VISC_K2_DRAG = 2, VISC_HYPER_DRAG = 3 };

// kpass mode bits
enum { KP_FWD = 1, KP_INV = 2, KP_CONJFIX = 4, KP_STAT = 8, KP_ITER = 16, KP_RAW = 32 };
// kpass epilogues
enum { EPI_FROM_STATE = 0, EPI_FROM_INPUT = 1, EPI_RHS = 2, EPI_RK1 = 3, EPI_RK2 = 4, EPI_RK3 = 5,
       EPI_RK4CN = 6, EPI_AB4 = 7 };
// rpass mode bits
// RP_CFL (with RP_NL): also the maximum over the rows of pi (|u|/dx + |v|/dy)
// into *cmax (GetTimestep's convective rate, solver.c:2425-2441)
enum { RP_NL = 1, RP_W = 2, RP_CFL = 4 };

// device status words (unsigned 64-bit)
enum { STAT_WMAX_CUR = 0, STAT_WMAX_LAST = 1, STAT_FIRST_BAD = 2, STAT_ITER = 3, STAT_CFL = 4, STAT_N = 5 };

// per-bin reduction quantities of the measures kernel
enum {
  Q_ENRG = 0, Q_ENST, Q_PALIN, Q_ENRG_DISS, Q_ENST_DISS, Q_DIV,
  Q_DERIV, Q_DISS, Q_DERIV_K2, Q_DISS_K2,
  Q_SB_DERIV, Q_SB_DISS, Q_SB_DERIV_K2, Q_SB_DISS_K2,
  NQ
};

struct PhysArgs {
  int nx, ny;
  double nu, hyper_pow, alpha, drag_pow;
  int filter;
  double kmax_sqr;      // (int)(Nx/3)^2, solver.c:1048
  int kolm_kill;        // 2/3 filter also zeroes kx=0 (KOLM, solver.c:1072)
  double houli_edge;    // Hou-Li (integer-division form) value on the Nyquist row/column
  double houli_corner;  // ... and on the corner
  double norm2;         // (1/(Nx*Ny))^2, solver.c:1005
  int force_on, force_k;
  double force_val;
  int ld;
  const double* filt_tab;  // FILTER_HOULI: per-mode filter, [Nx][ld]
};

struct KPassArgs {
  const double2* w_in;   // state read by the epilogue
  double2* w_out;        // new state (RK4CN stage 4, AB4) / ForceConjugacy target
  double2* acc;
  const double2* h_old;
  const double2* h_mid;
  const double2* h_new;
  double2* h_store;      // AB4: slot receiving this step's nonlinear term
  double2* hist_store;   // AB4 start-up: K1 of an RK4CN step
  const double2* nin;
  const double2* sin;
  double2* g[4];
  double2* kout;
  const double2* conj_src;
  double2* conj_dst;
  unsigned long long* stat;
  const double2* tw;
  double2* pairbuf;      // packed (0, Ny/2) column pair: held columns, [split][2][Nx]
  int nyf, ld;
  const double* lam;     // per-mode D_fac/dt of the CN update, [Nx][ld]
  int mode, epi;
  // 1: the last column group (one valid column, KS = 1 shapes) runs as two
  // work items with two of the four inverse fields each (set by the launcher)
  int last_split;
  double dt, coef;
  PhysArgs ph;
};

struct RPassArgs {
  const double2* g[4];
  int nx;
  double2* nout;
  double* wreal;
  unsigned long long* wmax;
  const double2* tw;
  int ld, mode;
  double wscale;
  // RP_CFL: u + iv = (row transform) * cfl_s, grid spacings, result bits
  unsigned long long* cmax;
  double cfl_s, cfl_dx, cfl_dy;
};

// adjoint kpass modes and epilogues
enum { KA_FWD = 1, KA_INV = 2 };
enum { ADJ_START = 0, ADJ_S4 = 1, ADJ_S3 = 2, ADJ_S2 = 3, ADJ_S1 = 4 };

struct KAdjArgs {
  const double2* bin[4];   // row-pass adjoint outputs B1..B4
  double2* lout;           // x-inverse of Wh .* M .* l_K for the next stage
  const double2* lamp_in;  // dJ/dw_{n+1}
  double2* lamp_out;       // dJ/dw_n (ADJ_S1)
  double2* lacc;           // running dJ/dw_n
  const double* lam;       // D_fac/dt table of this step
  const double* lam_prev;  // ... of the previous step (ADJ_S1 with KA_INV)
  const double2* tw;
  int nyf, ld, mode, epi;
  double dt;
  PhysArgs ph;
};

struct RAdjArgs {
  const double2* lin;      // x-transformed r2c^T field
  const double2* g[4];     // the stage's x-transformed G1..G4 (forward tape)
  double2* bout[4];
  const double2* tw;
  int nx, ld;
};

struct MeasArgs {
  const double2* w_hat;
  const double2* rhs;
  double* out;
  const double* lam;     // per-mode pre_fac (sys_msr.c:160-175), [Nx][ld]
  int nx, ny, ld, quirk_enst_diss;
  int nbins;             // spectrum bins (workgroups past the last one exit)
  double sbst_lo, sbst_hi;
  PhysArgs ph;
};

// blockIdx -> work item map that keeps consecutive work items on one XCD
// (blocks are dealt round-robin over the 8 XCDs, each with its own L2)
__device__ __forceinline__ int xcd_remap(int b, int n) {
  constexpr int NXCD = 8;
  if (n < NXCD) return b;
  const int q = n / NXCD, r = n % NXCD;
  const int xcd = b % NXCD, j = b / NXCD;
  return (xcd < r ? xcd * (q + 1) : r * (q + 1) + (xcd - r) * q) + j;
}

// Spectrum bin of measures workgroup g: chunks of MEAS_CHUNK consecutive bins
// dealt round-robin to the 8 XCDs (workgroup g runs on XCD g % 8), so that the
// bins sharing cache lines (neighbouring annuli) run on one XCD's L2 while
// every XCD gets inner and outer bins alike.  The grid is rounded up to
// 8 * MEAS_CHUNK workgroups; g maps onto [0, grid) one to one.
constexpr int MEAS_CHUNK = 32;
__host__ __device__ __forceinline__ int meas_bin(int g) {
  const int xcd = g % 8, j = g / 8;
  return ((j / MEAS_CHUNK) * 8 + xcd) * MEAS_CHUNK + j % MEAS_CHUNK;
}
inline int meas_grid(int nbins) { return (nbins + 8 * MEAS_CHUNK - 1) / (8 * MEAS_CHUNK) * (8 * MEAS_CHUNK); }

hipError_t launch_kpass(int nx, int variant, const KPassArgs& a, hipStream_t s);
int kpass_variant_count(int nx);
int kpass_default_variant(int nx, bool ab4);
hipError_t launch_kpass_batched(int nx, int nyf, int variant, const KPassArgs* tab, int members, hipStream_t s);
hipError_t launch_rpass_batched(int nx, int ny, int variant, const RPassArgs* tab, int members, hipStream_t s);
int kpass_variant_info(int nx, int idx, int* C, int* E, int* T, int* KS);
hipError_t launch_rpass(int nx, int ny, int variant, const RPassArgs& a, hipStream_t s);
int rpass_variant_count(int ny);
int rpass_variant_info(int ny, int idx, int* TLR, int* E, int* T, int* PAIRS);
int rpass_variant_extra(int ny, int idx, int* W, int* PS, int* OFF, int* PK);
int kpass_variant_extra(int nx, int idx, int* W, int* PS, int* OFF, int* PK);
unsigned long long shape_table_hash(int nx, int ny);
hipError_t launch_r2c_rows(int nx, int ny, const double* in, double2* out, const double2* tw, int ld,
                           hipStream_t s);
hipError_t launch_measures(int nbins, const MeasArgs& a, hipStream_t s);
hipError_t launch_relayout(const double2* src, double2* dst, int nx, int nyf, int ld, int dir, hipStream_t s);

// allocation-free measures: per-device scratch (ns2d_noalloc.hip) sized for
// the spectrum length of a 4096 x 4096 grid (sys_msr.c:394)
constexpr int NOALLOC_MAX_BINS = 2897;
constexpr size_t NOALLOC_SCRATCH_DOUBLES = (size_t)NOALLOC_MAX_BINS * (NQ + 8) + NQ;
hipError_t noalloc_scratch(int dev, double** ptr);
hipError_t launch_measures_finalize(const double* part, int nb, double cn, int with_flux, double* out, hipStream_t s);
hipError_t launch_kpass_adj(int nx, const KAdjArgs& a, hipStream_t s);
hipError_t launch_rpass_adj(int nx, int ny, const RAdjArgs& a, hipStream_t s);

}  // namespace ns2d
