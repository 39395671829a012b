// Host implementation of the ns2d C ABI (include/ns2d_gpu.h).
//
// Owns the device arrays of one solver instance and issues the kernel
// schedule of one time step.  RK4CN step (reference RK4Step,
// Solver/src/solver.c:625-873): four (rpass, kpass) pairs, the k-space pass of
// each stage finishing that stage's RHS, applying the stage epilogue and
// starting the next stage's inverse transforms.  CN-AB4 step (AB4Step,
// solver.c:261-350): one (rpass, kpass) pair.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <sys/stat.h>
#include <vector>

#include "../../include/ns2d_gpu.h"
#include "ns2d_fft.hpp"
#include "ns2d_internal.h"

using namespace ns2d;

namespace {

thread_local std::string g_err;

int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

#define HIPCHK(expr)                                                                      \
  do {                                                                                    \
    hipError_t e_ = (expr);                                                               \
    if (e_ != hipSuccess) return fail(NS2D_ERR_HIP, "%s: %s", #expr, hipGetErrorString(e_)); \
  } while (0)

bool pow2_in_range(int n) { return n >= 8 && n <= 4096 && (n & (n - 1)) == 0; }

}  // namespace

struct ns2d_handle {
  ns2d_params p;
  int nyf = 0, ld = 0, nspect = 0;
  size_t M = 0;  // nx * ld complex elements per spectral array
  hipStream_t stream = nullptr;
  bool own_stream = true;
  double2 *w[2] = {nullptr, nullptr};  // state ping-pong: w[wcur] is run_data->w_hat
  int wcur = 0;
  double2 *acc = nullptr, *g[4] = {nullptr, nullptr, nullptr, nullptr};
  double2 *nbuf = nullptr, *kout = nullptr, *sbuf = nullptr, *hist[4] = {nullptr, nullptr, nullptr, nullptr};
  int kvar = 0;            // kpass variant (tuning)
  int rvar = 0;            // rpass variant (tuning)
  double* lam_tab[4] = {nullptr, nullptr, nullptr, nullptr};  // per VISC_* kind
  double* filt_tab = nullptr;
  double2 *conj[2] = {nullptr, nullptr}, *tw = nullptr;
  double2* pairbuf = nullptr;  // column pass: held columns of the packed (0, Ny/2) pair, [split][2][Nx]
  unsigned long long* stat = nullptr;
  double* meas = nullptr;
  double* wreal = nullptr;
  double* rbuf = nullptr;
  double2* stage = nullptr;  // [Nx][Ny/2+1] staging between the host and the k-space layout (x-tiles only)
  bool inv_valid = false;  // G1..G4 and conj[cur] hold the transforms of the state
  int cur = 0;             // conj[cur] = ky=0 column of the current stage input
  long iters = 0;
  int hyper_eff = 0;       // HYPER_VISC_FLAG as mutated by the reference (Q3)
  PhysArgs ph{};
  // hipGraph of one period of steps (kernel launches only; see step_impl)
  bool use_graph = true;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  struct {
    double dt;
    int wcur, cur, ring, kvar, rvar, check, period;
    const double* lam;
  } gkey{};
  // adjoint buffers (allocated on first use)
  double2* tape = nullptr;     // the tape in use: [slot][4 fields][M], slot = 4*step + stage
  double2* tape_buf[2] = {nullptr, nullptr};  // second one only when it fits (overlap)
  long tape_slots = 0;
  hipStream_t stream2 = nullptr;  // recompute stream of the adjoint
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  double2* ckpt = nullptr;     // [segment][M]
  long ckpt_n = 0;
  double2 *lamp[2] = {nullptr, nullptr}, *lacc = nullptr, *lbuf = nullptr, *bbuf[4] = {nullptr, nullptr, nullptr, nullptr};
  // hipGraphs of one full adjoint segment (recompute + backward sweep), one
  // per parity of the state buffer and of the adjoint ping-pong at its start
  struct AdjGraph {
    hipGraphExec_t exec = nullptr;
    double dt = 0.0;
    long C = 0;
    int wcur = -1, lc = -1;
    const double* lam = nullptr;
    const double2* tape = nullptr;
  } adj_graph[2];
  // profiling
  bool profiling = false;
  std::vector<hipEvent_t> ev;
  std::vector<int> ev_kind;
};

namespace {

int visc_kind_rk(const ns2d_handle* h) {
  const bool hyp = h->hyper_eff == 1, drag = h->p.drag_flag == 1;
  if (hyp && drag) return VISC_HYPER_DRAG;
  if (!hyp && drag) return VISC_K2_DRAG;
  if (h->p.quirks & NS2D_QUIRK_VISC_ASSIGN) return VISC_HYPER;  // solver.c:843
  return hyp ? VISC_HYPER : VISC_K2;
}

int visc_kind_plain(const ns2d_handle* h) {
  const bool hyp = h->hyper_eff == 1, drag = h->p.drag_flag == 1;
  if (hyp && drag) return VISC_HYPER_DRAG;
  if (!hyp && drag) return VISC_K2_DRAG;
  if (hyp) return VISC_HYPER;
  return VISC_K2;
}

// Per-mode viscous / drag coefficient of the CN factor for one VISC_* kind,
// evaluated with the host libm pow exactly as the reference writes it
// (solver.c:835-850, sys_msr.c:160-175); built on first use.
const double* lam_table(ns2d_handle* h, int kind) {
  if (h->lam_tab[kind]) return h->lam_tab[kind];
  const int nx = h->p.nx, ld = h->ld;
  std::vector<double> t((size_t)nx * ld, 0.0);
  for (int i = 0; i < nx; ++i) {
    const long kx = i <= nx / 2 ? i : i - nx;
    for (int j = 0; j < h->nyf; ++j) {
      if (i == 0 && j == 0) continue;
      const double k_sqr = (double)(kx * kx + (long)j * j);
      const ns2d_params& p = h->p;
      double v;
      switch (kind) {
        case VISC_HYPER_DRAG: v = p.nu * std::pow(k_sqr, p.hyper_pow) + p.alpha * std::pow(k_sqr, p.drag_pow); break;
        case VISC_K2_DRAG: v = p.nu * k_sqr + p.alpha * std::pow(k_sqr, p.drag_pow); break;
        case VISC_HYPER: v = p.nu * std::pow(k_sqr, p.hyper_pow); break;
        default: v = p.nu * k_sqr; break;
      }
      t[kidx(i, j, ld, nx)] = v;
    }
  }
  double* d = nullptr;
  if (hipMalloc((void**)&d, sizeof(double) * t.size()) != hipSuccess) return nullptr;
  if (hipMemcpy(d, t.data(), sizeof(double) * t.size(), hipMemcpyHostToDevice) != hipSuccess) {
    (void)hipFree(d);
    return nullptr;
  }
  h->lam_tab[kind] = d;
  return d;
}

void after_rk_update(ns2d_handle* h) {
  if ((h->p.quirks & NS2D_QUIRK_VISC_ASSIGN) && h->p.drag_flag != 1) h->hyper_eff = 1;
}

KPassArgs kbase(ns2d_handle* h) {
  KPassArgs a{};
  a.w_in = h->w[h->wcur];
  a.w_out = h->w[h->wcur];
  a.acc = h->acc;
  a.nin = h->nbuf;
  for (int f = 0; f < 4; ++f) a.g[f] = h->g[f];
  a.kout = h->kout;
  a.stat = h->stat;
  a.tw = h->tw;
  a.pairbuf = h->pairbuf;
  a.nyf = h->nyf;
  a.ld = h->ld;
  a.ph = h->ph;
  return a;
}

RPassArgs rbase(ns2d_handle* h, int mode) {
  RPassArgs a{};
  for (int f = 0; f < 4; ++f) a.g[f] = h->g[f];
  a.nout = h->nbuf;
  a.wreal = nullptr;
  a.wmax = h->stat + STAT_WMAX_CUR;
  a.tw = h->tw;
  a.ld = h->ld;
  a.nx = h->p.nx;
  a.mode = mode;
  a.wscale = 1.0 / (double)((long)h->p.nx * h->p.ny);
  return a;
}

// A [Nx][Ny/2+1] complex array of a caller (host or device memory, `kind`)
// into / out of the handle's device layout.
int copy_in(ns2d_handle* h, double2* dst, const void* src, hipMemcpyKind kind) {
  const size_t rowb = h->nyf * sizeof(double2);
  if constexpr (XT_K == 1) {
    HIPCHK(hipMemcpy2DAsync(dst, h->ld * sizeof(double2), src, rowb, rowb, h->p.nx, kind, h->stream));
  } else {
    HIPCHK(hipMemcpyAsync(h->stage, src, rowb * h->p.nx, kind, h->stream));
    HIPCHK(launch_relayout(h->stage, dst, h->p.nx, h->nyf, h->ld, 0, h->stream));
  }
  return NS2D_OK;
}

int copy_out(ns2d_handle* h, void* dst, const double2* src, hipMemcpyKind kind) {
  const size_t rowb = h->nyf * sizeof(double2);
  if constexpr (XT_K == 1) {
    HIPCHK(hipMemcpy2DAsync(dst, rowb, src, h->ld * sizeof(double2), rowb, h->p.nx, kind, h->stream));
  } else {
    HIPCHK(launch_relayout(src, h->stage, h->p.nx, h->nyf, h->ld, 1, h->stream));
    HIPCHK(hipMemcpyAsync(dst, h->stage, rowb * h->p.nx, kind, h->stream));
  }
  return NS2D_OK;
}

enum { KIND_K = 0, KIND_R = 1, KIND_OTHER = 2 };

void ev_begin(ns2d_handle* h, int kind) {
  if (!h->profiling) return;
  hipEvent_t e;
  (void)hipEventCreate(&e);
  (void)hipEventRecord(e, h->stream);
  h->ev.push_back(e);
  h->ev_kind.push_back(kind);
}
void ev_end(ns2d_handle* h) {
  if (!h->profiling) return;
  hipEvent_t e;
  (void)hipEventCreate(&e);
  (void)hipEventRecord(e, h->stream);
  h->ev.push_back(e);
  h->ev_kind.push_back(-1);
}

int run_kpass(ns2d_handle* h, const KPassArgs& a) {
  ev_begin(h, KIND_K);
  hipError_t e = launch_kpass(h->p.nx, h->kvar, a, h->stream);
  ev_end(h);
  if (e != hipSuccess) return fail(NS2D_ERR_HIP, "kpass launch: %s", hipGetErrorString(e));
  return NS2D_OK;
}

int run_rpass(ns2d_handle* h, const RPassArgs& a) {
  ev_begin(h, KIND_R);
  hipError_t e = launch_rpass(h->p.nx, h->p.ny, h->rvar, a, h->stream);
  ev_end(h);
  if (e != hipSuccess) return fail(NS2D_ERR_HIP, "rpass launch: %s", hipGetErrorString(e));
  return NS2D_OK;
}

// G1..G4 (and conj[cur]) from the state: the inverse half of the first RHS
int prepare_inv(ns2d_handle* h) {
  if (h->inv_valid) return NS2D_OK;
  KPassArgs a = kbase(h);
  a.mode = KP_INV;
  a.epi = EPI_FROM_STATE;
  a.conj_dst = h->conj[h->cur];
  int rc = run_kpass(h, a);
  if (rc) return rc;
  h->inv_valid = true;
  return NS2D_OK;
}

int conj_mode(const ns2d_handle* h) { return (h->p.quirks & NS2D_QUIRK_CONJ_GLOBAL) ? KP_CONJFIX : 0; }

// Arguments of stage s (0..3) of an RK4CN step (RK4Step, solver.c:625-873)
// from the handle's current buffers.  The ky=0 column hand-off alternates
// conj[] every stage; the new state goes to the other state buffer.
int rk4cn_stage_args(ns2d_handle* h, double dt, bool check, int s, RPassArgs* ra, KPassArgs* ka) {
  static const int epis[4] = {EPI_RK1, EPI_RK2, EPI_RK3, EPI_RK4CN};
  const double coefs[4] = {dt * RK4_A21, dt * RK4_A32, dt * RK4_A43, 0.0};
  const long it = h->iters + 1;
  *ra = rbase(h, RP_NL | ((check && s == 0) ? RP_W : 0));
  KPassArgs a = kbase(h);
  a.mode = KP_FWD | KP_INV | conj_mode(h);
  if (check && s == 0) a.mode |= KP_STAT;
  if (s == 3) a.mode |= KP_ITER;
  a.epi = epis[s];
  a.coef = coefs[s];
  a.dt = dt;
  a.lam = lam_table(h, visc_kind_rk(h));
  if (!a.lam) return fail(NS2D_ERR_HIP, "coefficient table allocation failed");
  // AB4 start-up keeps K1 of its RK4CN steps (solver.c:283)
  a.hist_store = (s == 0 && h->p.integrator == NS2D_INT_AB4 && it <= 3) ? h->hist[it - 1] : nullptr;
  a.conj_src = h->conj[h->cur ^ (s & 1)];
  a.conj_dst = h->conj[h->cur ^ ((s + 1) & 1)];
  if (s == 3) a.w_out = h->w[h->wcur ^ 1];
  *ka = a;
  return NS2D_OK;
}

int rk4cn_step(ns2d_handle* h, double dt, bool check) {
  for (int s = 0; s < 4; ++s) {
    RPassArgs ra;
    KPassArgs ka;
    int rc = rk4cn_stage_args(h, dt, check, s, &ra, &ka);
    if (rc) return rc;
    rc = run_rpass(h, ra);
    if (rc) return rc;
    rc = run_kpass(h, ka);
    if (rc) return rc;
  }
  after_rk_update(h);
  h->iters += 1;
  h->wcur ^= 1;
  return NS2D_OK;
}

int ab4_step(ns2d_handle* h, double dt, bool check) {
  const long it = h->iters + 1;
  if (it <= 3) return rk4cn_step(h, dt, check);
  // AB iteration m: slots m%4, (m+1)%4, (m+2)%4 hold N(n-3), N(n-2), N(n-1);
  // N(n) goes to the fourth slot, so no slot is read and written in one pass
  const long m = it - 4;
  int rc = run_rpass(h, rbase(h, RP_NL | (check ? RP_W : 0)));
  if (rc) return rc;
  KPassArgs a = kbase(h);
  a.mode = KP_FWD | KP_INV | KP_ITER | conj_mode(h) | (check ? KP_STAT : 0);
  a.epi = EPI_AB4;
  a.dt = dt;
  a.lam = lam_table(h, visc_kind_plain(h));  // solver.c:315-330
  if (!a.lam) return fail(NS2D_ERR_HIP, "coefficient table allocation failed");
  a.h_old = h->hist[m % 4];
  a.h_mid = h->hist[(m + 1) % 4];
  a.h_new = h->hist[(m + 2) % 4];
  a.h_store = h->hist[(m + 3) % 4];
  a.conj_src = h->conj[h->cur];
  a.conj_dst = h->conj[h->cur ^ 1];
  a.w_out = h->w[h->wcur ^ 1];
  rc = run_kpass(h, a);
  if (rc) return rc;
  h->cur ^= 1;
  h->wcur ^= 1;
  h->iters = it;
  return NS2D_OK;
}

int read_stat(ns2d_handle* h, unsigned long long* st) {
  HIPCHK(hipMemcpyAsync(st, h->stat, sizeof(unsigned long long) * STAT_N, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  return NS2D_OK;
}

int reset_stat(ns2d_handle* h) {
  unsigned long long st[STAT_N] = {0ull, 0ull, ~0ull, 0ull};
  HIPCHK(hipMemcpyAsync(h->stat, st, sizeof st, hipMemcpyHostToDevice, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  return NS2D_OK;
}

// max|w| of the state into stat[WMAX_CUR] (after clearing it)
int state_wmax(ns2d_handle* h, double* out, double* wreal) {
  int rc = prepare_inv(h);
  if (rc) return rc;
  const unsigned long long zero = 0ull;
  HIPCHK(hipMemcpyAsync(h->stat + STAT_WMAX_CUR, &zero, sizeof zero, hipMemcpyHostToDevice, h->stream));
  RPassArgs a = rbase(h, RP_W);
  a.wreal = wreal;
  rc = run_rpass(h, a);
  if (rc) return rc;
  unsigned long long st[STAT_N];
  rc = read_stat(h, st);
  if (rc) return rc;
  double m;
  std::memcpy(&m, &st[STAT_WMAX_CUR], sizeof m);
  if (out) *out = m;
  return NS2D_OK;
}

// GetTimestep's convective rate of the state: the row pass of the state's
// fields with the CFL reduction (its N output is not used)
int state_cfl(ns2d_handle* h, double* out) {
  int rc = prepare_inv(h);
  if (rc) return rc;
  const unsigned long long zero = 0ull;
  HIPCHK(hipMemcpyAsync(h->stat + STAT_CFL, &zero, sizeof zero, hipMemcpyHostToDevice, h->stream));
  RPassArgs a = rbase(h, RP_NL | RP_CFL);
  a.cmax = h->stat + STAT_CFL;
  a.cfl_s = 1.0 / (double)((long)h->p.nx * h->p.ny);
  a.cfl_dx = 2.0 * M_PI / (double)h->p.nx;
  a.cfl_dy = 2.0 * M_PI / (double)h->p.ny;
  rc = run_rpass(h, a);
  if (rc) return rc;
  unsigned long long st[STAT_N];
  rc = read_stat(h, st);
  if (rc) return rc;
  std::memcpy(out, &st[STAT_CFL], sizeof *out);
  return NS2D_OK;
}

// K = NonlinearRHSBatch(stage input S) into kout, S given by prepared G's
// and conj[cur]; applies the ForceConjugacy write to the state when enabled.
int finish_rhs_into_kout(ns2d_handle* h, bool raw) {
  int rc = run_rpass(h, rbase(h, RP_NL));
  if (rc) return rc;
  KPassArgs a = kbase(h);
  a.mode = KP_FWD | (raw ? 0 : conj_mode(h));
  a.epi = EPI_RHS;
  a.conj_src = h->conj[h->cur];
  return run_kpass(h, a);
}

}  // namespace

extern "C" {

const char* ns2d_last_error(void) { return g_err.c_str(); }

int ns2d_variant_count(int pass, int n) {
  if (pass == NS2D_PASS_COLUMN) return kpass_variant_count(n);
  if (pass == NS2D_PASS_ROW) return rpass_variant_count(n);
  return 0;
}

int ns2d_variant_info(int pass, int n, int idx, int* shape4) {
  if (!shape4) return fail(NS2D_ERR_INVALID, "null argument");
  int rc = -1;
  if (pass == NS2D_PASS_COLUMN) rc = kpass_variant_info(n, idx, &shape4[0], &shape4[1], &shape4[2], &shape4[3]);
  if (pass == NS2D_PASS_ROW) rc = rpass_variant_info(n, idx, &shape4[0], &shape4[1], &shape4[2], &shape4[3]);
  if (rc) return fail(NS2D_ERR_INVALID, "no variant %d of pass %d for n=%d", idx, pass, n);
  return NS2D_OK;
}

int ns2d_variant_extra(int pass, int n, int idx, int* extra4) {
  if (!extra4) return fail(NS2D_ERR_INVALID, "null argument");
  int rc = -1;
  if (pass == NS2D_PASS_COLUMN) rc = kpass_variant_extra(n, idx, &extra4[0], &extra4[1], &extra4[2], &extra4[3]);
  if (pass == NS2D_PASS_ROW) rc = rpass_variant_extra(n, idx, &extra4[0], &extra4[1], &extra4[2], &extra4[3]);
  if (rc) return fail(NS2D_ERR_INVALID, "no variant %d of pass %d for n=%d", idx, pass, n);
  return NS2D_OK;
}

int ns2d_layout(int* xtile_k, int* xtile_g, int* xtile_n) {
  if (!xtile_k || !xtile_g || !xtile_n) return fail(NS2D_ERR_INVALID, "null argument");
  *xtile_k = XT_K;
  *xtile_g = XT_G;
  *xtile_n = XT_N;
  return NS2D_OK;
}

int ns2d_default_variant(int pass, int n, int integrator, int* idx) {
  if (!idx) return fail(NS2D_ERR_INVALID, "null argument");
  if (integrator != NS2D_INT_RK4CN && integrator != NS2D_INT_AB4)
    return fail(NS2D_ERR_INVALID, "unknown integrator %d", integrator);
  if (pass == NS2D_PASS_COLUMN) {
    if (kpass_variant_count(n) == 0) return fail(NS2D_ERR_INVALID, "no column-pass shape for Nx=%d", n);
    *idx = kpass_default_variant(n, integrator == NS2D_INT_AB4);
    return NS2D_OK;
  }
  if (pass == NS2D_PASS_ROW) {
    if (rpass_variant_count(n) == 0) return fail(NS2D_ERR_INVALID, "no row-pass shape for Ny=%d", n);
    *idx = 0;
    return NS2D_OK;
  }
  return fail(NS2D_ERR_INVALID, "unknown pass %d", pass);
}

int ns2d_set_variant(ns2d_t* h, int pass, int idx) {
  if (!h) return fail(NS2D_ERR_INVALID, "null handle");
  if (pass == NS2D_PASS_COLUMN) {
    if (idx < 0 || idx >= kpass_variant_count(h->p.nx))
      return fail(NS2D_ERR_INVALID, "no column-pass variant %d for Nx=%d", idx, h->p.nx);
    h->kvar = idx;
    return NS2D_OK;
  }
  if (pass == NS2D_PASS_ROW) {
    if (idx < 0 || idx >= rpass_variant_count(h->p.ny))
      return fail(NS2D_ERR_INVALID, "no row-pass variant %d for Ny=%d", idx, h->p.ny);
    h->rvar = idx;
    return NS2D_OK;
  }
  return fail(NS2D_ERR_INVALID, "unknown pass %d", pass);
}

int ns2d_params_default(ns2d_params* p, int nx, int ny) {
  if (!p) return fail(NS2D_ERR_INVALID, "null params");
  std::memset(p, 0, sizeof *p);
  // utils.c:51-85 defaults
  p->nx = nx;
  p->ny = ny;
  p->nu = 1.0;
  p->hyper_flag = 0;
  p->hyper_pow = 2.0;   // VISC_POW, utils.c:79
  p->alpha = 1.0;
  p->drag_flag = 0;
  p->drag_pow = -2.0;   // EKMN_POW
  p->filter = NS2D_FILTER_23;
  p->forcing = NS2D_FORCING_NONE;
  p->force_k = 0;
  p->force_scale = 1.0;
  p->integrator = NS2D_INT_RK4CN;
  p->quirks = NS2D_QUIRKS_REF;
  p->check_every = 1;
  p->device = 0;
  return NS2D_OK;
}

static int create_impl(const ns2d_params* p, hipStream_t ext_stream, ns2d_t** out);

int ns2d_create(const ns2d_params* p, ns2d_t** out) { return create_impl(p, nullptr, out); }

static int create_impl(const ns2d_params* p, hipStream_t ext_stream, ns2d_t** out) {
  if (!p || !out) return fail(NS2D_ERR_INVALID, "null argument");
  *out = nullptr;
  if (!pow2_in_range(p->nx) || !pow2_in_range(p->ny))
    return fail(NS2D_ERR_INVALID, "grid %dx%d: Nx and Ny must be powers of two in [8, 4096]", p->nx, p->ny);
  if (p->filter < 0 || p->filter > 3) return fail(NS2D_ERR_INVALID, "unknown filter %d", p->filter);
  if (p->integrator != NS2D_INT_RK4CN && p->integrator != NS2D_INT_AB4)
    return fail(NS2D_ERR_INVALID, "unknown integrator %d", p->integrator);
  if (p->forcing == NS2D_FORCING_KOLM && (p->force_k < 0 || p->force_k > p->ny / 2))
    return fail(NS2D_ERR_INVALID, "Kolmogorov forcing wavenumber %d outside [0, Ny/2]", p->force_k);
  if (p->forcing != NS2D_FORCING_NONE && p->forcing != NS2D_FORCING_KOLM)
    return fail(NS2D_ERR_INVALID, "unsupported forcing %d", p->forcing);

  int ndev = 0;
  HIPCHK(hipGetDeviceCount(&ndev));
  if (p->device < 0 || p->device >= ndev) return fail(NS2D_ERR_INVALID, "no HIP device %d (have %d)", p->device, ndev);
  HIPCHK(hipSetDevice(p->device));

  ns2d_handle* h = new ns2d_handle();
  h->p = *p;
  h->nyf = p->ny / 2 + 1;
  h->ld = (h->nyf + 7) / 8 * 8;
  h->M = (size_t)p->nx * h->ld;
  h->nspect = (int)std::lround(std::sqrt((p->nx / 2.0) * (p->nx / 2.0) + (p->ny / 2.0) * (p->ny / 2.0)) + 1);
  h->hyper_eff = p->hyper_flag;

  PhysArgs& ph = h->ph;
  ph.nx = p->nx;
  ph.ny = p->ny;
  ph.nu = p->nu;
  ph.hyper_pow = p->hyper_pow;
  ph.alpha = p->alpha;
  ph.drag_pow = p->drag_pow;
  ph.filter = p->filter;
  ph.kmax_sqr = std::pow((double)(int)(p->nx / 3.0), 2.0);                  // solver.c:1048
  ph.kolm_kill = (p->forcing == NS2D_FORCING_KOLM) && (p->quirks & NS2D_QUIRK_KOLM_KX0_KILL);
  ph.houli_edge = std::exp(-36.0 * std::pow(std::sqrt(1.0), 36.0));          // solver.c:1088
  ph.houli_corner = std::exp(-36.0 * std::pow(std::sqrt(2.0), 36.0));
  const double nf = 1.0 / (double)((long)p->nx * p->ny);
  ph.norm2 = std::pow(nf, 2.0);                                              // solver.c:1005
  ph.force_on = p->forcing == NS2D_FORCING_KOLM;
  ph.force_k = p->force_k;
  ph.force_val = -0.5 * p->force_scale * (double)p->force_k;                 // force.c:537
  if (!(p->quirks & NS2D_QUIRK_KOLM_UNSCALED)) ph.force_val *= (double)((long)p->nx * p->ny);

  auto cleanup = [&](int rc) {
    ns2d_destroy(h);
    return rc;
  };
  if (ext_stream) {
    h->stream = ext_stream;
    h->own_stream = false;
  } else if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) {
    return cleanup(fail(NS2D_ERR_HIP, "stream create failed"));
  }
  const size_t bytes = h->M * sizeof(double2);
  std::vector<double2**> bufs = {&h->w[0], &h->w[1], &h->acc, &h->g[0], &h->g[1], &h->g[2], &h->g[3],
                                 &h->nbuf, &h->kout, &h->sbuf};
  if (p->integrator == NS2D_INT_AB4)
    for (int i = 0; i < 4; ++i) bufs.push_back(&h->hist[i]);
  for (double2** b : bufs) {
    if (hipMalloc((void**)b, bytes) != hipSuccess)
      return cleanup(fail(NS2D_ERR_HIP, "hipMalloc of %zu bytes failed", bytes));
    if (hipMemsetAsync(*b, 0, bytes, h->stream) != hipSuccess) return cleanup(fail(NS2D_ERR_HIP, "memset failed"));
  }
  for (int i = 0; i < 2; ++i) {
    if (hipMalloc((void**)&h->conj[i], sizeof(double2) * p->nx) != hipSuccess)
      return cleanup(fail(NS2D_ERR_HIP, "hipMalloc failed"));
    (void)hipMemsetAsync(h->conj[i], 0, sizeof(double2) * p->nx, h->stream);
  }
  if (hipMalloc((void**)&h->pairbuf, sizeof(double2) * 8 * p->nx) != hipSuccess)
    return cleanup(fail(NS2D_ERR_HIP, "hipMalloc failed"));
  if (XT_K != 1 && hipMalloc((void**)&h->stage, sizeof(double2) * p->nx * h->nyf) != hipSuccess)
    return cleanup(fail(NS2D_ERR_HIP, "hipMalloc failed"));
  if (hipMalloc((void**)&h->stat, sizeof(unsigned long long) * STAT_N) != hipSuccess)
    return cleanup(fail(NS2D_ERR_HIP, "hipMalloc failed"));
  if (hipMalloc((void**)&h->meas, sizeof(double) * NQ * h->nspect) != hipSuccess)
    return cleanup(fail(NS2D_ERR_HIP, "hipMalloc failed"));
  // twiddles exp(-2 pi i m / NS2D_TW_LEN) in extended precision
  std::vector<double2> tw(NS2D_TW_LEN);
  for (int m = 0; m < NS2D_TW_LEN; ++m) {
    const long double ang = -2.0L * 3.141592653589793238462643383279502884L * (long double)m / (long double)NS2D_TW_LEN;
    tw[m] = make_double2((double)cosl(ang), (double)sinl(ang));
  }
  if (hipMalloc((void**)&h->tw, sizeof(double2) * NS2D_TW_LEN) != hipSuccess)
    return cleanup(fail(NS2D_ERR_HIP, "hipMalloc failed"));
  if (hipMemcpy(h->tw, tw.data(), sizeof(double2) * NS2D_TW_LEN, hipMemcpyHostToDevice) != hipSuccess)
    return cleanup(fail(NS2D_ERR_HIP, "twiddle upload failed"));
  // Hou-Li filter in its real-division form: per-mode table (solver.c:1088)
  ph.ld = h->ld;
  if (p->filter == NS2D_FILTER_HOULI) {
    std::vector<double> t((size_t)p->nx * h->ld, 0.0);
    for (int i = 0; i < p->nx; ++i) {
      const long kx = i <= p->nx / 2 ? i : i - p->nx;
      for (int j = 0; j < h->nyf; ++j)
        t[kidx(i, j, h->ld, p->nx)] = std::exp(-36.0 * std::pow(std::sqrt(std::pow(kx / (p->nx / 2.0), 2.0) +
                                                                        std::pow(j / (p->ny / 2.0), 2.0)), 36.0));
    }
    if (hipMalloc((void**)&h->filt_tab, sizeof(double) * t.size()) != hipSuccess ||
        hipMemcpy(h->filt_tab, t.data(), sizeof(double) * t.size(), hipMemcpyHostToDevice) != hipSuccess)
      return cleanup(fail(NS2D_ERR_HIP, "filter table upload failed"));
    ph.filt_tab = h->filt_tab;
  }
  if (const char* g = std::getenv("NS2D_GRAPH")) h->use_graph = std::atoi(g) != 0;
  // pass shapes: built-in defaults, or NS2D_KPASS_VARIANT_<Nx> / NS2D_RPASS_VARIANT_<Ny>
  {
    h->kvar = kpass_default_variant(p->nx, p->integrator == NS2D_INT_AB4);
    char name[64];
    std::snprintf(name, sizeof name, "NS2D_KPASS_VARIANT_%d", p->nx);
    const char* env = std::getenv(name);
    if (env) {
      const int v = std::atoi(env);
      if (v >= 0 && v < kpass_variant_count(p->nx)) h->kvar = v;
    }
    std::snprintf(name, sizeof name, "NS2D_RPASS_VARIANT_%d", p->ny);
    env = std::getenv(name);
    if (env) {
      const int v = std::atoi(env);
      if (v >= 0 && v < rpass_variant_count(p->ny)) h->rvar = v;
    }
  }
  // coefficient tables of every viscous kind the handle can reach (the
  // reference's flag assignment, solver.c:843, switches the kind once), so
  // that stepping and measuring allocate nothing after creation
  {
    const int hyp[2] = {p->hyper_flag, 1};
    const bool may_switch = (p->quirks & NS2D_QUIRK_VISC_ASSIGN) && p->drag_flag != 1;
    for (int i = 0; i < (may_switch ? 2 : 1); ++i) {
      h->hyper_eff = hyp[i];
      if (!lam_table(h, visc_kind_plain(h)) || !lam_table(h, visc_kind_rk(h)))
        return cleanup(fail(NS2D_ERR_HIP, "coefficient table allocation failed"));
    }
    h->hyper_eff = p->hyper_flag;
  }
  int rc = reset_stat(h);
  if (rc) return cleanup(rc);
  *out = h;
  return NS2D_OK;
}

static void drop_graph(ns2d_handle* h);

int ns2d_destroy(ns2d_t* h) {
  if (!h) return NS2D_OK;
  (void)hipSetDevice(h->p.device);
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  double2* bufs[] = {h->w[0], h->w[1], h->acc, h->g[0], h->g[1], h->g[2], h->g[3], h->nbuf, h->kout, h->sbuf,
                     h->hist[0], h->hist[1], h->hist[2], h->hist[3], h->conj[0], h->conj[1], h->tw, h->pairbuf};
  for (double2* b : bufs)
    if (b) (void)hipFree(b);
  for (double* t : h->lam_tab)
    if (t) (void)hipFree(t);
  if (h->stream2) (void)hipStreamSynchronize(h->stream2);
  double2* adj[] = {h->tape_buf[0], h->tape_buf[1], h->ckpt, h->lamp[0], h->lamp[1], h->lacc, h->lbuf, h->bbuf[0], h->bbuf[1], h->bbuf[2],
                    h->bbuf[3]};
  for (double2* b : adj)
    if (b) (void)hipFree(b);
  if (h->filt_tab) (void)hipFree(h->filt_tab);
  if (h->stat) (void)hipFree(h->stat);
  if (h->meas) (void)hipFree(h->meas);
  if (h->wreal) (void)hipFree(h->wreal);
  if (h->rbuf) (void)hipFree(h->rbuf);
  if (h->stage) (void)hipFree(h->stage);
  drop_graph(h);
  for (auto& g : h->adj_graph)
    if (g.exec) (void)hipGraphExecDestroy(g.exec);
  if (h->ev_fork) (void)hipEventDestroy(h->ev_fork);
  if (h->ev_join) (void)hipEventDestroy(h->ev_join);
  if (h->stream2) (void)hipStreamDestroy(h->stream2);
  if (h->stream && h->own_stream) (void)hipStreamDestroy(h->stream);
  delete h;
  return NS2D_OK;
}

static int set_state_impl(ns2d_t* h, const void* src, hipMemcpyKind kind) {
  if (!h || !src) return fail(NS2D_ERR_INVALID, "null argument");
  HIPCHK(hipSetDevice(h->p.device));
  int rc0 = copy_in(h, h->w[h->wcur], src, kind);
  if (rc0) return rc0;
  h->inv_valid = false;
  h->iters = 0;
  h->hyper_eff = h->p.hyper_flag;
  return reset_stat(h);
}

static int get_state_impl(ns2d_t* h, void* dst, hipMemcpyKind kind) {
  if (!h || !dst) return fail(NS2D_ERR_INVALID, "null argument");
  HIPCHK(hipSetDevice(h->p.device));
  int rc0 = copy_out(h, dst, h->w[h->wcur], kind);
  if (rc0) return rc0;
  HIPCHK(hipStreamSynchronize(h->stream));
  return NS2D_OK;
}

int ns2d_set_state(ns2d_t* h, const double* w_hat) { return set_state_impl(h, w_hat, hipMemcpyHostToDevice); }
int ns2d_get_state(ns2d_t* h, double* w_hat) { return get_state_impl(h, w_hat, hipMemcpyDeviceToHost); }
int ns2d_set_state_device(ns2d_t* h, const void* d) { return set_state_impl(h, d, hipMemcpyDeviceToDevice); }
int ns2d_get_state_device(ns2d_t* h, void* d) { return get_state_impl(h, d, hipMemcpyDeviceToDevice); }

static void drop_graph(ns2d_handle* h) {
  if (h->gexec) (void)hipGraphExecDestroy(h->gexec);
  if (h->graph) (void)hipGraphDestroy(h->graph);
  h->gexec = nullptr;
  h->graph = nullptr;
}

// One direct step of the configured integrator.
static int one_step(ns2d_handle* h, double dt, bool check) {
  return (h->p.integrator == NS2D_INT_AB4) ? ab4_step(h, dt, check) : rk4cn_step(h, dt, check);
}

// Steps repeat with period 2 (RK4CN: the state buffers alternate) or 4 (AB4:
// the four history slots rotate), after which every kernel argument is the
// same again, so one period is captured once as a hipGraph and replayed.  The
// capture holds kernel launches only.
static int graph_steps(ns2d_handle* h, double dt, bool check, int period) {
  const double* lam = lam_table(h, h->p.integrator == NS2D_INT_AB4 ? visc_kind_plain(h) : visc_kind_rk(h));
  if (!lam) return fail(NS2D_ERR_HIP, "coefficient table allocation failed");
  const int ring = h->p.integrator == NS2D_INT_AB4 ? (int)((h->iters + 1 - 4) % 4) : 0;
  const bool same = h->gexec && h->gkey.dt == dt && h->gkey.wcur == h->wcur && h->gkey.cur == h->cur &&
                    h->gkey.ring == ring && h->gkey.kvar == h->kvar && h->gkey.rvar == h->rvar &&
                    h->gkey.check == (int)check && h->gkey.period == period && h->gkey.lam == lam;
  if (!same) {
    drop_graph(h);
    const int cur0 = h->cur, wcur0 = h->wcur, hyp0 = h->hyper_eff;
    const long it0 = h->iters;
    HIPCHK(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
    int rc = NS2D_OK;
    for (int k = 0; k < period && rc == NS2D_OK; ++k) rc = one_step(h, dt, check);
    hipGraph_t g = nullptr;
    const hipError_t ec = hipStreamEndCapture(h->stream, &g);
    h->cur = cur0;
    h->wcur = wcur0;
    h->hyper_eff = hyp0;
    h->iters = it0;
    if (rc) {
      if (g) (void)hipGraphDestroy(g);
      return rc;
    }
    if (ec != hipSuccess) return fail(NS2D_ERR_HIP, "graph capture: %s", hipGetErrorString(ec));
    h->graph = g;
    HIPCHK(hipGraphInstantiate(&h->gexec, g, nullptr, nullptr, 0));
    h->gkey.dt = dt;
    h->gkey.wcur = h->wcur;
    h->gkey.cur = h->cur;
    h->gkey.ring = ring;
    h->gkey.kvar = h->kvar;
    h->gkey.rvar = h->rvar;
    h->gkey.check = check;
    h->gkey.period = period;
    h->gkey.lam = lam;
  }
  HIPCHK(hipGraphLaunch(h->gexec, h->stream));
  h->iters += period;  // cur and wcur are back to their values after a period
  return NS2D_OK;
}

static int step_impl(ns2d_t* h, double dt, long nsteps) {
  if (!h) return fail(NS2D_ERR_INVALID, "null handle");
  if (nsteps < 0) return fail(NS2D_ERR_INVALID, "nsteps < 0");
  if (!(dt > 0.0)) return fail(NS2D_ERR_INVALID, "dt must be positive");
  HIPCHK(hipSetDevice(h->p.device));
  if (nsteps == 0) return NS2D_OK;
  const bool check = h->p.check_every != 0;
  int rc = prepare_inv(h);
  if (rc) return rc;
  const bool ab4 = h->p.integrator == NS2D_INT_AB4;
  const int period = ab4 ? 4 : 2;
  long s = 0;
  while (s < nsteps) {
    // graphs only once the integrator's arguments are periodic: after the AB4
    // start-up and after the viscous-flag update of the first RK4CN step
    const bool settled = !(ab4 && h->iters < 3) &&
                         !((h->p.quirks & NS2D_QUIRK_VISC_ASSIGN) && h->p.drag_flag != 1 && h->hyper_eff != 1);
    if (h->use_graph && !h->profiling && settled && nsteps - s >= period) {
      rc = graph_steps(h, dt, check, period);
      s += period;
    } else {
      rc = one_step(h, dt, check);
      s += 1;
    }
    if (rc) return rc;
  }
  // the state's transforms are prepared by the last k-space pass
  h->inv_valid = true;
  if (check) {
    // SystemCheck on the final state (solver.c:2309)
    ev_begin(h, KIND_OTHER);
    RPassArgs a = rbase(h, RP_W);
    hipError_t e = launch_rpass(h->p.nx, h->p.ny, h->rvar, a, h->stream);
    ev_end(h);
    if (e != hipSuccess) return fail(NS2D_ERR_HIP, "rpass launch: %s", hipGetErrorString(e));
    unsigned long long st[STAT_N];
    rc = read_stat(h, st);
    if (rc) return rc;
    double m;
    std::memcpy(&m, &st[STAT_WMAX_CUR], sizeof m);
    if (!(m < MAX_VORT_LIM) && st[STAT_FIRST_BAD] == ~0ull) st[STAT_FIRST_BAD] = st[STAT_ITER];
    st[STAT_WMAX_LAST] = st[STAT_WMAX_CUR];
    st[STAT_WMAX_CUR] = 0ull;
    HIPCHK(hipMemcpyAsync(h->stat, st, sizeof st, hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    if (st[STAT_FIRST_BAD] != ~0ull)
      return fail(NS2D_ERR_BLOWUP, "System has reached maximum Vorticity limit at Iter: [%llu]",
                  (unsigned long long)st[STAT_FIRST_BAD]);
  }
  return NS2D_OK;
}

int ns2d_step(ns2d_t* h, double dt, long nsteps) { return step_impl(h, dt, nsteps); }

int ns2d_rhs(ns2d_t* h, const double* in, double* out) {
  if (!h || !in || !out) return fail(NS2D_ERR_INVALID, "null argument");
  HIPCHK(hipSetDevice(h->p.device));
  int rc = copy_in(h, h->sbuf, in, hipMemcpyHostToDevice);
  if (rc) return rc;
  KPassArgs a = kbase(h);
  a.mode = KP_INV;
  a.epi = EPI_FROM_INPUT;
  a.sin = h->sbuf;
  a.conj_dst = h->conj[h->cur];
  rc = run_kpass(h, a);
  if (rc) return rc;
  h->inv_valid = false;  // G1..G4 now hold the transforms of `in`
  rc = finish_rhs_into_kout(h, false);
  if (rc) return rc;
  rc = copy_out(h, out, h->kout, hipMemcpyDeviceToHost);
  if (rc) return rc;
  HIPCHK(hipStreamSynchronize(h->stream));
  return NS2D_OK;
}

}  // extern "C"

namespace {

// Measures kernel arguments for the handle's state; with_flux: the RHS is in kout.
MeasArgs measures_args(ns2d_handle* h, int with_flux, const double* lam, double* out) {
  MeasArgs a{};
  a.w_hat = h->w[h->wcur];
  a.rhs = with_flux ? h->kout : nullptr;
  a.out = out;
  a.nx = h->p.nx;
  a.ny = h->p.ny;
  a.ld = h->ld;
  a.lam = lam;
  a.quirk_enst_diss = (h->p.quirks & NS2D_QUIRK_ENST_DISS) ? 1 : 0;
  a.nbins = h->nspect;
  a.sbst_lo = 0.0;                   // LWR_SBST_LIM^2, data_types.h:175
  a.sbst_hi = 64.0 * 64.0;           // UPR_SBST_LIM^2, data_types.h:174
  a.ph = h->ph;
  return a;
}

// 4 pi^2 * 0.5 / (Nx Ny)^2 (sys_msr.c:43-44)
double measures_cn(const ns2d_handle* h) {
  const double norm_fac = 0.5 / std::pow((double)((long)h->p.nx * h->p.ny), 2.0);
  const double const_fac = 4.0 * std::pow(M_PI, 2.0);
  return const_fac * norm_fac;
}

// The reported totals from the bin-order sums tot[NQ] (sys_msr.c:340-372);
// wf is the state at the forced mode.
void measures_from_totals(const ns2d_handle* h, const double* tot, double2 wf, int with_flux, ns2d_measures_t* m) {
  const double cn = measures_cn(h);
  std::memset(m, 0, sizeof *m);
  m->energy = tot[Q_ENRG] * cn;
  m->enstrophy = tot[Q_ENST] * cn;
  m->palinstrophy = tot[Q_PALIN] * cn;
  m->divergence = tot[Q_DIV] * cn;
  m->enrg_diss = tot[Q_ENRG_DISS] * 2.0 * cn;
  m->enst_diss = tot[Q_ENST_DISS] * 2.0 * cn;
  if (h->ph.force_on) m->forcing = std::hypot(h->ph.force_val * wf.x, h->ph.force_val * wf.y) * cn;
  if (with_flux) {
    // sys_msr.c:200-225, normalised at :363-372 (the d/dt terms are not)
    m->d_enrg_dt_sbst = tot[Q_SB_DERIV_K2] + tot[Q_SB_DISS_K2];
    m->enrg_flux_sbst = tot[Q_SB_DERIV_K2] * cn;
    m->enrg_diss_sbst = tot[Q_SB_DISS_K2] * 2.0 * cn;
    m->d_enst_dt_sbst = tot[Q_SB_DERIV] + tot[Q_SB_DISS];
    m->enst_flux_sbst = tot[Q_SB_DERIV] * cn;
    m->enst_diss_sbst = tot[Q_SB_DISS] * 2.0 * cn;
  }
}

// sys_msr.c:125-133: NonlinearRHSBatch(w_hat) (with its ForceConjugacy write
// into w_hat) into kout; the forcing is subtracted inside the reduction
int measures_flux_rhs(ns2d_handle* h) {
  int rc = prepare_inv(h);
  if (rc) return rc;
  rc = finish_rhs_into_kout(h, false);
  if (rc) return rc;
  if (h->p.quirks & NS2D_QUIRK_CONJ_GLOBAL) h->inv_valid = false;
  return NS2D_OK;
}

}  // namespace

extern "C" {

int ns2d_measures(ns2d_t* h, int with_flux, ns2d_measures_t* m, double* enrg_spec, double* enst_spec,
                  double* flux_spec) {
  if (!h) return fail(NS2D_ERR_INVALID, "null handle");
  HIPCHK(hipSetDevice(h->p.device));
  int rc;
  if (with_flux) {
    rc = measures_flux_rhs(h);
    if (rc) return rc;
  }
  const double* lam = lam_table(h, visc_kind_plain(h));  // sys_msr.c:160-175
  if (!lam) return fail(NS2D_ERR_HIP, "coefficient table allocation failed");
  const MeasArgs a = measures_args(h, with_flux, lam, h->meas);
  hipError_t e = launch_measures(h->nspect, a, h->stream);
  if (e != hipSuccess) return fail(NS2D_ERR_HIP, "measures launch: %s", hipGetErrorString(e));
  std::vector<double> part((size_t)NQ * h->nspect);
  HIPCHK(hipMemcpyAsync(part.data(), h->meas, sizeof(double) * part.size(), hipMemcpyDeviceToHost, h->stream));
  double2 wf = make_double2(0.0, 0.0);
  if (h->ph.force_on)
    HIPCHK(hipMemcpyAsync(&wf, h->w[h->wcur] + kidx(0, h->p.force_k, h->ld, h->p.nx), sizeof wf, hipMemcpyDeviceToHost,
                          h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  const double cn = measures_cn(h);
  double tot[NQ] = {0};
  for (int b = 0; b < h->nspect; ++b)
    for (int q = 0; q < NQ; ++q) tot[q] += part[(size_t)b * NQ + q];
  if (m) measures_from_totals(h, tot, wf, with_flux, m);
  for (int b = 0; b < h->nspect; ++b) {
    const double* q = &part[(size_t)b * NQ];
    if (enrg_spec) enrg_spec[b] = q[Q_ENRG] * cn;
    if (enst_spec) enst_spec[b] = q[Q_ENST] * cn;
    if (flux_spec && with_flux) {
      const int n = h->nspect;
      flux_spec[0 * n + b] = (q[Q_DERIV] + q[Q_DISS]) * cn;
      flux_spec[1 * n + b] = q[Q_DISS] * cn;
      flux_spec[2 * n + b] = q[Q_DERIV] * cn;
      flux_spec[3 * n + b] = (q[Q_DERIV_K2] + q[Q_DISS_K2]) * cn;
      flux_spec[4 * n + b] = q[Q_DISS_K2] * cn;
      flux_spec[5 * n + b] = q[Q_DERIV_K2] * cn;
    }
  }
  return NS2D_OK;
}

// ns2d_measures without device allocation: partial sums, spectra and totals
// in the library's per-device scratch array (ns2d_noalloc.hip), the
// coefficient table built by ns2d_create.  Results equal ns2d_measures' bit
// for bit.  Calls that overlap on one GPU share the scratch and race.
int ns2d_measures_noalloc(ns2d_t* h, int with_flux, ns2d_measures_t* m, double* enrg_spec, double* enst_spec,
                          double* flux_spec) {
  if (!h) return fail(NS2D_ERR_INVALID, "null handle");
  if (h->nspect > NOALLOC_MAX_BINS)
    return fail(NS2D_ERR_INVALID, "spectrum of %d bins exceeds the allocation-free limit of %d bins (4096 x 4096)",
                h->nspect, NOALLOC_MAX_BINS);
  HIPCHK(hipSetDevice(h->p.device));
  double* scr = nullptr;
  HIPCHK(noalloc_scratch(h->p.device, &scr));
  const double* lam = h->lam_tab[visc_kind_plain(h)];
  if (!lam) return fail(NS2D_ERR_STATE, "coefficient table not built");
  int rc;
  if (with_flux) {
    rc = measures_flux_rhs(h);
    if (rc) return rc;
  }
  const int nb = h->nspect;
  double* out = scr + (size_t)NQ * NOALLOC_MAX_BINS;
  const MeasArgs a = measures_args(h, with_flux, lam, scr);
  hipError_t e = launch_measures(nb, a, h->stream);
  if (e != hipSuccess) return fail(NS2D_ERR_HIP, "measures launch: %s", hipGetErrorString(e));
  e = launch_measures_finalize(scr, nb, measures_cn(h), with_flux, out, h->stream);
  if (e != hipSuccess) return fail(NS2D_ERR_HIP, "measures finalize launch: %s", hipGetErrorString(e));
  double tot[NQ];
  HIPCHK(hipMemcpyAsync(tot, out + 8 * (size_t)nb, sizeof tot, hipMemcpyDeviceToHost, h->stream));
  if (enrg_spec) HIPCHK(hipMemcpyAsync(enrg_spec, out, sizeof(double) * nb, hipMemcpyDeviceToHost, h->stream));
  if (enst_spec) HIPCHK(hipMemcpyAsync(enst_spec, out + nb, sizeof(double) * nb, hipMemcpyDeviceToHost, h->stream));
  if (flux_spec && with_flux)
    HIPCHK(hipMemcpyAsync(flux_spec, out + 2 * (size_t)nb, sizeof(double) * 6 * nb, hipMemcpyDeviceToHost, h->stream));
  double2 wf = make_double2(0.0, 0.0);
  if (h->ph.force_on)
    HIPCHK(hipMemcpyAsync(&wf, h->w[h->wcur] + kidx(0, h->p.force_k, h->ld, h->p.nx), sizeof wf, hipMemcpyDeviceToHost,
                          h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  if (m) measures_from_totals(h, tot, wf, with_flux, m);
  return NS2D_OK;
}

int ns2d_max_vort(ns2d_t* h, double* out) {
  if (!h || !out) return fail(NS2D_ERR_INVALID, "null argument");
  HIPCHK(hipSetDevice(h->p.device));
  return state_wmax(h, out, nullptr);
}

int ns2d_cfl_rate(ns2d_t* h, double* conv) {
  if (!h || !conv) return fail(NS2D_ERR_INVALID, "null argument");
  HIPCHK(hipSetDevice(h->p.device));
  return state_cfl(h, conv);
}

int ns2d_real_vort(ns2d_t* h, double* w) {
  if (!h || !w) return fail(NS2D_ERR_INVALID, "null argument");
  HIPCHK(hipSetDevice(h->p.device));
  const size_t n = (size_t)h->p.nx * h->p.ny;
  if (!h->wreal) HIPCHK(hipMalloc((void**)&h->wreal, sizeof(double) * n));
  int rc = state_wmax(h, nullptr, h->wreal);
  if (rc) return rc;
  HIPCHK(hipMemcpyAsync(w, h->wreal, sizeof(double) * n, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  return NS2D_OK;
}

int ns2d_r2c(ns2d_t* h, const double* w, double* w_hat) {
  if (!h || !w || !w_hat) return fail(NS2D_ERR_INVALID, "null argument");
  HIPCHK(hipSetDevice(h->p.device));
  const size_t n = (size_t)h->p.nx * h->p.ny;
  if (!h->rbuf) HIPCHK(hipMalloc((void**)&h->rbuf, sizeof(double) * n));
  HIPCHK(hipMemcpyAsync(h->rbuf, w, sizeof(double) * n, hipMemcpyHostToDevice, h->stream));
  hipError_t e = launch_r2c_rows(h->p.nx, h->p.ny, h->rbuf, h->nbuf, h->tw, h->ld, h->stream);
  if (e != hipSuccess) return fail(NS2D_ERR_HIP, "r2c launch: %s", hipGetErrorString(e));
  KPassArgs a = kbase(h);
  a.mode = KP_FWD | KP_RAW;
  a.epi = EPI_RHS;
  int rc = run_kpass(h, a);
  if (rc) return rc;
  rc = copy_out(h, w_hat, h->kout, hipMemcpyDeviceToHost);
  if (rc) return rc;
  HIPCHK(hipStreamSynchronize(h->stream));
  return NS2D_OK;
}

int ns2d_c2r(ns2d_t* h, const double* w_hat, double* w) {
  if (!h || !w_hat || !w) return fail(NS2D_ERR_INVALID, "null argument");
  HIPCHK(hipSetDevice(h->p.device));
  const size_t n = (size_t)h->p.nx * h->p.ny;
  if (!h->wreal) HIPCHK(hipMalloc((void**)&h->wreal, sizeof(double) * n));
  int rc = copy_in(h, h->sbuf, w_hat, hipMemcpyHostToDevice);
  if (rc) return rc;
  KPassArgs a = kbase(h);
  a.mode = KP_INV;
  a.epi = EPI_FROM_INPUT;
  a.sin = h->sbuf;
  a.conj_dst = nullptr;  // the state's ky=0 hand-off is left alone
  rc = run_kpass(h, a);
  if (rc) return rc;
  h->inv_valid = false;  // G1..G4 now hold the transforms of w_hat
  RPassArgs r = rbase(h, RP_W);
  r.wreal = h->wreal;
  r.wscale = 1.0;
  rc = run_rpass(h, r);
  if (rc) return rc;
  // the row pass also folded max|w| into the step check's accumulator
  const unsigned long long zero = 0ull;
  HIPCHK(hipMemcpyAsync(h->stat + STAT_WMAX_CUR, &zero, sizeof zero, hipMemcpyHostToDevice, h->stream));
  HIPCHK(hipMemcpyAsync(w, h->wreal, sizeof(double) * n, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  return NS2D_OK;
}

int ns2d_get_info(ns2d_t* h, ns2d_info_t* info) {
  if (!h || !info) return fail(NS2D_ERR_INVALID, "null argument");
  HIPCHK(hipSetDevice(h->p.device));
  unsigned long long st[STAT_N];
  int rc = read_stat(h, st);
  if (rc) return rc;
  info->iters = h->iters;
  info->first_blowup_iter = st[STAT_FIRST_BAD] == ~0ull ? -1 : (long)st[STAT_FIRST_BAD];
  std::memcpy(&info->last_max_vort, &st[STAT_WMAX_LAST], sizeof(double));
  info->hyper_flag = h->hyper_eff;
  info->n_spect = h->nspect;
  info->ld = h->ld;
  return NS2D_OK;
}

int ns2d_sync(ns2d_t* h) {
  if (!h) return fail(NS2D_ERR_INVALID, "null handle");
  HIPCHK(hipSetDevice(h->p.device));
  HIPCHK(hipStreamSynchronize(h->stream));
  return NS2D_OK;
}

int ns2d_get_stream(ns2d_t* h, void** stream) {
  if (!h || !stream) return fail(NS2D_ERR_INVALID, "null argument");
  *stream = (void*)h->stream;
  return NS2D_OK;
}

int ns2d_profile_steps(ns2d_t* h, double dt, long nsteps, double* kpass_ms, double* rpass_ms, double* total_ms) {
  if (!h) return fail(NS2D_ERR_INVALID, "null handle");
  HIPCHK(hipSetDevice(h->p.device));
  HIPCHK(hipStreamSynchronize(h->stream));
  h->profiling = true;
  h->ev.clear();
  h->ev_kind.clear();
  hipEvent_t t0, t1;
  HIPCHK(hipEventCreate(&t0));
  HIPCHK(hipEventCreate(&t1));
  HIPCHK(hipEventRecord(t0, h->stream));
  int rc = step_impl(h, dt, nsteps);
  HIPCHK(hipEventRecord(t1, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  h->profiling = false;
  double ks = 0, rs = 0;
  long kn = 0, rn = 0;
  for (size_t i = 0; i + 1 < h->ev.size(); i += 2) {
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, h->ev[i], h->ev[i + 1]);
    if (h->ev_kind[i] == KIND_K) { ks += ms; ++kn; }
    if (h->ev_kind[i] == KIND_R) { rs += ms; ++rn; }
  }
  for (hipEvent_t e : h->ev) (void)hipEventDestroy(e);
  h->ev.clear();
  h->ev_kind.clear();
  float tot = 0.f;
  (void)hipEventElapsedTime(&tot, t0, t1);
  (void)hipEventDestroy(t0);
  (void)hipEventDestroy(t1);
  if (kpass_ms) *kpass_ms = kn ? ks / kn : 0.0;
  if (rpass_ms) *rpass_ms = rn ? rs / rn : 0.0;
  if (total_ms) *total_ms = tot;
  return rc;
}


// ---------------------------------------------------------------------------
// Ensemble: many independent solvers of possibly different sizes stepped
// together.  Members of one grid size form a class; every pass of a stage is
// one batched launch per class reading the members' arguments from a device
// table, and one period of two steps (all classes) is replayed as a hipGraph.
// ---------------------------------------------------------------------------
}  // extern "C"

struct ns2d_ensemble {
  hipStream_t stream = nullptr;
  std::vector<ns2d_handle*> m;
  std::vector<double> dt;
  struct Class {
    int nx = 0, ny = 0, nyf = 0, kvar = 0, rvar = 0;
    std::vector<int> idx;
    KPassArgs* dk = nullptr;   // [parity 2][stage 4][members]
    RPassArgs* dr = nullptr;   // [stage 4][members]
    hipStream_t stream = nullptr;  // the class's launches run concurrently
    hipEvent_t done = nullptr;
  };
  std::vector<Class> cls;
  hipEvent_t fork = nullptr;
  bool tuned = false;
  bool tables_ok = false;
  int tables_parity = -1;
  bool tables_check = false;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  int graph_parity = -1;
};

namespace {

void ens_drop(ns2d_ensemble* e) {
  if (e->gexec) (void)hipGraphExecDestroy(e->gexec);
  if (e->graph) (void)hipGraphDestroy(e->graph);
  e->gexec = nullptr;
  e->graph = nullptr;
  e->graph_parity = -1;
}

// (Re)build the per-class argument tables for the members' current buffers.
int ens_tables(ns2d_ensemble* e, bool check) {
  const int par = e->m[0]->wcur;
  if (e->tables_ok && e->tables_parity == par && e->tables_check == check) return NS2D_OK;
  ens_drop(e);
  for (auto& c : e->cls) {
    const size_t nm = c.idx.size();
    std::vector<KPassArgs> hk(2 * 4 * nm);
    std::vector<RPassArgs> hr(4 * nm);
    for (size_t j = 0; j < nm; ++j) {
      ns2d_handle* h = e->m[c.idx[j]];
      for (int p2 = 0; p2 < 2; ++p2) {
        const int w0 = h->wcur;
        h->wcur = par ^ p2;
        for (int st = 0; st < 4; ++st) {
          RPassArgs ra;
          KPassArgs ka;
          int rc = rk4cn_stage_args(h, e->dt[c.idx[j]], check, st, &ra, &ka);
          h->wcur = w0;
          if (rc) return rc;
          h->wcur = par ^ p2;
          hk[((size_t)p2 * 4 + st) * nm + j] = ka;
          if (p2 == 0) hr[(size_t)st * nm + j] = ra;
        }
        h->wcur = w0;
      }
    }
    if (!c.dk) HIPCHK(hipMalloc((void**)&c.dk, sizeof(KPassArgs) * hk.size()));
    if (!c.dr) HIPCHK(hipMalloc((void**)&c.dr, sizeof(RPassArgs) * hr.size()));
    HIPCHK(hipMemcpy(c.dk, hk.data(), sizeof(KPassArgs) * hk.size(), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c.dr, hr.data(), sizeof(RPassArgs) * hr.size(), hipMemcpyHostToDevice));
  }
  e->tables_ok = true;
  e->tables_parity = par;
  e->tables_check = check;
  return NS2D_OK;
}

// One batched RK4CN step of every member; p2 = parity relative to the tables.
// Size classes are independent: each runs its stages on its own stream,
// forked from and joined back into the ensemble stream (also under capture,
// where the fork/join becomes parallel branches of the graph).
int ens_launch_step(ns2d_ensemble* e, int p2) {
  const bool fan = e->cls.size() > 1;
  if (fan) HIPCHK(hipEventRecord(e->fork, e->stream));
  for (auto& c : e->cls) {
    hipStream_t cs = fan ? c.stream : e->stream;
    if (fan) HIPCHK(hipStreamWaitEvent(cs, e->fork, 0));
    for (int st = 0; st < 4; ++st) {
      hipError_t r = launch_rpass_batched(c.nx, c.ny, c.rvar, c.dr + (size_t)st * c.idx.size(), (int)c.idx.size(), cs);
      if (r != hipSuccess) return fail(NS2D_ERR_HIP, "batched rpass: %s", hipGetErrorString(r));
      r = launch_kpass_batched(c.nx, c.nyf, c.kvar, c.dk + ((size_t)p2 * 4 + st) * c.idx.size(), (int)c.idx.size(),
                               cs);
      if (r != hipSuccess) return fail(NS2D_ERR_HIP, "batched kpass: %s", hipGetErrorString(r));
    }
    if (fan) {
      HIPCHK(hipEventRecord(c.done, cs));
      HIPCHK(hipStreamWaitEvent(e->stream, c.done, 0));
    }
  }
  return NS2D_OK;
}

// The project's autotuner for the batched passes: per size class, time the
// stage-1 row pass and column pass of every compiled shape over all members
// of the class (HIP events, best of `reps`) and keep the fastest.  The column
// pass of stage 1 overwrites the members' transformed fields, which are
// rebuilt from their states afterwards; the states themselves are not touched.
int ens_autotune(ns2d_ensemble* e, int reps) {
  int rc = ens_tables(e, e->m[0]->p.check_every != 0);
  if (rc) return rc;
  // the step-check words (iteration count, blow-up record) survive the timing
  std::vector<unsigned long long> saved((size_t)STAT_N * e->m.size());
  for (size_t i = 0; i < e->m.size(); ++i) {
    rc = read_stat(e->m[i], &saved[i * STAT_N]);
    if (rc) return rc;
  }
  hipEvent_t t0 = nullptr, t1 = nullptr;
  if (hipEventCreate(&t0) != hipSuccess || hipEventCreate(&t1) != hipSuccess) {
    if (t0) (void)hipEventDestroy(t0);
    return fail(NS2D_ERR_HIP, "event create failed");
  }
  auto time_it = [&](auto&& launch) -> float {
    float best = 1e30f;
    for (int r = 0; r < reps; ++r) {
      if (hipEventRecord(t0, e->stream) != hipSuccess) return -1.0f;
      if (launch() != hipSuccess) return -1.0f;
      if (hipEventRecord(t1, e->stream) != hipSuccess || hipEventSynchronize(t1) != hipSuccess) return -1.0f;
      float ms = 0.0f;
      if (hipEventElapsedTime(&ms, t0, t1) != hipSuccess) return -1.0f;
      best = ms < best ? ms : best;
    }
    return best;
  };
  // each shape is timed over the four RK4CN stages of one step (the stage
  // epilogues read and write different arrays); the timing launches write only
  // scratch (accumulator, G fields, the next state buffer, the ky=0 hand-off),
  // which the rebuild below restores
  for (auto& c : e->cls) {
    const int nm = (int)c.idx.size();
    auto step_ms = [&](auto&& launch_stage) -> float {
      float sum = 0.0f;
      for (int st = 0; st < 4; ++st) {
        const float ms = time_it([&] { return launch_stage(st); });
        if (ms < 0.0f) return -1.0f;
        sum += ms;
      }
      return sum;
    };
    float best = 1e30f;
    int arg = c.rvar;
    for (int v = 0; v < rpass_variant_count(c.ny); ++v) {
      const float ms = step_ms([&](int st) {
        return launch_rpass_batched(c.nx, c.ny, v, c.dr + (size_t)st * nm, nm, e->stream);
      });
      if (ms >= 0.0f && ms < best) {
        best = ms;
        arg = v;
      }
    }
    c.rvar = arg;
    best = 1e30f;
    arg = c.kvar;
    for (int v = 0; v < kpass_variant_count(c.nx); ++v) {
      const float ms = step_ms([&](int st) {
        return launch_kpass_batched(c.nx, c.nyf, v, c.dk + (size_t)st * nm, nm, e->stream);
      });
      if (ms >= 0.0f && ms < best) {
        best = ms;
        arg = v;
      }
    }
    c.kvar = arg;
  }
  (void)hipEventDestroy(t0);
  (void)hipEventDestroy(t1);
  HIPCHK(hipGetLastError());
  // rebuild what the timing runs overwrote
  for (size_t i = 0; i < e->m.size(); ++i) {
    ns2d_handle* h = e->m[i];
    h->inv_valid = false;
    rc = prepare_inv(h);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(h->stat, &saved[i * STAT_N], sizeof(unsigned long long) * STAT_N, hipMemcpyHostToDevice,
                          h->stream));
  }
  HIPCHK(hipStreamSynchronize(e->stream));
  ens_drop(e);
  e->tuned = true;
  return NS2D_OK;
}

// ---- tuned schedule cache (JSON Lines, one record per size class) --------
// {"device": "<arch>/<name>/<CUs>", "tables": "<shape-table hash>", "nx": .., "ny": ..,
//  "members": .., "kpass": <shape index>, "rpass": <shape index>}
// File: NS2D_ENS_CACHE (empty: no cache), else $XDG_CACHE_HOME/ns2d/ens_schedule.jsonl
// or $HOME/.cache/ns2d/ens_schedule.jsonl.  The last matching record wins.
std::string ens_cache_path() {
  if (const char* p = std::getenv("NS2D_ENS_CACHE")) return p;
  std::string base;
  if (const char* x = std::getenv("XDG_CACHE_HOME"); x && *x) base = x;
  else if (const char* h = std::getenv("HOME"); h && *h) base = std::string(h) + "/.cache";
  return base.empty() ? std::string() : base + "/ns2d/ens_schedule.jsonl";
}

std::string ens_device_key(int dev) {
  hipDeviceProp_t pr{};
  if (hipGetDeviceProperties(&pr, dev) != hipSuccess) return "unknown";
  char buf[512];
  std::snprintf(buf, sizeof buf, "%s/%s/%d", pr.gcnArchName, pr.name, pr.multiProcessorCount);
  for (char* c = buf; *c; ++c)
    if (*c == '"' || *c == '\\') *c = '_';
  return buf;
}

// value of "key": in a record line (string without quotes, or a number)
bool jl_field(const std::string& line, const char* key, std::string* out) {
  const std::string k = std::string("\"") + key + "\":";
  size_t p = line.find(k);
  if (p == std::string::npos) return false;
  p += k.size();
  while (p < line.size() && line[p] == ' ') ++p;
  if (p < line.size() && line[p] == '"') {
    const size_t q = line.find('"', p + 1);
    if (q == std::string::npos) return false;
    *out = line.substr(p + 1, q - p - 1);
  } else {
    size_t q = p;
    while (q < line.size() && (line[q] == '-' || (line[q] >= '0' && line[q] <= '9'))) ++q;
    if (q == p) return false;
    *out = line.substr(p, q - p);
  }
  return true;
}

std::string ens_tables_key(const ns2d_ensemble::Class& c) {
  char buf[32];
  std::snprintf(buf, sizeof buf, "%016llx", shape_table_hash(c.nx, c.ny));
  return buf;
}

// Set the classes' shapes from the cache; true when every class was found.
bool ens_cache_load(ns2d_ensemble* e) {
  const std::string path = ens_cache_path();
  if (path.empty()) return false;
  FILE* f = std::fopen(path.c_str(), "r");
  if (!f) return false;
  const std::string dev = ens_device_key(e->m[0]->p.device);
  std::vector<int> kv(e->cls.size(), -1), rv(e->cls.size(), -1);
  char buf[1024];
  while (std::fgets(buf, sizeof buf, f)) {
    const std::string line(buf);
    std::string d, t, nx, ny, nm, k, r;
    if (!jl_field(line, "device", &d) || !jl_field(line, "tables", &t) || !jl_field(line, "nx", &nx) ||
        !jl_field(line, "ny", &ny) || !jl_field(line, "members", &nm) || !jl_field(line, "kpass", &k) ||
        !jl_field(line, "rpass", &r) || d != dev)
      continue;
    for (size_t i = 0; i < e->cls.size(); ++i) {
      const auto& c = e->cls[i];
      if (std::atoi(nx.c_str()) == c.nx && std::atoi(ny.c_str()) == c.ny &&
          std::atoi(nm.c_str()) == (int)c.idx.size() && t == ens_tables_key(c)) {
        const int ki = std::atoi(k.c_str()), ri = std::atoi(r.c_str());
        if (ki >= 0 && ki < kpass_variant_count(c.nx) && ri >= 0 && ri < rpass_variant_count(c.ny)) {
          kv[i] = ki;
          rv[i] = ri;
        }
      }
    }
  }
  std::fclose(f);
  for (size_t i = 0; i < e->cls.size(); ++i)
    if (kv[i] < 0) return false;
  for (size_t i = 0; i < e->cls.size(); ++i) {
    e->cls[i].kvar = kv[i];
    e->cls[i].rvar = rv[i];
  }
  return true;
}

// Append the classes' tuned shapes to the cache (best effort: an unwritable
// cache is skipped).
void ens_cache_store(ns2d_ensemble* e) {
  const std::string path = ens_cache_path();
  if (path.empty()) return;
  for (size_t p = 1; p < path.size(); ++p)
    if (path[p] == '/') (void)::mkdir(path.substr(0, p).c_str(), 0755);
  FILE* f = std::fopen(path.c_str(), "a");
  if (!f) return;
  const std::string dev = ens_device_key(e->m[0]->p.device);
  for (const auto& c : e->cls)
    std::fprintf(f,
                 "{\"device\": \"%s\", \"tables\": \"%s\", \"nx\": %d, \"ny\": %d, \"members\": %d, "
                 "\"kpass\": %d, \"rpass\": %d}\n",
                 dev.c_str(), ens_tables_key(c).c_str(), c.nx, c.ny, (int)c.idx.size(), c.kvar, c.rvar);
  std::fclose(f);
}

void ens_advance(ns2d_ensemble* e) {
  for (ns2d_handle* h : e->m) {
    after_rk_update(h);
    h->iters += 1;
    h->wcur ^= 1;
  }
}

}  // namespace

extern "C" {

int ns2d_ens_create(const ns2d_params* members, int n, ns2d_ens_t** out) {
  if (!members || !out || n <= 0) return fail(NS2D_ERR_INVALID, "bad ensemble arguments");
  *out = nullptr;
  for (int i = 0; i < n; ++i)
    if (members[i].integrator != NS2D_INT_RK4CN)
      return fail(NS2D_ERR_INVALID, "ensemble member %d: only RK4CN members are batched", i);
  HIPCHK(hipSetDevice(members[0].device));
  ns2d_ensemble* e = new ns2d_ensemble();
  if (hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&e->fork, hipEventDisableTiming) != hipSuccess) {
    if (e->stream) (void)hipStreamDestroy(e->stream);
    delete e;
    return fail(NS2D_ERR_HIP, "stream create failed");
  }
  for (int i = 0; i < n; ++i) {
    ns2d_t* h = nullptr;
    int rc = create_impl(&members[i], e->stream, &h);
    if (rc) {
      ns2d_ens_destroy(e);
      return rc;
    }
    e->m.push_back(h);
    e->dt.push_back(0.0);
    int ci = -1;
    for (size_t c = 0; c < e->cls.size(); ++c)
      if (e->cls[c].nx == members[i].nx && e->cls[c].ny == members[i].ny) ci = (int)c;
    if (ci < 0) {
      ns2d_ensemble::Class c;
      c.nx = members[i].nx;
      c.ny = members[i].ny;
      c.nyf = h->nyf;
      c.kvar = h->kvar;
      c.rvar = h->rvar;
      if (hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking) != hipSuccess ||
          hipEventCreateWithFlags(&c.done, hipEventDisableTiming) != hipSuccess) {
        e->cls.push_back(c);
        ns2d_ens_destroy(e);
        return fail(NS2D_ERR_HIP, "ensemble stream create failed");
      }
      e->cls.push_back(c);
      ci = (int)e->cls.size() - 1;
    }
    e->cls[ci].idx.push_back(i);
  }
  *out = e;
  return NS2D_OK;
}

int ns2d_ens_member(ns2d_ens_t* e, int i, ns2d_t** h) {
  if (!e || !h || i < 0 || i >= (int)e->m.size()) return fail(NS2D_ERR_INVALID, "no ensemble member %d", i);
  *h = e->m[i];
  return NS2D_OK;
}

int ns2d_ens_step(ns2d_ens_t* e, const double* dt, long nsteps) {
  if (!e || !dt || nsteps < 0) return fail(NS2D_ERR_INVALID, "bad ensemble step arguments");
  HIPCHK(hipSetDevice(e->m[0]->p.device));
  bool dt_changed = false;
  for (size_t i = 0; i < e->m.size(); ++i) {
    if (!(dt[i] > 0.0)) return fail(NS2D_ERR_INVALID, "member %zu: dt must be positive", i);
    if (dt[i] != e->dt[i]) dt_changed = true;
    e->dt[i] = dt[i];
  }
  if (dt_changed) e->tables_ok = false;
  if (nsteps == 0) return NS2D_OK;
  const bool check = e->m[0]->p.check_every != 0;
  for (ns2d_handle* h : e->m) {
    int rc = prepare_inv(h);
    if (rc) return rc;
    if (h->wcur != e->m[0]->wcur) return fail(NS2D_ERR_STATE, "ensemble members out of step");
  }
  long s = 0;
  // first steps one member at a time until the viscous flags have settled (Q3)
  while (s < nsteps) {
    bool settled = true;
    for (ns2d_handle* h : e->m)
      settled = settled && !((h->p.quirks & NS2D_QUIRK_VISC_ASSIGN) && h->p.drag_flag != 1 && h->hyper_eff != 1);
    if (settled) break;
    for (size_t i = 0; i < e->m.size(); ++i) {
      int rc = rk4cn_step(e->m[i], e->dt[i], check);
      if (rc) return rc;
    }
    e->tables_ok = false;
    ++s;
  }
  if (s < nsteps && !e->tuned) {
    // NS2D_ENS_TUNE: 0 keeps the plain solver's shapes, 2 re-tunes ignoring
    // the schedule cache, otherwise a cached schedule for these classes on
    // this device is used and a missing one is tuned and cached
    const char* t = std::getenv("NS2D_ENS_TUNE");
    const int mode = t ? std::atoi(t) : 1;
    if (mode != 0 && (mode == 2 || !ens_cache_load(e))) {
      int rc = ens_autotune(e, 3);
      if (rc) return rc;
      ens_cache_store(e);
    }
    e->tuned = true;
  }
  if (s < nsteps) {
    int rc = ens_tables(e, check);
    if (rc) return rc;
  }
  while (s < nsteps) {
    const int p2 = e->m[0]->wcur ^ e->tables_parity;
    if (nsteps - s >= 2 && p2 == 0) {
      if (!e->gexec) {
        HIPCHK(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
        int rc = ens_launch_step(e, 0);
        if (rc == NS2D_OK) rc = ens_launch_step(e, 1);
        hipGraph_t g = nullptr;
        const hipError_t ec = hipStreamEndCapture(e->stream, &g);
        if (rc) {
          if (g) (void)hipGraphDestroy(g);
          return rc;
        }
        if (ec != hipSuccess) return fail(NS2D_ERR_HIP, "ensemble graph capture: %s", hipGetErrorString(ec));
        e->graph = g;
        HIPCHK(hipGraphInstantiate(&e->gexec, g, nullptr, nullptr, 0));
      }
      HIPCHK(hipGraphLaunch(e->gexec, e->stream));
      ens_advance(e);
      ens_advance(e);
      s += 2;
    } else {
      int rc = ens_launch_step(e, p2);
      if (rc) return rc;
      ens_advance(e);
      s += 1;
    }
  }
  for (ns2d_handle* h : e->m) h->inv_valid = true;
  if (check) {
    // SystemCheck of every member's final state (solver.c:2309)
    for (ns2d_handle* h : e->m) {
      RPassArgs a = rbase(h, RP_W);
      hipError_t r = launch_rpass(h->p.nx, h->p.ny, h->rvar, a, e->stream);
      if (r != hipSuccess) return fail(NS2D_ERR_HIP, "rpass launch: %s", hipGetErrorString(r));
    }
    int bad = -1;
    for (size_t i = 0; i < e->m.size(); ++i) {
      ns2d_handle* h = e->m[i];
      unsigned long long st[STAT_N];
      int rc = read_stat(h, st);
      if (rc) return rc;
      double mx;
      std::memcpy(&mx, &st[STAT_WMAX_CUR], sizeof mx);
      if (!(mx < MAX_VORT_LIM) && st[STAT_FIRST_BAD] == ~0ull) st[STAT_FIRST_BAD] = st[STAT_ITER];
      st[STAT_WMAX_LAST] = st[STAT_WMAX_CUR];
      st[STAT_WMAX_CUR] = 0ull;
      HIPCHK(hipMemcpyAsync(h->stat, st, sizeof st, hipMemcpyHostToDevice, e->stream));
      if (st[STAT_FIRST_BAD] != ~0ull && bad < 0) bad = (int)i;
    }
    HIPCHK(hipStreamSynchronize(e->stream));
    if (bad >= 0) return fail(NS2D_ERR_BLOWUP, "ensemble member %d reached the maximum vorticity limit", bad);
  }
  return NS2D_OK;
}

int ns2d_ens_autotune(ns2d_ens_t* e, int reps) {
  if (!e || reps < 1) return fail(NS2D_ERR_INVALID, "bad autotune arguments");
  HIPCHK(hipSetDevice(e->m[0]->p.device));
  for (ns2d_handle* h : e->m) {
    int rc = prepare_inv(h);
    if (rc) return rc;
  }
  int rc = ens_autotune(e, reps);
  if (rc == NS2D_OK) ens_cache_store(e);
  return rc;
}

int ns2d_ens_variants(ns2d_ens_t* e, int member, int* kpass_idx, int* rpass_idx) {
  if (!e || !kpass_idx || !rpass_idx || member < 0 || member >= (int)e->m.size())
    return fail(NS2D_ERR_INVALID, "bad arguments");
  for (auto& c : e->cls)
    for (int i : c.idx)
      if (i == member) {
        *kpass_idx = c.kvar;
        *rpass_idx = c.rvar;
      }
  return NS2D_OK;
}

int ns2d_ens_sync(ns2d_ens_t* e) {
  if (!e) return fail(NS2D_ERR_INVALID, "null ensemble");
  HIPCHK(hipStreamSynchronize(e->stream));
  return NS2D_OK;
}

int ns2d_ens_destroy(ns2d_ens_t* e) {
  if (!e) return NS2D_OK;
  if (e->stream) (void)hipStreamSynchronize(e->stream);
  ens_drop(e);
  for (auto& c : e->cls) {
    if (c.stream) (void)hipStreamSynchronize(c.stream);
    if (c.dk) (void)hipFree(c.dk);
    if (c.dr) (void)hipFree(c.dr);
    if (c.done) (void)hipEventDestroy(c.done);
    if (c.stream) (void)hipStreamDestroy(c.stream);
  }
  for (ns2d_handle* h : e->m) ns2d_destroy(h);
  if (e->fork) (void)hipEventDestroy(e->fork);
  if (e->stream) (void)hipStreamDestroy(e->stream);
  delete e;
  return NS2D_OK;
}


// ---------------------------------------------------------------------------
// Discrete adjoint of the RK4CN map (no counterpart in the reference).
// ---------------------------------------------------------------------------
}  // extern "C"

namespace {

int adj_alloc(ns2d_handle* h, long slots, long nckpt) {
  const size_t fb = h->M * sizeof(double2);
  if (h->tape_slots < slots) {
    for (double2*& t : h->tape_buf) {
      if (t) (void)hipFree(t);
      t = nullptr;
    }
    h->tape_slots = 0;
    const size_t tb = fb * 4 * (size_t)slots;
    if (hipMalloc((void**)&h->tape_buf[0], tb) != hipSuccess)
      return fail(NS2D_ERR_HIP, "adjoint tape of %ld stage slots (%.1f GB) does not fit", slots, tb / 1e9);
    // a second tape lets the recompute of one segment overlap the adjoint
    // sweep of the next; taken only while it leaves half the free memory
    // (NS2D_ADJ_OVERLAP=0 keeps one tape and the sequential order)
    const char* ov = std::getenv("NS2D_ADJ_OVERLAP");
    size_t fr = 0, tot = 0;
    if (!(ov && ov[0] == '0') && hipMemGetInfo(&fr, &tot) == hipSuccess && tb <= fr / 2) {
      if (hipMalloc((void**)&h->tape_buf[1], tb) != hipSuccess) {
        (void)hipGetLastError();
        h->tape_buf[1] = nullptr;
      }
    }
    h->tape_slots = slots;
  }
  h->tape = h->tape_buf[0];
  if (h->tape_buf[1] && !h->stream2) {
    if (hipStreamCreateWithFlags(&h->stream2, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&h->ev_fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&h->ev_join, hipEventDisableTiming) != hipSuccess)
      return fail(NS2D_ERR_HIP, "adjoint stream create failed");
  }
  if (h->ckpt_n < nckpt) {
    if (h->ckpt) (void)hipFree(h->ckpt);
    h->ckpt = nullptr;
    h->ckpt_n = 0;
    if (hipMalloc((void**)&h->ckpt, fb * (size_t)nckpt) != hipSuccess)
      return fail(NS2D_ERR_HIP, "adjoint checkpoints do not fit");
    h->ckpt_n = nckpt;
  }
  double2** bufs[] = {&h->lamp[0], &h->lamp[1], &h->lacc, &h->lbuf, &h->bbuf[0], &h->bbuf[1], &h->bbuf[2], &h->bbuf[3]};
  for (double2** b : bufs) {
    if (*b) continue;
    if (hipMalloc((void**)b, fb) != hipSuccess) return fail(NS2D_ERR_HIP, "adjoint buffer allocation failed");
    HIPCHK(hipMemsetAsync(*b, 0, fb, h->stream));
  }
  return NS2D_OK;
}

double2* tape_slot(ns2d_handle* h, long slot, int f) { return h->tape + ((size_t)slot * 4 + f) * h->M; }

// RK4CN steps [0, cnt) from the current state with every stage's G1..G4 kept
// in tape slots 4*j + s (slot 4*cnt holds the final state's).
int taped_forward(ns2d_handle* h, double dt, long cnt) {
  KPassArgs a0 = kbase(h);
  a0.mode = KP_INV;
  a0.epi = EPI_FROM_STATE;
  a0.conj_dst = h->conj[h->cur];
  for (int f = 0; f < 4; ++f) a0.g[f] = tape_slot(h, 0, f);
  int rc = run_kpass(h, a0);
  if (rc) return rc;
  for (long j = 0; j < cnt; ++j) {
    for (int st = 0; st < 4; ++st) {
      RPassArgs ra;
      KPassArgs ka;
      rc = rk4cn_stage_args(h, dt, false, st, &ra, &ka);
      if (rc) return rc;
      for (int f = 0; f < 4; ++f) {
        ra.g[f] = tape_slot(h, 4 * j + st, f);
        ka.g[f] = tape_slot(h, 4 * j + st + 1, f);
      }
      rc = run_rpass(h, ra);
      if (rc) return rc;
      rc = run_kpass(h, ka);
      if (rc) return rc;
    }
    after_rk_update(h);
    h->iters += 1;
    h->wcur ^= 1;
  }
  h->inv_valid = false;
  return NS2D_OK;
}

// dJ/dw_0 for J depending on w_n = Phi^n(w_0) through dJ/dw_n (in lamp[0]);
// w_0 in w[wcur].  Result in lamp[h->cur ... ] returned through *res.
int vjp_impl(ns2d_handle* h, double dt, long n, long ckpt_every, double2** res) {
  if (h->p.quirks & NS2D_QUIRK_CONJ_GLOBAL)
    return fail(NS2D_ERR_INVALID, "the adjoint needs the ForceConjugacy quirk off (quirks & NS2D_QUIRK_CONJ_GLOBAL)");
  if (h->p.integrator != NS2D_INT_RK4CN) return fail(NS2D_ERR_INVALID, "the adjoint covers the RK4CN integrator");
  const long C = (ckpt_every <= 0 || ckpt_every > n) ? n : ckpt_every;
  const long S = (n + C - 1) / C;
  int rc = adj_alloc(h, 4 * C + 1, S);
  if (rc) return rc;
  const size_t fb = h->M * sizeof(double2);
  // D_fac tables of step 0 and of the later steps (the Q3 flag update)
  h->hyper_eff = h->p.hyper_flag;
  const double* lam0 = lam_table(h, visc_kind_rk(h));
  after_rk_update(h);
  const double* lam1 = lam_table(h, visc_kind_rk(h));
  h->hyper_eff = h->p.hyper_flag;
  if (!lam0 || !lam1) return fail(NS2D_ERR_HIP, "coefficient table allocation failed");
  auto lam_of = [&](long step) { return step == 0 ? lam0 : lam1; };
  // forward sweep: checkpoints, the last segment taped.  w_0 was just placed
  // in the state buffer, so G1..G4 of an earlier state must not be reused.
  // The untaped segments run as ordinary (graph-replayed) steps.
  h->inv_valid = false;
  h->iters = 0;
  const int check0 = h->p.check_every;
  h->p.check_every = 0;
  for (long seg = 0; seg < S; ++seg) {
    HIPCHK(hipMemcpyAsync(h->ckpt + (size_t)seg * h->M, h->w[h->wcur], fb, hipMemcpyDeviceToDevice, h->stream));
    const long cnt = (seg == S - 1) ? n - seg * C : C;
    if (seg == S - 1) h->tape = (h->tape_buf[1] && S > 1) ? h->tape_buf[seg & 1] : h->tape_buf[0];
    rc = (seg == S - 1) ? taped_forward(h, dt, cnt) : step_impl(h, dt, cnt);
    if (rc) {
      h->p.check_every = check0;
      return rc;
    }
    h->inv_valid = false;
  }
  h->p.check_every = check0;
  // backward sweep.  A segment's stages are recomputed onto a tape, then
  // swept backwards.  With two tapes the recompute of segment seg-1 (second
  // stream, state buffers) runs while the adjoint sweep of segment seg (own
  // stream, adjoint buffers) runs: they share no buffer but the tapes.
  int lc = 0;   // lamp[lc] holds dJ/dw at the current step boundary
  bool started = false;
  const bool dbl = h->tape_buf[1] != nullptr && S > 1;
  auto tape_of = [&](long seg) { return dbl ? h->tape_buf[seg & 1] : h->tape_buf[0]; };
  // segment seg's stages onto its tape, from its checkpoint already copied
  // into the state buffer (that copy stays outside the captured graphs: its
  // source address is the one thing that changes between segments)
  auto recompute = [&](long seg, hipStream_t st) -> int {
    const long cnt = (seg == S - 1) ? n - seg * C : C;
    hipStream_t keep = h->stream;
    h->stream = st;
    h->iters = seg * C;
    h->hyper_eff = h->p.hyper_flag;
    if (h->iters > 0) after_rk_update(h);
    h->tape = tape_of(seg);
    const int r = taped_forward(h, dt, cnt);
    h->stream = keep;
    return r;
  };
  auto adjoint = [&](long seg) -> int {
    const long cnt = (seg == S - 1) ? n - seg * C : C;
    h->tape = tape_of(seg);
    for (long j = cnt - 1; j >= 0; --j) {
      const long ng = seg * C + j;
      KAdjArgs ka{};
      ka.tw = h->tw;
      ka.nyf = h->nyf;
      ka.ld = h->ld;
      ka.dt = dt;
      ka.ph = h->ph;
      ka.lacc = h->lacc;
      ka.lout = h->lbuf;
      for (int f = 0; f < 4; ++f) ka.bin[f] = h->bbuf[f];
      ka.lam = lam_of(ng);
      if (!started) {
        KAdjArgs k0 = ka;
        k0.mode = KA_INV;
        k0.epi = ADJ_START;
        k0.lamp_in = h->lamp[lc];
        hipError_t e = launch_kpass_adj(h->p.nx, k0, h->stream);
        if (e != hipSuccess) return fail(NS2D_ERR_HIP, "adjoint kpass: %s", hipGetErrorString(e));
        started = true;
      }
      const int epis[4] = {ADJ_S1, ADJ_S2, ADJ_S3, ADJ_S4};
      for (int st = 3; st >= 0; --st) {
        RAdjArgs ra{};
        ra.lin = h->lbuf;
        for (int f = 0; f < 4; ++f) {
          ra.g[f] = tape_slot(h, 4 * j + st, f);
          ra.bout[f] = h->bbuf[f];
        }
        ra.tw = h->tw;
        ra.nx = h->p.nx;
        ra.ld = h->ld;
        hipError_t e = launch_rpass_adj(h->p.nx, h->p.ny, ra, h->stream);
        if (e != hipSuccess) return fail(NS2D_ERR_HIP, "adjoint rpass: %s", hipGetErrorString(e));
        KAdjArgs kk = ka;
        kk.mode = KA_FWD | ((st > 0 || ng > 0) ? KA_INV : 0);
        kk.epi = epis[st];
        kk.lamp_in = h->lamp[lc];
        if (st == 0) {
          kk.lamp_out = h->lamp[lc ^ 1];
          kk.lam_prev = ng > 0 ? lam_of(ng - 1) : nullptr;
        }
        e = launch_kpass_adj(h->p.nx, kk, h->stream);
        if (e != hipSuccess) return fail(NS2D_ERR_HIP, "adjoint kpass: %s", hipGetErrorString(e));
      }
      lc ^= 1;
    }
    return NS2D_OK;
  };
  // one backward iteration: the adjoint of seg and, ahead of it, segment
  // seg-1 onto its tape (overlapped when two tapes exist)
  auto iteration = [&](long seg) -> int {
    if (!dbl) {
      int r = seg != S - 1 ? recompute(seg, h->stream) : NS2D_OK;
      return r ? r : adjoint(seg);
    }
    // seg's tape is ready: the forward sweep taped S-1, each iteration the one before
    if (seg >= 1) {
      HIPCHK(hipEventRecord(h->ev_fork, h->stream));
      HIPCHK(hipStreamWaitEvent(h->stream2, h->ev_fork, 0));
      int r = recompute(seg - 1, h->stream2);
      if (r) return r;
      r = adjoint(seg);
      if (r) return r;
      HIPCHK(hipEventRecord(h->ev_join, h->stream2));
      HIPCHK(hipStreamWaitEvent(h->stream, h->ev_join, 0));
      return NS2D_OK;
    }
    return adjoint(seg);
  };
  for (long seg = S - 1; seg >= 0; --seg) {
    const long cnt = (seg == S - 1) ? n - seg * C : C;
    // the checkpoint this iteration recomputes from; with two tapes the
    // copy waits for nothing but the previous iteration's join
    const long rseg = dbl ? seg - 1 : (seg != S - 1 ? seg : -1);
    if (rseg >= 0)
      HIPCHK(hipMemcpyAsync(h->w[h->wcur], h->ckpt + (size_t)rseg * h->M, fb, hipMemcpyDeviceToDevice, h->stream));
    // middle iterations repeat the same launches: one capture per parity of
    // (tape, state buffer, adjoint buffer), replayed for the others
    const bool graphable = h->use_graph && !h->profiling && started && seg != S - 1 && C >= 2 &&
                           seg >= (dbl ? 2 : 1);
    if (!graphable) {
      rc = iteration(seg);
      if (rc) return rc;
      continue;
    }
    const double2* tkey = tape_of(seg);
    ns2d_handle::AdjGraph* g = nullptr;
    for (auto& c : h->adj_graph)
      if (c.exec && c.dt == dt && c.C == C && c.wcur == h->wcur && c.lc == lc && c.lam == lam1 && c.tape == tkey)
        g = &c;
    if (!g) {
      g = !h->adj_graph[0].exec ? &h->adj_graph[0] : !h->adj_graph[1].exec ? &h->adj_graph[1] : &h->adj_graph[seg & 1];
      if (g->exec) (void)hipGraphExecDestroy(g->exec);
      g->exec = nullptr;
      const int wcur0 = h->wcur, lc0 = lc;
      HIPCHK(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
      rc = iteration(seg);
      hipGraph_t gr = nullptr;
      const hipError_t ec = hipStreamEndCapture(h->stream, &gr);
      if (rc) {
        if (gr) (void)hipGraphDestroy(gr);
        return rc;
      }
      if (ec != hipSuccess) return fail(NS2D_ERR_HIP, "adjoint graph capture: %s", hipGetErrorString(ec));
      const hipError_t ei = hipGraphInstantiate(&g->exec, gr, nullptr, nullptr, 0);
      (void)hipGraphDestroy(gr);
      if (ei != hipSuccess) return fail(NS2D_ERR_HIP, "adjoint graph instantiate: %s", hipGetErrorString(ei));
      g->dt = dt;
      g->C = C;
      g->wcur = wcur0;
      g->lc = lc0;
      g->lam = lam1;
      g->tape = tkey;
      // the capture advanced the host-side parities; the replay does the work
      h->wcur = wcur0;
      lc = lc0;
    }
    HIPCHK(hipGraphLaunch(g->exec, h->stream));
    // host-side parities as the iteration leaves them: the recompute (of seg-1
    // with two tapes, of seg otherwise) flips the state buffer once per step,
    // the adjoint of seg flips lc once per step
    h->wcur ^= (int)(C & 1);
    lc ^= (int)(cnt & 1);
  }
  *res = h->lamp[lc];
  h->inv_valid = false;
  return NS2D_OK;
}

}  // namespace

extern "C" {

int ns2d_vjp_device(ns2d_t* h, const void* d_w0, double dt, long nsteps, const void* d_lam_T, void* d_grad_w0,
                    long ckpt_every) {
  if (!h || !d_w0 || !d_lam_T || !d_grad_w0) return fail(NS2D_ERR_INVALID, "null argument");
  if (nsteps < 1 || !(dt > 0.0)) return fail(NS2D_ERR_INVALID, "need nsteps >= 1 and dt > 0");
  HIPCHK(hipSetDevice(h->p.device));
  int rc = adj_alloc(h, 1, 1);
  if (rc) return rc;
  rc = copy_in(h, h->w[h->wcur], d_w0, hipMemcpyDeviceToDevice);
  if (rc == NS2D_OK) rc = copy_in(h, h->lamp[0], d_lam_T, hipMemcpyDeviceToDevice);
  if (rc) return rc;
  double2* res = nullptr;
  rc = vjp_impl(h, dt, nsteps, ckpt_every, &res);
  if (rc) return rc;
  rc = copy_out(h, d_grad_w0, res, hipMemcpyDeviceToDevice);
  if (rc) return rc;
  HIPCHK(hipStreamSynchronize(h->stream));
  return NS2D_OK;
}

int ns2d_vjp(ns2d_t* h, const double* w0, double dt, long nsteps, const double* lam_T, double* grad_w0,
             long ckpt_every) {
  if (!h || !w0 || !lam_T || !grad_w0) return fail(NS2D_ERR_INVALID, "null argument");
  if (nsteps < 1 || !(dt > 0.0)) return fail(NS2D_ERR_INVALID, "need nsteps >= 1 and dt > 0");
  HIPCHK(hipSetDevice(h->p.device));
  int rc = adj_alloc(h, 1, 1);
  if (rc) return rc;
  rc = copy_in(h, h->w[h->wcur], w0, hipMemcpyHostToDevice);
  if (rc == NS2D_OK) rc = copy_in(h, h->lamp[0], lam_T, hipMemcpyHostToDevice);
  if (rc) return rc;
  double2* res = nullptr;
  rc = vjp_impl(h, dt, nsteps, ckpt_every, &res);
  if (rc) return rc;
  rc = copy_out(h, grad_w0, res, hipMemcpyDeviceToHost);
  if (rc) return rc;
  HIPCHK(hipStreamSynchronize(h->stream));
  return NS2D_OK;
}

}  // extern "C"
