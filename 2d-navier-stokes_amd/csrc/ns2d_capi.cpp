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
#include <cstring>
#include <string>
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
  double2 *w_hat = nullptr, *acc = nullptr, *g[4] = {nullptr, nullptr, nullptr, nullptr};
  double2 *nbuf = nullptr, *kout = nullptr, *sbuf = nullptr, *hist[3] = {nullptr, nullptr, nullptr};
  double2 *conj[2] = {nullptr, nullptr}, *tw = nullptr;
  unsigned long long* stat = nullptr;
  double* meas = nullptr;
  double* wreal = nullptr;
  double* rbuf = nullptr;
  bool inv_valid = false;  // G1..G4 and conj[cur] hold the transforms of the state
  int cur = 0;             // conj[cur] = ky=0 column of the current stage input
  long iters = 0;
  int hyper_eff = 0;       // HYPER_VISC_FLAG as mutated by the reference (Q3)
  PhysArgs ph{};
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

void after_rk_update(ns2d_handle* h) {
  if ((h->p.quirks & NS2D_QUIRK_VISC_ASSIGN) && h->p.drag_flag != 1) h->hyper_eff = 1;
}

KPassArgs kbase(ns2d_handle* h) {
  KPassArgs a{};
  a.w_hat = h->w_hat;
  a.acc = h->acc;
  a.nin = h->nbuf;
  for (int f = 0; f < 4; ++f) a.g[f] = h->g[f];
  a.kout = h->kout;
  a.stat = h->stat;
  a.tw = h->tw;
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
  a.mode = mode;
  a.wscale = 1.0 / (double)((long)h->p.nx * h->p.ny);
  return a;
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
  hipError_t e = launch_kpass(h->p.nx, a, h->stream);
  ev_end(h);
  if (e != hipSuccess) return fail(NS2D_ERR_HIP, "kpass launch: %s", hipGetErrorString(e));
  return NS2D_OK;
}

int run_rpass(ns2d_handle* h, const RPassArgs& a) {
  ev_begin(h, KIND_R);
  hipError_t e = launch_rpass(h->p.nx, h->p.ny, a, h->stream);
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

int rk4cn_step(ns2d_handle* h, double dt, bool check) {
  const long it = h->iters + 1;
  double2* hist_store = nullptr;
  if (h->p.integrator == NS2D_INT_AB4 && it <= 3) hist_store = h->hist[it - 1];  // solver.c:283
  const int epis[4] = {EPI_RK1, EPI_RK2, EPI_RK3, EPI_RK4CN};
  const double coefs[4] = {dt * RK4_A21, dt * RK4_A32, dt * RK4_A43, 0.0};
  for (int s = 0; s < 4; ++s) {
    int rc = run_rpass(h, rbase(h, RP_NL | ((check && s == 0) ? RP_W : 0)));
    if (rc) return rc;
    KPassArgs a = kbase(h);
    a.mode = KP_FWD | KP_INV | conj_mode(h);
    if (check && s == 0) a.mode |= KP_STAT;
    if (s == 3) a.mode |= KP_ITER;
    a.epi = epis[s];
    a.coef = coefs[s];
    a.dt = dt;
    a.visc_kind = visc_kind_rk(h);
    a.hist_store = (s == 0) ? hist_store : nullptr;
    a.conj_src = h->conj[h->cur];
    a.conj_dst = h->conj[h->cur ^ 1];
    rc = run_kpass(h, a);
    if (rc) return rc;
    h->cur ^= 1;
  }
  after_rk_update(h);
  h->iters = it;
  return NS2D_OK;
}

int ab4_step(ns2d_handle* h, double dt, bool check) {
  const long it = h->iters + 1;
  if (it <= 3) return rk4cn_step(h, dt, check);
  const long m = it - 4;  // AB iteration: slots hold N(n-3), N(n-2), N(n-1)
  int rc = run_rpass(h, rbase(h, RP_NL | (check ? RP_W : 0)));
  if (rc) return rc;
  KPassArgs a = kbase(h);
  a.mode = KP_FWD | KP_INV | KP_ITER | conj_mode(h) | (check ? KP_STAT : 0);
  a.epi = EPI_AB4;
  a.dt = dt;
  a.visc_kind = visc_kind_plain(h);  // solver.c:315-330
  a.h_old = h->hist[m % 3];
  a.h_mid = h->hist[(m + 1) % 3];
  a.h_new = h->hist[(m + 2) % 3];
  a.conj_src = h->conj[h->cur];
  a.conj_dst = h->conj[h->cur ^ 1];
  rc = run_kpass(h, a);
  if (rc) return rc;
  h->cur ^= 1;
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

int ns2d_create(const ns2d_params* p, ns2d_t** out) {
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
  if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess)
    return cleanup(fail(NS2D_ERR_HIP, "stream create failed"));
  const size_t bytes = h->M * sizeof(double2);
  std::vector<double2**> bufs = {&h->w_hat, &h->acc, &h->g[0], &h->g[1], &h->g[2], &h->g[3], &h->nbuf, &h->kout, &h->sbuf};
  if (p->integrator == NS2D_INT_AB4)
    for (int i = 0; i < 3; ++i) bufs.push_back(&h->hist[i]);
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
  int rc = reset_stat(h);
  if (rc) return cleanup(rc);
  *out = h;
  return NS2D_OK;
}

int ns2d_destroy(ns2d_t* h) {
  if (!h) return NS2D_OK;
  (void)hipSetDevice(h->p.device);
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  double2* bufs[] = {h->w_hat, h->acc, h->g[0], h->g[1], h->g[2], h->g[3], h->nbuf, h->kout, h->sbuf,
                     h->hist[0], h->hist[1], h->hist[2], h->conj[0], h->conj[1], h->tw};
  for (double2* b : bufs)
    if (b) (void)hipFree(b);
  if (h->stat) (void)hipFree(h->stat);
  if (h->meas) (void)hipFree(h->meas);
  if (h->wreal) (void)hipFree(h->wreal);
  if (h->rbuf) (void)hipFree(h->rbuf);
  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
  return NS2D_OK;
}

static int set_state_impl(ns2d_t* h, const void* src, hipMemcpyKind kind) {
  if (!h || !src) return fail(NS2D_ERR_INVALID, "null argument");
  HIPCHK(hipSetDevice(h->p.device));
  HIPCHK(hipMemcpy2DAsync(h->w_hat, h->ld * sizeof(double2), src, h->nyf * sizeof(double2), h->nyf * sizeof(double2),
                          h->p.nx, kind, h->stream));
  h->inv_valid = false;
  h->iters = 0;
  h->hyper_eff = h->p.hyper_flag;
  return reset_stat(h);
}

static int get_state_impl(ns2d_t* h, void* dst, hipMemcpyKind kind) {
  if (!h || !dst) return fail(NS2D_ERR_INVALID, "null argument");
  HIPCHK(hipSetDevice(h->p.device));
  HIPCHK(hipMemcpy2DAsync(dst, h->nyf * sizeof(double2), h->w_hat, h->ld * sizeof(double2), h->nyf * sizeof(double2),
                          h->p.nx, kind, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  return NS2D_OK;
}

int ns2d_set_state(ns2d_t* h, const double* w_hat) { return set_state_impl(h, w_hat, hipMemcpyHostToDevice); }
int ns2d_get_state(ns2d_t* h, double* w_hat) { return get_state_impl(h, w_hat, hipMemcpyDeviceToHost); }
int ns2d_set_state_device(ns2d_t* h, const void* d) { return set_state_impl(h, d, hipMemcpyDeviceToDevice); }
int ns2d_get_state_device(ns2d_t* h, void* d) { return get_state_impl(h, d, hipMemcpyDeviceToDevice); }

static int step_impl(ns2d_t* h, double dt, long nsteps) {
  if (!h) return fail(NS2D_ERR_INVALID, "null handle");
  if (nsteps < 0) return fail(NS2D_ERR_INVALID, "nsteps < 0");
  if (!(dt > 0.0)) return fail(NS2D_ERR_INVALID, "dt must be positive");
  HIPCHK(hipSetDevice(h->p.device));
  if (nsteps == 0) return NS2D_OK;
  const bool check = h->p.check_every != 0;
  int rc = prepare_inv(h);
  if (rc) return rc;
  for (long s = 0; s < nsteps; ++s) {
    rc = (h->p.integrator == NS2D_INT_AB4) ? ab4_step(h, dt, check) : rk4cn_step(h, dt, check);
    if (rc) return rc;
  }
  // the state's transforms are prepared by the last k-space pass
  h->inv_valid = true;
  if (check) {
    // SystemCheck on the final state (solver.c:2309)
    ev_begin(h, KIND_OTHER);
    RPassArgs a = rbase(h, RP_W);
    hipError_t e = launch_rpass(h->p.nx, h->p.ny, a, h->stream);
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
  HIPCHK(hipMemcpy2DAsync(h->sbuf, h->ld * sizeof(double2), in, h->nyf * sizeof(double2), h->nyf * sizeof(double2),
                          h->p.nx, hipMemcpyHostToDevice, h->stream));
  KPassArgs a = kbase(h);
  a.mode = KP_INV;
  a.epi = EPI_FROM_INPUT;
  a.sin = h->sbuf;
  a.conj_dst = h->conj[h->cur];
  int rc = run_kpass(h, a);
  if (rc) return rc;
  h->inv_valid = false;  // G1..G4 now hold the transforms of `in`
  rc = finish_rhs_into_kout(h, false);
  if (rc) return rc;
  HIPCHK(hipMemcpy2DAsync(out, h->nyf * sizeof(double2), h->kout, h->ld * sizeof(double2), h->nyf * sizeof(double2),
                          h->p.nx, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  return NS2D_OK;
}

int ns2d_measures(ns2d_t* h, int with_flux, ns2d_measures_t* m, double* enrg_spec, double* enst_spec,
                  double* flux_spec) {
  if (!h) return fail(NS2D_ERR_INVALID, "null handle");
  HIPCHK(hipSetDevice(h->p.device));
  int rc;
  if (with_flux) {
    // sys_msr.c:125-133: NonlinearRHSBatch(w_hat) (with its ForceConjugacy
    // write into w_hat), minus the forcing
    rc = prepare_inv(h);
    if (rc) return rc;
    rc = finish_rhs_into_kout(h, false);
    if (rc) return rc;
    if (h->p.quirks & NS2D_QUIRK_CONJ_GLOBAL) h->inv_valid = false;
  }
  MeasArgs a{};
  a.w_hat = h->w_hat;
  a.rhs = with_flux ? h->kout : nullptr;
  a.out = h->meas;
  a.nx = h->p.nx;
  a.ny = h->p.ny;
  a.ld = h->ld;
  a.visc_kind = visc_kind_plain(h);  // sys_msr.c:160-175
  a.quirk_enst_diss = (h->p.quirks & NS2D_QUIRK_ENST_DISS) ? 1 : 0;
  a.sbst_lo = 0.0;                   // LWR_SBST_LIM^2, data_types.h:175
  a.sbst_hi = 64.0 * 64.0;           // UPR_SBST_LIM^2, data_types.h:174
  a.ph = h->ph;
  hipError_t e = launch_measures(h->nspect, a, h->stream);
  if (e != hipSuccess) return fail(NS2D_ERR_HIP, "measures launch: %s", hipGetErrorString(e));
  std::vector<double> part((size_t)NQ * h->nspect);
  HIPCHK(hipMemcpyAsync(part.data(), h->meas, sizeof(double) * part.size(), hipMemcpyDeviceToHost, h->stream));
  double2 wf = make_double2(0.0, 0.0);
  if (h->ph.force_on)
    HIPCHK(hipMemcpyAsync(&wf, h->w_hat + h->p.force_k, sizeof wf, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  const double norm_fac = 0.5 / std::pow((double)((long)h->p.nx * h->p.ny), 2.0);  // sys_msr.c:43
  const double const_fac = 4.0 * std::pow(M_PI, 2.0);
  const double cn = const_fac * norm_fac;
  double tot[NQ] = {0};
  for (int b = 0; b < h->nspect; ++b)
    for (int q = 0; q < NQ; ++q) tot[q] += part[(size_t)b * NQ + q];
  // with_flux: the forcing is subtracted from the RHS inside the reduction
  if (m) {
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

int ns2d_max_vort(ns2d_t* h, double* out) {
  if (!h || !out) return fail(NS2D_ERR_INVALID, "null argument");
  HIPCHK(hipSetDevice(h->p.device));
  return state_wmax(h, out, nullptr);
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
  HIPCHK(hipMemcpy2DAsync(w_hat, h->nyf * sizeof(double2), h->kout, h->ld * sizeof(double2), h->nyf * sizeof(double2),
                          h->p.nx, hipMemcpyDeviceToHost, h->stream));
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

}  // extern "C"
