/*
 * ns2d_cpu.c -- compiled CPU restatement of the reference solver's RK4CN time
 * step (EndCar808/2D-Navier-Stokes, Solver/src/solver.c), used as the CPU
 * baseline beside the GPU figures and checked against the numpy oracle.
 *
 * TEST AND MEASUREMENT INFRASTRUCTURE ONLY: only tests/ and bench.py's
 * cpu_baseline leg run it; the product path never does.
 *
 * What it computes, step by step as the reference does (single process, no
 * MPI; the reference's FFTW-MPI slab transforms become FFTW3 threaded 2-D
 * transforms, here provided by the FFTW3 interface of Intel MKL's libmkl_rt):
 *   NonlinearRHSBatch (solver.c:884-1032): u_hat = i ky/|k|^2 w_hat,
 *     v_hat = -i kx/|k|^2 w_hat (:914-915), w_x = i kx w_hat, w_y = i ky w_hat
 *     (:946-947), two batched c2r of two fields each (:932, :964), the product
 *     u w_x + v w_y with the reference's + sign (:981), r2c (:994), the
 *     (1/(Nx Ny))^2 scaling and (0,0) = 0 (:1005-1008), ApplyDealiasing
 *     (:1039-1099: 2/3 rule on (int)(Nx/3)^2, Hou-Li with the reference's
 *     integer division) and ForceConjugacy into the global state (:1031,
 *     :2074-2107);
 *   RK4Step, __RK4CN branch (solver.c:625-873): stage inputs w + dt a K
 *     (:690, :729, :768), viscous factor D = dt nu k^(2p) (:835-850, with the
 *     reference's HYPER_VISC_FLAG assignment), update (:853), (0,0) = 0 (:860);
 *   SystemCheck / GetMaxData("VORT") (solver.c:2171-2326): max |c2r(w_hat)|/(Nx Ny)
 *     after every step (switchable, to report the figure with and without it).
 *
 * Usage:
 *   ns2d_cpu Nx Ny nu hyper_pow filter dt nsteps check reps ic.bin [out.bin]
 *     filter: 23 | houli_ref;  check: 1 = per-step max|w| check, 0 = none;
 *     ic.bin: complex128 [Nx][Ny/2+1] (unnormalised FFTW units);
 *     out.bin: the state after the last repetition.
 * Prints one JSON line: steps/s of each repetition (CLOCK_MONOTONIC around
 * the stepping loop only: planning, the IC read and output are outside), the
 * median, the thread count.  Threads: OpenMP's (OMP_NUM_THREADS) for both the
 * pointwise sweeps and the transforms (MKL_THREADING_LAYER=GNU).
 */
#define _POSIX_C_SOURCE 200112L
#include <math.h>
#include <omp.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

/* FFTW3 interface (names and signatures of FFTW 3's public API, fftw3.h);
 * the implementation is MKL's. */
typedef double fftw_complex[2];
typedef struct fftw_plan_s* fftw_plan;
#define FFTW_MEASURE (0U)
#define FFTW_DESTROY_INPUT (1U << 0)
fftw_plan fftw_plan_many_dft_r2c(int rank, const int* n, int howmany, double* in, const int* inembed, int istride,
                                 int idist, fftw_complex* out, const int* onembed, int ostride, int odist,
                                 unsigned flags);
fftw_plan fftw_plan_many_dft_c2r(int rank, const int* n, int howmany, fftw_complex* in, const int* inembed,
                                 int istride, int idist, double* out, const int* onembed, int ostride, int odist,
                                 unsigned flags);
void fftw_execute(const fftw_plan p);
void fftw_destroy_plan(fftw_plan p);
int fftw_init_threads(void);
void fftw_plan_with_nthreads(int nthreads);
void* fftw_malloc(size_t n);
void fftw_free(void* p);

enum { FILT_23 = 0, FILT_HOULI_REF = 1 };

typedef struct {
  int nx, ny, nyf, filter;
  double nu, hyper_pow;
  long M, R;                 /* complex modes, real points */
  fftw_complex *w, *tmp, *k[4], *four;   /* state, stage input, K1..K4, 4 derivative fields */
  double* real4;             /* u, v, w_x, w_y in real space (planar) */
  double* nl;                /* the product */
  fftw_complex* nlh;         /* its r2c */
  fftw_complex* conj0;       /* ky = 0 column of the RHS argument */
  double *kxd, *dfac;        /* kx per row; dealias factor per mode */
  fftw_plan c2r_uv, c2r_ww, r2c_nl, c2r_w;
} solver_t;

static double now_s(void) {
  struct timespec t;
  clock_gettime(CLOCK_MONOTONIC, &t);
  return (double)t.tv_sec + 1e-9 * (double)t.tv_nsec;
}

static int kx_of(int i, int nx) { return i <= nx / 2 ? i : i - nx; }

static void setup(solver_t* s) {
  const int nx = s->nx, ny = s->ny, nyf = s->nyf;
  s->M = (long)nx * nyf;
  s->R = (long)nx * ny;
  s->w = fftw_malloc(sizeof(fftw_complex) * s->M);
  s->tmp = fftw_malloc(sizeof(fftw_complex) * s->M);
  for (int j = 0; j < 4; ++j) s->k[j] = fftw_malloc(sizeof(fftw_complex) * s->M);
  s->four = fftw_malloc(sizeof(fftw_complex) * 4 * s->M);
  s->real4 = fftw_malloc(sizeof(double) * 4 * s->R);
  s->nl = fftw_malloc(sizeof(double) * s->R);
  s->nlh = fftw_malloc(sizeof(fftw_complex) * s->M);
  s->conj0 = malloc(sizeof(fftw_complex) * nx);
  s->dfac = malloc(sizeof(double) * s->M);
  const int n[2] = {nx, ny};
  /* batched c2r of (u, v) and (w_x, w_y), r2c of the product (solver.c:2883-2895) */
  s->c2r_uv = fftw_plan_many_dft_c2r(2, n, 2, s->four, NULL, 1, (int)s->M, s->real4, NULL, 1, (int)s->R,
                                     FFTW_MEASURE | FFTW_DESTROY_INPUT);
  s->c2r_ww = fftw_plan_many_dft_c2r(2, n, 2, s->four + 2 * s->M, NULL, 1, (int)s->M, s->real4 + 2 * s->R, NULL,
                                     1, (int)s->R, FFTW_MEASURE | FFTW_DESTROY_INPUT);
  s->r2c_nl = fftw_plan_many_dft_r2c(2, n, 1, s->nl, NULL, 1, (int)s->R, s->nlh, NULL, 1, (int)s->M, FFTW_MEASURE);
  s->c2r_w = fftw_plan_many_dft_c2r(2, n, 1, s->tmp, NULL, 1, (int)s->M, s->nl, NULL, 1, (int)s->R,
                                    FFTW_MEASURE | FFTW_DESTROY_INPUT);
  /* ApplyDealiasing multiplier (solver.c:1039-1099) */
  const double kmax_sqr = pow((double)(int)(nx / 3.0), 2.0);
  const double e1 = exp(-36.0 * pow(sqrt(1.0), 36.0)), e2 = exp(-36.0 * pow(sqrt(2.0), 36.0));
  for (int i = 0; i < nx; ++i) {
    const int kx = kx_of(i, nx);
    for (int j = 0; j < nyf; ++j) {
      double f = 1.0;
      if (s->filter == FILT_23) {
        f = ((double)(kx * kx + j * j) >= kmax_sqr) ? 0.0 : 1.0;
      } else {
        const int q = (kx == nx / 2) + (j == ny / 2);   /* integer division k/(N/2) */
        f = q == 0 ? 1.0 : (q == 1 ? e1 : e2);
      }
      s->dfac[(long)i * nyf + j] = f;
    }
  }
}

static void teardown(solver_t* s) {
  fftw_destroy_plan(s->c2r_uv);
  fftw_destroy_plan(s->c2r_ww);
  fftw_destroy_plan(s->r2c_nl);
  fftw_destroy_plan(s->c2r_w);
  fftw_free(s->w);
  fftw_free(s->tmp);
  for (int j = 0; j < 4; ++j) fftw_free(s->k[j]);
  fftw_free(s->four);
  fftw_free(s->real4);
  fftw_free(s->nl);
  fftw_free(s->nlh);
  free(s->conj0);
  free(s->dfac);
}

/* NonlinearRHSBatch(arg) -> out, with ForceConjugacy(arg) into the state */
static void rhs(solver_t* s, const fftw_complex* arg, fftw_complex* out) {
  const int nx = s->nx, nyf = s->nyf;
  const long M = s->M, R = s->R;
  fftw_complex *u = s->four, *v = s->four + M, *wx = s->four + 2 * M, *wy = s->four + 3 * M;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < nx; ++i) {
    const double kx = (double)kx_of(i, nx);
    for (int j = 0; j < nyf; ++j) {
      const long id = (long)i * nyf + j;
      const double ky = (double)j;
      const double k2 = kx * kx + ky * ky;
      const double a = arg[id][0], b = arg[id][1];
      if (k2 != 0.0) {
        const double ik = 1.0 / k2;
        u[id][0] = -ky * ik * b;  u[id][1] = ky * ik * a;     /* i ky/|k|^2 w */
        v[id][0] = kx * ik * b;   v[id][1] = -kx * ik * a;    /* -i kx/|k|^2 w */
      } else {
        u[id][0] = u[id][1] = v[id][0] = v[id][1] = 0.0;
      }
      wx[id][0] = -kx * b; wx[id][1] = kx * a;                /* i kx w */
      wy[id][0] = -ky * b; wy[id][1] = ky * a;                /* i ky w */
    }
  }
  fftw_execute(s->c2r_uv);
  fftw_execute(s->c2r_ww);
  const double* ur = s->real4;
  const double *vr = ur + R, *wxr = ur + 2 * R, *wyr = ur + 3 * R;
#pragma omp parallel for schedule(static)
  for (long p = 0; p < R; ++p) s->nl[p] = ur[p] * wxr[p] + vr[p] * wyr[p];   /* + sign, solver.c:981 */
  fftw_execute(s->r2c_nl);
  const double norm = pow(1.0 / (double)((long)s->nx * s->ny), 2.0);
#pragma omp parallel for schedule(static)
  for (long id = 0; id < M; ++id) {
    out[id][0] = s->nlh[id][0] * norm * s->dfac[id];
    out[id][1] = s->nlh[id][1] * norm * s->dfac[id];
  }
  out[0][0] = out[0][1] = 0.0;
  /* ForceConjugacy(arg): w[kx<0, 0] = conj(arg[|kx|, 0]) on the global state */
  for (int i = 0; i < nx; ++i) {
    s->conj0[i][0] = arg[(long)i * nyf][0];
    s->conj0[i][1] = arg[(long)i * nyf][1];
  }
  for (int i = nx / 2 + 1; i < nx; ++i) {
    s->w[(long)i * nyf][0] = s->conj0[nx - i][0];
    s->w[(long)i * nyf][1] = -s->conj0[nx - i][1];
  }
}

static void stage_input(solver_t* s, double c, const fftw_complex* K) {
#pragma omp parallel for schedule(static)
  for (long id = 0; id < s->M; ++id) {
    s->tmp[id][0] = s->w[id][0] + c * K[id][0];
    s->tmp[id][1] = s->w[id][1] + c * K[id][1];
  }
}

static void rk4cn_step(solver_t* s, double dt) {
  rhs(s, s->w, s->k[0]);
  stage_input(s, dt * 0.5, s->k[0]);
  rhs(s, s->tmp, s->k[1]);
  stage_input(s, dt * 0.5, s->k[1]);
  rhs(s, s->tmp, s->k[2]);
  stage_input(s, dt * 1.0, s->k[2]);
  rhs(s, s->tmp, s->k[3]);
  const int nx = s->nx, nyf = s->nyf;
  const double b1 = 1.0 / 6.0, b2 = 1.0 / 3.0, b3 = 1.0 / 3.0, b4 = 1.0 / 6.0;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < nx; ++i) {
    const double kx = (double)kx_of(i, nx);
    for (int j = 0; j < nyf; ++j) {
      const long id = (long)i * nyf + j;
      const double k2 = kx * kx + (double)j * j;
      /* with the drag flag off the reference's branch condition assigns
         HYPER_VISC_FLAG = HYPER_VISC (solver.c:843): nu pow(k^2, p) from the
         first step on */
      const double D = dt * (s->nu * pow(k2, s->hyper_pow));
      const double A = (2.0 - D) / (2.0 + D), B = 2.0 * dt / (2.0 + D);
      for (int r = 0; r < 2; ++r) {
        const double sum = b1 * s->k[0][id][r] + b2 * s->k[1][id][r] + b3 * s->k[2][id][r] + b4 * s->k[3][id][r];
        s->w[id][r] = s->w[id][r] * A + B * sum;
      }
    }
  }
  s->w[0][0] = s->w[0][1] = 0.0;
}

static double max_vort(solver_t* s) {
  memcpy(s->tmp, s->w, sizeof(fftw_complex) * s->M);
  fftw_execute(s->c2r_w);
  const double sc = 1.0 / (double)s->R;
  double m = 0.0;
#pragma omp parallel for reduction(max : m) schedule(static)
  for (long p = 0; p < s->R; ++p) {
    const double a = fabs(s->nl[p] * sc);
    if (a > m) m = a;
  }
  return m;
}

static int cmp_d(const void* a, const void* b) {
  const double x = *(const double*)a, y = *(const double*)b;
  return (x > y) - (x < y);
}

int main(int argc, char** argv) {
  if (argc < 11) {
    fprintf(stderr, "usage: %s Nx Ny nu hyper_pow filter dt nsteps check reps ic.bin [out.bin]\n", argv[0]);
    return 2;
  }
  solver_t s;
  memset(&s, 0, sizeof s);
  s.nx = atoi(argv[1]);
  s.ny = atoi(argv[2]);
  s.nyf = s.ny / 2 + 1;
  s.nu = atof(argv[3]);
  s.hyper_pow = atof(argv[4]);
  s.filter = strcmp(argv[5], "houli_ref") == 0 ? FILT_HOULI_REF : FILT_23;
  const double dt = atof(argv[6]);
  const long nsteps = atol(argv[7]);
  const int check = atoi(argv[8]);
  int reps = atoi(argv[9]);
  if (reps < 1) reps = 1;
  /* MKL's transforms on the OpenMP runtime of the pointwise sweeps */
  setenv("MKL_THREADING_LAYER", "GNU", 0);
  const int nthr = omp_get_max_threads();
  fftw_init_threads();
  fftw_plan_with_nthreads(nthr);
  setup(&s);
  fftw_complex* ic = fftw_malloc(sizeof(fftw_complex) * s.M);
  FILE* f = fopen(argv[10], "rb");
  if (!f || fread(ic, sizeof(fftw_complex), (size_t)s.M, f) != (size_t)s.M) {
    fprintf(stderr, "cannot read %ld complex values from %s\n", s.M, argv[10]);
    return 1;
  }
  fclose(f);
  double rate[16];
  if (reps > 16) reps = 16;
  double wmax = 0.0;
  for (int r = 0; r < reps; ++r) {
    memcpy(s.w, ic, sizeof(fftw_complex) * s.M);
    const double t0 = now_s();
    for (long n = 0; n < nsteps; ++n) {
      rk4cn_step(&s, dt);
      if (check) {
        wmax = max_vort(&s);
        if (!(wmax < 1e100)) {
          fprintf(stderr, "maximum vorticity limit reached at step %ld\n", n + 1);
          return 3;
        }
      }
    }
    const double el = now_s() - t0;
    rate[r] = (double)nsteps / el;
  }
  if (argc > 11) {
    FILE* o = fopen(argv[11], "wb");
    if (!o || fwrite(s.w, sizeof(fftw_complex), (size_t)s.M, o) != (size_t)s.M) {
      fprintf(stderr, "cannot write %s\n", argv[11]);
      return 1;
    }
    fclose(o);
  }
  double sorted[16];
  memcpy(sorted, rate, sizeof(double) * reps);
  qsort(sorted, reps, sizeof(double), cmp_d);
  printf("{\"nx\": %d, \"ny\": %d, \"steps\": %ld, \"check\": %d, \"threads\": %d, \"steps_per_s\": [", s.nx, s.ny,
         nsteps, check, nthr);
  for (int r = 0; r < reps; ++r) printf("%s%.6g", r ? ", " : "", rate[r]);
  printf("], \"median_steps_per_s\": %.6g, \"last_max_vort\": %.17g}\n", sorted[reps / 2], wmax);
  fftw_free(ic);
  teardown(&s);
  return 0;
}
