/* Initial conditions, dealiasing, conjugacy and mean flows on the host.
 *
 * InitialConditions (Solver/src/solver.c:1107-2014) for one process holding
 * the whole grid: the random draws come from glibc rand() after
 * srand(123456789) (solver.c:1121), consumed in the reference's (i, j) loop
 * order and under the same conditions, so a run reproduces the reference's
 * single-rank field, or with ic_ranks = P the field of a P-rank run (each
 * rank's generator restarted on its own slab, the normalisation sums reduced
 * over the slabs; ic_slabs below).  The arithmetic of each formula is kept in the
 * reference's operation order (C99 complex, no FMA contraction on x86-64) so
 * that the spectral initial conditions agree bit for bit; the ones built in
 * real space go through the caller's transforms.
 *
 * Two reference formulas read past the end of the ky array and are therefore
 * undefined there: GAUSS_BLOB (solver.c:1892, k[1][j] for j < Ny) and the x
 * mean flow (sys_msr.c:339-347, k[1][j] for j < Nx).  Here GAUSS_BLOB uses
 * the collocation points the formula is written for and the x mean flow uses
 * each mode's own ky and |k|^2. */
#include <complex.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "ns2d_host.h"

/* initial-condition constants, data_types.h:141-165 */
#define IC_KAPPA 1.0
#define IC_SIGMA (15.0 / M_PI)
#define IC_DELTA 0.005
#define IC_BETA 1.0
#define IC_S 6.0
#define IC_DT_K0 6.0
#define IC_DT_E0 0.5
#define IC_DT2_K0 30
#define IC_DT2_E0 0.5
#define IC_DTEXP_K0 8.0
#define IC_DTEXP_E0 (343.0 / 96.0)
#define IC_RING_MIN_K 3.0
#define IC_RING_MAX_K 10.0
#define IC_EXTRM_ENS_MIN_K 1.5
#define IC_EXTRM_ENS_POW 1.5
#define IC_UNIF_MIN_K 2.5
#define IC_UNIF_MAX_K 9.5
#define IC_RAND_ENST0 1.0
#define IC_SEED 123456789u

static int say(char* msg, size_t n, const char* fmt, ...) {
  if (msg && n) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(msg, n, fmt, ap);
    va_end(ap);
  }
  return 1;
}

/* InitializeSpaceVariables (solver.c:2022-2068) for local_Nx = Nx */
static int kx_of(long i, long Nx) { return (int)(i <= Nx / 2 ? i : i - Nx); }
static double x_of(long i, long N) { return (double)i * 2.0 * M_PI / (double)N; }

static double urand(void) { return (double)rand() / (double)RAND_MAX; }

/* The MPI ranks of the reference run being reproduced.  Every rank calls
 * srand(123456789) (solver.c:1121) and walks only its own slab of kx rows, so
 * a random initial condition depends on the rank count.  The slabs are
 * FFTW's default block distribution (fftw_mpi_local_size_many with
 * FFTW_MPI_DEFAULT_BLOCK, solver.c:2615): ceil(Nx/P) rows per rank, the last
 * ranks short or empty.  A normalisation sum is formed per slab, from zero,
 * and the partial sums combined in the order of MPI_Allreduce's recursive
 * doubling for one double (the short-message algorithm of MPICH and Open
 * MPI): ranks below twice the excess over the largest power of two P2 <= P
 * first add in pairs (0+1, 2+3, ...), then the P2 values add as a pairwise
 * tree.  With P = 1 every generator is the single sequential walk. */
typedef struct ic_slabs {
  int P;
  long rows;    /* kx rows per rank */
  double* part; /* one partial sum per rank */
} ic_slabs;

/* reseed where rank i / rows starts its walk, then advance the generator by
 * the draws that rank made before this loop */
static void slab_reseed(const ic_slabs* s, long i, long skip) {
  if (i % s->rows) return;
  srand(IC_SEED);
  for (long q = 0; q < skip; ++q) (void)rand();
}

static void slab_clear(ic_slabs* s) {
  for (int r = 0; r < s->P; ++r) s->part[r] = 0.0;
}

static double* slab_part(ic_slabs* s, long i) { return &s->part[i / s->rows]; }

/* rank i / rows's rows: the draws one draw per mode leaves behind */
static long slab_len(const ic_slabs* s, long i, long Nx) { return Nx - i < s->rows ? Nx - i : s->rows; }

static double slab_allreduce(ic_slabs* s) {
  double* b = s->part;
  int p2 = 1;
  while (p2 * 2 <= s->P) p2 *= 2;
  const int rem = s->P - p2;
  /* every read index is at or above the index written, so in place is safe */
  for (int r = 0; r < p2; ++r) b[r] = r < rem ? b[2 * r] + b[2 * r + 1] : b[r + rem];
  for (int h = 1; h < p2; h *= 2)
    for (int r = 0; r + h < p2; r += 2 * h) b[r] += b[r + h];
  return b[0];
}

void ns2d_host_dealias(const ns2d_run_cfg* c, double* f_hat, int dim) {
  const long Nx = c->N[0], Ny = c->N[1], Nyf = Ny / 2 + 1;
  double complex* f = (double complex*)f_hat;
  const int houli = !strcmp(c->filter, "HOU_LI");
  const int kolm = !strcmp(c->forcing, "KOLM");
  const double kmax_sqr = pow((int)(Nx / 3.0), 2.0);  /* uses Nx for both axes */
  for (long i = 0; i < Nx; ++i) {
    const int k0 = kx_of(i, Nx);
    for (long j = 0; j < Nyf; ++j) {
      const int k1 = (int)j;
      const long idx = dim * (i * Nyf + j);
      if (houli) {
        /* integer division k/(N/2), solver.c:1088 */
        const double g = exp(-36.0 * pow(sqrt(pow((double)(k0 / (Nx / 2)), 2.0) + pow((double)(k1 / (Ny / 2)), 2.0)), 36.0));
        for (int l = 0; l < dim; ++l) f[idx + l] *= g;
      } else {
        const double k_sqr = (double)k0 * k0 + k1 * k1;
        if (k_sqr >= kmax_sqr || (kolm && k0 == 0))
          for (int l = 0; l < dim; ++l) f[idx + l] = 0.0;
      }
    }
  }
}

void ns2d_host_force_conjugacy(const ns2d_run_cfg* c, double* w_hat) {
  const long Nx = c->N[0], Nyf = c->N[1] / 2 + 1;
  double complex* w = (double complex*)w_hat;
  /* the gather reads the ky = 0 column before any row is written; the
   * positive rows it reads are never written, so in-place is the same */
  for (long i = 0; i < Nx; ++i) {
    const int k0 = kx_of(i, Nx);
    if (k0 < 0) w[i * Nyf] = conj(w[(long)abs(k0) * Nyf]);
  }
}

/* the energy sums of the decaying-turbulence initial conditions */
static double annulus_weight(long j, long Nyf) { return (j == 0 || j == Nyf - 1) ? 1.0 : 2.0; }

static int ic_real_space(const ns2d_run_cfg* c, double* w_hat, const ns2d_transforms* tf, char* msg, size_t n) {
  const long Nx = c->N[0], Ny = c->N[1];
  double* w = (double*)malloc(sizeof(double) * Nx * Ny);
  if (!w) return say(msg, n, "\n[ERROR] --- Unable to allocate memory for the [Real Space Vorticity]\n");
  for (long i = 0; i < Nx; ++i) {
    const double x = x_of(i, Nx);
    for (long j = 0; j < Ny; ++j) {
      const double y = x_of(j, Ny);
      double v;
      if (!strcmp(c->u0, "TG_VORT")) {
        v = 2.0 * IC_KAPPA * cos(IC_KAPPA * x) * cos(IC_KAPPA * y);
      } else if (!strcmp(c->u0, "DOUBLE_SHEAR_LAYER")) {
        v = IC_DELTA * cos(y) - IC_SIGMA / pow(cosh(IC_SIGMA * (x - 0.5 * M_PI)), 2.0);
        v += IC_DELTA * cos(y) + IC_SIGMA / pow(cosh(IC_SIGMA * (1.5 * M_PI - x)), 2.0);
      } else { /* GAUSS_BLOB */
        v = -exp(-(pow(x - M_PI, 2.0) + IC_BETA * pow(y - M_PI, 2.0)) / pow(2.0 * M_PI / IC_S, 2.0));
      }
      w[i * Ny + j] = v;
    }
  }
  int rc = tf->r2c(tf->ctx, w, w_hat);
  free(w);
  return rc ? say(msg, n, "\n[ERROR] --- forward transform of the initial condition failed\n") : 0;
}

static int ic_tg_vel(const ns2d_run_cfg* c, double* w_hat, const ns2d_transforms* tf, char* msg, size_t n) {
  const long Nx = c->N[0], Ny = c->N[1], Nyf = Ny / 2 + 1;
  double* u = (double*)malloc(sizeof(double) * Nx * Ny * 2);
  double complex* uh = (double complex*)malloc(sizeof(double complex) * Nx * Nyf * 2);
  if (!u || !uh) {
    free(u);
    free(uh);
    return say(msg, n, "\n[ERROR] --- Unable to allocate memory for the [Velocities]\n");
  }
  double* v = u + Nx * Ny;
  for (long i = 0; i < Nx; ++i)
    for (long j = 0; j < Ny; ++j) {
      const double x = x_of(i, Nx), y = x_of(j, Ny);
      u[i * Ny + j] = cos(IC_KAPPA * x) * sin(IC_KAPPA * y);
      v[i * Ny + j] = -sin(IC_KAPPA * x) * cos(IC_KAPPA * y);
    }
  double complex* vh = uh + Nx * Nyf;
  int rc = tf->r2c(tf->ctx, u, (double*)uh) || tf->r2c(tf->ctx, v, (double*)vh);
  free(u);
  if (rc) {
    free(uh);
    return say(msg, n, "\n[ERROR] --- forward transform of the initial velocity failed\n");
  }
  ns2d_host_dealias(c, (double*)uh, 1);
  ns2d_host_dealias(c, (double*)vh, 1);
  double complex* w = (double complex*)w_hat;
  for (long i = 0; i < Nx; ++i) {
    const int k0 = kx_of(i, Nx);
    for (long j = 0; j < Nyf; ++j) {
      const int k1 = (int)j;
      const long idx = i * Nyf + j;
      /* solver.c:1150: the condition also clears ky = 0, kx != 0 */
      if (k0 == 0 || k1 != 0) w[idx] = I * (k0 * vh[idx] - k1 * uh[idx]);
      else w[idx] = 0.0;
    }
  }
  free(uh);
  return 0;
}

static void ic_zero(const ns2d_run_cfg* c, double complex* w) {
  const long Nx = c->N[0], Nyf = c->N[1] / 2 + 1;
  const int kolm = !strcmp(c->u0, "ZERO_KOLM");
  for (long i = 0; i < Nx; ++i) {
    const int k0 = kx_of(i, Nx);
    for (long j = 0; j < Nyf; ++j) {
      const int k1 = (int)j;
      double complex v = 0.0;
      if (kolm) {
        if (k1 == 1 && k0 == 1) v = -0.5 + 0.0 * I;
        else if (k1 == 2 && (k0 == 3 || k0 == 4)) v = -4.0 + 1.0 * I;
        else if (k1 == 3 && (k0 == 4 || k0 == 5)) v = -2.0 + 3.0 * I;
        else if (k1 == 4 && (k0 == 5 || k0 == 6)) v = 8.0 - 6.0 * I;
      }
      w[i * Nyf + j] = v;
    }
  }
}

static void ic_decay_turb(const ns2d_run_cfg* c, double complex* w, ic_slabs* sl) {
  const long Nx = c->N[0], Ny = c->N[1], Nyf = Ny / 2 + 1;
  const int bb = !strcmp(c->u0, "DECAY_TURB_BB"), nb = !strcmp(c->u0, "DECAY_TURB_NB");
  for (long i = 0; i < Nx; ++i) {
    const int k0 = kx_of(i, Nx);
    slab_reseed(sl, i, 0);
    for (long j = 0; j < Nyf; ++j) {
      const int k1 = (int)j;
      if (k0 == 0 && k1 == 0) {
        w[i * Nyf + j] = 0.0;
        continue;
      }
      const double sqrt_k = sqrt((double)(k0 * k0 + k1 * k1));
      double spec_1d;
      if (bb) spec_1d = sqrt_k / (1.0 + pow(sqrt_k, 4.0) / IC_DT_K0);
      else if (nb) spec_1d = pow(sqrt_k, 6.0) / pow((1.0 + sqrt_k / (2.0 * IC_DT2_K0)), 18.0);
      else spec_1d = pow(sqrt_k, 7.0) / pow(IC_DTEXP_K0, 8.0) * cexp(-3.5 * pow(sqrt_k / IC_DTEXP_K0, 2.0));
      const double u1 = urand();
      if (!bb && !nb) w[i * Nyf + j] = sqrt(sqrt_k * spec_1d / (M_PI)) * cexp(2.0 * M_PI * u1 * I);
      else w[i * Nyf + j] = sqrt(spec_1d * (sqrt_k / (2.0 * M_PI))) * cexp(2.0 * M_PI * u1 * I);
    }
  }
  slab_clear(sl);
  for (long i = 0; i < Nx; ++i) {
    const int k0 = kx_of(i, Nx);
    double* enrg = slab_part(sl, i);
    for (long j = 0; j < Nyf; ++j) {
      const int k1 = (int)j;
      if (k0 != 0 || k1 != 0) {
        const double inv_k_sqr = 1.0 / (double)(k0 * k0 + k1 * k1);
        const double complex z = w[i * Nyf + j];
        if (j == 0 || j == Nyf - 1) *enrg += inv_k_sqr * cabs(z * conj(z));
        else *enrg += 2.0 * inv_k_sqr * cabs(z * conj(z));
      }
    }
  }
  /* solver.c:1331-1334: each rank scales its own sum before the reduction */
  for (int r = 0; r < sl->P; ++r) sl->part[r] *= (0.5 / pow(Nx * Ny, 2.0)) * 4.0 * pow(M_PI, 2.0);
  const double enrg = slab_allreduce(sl);
  const double e0 = bb ? IC_DT_E0 : nb ? IC_DT2_E0 : IC_DTEXP_E0;
  for (long i = 0; i < Nx; ++i) {
    const int k0 = kx_of(i, Nx);
    for (long j = 0; j < Nyf; ++j) {
      const int k1 = (int)j;
      /* solver.c:1344: the condition also clears ky = 0, kx != 0 */
      if (k0 == 0 || k1 != 0) w[i * Nyf + j] *= sqrt(e0 / enrg);
      else w[i * Nyf + j] = 0.0;
    }
  }
}

static int ic_decay_turb_alt(const ns2d_run_cfg* c, double complex* w, const ns2d_transforms* tf, ic_slabs* sl,
                             char* msg, size_t n) {
  const long Nx = c->N[0], Ny = c->N[1], Nyf = Ny / 2 + 1;
  double complex* psi_hat = (double complex*)malloc(sizeof(double complex) * Nx * Nyf);
  double* psi = (double*)malloc(sizeof(double) * Nx * Ny);
  if (!psi_hat || !psi) {
    free(psi_hat);
    free(psi);
    return say(msg, n, "\n[ERROR] --- Unable to allocate memory for the [Stream Function]\n");
  }
  for (long i = 0; i < Nx; ++i) {
    const int k0 = kx_of(i, Nx);
    slab_reseed(sl, i, 0);
    for (long j = 0; j < Nyf; ++j) {
      const int k1 = (int)j;
      const double u1 = urand();
      const double u2 = urand();
      const double z1 = sqrt(-2.0 * log(u1)) * cos(2.0 * M_PI * u2);
      const double z2 = sqrt(-2.0 * log(u1)) * sin(2.0 * M_PI * u2);
      if (k0 == 0 && k1 == 0) {
        psi_hat[i * Nyf + j] = 0.0;
      } else {
        const double k_sqrt = sqrt((double)(k0 * k0 + k1 * k1));
        const double spec_1d = 1.0 / (k_sqrt * (1.0 + pow(k_sqrt / 6.0, 4.0)));
        psi_hat[i * Nyf + j] = (z1 + z2 * I) * sqrt(spec_1d * pow(k_sqrt, 2.0) / (pow(k_sqrt, 3.0) * 2.0 * M_PI));
      }
    }
  }
  /* zero-mean stream function: c2r, normalise, subtract the mean, r2c */
  if (tf->c2r(tf->ctx, (const double*)psi_hat, psi)) {
    free(psi_hat);
    free(psi);
    return say(msg, n, "\n[ERROR] --- inverse transform of the stream function failed\n");
  }
  /* real-space rows are slabbed as the kx rows are */
  slab_clear(sl);
  for (long i = 0; i < Nx; ++i) {
    double* mean = slab_part(sl, i);
    for (long j = 0; j < Ny; ++j) {
      psi[i * Ny + j] /= (Nx * Ny);
      *mean += psi[i * Ny + j];
    }
  }
  const double mean = slab_allreduce(sl);
  for (long q = 0; q < Nx * Ny; ++q) psi[q] -= (mean / (Nx * Ny));
  int rc = tf->r2c(tf->ctx, psi, (double*)psi_hat);
  free(psi);
  if (rc) {
    free(psi_hat);
    return say(msg, n, "\n[ERROR] --- forward transform of the stream function failed\n");
  }
  slab_clear(sl);
  for (long i = 0; i < Nx; ++i) {
    const int k0 = kx_of(i, Nx);
    double* enrg = slab_part(sl, i);
    for (long j = 0; j < Nyf; ++j) {
      const int k1 = (int)j;
      const double k_sqrd = (double)(k0 * k0 + k1 * k1);
      const double complex z = psi_hat[i * Nyf + j];
      if (j == 0 || j == Nyf - 1) *enrg += k_sqrd * cabs(z * conj(z));
      else *enrg += 2.0 * k_sqrd * cabs(z * conj(z));
    }
  }
  /* solver.c:1487-1490: scaled per rank, then reduced */
  for (int r = 0; r < sl->P; ++r) sl->part[r] *= 4.0 * pow(M_PI, 2.0) * (0.5 / pow(Nx * Ny, 2.0));
  const double enrg = slab_allreduce(sl);
  for (long i = 0; i < Nx; ++i) {
    const int k0 = kx_of(i, Nx);
    for (long j = 0; j < Nyf; ++j) {
      const int k1 = (int)j;
      psi_hat[i * Nyf + j] *= sqrt(0.5 / enrg);
      const double k_sqrd = (double)(k0 * k0 + k1 * k1);
      if (k0 != 0 || k1 != 0) w[i * Nyf + j] = -k_sqrd * psi_hat[i * Nyf + j];
      else w[i * Nyf + j] = 0.0;
    }
  }
  free(psi_hat);
  return 0;
}

static int ic_gauss_decay_turb(const ns2d_run_cfg* c, double complex* w, ic_slabs* sl, char* msg, size_t n) {
  const long Nx = c->N[0], Ny = c->N[1], Nyf = Ny / 2 + 1;
  double complex* psi_hat = (double complex*)malloc(sizeof(double complex) * Nx * Nyf);
  if (!psi_hat) return say(msg, n, "\n[ERROR] --- Unable to allocate memory for the [Stream Function]\n");
  for (long i = 0; i < Nx; ++i) {
    const int k0 = kx_of(i, Nx);
    slab_reseed(sl, i, 0);
    for (long j = 0; j < Nyf; ++j) {
      const int k1 = (int)j;
      if (k0 == 0 && k1 == 0) {
        psi_hat[i * Nyf + j] = 0.0;
      } else {
        const double k_sqrt = sqrt((double)(k0 * k0 + k1 * k1));
        const double spec_1d = pow(k_sqrt, 6.0) / pow((1.0 + k_sqrt / 60.0), 18.0);
        const double u1 = urand();
        psi_hat[i * Nyf + j] = sqrt(spec_1d * (1.0 / k_sqrt)) * cexp(I * 2.0 * M_PI * u1);
      }
    }
  }
  slab_clear(sl);
  for (long i = 0; i < Nx; ++i) {
    const int k0 = kx_of(i, Nx);
    double* enrg = slab_part(sl, i);
    for (long j = 0; j < Nyf; ++j) {
      const int k1 = (int)j;
      if (k0 != 0 || k1 != 0) {
        const double k_sqrd = (double)(k0 * k0 + k1 * k1);
        const double complex z = psi_hat[i * Nyf + j];
        if (j == 0 || j == Nyf - 1) *enrg += k_sqrd * cabs(z * conj(z));
        else *enrg += 2.0 * k_sqrd * cabs(z * conj(z));
      }
    }
  }
  /* solver.c:1578-1581: scaled per rank, then reduced */
  for (int r = 0; r < sl->P; ++r) sl->part[r] *= 4.0 * pow(M_PI, 2.0) * (0.5 / pow(Nx * Ny, 2.0));
  const double enrg = slab_allreduce(sl);
  for (long i = 0; i < Nx; ++i) {
    const int k0 = kx_of(i, Nx);
    for (long j = 0; j < Nyf; ++j) {
      const int k1 = (int)j;
      psi_hat[i * Nyf + j] /= sqrt(enrg);
      psi_hat[i * Nyf + j] *= sqrt(0.5);
      const double k_sqrd = (double)(k0 * k0 + k1 * k1);
      if (k0 != 0 || k1 != 0) w[i * Nyf + j] = k_sqrd * psi_hat[i * Nyf + j];
      else w[i * Nyf + j] = 0.0;
    }
  }
  free(psi_hat);
  return 0;
}

static void ic_extrm_ens(const ns2d_run_cfg* c, double complex* w, ic_slabs* sl) {
  const long Nx = c->N[0], Nyf = c->N[1] / 2 + 1;
  slab_clear(sl);
  for (long i = 0; i < Nx; ++i) {
    const int k0 = kx_of(i, Nx);
    double* norm = slab_part(sl, i);
    for (long j = 0; j < Nyf; ++j) {
      const int k1 = (int)j;
      const double k_sqr = (double)(k0 * k0 + k1 * k1);
      /* edge columns once, interior twice; the factor 2 multiplies first */
      const double wgt = annulus_weight(j, Nyf);
      if ((k_sqr > 0) && (sqrt(k_sqr) > IC_EXTRM_ENS_MIN_K)) {
        const double p = pow(sqrt(k_sqr), -IC_EXTRM_ENS_POW);
        *norm += (wgt == 1.0) ? p : 2.0 * p;
      } else if (sqrt(k_sqr) <= IC_EXTRM_ENS_MIN_K) {
        const double p = pow(IC_EXTRM_ENS_MIN_K, -IC_EXTRM_ENS_POW);
        const double e = exp((sqrt(k_sqr) - IC_EXTRM_ENS_MIN_K) / 3.5);
        *norm += (wgt == 1.0) ? p * e : 2.0 * p * e;
      }
    }
  }
  const double norm = sqrt(IC_EXTRM_ENS_MIN_K / slab_allreduce(sl));
  for (long i = 0; i < Nx; ++i) {
    const int k0 = kx_of(i, Nx);
    slab_reseed(sl, i, 0);
    for (long j = 0; j < Nyf; ++j) {
      const int k1 = (int)j;
      const double k_sqr = (double)(k0 * k0 + k1 * k1);
      if ((k_sqr > 0) && (sqrt(k_sqr) > IC_EXTRM_ENS_MIN_K)) {
        const double u1 = urand();
        w[i * Nyf + j] = pow(sqrt(k_sqr), -IC_EXTRM_ENS_POW / 2.0) * norm * cexp(I * 2.0 * M_PI * u1);
      } else if (sqrt(k_sqr) <= IC_EXTRM_ENS_MIN_K) {
        const double u1 = urand();
        w[i * Nyf + j] = pow(IC_EXTRM_ENS_MIN_K, -IC_EXTRM_ENS_POW / 2.0) *
                         exp((sqrt(k_sqr) - IC_EXTRM_ENS_MIN_K) / 3.5) * norm * cexp(I * 2.0 * M_PI * u1);
      } else {
        w[i * Nyf + j] = 0.0;
      }
    }
  }
}

static void ic_ring(const ns2d_run_cfg* c, double complex* w, ic_slabs* sl) {
  const long Nx = c->N[0], Nyf = c->N[1] / 2 + 1;
  const double mid = (IC_RING_MIN_K + IC_RING_MAX_K) / 2.0;
  slab_clear(sl);
  for (long i = 0; i < Nx; ++i) {
    const int k0 = kx_of(i, Nx);
    double* norm = slab_part(sl, i);
    slab_reseed(sl, i, 0);
    for (long j = 0; j < Nyf; ++j) {
      const int k1 = (int)j;
      const double abs_k = sqrt((double)(k0 * k0 + k1 * k1));
      (void)urand();  /* solver.c:1690: drawn and discarded for every mode */
      if ((abs_k >= IC_RING_MIN_K && abs_k <= IC_RING_MAX_K) && (k1 != 0 || k0 > 0))
        *norm += exp(-pow(abs_k - mid, 2.0));
    }
  }
  const double norm = slab_allreduce(sl);
  for (long i = 0; i < Nx; ++i) {
    const int k0 = kx_of(i, Nx);
    /* each rank's generator continues after its slab's discarded draws */
    slab_reseed(sl, i, slab_len(sl, i, Nx) * Nyf);
    for (long j = 0; j < Nyf; ++j) {
      const int k1 = (int)j;
      const double abs_k = sqrt((double)(k0 * k0 + k1 * k1));
      if ((abs_k >= IC_RING_MIN_K && abs_k <= IC_RING_MAX_K) && (k1 != 0 || k0 > 0)) {
        const double u1 = urand();
        w[i * Nyf + j] = sqrt(72.0 * exp(-pow(abs_k - mid, 2.0)) / (M_PI * abs_k * norm)) * cexp(2.0 * M_PI * I * u1);
      } else {
        w[i * Nyf + j] = 0.0;
      }
    }
  }
}

static void ic_unif(const ns2d_run_cfg* c, double complex* w, ic_slabs* sl) {
  const long Nx = c->N[0], Nyf = c->N[1] / 2 + 1;
  slab_clear(sl);
  for (long i = 0; i < Nx; ++i) {
    const int k0 = kx_of(i, Nx);
    double* norm = slab_part(sl, i);
    for (long j = 0; j < Nyf; ++j) {
      const int k1 = (int)j;
      const double k_abs = sqrt((double)(k0 * k0 + k1 * k1));
      if (k_abs > IC_UNIF_MIN_K && k_abs < IC_UNIF_MAX_K) *norm += annulus_weight(j, Nyf);
    }
  }
  const double norm = sqrt(1.0 / slab_allreduce(sl));
  for (long i = 0; i < Nx; ++i) {
    const int k0 = kx_of(i, Nx);
    slab_reseed(sl, i, 0);
    for (long j = 0; j < Nyf; ++j) {
      const int k1 = (int)j;
      const double k_abs = sqrt((double)(k0 * k0 + k1 * k1));
      if (k_abs > IC_UNIF_MIN_K && k_abs < IC_UNIF_MAX_K) {
        const double u1 = urand();
        w[i * Nyf + j] = norm * cexp(2.0 * M_PI * I * u1);
      } else {
        w[i * Nyf + j] = 0.0;
      }
    }
  }
}

static void ic_random(const ns2d_run_cfg* c, double complex* w, ic_slabs* sl) {
  const long Nx = c->N[0], Nyf = c->N[1] / 2 + 1;
  /* solver.c:1911 leaves the sum uninitialised; it starts at zero here */
  slab_clear(sl);
  for (long i = 0; i < Nx; ++i) {
    const int k0 = kx_of(i, Nx);
    double* enst_norm = slab_part(sl, i);
    slab_reseed(sl, i, 0);
    for (long j = 0; j < Nyf; ++j) {
      const int k1 = (int)j;
      const double sqrt_k = sqrt(k0 * k0 + k1 * k1);
      if (sqrt_k > IC_UNIF_MIN_K && sqrt_k < IC_UNIF_MAX_K) {
        const double enrg_k = 1.0 / sqrt_k;
        const double enst_k = pow(sqrt_k, 2.0) * enrg_k;
        const double r1 = urand();
        w[i * Nyf + j] = sqrt(enst_k / enrg_k) * cexp(r1 * 2.0 * M_PI * I);
      } else {
        w[i * Nyf + j] = 0.0;
      }
      *enst_norm += pow(cabs(w[i * Nyf + j]), 2.0);
    }
  }
  const double enst_norm = slab_allreduce(sl) * 2.0;
  for (long i = 0; i < Nx; ++i) {
    const int k0 = kx_of(i, Nx);
    for (long j = 0; j < Nyf; ++j) {
      const int k1 = (int)j;
      if (k0 != 0 || k1 != 0) w[i * Nyf + j] *= sqrt(IC_RAND_ENST0 / enst_norm);
      else w[i * Nyf + j] = 0.0;
    }
  }
}

static void ic_testing(const ns2d_run_cfg* c, double complex* w) {
  const long Nx = c->N[0], Nyf = c->N[1] / 2 + 1;
  for (long i = 0; i < Nx; ++i) {
    const int k0 = kx_of(i, Nx);
    for (long j = 0; j < Nyf; ++j) {
      const int k1 = (int)j;
      if (k0 == 0 && k1 == 0) {
        w[i * Nyf + j] = 0.0;
      } else {
        const double inv_k_sqr = 1.0 / (double)(k0 * k0 + k1 * k1);
        if (j == 0 && k0 < 0) w[i * Nyf + j] = inv_k_sqr * cexp(-I * M_PI / 4.0);
        else w[i * Nyf + j] = inv_k_sqr * cexp(I * M_PI / 4.0);
      }
    }
  }
}

static void ic_none(const ns2d_run_cfg* c, double complex* w, ic_slabs* sl) {
  const long Nx = c->N[0], Nyf = c->N[1] / 2 + 1;
  for (long i = 0; i < Nx; ++i) {
    slab_reseed(sl, i, 0);
    for (long j = 0; j < Nyf; ++j) {
      /* solver.c:1986: amplitude drawn before the phase (gcc evaluates the
       * left operand of the product first) */
      const double amp = urand();
      const double ph = urand();
      w[i * Nyf + j] = amp * cexp(ph * 2.0 * M_PI * I);
    }
  }
}

static int ic_max_palin(const ns2d_run_cfg* c, double complex* w, const ns2d_transforms* tf, char* msg, size_t n) {
  const long Nx = c->N[0], Ny = c->N[1], Nyf = Ny / 2 + 1;
  char path[1200];
  snprintf(path, sizeof path, "%sMaxPalinstrophy/maxdpdt_P10000_N%ld_IG0_psi_f.h5", c->input_dir, Nx);
  double* psi = (double*)malloc(sizeof(double) * Nx * Ny);
  if (!psi) return say(msg, n, "\n[ERROR] --- Unable to allocate memory for the [Stream Function]\n");
  if (ns2d_h5_read_psi(path, psi, Nx * Ny, msg, n)) {
    free(psi);
    return 1;
  }
  int rc = tf->r2c(tf->ctx, psi, (double*)w);
  free(psi);
  if (rc) return say(msg, n, "\n[ERROR] --- forward transform of the stream function failed\n");
  for (long i = 0; i < Nx; ++i) {
    const int k0 = kx_of(i, Nx);
    for (long j = 0; j < Nyf; ++j) {
      const int k1 = (int)j;
      w[i * Nyf + j] = -(double)(k0 * k0 + k1 * k1) * w[i * Nyf + j];
    }
  }
  return 0;
}

int ns2d_initial_condition(const ns2d_run_cfg* c, double* w_hat, const ns2d_transforms* tf, char* msg, size_t n) {
  double complex* w = (double complex*)w_hat;
  const char* u0 = c->u0;
  const int P = c->ic_ranks;
  if (P < 1 || P > NS2D_IC_MAX_RANKS)
    return say(msg, n, "\n[ERROR] --- emulated rank count [%d] must lie in 1..%d\n", P, NS2D_IC_MAX_RANKS);
  ic_slabs sl = {P, (c->N[0] + P - 1) / P, (double*)calloc((size_t)P, sizeof(double))};
  if (!sl.part) return say(msg, n, "\n[ERROR] --- Unable to allocate memory for the rank sums\n");
  srand(IC_SEED);
  int rc = 0;
  if (!strcmp(u0, "TG_VEL")) rc = ic_tg_vel(c, w_hat, tf, msg, n);
  else if (!strcmp(u0, "TG_VORT") || !strcmp(u0, "DOUBLE_SHEAR_LAYER") || !strcmp(u0, "GAUSS_BLOB"))
    rc = ic_real_space(c, w_hat, tf, msg, n);
  else if (!strcmp(u0, "ZERO") || !strcmp(u0, "ZERO_KOLM")) ic_zero(c, w);
  else if (!strcmp(u0, "DECAY_TURB_BB") || !strcmp(u0, "DECAY_TURB_NB") || !strcmp(u0, "DECAY_TURB_EXP"))
    ic_decay_turb(c, w, &sl);
  else if (!strcmp(u0, "DECAY_TURB_ALT")) rc = ic_decay_turb_alt(c, w, tf, &sl, msg, n);
  else if (!strcmp(u0, "GAUSS_DECAY_TURB")) rc = ic_gauss_decay_turb(c, w, &sl, msg, n);
  else if (!strcmp(u0, "EXTRM_ENS")) ic_extrm_ens(c, w, &sl);
  else if (!strcmp(u0, "RING")) ic_ring(c, w, &sl);
  else if (!strcmp(u0, "UNIF")) ic_unif(c, w, &sl);
  else if (!strcmp(u0, "MAX_PALIN")) rc = ic_max_palin(c, w, tf, msg, n);
  else if (!strcmp(u0, "RANDOM")) ic_random(c, w, &sl);
  else if (!strcmp(u0, "TESTING")) ic_testing(c, w);
  else ic_none(c, w, &sl);
  free(sl.part);
  if (rc) return rc;
  ns2d_host_dealias(c, w_hat, 1);
  ns2d_host_force_conjugacy(c, w_hat);
  return 0;
}

void ns2d_mean_flows(const ns2d_run_cfg* c, const double* w_hat, double* mean_flow_y, double* mean_flow_x) {
  const long Nx = c->N[0], Nyf = c->N[1] / 2 + 1;
  const double complex* w = (const double complex*)w_hat;
  for (long i = 0; i < Nx; ++i) {
    const int k0 = kx_of(i, Nx);
    for (long j = 0; j < Nyf; ++j) {
      const int k1 = (int)j;
      if (k0 == 0 && k1 == 0) continue;
      const double k_sqr = (double)(k0 * k0 + k1 * k1);
      const double complex v_z = -I * ((double)k0) * w[i * Nyf + j] / k_sqr;
      const double complex u_z = I * ((double)k1) * w[i * Nyf + j] / k_sqr;
      const double wgt = annulus_weight(j, Nyf);
      if (mean_flow_y) mean_flow_y[i] += (wgt == 1.0) ? cabs(v_z) : 2.0 * cabs(v_z);
      if (mean_flow_x) mean_flow_x[j] += (wgt == 1.0) ? cabs(u_z) : 2.0 * cabs(u_z);
    }
  }
}
