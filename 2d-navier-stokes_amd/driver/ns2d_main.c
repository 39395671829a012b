/* ns2d_solver: the reference's solver executable on one MI355X.
 *
 * Same command line as Solver/bin/main (GetCMLArgs, utils.c:34-461), same
 * initial conditions (solver.c:1107-2014), the same time loop with its save,
 * transient and print logic (SpectralSolve, solver.c:54-251) and the same
 * HDF5 output (hdf5_funcs.c); the time steps, the per-step blow-up check
 * and the measures run on the GPU through include/ns2d_gpu.h.
 *
 * Between two saves the driver hands the whole run of steps to ns2d_step,
 * which checks max|w| after every step on the device; a blow-up is reported
 * at the step the reference would have stopped at. */
#define _GNU_SOURCE
#include <complex.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "ns2d_gpu.h"
#include "ns2d_host.h"

#define MIN_STEP_SIZE 1e-10  /* data_types.h:170 */
#define MAX_ITERS 1e+12      /* data_types.h:171 */

static char g_msg[4096];

static void die(const char* what) {
  fprintf(stderr, "%s", what);
  if (what[0] && what[strlen(what) - 1] != '\n') fputc('\n', stderr);
  exit(1);
}

static void gpu_check(int rc, const char* where) {
  if (rc == NS2D_OK) return;
  fprintf(stderr, "\n[ERROR] --- %s: %s\n-->> Exiting!!!\n", where, ns2d_last_error());
  exit(1);
}

static int tf_r2c(void* ctx, const double* in, double* out) { return ns2d_r2c((ns2d_t*)ctx, in, out); }
static int tf_c2r(void* ctx, const double* in, double* out) { return ns2d_c2r((ns2d_t*)ctx, in, out); }

typedef struct run {
  ns2d_run_cfg c;
  ns2d_t* h;
  long Nx, Ny, Nyf;
  int n_spect;
  double *w_hat, *w, *tg, *enrg_spect, *enst_spect, *flux_spect;
  ns2d_series s;
  long num_t_steps, num_print_steps, trans_steps;
  long num_sys_msr_counts;
} run;

static double* alloc_d(long n) {
  double* p = (double*)calloc((size_t)(n > 0 ? n : 1), sizeof(double));
  if (!p) die("\n[ERROR] --- Unable to allocate host memory\n-->> Exiting!!!\n");
  return p;
}

/* ComputeSystemMeasurables (sys_msr.c:32-373) into series slot idx */
static void measure(run* r, double t, long idx) {
  ns2d_measures_t m;
  gpu_check(ns2d_measures(r->h, 1, &m, r->enrg_spect, r->enst_spect, r->flux_spect), "system measures");
  ++r->num_sys_msr_counts;
  if (idx < r->num_print_steps) {
    ns2d_series* s = &r->s;
    s->time[idx] = t;
    s->tot_energy[idx] = m.energy;
    s->tot_enstr[idx] = m.enstrophy;
    s->tot_palin[idx] = m.palinstrophy;
    s->tot_forc[idx] = m.forcing;
    s->tot_div[idx] = m.divergence;
    s->enrg_diss[idx] = m.enrg_diss;
    s->enst_diss[idx] = m.enst_diss;
    s->d_enst_dt_sbst[idx] = m.d_enst_dt_sbst;
    s->enst_flux_sbst[idx] = m.enst_flux_sbst;
    s->enst_diss_sbst[idx] = m.enst_diss_sbst;
    s->d_enrg_dt_sbst[idx] = m.d_enrg_dt_sbst;
    s->enrg_flux_sbst[idx] = m.enrg_flux_sbst;
    s->enrg_diss_sbst[idx] = m.enrg_diss_sbst;
  } else if (idx % 100 == 0) {
    printf("\n[WARNING] --- Unable to write system measures at Indx: [%ld] t: [%lf] ---- Number of intergration steps "
           "is now greater then memory allocated\n",
           idx, t);
  }
  gpu_check(ns2d_get_state(r->h, r->w_hat), "state download");
  ns2d_mean_flows(&r->c, r->w_hat, idx < r->num_print_steps ? r->s.mean_flow_y : NULL, r->s.mean_flow_x);
}

/* TaylorGreenSoln (solver.c:2560-2596) */
static void tg_solution(run* r, double t) {
  const int vel = !strcmp(r->c.u0, "TG_VEL");
  for (long i = 0; i < r->Nx; ++i)
    for (long j = 0; j < r->Ny; ++j) {
      const double x = (double)i * 2.0 * M_PI / (double)r->Nx, y = (double)j * 2.0 * M_PI / (double)r->Ny;
      r->tg[i * r->Ny + j] = vel ? -2.0 * sin(x) * sin(y) * exp(-2.0 * r->c.nu * t)
                                 : 2.0 * cos(x) * cos(y) * exp(-2.0 * r->c.nu * t);
    }
}

static int tg_run(const run* r) {
  return r->c.testing && (!strcmp(r->c.u0, "TG_VEL") || !strcmp(r->c.u0, "TG_VORT"));
}

/* WriteDataToFile (hdf5_funcs.c:650-1045) */
static void write_save(run* r, ns2d_h5* f, double t, double dt, long idx) {
  gpu_check(ns2d_get_state(r->h, r->w_hat), "state download");
  gpu_check(ns2d_real_vort(r->h, r->w), "real-space vorticity");
  ns2d_save sv;
  memset(&sv, 0, sizeof sv);
  sv.t = t;
  sv.dt = dt;
  sv.index = (int)idx;
  sv.w = r->w;
  sv.w_hat = r->w_hat;
  if (tg_run(r)) {
    tg_solution(r, t);
    sv.tg_soln = r->tg;
  }
  sv.enst_spect = r->enst_spect;
  sv.enrg_spect = r->enrg_spect;
  sv.flux_spect = r->flux_spect;
  if (ns2d_h5_write_save(f, &sv, g_msg, sizeof g_msg)) die(g_msg);
}

/* PrintUpdateToTerminal (solver.c:2336-2370) */
static void print_update(run* r, long iters, double t, double dt, double T, long idx) {
  double max_vort = 0.0;
  gpu_check(ns2d_max_vort(r->h, &max_vort), "max vorticity");
  const ns2d_series* s = &r->s;
  if (idx >= r->num_print_steps) idx = r->num_print_steps - 1;
  if (tg_run(r)) {
    /* TestTaylorGreenVortex (solver.c:2487-2540) */
    gpu_check(ns2d_real_vort(r->h, r->w), "real-space vorticity");
    tg_solution(r, t);
    double l2 = 0.0, linf = 0.0;
    for (long q = 0; q < r->Nx * r->Ny; ++q) {
      const double e = fabs(r->w[q] - r->tg[q]);
      l2 += e * e;
      linf = fmax(e, linf);
    }
    l2 = sqrt(l2 / (double)(r->Nx * r->Ny));
    printf("Iter: %ld\tt: %1.6lf/%1.3lf\tdt: %1.6g \tMax Vort: %1.5g \tKE: %1.5g\tENS: %1.5g\tPAL: %1.5g\tL2: %1.5g\tLinf: "
           "%1.5g\n",
           iters, t, T, dt, max_vort, s->tot_energy[idx], s->tot_enstr[idx], s->tot_palin[idx], l2, linf);
  } else {
    printf("Iter: %ld/%ld\tt: %1.6lf/%1.3lf\tdt: %1.6g \tMax Vort: %1.5g \tKE: %1.5g\tENS: %1.5g\t PAL: %1.5g\t E_Diss: "
           "%1.5g\tEns_Diss: %1.5g\n",
           iters, r->num_t_steps, t, T, dt, max_vort, s->tot_energy[idx], s->tot_enstr[idx], s->tot_palin[idx],
           s->enrg_diss[idx], s->enst_diss[idx]);
  }
  fflush(stdout);
}

/* GetTimestep (solver.c:2376-2462).  The CFL branch uses the velocity of
 * every mode (the reference's row offset at solver.c:2394 drops it); its
 * convective rate max pi (|u|/dx + |v|/dy) is reduced on the GPU. */
static double next_dt(run* r, double dt, double w_max_init) {
  const ns2d_run_cfg* c = &r->c;
  if (c->cfl_cond_flag == 1) {
    const double dx = 2.0 * M_PI / (double)r->Nx, dy = 2.0 * M_PI / (double)r->Ny;
    double conv = 0.0;
    gpu_check(ns2d_cfl_rate(r->h, &conv), "CFL rate");
    const double diff = pow(M_PI, 2.0) * (c->nu / pow(dx, 2.0) + c->nu / pow(dy, 2.0));
    return c->cfl_const / (conv + diff);
  }
  double w_max = 0.0;
  gpu_check(ns2d_max_vort(r->h, &w_max), "max vorticity");
  const double dt_new = c->dt * (w_max_init / w_max);
  return dt_new > 2.0 * dt ? 2.0 * dt : dt_new;
}

static void write_details(const run* r, const ns2d_out_paths* p, int argc, char** argv, double secs) {
  char path[1300];
  snprintf(path, sizeof path, "%sSimulationDetails.txt", p->dir);
  FILE* f = fopen(path, "w");
  if (!f) return;
  fprintf(f, "Executing Command:\n\n\t");
  for (int i = 0; i < argc; ++i) fprintf(f, "%s ", argv[i]);
  fprintf(f, "\n\nSystem Type: NAVIER\nSolver Type: %s\nModel Type: FULL\n", r->c.integrator);
  fprintf(f, "Viscosity: %1.6lf\nRe: %5.1lf\n", r->c.nu, 1.0 / r->c.nu);
  fprintf(f, "Ekman Alpha: %1.4lf\nEkman Drag Flag: %d\nEkman Power: %1.1lf\n", r->c.ekmn_alpha, r->c.ekmn_drag_flag,
          r->c.ekmn_drag_pow);
  fprintf(f, "Hyperviscosity Flag: %d\nHyperviscosity Power: %1.1lf\n", r->c.hyper_visc_flag, r->c.hyper_visc_pow);
  fprintf(f, "\nCollocation Points: [%ld, %ld]\n", r->Nx, r->Ny);
  fprintf(f, "Spatial Increment: [%1.4lf, %1.4lf]\n", 2.0 * M_PI / r->Nx, 2.0 * M_PI / r->Ny);
  fprintf(f, "Fourier Modes: [%ld, %ld]\n\n", r->Nx, r->Nyf);
  fprintf(f, "Initial Conditions: %s\nForcing: %s [%d, %g]\nDealiasing: %s\n", r->c.u0, r->c.forcing, r->c.force_k,
          r->c.force_scale_var, r->c.filter);
  fprintf(f, "\nStart Time: %1.6lf\nEnd Time: %1.6lf\nTimestep: %1.6g\nTotal Iterations: %ld\nSaving Steps: %ld\n", r->c.t0,
          r->c.T, r->c.dt, r->num_t_steps, r->num_print_steps);
  fprintf(f, "\nDevice: MI355X (HIP)\nExecution Time: %5.10lf s\n", secs);
  fclose(f);
}

int main(int argc, char** argv) {
  struct timespec t_begin, t_end;
  clock_gettime(CLOCK_MONOTONIC, &t_begin);
  run R;
  memset(&R, 0, sizeof R);
  run* r = &R;
  ns2d_cli_defaults(&r->c);
  if (ns2d_cli_parse(argc, argv, &r->c, g_msg, sizeof g_msg)) {
    fprintf(stderr, "%s", g_msg);
    fprintf(stderr, "\n[ERROR]: Error in reading in command line aguments, check utils.c file for details\n");
    return 1;
  }
  ns2d_run_cfg* c = &r->c;
  if (strcmp(c->forcing, "NONE") && strcmp(c->forcing, "KOLM")) {
    fprintf(stderr, "\n[ERROR] --- Forcing [%s] is not provided by this solver (NONE and KOLM are)\n-->> Exiting!!!\n",
            c->forcing);
    return 1;
  }
  r->Nx = c->N[0];
  r->Ny = c->N[1];
  r->Nyf = r->Ny / 2 + 1;

  ns2d_params p;
  gpu_check(ns2d_params_default(&p, (int)r->Nx, (int)r->Ny), "parameters");
  p.nu = c->nu;
  p.hyper_flag = c->hyper_visc_flag;
  p.hyper_pow = c->hyper_visc_pow;
  p.alpha = c->ekmn_alpha;
  p.drag_flag = c->ekmn_drag_flag;
  p.drag_pow = c->ekmn_drag_pow;
  p.filter = !strcmp(c->filter, "HOU_LI") ? NS2D_FILTER_HOULI_REF : NS2D_FILTER_23;
  p.forcing = !strcmp(c->forcing, "KOLM") ? NS2D_FORCING_KOLM : NS2D_FORCING_NONE;
  p.force_k = c->force_k;
  p.force_scale = c->force_scale_var;
  p.integrator = !strcmp(c->integrator, "AB4") ? NS2D_INT_AB4 : NS2D_INT_RK4CN;
  p.quirks = c->quirks;
  p.check_every = 1;
  gpu_check(ns2d_create(&p, &r->h), "solver creation");
  ns2d_info_t info;
  gpu_check(ns2d_get_info(r->h, &info), "solver info");
  r->n_spect = info.n_spect;

  r->w_hat = alloc_d(2 * r->Nx * r->Nyf);
  r->w = alloc_d(r->Nx * r->Ny);
  r->tg = alloc_d(r->Nx * r->Ny);
  r->enrg_spect = alloc_d(r->n_spect);
  r->enst_spect = alloc_d(r->n_spect);
  r->flux_spect = alloc_d(6L * r->n_spect);

  /* InitialConditions + InitializeForcing (solver.c:89-92) */
  ns2d_transforms tf = {r->h, tf_r2c, tf_c2r};
  if (ns2d_initial_condition(c, r->w_hat, &tf, g_msg, sizeof g_msg)) die(g_msg);
  gpu_check(ns2d_set_state(r->h, r->w_hat), "state upload");
  double w_max_init = 0.0;
  gpu_check(ns2d_max_vort(r->h, &w_max_init), "max vorticity");

  /* InitializeIntegrationVariables (solver.c:2116-2166) */
  if (c->adapt_step_flag == 1) c->dt = next_dt(r, c->dt, w_max_init);
  const double t0 = c->t0, T = c->T;
  double t = c->t0, dt = c->dt;
  r->num_t_steps = (long)((T - t0) / dt);
  if (c->trans_iters_flag == 1) {
    r->trans_steps = (long)(c->trans_iters_frac * r->num_t_steps);
    r->num_print_steps = (r->num_t_steps >= c->save_every) ? (r->num_t_steps - r->trans_steps) / c->save_every
                                                           : r->num_t_steps - r->trans_steps;
    printf("Total Iters: %ld\t Saving Iters: %ld\t Transient Steps: %ld\n", r->num_t_steps, r->num_print_steps,
           r->trans_steps);
  } else {
    r->trans_steps = 0;
    r->num_print_steps = (r->num_t_steps >= c->save_every) ? r->num_t_steps / c->save_every + 1 : r->num_t_steps + 1;
    printf("Total Iters: %ld\t Saving Iters: %ld\n", r->num_t_steps, r->num_print_steps);
  }
  /* InitializeSystemMeasurables (sys_msr.c:378-389) */
  if (c->adapt_step_flag == 1) r->num_print_steps *= 2;
  if (r->num_print_steps < 1) r->num_print_steps = 1;
  ns2d_series* s = &r->s;
  s->n = r->num_print_steps;
  double** arrays[] = {&s->time,           &s->tot_energy,     &s->tot_enstr,      &s->tot_palin,     &s->tot_forc,
                       &s->tot_div,        &s->enrg_diss,      &s->enst_diss,      &s->d_enst_dt_sbst, &s->enst_flux_sbst,
                       &s->enst_diss_sbst, &s->d_enrg_dt_sbst, &s->enrg_flux_sbst, &s->enrg_diss_sbst};
  for (size_t q = 0; q < sizeof arrays / sizeof arrays[0]; ++q) *arrays[q] = alloc_d(s->n);
  s->mean_flow_y = alloc_d(r->Nx);
  s->mean_flow_x = alloc_d(r->Nyf);
  measure(r, 0.0, 0);

  /* CreateOutputFilesWriteICs (hdf5_funcs.c:33-420) */
  ns2d_out_paths paths;
  if (ns2d_output_paths(c, 1, &paths, g_msg, sizeof g_msg)) die(g_msg);
  printf("\nMain Output File: %s\n\n", paths.main_file);
  ns2d_h5* f = ns2d_h5_create(c, &paths, r->n_spect, g_msg, sizeof g_msg);
  if (!f) die(g_msg);
  if (c->trans_iters_flag != 1) write_save(r, f, 0.0, dt, 0);
  print_update(r, 0, t0, dt, T, 0);

  /* SpectralSolve loop (solver.c:142-236) */
  t += dt;
  long iters = 1;
  long save_idx = (c->trans_iters_flag == 1) ? 0 : 1;
  const int adaptive = c->adapt_step_flag == 1;
  while (t <= T) {
    /* steps up to and including the next save step (one at a time when the
     * step size adapts) */
    long n = 1;
    if (!adaptive)
      while (iters + n - 1 < (long)MAX_ITERS && (iters + n - 1) % c->save_every != 0 && (iters + n) * dt <= T) ++n;
    const long last = iters + n - 1;
    int rc = ns2d_step(r->h, dt, n);
    long bad = -1;
    if (rc == NS2D_ERR_BLOWUP) {
      gpu_check(ns2d_get_info(r->h, &info), "solver info");
      bad = info.first_blowup_iter;
    } else {
      gpu_check(rc, "time step");
    }
    iters = last;
    t = adaptive ? t : iters * dt;
    if (bad >= 0 && bad < last) {
      /* the reference stops at SystemCheck right after step `bad` */
      fprintf(stderr, "\n[SOVLER FAILURE] --- System has reached maximum Vorticity limt at Iter: [%ld]\n-->> Exiting!!!n",
              bad + 1);
      ns2d_h5_close(f);
      return 1;
    }
    if (iters % c->save_every == 0) {
      measure(r, t, save_idx);
      if (iters > r->trans_steps) {
        write_save(r, f, t, dt, save_idx);
        ++save_idx;
      }
    }
    if (c->trans_iters_flag == 1 && iters == r->trans_steps) printf("\n\n...Transient Iterations Complete!\n\n");
    if (iters % c->save_every == 0) print_update(r, iters, t, dt, T, iters <= r->trans_steps ? 0 : save_idx - 1);
    ++iters;
    if (adaptive) {
      dt = next_dt(r, dt, w_max_init);
      t += dt;
    } else {
      t = iters * dt;
    }
    /* SystemCheck (solver.c:2301-2326) */
    if (bad >= 0) {
      fprintf(stderr, "\n[SOVLER FAILURE] --- System has reached maximum Vorticity limt at Iter: [%ld]\n-->> Exiting!!!n",
              iters);
      ns2d_h5_close(f);
      return 1;
    }
    if (dt <= MIN_STEP_SIZE) {
      fprintf(stderr, "\n[SOVLER FAILURE]--- Timestep has become too small to continue at Iter: [%ld]\n-->> Exiting!!!\n",
              iters);
      ns2d_h5_close(f);
      return 1;
    }
    if (iters >= MAX_ITERS) {
      fprintf(stderr,
              "\n[SOVLER FAILURE]--- The maximum number of iterations has been reached at Iter: [%ld]\n-->> Exiting!!!\n",
              iters);
      ns2d_h5_close(f);
      return 1;
    }
  }

  /* FinalWriteAndCloseOutputFiles (hdf5_funcs.c:1370-1836) */
  const double cnt = (double)(r->num_sys_msr_counts > 0 ? r->num_sys_msr_counts : 1);
  for (long i = 0; i < r->Nx; ++i) s->mean_flow_y[i] /= cnt;
  for (long j = 0; j < r->Nyf; ++j) s->mean_flow_x[j] /= cnt;
  if (ns2d_h5_final(f, s, g_msg, sizeof g_msg)) die(g_msg);
  ns2d_h5_close(f);
  gpu_check(ns2d_destroy(r->h), "solver destruction");

  clock_gettime(CLOCK_MONOTONIC, &t_end);
  const double secs = (double)(t_end.tv_sec - t_begin.tv_sec) + 1e-9 * (double)(t_end.tv_nsec - t_begin.tv_nsec);
  const int hh = (int)secs / 3600, mm = ((int)secs - hh * 3600) / 60, ss = (int)secs - hh * 3600 - mm * 60;
  printf("\n\nTotal Execution Time: [%5.10lf] --> %d hrs : %d mins : %d secs\n\n", secs, hh, mm, ss);
  if (!c->file_only) write_details(r, &paths, argc, argv, secs);
  for (size_t q = 0; q < sizeof arrays / sizeof arrays[0]; ++q) free(*arrays[q]);
  free(s->mean_flow_y);
  free(s->mean_flow_x);
  free(r->w_hat);
  free(r->w);
  free(r->tg);
  free(r->enrg_spect);
  free(r->enst_spect);
  free(r->flux_spect);
  return 0;
}
