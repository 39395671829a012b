/* Host side of the ns2d solver driver: command line, initial conditions and
 * HDF5 output.  Plain C99, no HIP: the GPU is reached only through
 * include/ns2d_gpu.h.  The pieces mirror the reference driver
 * (Solver/src/main.c, utils.c, solver.c InitialConditions, hdf5_funcs.c)
 * for a single process holding the whole grid (local_Nx = Nx). */
#ifndef NS2D_HOST_H
#define NS2D_HOST_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Everything the command line sets (GetCMLArgs, utils.c:34-461), with the
 * reference's defaults (utils.c:48-88), plus the choices the reference makes
 * at compile time (data_types.h:49-60, Solver/Makefile DSET_FLAGS), which this
 * driver takes as extra flags: -I integrator, -D filter, -Q quirk mask,
 * -X (the TESTING build's Taylor-Green solution and error norms), and the
 * process count of `mpirun -n P` (run_solver.py:270), on which the reference's
 * random initial conditions depend: -R P. */
typedef struct ns2d_run_cfg {
  char input_file_name[512];
  char input_dir[512];
  char output_dir[512];
  char output_tag[64];
  int file_only;
  long N[2];
  double t0, dt, T;
  double cfl_const;
  int adapt_step_flag;
  int cfl_cond_flag;
  char u0[64];
  char forcing[64];
  int force_k;
  double force_scale_var;
  int trans_iters_flag;
  double trans_iters_frac;
  double nu;
  int hyper_visc_flag;
  double hyper_visc_pow;
  double ekmn_alpha;
  int ekmn_drag_flag;
  double ekmn_drag_pow;
  int save_every;
  /* compile-time choices of the reference */
  char integrator[16]; /* "RK4CN" (data_types.h:54 default) or "AB4"         */
  char filter[16];     /* "23" (__DEALIAS_23, default) or "HOU_LI"            */
  unsigned quirks;     /* NS2D_QUIRK_* bits, default NS2D_QUIRKS_REF            */
  int testing;         /* -X: TGSoln dataset and L2/Linf print (TESTING build) */
  int ic_ranks;        /* -R P: the initial condition of a P-rank run, default 1 */
} ns2d_run_cfg;

/* the largest rank count -R accepts */
#define NS2D_IC_MAX_RANKS 65536

void ns2d_cli_defaults(ns2d_run_cfg* c);

/* getopt parse with the reference's flags and repeated-flag semantics
 * (-n Nx -n Ny; -h dt -h adapt; -v nu -v hyper_flag -v hyper_pow;
 * -d alpha -d drag_flag -d drag_pow; -c cfl_flag -c cfl; -T flag -T frac;
 * -f type -f k -f scale; -o dir [-o x: file only]).  Returns 0, or 1 with
 * the reference's error text in msg. */
int ns2d_cli_parse(int argc, char** argv, ns2d_run_cfg* c, char* msg, size_t msglen);

/* Output naming of GetOutputDirPath (hdf5_funcs.c:440-640): fills the main,
 * spectra and system-measures file names and, unless file_only, appends the
 * SIM_DATA_... folder to output_dir.  mkdirs != 0 creates the folders. */
typedef struct ns2d_out_paths {
  char dir[1200];
  char main_file[1240];
  char spectra_file[1240];
  char sys_msr_file[1240];
} ns2d_out_paths;
int ns2d_output_paths(const ns2d_run_cfg* c, int mkdirs, ns2d_out_paths* out, char* msg, size_t msglen);

/* Unnormalised 2D transforms the initial conditions need, supplied by the
 * caller (the GPU library in the driver, numpy in the host tests):
 * r2c: real [Nx][Ny] -> complex [Nx][Ny/2+1]; c2r: the inverse, both
 * with FFTW's conventions. */
typedef struct ns2d_transforms {
  void* ctx;
  int (*r2c)(void* ctx, const double* in, double* out);
  int (*c2r)(void* ctx, const double* in, double* out);
} ns2d_transforms;

/* InitialConditions (solver.c:1107-2014) for a single process: fills w_hat
 * ([Nx][Ny/2+1] interleaved complex, FFTW units) for c->u0, draws its random
 * numbers from glibc rand() after srand(123456789) in the reference's (i, j)
 * order, then applies ApplyDealiasing and ForceConjugacy as the reference
 * does (solver.c:1996-2000).  Returns 0 or 1 with msg. */
int ns2d_initial_condition(const ns2d_run_cfg* c, double* w_hat, const ns2d_transforms* tf, char* msg,
                           size_t msglen);

/* ApplyDealiasing (solver.c:1039-1099) on dim interleaved complex fields. */
void ns2d_host_dealias(const ns2d_run_cfg* c, double* f_hat, int dim);
/* ForceConjugacy(w_hat, N) (solver.c:2074-2107) with the array as argument. */
void ns2d_host_force_conjugacy(const ns2d_run_cfg* c, double* w_hat);

/* Mean flows accumulated by ComputeSystemMeasurables (sys_msr.c:196-230,
 * 339-347) for one state; the caller sums them over measurement calls. */
void ns2d_mean_flows(const ns2d_run_cfg* c, const double* w_hat, double* mean_flow_y /* Nx */,
                     double* mean_flow_x /* Ny/2+1 */);

/* ---------------------------------------------------------------- HDF5 */
typedef struct ns2d_h5 ns2d_h5;

/* One save: the state, its real-space vorticity, spectra and (testing) the
 * Taylor-Green solution.  Any pointer may be NULL to skip that dataset. */
typedef struct ns2d_save {
  double t, dt;
  int index;               /* the group is "/Iter_%05d" of this index            */
  const double* w;         /* [Nx][Ny]                                          */
  const double* w_hat;     /* [Nx][Ny/2+1] complex                              */
  const double* tg_soln;   /* [Nx][Ny]                                          */
  const double* enst_spect, *enrg_spect; /* n_spect each                        */
  const double* flux_spect;              /* 6 * n_spect, ns2d_measures order    */
} ns2d_save;

/* Per-save series written at the end (FinalWriteAndCloseOutputFiles,
 * hdf5_funcs.c:1370-1836), num_print_steps entries each. */
typedef struct ns2d_series {
  long n;
  double *time, *tot_energy, *tot_enstr, *tot_palin, *tot_forc, *tot_div, *enrg_diss, *enst_diss;
  double *d_enst_dt_sbst, *enst_flux_sbst, *enst_diss_sbst, *d_enrg_dt_sbst, *enrg_flux_sbst, *enrg_diss_sbst;
  double* mean_flow_y; /* Nx, already divided by the number of measurements  */
  double* mean_flow_x; /* Ny/2+1, likewise                                   */
} ns2d_series;

ns2d_h5* ns2d_h5_create(const ns2d_run_cfg* c, const ns2d_out_paths* p, int n_spect, char* msg, size_t msglen);
int ns2d_h5_write_save(ns2d_h5* f, const ns2d_save* s, char* msg, size_t msglen);
int ns2d_h5_final(ns2d_h5* f, const ns2d_series* s, char* msg, size_t msglen);
void ns2d_h5_close(ns2d_h5* f);

/* MAX_PALIN ingest: "/psi" ([Nx][Ny] doubles) of an HDF5 file. */
int ns2d_h5_read_psi(const char* path, double* psi, long n, char* msg, size_t msglen);

#ifdef __cplusplus
}
#endif
#endif
