/*
 * ns2d_gpu.h -- C ABI of the MI355X (gfx950) pseudospectral 2-D Navier-Stokes
 * vorticity solver: the right-hand side and the RK4-CN / AB4 time integrators
 * of EndCar808/2D-Navier-Stokes, plus its system measures and max-vorticity
 * check, computed in FP64 on one GPU.
 *
 * Every entry point replaces a function of the reference solver
 * (Solver/src/...); the reference exchanges fftw_complex arrays in the layout
 * w_hat[Nx][Ny/2+1] (x slow, interleaved re/im, unnormalised FFTW units:
 * w_hat = r2c(w)).  The host buffers taken and returned here use exactly that
 * layout, so a caller can pass run_data->w_hat directly.
 *
 * Conventions: all functions return NS2D_OK (0) or an NS2D_ERR_* code and set a
 * thread-local message readable with ns2d_last_error().  The caller owns host
 * buffers; a handle owns its device memory and one HIP stream (default
 * priority).  A handle is not re-entrant: use one handle per host thread.
 */
#ifndef NS2D_GPU_H
#define NS2D_GPU_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NS2D_OK 0
#define NS2D_ERR_INVALID 1     /* bad argument / unsupported size            */
#define NS2D_ERR_HIP 2         /* HIP runtime failure                         */
#define NS2D_ERR_BLOWUP 3      /* max|w| >= 1e100 (reference SystemCheck)     */
#define NS2D_ERR_STATE 4       /* call not valid in the handle's state        */

/* dealiasing filter (reference ApplyDealiasing, Solver/src/solver.c:1039-1099) */
#define NS2D_FILTER_23 0         /* 2/3 isotropic cutoff (reference default)   */
#define NS2D_FILTER_HOULI_REF 1  /* Hou-Li with the reference's integer division */
#define NS2D_FILTER_HOULI 2      /* Hou-Li exp(-36 (|k|/kmax)^36)               */
#define NS2D_FILTER_NONE 3

/* forcing (reference force.c) */
#define NS2D_FORCING_NONE 0
#define NS2D_FORCING_KOLM 1      /* f_hat(0, k_f) = -0.5 F k_f, force.c:535-538 */

/* integrators (reference RK4Step / AB4Step) */
#define NS2D_INT_RK4CN 0         /* RK4 + Crank-Nicolson viscous term, solver.c:625-873 */
#define NS2D_INT_AB4 1           /* CN-AB4 with 3 RK4CN start-up steps, solver.c:261-350 */

/* reference quirks (bit set = reproduce the reference's behaviour) */
#define NS2D_QUIRK_CONJ_GLOBAL 0x01    /* ForceConjugacy writes the global w_hat, solver.c:2089-2101 */
#define NS2D_QUIRK_VISC_ASSIGN 0x02    /* HYPER_VISC_FLAG = HYPER_VISC assignment, solver.c:843   */
#define NS2D_QUIRK_KOLM_UNSCALED 0x04  /* forcing added without the N^2 scale, force.c:537       */
#define NS2D_QUIRK_KOLM_KX0_KILL 0x08  /* 2/3 filter zeroes kx=0 under KOLM, solver.c:1072-1078   */
#define NS2D_QUIRK_ENST_DISS 0x10      /* enst_diss drops pre_fac off the edge, sys_msr.c:225     */
#define NS2D_QUIRKS_REF 0x1f

typedef struct ns2d_params {
  int nx, ny;            /* -n Nx -n Ny (powers of two, 8..4096)               */
  double nu;             /* -v nu                                              */
  int hyper_flag;        /* -v (2nd): HYPER_VISC_FLAG                          */
  double hyper_pow;      /* -v (3rd): HYPER_VISC_POW                           */
  double alpha;          /* -d alpha: EKMN_ALPHA                               */
  int drag_flag;         /* -d (2nd): EKMN_DRAG_FLAG                           */
  double drag_pow;       /* -d (3rd): EKMN_DRAG_POW                            */
  int filter;            /* NS2D_FILTER_*                                      */
  int forcing;           /* NS2D_FORCING_*                                     */
  int force_k;           /* -f (2nd): forced wavenumber k_f                    */
  double force_scale;    /* -f (1st): forcing amplitude F                      */
  int integrator;        /* NS2D_INT_*                                         */
  unsigned int quirks;   /* NS2D_QUIRK_* bits                                  */
  int check_every;       /* 1: per-step max|w| blow-up check like SystemCheck  */
  int device;            /* HIP device ordinal                                 */
} ns2d_params;

/* Totals of ComputeSystemMeasurables (sys_msr.c:32-373), already normalised
 * as the reference normalises them (sys_msr.c:340-372). */
typedef struct ns2d_measures {
  double energy, enstrophy, palinstrophy, divergence;
  double enrg_diss, enst_diss, forcing;
  /* flux subsets (LWR_SBST_LIM <= |k| < UPR_SBST_LIM), filled when requested */
  double d_enrg_dt_sbst, enrg_flux_sbst, enrg_diss_sbst;
  double d_enst_dt_sbst, enst_flux_sbst, enst_diss_sbst;
} ns2d_measures_t;

typedef struct ns2d_info {
  long iters;              /* completed steps since the last ns2d_set_state       */
  long first_blowup_iter;  /* step after which max|w| >= 1e100 was seen, or -1    */
  double last_max_vort;    /* max|w| of the last checked state                    */
  int hyper_flag;          /* HYPER_VISC_FLAG as the reference would now hold it  */
  int n_spect;             /* spectrum length, sys_msr.c:394                      */
  int ld;                  /* device row stride (complex elements)                */
} ns2d_info_t;

typedef struct ns2d_handle ns2d_t;

/* Fill *p with the reference's CLI defaults (utils.c:51-85) for an nx x ny grid. */
int ns2d_params_default(ns2d_params* p, int nx, int ny);

/* Replaces AllocateMemory + InitializeFFTWPlans + InitializeSpaceVariables +
 * InitializeForcing (solver.c:2603-2901, 2022-2068, force.c:28-102). */
int ns2d_create(const ns2d_params* p, ns2d_t** out);
int ns2d_destroy(ns2d_t* h);

/* Load / read the state run_data->w_hat (host memory, [Nx][Ny/2+1] complex).
 * set_state resets the step counter, AB4 history and blow-up record. */
int ns2d_set_state(ns2d_t* h, const double* w_hat);
int ns2d_get_state(ns2d_t* h, double* w_hat);
/* Same, with a device pointer to a [Nx][Ny/2+1] complex array. */
int ns2d_set_state_device(ns2d_t* h, const void* d_w_hat);
int ns2d_get_state_device(ns2d_t* h, void* d_w_hat);

/* Advance nsteps steps of size dt with the configured integrator: replaces
 * the RK4Step / AB4Step calls of the SpectralSolve loop (solver.c:142-236)
 * together with the per-step SystemCheck (solver.c:2301-2326) when
 * check_every is set.  Returns NS2D_ERR_BLOWUP if max|w| >= 1e100 was seen
 * (ns2d_info_t.first_blowup_iter tells after which step). */
int ns2d_step(ns2d_t* h, double dt, long nsteps);

/* NonlinearRHSBatch (solver.c:884-1032): out = dealiased, forced nonlinear
 * term of `in`, both host [Nx][Ny/2+1] complex.  As in the reference, with
 * NS2D_QUIRK_CONJ_GLOBAL the handle's state gets w_hat[kx<0,0] = conj(in[|kx|,0]). */
int ns2d_rhs(ns2d_t* h, const double* in, double* out);

/* ComputeSystemMeasurables (sys_msr.c:32-373) on the current state.
 * enrg_spec / enst_spec: n_spect doubles each (or NULL).  flux_spec: NULL or
 * 6*n_spect doubles in the order d_enst_dt, enst_diss, enst_flux, d_enrg_dt,
 * enrg_diss, enrg_flux (hdf5_funcs.c:323-336).  with_flux != 0 evaluates the
 * extra nonlinear term the flux terms need (sys_msr.c:125-133). */
int ns2d_measures(ns2d_t* h, int with_flux, ns2d_measures_t* m, double* enrg_spec, double* enst_spec,
                  double* flux_spec);

/* ns2d_measures for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * No counterpart in the reference.  Same arguments and results (bit for bit)
 * as ns2d_measures; calls no hipMalloc/hipFree: the per-bin partial sums and
 * spectra live in a file-scope device array of the library (one per GPU,
 * 510 KB, sized for n_spect <= 2897, i.e. grids up to 4096 x 4096; larger
 * spectra are rejected with NS2D_ERR_INVALID).  Every call on a GPU shares
 * that GPU's array: calls that overlap on one GPU (other threads, other
 * handles) race. */
int ns2d_measures_noalloc(ns2d_t* h, int with_flux, ns2d_measures_t* m, double* enrg_spec, double* enst_spec,
                          double* flux_spec);

/* GetMaxData("VORT") (solver.c:2171-2295): max |c2r(w_hat)| / (Nx*Ny). */
int ns2d_max_vort(ns2d_t* h, double* out);

/* GetTimestep's convective rate for the CFL step (solver.c:2381-2441): the
 * maximum over the grid of pi (|u|/dx + |v|/dy) of the current state, u and v
 * from u_hat = i ky/|k|^2 w_hat, v_hat = -i kx/|k|^2 w_hat, c2r / (Nx*Ny),
 * dx = 2 pi/Nx, dy = 2 pi/Ny (every mode's own velocity: the reference's row
 * offset at solver.c:2394 is not reproduced).  The caller forms
 * dt = cfl / (conv + pi^2 nu (1/dx^2 + 1/dy^2)) as solver.c:2443-2447 does.
 * Reduced on the device: no host copy of the fields. */
int ns2d_cfl_rate(ns2d_t* h, double* conv);

/* The real-space vorticity written as "w" (hdf5_funcs.c:174,796):
 * w = c2r(w_hat)/(Nx*Ny), host [Nx][Ny] doubles. */
int ns2d_real_vort(ns2d_t* h, double* w);

/* Unnormalised forward transform of a real field, as the ICs use
 * (fftw_mpi_execute_dft_r2c, solver.c:1178): w [Nx][Ny] -> w_hat [Nx][Ny/2+1]. */
int ns2d_r2c(ns2d_t* h, const double* w, double* w_hat);

/* Unnormalised inverse transform of a host spectrum, as
 * fftw_mpi_execute_dft_c2r (solver.c:1407 DECAY_TURB_ALT, :2409 GetTimestep):
 * w_hat [Nx][Ny/2+1] -> w [Nx][Ny]; the imaginary parts of the ky = 0 and
 * ky = Ny/2 entries enter as FFTW's c2r lets them.  Leaves the state alone. */
int ns2d_c2r(ns2d_t* h, const double* w_hat, double* w);

int ns2d_get_info(ns2d_t* h, ns2d_info_t* info);
int ns2d_sync(ns2d_t* h);
/* The handle's hipStream_t, for callers that order their own work after it. */
int ns2d_get_stream(ns2d_t* h, void** stream);

/* Run nsteps steps with HIP events around every kernel launch and report the
 * mean duration (ms) of the column pass (kpass) and the row pass (rpass) and
 * the total elapsed time (ms).  The state advances as with ns2d_step. */
int ns2d_profile_steps(ns2d_t* h, double dt, long nsteps, double* kpass_ms, double* rpass_ms,
                       double* total_ms);

/* Tuning.  Each pass has several compiled shapes per transform length:
 *   column pass (Nx): {columns per workgroup, rows per thread, threads,
 *                      workgroups sharing a column group}
 *   row pass (Ny):    {threads per row pair, points per thread, threads,
 *                      row pairs per workgroup}
 * The project's autotuner times them and picks one per grid; a new handle
 * starts from ns2d_default_variant's shape (index 0 for both integrators;
 * the integrator argument lets a table name different ones).  No counterpart in the reference (FFTW_MEASURE
 * planning, solver.c:2883-2895, plays this role there). */
#define NS2D_PASS_COLUMN 0
#define NS2D_PASS_ROW 1
int ns2d_variant_count(int pass, int n);
int ns2d_variant_info(int pass, int n, int idx, int* shape4);
/* Further compile-time choices of a shape: extra4 = {minimum waves per SIMD
 * its register budget allows (1: no constraint), LDS padding shift (one slot
 * per 2^shift elements), 16-B slots between neighbouring lines, flags}; flag
 * bit 0: the column pass carries the ky = 0 and ky = Ny/2 columns as one
 * packed complex line (Ny/2 column slots instead of Ny/2+1); bit 1: the
 * column pass issues its RK epilogue loads before the forward transform;
 * bit 2: the packed pair's second column stays in registers; bit 3: the row
 * pass runs one row per line group instead of a row pair; bit 4: that row's
 * real product is transformed forward as a complex line of half the length;
 * bit 5: the row transforms multiply by twiddle powers preloaded from the
 * table instead of forming them by products (row pass only; bit 1 then means
 * its G fields are loaded up front). */
int ns2d_variant_extra(int pass, int n, int idx, int* extra4);
/* The shape a new handle with this integrator (NS2D_INT_*) runs for a
 * transform length n unless NS2D_KPASS_VARIANT_<Nx> / NS2D_RPASS_VARIANT_<Ny>
 * name another.  Needs no GPU. */
int ns2d_default_variant(int pass, int n, int integrator, int* idx);
int ns2d_set_variant(ns2d_t* h, int pass, int idx);
/* Device storage of the spectral arrays this library was compiled with:
 * *xtile_k for the k-space arrays (state, stage fields, mode tables), 0 =
 * column-major [Ny/2+1][Nx], else rows of x-tiles of that many x; *xtile_g
 * for the mixed-space fields the column pass writes for the row pass, *xtile_n
 * for those the row pass writes for the column pass (x-tile rows).  The host
 * side always sees [Nx][Ny/2+1].  Needs no GPU.  No counterpart in the
 * reference (FFTW's layout is fixed by its plans). */
int ns2d_layout(int* xtile_k, int* xtile_g, int* xtile_n);

/* Ensemble (no counterpart in the reference, which runs one realisation per
 * process): n independent RK4CN solvers of possibly different grid sizes,
 * stepped together.  Members of one size share batched kernel launches; a
 * two-step period of all members is replayed as one hipGraph.  Member
 * handles (ns2d_ens_member) accept every single-solver call (set_state,
 * get_state, measures, ...) and must not be destroyed individually. */
typedef struct ns2d_ensemble ns2d_ens_t;
int ns2d_ens_create(const ns2d_params* members, int n, ns2d_ens_t** out);
int ns2d_ens_member(ns2d_ens_t* e, int i, ns2d_t** h);
/* nsteps RK4CN steps of every member, member i with step dt[i]. */
int ns2d_ens_step(ns2d_ens_t* e, const double* dt, long nsteps);
/* The batched passes' autotuner: times every compiled row- and column-pass
 * shape on all members of each size class (best of reps) and keeps the
 * fastest; member states are preserved.  ns2d_ens_step runs it once on its
 * first batched step unless NS2D_ENS_TUNE=0.  ns2d_ens_variants reports the
 * shapes (ns2d_variant_info indices) a member's class uses. */
int ns2d_ens_autotune(ns2d_ens_t* e, int reps);
int ns2d_ens_variants(ns2d_ens_t* e, int member, int* kpass_idx, int* rpass_idx);
int ns2d_ens_sync(ns2d_ens_t* e);
int ns2d_ens_destroy(ns2d_ens_t* e);

/* Discrete adjoint (no counterpart in the reference): the vector-Jacobian
 * product of nsteps RK4CN steps w_n = Phi^n(w_0).  Given lam_T = dJ/dw_n it
 * returns grad_w0 = dJ/dw_0, both packed as dJ/dRe + i dJ/dIm in the
 * [Nx][Ny/2+1] layout.  The forward map is re-run from w_0 with checkpoints
 * every ckpt_every steps (<= 0: one segment, i.e. every stage of every step
 * kept on a device tape of 64*M*(4*nsteps+1) bytes); each segment is replayed
 * once with its stages taped, then swept backwards.  Requires the
 * ForceConjugacy quirk off.  Leaves the handle's state undefined. */
int ns2d_vjp(ns2d_t* h, const double* w0, double dt, long nsteps, const double* lam_T, double* grad_w0,
             long ckpt_every);
int ns2d_vjp_device(ns2d_t* h, const void* d_w0, double dt, long nsteps, const void* d_lam_T, void* d_grad_w0,
                    long ckpt_every);

const char* ns2d_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* NS2D_GPU_H */
