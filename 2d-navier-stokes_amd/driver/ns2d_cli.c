/* Command line and output naming of the ns2d driver.
 *
 * Flags, defaults, repeated-flag semantics and error texts follow GetCMLArgs
 * (Solver/src/utils.c:34-461) so that the reference's run scripts
 * (run_solver.py:232-245) drive this binary unchanged; output folder and file
 * names follow GetOutputDirPath (Solver/src/hdf5_funcs.c:440-640). */
#define _GNU_SOURCE
#include <errno.h>
#include <getopt.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>
#include <unistd.h>

#include "ns2d_gpu.h"
#include "ns2d_host.h"

static int say(char* msg, size_t n, const char* fmt, ...) {
  if (msg && n) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(msg, n, fmt, ap);
    va_end(ap);
  }
  return 1;
}

static void copy_str(char* dst, size_t cap, const char* src) {
  strncpy(dst, src, cap - 1);
  dst[cap - 1] = '\0';
}

void ns2d_cli_defaults(ns2d_run_cfg* c) {
  memset(c, 0, sizeof *c);
  copy_str(c->input_file_name, sizeof c->input_file_name, "NONE");
  copy_str(c->input_dir, sizeof c->input_dir, "./Data/Solver/InitialConditions/");
  copy_str(c->output_dir, sizeof c->output_dir, "./Data/Tmp/");
  copy_str(c->output_tag, sizeof c->output_tag, "No-Tag");
  c->file_only = 0;
  c->N[0] = 64;
  c->N[1] = 64;
  c->t0 = 0.0;
  c->dt = 1e-4;
  c->T = 1.0;
  c->cfl_const = 0.9;
  c->adapt_step_flag = 0;
  c->cfl_cond_flag = 1;
  copy_str(c->u0, sizeof c->u0, "TG_VORT");
  copy_str(c->forcing, sizeof c->forcing, "NONE");
  c->force_k = 0;
  c->force_scale_var = 1.0;
  c->trans_iters_flag = 0;
  c->trans_iters_frac = 0.2;  /* TRANSIENT_FRAC, data_types.h:72 */
  c->nu = 0.0001;
  c->hyper_visc_flag = 0;
  c->hyper_visc_pow = 2.0;    /* VISC_POW, data_types.h:67 */
  c->ekmn_alpha = 0.0;
  c->ekmn_drag_flag = 0;
  c->ekmn_drag_pow = -2.0;    /* EKMN_POW, data_types.h:69 */
  c->save_every = 100;
  copy_str(c->integrator, sizeof c->integrator, "RK4CN");
  copy_str(c->filter, sizeof c->filter, "23");
  c->quirks = NS2D_QUIRKS_REF;
  c->testing = 0;
  c->ic_ranks = 1;
}

/* initial conditions the reference accepts for -i (utils.c:262-345); any
 * other name selects the "NONE" random field */
static const char* const k_ics[] = {"TG_VEL",         "TG_VORT",          "DOUBLE_SHEAR_LAYER", "TESTING",
                                    "DECAY_TURB_BB",  "DECAY_TURB_NB",    "DECAY_TURB_ALT",     "DECAY_TURB_EXP",
                                    "GAUSS_DECAY_TURB", "EXTRM_ENS",      "MAX_PALIN",          "RING",
                                    "UNIF",           "RANDOM",           "GAUSS_BLOB",         "ZERO",
                                    "ZERO_KOLM"};
static const char* const k_forcings[] = {"ZERO", "CONST_GAUSS", "KOLM", "STOC", "STOC_GAUSS", "NONE"};

#define PFX "\n[ERROR] Parsing of Command Line Arguements Failed: "

int ns2d_cli_parse(int argc, char** argv, ns2d_run_cfg* c, char* msg, size_t n) {
  int dim_i = 0, force_i = 0, cfl_i = 0, visc_i = 0, drag_i = 0, out_i = 0, trans_i = 0, dt_i = 0;
  int opt;
  optind = 1;
  opterr = 0;
  while ((opt = getopt(argc, argv, "o:h:n:d:s:e:t:v:i:c:p:f:z:T:I:D:Q:R:X")) != -1) {
    switch (opt) {
      case 'o':
        if (out_i == 0) copy_str(c->output_dir, sizeof c->output_dir, optarg);
        else c->file_only = 1;  /* a second -o selects file-only output */
        ++out_i;
        break;
      case 'n': {
        if (dim_i > 1) return say(msg, n, PFX "more than two -n dimensions given\n");
        long v = atol(optarg);
        if (v <= 2)
          return say(msg, n, PFX "The minimum dimension size of [%ld] must be greater than 2\n-->> Exiting!\n", v);
        if (v % 2 != 0)
          return say(msg, n, PFX "The system dimension of [%ld] must be a multiple of 2\n-->> Exiting!\n", v);
        c->N[dim_i++] = v;
        break;
      }
      case 's':
        c->t0 = atof(optarg);
        if (c->t0 < 0)
          return say(msg, n, PFX "The integration start time [%lf] must be a positive\n-->> Exiting!\n", c->t0);
        break;
      case 'e':
        c->T = atof(optarg);
        if (c->T < 0)
          return say(msg, n, PFX "The integration end time [%lf] must be a positive\n-->> Exiting!\n", c->T);
        if (c->T < c->t0)
          return say(msg, n,
                     PFX "The provided end time: [%lf] must be greater than the initial time: [%lf]\n-->> Exiting!\n",
                     c->T, c->t0);
        break;
      case 'h':
        if (dt_i == 0) {
          c->dt = atof(optarg);
          if (c->dt <= 0)
            return say(msg, n, PFX "The provided timestep: [%lf] must be strictly positive\n-->> Exiting!\n", c->dt);
          dt_i = 1;
        } else {
          c->adapt_step_flag = atoi(optarg);
          if (c->adapt_step_flag != 0 && c->adapt_step_flag != 1)
            return say(msg, n, PFX "Incorrect Adaptive stepping flag: [%d] Must be either 0 or 1\n-->> Exiting!\n",
                       c->adapt_step_flag);
        }
        break;
      case 'c':
        if (cfl_i == 0) {
          c->cfl_cond_flag = atoi(optarg);
          if (c->cfl_cond_flag != 0 && c->cfl_cond_flag != 1)
            return say(msg, n, PFX "Incorrect CFL condition flag: [%d] Must be either 0 or 1\n-->> Exiting!\n",
                       c->cfl_cond_flag);
          cfl_i = 1;
        } else {
          c->cfl_const = atof(optarg);
          if (c->cfl_const <= 0)
            return say(msg, n, PFX "The provided CFL Constant: [%lf] must be strictly positive\n-->> Exiting!\n",
                       c->cfl_const);
        }
        break;
      case 'v':
        if (visc_i == 0) {
          c->nu = atof(optarg);
          if (c->nu < 0)
            return say(msg, n, PFX "The provided viscosity: [%lf] must be positive\n-->> Exiting!\n", c->nu);
          visc_i = 1;
        } else if (visc_i == 1) {
          c->hyper_visc_flag = atoi(optarg);
          if (c->hyper_visc_flag != 0 && c->hyper_visc_flag != 1)
            return say(msg, n, PFX "Incorrect Hyperviscosity flag: [%d] Must be either 0 or 1\n-->> Exiting!\n",
                       c->hyper_visc_flag);
          visc_i = 2;
        } else {
          c->hyper_visc_pow = atof(optarg);
          if (c->hyper_visc_pow <= 0.0)
            return say(msg, n, PFX "The provided hyperviscosity power: [%lf] must be strictly positive\n-->> Exiting!\n",
                       c->hyper_visc_pow);
        }
        break;
      case 'd':
        if (drag_i == 0) {
          c->ekmn_alpha = atof(optarg);
          if (c->ekmn_alpha < 0)
            return say(msg, n, PFX "The provided Ekman Drag: [%lf] must be positive\n-->> Exiting!\n", c->ekmn_alpha);
          drag_i = 1;
        } else if (drag_i == 1) {
          c->ekmn_drag_flag = atoi(optarg);
          if (c->ekmn_drag_flag != 0 && c->ekmn_drag_flag != 1)
            return say(msg, n, PFX "Incorrect Ekman Drag flag: [%d] Must be either 0 or 1\n-->> Exiting!\n",
                       c->ekmn_drag_flag);
          drag_i = 2;
        } else {
          c->ekmn_drag_pow = atof(optarg);
          if (c->ekmn_drag_pow > 0.0)
            return say(msg, n, PFX "The provided Hypodiffusivity power: [%lf] must be negative or zero\n-->> Exiting!\n",
                       c->ekmn_drag_pow);
        }
        break;
      case 'i': {
        const char* sel = "NONE";
        for (size_t k = 0; k < sizeof k_ics / sizeof k_ics[0]; ++k)
          if (!strcmp(optarg, k_ics[k])) sel = k_ics[k];
        copy_str(c->u0, sizeof c->u0, sel);
        break;
      }
      case 't':
        copy_str(c->output_tag, sizeof c->output_tag, optarg);
        break;
      case 'T':
        if (trans_i == 0) {
          c->trans_iters_flag = atoi(optarg);
          if (c->trans_iters_flag != 0 && c->trans_iters_flag != 1)
            return say(msg, n, PFX "Incorrect transient iterations flag: [%d] Must be either 0 or 1\n-->> Exiting!\n",
                       c->trans_iters_flag);
          trans_i = 1;
        } else {
          c->trans_iters_frac = atof(optarg);
          if (c->trans_iters_frac >= 1.0 || c->trans_iters_frac <= 0.0)
            return say(msg, n, PFX "Incorrect transient iterations fraction: [%lf] Must be between 0 and 1\n-->> Exiting!\n",
                       c->trans_iters_frac);
        }
        break;
      case 'p':
        c->save_every = atoi(optarg);
        break;
      case 'f':
        /* first -f: forcing type (unknown names are ignored, as in the
         * reference); second: forced wavenumber; third and later: scale */
        if (force_i == 0) {
          for (size_t k = 0; k < sizeof k_forcings / sizeof k_forcings[0]; ++k)
            if (!strcmp(optarg, k_forcings[k])) {
              copy_str(c->forcing, sizeof c->forcing, k_forcings[k]);
              force_i = 1;
            }
        } else if (force_i == 1) {
          c->force_k = atoi(optarg);
          force_i = 2;
        } else {
          c->force_scale_var = atof(optarg);
        }
        break;
      case 'z':
        copy_str(c->input_file_name, sizeof c->input_file_name, optarg);
        if (access(c->input_file_name, F_OK) != 0)
          return say(msg, n,
                     PFX "The input file [%s] cannot be found, please ensure correct path to file is specified.\n",
                     c->input_file_name);
        break;
      case 'I':
        if (strcmp(optarg, "RK4CN") && strcmp(optarg, "AB4"))
          return say(msg, n, PFX "Unknown integrator [%s]: use RK4CN or AB4\n", optarg);
        copy_str(c->integrator, sizeof c->integrator, optarg);
        break;
      case 'D':
        if (strcmp(optarg, "23") && strcmp(optarg, "HOU_LI"))
          return say(msg, n, PFX "Unknown dealiasing filter [%s]: use 23 or HOU_LI\n", optarg);
        copy_str(c->filter, sizeof c->filter, optarg);
        break;
      case 'Q': {
        char* end = NULL;
        unsigned long q = strtoul(optarg, &end, 0);
        if (!end || *end || q > NS2D_QUIRKS_REF) return say(msg, n, PFX "Bad quirk mask [%s]\n", optarg);
        c->quirks = (unsigned)q;
        break;
      }
      case 'R': {
        char* end = NULL;
        long p = strtol(optarg, &end, 10);
        if (!end || *end || p < 1 || p > NS2D_IC_MAX_RANKS)
          return say(msg, n, PFX "The rank count [%s] must be an integer in 1..%d\n", optarg, NS2D_IC_MAX_RANKS);
        c->ic_ranks = (int)p;
        break;
      }
      case 'X':
        c->testing = 1;
        break;
      default:
        return say(msg, n,
                   "\n[ERROR] Incorrect command line flag encountered\n"
                   "Use -o to specify the output directory\n"
                   "Use -n to specify the size of each dimension in the system\n"
                   "Use -s to specify the start time of the simulation\n"
                   "Use -e to specify the end time of the simulation\n"
                   "Use -h to specify the timestep\n"
                   "Use -c to specify the CFL constant for the adaptive stepping\n"
                   "Use -v to specify the system viscosity\n"
                   "Use -i to specify the initial condition\n"
                   "Use -t to specify the tag name to be used in the output file directory\n"
                   "Use -f to specify the forcing type\n"
                   "Use -T to specify if transient iterations are to be performed\n"
                   "Use -p to specify how often to print to file\n"
                   "Use -z to specify an input file to read parameters from\n"
                   "Use -I / -D / -Q to select the integrator (RK4CN, AB4), the dealiasing filter (23, HOU_LI)\n"
                   "and the reference-quirk mask; -X writes the Taylor-Green solution and error norms;\n"
                   "-R P builds the initial condition of a run on P MPI ranks\n"
                   "-->> Now Exiting!\n");
    }
  }
  return 0;
}

static int mkdir_if_missing(const char* d, char* msg, size_t n, const char* what) {
  struct stat st;
  if (stat(d, &st) == -1 && mkdir(d, 0700) == -1 && errno != EEXIST)
    return say(msg, n, "\n[ERROR] --- Unable to create %s [%s]\n-->> Exiting...\n", what, d);
  return 0;
}

int ns2d_output_paths(const ns2d_run_cfg* c, int mkdirs, ns2d_out_paths* out, char* msg, size_t n) {
  memset(out, 0, sizeof *out);
  if (mkdirs && mkdir_if_missing(c->output_dir, msg, n, "provided output directory")) return 1;
  char params[640];
  int pn;
  if (c->file_only) {
    /* hdf5_funcs.c:463-535: NAV/EUL, short model tags, files beside each other */
    pn = snprintf(params, sizeof params,
             "_SIM[%s-%s-%s]_N[%ld,%ld]_T[%1.1lf,%g,%1.3lf]_NU[%g,%d,%1.1lf]_DRAG[%g,%d,%1.1lf]_CFL[%1.2lf]_FORC[%s,%d,%g]_"
             "u0[%s].h5",
             "NAV", c->integrator, "FULL", c->N[0], c->N[1], c->t0, c->dt, c->T, c->nu, c->hyper_visc_flag,
             c->hyper_visc_pow, c->ekmn_alpha, c->ekmn_drag_flag, c->ekmn_drag_pow, c->cfl_const, c->forcing, c->force_k,
             c->force_scale_var, c->u0);
    if (pn < 0 || pn >= (int)sizeof params) return say(msg, n, "\n[ERROR] --- output file name too long\n");
    snprintf(out->dir, sizeof out->dir, "%s", c->output_dir);
    snprintf(out->main_file, sizeof out->main_file, "%sMain_HDF_Data%s", c->output_dir, params);
    snprintf(out->spectra_file, sizeof out->spectra_file, "%sSpectra_HDF_Data%s", c->output_dir, params);
    /* the reference leaves the measures file unnamed in this mode; they go
     * into the main file, where Plotting/functions.py:499-579 also looks */
    snprintf(out->sys_msr_file, sizeof out->sys_msr_file, "%s", out->main_file);
  } else {
    pn = snprintf(params, sizeof params,
             "SIM_DATA_%s_%s_%s_N[%ld,%ld]_T[%1.1lf,%g,%1.3lf]_NU[%g,%d,%1.1lf]_DRAG[%g,%d,%1.1lf]_CFL[%1.2lf]_FORC[%s,%d,%"
             "g]_u0[%s]_TAG[%s]/",
             "NAVIER", c->integrator, "FULL", c->N[0], c->N[1], c->t0, c->dt, c->T, c->nu, c->hyper_visc_flag,
             c->hyper_visc_pow, c->ekmn_alpha, c->ekmn_drag_flag, c->ekmn_drag_pow, c->cfl_const, c->forcing, c->force_k,
             c->force_scale_var, c->u0, c->output_tag);
    if (pn < 0 || pn >= (int)sizeof params) return say(msg, n, "\n[ERROR] --- output folder name too long\n");
    snprintf(out->dir, sizeof out->dir, "%s%s", c->output_dir, params);
    if (mkdirs && mkdir_if_missing(out->dir, msg, n, "folder for output files")) return 1;
    snprintf(out->main_file, sizeof out->main_file, "%sMain_HDF_Data.h5", out->dir);
    snprintf(out->spectra_file, sizeof out->spectra_file, "%sSpectra_HDF_Data.h5", out->dir);
    snprintf(out->sys_msr_file, sizeof out->sys_msr_file, "%sSystemMeasures_HDF_Data.h5", out->dir);
  }
  return 0;
}
