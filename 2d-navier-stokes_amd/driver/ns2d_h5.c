/* HDF5 output of the ns2d driver, serial HDF5 (one process holds the grid).
 *
 * Layout as written by the reference (Solver/src/hdf5_funcs.c):
 *   Main_HDF_Data.h5     /Iter_%05d groups with attributes TimeValue and
 *                        TimeStep (double[1], :1058-1115) holding "w"
 *                        (double [Nx][Ny], :764-797), "w_hat" ([Nx][Ny/2+1]
 *                        compound {r, i} of doubles, :799-811, :1837-1868)
 *                        and, in testing runs, "TGSoln" (:930-941);
 *   Spectra_HDF_Data.h5  the same groups holding the n_spect-long spectra
 *                        (:944-980);
 *   SystemMeasures_HDF_Data.h5  kx, ky (int), x, y, Time, MeanFlow_y,
 *                        MeanFlow_x and the per-save series (:1370-1680).
 * The reference opens and closes the files at every save; here they stay
 * open and are flushed after each save, which leaves the same files. */
#include <hdf5.h>
#include <hdf5_hl.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "ns2d_host.h"

struct ns2d_h5 {
  hid_t main_f, spec_f, cplx;
  long Nx, Ny;
  int n_spect;
  char sys_msr_file[1240];
  int sys_msr_in_main;
};

static int say(char* msg, size_t n, const char* fmt, ...) {
  if (msg && n) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(msg, n, fmt, ap);
    va_end(ap);
  }
  return 1;
}

/* CreateComplexDatatype, hdf5_funcs.c:1837-1868: compound {"r", "i"} */
static hid_t complex_type(void) {
  hid_t t = H5Tcreate(H5T_COMPOUND, 2 * sizeof(double));
  if (t < 0) return t;
  if (H5Tinsert(t, "r", 0, H5T_NATIVE_DOUBLE) < 0 || H5Tinsert(t, "i", sizeof(double), H5T_NATIVE_DOUBLE) < 0) {
    H5Tclose(t);
    return -1;
  }
  return t;
}

ns2d_h5* ns2d_h5_create(const ns2d_run_cfg* c, const ns2d_out_paths* p, int n_spect, char* msg, size_t n) {
  ns2d_h5* f = (ns2d_h5*)calloc(1, sizeof *f);
  if (!f) {
    say(msg, n, "\n[ERROR] --- out of memory\n");
    return NULL;
  }
  f->Nx = c->N[0];
  f->Ny = c->N[1];
  f->n_spect = n_spect;
  f->main_f = f->spec_f = f->cplx = -1;
  snprintf(f->sys_msr_file, sizeof f->sys_msr_file, "%s", p->sys_msr_file);
  f->sys_msr_in_main = !strcmp(p->sys_msr_file, p->main_file);
  f->cplx = complex_type();
  f->main_f = H5Fcreate(p->main_file, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
  if (f->main_f < 0) {
    say(msg, n, "\n[ERROR] --- Unable to create output file [%s]\n-->> Exiting...\n", p->main_file);
    ns2d_h5_close(f);
    return NULL;
  }
  f->spec_f = H5Fcreate(p->spectra_file, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
  if (f->spec_f < 0 || f->cplx < 0) {
    say(msg, n, "\n[ERROR] --- Unable to create spectra file [%s]\n-->> Exiting...\n", p->spectra_file);
    ns2d_h5_close(f);
    return NULL;
  }
  return f;
}

/* CreateGroup, hdf5_funcs.c:1058-1115 */
static hid_t iter_group(hid_t file, int index, double t, double dt) {
  char name[64];
  snprintf(name, sizeof name, "/Iter_%05d", index);
  if (H5Lexists(file, name, H5P_DEFAULT) > 0) return H5Gopen2(file, name, H5P_DEFAULT);
  hid_t g = H5Gcreate2(file, name, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
  if (g < 0) return g;
  hsize_t one = 1;
  hid_t sp = H5Screate_simple(1, &one, NULL);
  hid_t a = H5Acreate2(g, "TimeValue", H5T_NATIVE_DOUBLE, sp, H5P_DEFAULT, H5P_DEFAULT);
  herr_t e1 = H5Awrite(a, H5T_NATIVE_DOUBLE, &t);
  H5Aclose(a);
  a = H5Acreate2(g, "TimeStep", H5T_NATIVE_DOUBLE, sp, H5P_DEFAULT, H5P_DEFAULT);
  herr_t e2 = H5Awrite(a, H5T_NATIVE_DOUBLE, &dt);
  H5Aclose(a);
  H5Sclose(sp);
  if (e1 < 0 || e2 < 0) {
    H5Gclose(g);
    return -1;
  }
  return g;
}

static herr_t write_ds(hid_t loc, const char* name, hid_t type, int rank, const hsize_t* dims, const void* data) {
  hid_t sp = H5Screate_simple(rank, dims, NULL);
  if (sp < 0) return -1;
  hid_t d = H5Dcreate2(loc, name, type, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
  herr_t e = d < 0 ? -1 : H5Dwrite(d, type, H5S_ALL, H5S_ALL, H5P_DEFAULT, data);
  if (d >= 0) H5Dclose(d);
  H5Sclose(sp);
  return e;
}

int ns2d_h5_write_save(ns2d_h5* f, const ns2d_save* s, char* msg, size_t n) {
  hid_t g = iter_group(f->main_f, s->index, s->t, s->dt);
  if (g < 0) return say(msg, n, "\n[ERROR] --- Unable to create group Iter_%05d in the main file\n", s->index);
  const hsize_t dr[2] = {(hsize_t)f->Nx, (hsize_t)f->Ny};
  const hsize_t df[2] = {(hsize_t)f->Nx, (hsize_t)(f->Ny / 2 + 1)};
  herr_t e = 0;
  if (s->w && e >= 0) e = write_ds(g, "w", H5T_NATIVE_DOUBLE, 2, dr, s->w);
  if (s->w_hat && e >= 0) e = write_ds(g, "w_hat", f->cplx, 2, df, s->w_hat);
  if (s->tg_soln && e >= 0) e = write_ds(g, "TGSoln", H5T_NATIVE_DOUBLE, 2, dr, s->tg_soln);
  H5Gclose(g);
  if (e < 0) return say(msg, n, "\n[ERROR] --- Unable to write data to the main file at Iter = [%d]\n", s->index);
  if (s->enst_spect || s->enrg_spect || s->flux_spect) {
    g = iter_group(f->spec_f, s->index, s->t, s->dt);
    if (g < 0) return say(msg, n, "\n[ERROR] --- Unable to create group Iter_%05d in the spectra file\n", s->index);
    const hsize_t ds = (hsize_t)f->n_spect;
    if (s->enst_spect && e >= 0) e = write_ds(g, "EnstrophySpectrum", H5T_NATIVE_DOUBLE, 1, &ds, s->enst_spect);
    if (s->enrg_spect && e >= 0) e = write_ds(g, "EnergySpectrum", H5T_NATIVE_DOUBLE, 1, &ds, s->enrg_spect);
    if (s->flux_spect) {
      static const char* const names[6] = {"TimeDerivativeEnstrophySpectrum", "EnstrophyDissipationSpectrum",
                                           "EnstrophyFluxSpectrum",           "TimeDerivativeEnergySpectrum",
                                           "EnergyDissipationSpectrum",       "EnergyFluxSpectrum"};
      for (int q = 0; q < 6 && e >= 0; ++q)
        e = write_ds(g, names[q], H5T_NATIVE_DOUBLE, 1, &ds, s->flux_spect + (size_t)q * f->n_spect);
    }
    H5Gclose(g);
    if (e < 0) return say(msg, n, "\n[ERROR] --- Unable to write spectra at Iter = [%d]\n", s->index);
  }
  H5Fflush(f->main_f, H5F_SCOPE_LOCAL);
  H5Fflush(f->spec_f, H5F_SCOPE_LOCAL);
  return 0;
}

int ns2d_h5_final(ns2d_h5* f, const ns2d_series* s, char* msg, size_t n) {
  hid_t out = f->sys_msr_in_main ? f->main_f : H5Fcreate(f->sys_msr_file, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
  if (out < 0) return say(msg, n, "\n[ERROR] --- Unable to create system measures file [%s]\n", f->sys_msr_file);
  const long Nx = f->Nx, Ny = f->Ny, Nyf = Ny / 2 + 1;
  int* kx = (int*)malloc(sizeof(int) * Nx);
  int* ky = (int*)malloc(sizeof(int) * Nyf);
  double* x = (double*)malloc(sizeof(double) * Nx);
  double* y = (double*)malloc(sizeof(double) * Ny);
  herr_t e = (kx && ky && x && y) ? 0 : -1;
  if (e >= 0) {
    for (long i = 0; i < Nx; ++i) {
      kx[i] = (int)(i <= Nx / 2 ? i : i - Nx);
      x[i] = (double)i * 2.0 * M_PI / (double)Nx;
    }
    for (long j = 0; j < Ny; ++j) {
      if (j < Nyf) ky[j] = (int)j;
      y[j] = (double)j * 2.0 * M_PI / (double)Ny;
    }
    hsize_t d = (hsize_t)Nx;
    e = H5LTmake_dataset(out, "kx", 1, &d, H5T_NATIVE_INT, kx);
    d = (hsize_t)Nyf;
    if (e >= 0) e = H5LTmake_dataset(out, "ky", 1, &d, H5T_NATIVE_INT, ky);
    d = (hsize_t)Nx;
    if (e >= 0) e = H5LTmake_dataset(out, "x", 1, &d, H5T_NATIVE_DOUBLE, x);
    d = (hsize_t)Ny;
    if (e >= 0) e = H5LTmake_dataset(out, "y", 1, &d, H5T_NATIVE_DOUBLE, y);
    d = (hsize_t)s->n;
    if (e >= 0) e = H5LTmake_dataset(out, "Time", 1, &d, H5T_NATIVE_DOUBLE, s->time);
    d = (hsize_t)Nx;
    if (e >= 0) e = H5LTmake_dataset(out, "MeanFlow_y", 1, &d, H5T_NATIVE_DOUBLE, s->mean_flow_y);
    d = (hsize_t)Nyf;
    if (e >= 0) e = H5LTmake_dataset(out, "MeanFlow_x", 1, &d, H5T_NATIVE_DOUBLE, s->mean_flow_x);
    struct {
      const char* name;
      const double* v;
    } series[] = {{"TotalEnergy", s->tot_energy},
                  {"TotalEnstrophy", s->tot_enstr},
                  {"TotalPalinstrophy", s->tot_palin},
                  {"TotalForcing", s->tot_forc},
                  {"TotalDivergence", s->tot_div},
                  {"EnergyDissipation", s->enrg_diss},
                  {"EnstrophyDissipation", s->enst_diss},
                  {"TimeDerivativeEnstrophySubset", s->d_enst_dt_sbst},
                  {"EnstrophyFluxSubset", s->enst_flux_sbst},
                  {"EnstrophyDissSubset", s->enst_diss_sbst},
                  {"TimeDerivativeEnergySubset", s->d_enrg_dt_sbst},
                  {"EnergyFluxSubset", s->enrg_flux_sbst},
                  {"EnergyDissSubset", s->enrg_diss_sbst}};
    d = (hsize_t)s->n;
    for (size_t q = 0; q < sizeof series / sizeof series[0] && e >= 0; ++q)
      e = H5LTmake_dataset(out, series[q].name, 1, &d, H5T_NATIVE_DOUBLE, series[q].v);
  }
  free(kx);
  free(ky);
  free(x);
  free(y);
  if (!f->sys_msr_in_main) H5Fclose(out);
  else H5Fflush(out, H5F_SCOPE_LOCAL);
  return e < 0 ? say(msg, n, "\n[ERROR] --- Failed to write the system measures to [%s]\n", f->sys_msr_file) : 0;
}

void ns2d_h5_close(ns2d_h5* f) {
  if (!f) return;
  if (f->main_f >= 0) H5Fclose(f->main_f);
  if (f->spec_f >= 0) H5Fclose(f->spec_f);
  if (f->cplx >= 0) H5Tclose(f->cplx);
  free(f);
}

int ns2d_h5_read_psi(const char* path, double* psi, long count, char* msg, size_t n) {
  hid_t f = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT);
  if (f < 0)
    return say(msg, n, "\n[ERROR] --- Unable to open input file [%s] for the initial condition [MAX_PALIN]\n-->> Exiting...\n",
               path);
  int rank = 0;
  herr_t e = H5LTget_dataset_ndims(f, "/psi", &rank);
  hsize_t dims[8] = {0};
  if (e >= 0 && rank > 0 && rank <= 8) e = H5LTget_dataset_info(f, "/psi", dims, NULL, NULL);
  long total = 1;
  for (int r = 0; r < rank; ++r) total *= (long)dims[r];
  if (e < 0 || total != count) {
    H5Fclose(f);
    return say(msg, n, "\n[ERROR] --- Unable to read dataset [/psi] of %ld doubles for the initial condition [MAX_PALIN]\n",
               count);
  }
  e = H5LTread_dataset_double(f, "/psi", psi);
  H5Fclose(f);
  return e < 0 ? say(msg, n, "\n[ERROR] --- Unable to read dataset [/psi] for the initial condition [MAX_PALIN]\n") : 0;
}
