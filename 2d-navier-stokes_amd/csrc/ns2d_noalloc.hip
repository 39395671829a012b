// Device scratch of the allocation-free measures entry point
// (ns2d_measures_noalloc): a file-scope device array holding the per-bin
// partial sums of the measures kernel and the spectra formed from them, so that
// the entry point needs no device allocation.  Each GPU holds its own copy
// until the process ends; its address is looked up once per device with
// hipGetSymbolAddress (a host-side use of the array's name is not a device
// address).  Calls that overlap on one GPU share the array and race.
#include <hip/hip_runtime.h>

#include <mutex>

#include "ns2d_internal.h"

namespace ns2d {

// [0, NQ*MAXB): per-bin partials of measures_kernel, bin-major
// [NQ*MAXB, NQ*MAXB + 8*nb): energy spectrum, enstrophy spectrum, six flux
//                            spectra (nb each), then NQ bin-order totals
static __device__ __align__(256) double scratch[NOALLOC_SCRATCH_DOUBLES];

namespace {
constexpr int MAX_DEV = 64;
std::mutex g_mu;
double* g_ptr[MAX_DEV] = {};
}  // namespace

// Address of this device's scratch (the current device must be `dev`).
hipError_t noalloc_scratch(int dev, double** ptr) {
  if (dev < 0 || dev >= MAX_DEV) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_ptr[dev] = static_cast<double*>(p);
  }
  *ptr = g_ptr[dev];
  return hipSuccess;
}

// Spectra (blocks 0 .. nblk-1, one thread per bin) and the bin-order totals
// (block nblk, one thread per quantity) from the per-bin partials, with the
// same operations as the host code of ns2d_measures, so the results agree
// bit for bit.
__global__ __launch_bounds__(256) void measures_finalize_kernel(const double* __restrict__ part, int nb, double cn,
                                                                 int with_flux, double* __restrict__ out) {
  const int nblk = (nb + 255) / 256;
  if ((int)blockIdx.x < nblk) {
    const int b = (int)blockIdx.x * 256 + (int)threadIdx.x;
    if (b >= nb) return;
    const double* q = part + (size_t)b * NQ;
    out[b] = q[Q_ENRG] * cn;
    out[nb + b] = q[Q_ENST] * cn;
    if (with_flux) {
      double* f = out + 2 * (size_t)nb;
      f[0 * (size_t)nb + b] = (q[Q_DERIV] + q[Q_DISS]) * cn;
      f[1 * (size_t)nb + b] = q[Q_DISS] * cn;
      f[2 * (size_t)nb + b] = q[Q_DERIV] * cn;
      f[3 * (size_t)nb + b] = (q[Q_DERIV_K2] + q[Q_DISS_K2]) * cn;
      f[4 * (size_t)nb + b] = q[Q_DISS_K2] * cn;
      f[5 * (size_t)nb + b] = q[Q_DERIV_K2] * cn;
    }
  } else if ((int)threadIdx.x < NQ) {
    const int qi = (int)threadIdx.x;
    double t = 0.0;
    for (int b = 0; b < nb; ++b) t += part[(size_t)b * NQ + qi];
    out[8 * (size_t)nb + qi] = t;
  }
}

hipError_t launch_measures_finalize(const double* part, int nb, double cn, int with_flux, double* out, hipStream_t s) {
  const int nblk = (nb + 255) / 256;
  hipLaunchKernelGGL(measures_finalize_kernel, dim3(nblk + 1), dim3(256), 0, s, part, nb, cn, with_flux, out);
  return hipGetLastError();
}

}  // namespace ns2d
