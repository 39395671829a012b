// Host-side launchers of the fused kernels: the pass-shape registries
// (instantiated per transform length in ns2d_pass_tu.hip), the row/column pass
// launches, the real-field forward transform and the measures reduction.
#include "ns2d_kernels.hpp"

namespace ns2d {
// ---------------------------------------------------------------------------
// host-side launchers (template dispatch on the transform lengths)
// ---------------------------------------------------------------------------

// the shape instantiations, one function per translation unit of
// ns2d_pass_tu.hip (built once per length range, ascending)
#define NS2D_PASS_TUS(X) X(8, 128) X(256, 256) X(512, 512) X(1024, 1024) X(2048, 2048) X(4096, 4096)
#define NS2D_TU_DECL(LO, HI)                           \
  void kv_add_##LO##_##HI(std::vector<KVariant>& v); \
  void rv_add_##LO##_##HI(std::vector<RVariant>& v);
NS2D_PASS_TUS(NS2D_TU_DECL)
#undef NS2D_TU_DECL

const std::vector<KVariant>& kvariant_table() {
  static const std::vector<KVariant> t = [] {
    std::vector<KVariant> v;
#define NS2D_TU_CALL(LO, HI) kv_add_##LO##_##HI(v);
    NS2D_PASS_TUS(NS2D_TU_CALL)
#undef NS2D_TU_CALL
    return v;
  }();
  return t;
}

// FNV-1a hash of the compiled column-pass shapes of length nx and row-pass
// shapes of length ny (a tuned schedule names shapes by index, valid only for
// the same tables)
unsigned long long shape_table_hash(int nx, int ny) {
  unsigned long long h = 1469598103934665603ull;
  auto mix = [&](int x) {
    for (int b = 0; b < 4; ++b) {
      h ^= (unsigned long long)((x >> (8 * b)) & 0xff);
      h *= 1099511628211ull;
    }
  };
  for (const KVariant& v : kvariant_table())
    if (v.L == nx) {
      const int f[] = {v.L, v.C, v.T, v.KS, v.PS, v.OFF, v.PK, v.W, v.PF};
      for (int x : f) mix(x);
    }
  mix(-1);
  for (const RVariant& v : rvariant_table())
    if (v.NY == ny) {
      const int f[] = {v.NY, v.TLR, v.PAIRS, v.W, v.PFR};
      for (int x : f) mix(x);
    }
  return h;
}

const std::vector<RVariant>& rvariant_table() {
  static const std::vector<RVariant> t = [] {
    std::vector<RVariant> v;
#define NS2D_TU_CALL(LO, HI) rv_add_##LO##_##HI(v);
    NS2D_PASS_TUS(NS2D_TU_CALL)
#undef NS2D_TU_CALL
    return v;
  }();
  return t;
}

static const KVariant* find_kvariant(int nx, int idx) {
  int n = 0;
  for (const KVariant& v : kvariant_table()) {
    if (v.L != nx) continue;
    if (n++ == idx) return &v;
  }
  return nullptr;
}

// Every member of the batch must run a pass with inverse transforms (KS as
// compiled) and share nyf.
hipError_t launch_kpass_batched(int nx, int nyf, int variant, const KPassArgs* tab, int members, hipStream_t s) {
  const KVariant* v = find_kvariant(nx, variant);
  if (!v) return hipErrorInvalidValue;
  const int per = v->groups(nyf) * v->KS;
  hipLaunchKernelGGL(v->batched, dim3(per * members), dim3(v->T), 0, s, tab, per);
  return hipGetLastError();
}

// Default column-pass shape: index 0 of its length for both integrators (at
// 4096 the unpacked 1024-thread shape with the 16-slot padding: 320.5
// against 326.9 us per RK4CN launch for the 512-thread one, 444.2 against
// 445.0 us for CN-AB4's heavier epilogue, in alternating bench runs).
int kpass_default_variant(int nx, bool ab4) {
  (void)nx;
  (void)ab4;
  return 0;
}

int kpass_variant_count(int nx) {
  int n = 0;
  for (const KVariant& v : kvariant_table()) n += v.L == nx;
  return n;
}

int kpass_variant_info(int nx, int idx, int* C, int* E, int* T, int* KS) {
  const KVariant* v = find_kvariant(nx, idx);
  if (!v) return -1;
  *C = v->C;
  *E = v->L * v->C / v->T;
  *T = v->T;
  *KS = v->KS;
  return 0;
}

int kpass_variant_extra(int nx, int idx, int* W, int* PS, int* OFF, int* PK) {
  const KVariant* v = find_kvariant(nx, idx);
  if (!v) return -1;
  *W = v->W;
  *PS = v->PS;
  *OFF = v->OFF;
  *PK = (v->PK != 0) | (v->PF << 1) | ((v->PK == 2) << 2);
  return 0;
}

// Launch the kpass for an nx-point column transform.  A pass without inverse
// transforms (an RHS evaluation) uses a KS = 1 sibling of the same shape:
// split workgroups exist only to share the inverse fields.
hipError_t launch_kpass(int nx, int variant, const KPassArgs& a, hipStream_t s) {
  const KVariant* v = find_kvariant(nx, variant);
  if (!v) return hipErrorInvalidValue;
  if (!(a.mode & KP_INV) && v->KS != 1) {
    const KVariant* w = nullptr;
    for (const KVariant& u : kvariant_table())
      if (u.L == nx && u.KS == 1 && u.C == v->C && u.T == v->T && u.PS == v->PS && u.OFF == v->OFF &&
          u.PK == v->PK && u.W == v->W && u.PF == v->PF) {
        w = &u;
        break;
      }
    if (!w)
      for (const KVariant& u : kvariant_table())
        if (u.L == nx && u.KS == 1) { w = &u; break; }
    if (!w) return hipErrorInvalidValue;
    v = w;
  }
  const int groups = v->groups(a.nyf);
  // A last column group holding a single column (unpacked, KS = 1, more than
  // one column per group) runs as two half items: it would otherwise be the
  // one workgroup left when every CU has finished its share
  KPassArgs b = a;
  b.last_split = (v->KS == 1 && v->PK == 0 && v->C > 1 && (a.mode & KP_INV) && a.nyf % v->C == 1) ? 1 : 0;
  hipLaunchKernelGGL(v->kern, dim3(groups * v->KS + b.last_split), dim3(v->T), 0, s, b);
  return hipGetLastError();
}



static const RVariant* find_rvariant(int ny, int idx) {
  int n = 0;
  for (const RVariant& v : rvariant_table()) {
    if (v.NY != ny) continue;
    if (n++ == idx) return &v;
  }
  return nullptr;
}

hipError_t launch_rpass_batched(int nx, int ny, int variant, const RPassArgs* tab, int members, hipStream_t s) {
  const RVariant* v = find_rvariant(ny, variant);
  if (!v) return hipErrorInvalidValue;
  const int units = (v->PFR & 7) >= 2 ? nx : nx / 2;   // rows or row pairs
  const int per = (units + v->PAIRS - 1) / v->PAIRS;
  hipLaunchKernelGGL(v->batched, dim3(per * members), dim3(v->TLR * v->PAIRS), 0, s, tab, per);
  return hipGetLastError();
}

int rpass_variant_count(int ny) {
  int n = 0;
  for (const RVariant& v : rvariant_table()) n += v.NY == ny;
  return n;
}

int rpass_variant_info(int ny, int idx, int* TLR, int* E, int* T, int* PAIRS) {
  const RVariant* v = find_rvariant(ny, idx);
  if (!v) return -1;
  *TLR = v->TLR;
  *E = v->NY / v->TLR;
  *T = v->TLR * v->PAIRS;
  *PAIRS = v->PAIRS;
  return 0;
}

int rpass_variant_extra(int ny, int idx, int* W, int* PS, int* OFF, int* PK) {
  const RVariant* v = find_rvariant(ny, idx);
  if (!v) return -1;
  *W = v->W;
  *PS = 3;
  *OFF = 1;
  const int pm = v->PFR & 7;
  *PK = ((pm == 1 || pm == 3 || pm == 4) << 1) | ((pm >= 2) << 3) | ((pm >= 4) << 4) | (((v->PFR & 8) != 0) << 5);
  return 0;
}

hipError_t launch_rpass(int nx, int ny, int variant, const RPassArgs& a, hipStream_t s) {
  const RVariant* v = find_rvariant(ny, variant);
  if (!v) return hipErrorInvalidValue;
  const int units = (v->PFR & 7) >= 2 ? nx : nx / 2;   // rows or row pairs
  hipLaunchKernelGGL((a.mode & RP_CFL) ? v->cfl : v->kern, dim3((units + v->PAIRS - 1) / v->PAIRS),
                     dim3(v->TLR * v->PAIRS), 0, s, a);
  return hipGetLastError();
}

template <int NY>
struct R2CCfg {
  static constexpr int T = NY >= 1024 ? 256 : (NY / 4 < 64 ? 64 : NY / 4);
};

template <int NY>
static hipError_t launch_r2c_rows_t(int nx, const double* in, double2* out, const double2* tw, int ld,
                                    hipStream_t s) {
  constexpr int T = R2CCfg<NY>::T;
  hipLaunchKernelGGL((r2c_rows_kernel<NY, T>), dim3(nx), dim3(T), 0, s, in, out, tw, ld);
  return hipGetLastError();
}

hipError_t launch_r2c_rows(int nx, int ny, const double* in, double2* out, const double2* tw, int ld,
                           hipStream_t s) {
  switch (ny) {
    case 8: return launch_r2c_rows_t<8>(nx, in, out, tw, ld, s);
    case 16: return launch_r2c_rows_t<16>(nx, in, out, tw, ld, s);
    case 32: return launch_r2c_rows_t<32>(nx, in, out, tw, ld, s);
    case 64: return launch_r2c_rows_t<64>(nx, in, out, tw, ld, s);
    case 128: return launch_r2c_rows_t<128>(nx, in, out, tw, ld, s);
    case 256: return launch_r2c_rows_t<256>(nx, in, out, tw, ld, s);
    case 512: return launch_r2c_rows_t<512>(nx, in, out, tw, ld, s);
    case 1024: return launch_r2c_rows_t<1024>(nx, in, out, tw, ld, s);
    case 2048: return launch_r2c_rows_t<2048>(nx, in, out, tw, ld, s);
    case 4096: return launch_r2c_rows_t<4096>(nx, in, out, tw, ld, s);
    default: return hipErrorInvalidValue;
  }
}

// Between the host layout [Nx][nyf] (contiguous rows, what callers pass) and
// the device k-space layout (kidx, padding columns ky >= nyf zero): dir 0 packs
// src (host layout) into dst, dir 1 unpacks src into dst (host layout).
__global__ __launch_bounds__(256) void relayout_kernel(const double2* __restrict__ src, double2* __restrict__ dst,
                                                       int nyf, int ld, int nx, int dir) {
  const int x = blockIdx.y, ky = blockIdx.x * 256 + threadIdx.x;
  if (dir == 0) {
    if (ky < ld) dst[kidx(x, ky, ld, nx)] = ky < nyf ? src[(size_t)x * nyf + ky] : make_double2(0.0, 0.0);
  } else if (ky < nyf) {
    dst[(size_t)x * nyf + ky] = src[kidx(x, ky, ld, nx)];
  }
}

hipError_t launch_relayout(const double2* src, double2* dst, int nx, int nyf, int ld, int dir, hipStream_t s) {
  hipLaunchKernelGGL(relayout_kernel, dim3((ld + 255) / 256, nx), dim3(256), 0, s, src, dst, nyf, ld, nx, dir);
  return hipGetLastError();
}

hipError_t launch_measures(int nbins, const MeasArgs& a, hipStream_t s) {
  if (a.nbins != nbins) return hipErrorInvalidValue;
  hipLaunchKernelGGL((measures_kernel<256>), dim3(meas_grid(nbins)), dim3(256), 0, s, a);
  return hipGetLastError();
}

}  // namespace ns2d
