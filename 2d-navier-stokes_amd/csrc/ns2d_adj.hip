// Shape tables and launchers of the discrete-adjoint passes (rpass_adj,
// kpass_adj), instantiated in a translation unit of their own.
#include "ns2d_kernels.hpp"

namespace ns2d {

// adjoint pass shapes per transform length; index 0 is the default, others
// are selectable (NS2D_KADJ_VARIANT_<N>, NS2D_RADJ_VARIANT_<N>) for tuning.
// Row pass {NY, TLR, PAIRS, SR}: SR = 1 runs one row per TLR threads.
#define NS2D_RADJ_VARIANTS(X) \
  X(8, 1, 64, 0)       \
  X(16, 1, 64, 0)      \
  X(32, 4, 16, 0)      \
  X(64, 8, 8, 0)       \
  X(128, 32, 2, 0)     \
  X(128, 16, 4, 0)     \
  X(256, 64, 1, 1)     \
  X(256, 64, 1, 0)     \
  X(256, 16, 4, 0)     \
  X(256, 32, 2, 0)     \
  X(512, 64, 1, 1)     \
  X(512, 64, 1, 0)     \
  X(512, 128, 1, 0)    \
  X(512, 32, 2, 0)     \
  X(1024, 256, 1, 1)   \
  X(1024, 256, 1, 0)   \
  X(1024, 64, 1, 0)    \
  X(1024, 128, 1, 0)   \
  X(2048, 256, 1, 0)   \
  X(2048, 128, 1, 0)   \
  X(2048, 512, 1, 0)   \
  X(4096, 512, 1, 0)   \
  X(4096, 256, 1, 0)  

#define NS2D_KADJ_VARIANTS(X) \
  X(8, 16, 64)       \
  X(16, 16, 64)      \
  X(32, 16, 64)      \
  X(64, 16, 64)      \
  X(128, 16, 128)    \
  X(128, 8, 128)     \
  X(256, 4, 256)     \
  X(256, 4, 128)     \
  X(256, 8, 128)     \
  X(256, 1, 64)      \
  X(256, 1, 128)     \
  X(512, 2, 256)     \
  X(512, 2, 128)     \
  X(512, 4, 128)     \
  X(512, 4, 256)     \
  X(512, 1, 128)     \
  X(512, 1, 64)      \
  X(1024, 2, 512)    \
  X(1024, 2, 256)    \
  X(1024, 2, 128)    \
  X(1024, 1, 256)    \
  X(1024, 1, 128)    \
  X(2048, 2, 512)    \
  X(2048, 2, 256)    \
  X(4096, 2, 1024)   \
  X(4096, 1, 256)    \
  X(4096, 2, 512)   

struct RAdjVariant {
  int NY, TLR, PAIRS, SR;
  void (*kern)(RAdjArgs);
};
struct KAdjVariant {
  int L, C, T;
  void (*kern)(KAdjArgs);
};
#define NS2D_RADJ_ENTRY(NY_, TLR_, PAIRS_, SR_) {NY_, TLR_, PAIRS_, SR_, rpass_adj_kernel<NY_, TLR_, PAIRS_, SR_>},
#define NS2D_KADJ_ENTRY(L_, C_, T_) {L_, C_, T_, kpass_adj_kernel<L_, C_, T_>},
static const RAdjVariant radj_variants[] = {NS2D_RADJ_VARIANTS(NS2D_RADJ_ENTRY)};
static const KAdjVariant kadj_variants[] = {NS2D_KADJ_VARIANTS(NS2D_KADJ_ENTRY)};
#undef NS2D_RADJ_ENTRY
#undef NS2D_KADJ_ENTRY

// index chosen through the environment (read once per length), default 0
static int adj_choice(const char* prefix, int n) {
  char name[64];
  std::snprintf(name, sizeof name, "%s%d", prefix, n);
  const char* v = std::getenv(name);
  return v ? std::atoi(v) : 0;
}

hipError_t launch_rpass_adj(int nx, int ny, const RAdjArgs& a, hipStream_t s) {
  static int idx[16] = {-1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1};
  const int lg = ny > 0 ? 31 - __builtin_clz((unsigned)ny) : 0;
  if (lg >= 16) return hipErrorInvalidValue;
  if (idx[lg] < 0) idx[lg] = adj_choice("NS2D_RADJ_VARIANT_", ny);
  const RAdjVariant* v = nullptr;
  int k = 0;
  for (const RAdjVariant& u : radj_variants) {
    if (u.NY != ny) continue;
    if (!v) v = &u;
    if (k++ == idx[lg]) {
      v = &u;
      break;
    }
  }
  if (!v) return hipErrorInvalidValue;
  const int units = v->SR ? nx : nx / 2;   // rows or row pairs
  hipLaunchKernelGGL(v->kern, dim3((units + v->PAIRS - 1) / v->PAIRS), dim3(v->TLR * v->PAIRS), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_kpass_adj(int nx, const KAdjArgs& a, hipStream_t s) {
  static int idx[16] = {-1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1};
  const int lg = nx > 0 ? 31 - __builtin_clz((unsigned)nx) : 0;
  if (lg >= 16) return hipErrorInvalidValue;
  if (idx[lg] < 0) idx[lg] = adj_choice("NS2D_KADJ_VARIANT_", nx);
  const KAdjVariant* v = nullptr;
  int k = 0;
  for (const KAdjVariant& u : kadj_variants) {
    if (u.L != nx) continue;
    if (!v) v = &u;
    if (k++ == idx[lg]) {
      v = &u;
      break;
    }
  }
  if (!v) return hipErrorInvalidValue;
  hipLaunchKernelGGL(v->kern, dim3((a.nyf + v->C - 1) / v->C), dim3(v->T), 0, s, a);
  return hipGetLastError();
}

}  // namespace ns2d
