// HBM bandwidth of the two access shapes of the pseudospectral passes on a
// 4096 x (4096/2+1) FP64-complex grid stored [Nx][ld] (kx rows, ky columns),
// moving the column pass's traffic shape: read one field, write four.
//
//   column shape  a workgroup owns C adjacent ky columns and walks all Nx
//                 rows (C x 16 contiguous bytes per row), as the column pass
//                 does for its x transforms; with or without the XCD-aware
//                 workgroup order the column pass uses (neighbouring column
//                 groups, which share cache lines, on one XCD and its L2);
//   row shape     a workgroup owns one kx row (contiguous), as the row pass.
//
// Prints the time per launch (HIP events, 10 launches after a warm-up) and
// the bandwidth over the 5 x Nx x (Ny/2+1) x 16 bytes moved.  Built in-tree
// by build_lib.py (probes/column_bandwidth); run it directly or under
// rocprofv3 --kernel-trace --stats.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

#define CK(x)                                                                     \
  do {                                                                            \
    hipError_t e_ = (x);                                                          \
    if (e_ != hipSuccess) {                                                       \
      std::printf("HIP error %s at line %d\n", hipGetErrorString(e_), __LINE__);  \
      std::exit(1);                                                               \
    }                                                                             \
  } while (0)

constexpr int NX = 4096, NYF = 2049, LD = 2056;

// C columns per workgroup, T threads, E rows per thread per batch; every
// configuration below has NX divisible by (T / C) * E
// workgroup b -> work item: consecutive work items on one XCD (workgroups are
// dealt round-robin over the 8 XCDs)
__device__ __forceinline__ int xcd_order(int b, int n) {
  constexpr int NXCD = 8;
  if (n < NXCD) return b;
  const int q = n / NXCD, r = n % NXCD;
  const int xcd = b % NXCD, j = b / NXCD;
  return (xcd < r ? xcd * (q + 1) : r * (q + 1) + (xcd - r) * q) + j;
}

template <int C, int T, int E, bool XCD>
__global__ __launch_bounds__(T) void column_shape(const double2* __restrict__ in, double2* __restrict__ o0,
                                                  double2* __restrict__ o1, double2* __restrict__ o2,
                                                  double2* __restrict__ o3) {
  constexpr int TL = T / C;
  static_assert(NX % (TL * E) == 0, "rows must tile");
  const int c = threadIdx.x % C, l = threadIdx.x / C;
  const int item = XCD ? xcd_order((int)blockIdx.x, (int)gridDim.x) : (int)blockIdx.x;
  const int ky = item * C + c;
  if (ky >= NYF) return;
  for (int x0 = 0; x0 < NX; x0 += TL * E) {
    double2 v[E];
#pragma unroll
    for (int e = 0; e < E; ++e) v[e] = in[(size_t)(x0 + l + TL * e) * LD + ky];
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const size_t i = (size_t)(x0 + l + TL * e) * LD + ky;
      o0[i] = v[e];
      o1[i] = make_double2(v[e].y, v[e].x);
      o2[i] = make_double2(-v[e].x, v[e].y);
      o3[i] = make_double2(v[e].x, -v[e].y);
    }
  }
}

template <int T>
__global__ __launch_bounds__(T) void row_shape(const double2* __restrict__ in, double2* __restrict__ o0,
                                               double2* __restrict__ o1, double2* __restrict__ o2,
                                               double2* __restrict__ o3) {
  const int x = blockIdx.x;
  for (int ky = threadIdx.x; ky < NYF; ky += T) {
    const size_t i = (size_t)x * LD + ky;
    const double2 v = in[i];
    o0[i] = v;
    o1[i] = make_double2(v.y, v.x);
    o2[i] = make_double2(-v.x, v.y);
    o3[i] = make_double2(v.x, -v.y);
  }
}

template <typename F>
static float time_launch(F launch) {
  hipEvent_t a, b;
  CK(hipEventCreate(&a));
  CK(hipEventCreate(&b));
  launch();
  CK(hipDeviceSynchronize());
  CK(hipEventRecord(a));
  for (int r = 0; r < 10; ++r) launch();
  CK(hipEventRecord(b));
  CK(hipEventSynchronize(b));
  float ms = 0.f;
  CK(hipEventElapsedTime(&ms, a, b));
  CK(hipEventDestroy(a));
  CK(hipEventDestroy(b));
  return ms / 10;
}

int main() {
  const size_t n = (size_t)NX * LD;
  double2 *in, *o[4];
  CK(hipMalloc(&in, n * sizeof(double2)));
  for (auto& p : o) CK(hipMalloc(&p, n * sizeof(double2)));
  CK(hipMemset(in, 0, n * sizeof(double2)));
  const double bytes = 5.0 * NX * NYF * sizeof(double2);
  auto report = [&](const char* what, float ms) {
    std::printf("%-28s %8.1f us  %6.0f GB/s\n", what, ms * 1e3, bytes / ms / 1e6);
    std::fflush(stdout);
  };
#define COLUMN(C, T, E, XCD)                                                                               \
  report("column C=" #C " T=" #T " xcd=" #XCD, time_launch([&] {                                            \
           hipLaunchKernelGGL((column_shape<C, T, E, XCD>), dim3((NYF + C - 1) / C), dim3(T), 0, 0, in, o[0], \
                              o[1], o[2], o[3]);                                                             \
         }));
  COLUMN(1, 256, 16, false)
  COLUMN(1, 256, 16, true)
  COLUMN(2, 512, 16, false)
  COLUMN(2, 512, 16, true)
  COLUMN(4, 512, 16, false)
  COLUMN(4, 512, 16, true)
  COLUMN(8, 512, 8, false)
  COLUMN(8, 512, 8, true)
  COLUMN(16, 512, 8, true)
  report("row T=256", time_launch([&] {
           hipLaunchKernelGGL((row_shape<256>), dim3(NX), dim3(256), 0, 0, in, o[0], o[1], o[2], o[3]);
         }));
  report("row T=512", time_launch([&] {
           hipLaunchKernelGGL((row_shape<512>), dim3(NX), dim3(512), 0, 0, in, o[0], o[1], o[2], o[3]);
         }));
  CK(hipFree(in));
  for (auto p : o) CK(hipFree(p));
  return 0;
}
