// Runtime of the CPU stand-in for HIP (tests/cpuemu/hip/hip_runtime.h):
// launches every workgroup of a grid with one host thread per GPU thread
// (workgroups in turn, so __shared__ statics are per workgroup), with
// workgroup and wave barriers.  Test infrastructure only.
#include <hip/hip_runtime.h>

#include <condition_variable>

thread_local dim3 threadIdx, blockIdx, gridDim, blockDim;

namespace {
struct WaveBar {
  std::mutex m;
  std::condition_variable cv;
  int expected = 0, arrived = 0, live = 0;
  long gen = 0;
};
WaveBar g_wave[16];
pthread_barrier_t g_block;
std::mutex g_launch_mu;

// this thread stops counting towards its wave's barrier (waiting threads that
// only waited for it are released)
void wave_leave(unsigned w, bool finished) {
  WaveBar& b = g_wave[w];
  std::lock_guard<std::mutex> lk(b.m);
  --b.expected;
  if (finished) --b.live;
  if (b.arrived > 0 && b.arrived >= b.expected) {
    b.arrived = 0;
    ++b.gen;
    b.cv.notify_all();
  }
}

void block_barrier_and_reset(unsigned t) {
  pthread_barrier_wait(&g_block);
  if (t % 64 == 0) {
    WaveBar& b = g_wave[t / 64];
    std::lock_guard<std::mutex> lk(b.m);
    b.expected = b.live;
    b.arrived = 0;
  }
  pthread_barrier_wait(&g_block);
}
}  // namespace

double emu_shfl_d[1024];
unsigned long long emu_shfl_u[1024];
int emu_dpp_i[1024];
std::mutex emu_atomic_mu;
thread_local emu_graph* emu_capturing = nullptr;

void emu_wave_arrive(unsigned w) {
  WaveBar& b = g_wave[w];
  std::unique_lock<std::mutex> lk(b.m);
  const long g = b.gen;
  if (++b.arrived >= b.expected) {
    b.arrived = 0;
    ++b.gen;
    b.cv.notify_all();
  } else {
    b.cv.wait(lk, [&] { return b.gen != g; });
  }
}

void emu_syncthreads() {
  const unsigned t = threadIdx.x;
  wave_leave(t / 64, false);
  block_barrier_and_reset(t);
}

// A pool of 1024 host threads, created once; a launch runs the grid's
// workgroups in turn on the first blockDim.x of them.
namespace {
struct Pool {
  std::mutex m;
  std::condition_variable go[16], done;
  long epoch = 0;
  int running = 0;
  unsigned T = 0;
  dim3 grid, block;
  std::function<void()>* body = nullptr;
  std::vector<std::thread> th;

  void worker(unsigned t) {
    long seen = 0;
    for (;;) {
      {
        std::unique_lock<std::mutex> lk(m);
        go[t / 64].wait(lk, [&] { return epoch != seen && t < T; });
        seen = epoch;
      }
      threadIdx = dim3(t);
      blockDim = block;
      gridDim = grid;
      for (unsigned by = 0; by < grid.y; ++by)
        for (unsigned bx = 0; bx < grid.x; ++bx) {
          blockIdx = dim3(bx, by);
          (*body)();
          wave_leave(t / 64, true);
          pthread_barrier_wait(&g_block);
          if (t % 64 == 0) {
            const unsigned w = t / 64;
            const int n = (w + 1) * 64 <= T ? 64 : (int)(T - w * 64);
            g_wave[w].expected = g_wave[w].live = n;
            g_wave[w].arrived = 0;
          }
          pthread_barrier_wait(&g_block);
        }
      std::lock_guard<std::mutex> lk(m);
      if (--running == 0) done.notify_all();
    }
  }
};
Pool* g_pool = nullptr;
}  // namespace

void emu_run(dim3 grid, dim3 block, std::function<void()> body) {
  std::lock_guard<std::mutex> lk(g_launch_mu);
  const unsigned T = block.x;
  if (T == 0 || T > 1024 || (T % 64 != 0 && T > 64)) {
    std::fprintf(stderr, "emu: bad block size %u\n", T);
    std::abort();
  }
  if (!g_pool) {
    g_pool = new Pool();
    for (unsigned t = 0; t < 1024; ++t) g_pool->th.emplace_back([t] { g_pool->worker(t); });
    for (auto& x : g_pool->th) x.detach();
  }
  pthread_barrier_init(&g_block, nullptr, T);
  const unsigned nw = (T + 63) / 64;
  for (unsigned w = 0; w < nw; ++w) {
    const int n = (w + 1) * 64 <= T ? 64 : (int)(T - w * 64);
    g_wave[w].expected = g_wave[w].live = n;
    g_wave[w].arrived = 0;
  }
  Pool& p = *g_pool;
  {
    std::unique_lock<std::mutex> pl(p.m);
    p.T = T;
    p.grid = grid;
    p.block = block;
    p.body = &body;
    p.running = (int)T;
    ++p.epoch;
    for (unsigned w = 0; w < nw; ++w) p.go[w].notify_all();
    p.done.wait(pl, [&] { return p.running == 0; });
  }
  pthread_barrier_destroy(&g_block);
}
