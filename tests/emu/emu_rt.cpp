// Block scheduler of the CPU emulation: the blocks of a launch run one after
// another, each on a pool of threads, one per GPU thread.  LDS is a static array poisoned
// with 0xFF before each block; a write past the launch's dynamic LDS size aborts.
#include <cstdio>
#include <cstdlib>
#include <hip/hip_runtime.h>

#include "ssg_mfma.hpp"

thread_local emu_idx threadIdx, blockIdx;
dim3 blockDim, gridDim;
std::barrier<>* g_bar = nullptr;
std::barrier<>* g_wbar[16];
double g_shfl[16][64];
namespace ssg {
alignas(16) unsigned char smraw[160 * 1024];
WaveX g_wx[16];
}  // namespace ssg

namespace {
// 256 worker threads that run the blocks of every launch in turn (thread t of a
// block on worker t); the host thread releases a block and waits for it
struct Pool {
  std::vector<std::thread> th;
  std::barrier<> go{257}, done{257};
  const std::function<void()>* body = nullptr;
  int nt = 0;
  unsigned blk = 0;
  bool quit = false;
  Pool() {
    for (int t = 0; t < 256; t++)
      th.emplace_back([this, t]() {
        for (;;) {
          go.arrive_and_wait();
          if (quit) return;
          if (t < nt) {
            threadIdx.x = unsigned(t);
            blockIdx.x = blk;
            (*body)();
          }
          done.arrive_and_wait();
        }
      });
  }
  ~Pool() {
    quit = true;
    go.arrive_and_wait();
    for (auto& x : th) x.join();
  }
};
}  // namespace

void emu_launch(dim3 grid, dim3 block, size_t lds, const std::function<void()>& body) {
  static Pool pool;
  blockDim = block;
  gridDim = grid;
  const int nt = int(block.x);
  if (nt > 256) { fprintf(stderr, "blocks of more than 256 threads are not emulated\n"); abort(); }
  for (unsigned b = 0; b < grid.x; b++) {
    std::barrier<> bar(nt);
    g_bar = &bar;
    for (int w = 0; w < (nt + 63) / 64; w++) g_wbar[w] = new std::barrier<>(std::min(64, nt - 64 * w));
    std::memset(ssg::smraw, 0xFF, sizeof(ssg::smraw));
    pool.body = &body;
    pool.nt = nt;
    pool.blk = b;
    pool.go.arrive_and_wait();
    pool.done.arrive_and_wait();
    for (size_t i = lds; i < sizeof(ssg::smraw); i++)
      if (ssg::smraw[i] != 0xFF) {
        fprintf(stderr, "LDS write past the launch's %zu bytes at byte %zu (block %u)\n", lds, i, b);
        abort();
      }
    for (int w = 0; w < (nt + 63) / 64; w++) delete g_wbar[w];
  }
}
