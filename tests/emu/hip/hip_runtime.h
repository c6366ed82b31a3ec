// CPU emulation of the HIP subset that the chirp-z ring kernels use, for host-only
// checks of their indexing, ordering and LDS use: one std::thread per GPU thread of
// a block, std::barrier for __syncthreads and for the wave-level ordering points,
// blocks run one after another.  A kernel compiled against this header runs its
// own device code on the host, so AddressSanitizer sees every global access.
#pragma once
#include <algorithm>
#include <barrier>
#include <cmath>
#include <cstddef>
#include <cstring>
#include <functional>
#include <thread>
#include <vector>
#define __device__
#define __host__
#define __global__
#define __forceinline__ inline
#define __shared__
#define __restrict__
#define __align__(n) __attribute__((aligned(n)))
#define __launch_bounds__(...)
#define amdgpu_waves_per_eu(...) unused
struct dim3 {
  unsigned x = 1, y = 1, z = 1;
  dim3(unsigned a = 1, unsigned b = 1, unsigned c = 1) : x(a), y(b), z(c) {}
};
struct emu_idx { unsigned x = 0, y = 0, z = 0; };
extern thread_local emu_idx threadIdx, blockIdx;
extern dim3 blockDim, gridDim;
struct double2 { double x, y; };
struct float2 { float x, y; };
struct alignas(16) float4 { float x, y, z, w; };
inline float2 make_float2(float a, float b) { return float2{a, b}; }
struct uint4 { unsigned x, y, z, w; };
inline double2 make_double2(double a, double b) { return double2{a, b}; }
inline uint4 make_uint4(unsigned a, unsigned b, unsigned c, unsigned d) { return uint4{a, b, c, d}; }
typedef void* hipStream_t;
typedef int hipError_t;
#define hipSuccess 0
extern std::barrier<>* g_bar;
extern std::barrier<>* g_wbar[16];
inline void __syncthreads() { g_bar->arrive_and_wait(); }
// the threads of a wave run independently here, so a wave-level ordering point is a
// real barrier of the wave's threads
inline void emu_wave_barrier() { g_wbar[threadIdx.x >> 6]->arrive_and_wait(); }
#define __builtin_amdgcn_wave_barrier() emu_wave_barrier()
#define __builtin_amdgcn_fence(a, b) ((void)0)
#define __builtin_amdgcn_sched_barrier(x) ((void)0)
namespace ssg { extern unsigned char smraw[]; }
using std::ilogb;
using std::isfinite;
using std::ldexp;
// cross-lane exchange within a wave (every lane of the wave takes part)
extern double g_shfl[16][64];
inline double __shfl_xor(double v, int mask) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  g_shfl[w][lane] = v;
  g_wbar[w]->arrive_and_wait();
  const double r = g_shfl[w][lane ^ mask];
  g_wbar[w]->arrive_and_wait();
  return r;
}
inline int min(int a, int b) { return a < b ? a : b; }
inline int max(int a, int b) { return a > b ? a : b; }
inline long min(long a, long b) { return a < b ? a : b; }
inline unsigned min(unsigned a, unsigned b) { return a < b ? a : b; }
void emu_launch(dim3 grid, dim3 block, size_t lds, const std::function<void()>& body);
#define hipLaunchKernelGGL(K, G, B, LDS, ST, ...) emu_launch(G, B, LDS, [&]() { K(__VA_ARGS__); })
inline hipError_t hipGetLastError() { return 0; }
