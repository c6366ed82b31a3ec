// CPU emulation of the 16x16x4 MFMA products (stands in for csrc/ssg_mfma.hpp): the
// lanes of a wave exchange their operands through a per-wave buffer.  Layout as on
// gfx950: A[i][k] in lane i + 16k, B[k][j] in lane j + 16k; result value r of lane l
// is row (l / 16) + 4r (FP64) or 4 (l / 16) + r (FP32), column l % 16.
#pragma once
#include <hip/hip_runtime.h>
namespace ssg {
struct WaveX { double a[64], b[64], c[64][4]; };
extern WaveX g_wx[16];
template <typename R> struct Mfma;
template <typename R, typename V, int (*ROW)(int, int)>
inline V emu_mma(R a, R b, V c) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  WaveX& X = g_wx[w];
  X.a[lane] = a;
  X.b[lane] = b;
  for (int r = 0; r < 4; r++) X.c[lane][r] = c[r];
  g_wbar[w]->arrive_and_wait();
  V out;
  for (int r = 0; r < 4; r++) {
    const int i = ROW(lane, r), j = lane & 15;
    double s = X.c[lane][r];
    for (int k = 0; k < 4; k++) s += X.a[i + 16 * k] * X.b[j + 16 * k];
    out[r] = R(s);
  }
  g_wbar[w]->arrive_and_wait();
  return out;
}
inline int row_f64(int lane, int r) { return (lane >> 4) + 4 * r; }
inline int row_f32(int lane, int r) { return 4 * (lane >> 4) + r; }
template <> struct Mfma<double> {
  typedef double v4 __attribute__((ext_vector_type(4)));
  static v4 mma(double a, double b, v4 c) { return emu_mma<double, v4, row_f64>(a, b, c); }
  static int row(int lane, int r) { return row_f64(lane, r); }
};
template <> struct Mfma<float> {
  typedef float v4 __attribute__((ext_vector_type(4)));
  static v4 mma(float a, float b, v4 c) { return emu_mma<float, v4, row_f32>(a, b, c); }
  static int row(int lane, int r) { return row_f32(lane, r); }
};
}  // namespace ssg
