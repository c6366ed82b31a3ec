// ssg_regular.hip -- evaluation of a grid's spectral coefficients on a
// tensor-product set of points (regularGridTransform, RL: r x lambda,
// src/transforms_cylindrical.jl:929-1020; RLZ: r x lambda x z, 1168-1287;
// RZ: x x z, src/transforms_cartesian.jl:946-1001).
//
// The reference runs, per variable, SA on every spectral line, SI at the
// requested radii, a Fourier sum at every (r, lambda) and a Chebyshev sum at
// every z.  Here the banded solve is the grid transform's own kernel; the
// evaluation at arbitrary radii is k_si_pts (four taps per point, weights from
// the host); the Fourier and Chebyshev sums are dense contractions with tables
// the host builds from the points (cos / sin of k lambda with the R2HC factor
// 2 and the derivative factors folded in; the Chebyshev series and its z
// derivatives at each z composed with the CAtransform! fold), run by
// k_gemm_f64 on the FP64 matrix cores.
#include "ssg_kernels.hpp"
#include "ssg_mfma.hpp"

namespace ssg {

// ----------------------------------------------------------------------------
// SI at arbitrary radii.  One lane per banded-solve lane (RL / RLZ: the
// (transform, k, part) lanes of k_sa_ring; RZ: the lines of k_sa_si), x read
// from y and folded to a = G^T x on the fly.  Point ip has first tap m = pxm[ip]
// and weights pw[(d * npt + ip) * 4 + t] = phi^(d)_{m+t}(x_ip), zero for taps
// outside -1 .. nc+1 (SItransform, src/CubicBSpline.jl:1554-1567).
// Output S[d][ip][line]: RL / RLZ line = t * (2K + 1) + ks with ks = 0 (k = 0),
// 2k - 1 (cosine line of k), 2k (sine line); RZ line = var * nzb + zb.
// ----------------------------------------------------------------------------
template <bool Ring>
__global__ void __launch_bounds__(256) k_si_pts(DevTables T, const double* __restrict__ y, int npt,
                                                const int* __restrict__ pxm, const double* __restrict__ pw,
                                                double* __restrict__ S) {
  const long nlanes = Ring ? 2 * long(T.ntr) * (T.K + 1) : T.nlines;
  const long u = long(blockIdx.x) * blockDim.x + threadIdx.x;
  if (u >= nlanes) return;
  const int nks = 2 * T.K + 1;
  long line;
  int v;
  bool store = true;
  if (Ring) {
    const int part = int(u & 1);
    const long pu = u >> 1;
    const int k = int(pu % (T.K + 1));
    const int t = int(pu / (T.K + 1));
    v = t / T.nzb;
    line = long(t) * nks + (k == 0 ? 0 : 2 * k - 1 + part);
    store = !(k == 0 && part == 1);                 // both lanes of k = 0 solve line 0
  } else {
    line = u;
    v = int(u / T.nzb);
  }
  const long nline = Ring ? long(T.ntr) * nks : T.nlines;
  const int* gi = T.gint + 8 * v;
  const double* gd = T.gdbl + 4 * v;
  const int Mdim = gi[0], Min = gi[1], rL = gi[2], lk = gi[4], rk = gi[5];
  const double aL = gd[0], bL = gd[1], aR = gd[2], bR = gd[3];
  const double xf0 = y[u], xf1 = y[nlanes + u];
  const double xl1 = y[long(Min - 1) * nlanes + u], xl2 = y[long(Min - 2) * nlanes + u];
  auto a_at = [&](int j) -> double {
    const bool in = j >= rL && j < rL + Min;
    const double yj = y[long(in ? j - rL : 0) * nlanes + u];
    double a = in ? yj : 0.0;
    a = (lk == 1 && j == 0) ? aL * xf0 + bL * xf1 : a;
    a = (lk == 2 && j == 0) ? aL * xf0 : a;
    a = (lk == 2 && j == 1) ? bL * xf0 : a;
    a = (rk == 1 && j == Mdim - 1) ? aR * xl1 + bR * xl2 : a;
    a = (rk == 2 && j == Mdim - 2) ? bR * xl1 : a;
    a = (rk == 2 && j == Mdim - 1) ? aR * xl1 : a;
    return a;
  };
  // points are usually sorted: the four coefficients are reloaded only when
  // the first tap moves
  int cur = -1000000;
  double w0 = 0, w1 = 0, w2 = 0, w3 = 0;
  for (int ip = 0; ip < npt; ip++) {
    const int m = pxm[ip];
    if (m != cur) {
      cur = m;
      w0 = a_at(m + 1); w1 = a_at(m + 2); w2 = a_at(m + 3); w3 = a_at(m + 4);   // a index = m + 1 (0-based)
    }
#pragma unroll
    for (int d = 0; d < 3; d++) {
      const double* w = pw + (long(d) * npt + ip) * 4;
      double s = 0.0;
      s += w[0] * w0;
      s += w[1] * w1;
      s += w[2] * w2;
      s += w[3] * w3;
      if (store) S[(long(d) * npt + ip) * nline + line] = s;
    }
  }
}

void launch_si_pts(const DevTables& T, bool ring, const double* y, int npt, const int* pxm, const double* pw,
                   double* S, hipStream_t st) {
  const long nlanes = ring ? 2 * long(T.ntr) * (T.K + 1) : T.nlines;
  const dim3 grid(unsigned((nlanes + 255) / 256));
  if (ring) hipLaunchKernelGGL(k_si_pts<true>, grid, dim3(256), 0, st, T, y, npt, pxm, pw, S);
  else hipLaunchKernelGGL(k_si_pts<false>, grid, dim3(256), 0, st, T, y, npt, pxm, pw, S);
}

// ----------------------------------------------------------------------------
// The Fourier rows at the requested angles: the sum of regularGridTransform
// (src/transforms_cylindrical.jl:990-999) is f = a0 + 2 sum_k (rc cos k lambda
// + rs sin k lambda); its lambda derivatives bring down k and -k^2.  The
// argument is the FP64 product k * lambda, as there.  One thread per (ks, j).
// ----------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_trig_table(const double* __restrict__ lam, int nj, int K,
                                                    double* __restrict__ trig) {
  const int nks = 2 * K + 1;
  const long u = long(blockIdx.x) * blockDim.x + threadIdx.x;
  if (u >= long(nks) * nj) return;
  const int j = int(u % nj), ks = int(u / nj);
  const long plane = long(nks) * nj;
  double v0, v1, v2;
  if (ks == 0) {
    v0 = 1.0; v1 = 0.0; v2 = 0.0;
  } else {
    const int k = (ks + 1) >> 1;
    const double kk = double(k);
    double s, c;
    sincos(kk * lam[j], &s, &c);
    if (ks & 1) {                                     // cosine line
      v0 = 2.0 * c; v1 = -2.0 * kk * s; v2 = -2.0 * (kk * kk) * c;
    } else {                                          // sine line
      v0 = 2.0 * s; v1 = 2.0 * kk * c; v2 = -2.0 * (kk * kk) * s;
    }
  }
  trig[u] = v0;
  trig[plane + u] = v1;
  trig[2 * plane + u] = v2;
}

void launch_trig_table(const double* lam, int nj, int K, double* trig, hipStream_t st) {
  const long n = long(2 * K + 1) * nj;
  hipLaunchKernelGGL(k_trig_table, dim3(unsigned((n + 255) / 256)), dim3(256), 0, st, lam, nj, K, trig);
}

// ----------------------------------------------------------------------------
// Batched FP64 GEMM C = A B with element strides (any of the layouts the
// regular-grid stages use) and one batch stride per operand.  Block tile 64 x 64,
// four waves of 32 x 32 (2 x 2 tiles of v_mfma_f64_16x16x4), k staged 16 at a
// time through LDS with the next chunk's loads in flight during the products.
// Loads run along whichever of a tile's two dimensions is contiguous.  Column
// tiles run fastest over blockIdx.x (one linear index over all tiles), so the
// blocks that share a row panel of A are resident together and A streams from
// HBM once.
// ----------------------------------------------------------------------------
static constexpr int kGT = 64, kGK = 16;

__global__ void __launch_bounds__(256) k_gemm_f64(GemmDesc g) {
  typedef Mfma<double>::v4 v4;
  __shared__ double As[kGK][kGT + 2];                // As[k][i]
  __shared__ double Bs[kGK][kGT + 2];                // Bs[k][j]
  const long b = blockIdx.z;
  const int ntj = (g.N + kGT - 1) / kGT;
  const int i0 = int(blockIdx.x / ntj) * kGT, j0 = int(blockIdx.x % ntj) * kGT;
  const double* A = g.A + b * g.sA;
  const double* B = g.B + b * g.sB;
  double* C = g.C + b * g.sC;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int wi = w & 1, wj = w >> 1;
  const bool a_kfast = g.ak == 1, b_jfast = g.bj == 1;
  double ra[4], rb[4];
  auto a_ik = [&](int e, int& i, int& k) { if (a_kfast) { k = e & 15; i = e >> 4; } else { i = e & 63; k = e >> 6; } };
  auto b_kj = [&](int e, int& k, int& j) { if (b_jfast) { j = e & 63; k = e >> 6; } else { k = e & 15; j = e >> 4; } };
  auto load = [&](int k0) {
#pragma unroll
    for (int s = 0; s < 4; s++) {
      const int e = tid + 256 * s;
      int i, k, j, kb;
      a_ik(e, i, k);
      ra[s] = (i0 + i < g.M && k0 + k < g.K) ? A[long(i0 + i) * g.ai + long(k0 + k) * g.ak] : 0.0;
      b_kj(e, kb, j);
      rb[s] = (k0 + kb < g.K && j0 + j < g.N) ? B[long(k0 + kb) * g.bk + long(j0 + j) * g.bj] : 0.0;
    }
  };
  auto put = [&]() {
#pragma unroll
    for (int s = 0; s < 4; s++) {
      const int e = tid + 256 * s;
      int i, k, j, kb;
      a_ik(e, i, k);
      As[k][i] = ra[s];
      b_kj(e, kb, j);
      Bs[kb][j] = rb[s];
    }
  };
  v4 acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int c = 0; c < 2; c++) acc[a][c] = v4{0, 0, 0, 0};
  load(0);
  for (int k0 = 0; k0 < g.K; k0 += kGK) {
    __syncthreads();                                 // the previous chunk's products are done
    put();
    __syncthreads();
    if (k0 + kGK < g.K) load(k0 + kGK);
#pragma unroll
    for (int ks = 0; ks < kGK / 4; ks++) {
      const int kr = ks * 4 + (lane >> 4);
      double av[2], bv[2];
#pragma unroll
      for (int a = 0; a < 2; a++) av[a] = As[kr][wi * 32 + a * 16 + (lane & 15)];
#pragma unroll
      for (int c = 0; c < 2; c++) bv[c] = Bs[kr][wj * 32 + c * 16 + (lane & 15)];
#pragma unroll
      for (int a = 0; a < 2; a++)
#pragma unroll
        for (int c = 0; c < 2; c++) acc[a][c] = Mfma<double>::mma(av[a], bv[c], acc[a][c]);
    }
  }
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int c = 0; c < 2; c++) {
      const int j = j0 + wj * 32 + c * 16 + (lane & 15);
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int i = i0 + wi * 32 + a * 16 + Mfma<double>::row(lane, r);
        const long ci = g.msplit > 0 ? long(i % g.msplit) * g.ci + long(i / g.msplit) * g.ci2 : long(i) * g.ci;
        if (i < g.M && j < g.N) C[ci + long(j) * g.cj] = acc[a][c][r];
      }
    }
}

void launch_gemm_f64(const GemmDesc& g, hipStream_t st) {
  if (g.M <= 0 || g.N <= 0 || g.batch <= 0) return;
  const long tiles = long((g.N + kGT - 1) / kGT) * ((g.M + kGT - 1) / kGT);
  const dim3 grid(unsigned(tiles), 1u, unsigned(g.batch));
  hipLaunchKernelGGL(k_gemm_f64, grid, dim3(256), 0, st, g);
}

}  // namespace ssg
