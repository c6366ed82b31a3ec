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
// The Gamma^T fold of one solve lane: a_j from x (in y, lane stride nlanes) for
// variable v; the fold is Gamma^T x of SAtransform! (src/CubicBSpline.jl:1388-1393).
struct LaneFold {
  int Mdim, Min, rL, lk, rk;
  double aL, bL, aR, bR, xf0, xf1, xl1, xl2;
  const double* y;
  long nlanes, u;
  __device__ LaneFold(const DevTables& T, int v, const double* yy, long nl, long uu) : y(yy), nlanes(nl), u(uu) {
    const int* gi = T.gint + 8 * v;
    const double* gd = T.gdbl + 4 * v;
    Mdim = gi[0]; Min = gi[1]; rL = gi[2]; lk = gi[4]; rk = gi[5];
    aL = gd[0]; bL = gd[1]; aR = gd[2]; bR = gd[3];
    xf0 = y[u]; xf1 = y[nlanes + u];
    xl1 = y[long(Min - 1) * nlanes + u]; xl2 = y[long(Min - 2) * nlanes + u];
  }
  __device__ double operator()(int j) const {
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
  }
};

// (solve lane u) -> (line, variable, whether the lane stores): RL / RLZ lanes are
// (transform t, k, part) with line t * (2K + 1) + ks, ks = 0 (k = 0), 2k - 1
// (cosine), 2k (sine), both lanes of k = 0 solving line 0; RZ lanes are the lines
__device__ __forceinline__ void lane_line(const DevTables& T, bool ring, long u, long& line, int& v, bool& store) {
  if (ring) {
    const int part = int(u & 1);
    const long pu = u >> 1;
    const int k = int(pu % (T.K + 1));
    const int t = int(pu / (T.K + 1));
    v = t / T.nzb;
    line = long(t) * (2 * T.K + 1) + (k == 0 ? 0 : 2 * k - 1 + part);
    store = !(k == 0 && part == 1);
  } else {
    line = u;
    v = int(u / T.nzb);
    store = true;
  }
}

template <bool Ring>
__global__ void __launch_bounds__(256) k_si_pts(DevTables T, const double* __restrict__ y, int npt,
                                                const int* __restrict__ pxm, const double* __restrict__ pw,
                                                double* __restrict__ S) {
  const long nlanes = Ring ? 2 * long(T.ntr) * (T.K + 1) : T.nlines;
  const long u = long(blockIdx.x) * blockDim.x + threadIdx.x;
  if (u >= nlanes) return;
  long line;
  int v;
  bool store;
  lane_line(T, Ring, u, line, v, store);
  const long nline = Ring ? long(T.ntr) * (2 * T.K + 1) : T.nlines;
  const LaneFold a_at(T, v, y, nlanes, u);
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
// Evaluation at unstructured points (evaluate_unstructured,
// src/interpolation.jl:1735-1822; RL 1628-1668, RZ 1376-1402, RLZ 1670-1733).
// The reference runs SAtransform! on every line, SItransform at each point's
// radius for every line, the Fourier sum at its angle (RL, RLZ) and, per point,
// CAtransform! and the Chebyshev series at its height (RZ, RLZ).
//
// Here k_fold_lines writes the folded coefficients a = Gamma^T x of every line
// once, as A2[(j * nks + ks) * ntr + t] (j the B-spline index, ks the line of
// a wavenumber, t the (variable, z mode) transform), so that the coefficients a
// cell needs form one contiguous slab.  The host sorts the points by their
// first tap and cuts them into groups of at most 64 with the same tap; for a
// group, k_unstr_gemm evaluates the four tap weights of each point and forms C[n][t] = sum over (tap, ks) of W[n][(tap, ks)]
// A2[(xm + 1 + tap, ks)][t] on the FP64 matrix cores, with W[n][(tap, ks)] =
// w_tap(r_n) F_ks(lambda_n) generated in the kernel (F_0 = 1, F_{2k-1} =
// 2 cos(k lambda), F_{2k} = 2 sin(k lambda): the reference's +rs sin sum).
// k_unstr_cheb folds each point's z modes with CAtransform!'s matrix and sums
// the series at its height (_cheb_eval_single, angle clamped to [-1, 1]).
// ----------------------------------------------------------------------------
template <bool Ring>
__global__ void __launch_bounds__(256) k_fold_lines(DevTables T, const double* __restrict__ y,
                                                    double* __restrict__ A2) {
  const long nlanes = Ring ? 2 * long(T.ntr) * (T.K + 1) : T.nlines;
  const long u = long(blockIdx.x) * blockDim.x + threadIdx.x;
  if (u >= nlanes) return;
  long line;
  int v;
  bool store;
  lane_line(T, Ring, u, line, v, store);
  if (!store) return;
  const int nks = Ring ? 2 * T.K + 1 : 1;
  const long ntr = Ring ? long(T.ntr) : T.nlines;
  const long t = line / nks, ks = line % nks;
  const LaneFold a_at(T, v, y, nlanes, u);
  for (int j = 0; j < T.b_iDim; j++) A2[(long(j) * nks + ks) * ntr + t] = a_at(j);
}

void launch_fold_lines(const DevTables& T, bool ring, const double* y, double* A2, hipStream_t st) {
  const long nlanes = ring ? 2 * long(T.ntr) * (T.K + 1) : T.nlines;
  const dim3 grid(unsigned((nlanes + 255) / 256));
  if (ring) hipLaunchKernelGGL(k_fold_lines<true>, grid, dim3(256), 0, st, T, y, A2);
  else hipLaunchKernelGGL(k_fold_lines<false>, grid, dim3(256), 0, st, T, y, A2);
}

// Block: one group (<= 64 points, one first tap) x NT transforms.  NT = 64:
// four waves of 2 x 2 v_mfma_f64_16x16x4 tiles (32 x 32 each); NT = 16: four
// waves of one tile, stacked along the points.  The (tap, ks) sum runs over
// q = 0 .. nks - 1 in chunks of 16, with ks = q + 1 for q < nks - 1 (so the
// cosine and sine lines of a wavenumber share a chunk and one sincos) and the
// k = 0 line at q = nks - 1.  Per chunk the block computes F for 64 points x 8
// wavenumbers and stages the 4 x 16 x NT coefficients in LDS, then each wave
// runs 4 taps x 4 MFMA k-steps with A[i][q] = w_tap(point i) F[q][point i].
// Cubic B-spline of node m at x, value only (basis, src/CubicBSpline.jl:585-628)
__device__ __forceinline__ double bspline0(double xmin, double DX, double DXr, int m, double x) {
  const double delta = (x - (xmin + m * DX)) * DXr;
  const double t = fabs(delta);
  if (!(t < 2.0)) return 0.0;
  const double w = 2.0 - t;
  double v = w * w * w / 6.0;
  if (t < 1.0) {
    const double u = 1.0 - t;
    v -= 4.0 * u * u * u / 6.0;
  }
  return v;
}

template <int NT>
__global__ void __launch_bounds__(256) k_unstr_gemm(UnstrDesc d) {
  typedef Mfma<double>::v4 v4;
  constexpr int TI = NT == 64 ? 2 : 1;               // 16-row tiles per wave
  constexpr int TJ = NT == 64 ? 2 : 1;               // 16-column tiles per wave
  constexpr int WI = 64 / (16 * TI);                 // waves along the points
  __shared__ double Fs[16][64 + 1];                  // Fs[ks - ks0][point]
  __shared__ double Bs[4][16][NT + 1];               // Bs[tap][ks - ks0][t - t0]
  const int g = blockIdx.x, t0 = blockIdx.y * NT;
  const int gs = d.gstart[g], gn = d.gcount[g], xm = d.gxm[g];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int wi = w % WI, wj = w / WI;
  double wt[TI][4];
#pragma unroll
  for (int a = 0; a < TI; a++) {
    const int i = wi * 16 * TI + a * 16 + (lane & 15);
    const double x = i < gn ? d.x[gs + i] : 0.0;
#pragma unroll
    for (int tap = 0; tap < 4; tap++) {
      const int m = xm + tap;                        // taps outside -1 .. nc + 1 carry no weight
      wt[a][tap] = (i < gn && m >= -1 && m <= d.nc + 1) ? bspline0(d.xmin, d.DX, d.DXr, m, x) : 0.0;
    }
  }
  v4 acc[TI][TJ];
#pragma unroll
  for (int a = 0; a < TI; a++)
#pragma unroll
    for (int c = 0; c < TJ; c++) acc[a][c] = v4{0, 0, 0, 0};
  for (int ks0 = 0; ks0 < d.nks; ks0 += 16) {
    __syncthreads();                                 // the previous chunk's products are done
#pragma unroll
    for (int s = 0; s < 2; s++) {
      const int e = tid + 256 * s, pt = e & 63, kp = e >> 6, q = ks0 + 2 * kp;   // q even: cosine line of k
      double fc = 0.0, fs = 0.0;
      if (pt < gn && q < d.nks) {
        if (q == d.nks - 1) fc = 1.0;                 // the k = 0 line
        else {
          const double kd = double((q >> 1) + 1);
          double sn, cs;
          sincos(kd * d.lam[gs + pt], &sn, &cs);
          fc = 2.0 * cs;
          fs = 2.0 * sn;
        }
      }
      Fs[2 * kp][pt] = fc;
      Fs[2 * kp + 1][pt] = fs;
    }
#pragma unroll
    for (int s = 0; s < (4 * 16 * NT) / 256; s++) {
      const int e = tid + 256 * s, col = e % NT, rest = e / NT, kk = rest & 15, tap = rest >> 4;
      const int q = ks0 + kk, ks = q == d.nks - 1 ? 0 : q + 1, t = t0 + col;
      int j = xm + 1 + tap;                          // taps outside the basis carry zero weight
      j = j < 0 ? 0 : (j >= d.b_iDim ? d.b_iDim - 1 : j);
      Bs[tap][kk][col] = (q < d.nks && t < d.ntr) ? d.A2[(long(j) * d.nks + ks) * d.ntr + t] : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int tap = 0; tap < 4; tap++)
#pragma unroll
      for (int kq = 0; kq < 4; kq++) {
        const int kr = kq * 4 + (lane >> 4);
        double av[TI], bv[TJ];
#pragma unroll
        for (int a = 0; a < TI; a++) av[a] = wt[a][tap] * Fs[kr][wi * 16 * TI + a * 16 + (lane & 15)];
#pragma unroll
        for (int c = 0; c < TJ; c++) bv[c] = Bs[tap][kr][wj * 16 * TJ + c * 16 + (lane & 15)];
#pragma unroll
        for (int a = 0; a < TI; a++)
#pragma unroll
          for (int c = 0; c < TJ; c++) acc[a][c] = Mfma<double>::mma(av[a], bv[c], acc[a][c]);
      }
  }
#pragma unroll
  for (int a = 0; a < TI; a++)
#pragma unroll
    for (int c = 0; c < TJ; c++) {
      const int t = t0 + wj * 16 * TJ + c * 16 + (lane & 15);
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int i = wi * 16 * TI + a * 16 + Mfma<double>::row(lane, r);
        if (i < gn && t < d.ntr) {
          const long n = gs + i;
          if (d.C) d.C[(n - d.cbase) * d.ntr + t] = acc[a][c][r];
          else d.out[long(t) * d.nall + d.perm[n]] = acc[a][c][r];
        }
      }
    }
}

void launch_unstr_gemm(const UnstrDesc& d, int ngroups, hipStream_t st) {
  if (ngroups <= 0) return;
  if (d.ntr <= 16) {
    hipLaunchKernelGGL(k_unstr_gemm<16>, dim3(unsigned(ngroups), unsigned((d.ntr + 15) / 16)), dim3(256), 0, st, d);
  } else {
    hipLaunchKernelGGL(k_unstr_gemm<64>, dim3(unsigned(ngroups), unsigned((d.ntr + 63) / 64)), dim3(256), 0, st, d);
  }
}

// One thread per (point, variable), a block's threads share the variable:
// a = CA b from the point's z modes (held in registers, NZB >= nzb), then
// a_0 + 2 sum_{k=1}^{N-2} a_k cos(k t) + a_{N-1} cos((N-1) t), t = acos(xi)
// with xi clamped to [-1, 1] (_cheb_eval_single, src/interpolation.jl:1271-1285).
template <int NZB>
__global__ void __launch_bounds__(256) k_unstr_cheb(UnstrCheb d) {
  const long n = long(blockIdx.x) * blockDim.x + threadIdx.x;
  const int v = blockIdx.y;
  if (n >= d.npt) return;
  const double scale = -0.5 * (d.zmax - d.zmin), offset = 0.5 * (d.zmin + d.zmax);
  double xi = (d.z[d.cbase + n] - offset) / scale;
  xi = xi < -1.0 ? -1.0 : (xi > 1.0 ? 1.0 : xi);
  const double t = acos(xi);
  const double* bn = d.C + n * d.ntr + long(v) * d.nzb;
  double b[NZB];
#pragma unroll
  for (int zb = 0; zb < NZB; zb++) b[zb] = zb < d.nzb ? bn[zb] : 0.0;
  const double* ca = d.ca + long(v) * d.kDim * d.nzb;
  double val = 0.0;
  for (int k = 0; k < d.kDim; k++) {
    double ak = 0.0;
#pragma unroll
    for (int zb = 0; zb < NZB; zb++)
      if (zb < d.nzb) ak += ca[long(k) * d.nzb + zb] * b[zb];
    if (k == 0) val = ak;
    else if (k < d.kDim - 1) val += 2.0 * ak * cos(double(k) * t);
    else val += ak * cos(double(k) * t);
  }
  d.out[long(v) * d.nall + d.perm[d.cbase + n]] = val;
}

void launch_unstr_cheb(const UnstrCheb& d, hipStream_t st) {
  if (d.npt <= 0) return;
  const dim3 grid(unsigned((d.npt + 255) / 256), unsigned(d.nv));
  if (d.nzb <= 8) hipLaunchKernelGGL(k_unstr_cheb<8>, grid, dim3(256), 0, st, d);
  else if (d.nzb <= 16) hipLaunchKernelGGL(k_unstr_cheb<16>, grid, dim3(256), 0, st, d);
  else if (d.nzb <= 24) hipLaunchKernelGGL(k_unstr_cheb<24>, grid, dim3(256), 0, st, d);
  else if (d.nzb <= 32) hipLaunchKernelGGL(k_unstr_cheb<32>, grid, dim3(256), 0, st, d);
  else hipLaunchKernelGGL(k_unstr_cheb<48>, grid, dim3(256), 0, st, d);
}

// out[v * nall + idx[i]] = fill for the points outside the domain
__global__ void __launch_bounds__(256) k_unstr_fill(const long* __restrict__ idx, long n, int nv, long nall,
                                                    double fill, double* __restrict__ out) {
  const long u = long(blockIdx.x) * blockDim.x + threadIdx.x;
  if (u >= n * nv) return;
  out[(u / n) * nall + idx[u % n]] = fill;
}

void launch_unstr_fill(const long* idx, long n, int nv, long nall, double fill, double* out, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_unstr_fill, dim3(unsigned((n * nv + 255) / 256)), dim3(256), 0, st, idx, n, nv, nall, fill,
                     out);
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
