// ssg_setup.hpp -- host FP64 set-up of one transform plan: every table the
// kernels read, derived from an ssg_desc.  Pure C++ (no HIP), so it is also
// built and unit-checked on machines without a GPU.
#pragma once

#include <string>
#include <vector>

#include "springsteel_hip.h"

namespace ssg {

// Gamma fold of one spline (src/CubicBSpline.jl:321-500): out = G b keeps
// b[i + rankL] plus at most two corrections per side; G^T is its adjoint.
struct GammaFold {
  int Mdim = 0, Min = 0, rankL = 0, rankR = 0, lk = 0, rk = 0;
  double aL = 0, bL = 0, aR = 0, bR = 0;
};

struct SplineVar {
  GammaFold g;
  bool r3x = false;          // an R3X side: SAtransform! takes the background-coefficient branch
  std::vector<double> pq;    // Mdim x 4 band of the unfolded P + Q, pq[i*4 + o] = (P+Q)(i, i+o)
  std::vector<double> L;     // 4 x Min band of the LL^T factor of G (P+Q) G^T, column-major
  std::vector<double> Linv;  // 1 / L[0, i]
};

struct ChebVar {
  std::vector<double> fwd;   // b_kDim x kDim: filter * DCT-I / (2(N-1)), row-major [zb][z]
  std::vector<double> inv0;  // kDim x b_kDim: DCT-I * (bfill + gamma fold),  row-major [z][zb]
  std::vector<double> inv1;  // d/dz of inv0 (CIxcoefficients applied once)
  std::vector<double> inv2;  // d2/dz2 (twice)
  std::vector<double> ca;    // kDim x b_kDim: a = (bfill + gamma fold) b, row-major [k][zb] (CAtransform!)
};

struct Setup {
  ssg_desc d{};              // copy of the scalar fields
  int geometry = 0, nvars = 0, nc = 0, mubar = 0;   // SL / SLZ run as RL / RLZ (spherical set)
  bool spherical = false;    // rings of 4 round(sin(theta) max_ri) points
  bool any_r3x = false;      // some variable has an R3X side
  int iDim = 0, b_iDim = 0, K = 0, kDim = 0, b_kDim = 0, jDim = 0, b_jDim = 0;
  int nzb = 1;               // Chebyshev modes carried by the line stages (1 for RL)
  int nslots = 5;
  long npts = 0, nspec = 0;
  int maxN = 0, maxMin = 0;
  double DX = 0, DXrecip = 0;

  // per-cell B-spline tables: sbw[(mc*mubar + mu)*4 + t] = DX w_mu phi_{mc-1+t}(x_{mc,mu});
  // siw[d][(mc*mubar + mu)*4 + t] = phi^{(d)}_{mc-1+t}(x_{mc,mu})
  std::vector<double> sbw;
  std::vector<double> siw[3];
  std::vector<double> mish;  // radial points

  std::vector<SplineVar> spl;   // [nvars]
  std::vector<ChebVar> cheb;    // [nvars] (RZ, RLZ)

  // rings (RL, RLZ): r = 0..iDim-1
  std::vector<int> ringN, ringRi, ringOff, coefOff, phaseOff, twOff;  // [iDim (+1 for offsets)]
  std::vector<int> kmax;        // [nvars * iDim] effective wavenumber cap per (var, ring)
  std::vector<double> phaseC, phaseS;   // cos(k ymin), sin(k ymin), k = 1..ri per ring
  std::vector<double> twid;     // cos(2 pi m / N), m = 0..N-1 per ring

  // chirp-z (Bluestein) ring synthesis: ring r evaluates its 2 ri + 1 wavenumbers at
  // its N points as a circular convolution of length M = 2^fftL[r] >= N + 2 ri
  std::vector<int> fftL;        // [iDim] log2 M (0: the ring needs M above kFftMaxM)
  std::vector<int> fftHOff;     // [iDim + 1] complex offset of the ring's H
  std::vector<double> fftChirp; // (re, im) of w(n) = exp(i pi n^2 / N), n < N, at ringOff[r]
  std::vector<double> fftH;     // (re, im) of FFT_M(conj chirp arranged circularly) / M
  std::vector<double> fftTw;    // (re, im) of exp(-2 pi i n / kFftMaxM), n < kFftMaxM
};

// largest chirp-z transform of the ring synthesis
static constexpr int kFftMaxLog = 12;
static constexpr int kFftMaxM = 1 << kFftMaxLog;
// log2 of the chirp-z length of a ring of N points and ri wavenumbers (at least 16)
inline int fft_log_for(int N, int ri) {
  int L = 4;
  while ((1 << L) < N + 2 * ri) L++;
  return L;
}

// Returns SSG_OK or an SSG_E* code; err receives the message.
int build_setup(const ssg_desc& desc, Setup& s, std::string& err);

// B-spline basis (value, d/dx, d2/dx2, d3/dx3) of node m at x, as defined by
// src/CubicBSpline.jl:585-628, written as the piecewise cubic in |delta|.
double bspline(double xmin, double DX, double DXrecip, int m, double x, int derivative);

}  // namespace ssg
