// ssg_setup.cpp -- host FP64 set-up of a transform plan.
//
// Everything here runs once per plan on the CPU.  The quantities are the ones
// createGrid derives in the reference (src/factory.jl:461-502 RZ, 646-685 RL,
// 688-740 RLZ); each block names the reference lines whose definition it
// implements.  Matrices that the reference applies as a chain of steps per
// call (Chebyshev filter + DCT-I, Gamma fold + DCT-I + derivative recurrence)
// are composed here once, in long double, into dense matrices.
#include "ssg_setup.hpp"

#include <cmath>
#include <cstdio>

namespace ssg {

static const long double PI_L = 3.141592653589793238462643383279502884L;

double bspline(double xmin, double DX, double DXrecip, int m, double x, int derivative) {
  // Cubic B-spline of node m (support |delta| < 2), src/CubicBSpline.jl:585-628.
  const double delta = (x - (xmin + m * DX)) * DXrecip;
  const double t = std::fabs(delta);
  if (!(t < 2.0)) return 0.0;
  const double w = 2.0 - t;       // distance to the support edge
  const double sgn = delta > 0 ? -1.0 : 1.0;
  switch (derivative) {
    case 0: {
      double v = w * w * w / 6.0;
      if (t < 1.0) { const double u = 1.0 - t; v -= 4.0 * u * u * u / 6.0; }
      return v;
    }
    case 1: {
      double v = w * w / 6.0;
      if (t < 1.0) { const double u = 1.0 - t; v -= 4.0 * u * u / 6.0; }
      return v * sgn * 3.0 * DXrecip;
    }
    case 2: {
      double v = w;
      if (t < 1.0) v -= 4.0 * (1.0 - t);
      return v * DXrecip * DXrecip;
    }
    default: {
      const double v = t > 1.0 ? 1.0 : (t < 1.0 ? -3.0 : 0.0);
      return v * sgn * DXrecip * DXrecip * DXrecip;
    }
  }
}

// Gauss-Legendre nodes/weights on [-1,1] mapped to [0,1] with weights summing
// to 1, or the equispaced midpoint rule (src/CubicBSpline.jl:128-200).
static void quadrature(int mb, int q, std::vector<double>& pts, std::vector<double>& wts) {
  pts.assign(mb, 0.0);
  wts.assign(mb, 0.0);
  if (q == 0) {
    std::vector<double> x, w;
    switch (mb) {
      case 1: x = {0.0}; w = {2.0}; break;
      case 2: x = {-1.0 / std::sqrt(3.0), 1.0 / std::sqrt(3.0)}; w = {1.0, 1.0}; break;
      case 3: x = {-std::sqrt(0.6), 0.0, std::sqrt(0.6)}; w = {5.0 / 9.0, 8.0 / 9.0, 5.0 / 9.0}; break;
      case 4: {
        const double a = std::sqrt(3.0 / 7.0 + 2.0 / 7.0 * std::sqrt(6.0 / 5.0));
        const double b = std::sqrt(3.0 / 7.0 - 2.0 / 7.0 * std::sqrt(6.0 / 5.0));
        const double wa = (18.0 - std::sqrt(30.0)) / 36.0, wb = (18.0 + std::sqrt(30.0)) / 36.0;
        x = {-a, -b, b, a}; w = {wa, wb, wb, wa};
        break;
      }
      default: {
        const double a = std::sqrt(5.0 + 2.0 * std::sqrt(10.0 / 7.0)) / 3.0;
        const double b = std::sqrt(5.0 - 2.0 * std::sqrt(10.0 / 7.0)) / 3.0;
        const double wa = (322.0 - 13.0 * std::sqrt(70.0)) / 900.0, wb = (322.0 + 13.0 * std::sqrt(70.0)) / 900.0;
        x = {-a, -b, 0.0, b, a}; w = {wa, wb, 128.0 / 225.0, wb, wa};
        break;
      }
    }
    for (int i = 0; i < mb; i++) { pts[i] = (x[i] + 1.0) / 2.0; wts[i] = w[i] / 2.0; }
  } else {
    for (int k = 1; k <= mb; k++) { pts[k - 1] = double(2 * k - 1) / double(2 * mb); wts[k - 1] = 1.0 / mb; }
  }
}

// Sparse rows of Gamma (Min x Mdim): identity shifted by rankL plus the
// boundary corrections of src/CubicBSpline.jl:321-500 (GammaBC getindex).
struct GEntry { int col; double val; };
static std::vector<std::vector<GEntry>> gamma_rows(const GammaFold& g) {
  std::vector<std::vector<GEntry>> rows(g.Min);
  for (int i = 0; i < g.Min; i++) rows[i].push_back({i + g.rankL, 1.0});
  if (g.lk == SSG_BC_ALPHA1) { rows[0].push_back({0, g.aL}); if (g.Min > 1) rows[1].push_back({0, g.bL}); }
  else if (g.lk == SSG_BC_ALPHA2) { rows[0].push_back({0, g.aL}); rows[0].push_back({1, g.bL}); }
  if (g.rk == SSG_BC_ALPHA1) { rows[g.Min - 1].push_back({g.Mdim - 1, g.aR}); if (g.Min > 1) rows[g.Min - 2].push_back({g.Mdim - 1, g.bR}); }
  else if (g.rk == SSG_BC_ALPHA2) { rows[g.Min - 1].push_back({g.Mdim - 1, g.aR}); rows[g.Min - 1].push_back({g.Mdim - 2, g.bR}); }
  return rows;
}

static int rank_of(int kind) {
  switch (kind) {
    case SSG_BC_R0: return 0;
    case SSG_BC_ALPHA1: return 1;
    case SSG_BC_ALPHA2: return 2;
    default: return 3;
  }
}

// P + Q of the Ooyama (2002) variational problem folded by Gamma and factored
// as a half-bandwidth-3 LL^T (src/CubicBSpline.jl:733-783, src/banded_cholesky.jl:166-197).
static int build_spline_var(const Setup& s, const std::vector<double>& qp, const std::vector<double>& qw, double l_q,
                            SplineVar& out, std::string& err) {
  const int nc = s.nc, mb = s.mubar, Mdim = nc + 3;
  const double DX = s.DX, xmin = s.d.iMin;
  const double eps_q = std::pow((l_q * DX) / (2 * M_PI), 6);
  // band storage A[i][o] = A(i, i + o), o = 0..3, separately for P and Q
  std::vector<double> P(size_t(Mdim) * 4, 0.0), Q(size_t(Mdim) * 4, 0.0);
  for (int mc = 0; mc < nc; mc++)
    for (int mu = 0; mu < mb; mu++) {
      const double x = xmin + mc * DX + qp[mu] * DX;
      double phi[4], d3[4];
      for (int t = 0; t < 4; t++) {
        phi[t] = bspline(xmin, DX, s.DXrecip, mc - 1 + t, x, 0);
        d3[t] = bspline(xmin, DX, s.DXrecip, mc - 1 + t, x, 3);
      }
      const double wpx = DX * qw[mu];
      for (int a = 0; a < 4; a++)
        for (int b = a; b < 4; b++) {
          P[size_t(mc + a) * 4 + (b - a)] += wpx * phi[a] * phi[b];
          Q[size_t(mc + a) * 4 + (b - a)] += wpx * eps_q * d3[a] * d3[b];
        }
    }
  // the unfolded P + Q (SAtransform!'s pq, src/CubicBSpline.jl:772): the R3X branch subtracts pq ahat
  out.pq.assign(size_t(Mdim) * 4, 0.0);
  for (size_t e = 0; e < out.pq.size(); e++) out.pq[e] = P[e] + Q[e];
  auto band = [&](const std::vector<double>& A, int i, int j) -> double {
    if (i > j) std::swap(i, j);
    if (j - i > 3 || i < 0 || j >= Mdim) return 0.0;
    return A[size_t(i) * 4 + (j - i)];
  };
  const GammaFold& g = out.g;
  const int Min = g.Min;
  auto rows = gamma_rows(g);
  // PQopen = G P G^T + G Q G^T, band part (the fold keeps bandwidth 3)
  std::vector<double> Lb(size_t(Min) * 4, 0.0);  // column-major 4 x Min: Lb[j*4 + (i-j)] = A(i, j), i >= j
  for (int j = 0; j < Min; j++)
    for (int i = j; i < Min && i <= j + 3; i++) {
      double fp = 0.0, fq = 0.0;
      for (const GEntry& ei : rows[i])
        for (const GEntry& ej : rows[j]) {
          fp += ei.val * band(P, ei.col, ej.col) * ej.val;
          fq += ei.val * band(Q, ei.col, ej.col) * ej.val;
        }
      Lb[size_t(j) * 4 + (i - j)] = fp + fq;
    }
  // LL^T in place, column by column
  for (int j = 0; j < Min; j++) {
    double dsum = Lb[size_t(j) * 4];
    for (int k = std::max(0, j - 3); k < j; k++) { const double ljk = Lb[size_t(k) * 4 + (j - k)]; dsum -= ljk * ljk; }
    if (!(dsum > 0.0)) {
      char m[160];
      std::snprintf(m, sizeof m, "PosDefException: (P+Q) not positive definite at row %d", j + 1);
      err = m;
      return SSG_EPOSDEF;
    }
    const double ljj = std::sqrt(dsum);
    Lb[size_t(j) * 4] = ljj;
    for (int i = j + 1; i < Min && i <= j + 3; i++) {
      double v = Lb[size_t(j) * 4 + (i - j)];
      for (int k = std::max(0, i - 3); k < j; k++) v -= Lb[size_t(k) * 4 + (i - k)] * Lb[size_t(k) * 4 + (j - k)];
      Lb[size_t(j) * 4 + (i - j)] = v / ljj;
    }
  }
  out.L = Lb;
  out.Linv.resize(Min);
  for (int i = 0; i < Min; i++) out.Linv[i] = 1.0 / Lb[size_t(i) * 4];
  return SSG_OK;
}

// Chebyshev column matrices (src/Chebyshev.jl:364-378 filter, 417-426 CB,
// 473-499 CA, 530-535 CI, 636-722 CIx/CIxx, 991-1177 calcGammaBC).
static int build_cheb_var(const Setup& s, int bcb, int bct, ChebVar& cv, std::string& err) {
  const int N = s.kDim, B = s.b_kDim;
  typedef long double ld;
  // DCT-I (REDFT00, unnormalised): D[k][j]
  std::vector<ld> D(size_t(N) * N);
  for (int k = 0; k < N; k++)
    for (int j = 0; j < N; j++) {
      ld v;
      if (j == 0) v = 1.0L;
      else if (j == N - 1) v = (k & 1) ? -1.0L : 1.0L;
      else {
        const long e = (long(j) * k) % (2L * (N - 1));
        v = 2.0L * std::cos(PI_L * ld(e) / ld(N - 1));
      }
      D[size_t(k) * N + j] = v;
    }
  // forward: b = F * (D u) / (2(N-1)), F = truncation rows or exp filter
  cv.fwd.assign(size_t(B) * N, 0.0);
  for (int zb = 0; zb < B; zb++) {
    ld f = 1.0L;
    if (B >= N) f = std::exp(-36.0L * std::pow(ld(zb + 1) / ld(N), 36));
    for (int z = 0; z < N; z++) cv.fwd[size_t(zb) * N + z] = double(f * D[size_t(zb) * N + z] / ld(2 * (N - 1)));
  }
  // Gamma: a = bfill + gamma^T bfill (matrix) or bfill .+ (gamma . bfill) (vector)
  const int R0 = SSG_CBC_R0, T0 = SSG_CBC_R1T0, T1 = SSG_CBC_R1T1;
  if (bcb > T1 || bct > T1 || bcb < 0 || bct < 0) {
    err = "DomainError: Unrecognized Chebyshev BC dict keys";
    return SSG_EDOMAIN;
  }
  std::vector<ld> gm(size_t(N) * N, 0.0L);   // gm[i][j] = gamma(i, j) (matrix form)
  std::vector<ld> gv(N, 0.0L);               // vector form
  bool is_vec = false;
  auto cfac = [&](int i) -> ld { return (i == 0 || i == N - 1) ? 2.0L : 1.0L; };   // c[1] = c[N] = 2
  auto neumann_bottom = [&](std::vector<ld>& g) {  // Wang et al. (1993) global method, bottom
    ld sf = 0.0L;
    for (int i = 0; i < N; i++) sf += -ld(i) * ld(i) / (cfac(i) * (N - 1));
    for (int i = 0; i < N; i++)
      for (int j = 0; j < N; j++) g[size_t(i) * N + j] += ld(i) * ld(i) / (sf * cfac(j) * (N - 1));
  };
  auto neumann_top = [&](std::vector<ld>& g) {
    ld sf = 0.0L;
    for (int i = 0; i < N; i++) sf += ld(i) * ld(i) / (cfac(i) * (N - 1));   // -(n^2)(-1)^n(-1)^(n+1) = n^2
    for (int i = 0; i < N; i++)
      for (int j = 0; j < N; j++) {
        const ld sj = ((j + 2) & 1) ? -1.0L : 1.0L;   // (-1)^(j+1) with 1-based j -> (-1)^(j0+2)
        const ld sn = ((i + 1) & 1) ? -1.0L : 1.0L;   // (-1)^(n+1)
        g[size_t(i) * N + j] += sj * sn * ld(i) * ld(i) / (sf * cfac(j) * (N - 1));
      }
  };
  auto dirichlet_top = [&](std::vector<ld>& g) {
    for (int i = 0; i < N; i++)
      for (int j = 0; j < N; j++) {
        ld v = -1.0L * (((i + j) & 1) ? -1.0L : 1.0L) / ld(N - 1);
        if (i == 0 || i == N - 1) v *= 0.5L;
        g[size_t(i) * N + j] += v;
      }
  };
  auto dirichlet_bottom_vec = [&](std::vector<ld>& v) {
    for (int i = 0; i < N; i++) v[i] = ((i == 0 || i == N - 1) ? 1.0L : 2.0L) * (-0.5L / ld(N - 1));
  };
  if (bcb == R0 && bct == R0) {
    is_vec = true;
  } else if (bcb == T0 && bct == R0) {
    is_vec = true; dirichlet_bottom_vec(gv);
  } else if (bcb == T1 && bct == R0) {
    neumann_bottom(gm);
  } else if (bcb == R0 && bct == T0) {
    dirichlet_top(gm);
  } else if (bcb == R0 && bct == T1) {
    neumann_top(gm);
  } else if (bcb == T0 && bct == T0) {
    dirichlet_top(gm);
    std::vector<ld> v(N); dirichlet_bottom_vec(v);
    for (int i = 0; i < N; i++) for (int j = 0; j < N; j++) gm[size_t(i) * N + j] += v[i];
  } else if (bcb == T1 && bct == T1) {
    neumann_bottom(gm); neumann_top(gm);
  } else if (bcb == T0 && bct == T1) {
    neumann_top(gm);
    std::vector<ld> v(N); dirichlet_bottom_vec(v);
    for (int i = 0; i < N; i++) for (int j = 0; j < N; j++) gm[size_t(i) * N + j] += v[i];
  } else {  // T1 bottom, T0 top
    neumann_bottom(gm); dirichlet_top(gm);
  }
  // A (N x B): a = A b
  std::vector<ld> A(size_t(N) * B, 0.0L);
  for (int i = 0; i < N; i++)
    for (int j = 0; j < B; j++) {
      ld v = (i == j) ? 1.0L : 0.0L;
      v += is_vec ? gv[j] : gm[size_t(j) * N + i];
      A[size_t(i) * B + j] = v;
    }
  // derivative-coefficient map (CIxcoefficients): ax[N-1] = 0, ax[N-2] = 2(N-1) a[N-1],
  // ax[j] = 2(j+1) a[j+1] + ax[j+2]; then ax /= -(zmax - zmin)/2
  const ld den = -0.5L * (ld(s.d.kMax) - ld(s.d.kMin));
  auto dz = [&](const std::vector<ld>& in, std::vector<ld>& o) {   // in, o: N x B
    o.assign(size_t(N) * B, 0.0L);
    for (int c = 0; c < B; c++) {
      std::vector<ld> ax(N, 0.0L);
      if (N >= 2) ax[N - 2] = 2.0L * (N - 1) * in[size_t(N - 1) * B + c];
      for (int j = N - 3; j >= 0; j--) ax[j] = 2.0L * (j + 1) * in[size_t(j + 1) * B + c] + ax[j + 2];
      for (int j = 0; j < N; j++) o[size_t(j) * B + c] = ax[j] / den;
    }
  };
  auto apply_D = [&](const std::vector<ld>& in, std::vector<double>& o) {
    o.assign(size_t(N) * B, 0.0);
    for (int z = 0; z < N; z++)
      for (int c = 0; c < B; c++) {
        ld acc = 0.0L;
        for (int k = 0; k < N; k++) acc += D[size_t(z) * N + k] * in[size_t(k) * B + c];
        o[size_t(z) * B + c] = double(acc);
      }
  };
  cv.ca.assign(A.begin(), A.end());
  std::vector<ld> A1, A2;
  dz(A, A1);
  dz(A1, A2);
  apply_D(A, cv.inv0);
  apply_D(A1, cv.inv1);
  apply_D(A2, cv.inv2);
  return SSG_OK;
}

// In-place radix-2 FFT of length M (a power of two), sign -1, in long double.
static void fft_ld(std::vector<long double>& re, std::vector<long double>& im) {
  const size_t M = re.size();
  for (size_t i = 1, j = 0; i < M; i++) {
    size_t bit = M >> 1;
    for (; j & bit; bit >>= 1) j ^= bit;
    j ^= bit;
    if (i < j) { std::swap(re[i], re[j]); std::swap(im[i], im[j]); }
  }
  std::vector<long double> c(M / 2), sn(M / 2);
  for (size_t k = 0; k < M / 2; k++) {
    c[k] = std::cos(2.0L * PI_L * (long double)k / (long double)M);
    sn[k] = -std::sin(2.0L * PI_L * (long double)k / (long double)M);
  }
  for (size_t len = 2; len <= M; len <<= 1) {
    const size_t half = len >> 1, step = M / len;
    for (size_t b = 0; b < M; b += len)
      for (size_t k = 0; k < half; k++) {
        const long double wr = c[k * step], wi = sn[k * step];
        const size_t i = b + k, j = i + half;
        const long double xr = re[j] * wr - im[j] * wi, xi = re[j] * wi + im[j] * wr;
        re[j] = re[i] - xr; im[j] = im[i] - xi;
        re[i] += xr; im[i] += xi;
      }
  }
}

// Tables of the chirp-z ring synthesis.  With W = exp(2 pi i / N) and w(n) =
// W^(n^2 / 2) = exp(i pi n^2 / N), kl = (k^2 + l^2 - (l - k)^2) / 2 turns
//   y(l) = sum_{k=-K}^{K} a_k W^(k l)
// into y(l) = w(l) sum_k [a_k w(k)] conj(w(l - k)), a linear convolution whose
// arguments l - k span [-K, N + K - 1]; a circular one of length M >= N + 2K
// holds it exactly.  n^2 is reduced mod 2N before the angle is formed, and H
// comes from a long double FFT, so the tables are correctly rounded doubles.
static void build_fft_tables(Setup& s) {
  const int I = s.iDim;
  s.fftL.assign(I, 0);
  s.fftHOff.assign(I + 1, 0);
  s.fftChirp.assign(2 * size_t(s.ringOff[I]), 0.0);
  auto wang = [](long n, int N) { return PI_L * (long double)((n * n) % (2L * N)) / (long double)N; };
  for (int r = 0; r < I; r++) {
    const int N = s.ringN[r], K = s.ringRi[r];
    for (int n = 0; n < N; n++) {
      const long double a = wang(n, N);
      s.fftChirp[2 * (size_t(s.ringOff[r]) + n)] = double(std::cos(a));
      s.fftChirp[2 * (size_t(s.ringOff[r]) + n) + 1] = double(std::sin(a));
    }
    const int L = fft_log_for(N, K);
    // the synthesis's first pass takes its 2K + 1 inputs from the first 6 of 16 rows
    // of 2^L / 16 entries (true whenever N >= 4K, as on every ring here)
    s.fftL[r] = L <= kFftMaxLog && 2 * K + 1 <= 6 * ((1 << L) / 16) ? L : 0;
    s.fftHOff[r + 1] = s.fftHOff[r] + (s.fftL[r] ? (1 << L) : 0);
  }
  s.fftH.assign(2 * size_t(s.fftHOff[I]), 0.0);
  for (int r = 0; r < I; r++) {
    if (!s.fftL[r]) continue;
    const int N = s.ringN[r], K = s.ringRi[r], M = 1 << s.fftL[r];
    std::vector<long double> re(M, 0.0L), im(M, 0.0L);
    for (int n = -K; n <= N + K - 1; n++) {
      const long double a = wang(std::abs(n), N);
      const int j = n >= 0 ? n : n + M;
      re[j] = std::cos(a);
      im[j] = -std::sin(a);
    }
    fft_ld(re, im);
    for (int j = 0; j < M; j++) {
      s.fftH[2 * (size_t(s.fftHOff[r]) + j)] = double(re[j] / (long double)M);
      s.fftH[2 * (size_t(s.fftHOff[r]) + j) + 1] = double(im[j] / (long double)M);
    }
  }
  s.fftTw.assign(2 * size_t(kFftMaxM), 0.0);
  for (int n = 0; n < kFftMaxM; n++) {
    s.fftTw[2 * n] = double(std::cos(2.0L * PI_L * (long double)n / (long double)kFftMaxM));
    s.fftTw[2 * n + 1] = double(-std::sin(2.0L * PI_L * (long double)n / (long double)kFftMaxM));
  }
}

int build_setup(const ssg_desc& d, Setup& s, std::string& err) {
  s.d = d;
  if (d.geometry < SSG_RL || d.geometry > SSG_SLZ) { err = "DomainError: unknown geometry"; return SSG_EDOMAIN; }
  // SL / SLZ are the RL / RLZ transforms on spherical rings
  // (src/transforms_spherical.jl: "identical in structure ... with sin(theta)-dependent ring sizes")
  s.spherical = d.geometry == SSG_SL || d.geometry == SSG_SLZ;
  s.geometry = d.geometry == SSG_SL ? SSG_RL : d.geometry == SSG_SLZ ? SSG_RLZ : d.geometry;
  s.d.geometry = s.geometry;
  if (d.num_cells < 1 || d.nvars < 1 || !(d.iMax > d.iMin)) { err = "ArgumentError: num_cells, nvars >= 1 and iMax > iMin required"; return SSG_EINVAL; }
  if (d.mubar < 1 || (d.quadrature == 0 && d.mubar > 5) || d.quadrature < 0 || d.quadrature > 1) {
    err = "ArgumentError: mubar must be 1..5 for :gauss (>= 1 for :regular)"; return SSG_EINVAL;
  }
  if (d.patchOffsetL < 0) { err = "ArgumentError: patchOffsetL must be >= 0"; return SSG_EINVAL; }
  if (d.precision != SSG_FP64 && d.precision != SSG_FP32_REFINE) { err = "ArgumentError: precision"; return SSG_EINVAL; }
  if (d.filter_flags) { err = "ENOTSUP: filter weights are passed with ssg_plan_set_filter, not in the descriptor"; return SSG_ENOTSUP; }
  if (d.wavenumber_ahat) { err = "ENOTSUP: multipatch per-wavenumber ahat is not run by this library"; return SSG_ENOTSUP; }
  if (!d.l_q || !d.bcl_kind || !d.bcr_kind || !d.bcl_ab || !d.bcr_ab) { err = "ArgumentError: per-variable arrays missing"; return SSG_EINVAL; }
  s.nvars = d.nvars; s.nc = d.num_cells; s.mubar = d.mubar;
  s.iDim = d.num_cells * d.mubar;
  s.b_iDim = d.num_cells + 3;
  s.K = s.iDim + d.patchOffsetL;
  s.DX = (d.iMax - d.iMin) / d.num_cells;
  s.DXrecip = 1.0 / s.DX;
  const bool cheb = s.geometry != SSG_RL;
  if (cheb) {
    if (d.kDim < 2) { err = "ArgumentError: zDim must be >= 2 for a DCT-I"; return SSG_EINVAL; }
    if (d.b_kDim < 1 || d.b_kDim > d.kDim) { err = "ArgumentError: 1 <= b_kDim <= kDim required"; return SSG_EINVAL; }
    if (s.geometry == SSG_RLZ && d.b_kDim < 3) { err = "BoundsError: RLZ reads ibasis rows 2 and 3, b_kDim >= 3 required"; return SSG_EBOUNDS; }
    if (!d.bcb_kind || !d.bct_kind) { err = "ArgumentError: Chebyshev BC arrays missing"; return SSG_EINVAL; }
    if (!(d.kMax != d.kMin)) { err = "ArgumentError: kMax == kMin"; return SSG_EINVAL; }
    s.kDim = d.kDim; s.b_kDim = d.b_kDim;
  }
  for (int v = 0; v < d.nvars; v++) {
    const int kinds[2] = {d.bcl_kind[v], d.bcr_kind[v]};
    for (int kk : kinds) {
      if (kk == SSG_BC_PERIODIC) { err = "ENOTSUP: periodic spline BC (dense fallback factor)"; return SSG_ENOTSUP; }
      if (kk < 0 || kk > SSG_BC_R3X) { err = "ArgumentError: unknown spline BC kind"; return SSG_EINVAL; }
    }
    if (!(d.l_q[v] > 0.0)) { err = "ArgumentError: l_q must be > 0"; return SSG_EINVAL; }
  }
  s.nzb = (s.geometry == SSG_RL) ? 1 : s.b_kDim;
  s.nslots = (s.geometry == SSG_RLZ) ? 7 : 5;

  std::vector<double> qp, qw;
  quadrature(d.mubar, d.quadrature, qp, qw);
  // mish points and per-cell tables (src/CubicBSpline.jl:887-897, 1000-1030, 1554-1567)
  s.mish.resize(s.iDim);
  s.sbw.assign(size_t(s.iDim) * 4, 0.0);
  for (int dd = 0; dd < 3; dd++) s.siw[dd].assign(size_t(s.iDim) * 4, 0.0);
  for (int mc = 0; mc < s.nc; mc++)
    for (int mu = 0; mu < s.mubar; mu++) {
      const double x = d.iMin + mc * s.DX + qp[mu] * s.DX;
      const int i = mc * s.mubar + mu;
      s.mish[i] = x;
      for (int t = 0; t < 4; t++) {
        const int m = mc - 1 + t;
        s.sbw[size_t(i) * 4 + t] = s.DX * qw[mu] * bspline(d.iMin, s.DX, s.DXrecip, m, x, 0);
        for (int dd = 0; dd < 3; dd++) s.siw[dd][size_t(i) * 4 + t] = bspline(d.iMin, s.DX, s.DXrecip, m, x, dd);
      }
    }
  // splines per variable
  s.spl.resize(d.nvars);
  for (int v = 0; v < d.nvars; v++) {
    GammaFold& g = s.spl[v].g;
    g.Mdim = s.b_iDim;
    // R3X folds as R3, the interior selector (src/CubicBSpline.jl:392, 411); its
    // background coefficients enter the solve (ssg_plan_set_ahat)
    s.spl[v].r3x = d.bcl_kind[v] == SSG_BC_R3X || d.bcr_kind[v] == SSG_BC_R3X;
    s.any_r3x = s.any_r3x || s.spl[v].r3x;
    g.lk = d.bcl_kind[v] == SSG_BC_R3X ? SSG_BC_R3 : d.bcl_kind[v];
    g.rk = d.bcr_kind[v] == SSG_BC_R3X ? SSG_BC_R3 : d.bcr_kind[v];
    g.rankL = rank_of(g.lk); g.rankR = rank_of(g.rk);
    g.aL = d.bcl_ab[2 * v]; g.bL = d.bcl_ab[2 * v + 1];
    g.aR = d.bcr_ab[2 * v]; g.bR = d.bcr_ab[2 * v + 1];
    g.Min = g.Mdim - g.rankL - g.rankR;
    if (g.Min < 2) { err = "ArgumentError: too few interior spline coefficients"; return SSG_EINVAL; }
    const int rc = build_spline_var(s, qp, qw, d.l_q[v], s.spl[v], err);
    if (rc != SSG_OK) return rc;
    s.maxMin = std::max(s.maxMin, g.Min);
  }
  // Chebyshev per variable
  if (cheb) {
    s.cheb.resize(d.nvars);
    for (int v = 0; v < d.nvars; v++) {
      const int rc = build_cheb_var(s, d.bcb_kind[v], d.bct_kind[v], s.cheb[v], err);
      if (rc != SSG_OK) return rc;
    }
  }
  // rings (src/factory.jl:153-175, 608-624; src/phase_filter.jl:63-71).  Cylindrical:
  // ri = r + patchOffsetL wavenumbers on 4 + 4 ri points.  Spherical (159-163,
  // 626-643): kr = max(1, round(sin(theta_r) max_ri)) wavenumbers on 4 kr points,
  // Julia's round being ties-to-even as nearbyint in the default rounding mode.
  // A ring's coefficients here always span its uncapped wavenumbers; a cap zeroes
  // the wavenumbers above it, which is what the reference's capped bDim and
  // its spline-stage limit (k <= the ring's kmax) give.
  if (s.geometry != SSG_RZ) {
    const int I = s.iDim;
    s.ringN.resize(I); s.ringRi.resize(I);
    s.ringOff.assign(I + 1, 0); s.coefOff.assign(I + 1, 0); s.phaseOff.assign(I + 1, 0); s.twOff.assign(I + 1, 0);
    for (int r = 0; r < I; r++) {
      int ri = r + 1 + d.patchOffsetL;
      int N = 4 + 4 * ri;
      if (s.spherical) {
        ri = std::max(1, int(std::nearbyint(std::sin(s.mish[r]) * double(s.K))));
        N = 4 * ri;
      }
      s.ringN[r] = N; s.ringRi[r] = ri;
      s.ringOff[r + 1] = s.ringOff[r] + N;
      s.coefOff[r + 1] = s.coefOff[r] + 2 * ri + 1;
      s.phaseOff[r + 1] = s.phaseOff[r] + ri;
      s.twOff[r + 1] = s.twOff[r] + N;
      s.maxN = std::max(s.maxN, N);
    }
    s.jDim = s.ringOff[I]; s.b_jDim = s.coefOff[I];
    s.phaseC.resize(s.phaseOff[I]); s.phaseS.resize(s.phaseOff[I]);
    s.twid.resize(s.twOff[I]);
    for (int r = 0; r < I; r++) {
      const int ri = s.ringRi[r], N = s.ringN[r];
      const double dl = 2 * M_PI / N;
      const double ymin = s.spherical ? 0.5 * dl * (r + 1) : 0.5 * dl * (ri - 1);
      for (int k = 1; k <= ri; k++) {
        s.phaseC[s.phaseOff[r] + k - 1] = std::cos(k * ymin);
        s.phaseS[s.phaseOff[r] + k - 1] = std::sin(k * ymin);
      }
      for (int m = 0; m < N; m++) s.twid[s.twOff[r] + m] = double(std::cos(2.0L * PI_L * (long double)m / (long double)N));
    }
    build_fft_tables(s);
    s.kmax.resize(size_t(d.nvars) * I);
    for (int v = 0; v < d.nvars; v++) {
      int cap = -1;
      if (s.geometry == SSG_RL) cap = d.max_wavenumber ? d.max_wavenumber[v] : -1;
      else cap = d.ring_kmax;
      for (int r = 0; r < I; r++) {
        const int ri = s.ringRi[r];
        s.kmax[size_t(v) * I + r] = cap >= 0 ? std::min(cap, ri) : ri;
      }
    }
  }
  if (s.geometry == SSG_RL) { s.npts = s.jDim; s.nspec = long(s.b_iDim) * (1 + 2 * s.K); }
  else if (s.geometry == SSG_RZ) { s.npts = long(s.iDim) * s.kDim; s.nspec = long(s.b_kDim) * s.b_iDim; }
  else { s.npts = long(s.kDim) * s.jDim; s.nspec = long(s.b_kDim) * s.b_iDim * (1 + 2 * s.K); }
  return SSG_OK;
}

}  // namespace ssg
