// CPU emulation check of the chirp-z ring analysis (k_ana_fft) against a direct
// R2HC / N with the forward phase filter, in long double.
// usage: emu_ana GEOM NC NVARS OFF CAP   (column magnitudes over 12 decades, some
// columns exactly zero)
#include <cstdio>
#include <random>

#include "emu_grid.hpp"
using namespace ssg;

int main(int argc, char** argv) {
  if (argc < 6) return 2;
  EmuGrid G;
  const int nvars = atoi(argv[3]), cap = atoi(argv[5]);
  emu_grid(G, emu_geom(argv[1]), atoi(argv[2]), nvars, 8, 7, cap, atoi(argv[4]));
  const Setup& S = G.S;
  const bool rlz = S.geometry == SSG_RLZ;
  const int nzb = S.nzb;
  std::mt19937_64 rng(11);
  std::normal_distribution<double> nd;
  std::uniform_real_distribution<double> ud(-12, 0);
  // input: RL physical slot 1 (npts per variable), RLZ cm[var][zb][ring point]
  std::vector<double> in(rlz ? size_t(nvars) * nzb * S.jDim : size_t(nvars) * S.npts);
  for (auto& x : in) x = nd(rng);
  const int ncol = nvars * nzb;
  auto at = [&](int c, long pt) -> double& {
    const int v = c / nzb, zb = c % nzb;
    return rlz ? in[(size_t(v) * nzb + zb) * S.jDim + pt] : in[size_t(v) * S.npts + pt];
  };
  for (int c = 0; c < ncol; c++) {
    const double sc = c % 5 == 3 ? 0.0 : std::pow(10.0, ud(rng));
    for (long pt = 0; pt < S.jDim; pt++) at(c, pt) *= sc;
  }
  FftSynBlocks B = fft_blocks(S.fftL, S.iDim, ana_fft_pairs(nvars, nzb, rlz));
  std::vector<double> rc(size_t(ncol) * S.b_jDim, -7.0);
  launch_ana_fft<double>(G.T, B.blk.data(), B.nblocks, in.data(), rc.data(), nullptr);
  double worst = 0;
  for (int c = 0; c < ncol; c++) {
    const int v = c / nzb;
    double e2 = 0, r2 = 0;
    for (int r = 0; r < S.iDim; r++) {
      const int N = S.ringN[r], K = S.ringRi[r], bDim = 2 * K + 1, km = S.kmax[v * S.iDim + r];
      std::vector<long double> ref(bDim, 0.0L), ct(N), st(N);   // cos, sin of 2 pi m / N
      for (int m = 0; m < N; m++) {
        ct[m] = cosl(2.0L * 3.14159265358979323846264338327950288L * m / N);
        st[m] = sinl(2.0L * 3.14159265358979323846264338327950288L * m / N);
      }
      for (int k = 0; k <= K && k <= km; k++) {
        long double re = 0, im = 0;
        for (int l = 0; l < N; l++) {
          const long m = (long(k) * l) % N;
          re += at(c, S.ringOff[r] + l) * ct[m];
          im -= at(c, S.ringOff[r] + l) * st[m];
        }
        re /= N;
        im /= N;
        if (k == 0) { ref[0] = re; continue; }
        const long double cc = S.phaseC[S.phaseOff[r] + k - 1], ss = S.phaseS[S.phaseOff[r] + k - 1];
        ref[2 * k - 1] = cc * re + ss * im;   // rc's ring layout: spectral line order
        ref[2 * k] = -ss * re + cc * im;
      }
      for (int i = 0; i < bDim; i++) {
        const double g = rc[size_t(c) * S.b_jDim + S.coefOff[r] + i];
        e2 += (g - double(ref[i])) * (g - double(ref[i]));
        r2 += double(ref[i]) * double(ref[i]);
      }
    }
    if (r2 == 0 && e2 != 0) { printf("zero column %d not zero\n", c); return 1; }
    worst = std::max(worst, r2 > 0 ? std::sqrt(e2 / r2) : 0.0);
  }
  printf("ana %s nc %s nvars %d off %s cap %d: %d blocks, worst column relL2 %.3e\n", argv[1], argv[2], nvars, argv[4],
         cap, B.nblocks, worst);
  return worst < 2e-15 ? 0 : 1;
}
