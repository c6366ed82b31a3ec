// CPU emulation check of the chirp-z ring synthesis (k_syn_fft) against a direct
// HC2R sum in long double, and of the RLZ inverse Chebyshev stage (k_syn_cheb)
// against a direct contraction of the synthesis output.
// usage: emu_syn GEOM NC NVARS OFF SPREAD [F32]   (SPREAD 1: column magnitudes over
// 12 decades and exactly zero columns; F32 1: the FP32 form's transforms on float
// input, output and tables, held to 2e-6 per column)
#include <cstdio>
#include <random>

#include "emu_grid.hpp"
using namespace ssg;

int main(int argc, char** argv) {
  if (argc < 6) return 2;
  EmuGrid G;
  emu_grid(G, emu_geom(argv[1]), atoi(argv[2]), atoi(argv[3]), 8, 6, -1, atoi(argv[4]));
  const bool spread = atoi(argv[5]) != 0;
  const bool f32 = argc > 6 && atoi(argv[6]) != 0;
  const Setup& S = G.S;
  DevTables& T = G.T;
  const int ntb = T.ntb;
  std::vector<double> X(size_t(6) * T.ntr * T.Ktot);
  std::mt19937_64 rng(7);
  std::normal_distribution<double> nd;
  for (auto& x : X) x = nd(rng);
  if (spread) {
    // one scale per (P | Q, order, transform) run pair: P and Q of a run share it
    for (int vb = 0; vb < T.nvb; vb++)
      for (int r = 0; r < S.iDim; r++) {
        const int kl = S.ringRi[r] + 1;
        double* Xr = X.data() + (long(vb) * T.Ktot + S.phaseOff[r] + r) * 6 * ntb;
        for (int run = 0; run < 6 * ntb; run++) {
          const int rem = run % (3 * ntb);
          std::mt19937_64 r2(1000003 * (vb * 7919 + r) + rem);
          std::uniform_real_distribution<double> u2(-12.0, 0.0);
          const double sc = (rem * 7 + r) % 5 == 0 ? 0.0 : std::pow(10.0, u2(r2));
          for (int k = 0; k < kl; k++) Xr[run * kl + k] *= sc;
        }
      }
  }
  FftSynBlocks B = fft_blocks(S.fftL, S.iDim, syn_fft_pairs(S.nvars, S.nzb));
  for (int r = 0; r < S.iDim; r++)
    if (!S.fftL[r]) { printf("ring %d outside the chirp-z sizes\n", r); return 1; }
  const bool rlz = S.geometry == SSG_RLZ;
  const size_t nout = rlz ? size_t(S.nvars) * 5 * S.nzb * S.jDim : size_t(5) * S.nvars * S.npts;
  std::vector<double> out(nout, -7.0);
  std::vector<float> X32, out32;
  if (f32) {
    // the reference then reads the rounded input
    X32.assign(X.begin(), X.end());
    for (size_t i = 0; i < X.size(); i++) X[i] = X32[i];
    out32.assign(nout, -7.0f);
    launch_syn_fft<float, float>(T, B.blk.data(), B.nblocks, X32.data(), out32.data(), nullptr);
    for (size_t i = 0; i < nout; i++) out[i] = out32[i];
  } else {
    launch_syn_fft<double, double>(T, B.blk.data(), B.nblocks, X.data(), out.data(), nullptr);
  }
  const double bar = f32 ? 2e-6 : 2e-15;
  double worst = 0;
  for (int vb = 0; vb < T.nvb; vb++)
    for (int r = 0; r < S.iDim; r++) {
      const int N = S.ringN[r], K = S.ringRi[r], kl = K + 1;
      const double* Xr = X.data() + (long(vb) * T.Ktot + S.phaseOff[r] + r) * 6 * ntb;
      std::vector<long double> ct(N), st(N);   // cos, sin of 2 pi m / N
      for (int m = 0; m < N; m++) {
        ct[m] = cosl(2.0L * 3.14159265358979323846264338327950288L * m / N);
        st[m] = sinl(2.0L * 3.14159265358979323846264338327950288L * m / N);
      }
      for (int c = 0; c < 5 * ntb; c++) {
        double cr2 = 0, ce2 = 0;
        const int f = c / ntb, tl = c % ntb, dd = f < 3 ? 0 : f - 2;
        const double* P = Xr + (dd * ntb + tl) * kl;
        const double* Q = Xr + ((3 + dd) * ntb + tl) * kl;
        const int km = S.kmax[(rlz ? vb : tl) * S.iDim + r];
        const int slot = f == 0 ? 0 : f == 1 ? 3 : f == 2 ? 4 : f == 3 ? 1 : 2;
        const double* o = rlz ? out.data() + (long(vb) * 5 * ntb + c) * S.jDim + S.ringOff[r]
                              : out.data() + (long(slot) * S.nvars + tl) * S.npts + S.ringOff[r];
        for (int l = 0; l < N; l++) {
          long double y = (f == 1 || f == 2) ? 0.0L : P[0];
          for (int k = 1; k <= km; k++) {
            const long double p = P[k], q = Q[k];
            long double pp = p, qq = q;
            if (f == 1) { pp = k * q; qq = -k * p; }
            if (f == 2) { pp = -(long double)k * k * p; qq = -(long double)k * k * q; }
            const long m = (long(k) * l) % N;
            y += 2.0L * (pp * ct[m] + qq * st[m]);
          }
          const double e = std::fabs(double(y) - o[l]);
          cr2 += double(y) * double(y);
          ce2 += e * e;
        }
        if (cr2 == 0 && ce2 != 0) { printf("zero column %d of ring %d not zero\n", c, r); return 1; }
        worst = std::max(worst, cr2 > 0 ? std::sqrt(ce2 / cr2) : 0.0);
      }
    }
  printf("syn%s %s nc %s nvars %s off %s: %d blocks, worst column relL2 %.3e\n", f32 ? " fp32" : "", argv[1],
         argv[2], argv[3], argv[4], B.nblocks, worst);
  if (!(worst < bar)) return 1;
  if (rlz) {
    const int bk = S.nzb, kz = S.kDim;
    std::vector<double> phys(size_t(7) * S.nvars * S.npts, -9.0);
    if (f32) {
      std::vector<float> p32(phys.size(), -9.0f);
      launch_syn_cheb<float>(T, out32.data(), p32.data(), nullptr);
      for (size_t i = 0; i < phys.size(); i++) phys[i] = p32[i];
    } else {
      launch_syn_cheb<double>(T, out.data(), phys.data(), nullptr);
    }
    double e2 = 0, r2 = 0;
    for (int v = 0; v < S.nvars; v++)
      for (int f = 0; f < 5; f++)
        for (int m = 0; m < (f == 0 ? 3 : 1); m++) {
          const int slot = f == 0 ? (m == 0 ? 0 : m == 1 ? 5 : 6) : (f == 1 ? 3 : f == 2 ? 4 : f == 3 ? 1 : 2);
          const std::vector<double>& Mm = m == 0 ? S.cheb[v].inv0 : m == 1 ? S.cheb[v].inv1 : S.cheb[v].inv2;
          for (long pt = 0; pt < S.jDim; pt++)
            for (int zz = 0; zz < kz; zz++) {
              long double acc = 0;
              for (int zb = 0; zb < bk; zb++)
                acc += (long double)Mm[size_t(zz) * bk + zb] * out[(size_t(v) * 5 * bk + f * bk + zb) * S.jDim + pt];
              const double g = phys[(size_t(slot) * S.nvars + v) * S.npts + pt * kz + zz];
              e2 += (g - double(acc)) * (g - double(acc));
              r2 += double(acc) * double(acc);
            }
        }
    printf("cheb stage relL2 %.3e\n", std::sqrt(e2 / r2));
    if (!(std::sqrt(e2 / r2) < (f32 ? 1e-6 : 1e-15))) return 1;
  }
  return 0;
}
