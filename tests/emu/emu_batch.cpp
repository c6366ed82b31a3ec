// CPU emulation check of the batched chirp-z launches (k_syn_fft_batch,
// k_ana_fft_batch, k_syn_cheb_batch): a batch of RL, RLZ, SL and SLZ grids whose
// job ranges are padded to multiples of 8 blocks, as the batching form lays them
// out, so that every launch has padding blocks past a job's own blocks.  Every
// array is allocated at its exact size (AddressSanitizer reports any access past
// it), and each grid's batched results must be bitwise its single-grid results.
#include <cstdio>
#include <cstring>
#include <random>

#include "emu_grid.hpp"
using namespace ssg;

// job ranges of one batched launch: job g owns blocks [start[g], start[g + 1]),
// each range rounded up to a multiple of 8 (the XCD of a block is that of its
// local index)
static void add_range(std::vector<int>& start, int blocks) { start.push_back(start.back() + (blocks + 7) / 8 * 8); }

template <typename V>
static bool same(const V& a, const V& b) { return a.size() == b.size() && !std::memcmp(a.data(), b.data(), a.size() * sizeof(a[0])); }

int main() {
  struct Spec { const char* g; int nc, nvars, kDim, bk, off; };
  const Spec specs[] = {{"RL", 6, 2, 0, 0, 0}, {"RLZ", 5, 2, 8, 6, 3}, {"SL", 7, 1, 0, 0, 0}, {"SLZ", 5, 3, 10, 5, 0}};
  const int ng = 4;
  std::vector<EmuGrid> G(ng);
  std::vector<FftSynBlocks> sb(ng), ab(ng);
  std::vector<std::vector<double>> X(ng), Ys(ng), Yb(ng), src(ng), rcs(ng), rcb(ng), rcf(ng), ps(ng), pb(ng);
  std::vector<std::vector<float>> srcf(ng);
  std::mt19937_64 rng(3);
  std::normal_distribution<double> nd;
  for (int i = 0; i < ng; i++) {
    emu_grid(G[i], emu_geom(specs[i].g), specs[i].nc, specs[i].nvars, specs[i].kDim, specs[i].bk, -1, specs[i].off);
    const Setup& S = G[i].S;
    const DevTables& T = G[i].T;
    const bool rlz = S.geometry == SSG_RLZ;
    sb[i] = fft_blocks(S.fftL, S.iDim, syn_fft_pairs(S.nvars, S.nzb));
    ab[i] = fft_blocks(S.fftL, S.iDim, ana_fft_pairs(S.nvars, S.nzb, rlz));
    X[i].resize(size_t(6) * T.ntr * T.Ktot);
    for (auto& x : X[i]) x = nd(rng);
    const size_t ny = rlz ? size_t(S.nvars) * 5 * S.nzb * S.jDim : size_t(5) * S.nvars * S.npts;
    Ys[i].assign(ny, -7.0);
    Yb[i].assign(ny, -7.0);
    src[i].resize(rlz ? size_t(S.nvars) * S.nzb * S.jDim : size_t(S.nvars) * S.npts);
    for (auto& x : src[i]) x = nd(rng);
    srcf[i].resize(src[i].size());
    for (size_t k = 0; k < src[i].size(); k++) srcf[i][k] = float(src[i][k]);
    rcs[i].assign(size_t(T.ntr) * S.b_jDim, -7.0);
    rcb[i].assign(rcs[i].size(), -7.0);
    rcf[i].assign(rcs[i].size(), -7.0);
    ps[i].assign(size_t(7) * S.nvars * S.npts, -9.0);
    pb[i].assign(ps[i].size(), -9.0);
    printf("grid %d %s: %d rings, synthesis %d blocks, analysis %d blocks, Chebyshev %d blocks\n", i, specs[i].g,
           S.iDim, sb[i].nblocks, ab[i].nblocks, rlz ? syn_cheb_blocks(T) : 0);
  }
  // single-grid launches
  for (int i = 0; i < ng; i++) {
    const DevTables& T = G[i].T;
    launch_syn_fft<double, double>(T, sb[i].blk.data(), sb[i].nblocks, X[i].data(), Ys[i].data(), nullptr);
    launch_ana_fft<double>(T, ab[i].blk.data(), ab[i].nblocks, src[i].data(), rcs[i].data(), nullptr);
    if (G[i].S.geometry == SSG_RLZ) launch_syn_cheb<double>(T, Ys[i].data(), ps[i].data(), nullptr);
  }
  // batched launches over all grids
  std::vector<FftJob<double, double>> fj;
  std::vector<AnaFftJob<double>> aj;
  std::vector<AnaFftJob<float>> af;
  std::vector<ChebSynJob<double>> cj;
  std::vector<int> fs{0}, as{0}, cs{0};
  size_t lds = 0;
  for (int i = 0; i < ng; i++) {
    const DevTables& T = G[i].T;
    fj.push_back({T, sb[i].blk.data(), X[i].data(), Yb[i].data()});
    add_range(fs, sb[i].nblocks);
    aj.push_back({T, ab[i].blk.data(), src[i].data(), rcb[i].data()});
    af.push_back({T, ab[i].blk.data(), srcf[i].data(), rcf[i].data()});
    add_range(as, ab[i].nblocks);
    if (G[i].S.geometry == SSG_RLZ) {
      cj.push_back({T, Yb[i].data(), pb[i].data()});
      add_range(cs, syn_cheb_blocks(T));
      lds = std::max(lds, syn_cheb_lds(T, sizeof(double)));
    }
  }
  launch_syn_fft_batch<double, double>(fj.data(), fs.data(), int(fj.size()), fs.back(), nullptr);
  launch_ana_fft_batch<double>(aj.data(), as.data(), int(aj.size()), as.back(), nullptr);
  launch_ana_fft_batch<float>(af.data(), as.data(), int(af.size()), as.back(), nullptr);
  // the Chebyshev jobs here share one kernel class (b_kDim up to 16), as the batching
  // form groups them
  for (auto& j : cj)
    if (syn_cheb_class(j.T) != syn_cheb_class(cj[0].T)) { printf("mixed Chebyshev classes\n"); return 1; }
  launch_syn_cheb_batch<double>(cj.data(), cs.data(), int(cj.size()), cs.back(), lds, syn_cheb_class(cj[0].T), nullptr);
  // the FP32 form's synthesis (FP32 transforms, float output) and float Chebyshev stage
  std::vector<std::vector<float>> X32(ng), Y32s(ng), Y32b(ng), p32s(ng), p32b(ng);
  std::vector<FftJob<float, float>> fj32;
  std::vector<ChebSynJob<float>> cj32;
  std::vector<int> cs32{0};
  size_t lds32 = 0;
  for (int i = 0; i < ng; i++) {
    const DevTables& T = G[i].T;
    X32[i].assign(X[i].begin(), X[i].end());
    Y32s[i].assign(Ys[i].size(), -7.0f);
    Y32b[i].assign(Ys[i].size(), -7.0f);
    p32s[i].assign(ps[i].size(), -9.0f);
    p32b[i].assign(ps[i].size(), -9.0f);
    launch_syn_fft<float, float>(T, sb[i].blk.data(), sb[i].nblocks, X32[i].data(), Y32s[i].data(), nullptr);
    fj32.push_back({T, sb[i].blk.data(), X32[i].data(), Y32b[i].data()});
    if (G[i].S.geometry == SSG_RLZ) {
      launch_syn_cheb<float>(T, Y32s[i].data(), p32s[i].data(), nullptr);
      cj32.push_back({T, Y32b[i].data(), p32b[i].data()});
      add_range(cs32, syn_cheb_blocks(T));
      lds32 = std::max(lds32, syn_cheb_lds(T, sizeof(float)));
    }
  }
  launch_syn_fft_batch<float, float>(fj32.data(), fs.data(), int(fj32.size()), fs.back(), nullptr);
  launch_syn_cheb_batch<float>(cj32.data(), cs32.data(), int(cj32.size()), cs32.back(), lds32, syn_cheb_class(cj32[0].T),
                               nullptr);
  int bad = 0;
  for (int i = 0; i < ng; i++) {
    const bool ok32 = same(Y32s[i], Y32b[i]) && (G[i].S.geometry != SSG_RLZ || same(p32s[i], p32b[i]));
    printf("grid %d %s: FP32 synthesis and Chebyshev %s\n", i, specs[i].g, ok32 ? "bitwise" : "DIFFERS");
    bad += !ok32;
  }
  for (int i = 0; i < ng; i++) {
    const bool ok_syn = same(Ys[i], Yb[i]), ok_ana = same(rcs[i], rcb[i]);
    const bool ok_cheb = G[i].S.geometry != SSG_RLZ || same(ps[i], pb[i]);
    // the float-source analysis reads the rounded input: compare with a single-grid float launch
    std::vector<double> rf(rcs[i].size(), -7.0);
    launch_ana_fft<float>(G[i].T, ab[i].blk.data(), ab[i].nblocks, srcf[i].data(), rf.data(), nullptr);
    const bool ok_f = same(rf, rcf[i]);
    printf("grid %d %s: synthesis %s, analysis %s, float analysis %s, Chebyshev %s\n", i, specs[i].g,
           ok_syn ? "bitwise" : "DIFFERS", ok_ana ? "bitwise" : "DIFFERS", ok_f ? "bitwise" : "DIFFERS",
           ok_cheb ? "bitwise" : "DIFFERS");
    bad += !ok_syn + !ok_ana + !ok_f + !ok_cheb;
  }
  printf("padded launches: synthesis %d blocks, analysis %d, Chebyshev %d\n", fs.back(), as.back(), cs.back());
  return bad ? 1 : 0;
}
