// Small grids for the CPU emulation drivers: the plan's host tables (build_setup)
// viewed through DevTables the way ssg_plan_create fills them for the chirp-z ring
// kernels and the RLZ inverse Chebyshev stage.
#pragma once
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "ssg_kernels.hpp"
#include "ssg_setup.hpp"

struct EmuGrid {
  ssg::Setup S;
  ssg::DevTables T{};
  std::vector<double> invt;   // [v][m][zb][z] = M_m[z][zb], zero padded to (bkp, kzp)
  std::vector<float> chirp32, h32, tw32, invt32;   // the FP32 form's float copies
};

// geom SSG_RL / SSG_SL / SSG_RLZ / SSG_SLZ; kDim, b_kDim for the z grids; cap -1 = ring
// specific; off = patchOffsetL (rings ri = r + 1 + off on cylindrical grids)
inline void emu_grid(EmuGrid& G, int geom, int nc, int nvars, int kDim, int bk, int cap, int off) {
  using namespace ssg;
  const bool z = geom == SSG_RLZ || geom == SSG_SLZ;
  static std::vector<std::vector<double>> keep_d;
  static std::vector<std::vector<int>> keep_i;
  keep_d.push_back(std::vector<double>(nvars, 2.0));
  const double* lq = keep_d.back().data();
  keep_d.push_back(std::vector<double>(2 * nvars, 0.0));
  const double* ab = keep_d.back().data();
  keep_i.push_back(std::vector<int>(nvars, SSG_BC_R0));
  const int* bc = keep_i.back().data();
  keep_i.push_back(std::vector<int>(nvars, SSG_CBC_R0));
  const int* cb = keep_i.back().data();
  keep_i.push_back(std::vector<int>(nvars, cap));
  ssg_desc d{};
  d.geometry = geom; d.num_cells = nc; d.iMin = 0; d.iMax = geom == SSG_SL || geom == SSG_SLZ ? 3.141592653589793 : 100;
  d.mubar = 3; d.quadrature = 0; d.nvars = nvars;
  d.kDim = z ? kDim : 0; d.b_kDim = z ? bk : 0; d.kMin = 0; d.kMax = 1;
  d.ring_kmax = z ? cap : -1;
  d.max_wavenumber = z ? nullptr : keep_i.back().data();
  d.patchOffsetL = off;
  d.l_q = lq; d.bcl_kind = bc; d.bcr_kind = bc; d.bcl_ab = ab; d.bcr_ab = ab; d.bcb_kind = cb; d.bct_kind = cb;
  std::string err;
  if (build_setup(d, G.S, err)) { fprintf(stderr, "setup: %s\n", err.c_str()); exit(2); }
  const Setup& S = G.S;
  DevTables& T = G.T;
  T.geometry = S.geometry; T.nvars = S.nvars; T.iDim = S.iDim; T.K = S.K; T.kDim = S.kDim; T.nzb = S.nzb;
  T.jDim = S.jDim; T.npts = S.npts; T.b_jDim = S.b_jDim;
  T.ntr = S.nvars * S.nzb;
  T.ntb = S.geometry == SSG_RLZ ? S.nzb : S.nvars;
  T.nvb = T.ntr / T.ntb;
  T.Ktot = S.phaseOff[S.iDim] + S.iDim;
  T.ringN = S.ringN.data(); T.ringRi = S.ringRi.data(); T.ringOff = S.ringOff.data(); T.phaseOff = S.phaseOff.data();
  T.coefOff = S.coefOff.data(); T.kmax = S.kmax.data(); T.phaseC = S.phaseC.data(); T.phaseS = S.phaseS.data();
  T.fftL = S.fftL.data(); T.fftHOff = S.fftHOff.data(); T.fftChirp = S.fftChirp.data();
  T.fftH = S.fftH.data(); T.fftTw = S.fftTw.data();
  if (S.geometry == SSG_RLZ) {
    const int b = S.nzb, kz = S.kDim, bkp = syn_bkp(b), kzp = syn_kzp(kz);
    G.invt.assign(size_t(S.nvars) * 3 * bkp * kzp, 0.0);
    for (int v = 0; v < S.nvars; v++)
      for (int m = 0; m < 3; m++) {
        const std::vector<double>& M = m == 0 ? S.cheb[v].inv0 : m == 1 ? S.cheb[v].inv1 : S.cheb[v].inv2;
        for (int zb = 0; zb < b; zb++)
          for (int zz = 0; zz < kz; zz++) G.invt[((size_t(v) * 3 + m) * bkp + zb) * kzp + zz] = M[size_t(zz) * b + zb];
      }
    T.invt = G.invt.data();
  }
  G.chirp32.assign(S.fftChirp.begin(), S.fftChirp.end());
  G.h32.assign(S.fftH.begin(), S.fftH.end());
  G.tw32.assign(S.fftTw.begin(), S.fftTw.end());
  G.invt32.assign(G.invt.begin(), G.invt.end());
  T.f32.fftChirp = G.chirp32.data();
  T.f32.fftH = G.h32.data();
  T.f32.fftTw = G.tw32.data();
  T.f32.invt = G.invt32.data();
}

inline int emu_geom(const char* s) {
  const std::string g(s);
  return g == "RL" ? SSG_RL : g == "SL" ? SSG_SL : g == "RLZ" ? SSG_RLZ : g == "SLZ" ? SSG_SLZ : -1;
}
