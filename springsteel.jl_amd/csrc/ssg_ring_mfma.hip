// ssg_ring_mfma.hip -- ring synthesis (HC2R with azimuthal derivatives) as
// FP64 MFMA GEMMs, fused for RLZ with the inverse Chebyshev stage.
//
// Ring r has N = 4 + 4 ri points at phi_l = 2 pi l / N.  After the inverse
// phase rotation (src/phase_filter.jl:154-190, applied by k_sa_si_ring) every
// output of one family is
//   y(phi) = dc + 2 sum_{k=1}^{kmax} [P_k cos(k phi) + Q_k sin(k phi)]
// (HC2R, src/Fourier.jl:463-468), with (P, Q) = (a_re, -a_im) for values,
// (-k a_im, -k a_re) = (k Q, -k P) for d/dlambda and (-k^2 a_re, k^2 a_im) =
// (-k^2 P, -k^2 Q) for d2/dlambda2 (FIxcoefficients, src/Fourier.jl:495-503).
// The four points l, N/2 - l, N/2 + l, N - l share |cos|, |sin|; splitting k
// by parity gives all four from PCe, PCo (sum P cos over even / odd k) and
// QSe, QSo:
//   y(l)       = dc + 2 (PCe + PCo + QSe + QSo)
//   y(N - l)   = dc + 2 (PCe + PCo - QSe - QSo)
//   y(N/2 - l) = dc + 2 (PCe - PCo - QSe + QSo)
//   y(N/2 + l) = dc + 2 (PCe - PCo + QSe - QSo)
// Each partial sum is a GEMM [quads x k] x [k x columns]: A = cos/sin(k phi_l)
// from the ring's cos table in LDS, B = (P, Q) staged in LDS per chunk of 16
// wavenumbers, accumulated with v_mfma_f64_16x16x4f64.  A column is one
// (family, transform) pair, family f: 0 = values (d0), 1 = d/dlambda (d0),
// 2 = d2/dlambda2 (d0), 3 = d/dr values (d1), 4 = d2/dr2 values (d2).
//
// Input X[vb][ring][P|Q][d][tl][k] (k_sa_si_ring): per ring 6 ntb runs of
// ri + 1 wavenumbers, k = 0 holding the dc values.  Each wave owns CT column tiles and QT
// quad tiles of 16; every tile it owns is valid (columns padded with zeros,
// quad rows past N/4 computed and dropped), so the inner loop has no branches.
//
// RL: a block covers all variables; outputs go straight to physical slots
// 1, 4, 5, 2, 3.  RLZ: a block covers one variable and all Chebyshev modes;
// per family the synthesised values are staged in LDS and contracted with the
// variable's (kDim x b_kDim) inverse Chebyshev matrices (a second MFMA GEMM)
// into physical slots: f0 -> 1 (M0), 6 (M1), 7 (M2); f1 -> 4; f2 -> 5;
// f3 -> 2; f4 -> 3 (src/transforms_cylindrical.jl:780-844).
#include "ssg_kernels.hpp"
#include "ssg_mfma.hpp"

namespace ssg {


static constexpr int kKC = 16;       // wavenumbers staged per chunk (two k-steps of 8)
static constexpr int kPre = 1;       // staged elements per thread prefetched across a chunk

// Family of staged column group g = col / ntb: the groups hold f0, f3, f4
// (orders d = 0, 1, 2 of the X input, in X's own order) and then f1, f2.
__device__ __forceinline__ int fam_of_group(int g) { return g == 0 ? 0 : g <= 2 ? g + 2 : g - 2; }


template <typename R, int CT, int QT>
__device__ __forceinline__ void ring_syn_body(const DevTables& T, const SynLaunch& P, int lin,
                                              const R* __restrict__ X, R* __restrict__ phys) {
  typedef typename Mfma<R>::v4 v4;
  extern __shared__ __align__(16) unsigned char smraw[];
  R* sm = reinterpret_cast<R*>(smraw);
  const RTab<R> tb = rtab<R>(T);
  // unit = (ring, variable block), largest rings first; the quad tiles of a
  // unit share its staged coefficients and run on one XCD
  int unit, qt;
  xcd_unit_tile(lin, P.nqt, unit, qt);
  if (unit >= T.iDim * T.nvb) return;
  const int r = T.iDim - 1 - unit / T.nvb;
  const int vb = unit - (unit / T.nvb) * T.nvb;
  const int N = T.ringN[r], ri = T.ringRi[r], N4 = N / 4;
  const int qbase = qt * P.qpb;
  if (qbase > N4) return;                            // block-uniform
  const bool rlz = T.geometry == 3;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int qg = w % P.nwq, cg = w / P.nwq;
  const int nthr = blockDim.x;
  const int ntb = P.ntb, kl = ri + 1;
  // this ring's block: run (P|Q, d, tl) at Xr + ((pq * 3 + d) * ntb + tl) * kl, wavenumber k at offset k
  const R* Xr = X + (long(vb) * T.Ktot + T.phaseOff[r] + r) * 6 * ntb;

  // block-wide maximum wavenumber over the variables it covers
  int kmb = 0;
  if (rlz) kmb = T.kmax[(vb * ntb / T.nzb) * T.iDim + r];
  else for (int v = 0; v < T.nvars; v++) kmb = max(kmb, T.kmax[v * T.iDim + r]);

  R* tw = sm;                                        // [2N] cos(2 pi m / N), m < 2N
  R* stP = sm + 2 * (4 + 4 * T.K);                   // [kKC][ldp]
  R* stQ = stP + kKC * P.ldp;
  {
    const R* tws = tb.twid + T.twOff[r];
    for (int m = threadIdx.x; m < 2 * N; m += nthr) tw[m] = tws[m < N ? m : m - N];
    // padding columns stay zero for the whole main loop
    const int c0 = P.ncols, cw = P.nctp * 16 - P.ncols;
    for (int e = threadIdx.x; e < kKC * cw; e += nthr) {
      const int kk = e / cw, c = c0 + e - (e / cw) * cw;
      stP[kk * P.ldp + c] = 0.0;
      stQ[kk * P.ldp + c] = 0.0;
    }
  }

  // Staging: element e = threadIdx.x + s * nthr of a chunk is (rem, kk) with
  // rem = d * ntb + tl and kk fastest, so a wave reads runs of kKC wavenumbers.
  // Staged columns: rem itself for every order d (f0, f3, f4 in the first
  // 3 ntb columns), and f1 = (k Q, -k P), f2 = (-k^2 P, -k^2 Q) of the values
  // at rem + 3 ntb and rem + 4 ntb.  Offsets are formed from the thread index
  // at each use (unsigned, kKC a power of two) rather than kept in registers.
  const unsigned nrem = 3u * unsigned(ntb);
  const unsigned qoff = nrem * unsigned(kl);          // Q runs follow the P runs
  const int qst = kKC * P.ldp;                        // stQ - stP
  auto load = [&](int s, int k0, R& pv, R& qv) {
    const unsigned e = threadIdx.x + unsigned(s * nthr);
    const unsigned rem = e / kKC;
    const int k = k0 + int(e % kKC);
    pv = R(0); qv = R(0);
    if (rem < nrem && k >= 1 && k <= kmb) {
      const unsigned o = rem * unsigned(kl) + unsigned(k);
      pv = Xr[o];
      qv = Xr[o + qoff];
    }
  };
  auto put = [&](int s, int k0, R pv, R qv) {
    const unsigned e = threadIdx.x + unsigned(s * nthr);
    const unsigned rem = e / kKC, kk = e % kKC;
    if (rem >= nrem) return;
    R* x = stP + (kk * unsigned(P.ldp) + rem);
    x[0] = pv;
    x[qst] = qv;
    if (rem < unsigned(ntb)) {
      const R kd = R(k0 + int(kk)), k2 = kd * kd;
      x[nrem] = kd * qv;
      x[nrem + qst] = -kd * pv;
      x[nrem + ntb] = -k2 * pv;
      x[nrem + ntb + qst] = -k2 * qv;
    }
  };
  for (int s = 0; s < P.nslot; s++) {
    R pv, qv;
    load(s, 0, pv, qv);
    put(s, 0, pv, qv);
  }

  v4 acc[QT][CT][4];
#pragma unroll
  for (int j = 0; j < QT; j++)
#pragma unroll
    for (int c = 0; c < CT; c++)
#pragma unroll
      for (int s = 0; s < 4; s++) acc[j][c][s] = v4{0, 0, 0, 0};

  // A operand: row = quad q of this lane, k-group g = lane / 16 holds k = kb + 2g
  // (even) and kb + 2g + 1 (odd): (cos, sin)(k theta_q), theta_q = 2 pi q / N.  The
  // even-k pair is read from the table at the start of each chunk (index (k q) mod N,
  // advanced by 16q per chunk; sin(x) = cos(x - pi/2) at index + 3N/4) and rotated by
  // 8 theta_q for the chunk's second k-step; the odd-k pair is the even one rotated by
  // theta_q.  So a table entry is at most two complex rotations from its exact value,
  // and the table reads, 64-lane gathers, are a quarter of one per entry.  Rows past
  // N/4 read valid table entries and are never stored.
  // wrap(x) for x < 2N: min over unsigned of x and x - N.
  auto wrap = [N](int x) { return int(min(unsigned(x), unsigned(x - N))); };
  const int q3 = 3 * N4;
  int qm[QT], step16[QT], ie[QT];
  R c1[QT], s1[QT], c8[QT], s8[QT];
#pragma unroll
  for (int j = 0; j < QT; j++) {
    qm[j] = (qbase + (qg * QT + j) * 16 + (lane & 15)) % N;
    step16[j] = (16 * qm[j]) % N;
    ie[j] = (2 * (lane >> 4) * qm[j]) % N;
  }
  const int colb = cg * CT * 16 + (lane & 15);       // B column of tile c: colb + 16 c
  __syncthreads();                                   // the table and the first chunk are in LDS
#pragma unroll
  for (int j = 0; j < QT; j++) {
    const int i8 = (8 * qm[j]) % N;
    c1[j] = tw[qm[j]]; s1[j] = tw[wrap(qm[j] + q3)];
    c8[j] = tw[i8]; s8[j] = tw[wrap(i8 + q3)];
  }

  for (int k0 = 0; k0 <= kmb; k0 += kKC) {
    const bool more = k0 + kKC <= kmb;
    R pr[kPre], pi[kPre];
#pragma unroll
    for (int s = 0; s < kPre; s++) {
      pr[s] = R(0); pi[s] = R(0);
      if (more && s < P.nslot) load(s, k0 + kKC, pr[s], pi[s]);
    }
    static_assert(kKC == 16, "the A operand is anchored once per chunk of two k-steps");
    R ce[QT], se[QT];
#pragma unroll
    for (int j = 0; j < QT; j++) {
      ce[j] = tw[ie[j]]; se[j] = tw[wrap(ie[j] + q3)];
      ie[j] = wrap(ie[j] + step16[j]);
    }
#pragma unroll
    for (int ks = 0; ks < kKC / 8; ks++) {
      if (k0 + ks * 8 > kmb) break;                  // uniform
      R co[QT], so[QT];
#pragma unroll
      for (int j = 0; j < QT; j++) {
        if (ks > 0) {                                // (ce, se) rotated by 8 theta_q
          const R cn = ce[j] * c8[j] - se[j] * s8[j];
          const R sn = se[j] * c8[j] + ce[j] * s8[j];
          ce[j] = cn; se[j] = sn;
        }
        co[j] = ce[j] * c1[j] - se[j] * s1[j];       // odd k: rotated by theta_q
        so[j] = se[j] * c1[j] + ce[j] * s1[j];
      }
      const int kr = ks * 8 + 2 * (lane >> 4);       // staged row of the even k
#pragma unroll
      for (int c = 0; c < CT; c++) {
        const int col = colb + 16 * c;
        const R pe = stP[kr * P.ldp + col], po = stP[(kr + 1) * P.ldp + col];
        const R qe = stQ[kr * P.ldp + col], qo = stQ[(kr + 1) * P.ldp + col];
#pragma unroll
        for (int j = 0; j < QT; j++) {
          acc[j][c][0] = Mfma<R>::mma(ce[j], pe, acc[j][c][0]);
          acc[j][c][1] = Mfma<R>::mma(co[j], po, acc[j][c][1]);
          acc[j][c][2] = Mfma<R>::mma(se[j], qe, acc[j][c][2]);
          acc[j][c][3] = Mfma<R>::mma(so[j], qo, acc[j][c][3]);
        }
      }
    }
    if (!more) break;
    __syncthreads();                                 // every wave is done with the chunk
#pragma unroll
    for (int s = 0; s < kPre; s++)
      if (s < P.nslot) put(s, k0 + kKC, pr[s], pi[s]);
    for (int s = kPre; s < P.nslot; s++) {
      R pv, qv;
      load(s, k0 + kKC, pv, qv);
      put(s, k0 + kKC, pv, qv);
    }
    __syncthreads();
  }

  // result row of lane in a tile: Mfma<R>::row(lane, rr); quad of tile j: qbase + (qg * QT + j) * 16 + row
  if (!rlz) {
    // RL: write the four points of every (quad, column) straight to physical
    const int slot_of[5] = {0, 3, 4, 1, 2};
#pragma unroll
    for (int c = 0; c < CT; c++) {
      const int col = colb + 16 * c;
      if (col >= P.ncols) continue;
      const int g = int(col >= ntb) + int(col >= 2 * ntb) + int(col >= 3 * ntb) + int(col >= 4 * ntb);
      const int f = fam_of_group(g);
      const int v = col - g * ntb;
      const int d = f < 3 ? 0 : f - 2;
      const double dc = (f == 1 || f == 2) ? 0.0 : double(Xr[(d * ntb + v) * kl]);
      R* out = phys + (long(slot_of[f]) * T.nvars + v) * T.npts + T.ringOff[r];
#pragma unroll
      for (int j = 0; j < QT; j++)
#pragma unroll
        for (int rr = 0; rr < 4; rr++) {
          const int l = qbase + (qg * QT + j) * 16 + Mfma<R>::row(lane, rr);
          if (l > N4) continue;
          const double pce = acc[j][c][0][rr], pco = acc[j][c][1][rr], qse = acc[j][c][2][rr], qso = acc[j][c][3][rr];
          out[l] = R(dc + 2.0 * (pce + pco + qse + qso));
          if (l < N4) out[N / 2 - l] = R(dc + 2.0 * (pce - pco - qse + qso));
          if (l > 0) out[N / 2 + l] = R(dc + 2.0 * (pce - pco + qse - qso));
          if (l > 0 && l < N4) out[N - l] = R(dc + 2.0 * (pce + pco - qse - qso));
        }
    }
    return;
  }

  // RLZ epilogue: per family, stage the synthesised values and contract with
  // the inverse Chebyshev matrices.  Two staging tiles alternate between
  // families, so one barrier per family separates the writes of a tile from
  // the contraction that reads it (the barrier of family f + 1 also orders the
  // contraction of f - 1 before the rewrite of its tile at f + 1).
  const int bk = T.nzb, kDim = T.kDim;
  const int npl = 4 * P.qpb;                         // local points: which * qpb + ql
  __syncthreads();                                   // tw / staging no longer needed
  R* Ys = sm;                                        // [2][npl][ys]
  R* Mt = Ys + 2 * npl * P.ys;                       // [3][bkp][kzp]: Mt[m][zb][z] = M_m[z][zb]
  R* dcs = Mt + 3 * P.bkp * P.kzp;                   // [3][bk]
  {
    // the table is stored padded in this layout: a straight copy, four loads in flight per thread
    const R* Mg = tb.invt + long(vb) * 3 * P.bkp * P.kzp;
    const int nmt = 3 * P.bkp * P.kzp;
    for (int e0 = threadIdx.x; e0 < nmt; e0 += 4 * nthr) {
      R m4[4];
#pragma unroll
      for (int u = 0; u < 4; u++) m4[u] = e0 + u * nthr < nmt ? Mg[e0 + u * nthr] : R(0);
#pragma unroll
      for (int u = 0; u < 4; u++)
        if (e0 + u * nthr < nmt) Mt[e0 + u * nthr] = m4[u];
    }
  }
  for (int e = threadIdx.x; e < 3 * bk; e += nthr) dcs[e] = Xr[e * kl];   // k = 0: P of (d, zb)
  // columns bk <= zb < bkp are read by the contraction and never written: zero them in both tiles
  const int npad = P.bkp - bk;
  for (int e = threadIdx.x; e < 2 * npl * npad; e += nthr) {
    const int row = e / npad;
    Ys[row * P.ys + bk + (e - row * npad)] = R(0);
  }
  // this lane's family and mode for each owned column tile (-1: padding column)
  int fam[CT], zbc[CT];
#pragma unroll
  for (int c = 0; c < CT; c++) {
    const int col = colb + 16 * c;
    const int g = int(col >= ntb) + int(col >= 2 * ntb) + int(col >= 3 * ntb) + int(col >= 4 * ntb);
    fam[c] = col < P.ncols ? fam_of_group(g) : -1;
    zbc[c] = col - g * ntb;
  }
  __syncthreads();
  const int ztiles = P.kzp >> 4, ntiles = (npl >> 4) * ztiles;
  const int nwaves = nthr >> 6;
  R* physv = phys + long(vb) * T.npts;
  for (int f = 0; f < 5; f++) {
    R* Y = Ys + (f & 1) * npl * P.ys;
#pragma unroll
    for (int c = 0; c < CT; c++) {
      if (fam[c] != f) continue;
      const int zb = zbc[c];
      const int d = f < 3 ? 0 : f - 2;
      const R dc = (f == 1 || f == 2) ? R(0) : dcs[d * bk + zb];
#pragma unroll
      for (int j = 0; j < QT; j++)
#pragma unroll
        for (int rr = 0; rr < 4; rr++) {
          const int ql = (qg * QT + j) * 16 + Mfma<R>::row(lane, rr);
          const R pce = acc[j][c][0][rr], pco = acc[j][c][1][rr], qse = acc[j][c][2][rr], qso = acc[j][c][3][rr];
          Y[(0 * P.qpb + ql) * P.ys + zb] = dc + R(2) * (pce + pco + qse + qso);   // l
          Y[(1 * P.qpb + ql) * P.ys + zb] = dc + R(2) * (pce - pco - qse + qso);   // N/2 - l
          Y[(2 * P.qpb + ql) * P.ys + zb] = dc + R(2) * (pce - pco + qse - qso);   // N/2 + l
          Y[(3 * P.qpb + ql) * P.ys + zb] = dc + R(2) * (pce + pco - qse - qso);   // N - l
        }
    }
    __syncthreads();
    // family 0 feeds slots 1, 6, 7 (M0, M1, M2); families 1-4 feed slots 4, 5, 2, 3 (M0)
    const int nsl = f == 0 ? 3 : 1;
    for (int si = 0; si < nsl; si++) {
      const int slot = f == 0 ? (si == 0 ? 0 : si == 1 ? 5 : 6) : (f == 1 ? 3 : f == 2 ? 4 : f == 3 ? 1 : 2);
      const R* Mm = Mt + si * P.bkp * P.kzp;
      R* outs = physv + long(slot) * T.nvars * T.npts + long(T.ringOff[r]) * kDim;
      for (int t = w; t < ntiles; t += nwaves) {
        const int pt = t / ztiles, zt = t - pt * ztiles;
        const R* ya = Y + ((pt << 4) + (lane & 15)) * P.ys + (lane >> 4);
        const R* ma = Mm + (lane >> 4) * P.kzp + (zt << 4) + (lane & 15);
        v4 o = v4{0, 0, 0, 0};
#pragma unroll 2
        for (int kb = 0; kb < P.bkp; kb += 4) o = Mfma<R>::mma(ya[kb], ma[kb * P.kzp], o);
        const int z = (zt << 4) + (lane & 15);
        if (z >= kDim) continue;
        // a point tile lies in one group which = (pt * 16) / qpb (uniform):
        // point = base + sgn * l for l in [lmin, lmax]
        const int which = (pt << 4) >> P.lqpb;
        const int ql0 = (pt << 4) - (which << P.lqpb);
        const int lmin = which >= 2 ? 1 : 0, lmax = (which & 1) ? N4 - 1 : N4;
        const int base = which == 0 ? 0 : which == 3 ? N : N / 2;
        const int sgn = (which & 1) ? -1 : 1;
#pragma unroll
        for (int rr = 0; rr < 4; rr++) {
          const int l = qbase + ql0 + Mfma<R>::row(lane, rr);
          if (l >= lmin && l <= lmax) outs[long(base + sgn * l) * kDim + z] = o[rr];
        }
      }
    }
  }
}

template <typename R, int CT, int QT>
__global__ void __launch_bounds__(512, 4) k_ring_syn(DevTables T, SynLaunch P, const R* __restrict__ X,
                                                      R* __restrict__ phys) {
  ring_syn_body<R, CT, QT>(T, P, int(blockIdx.x), X, phys);
}

template <typename R, int CT, int QT>
__global__ void __launch_bounds__(512, 4) k_ring_syn_batch(const SynJob<R>* __restrict__ jobs,
                                                            const int* __restrict__ start, int njobs) {
  int lin;
  const SynJob<R>& j = jobs[batch_job(start, njobs, lin)];
  ring_syn_body<R, CT, QT>(j.T, j.P, lin, j.X, j.phys);
}

SynLaunch plan_syn(const DevTables& T, size_t esize) {
  SynLaunch P{};
  const bool rlz = T.geometry == 3;
  P.ntb = T.ntb;
  P.ncols = 5 * P.ntb;
  const int nct = (P.ncols + 15) / 16;
  // each wave: one column tile and two quad tiles (default up to 8 column
  // tiles), else two column tiles and one quad tile; the tuner may force either
  // shape 3 (two column tiles x two quad tiles) only for FP32, whose accumulators take half the registers;
  // without tuning: shape 1 for one or two column tiles, else shape 2 (the tuner's usual choice)
  const int shape = (T.tune.syn_ct == 1 || T.tune.syn_ct == 2 || (T.tune.syn_ct == 3 && esize == sizeof(float)))
                        ? T.tune.syn_ct : (nct <= 2 ? 1 : 2);
  if (shape == 1) { P.ct = 1; P.qt = 2; } else if (shape == 2) { P.ct = 2; P.qt = 1; } else { P.ct = 2; P.qt = 2; }
  P.nwc = (nct + P.ct - 1) / P.ct;
  P.nctp = P.nwc * P.ct;
  P.ldp = P.nctp * 16 + 2;
  P.bkp = (T.nzb + 3) / 4 * 4;
  P.kzp = rlz ? (T.kDim + 15) / 16 * 16 : 16;
  P.ys = P.bkp + 1;
  const size_t maxN = size_t(4 + 4 * T.K);
  const size_t phaseA = (2 * maxN + 2 * kKC * size_t(P.ldp)) * esize;
  // two blocks per CU: at most 512 threads and 80 KB of LDS each
  const int nwq0 = T.tune.syn_nwq == 1 || T.tune.syn_nwq == 2 || T.tune.syn_nwq == 4 ? T.tune.syn_nwq
                                                                                     : (shape == 1 ? 4 : 2);
  for (P.nwq = nwq0; P.nwq >= 1; P.nwq /= 2) {
    P.qpb = 16 * P.qt * P.nwq;
    P.lqpb = 31 - __builtin_clz(unsigned(P.qpb));
    size_t phaseB = 0;
    if (rlz) phaseB = (2 * size_t(4 * P.qpb) * P.ys + 3 * size_t(P.bkp) * P.kzp + 3 * size_t(T.nzb)) * esize;
    P.lds = phaseA > phaseB ? phaseA : phaseB;
    if (P.lds <= 80 * 1024 && 64 * P.nwq * P.nwc <= 512) break;
    if (P.nwq == 1) break;
  }
  const int nthr = 64 * P.nwq * P.nwc;
  P.nthr = nthr;
  P.nslot = (kKC * 3 * P.ntb + nthr - 1) / nthr;
  P.nqt = (T.K + 2 + P.qpb - 1) / P.qpb;         // quad tiles of the largest ring
  P.nblocks = int(xcd_unit_grid(P.nqt, T.iDim * T.nvb));
  P.ok = nthr <= 512 && P.lds <= 160 * 1024;
  return P;
}

template <typename R>
void launch_ring_syn(const DevTables& T, const R* X, R* phys, hipStream_t st) {
  const SynLaunch P = plan_syn(T, sizeof(R));
  const dim3 grid(P.nblocks), block(P.nthr);
  if (P.ct == 1) {
    hipLaunchKernelGGL((k_ring_syn<R, 1, 2>), grid, block, P.lds, st, T, P, X, phys);
  } else if (P.qt == 1) {
    hipLaunchKernelGGL((k_ring_syn<R, 2, 1>), grid, block, P.lds, st, T, P, X, phys);
  } else if constexpr (sizeof(R) == sizeof(float)) {
    hipLaunchKernelGGL((k_ring_syn<R, 2, 2>), grid, block, P.lds, st, T, P, X, phys);
  }
}
template void launch_ring_syn<double>(const DevTables&, const double*, double*, hipStream_t);
template void launch_ring_syn<float>(const DevTables&, const float*, float*, hipStream_t);

template <typename R>
void launch_ring_syn_batch(const SynJob<R>* jobs, const int* start, int njobs, int nblocks, int ct, int nthr,
                           size_t lds, hipStream_t st) {
  if (ct == 1) {
    hipLaunchKernelGGL((k_ring_syn_batch<R, 1, 2>), dim3(nblocks), dim3(nthr), lds, st, jobs, start, njobs);
  } else if (ct == 2) {
    hipLaunchKernelGGL((k_ring_syn_batch<R, 2, 1>), dim3(nblocks), dim3(nthr), lds, st, jobs, start, njobs);
  } else if constexpr (sizeof(R) == sizeof(float)) {
    hipLaunchKernelGGL((k_ring_syn_batch<R, 2, 2>), dim3(nblocks), dim3(nthr), lds, st, jobs, start, njobs);
  }
}
template void launch_ring_syn_batch<double>(const SynJob<double>*, const int*, int, int, int, int, size_t, hipStream_t);
template void launch_ring_syn_batch<float>(const SynJob<float>*, const int*, int, int, int, int, size_t, hipStream_t);

}  // namespace ssg
