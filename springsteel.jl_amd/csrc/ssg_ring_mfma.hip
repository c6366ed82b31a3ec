// ssg_ring_mfma.hip -- ring synthesis (inverse phase filter + HC2R with
// azimuthal derivatives) as FP64 MFMA GEMMs, fused for RLZ with the inverse
// Chebyshev stage.
//
// Ring r has N = 4 + 4 ri points at phi_l = 2 pi l / N.  After the inverse
// phase rotation (src/phase_filter.jl:154-190) every output of one family is
//   y(phi) = dc + 2 sum_{k=1}^{kmax} [P_k cos(k phi) + Q_k sin(k phi)]
// (HC2R, src/Fourier.jl:463-468), with (P, Q) = (a_re, -a_im) for values,
// (-k a_im, -k a_re) for d/dlambda and (-k^2 a_re, k^2 a_im) for d2/dlambda2
// (FIxcoefficients, src/Fourier.jl:495-503).  The four points l, N/2 - l,
// N/2 + l, N - l share |cos|, |sin|; splitting k by parity gives all four from
// PCe, PCo (sum P cos over even / odd k) and QSe, QSo:
//   y(l)       = dc + 2 (PCe + PCo + QSe + QSo)
//   y(N - l)   = dc + 2 (PCe + PCo - QSe - QSo)
//   y(N/2 - l) = dc + 2 (PCe - PCo - QSe + QSo)
//   y(N/2 + l) = dc + 2 (PCe - PCo + QSe - QSo)
// Each partial sum is a GEMM [quads x k] x [k x columns]: A = cos/sin(k phi_l)
// generated from the ring's cos table in LDS, B = (P, Q) staged in LDS,
// accumulated with v_mfma_f64_16x16x4f64.  A column is one (family, transform)
// pair, family f: 0 = values (d0), 1 = d/dlambda (d0), 2 = d2/dlambda2 (d0),
// 3 = d/dr values (d1), 4 = d2/dr2 values (d2).
//
// RL: each block covers all variables; outputs go straight to physical slots
// 1, 4, 5, 2, 3.  RLZ: a block covers one variable and all Chebyshev modes;
// per family the synthesised values are staged in LDS and contracted with the
// variable's (kDim x b_kDim) inverse Chebyshev matrices (a second MFMA GEMM)
// into physical slots: f0 -> 1 (M0), 6 (M1), 7 (M2); f1 -> 4; f2 -> 5;
// f3 -> 2; f4 -> 3 (src/transforms_cylindrical.jl:780-844).
#include "ssg_kernels.hpp"

namespace ssg {

typedef double d4 __attribute__((ext_vector_type(4)));

static constexpr int kMaxCT = 2;     // column tiles per wave
static constexpr int kKC = 16;       // k values staged per chunk (two k-steps of 8)
static constexpr int kPre = 2;       // staged elements per thread held in registers across a chunk

struct SynLaunch {
  int ntb;        // transforms per block: RL nvars, RLZ b_kDim
  int ncols;      // 5 * ntb
  int nct;        // column tiles of 16
  int ldp;        // staged row stride (doubles)
  int nwq;        // quad sub-tiles (16 quads each) per block
  int nwc;        // wave groups over column tiles
  int qpb;        // quads per block = 16 * nwq
  int bkp;        // b_kDim padded to a multiple of 4 (RLZ)
  int kzp;        // kDim padded to a multiple of 16 (RLZ)
  int ys;         // row stride of the synthesised-value tile (doubles)
  int nslot;      // staged elements per thread per chunk
  size_t lds;     // bytes of dynamic LDS
};

__device__ __forceinline__ d4 mfma64(double a, double b, d4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

__global__ void __launch_bounds__(512, 4) k_ring_syn(DevTables T, SynLaunch P, const double* __restrict__ rcd,
                                                   double* __restrict__ phys) {
  extern __shared__ double sm[];
  const int r = T.iDim - 1 - int(blockIdx.y);       // largest rings first
  const int N = T.ringN[r], ri = T.ringRi[r], N4 = N / 4;
  const int qbase = int(blockIdx.x) * P.qpb;
  if (qbase > N4) return;                            // block-uniform
  const bool rlz = T.geometry == 3;
  const int v_blk = rlz ? int(blockIdx.z) : 0;
  const int t0 = rlz ? v_blk * T.nzb : 0;            // first transform of the block
  const int ntr = T.nvars * T.nzb;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int qs = w % P.nwq, cg = w / P.nwq;
  const int bDim = 2 * ri + 1;
  const int coef0 = T.coefOff[r];
  const int nthr = blockDim.x;

  double* tw = sm;                                   // [N] cos(2 pi m / N)
  double* phc = sm + (4 + 4 * T.K);                  // [K] cos(k ymin), k = 1..ri
  double* phs = phc + T.K;                           // [K] sin(k ymin)
  int* kml = reinterpret_cast<int*>(phs + T.K);      // [ntb] wavenumber cap per transform
  double* stP = phs + T.K + (P.ntb + 1) / 2;         // [kKC][ldp]
  double* stQ = stP + kKC * P.ldp;
  {
    const double* tws = T.twid + T.twOff[r];
    for (int m = threadIdx.x; m < N; m += nthr) tw[m] = tws[m];
    const double* pc = T.phaseC + T.phaseOff[r];
    const double* ps = T.phaseS + T.phaseOff[r];
    for (int k = threadIdx.x; k < ri; k += nthr) { phc[k] = pc[k]; phs[k] = ps[k]; }
    for (int tl = threadIdx.x; tl < P.ntb; tl += nthr) kml[tl] = T.kmax[(rlz ? v_blk : tl) * T.iDim + r];
  }
  // block-wide maximum wavenumber over the variables it covers
  int kmb = 0;
  if (rlz) kmb = T.kmax[v_blk * T.iDim + r];
  else for (int v = 0; v < T.nvars; v++) kmb = max(kmb, T.kmax[v * T.iDim + r]);
  __syncthreads();

  // Staging of one chunk: element e = threadIdx.x + s * nthr is column c = e / kKC,
  // wavenumber k = k0 + e % kKC (consecutive threads read consecutive k).
  const int nel = kKC * P.nct * 16;
  auto fam = [&](int c) { return int(c >= P.ntb) + int(c >= 2 * P.ntb) + int(c >= 3 * P.ntb) + int(c >= 4 * P.ntb); };
  auto raw = [&](int e, int k0, double& br, double& bi) {
    br = 0.0; bi = 0.0;
    const int c = e / kKC, k = k0 + (e - (e / kKC) * kKC);
    if (e < nel && c < P.ncols && k >= 1) {
      const int f = fam(c), tl = c - f * P.ntb;
      if (k <= kml[tl]) {
        const int d = f < 3 ? 0 : f - 2;
        const double* b = rcd + (long(d) * ntr + t0 + tl) * T.b_jDim + coef0;
        br = b[k]; bi = b[bDim - k];
      }
    }
  };
  auto put = [&](int e, int k0, double br, double bi) {
    if (e >= nel) return;
    const int c = e / kKC, kk = e - c * kKC, k = k0 + kk;
    double pv = 0.0, qv = 0.0;
    if (c < P.ncols && k >= 1 && k <= ri) {
      const int f = fam(c);
      const double cc = phc[k - 1], ss = phs[k - 1];
      const double are = cc * br - ss * bi, aim = ss * br + cc * bi;
      // family weights as selects: hipcc 7.2 structurises an if / else-if /
      // else chain here so that the final else runs only for f <= 1 lanes
      const double kd = double(k), k2 = kd * kd;
      const bool d1 = f == 1, d2 = f == 2;
      pv = d1 ? -kd * aim : (d2 ? -k2 * are : are);
      qv = d1 ? -kd * are : (d2 ? k2 * aim : -aim);
    }
    stP[kk * P.ldp + c] = pv;
    stQ[kk * P.ldp + c] = qv;
  };
  for (int s = 0; s < P.nslot; s++) {
    double br, bi;
    raw(threadIdx.x + s * nthr, 0, br, bi);
    put(threadIdx.x + s * nthr, 0, br, bi);
  }

  d4 acc[kMaxCT][4];
#pragma unroll
  for (int c = 0; c < kMaxCT; c++)
#pragma unroll
    for (int s = 0; s < 4; s++) acc[c][s] = d4{0.0, 0.0, 0.0, 0.0};

  // A operand: row = quad q of this lane, k-group g = lane / 16 holds k = kb + 2g
  // (even) and kb + 2g + 1 (odd).  Table index (k q) mod N advances by 8q per k-step.
  const int q = qbase + qs * 16 + (lane & 15);
  const bool qok = q <= N4;
  const int qm = qok ? q : 0;
  const int q3 = 3 * N4;
  const int step8 = (8 * qm) % N;
  int ie = (2 * (lane >> 4) * qm) % N;
  __syncthreads();

  for (int k0 = 0; k0 <= kmb; k0 += kKC) {
    const bool more = k0 + kKC <= kmb;
    double pr[kPre > 0 ? kPre : 1], pi[kPre > 0 ? kPre : 1];
#pragma unroll
    for (int s = 0; s < kPre; s++) {
      pr[s] = 0.0; pi[s] = 0.0;
      if (more && s < P.nslot) raw(threadIdx.x + s * nthr, k0 + kKC, pr[s], pi[s]);
    }
#pragma unroll
    for (int ks = 0; ks < kKC / 8; ks++) {
      if (k0 + ks * 8 > kmb) break;                  // uniform
      int io = ie + qm; if (io >= N) io -= N;
      int je = ie + q3; if (je >= N) je -= N;        // sin(x) = cos(x - pi/2)
      int jo = io + q3; if (jo >= N) jo -= N;
      const double ce = qok ? tw[ie] : 0.0, se = qok ? tw[je] : 0.0;
      const double co = qok ? tw[io] : 0.0, so = qok ? tw[jo] : 0.0;
      ie += step8; if (ie >= N) ie -= N;
      const int kr = ks * 8 + 2 * (lane >> 4);       // staged row of the even k
#pragma unroll
      for (int c = 0; c < kMaxCT; c++) {
        const int ct = cg + c * P.nwc;
        if (ct < P.nct) {
          const int col = ct * 16 + (lane & 15);
          const double pe = stP[kr * P.ldp + col], po = stP[(kr + 1) * P.ldp + col];
          const double qe = stQ[kr * P.ldp + col], qo = stQ[(kr + 1) * P.ldp + col];
          acc[c][0] = mfma64(ce, pe, acc[c][0]);
          acc[c][1] = mfma64(co, po, acc[c][1]);
          acc[c][2] = mfma64(se, qe, acc[c][2]);
          acc[c][3] = mfma64(so, qo, acc[c][3]);
        }
      }
    }
    if (!more) break;
    __syncthreads();                                 // every wave is done with the chunk
#pragma unroll
    for (int s = 0; s < kPre; s++)
      if (s < P.nslot) put(threadIdx.x + s * nthr, k0 + kKC, pr[s], pi[s]);
    for (int s = kPre; s < P.nslot; s++) {
      double br, bi;
      raw(threadIdx.x + s * nthr, k0 + kKC, br, bi);
      put(threadIdx.x + s * nthr, k0 + kKC, br, bi);
    }
    __syncthreads();
  }

  // quad of result row i of this lane: i = (lane >> 4) + 4 * rr
  if (!rlz) {
    // RL: write the four points of every (quad, column) straight to physical
    const int slot_of[5] = {0, 3, 4, 1, 2};
#pragma unroll
    for (int c = 0; c < kMaxCT; c++) {
      const int ct = cg + c * P.nwc;
      if (ct >= P.nct) continue;
      const int col = ct * 16 + (lane & 15);
      if (col >= P.ncols) continue;
      const int f = col / P.ntb, tl = col - f * P.ntb, v = tl;
      const int d = f < 3 ? 0 : f - 2;
      const double dc = (f == 1 || f == 2) ? 0.0 : rcd[(long(d) * ntr + tl) * T.b_jDim + coef0];
      double* out = phys + (long(slot_of[f]) * T.nvars + v) * T.npts + T.ringOff[r];
#pragma unroll
      for (int rr = 0; rr < 4; rr++) {
        const int l = qbase + qs * 16 + (lane >> 4) + 4 * rr;
        if (l > N4) continue;
        const double pce = acc[c][0][rr], pco = acc[c][1][rr], qse = acc[c][2][rr], qso = acc[c][3][rr];
        out[l] = dc + 2.0 * (pce + pco + qse + qso);
        if (l < N4) out[N / 2 - l] = dc + 2.0 * (pce - pco - qse + qso);
        if (l > 0) out[N / 2 + l] = dc + 2.0 * (pce - pco + qse - qso);
        if (l > 0 && l < N4) out[N - l] = dc + 2.0 * (pce + pco - qse - qso);
      }
    }
    return;
  }

  // RLZ epilogue: per family, stage the synthesised values and contract with
  // the inverse Chebyshev matrices.
  const int bk = T.nzb, kDim = T.kDim;
  const int npl = 4 * P.qpb;                         // local points: which * qpb + ql
  __syncthreads();                                   // tw / staging no longer needed
  double* Y = sm;                                    // [npl][ys]
  double* Mt = Y + npl * P.ys;                       // [3][bkp][kzp]: Mt[m][zb][z] = M_m[z][zb]
  double* dcs = Mt + 3 * P.bkp * P.kzp;              // [3][bk]
  const double* Mg = T.inv + long(v_blk) * 3 * kDim * bk;
  for (int e = threadIdx.x; e < 3 * P.bkp * P.kzp; e += nthr) {
    const int m = e / (P.bkp * P.kzp), rem = e - m * P.bkp * P.kzp, zb = rem / P.kzp, z = rem - zb * P.kzp;
    Mt[e] = (zb < bk && z < kDim) ? Mg[(long(m) * kDim + z) * bk + zb] : 0.0;
  }
  for (int e = threadIdx.x; e < 3 * bk; e += nthr) {
    const int d = e / bk, zb = e - d * bk;
    dcs[e] = rcd[(long(d) * ntr + t0 + zb) * T.b_jDim + coef0];
  }
  const int ztiles = P.kzp / 16, ptiles = npl / 16;
  const int nwaves = nthr / 64;
  double* physv = phys + long(v_blk) * T.npts;
  for (int f = 0; f < 5; f++) {
    __syncthreads();
    // zero the tile (padding rows / columns must contribute nothing)
    for (int e = threadIdx.x; e < npl * P.ys; e += nthr) Y[e] = 0.0;
    __syncthreads();
#pragma unroll
    for (int c = 0; c < kMaxCT; c++) {
      const int ct = cg + c * P.nwc;
      if (ct >= P.nct) continue;
      const int col = ct * 16 + (lane & 15);
      if (col >= P.ncols || col / P.ntb != f) continue;
      const int zb = col - f * P.ntb;
      const int d = f < 3 ? 0 : f - 2;
      const double dc = (f == 1 || f == 2) ? 0.0 : dcs[d * bk + zb];
#pragma unroll
      for (int rr = 0; rr < 4; rr++) {
        const int ql = qs * 16 + (lane >> 4) + 4 * rr;
        const double pce = acc[c][0][rr], pco = acc[c][1][rr], qse = acc[c][2][rr], qso = acc[c][3][rr];
        Y[(0 * P.qpb + ql) * P.ys + zb] = dc + 2.0 * (pce + pco + qse + qso);   // l
        Y[(1 * P.qpb + ql) * P.ys + zb] = dc + 2.0 * (pce - pco - qse + qso);   // N/2 - l
        Y[(2 * P.qpb + ql) * P.ys + zb] = dc + 2.0 * (pce - pco + qse - qso);   // N/2 + l
        Y[(3 * P.qpb + ql) * P.ys + zb] = dc + 2.0 * (pce + pco - qse - qso);   // N - l
      }
    }
    __syncthreads();
    const int nsl = f == 0 ? 3 : 1;
    for (int si = 0; si < nsl; si++) {
      const int slot = f == 0 ? (si == 0 ? 0 : si == 1 ? 5 : 6) : (f == 1 ? 3 : f == 2 ? 4 : f == 3 ? 1 : 2);
      const int m = f == 0 ? si : 0;
      const double* Mm = Mt + m * P.bkp * P.kzp;
      double* outs = physv + long(slot) * T.nvars * T.npts + long(T.ringOff[r]) * kDim;
      for (int tile = w; tile < ptiles * ztiles; tile += nwaves) {
        const int pt = tile / ztiles, zt = tile - pt * ztiles;
        d4 o = d4{0.0, 0.0, 0.0, 0.0};
        for (int kb4 = 0; kb4 < P.bkp; kb4 += 4) {
          const int zb = kb4 + (lane >> 4);
          const double a = Y[(pt * 16 + (lane & 15)) * P.ys + zb];
          const double bb = Mm[zb * P.kzp + zt * 16 + (lane & 15)];
          o = mfma64(a, bb, o);
        }
        const int z = zt * 16 + (lane & 15);
        if (z >= kDim) continue;
#pragma unroll
        for (int rr = 0; rr < 4; rr++) {
          const int pl = pt * 16 + (lane >> 4) + 4 * rr;
          const int which = pl / P.qpb, ql = pl - which * P.qpb;
          const int l = qbase + ql;
          // point and validity as selects (see the staging note above)
          const bool ok = l <= N4 && (which == 0 || (which == 1 && l < N4) || (which == 2 && l > 0) ||
                                      (which == 3 && l > 0 && l < N4));
          const int point = which == 0 ? l : which == 1 ? N / 2 - l : which == 2 ? N / 2 + l : N - l;
          if (ok) outs[long(point) * kDim + z] = o[rr];
        }
      }
    }
  }
}

static SynLaunch plan_syn(const DevTables& T) {
  SynLaunch P{};
  const bool rlz = T.geometry == 3;
  P.ntb = rlz ? T.nzb : T.nvars;
  P.ncols = 5 * P.ntb;
  P.nct = (P.ncols + 15) / 16;
  P.ldp = P.nct * 16 + 2;
  P.nwc = (P.nct + kMaxCT - 1) / kMaxCT;
  P.bkp = (T.nzb + 3) / 4 * 4;
  P.kzp = rlz ? (T.kDim + 15) / 16 * 16 : 16;
  P.ys = P.bkp + 1;
  const size_t maxN = size_t(4 + 4 * T.K);
  const size_t phaseA = (maxN + 2 * size_t(T.K) + size_t(P.ntb + 1) / 2 + 2 * kKC * size_t(P.ldp)) * sizeof(double);
  // two blocks per CU: at most 512 threads and 80 KB of LDS each
  for (P.nwq = 4; P.nwq >= 1; P.nwq /= 2) {
    P.qpb = 16 * P.nwq;
    size_t phaseB = 0;
    if (rlz) phaseB = (size_t(4 * P.qpb) * P.ys + 3 * size_t(P.bkp) * P.kzp + 3 * size_t(T.nzb)) * sizeof(double);
    P.lds = phaseA > phaseB ? phaseA : phaseB;
    if (P.lds <= 80 * 1024 && 64 * P.nwq * P.nwc <= 512) break;
    if (P.nwq == 1) break;
  }
  const int nthr = 64 * P.nwq * P.nwc;
  P.nslot = (kKC * P.nct * 16 + nthr - 1) / nthr;
  return P;
}

size_t ring_syn_lds(const DevTables& T) { return plan_syn(T).lds; }

void launch_ring_syn(const DevTables& T, const double* rcd, double* phys, hipStream_t st) {
  const SynLaunch P = plan_syn(T);
  const int maxq = T.K + 2;                       // quads of the largest ring
  dim3 grid((maxq + P.qpb - 1) / P.qpb, T.iDim, T.geometry == 3 ? T.nvars : 1);
  dim3 block(64 * P.nwq * P.nwc);
  hipLaunchKernelGGL(k_ring_syn, grid, block, P.lds, st, T, P, rcd, phys);
}

}  // namespace ssg
