// ssg_ring_ana.hip -- forward ring DFT (R2HC / N + forward phase filter) as
// FP64 MFMA GEMMs, fused for RLZ with the forward Chebyshev stage.
//
// For ring r (N = 4 + 4 ri points) the pruned real DFT is
//   re_k = sum_j x_j cos(2 pi jk/N),  im_k = -sum_j x_j sin(2 pi jk/N),  k <= kmax
// (FBtransform!, src/Fourier.jl:388-395, FFTW R2HC definition).  The points of
// quad q (q, N/2 - q, N/2 + q, N - q) share |cos|, |sin|, so with the folded
// inputs
//   E(+/-)_q = (x_q + x_{N-q}) +/- (x_{N/2-q} + x_{N/2+q})
//   O(+/-)_q = (x_q - x_{N-q}) +/- (x_{N/2+q} - x_{N/2-q})
// (q = 0: E = x_0 +/- x_{N/2}, O = 0; q = N/4: E = O = x_{N/4} +/- x_{3N/4})
// one gets re_k = sum_q cos(k phi_q) E(par k)_q and im_k = -sum_q sin(k phi_q)
// O(par k)_q, phi_q = 2 pi q / N: a GEMM [k rows of one parity x quads] x
// [quads x columns] on v_mfma_f64_16x16x4f64 with A = cos / sin from the
// ring's cos table in LDS.  Then s = (re, im) / N is rotated by -k ymin
// (apply_phasefilter_forward!, src/phase_filter.jl:105-143) and stored in the
// reference's b layout (DC, re 1..ri, im ri..1, zero above kmax).
//
// RLZ: a block covers one variable; each chunk of 16 quads (64 physical
// columns) is first contracted with the variable's (b_kDim x kDim) forward
// Chebyshev matrix (CBtransform!, src/Chebyshev.jl:417-426; a second MFMA
// GEMM), so the columns of the ring GEMM are the Chebyshev modes
// (src/transforms_cylindrical.jl:541-566).  RL: columns are the variables.
// Large rings split their k rows over several blocks.
#include "ssg_kernels.hpp"

namespace ssg {

typedef double d4 __attribute__((ext_vector_type(4)));

static constexpr int kAnaWaves = 8;
static constexpr int kQC = 16;        // quads per chunk (four k-steps of 4 quads)

struct AnaLaunch {
  int ncols;      // RLZ b_kDim, RL nvars
  int nct;        // column tiles
  int ldc;        // row stride of the folded tiles (doubles)
  int kzp;        // kDim padded to a multiple of 4 (RLZ)
  int ldu;        // row stride of the staged physical tile (doubles)
  int nkb_max;    // k-blocks of the largest ring
  size_t lds;
};

__device__ __forceinline__ d4 mfma64a(double a, double b, d4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

template <int NCT>
__global__ void __launch_bounds__(64 * kAnaWaves) k_ring_ana(DevTables T, AnaLaunch P, const double* __restrict__ phys,
                                                             double* __restrict__ rc) {
  constexpr int MR = 8 / NCT;                        // row tiles per wave
  extern __shared__ double sm[];
  const int r = T.iDim - 1 - int(blockIdx.y);       // largest rings first
  const int N = T.ringN[r], ri = T.ringRi[r], N4 = N / 4, N2 = N / 2;
  const bool rlz = T.geometry == 3;
  const int v_blk = rlz ? int(blockIdx.z) : 0;
  int km = 0;
  if (rlz) km = T.kmax[v_blk * T.iDim + r];
  else for (int v = 0; v < T.nvars; v++) km = max(km, T.kmax[v * T.iDim + r]);
  const int ne = (km / 2 + 1 + 15) / 16;            // even-k row tiles (k = 0, 2, ...)
  const int no = ((km + 1) / 2 + 15) / 16;          // odd-k row tiles
  const int nrt = ne + no;
  const int per_block = kAnaWaves * MR;
  const int nkb = (nrt + per_block - 1) / per_block;
  if (int(blockIdx.x) >= nkb) return;               // block-uniform
  const int rt0 = int(blockIdx.x) * nrt / nkb, rt1 = (int(blockIdx.x) + 1) * nrt / nkb;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int nthr = blockDim.x;
  const int ntr = T.nvars * T.nzb;
  const int t0 = rlz ? v_blk * T.nzb : 0;
  const int kDim = T.kDim;

  // LDS: tw[N] | u[64][ldu] | fw[kzp][ldc] | vals[64][ldc] | EO[4][16][ldc]
  double* tw = sm;
  double* ut = tw + (4 + 4 * T.K);
  double* fw = ut + 64 * P.ldu;
  double* vals = fw + P.kzp * P.ldc;
  double* EO = vals + 64 * P.ldc;
  const double* tws = T.twid + T.twOff[r];
  for (int m = threadIdx.x; m < N; m += nthr) tw[m] = tws[m];
  if (rlz) {
    const double* F = T.fwd + long(v_blk) * T.nzb * kDim;   // [zb][z]
    for (int e = threadIdx.x; e < P.kzp * P.ldc; e += nthr) {
      const int z = e / P.ldc, zb = e - z * P.ldc;
      fw[e] = (z < kDim && zb < T.nzb) ? F[zb * kDim + z] : 0.0;
    }
  }

  // rows of this wave: row tiles rt0 + w + j * kAnaWaves, j < MR
  int kr[MR];          // k of this lane's A row, per row tile (-1 = none)
  int idx[MR], inc[MR];
  bool par[MR];
#pragma unroll
  for (int j = 0; j < MR; j++) {
    const int rt = rt0 + w + j * kAnaWaves;
    kr[j] = -1; idx[j] = 0; inc[j] = 0; par[j] = false;
    if (rt < rt1) {
      const bool odd = rt >= ne;
      const int ti = odd ? rt - ne : rt;
      const int k = (odd ? 1 : 0) + 2 * (ti * 16 + (lane & 15));
      par[j] = odd;
      if (k <= km) {
        kr[j] = k;
        const int q = lane >> 4;                     // first quad of this lane (k-step 0)
        idx[j] = (k * q) % N;
        inc[j] = (4 * k) % N;
      }
    }
  }
  d4 acr[MR][NCT], aci[MR][NCT];
#pragma unroll
  for (int j = 0; j < MR; j++)
#pragma unroll
    for (int c = 0; c < NCT; c++) { acr[j][c] = d4{0.0, 0.0, 0.0, 0.0}; aci[j][c] = d4{0.0, 0.0, 0.0, 0.0}; }
  const int q3 = 3 * N4;
  const double* xbase = rlz ? phys + long(v_blk) * T.npts + long(T.ringOff[r]) * kDim : phys + T.ringOff[r];

  for (int q0 = 0; q0 <= N4; q0 += kQC) {
    __syncthreads();
    // local point pl = which * 16 + (q - q0): which 0: q, 1: N/2 - q, 2: N/2 + q, 3: N - q (mod N)
    if (rlz) {
      for (int e = threadIdx.x; e < 64 * kDim; e += nthr) {
        const int pl = e / kDim, z = e - pl * kDim;
        const int which = pl >> 4, q = q0 + (pl & 15);
        int pt = which == 0 ? q : which == 1 ? N2 - q : which == 2 ? N2 + q : N - q;
        pt = pt >= N ? pt - N : pt;
        ut[pl * P.ldu + z] = q <= N4 ? xbase[long(pt) * kDim + z] : 0.0;
      }
      __syncthreads();
      // vals[pl][zb] = sum_z u[pl][z] F[zb][z]: 4 point tiles x nct column tiles
      for (int tile = w; tile < 4 * P.nct; tile += kAnaWaves) {
        const int pt4 = tile / P.nct, ct = tile - pt4 * P.nct;
        d4 o = d4{0.0, 0.0, 0.0, 0.0};
        for (int k4 = 0; k4 < P.kzp; k4 += 4) {
          const int z = k4 + (lane >> 4);
          const double a = z < kDim ? ut[(pt4 * 16 + (lane & 15)) * P.ldu + z] : 0.0;
          const double b = fw[z * P.ldc + ct * 16 + (lane & 15)];
          o = mfma64a(a, b, o);
        }
#pragma unroll
        for (int rr = 0; rr < 4; rr++)
          vals[(pt4 * 16 + (lane >> 4) + 4 * rr) * P.ldc + ct * 16 + (lane & 15)] = o[rr];
      }
    } else {
      for (int e = threadIdx.x; e < 64 * P.nct * 16; e += nthr) {
        const int pl = e / (P.nct * 16), col = e - pl * (P.nct * 16);
        const int which = pl >> 4, q = q0 + (pl & 15);
        int pt = which == 0 ? q : which == 1 ? N2 - q : which == 2 ? N2 + q : N - q;
        pt = pt >= N ? pt - N : pt;
        vals[pl * P.ldc + col] = (q <= N4 && col < T.nvars) ? xbase[long(col) * T.npts + pt] : 0.0;
      }
    }
    __syncthreads();
    // fold: EO[combo][ql][col], combo 0 = E+, 1 = E-, 2 = O+, 3 = O-
    for (int e = threadIdx.x; e < 16 * P.nct * 16; e += nthr) {
      const int ql = e / (P.nct * 16), col = e - ql * (P.nct * 16);
      const int q = q0 + ql;
      const double xa = vals[(0 * 16 + ql) * P.ldc + col];   // x_q
      const double xb = vals[(1 * 16 + ql) * P.ldc + col];   // x_{N/2-q}
      const double xc = vals[(2 * 16 + ql) * P.ldc + col];   // x_{N/2+q}
      const double xd = vals[(3 * 16 + ql) * P.ldc + col];   // x_{N-q}
      const bool first = q == 0, quarter = q == N4, out = q > N4;
      const double s1 = xa + xd, s2 = xb + xc, d1 = xa - xd, d2 = xc - xb;
      // q = 0: x_0 +/- x_{N/2} and no sine part; q = N/4: x_{N/4} +/- x_{3N/4} for both
      const double ep = out ? 0.0 : first ? xa + xb : quarter ? xa + xc : s1 + s2;
      const double em = out ? 0.0 : first ? xa - xb : quarter ? xa - xc : s1 - s2;
      const double op = (out || first) ? 0.0 : quarter ? xa + xc : d1 + d2;
      const double om = (out || first) ? 0.0 : quarter ? xa - xc : d1 - d2;
      EO[(0 * 16 + ql) * P.ldc + col] = ep;
      EO[(1 * 16 + ql) * P.ldc + col] = em;
      EO[(2 * 16 + ql) * P.ldc + col] = op;
      EO[(3 * 16 + ql) * P.ldc + col] = om;
    }
    __syncthreads();
#pragma unroll
    for (int ks = 0; ks < kQC / 4; ks++) {
      const int qrow = ks * 4 + (lane >> 4);         // quad (within the chunk) of this lane's B row
#pragma unroll
      for (int j = 0; j < MR; j++) {
        if (rt0 + w + j * kAnaWaves >= rt1) continue;   // wave-uniform
        double ca = 0.0, sa = 0.0;
        if (kr[j] >= 0) {
          int is = idx[j] + q3;
          if (is >= N) is -= N;
          ca = tw[idx[j]];
          sa = tw[is];
        }
        idx[j] += inc[j];
        if (idx[j] >= N) idx[j] -= N;
        const int ce = par[j] ? 1 : 0;               // E- / O- for odd k
#pragma unroll
        for (int c = 0; c < NCT; c++) {
          const int col = c * 16 + (lane & 15);
          const double be = EO[(ce * 16 + qrow) * P.ldc + col];
          const double bo = EO[((2 + ce) * 16 + qrow) * P.ldc + col];
          acr[j][c] = mfma64a(ca, be, acr[j][c]);
          aci[j][c] = mfma64a(sa, bo, aci[j][c]);
        }
      }
    }
  }

  // epilogue: scale, rotate by -k ymin, store in the b layout
  const double invN = 1.0 / N;
  const int bDim = 2 * ri + 1;
  const int coef0 = T.coefOff[r];
  const double* pc = T.phaseC + T.phaseOff[r];
  const double* ps = T.phaseS + T.phaseOff[r];
#pragma unroll
  for (int j = 0; j < MR; j++) {
    const int rt = rt0 + w + j * kAnaWaves;
    if (rt >= rt1) continue;
    const bool odd = rt >= ne;
    const int ti = odd ? rt - ne : rt;
#pragma unroll
    for (int c = 0; c < NCT; c++) {
      const int col = c * 16 + (lane & 15);
      if (col >= P.ncols) continue;
      const int t = t0 + col;
      const int v = rlz ? v_blk : col;
      const int kmv = T.kmax[v * T.iDim + r];
      double* out = rc + long(t) * T.b_jDim + coef0;
#pragma unroll
      for (int rr = 0; rr < 4; rr++) {
        const int k = (odd ? 1 : 0) + 2 * (ti * 16 + (lane >> 4) + 4 * rr);
        if (k > kmv) continue;
        const double re = acr[j][c][rr] * invN, im = -aci[j][c][rr] * invN;
        if (k == 0) { out[0] = re; continue; }
        const double cc = pc[k - 1], ss = ps[k - 1];
        out[k] = cc * re + ss * im;
        out[bDim - k] = -ss * re + cc * im;
      }
    }
  }
  // zero-fill kmax < k <= ri (only the first k-block)
  if (blockIdx.x == 0) {
    for (int e = threadIdx.x; e < P.ncols * ri; e += nthr) {
      const int col = e / ri, k = 1 + e - col * ri;
      const int v = rlz ? v_blk : col;
      if (k > T.kmax[v * T.iDim + r]) {
        double* out = rc + long(t0 + col) * T.b_jDim + coef0;
        out[k] = 0.0;
        out[bDim - k] = 0.0;
      }
    }
  }
}

static AnaLaunch plan_ana(const DevTables& T) {
  AnaLaunch P{};
  const bool rlz = T.geometry == 3;
  P.ncols = rlz ? T.nzb : T.nvars;
  P.nct = (P.ncols + 15) / 16;
  P.ldc = P.nct * 16 + 2;
  P.kzp = rlz ? (T.kDim + 3) / 4 * 4 : 4;
  P.ldu = rlz ? T.kDim + 1 : 1;
  const int MR = 8 / P.nct;
  const int km = T.K;                          // largest ring, uncapped bound
  const int nrt = (km / 2 + 1 + 15) / 16 + ((km + 1) / 2 + 15) / 16;
  P.nkb_max = (nrt + kAnaWaves * MR - 1) / (kAnaWaves * MR);
  const size_t dbl = size_t(4 + 4 * T.K) + 64 * size_t(P.ldu) + size_t(P.kzp) * P.ldc + 64 * size_t(P.ldc) +
                     64 * size_t(P.ldc);
  P.lds = dbl * sizeof(double);
  return P;
}

bool ring_ana_supported(const DevTables& T) { return plan_ana(T).nct <= 3; }

void launch_ring_ana(const DevTables& T, const double* phys, double* rc, hipStream_t st) {
  const AnaLaunch P = plan_ana(T);
  dim3 grid(P.nkb_max, T.iDim, T.geometry == 3 ? T.nvars : 1);
  dim3 block(64 * kAnaWaves);
  switch (P.nct) {
    case 1: hipLaunchKernelGGL(k_ring_ana<1>, grid, block, P.lds, st, T, P, phys, rc); break;
    case 2: hipLaunchKernelGGL(k_ring_ana<2>, grid, block, P.lds, st, T, P, phys, rc); break;
    default: hipLaunchKernelGGL(k_ring_ana<3>, grid, block, P.lds, st, T, P, phys, rc); break;
  }
}

}  // namespace ssg
