// ssg_ring_ana.hip -- forward ring DFT (R2HC / N + forward phase filter) as
// MFMA GEMMs; for RLZ preceded by the forward Chebyshev contraction of every
// physical column on MFMA.  Templated on the arithmetic type R: double (plain
// form, v_mfma_f64_16x16x4f64); the FP32 form runs the same FP64 instances on
// its FP32 physical input (the forward path feeds the d/dr and d2/dr2
// evaluations, which amplify rounding of its outputs by (L/DX)^2).
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
// (apply_phasefilter_forward!, src/phase_filter.jl:105-143) and stored in line
// order (DC, re 1, im 1, re 2, ...; zero above kmax): the reference's b layout
// (DC, re 1..ri, im ri..1) holds the same values, reordered so that the
// projection reads consecutive spectral lines as one run.
//
// Columns: RL the variables of physical slot 1; RLZ the Chebyshev modes of one
// variable, produced first by k_cheb_fwd_pts (CBtransform!,
// src/Chebyshev.jl:417-426, on every ring point: src/transforms_cylindrical.jl:
// 541-566) into cm[var][zb][point].  Large rings split their k rows over
// several blocks; each block streams all quads of its ring.
#include <algorithm>

#include "ssg_kernels.hpp"
#include "ssg_mfma.hpp"

namespace ssg {

static constexpr int kAnaWaves = 8;
static constexpr int kQC = 16;        // quads per chunk (four k-steps of 4 quads)


// ----------------------------------------------------------------------------
// Forward Chebyshev contraction of every ring point of the RLZ grid:
// cm[v][zb][p] = sum_z F_v[zb][z] u_v[p][z] (F = filter * DCT-I / (2(N-1)),
// composed on the host).  Each block walks tiles of PT points of one variable
// (64, or 32 for more than 32 heights, so that the tile and the matrix leave
// LDS for three blocks per CU): the next tile's inputs are loaded into registers
// while the current tile is contracted on MFMA from LDS, and each accumulator
// is stored straight to cm (16 consecutive points of four modes per store).
// ----------------------------------------------------------------------------
static constexpr int kPtThreads = 256;
__host__ __device__ inline int cheb_fwd_pt_tile(int kDim) { return kDim <= 32 ? 64 : 32; }

// blocks: nbx per variable, block (bx, v) walks tiles bx, bx + nbx, ...
// KL: input loads per thread per tile, PT / (256 / kDim): 8 (kDim <= 32 with PT 64,
// kDim <= 64 with PT 32)
// In: the type of physical (FP64, or FP32 storage in the FP32 form); the
// contraction and cm are FP64 in both forms.
template <typename In, int KL, int PT>
__device__ __forceinline__ void cheb_fwd_pts_body(const DevTables& T, const In* __restrict__ phys,
                                                  double* __restrict__ cm, int kzp, int bkp, int nbx, int lin) {
  typedef double R;
  typedef typename Mfma<R>::v4 v4;
  extern __shared__ __align__(16) unsigned char smraw[];
  R* sm = reinterpret_cast<R*>(smraw);
  const int v = lin / nbx, bx = lin - v * nbx;
  const int kDim = T.kDim, bk = T.nzb;
  const int ldu = kzp + 1;
  R* ut = sm;                                       // [PT][ldu]
  R* ft = ut + PT * ldu;                            // [kzp][bkp]: ft[z][zb] = F[zb][z]
  const R* F = rtab<R>(T).fwd + long(v) * bk * kDim;
  for (int e = threadIdx.x; e < kzp * bkp; e += blockDim.x) {
    const int z = e / bkp, zb = e - z * bkp;
    ft[e] = (z < kDim && zb < bk) ? F[zb * kDim + z] : R(0);
  }
  // padding z columns of the input tile stay zero
  for (int e = threadIdx.x; e < PT * ldu; e += blockDim.x) ut[e] = R(0);
  const int ntile = (T.jDim + PT - 1) / PT;
  const In* src = phys + long(v) * T.npts;
  R* dstv = cm + long(v) * T.jDim * bk;
  // thread (pt_, z_) loads input z_ of points pt_ + s * rpp: a pass of the
  // block reads rpp whole points, contiguous in memory.  Float input with an even
  // number of heights is loaded two heights per thread (8 bytes), so that a load
  // moves as many bytes as an FP64 one; the tile then takes KL / 2 passes.
  const bool two = sizeof(In) == sizeof(float) && (kDim & 1) == 0 && (reinterpret_cast<size_t>(src) & 7) == 0;
  const int zl = two ? kDim / 2 : kDim;             // loading threads per point
  const int rpp = kPtThreads / zl;
  const int pt_ = int(threadIdx.x) / zl, z_ = (int(threadIdx.x) - pt_ * zl) * (two ? 2 : 1);
  const bool act = pt_ < rpp;
  double xr[KL];
  auto load = [&](int tile) {
    const long p0 = long(tile) * PT;
    const int npl = int(min(long(PT), long(T.jDim) - p0));
    const In* s0 = src + p0 * kDim + z_;
    if constexpr (sizeof(In) == sizeof(float)) {
      if (two) {
        // heights z_ and z_ + 1 of point pl in xr[2 s] and xr[2 s + 1]
#pragma unroll
        for (int s = 0; s < KL / 2; s++) {
          const int pl = pt_ + s * rpp;
          float2 f = make_float2(0.0f, 0.0f);
          if (act && pl < npl) f = *reinterpret_cast<const float2*>(s0 + long(pl) * kDim);
          xr[2 * s] = f.x;
          xr[2 * s + 1] = f.y;
        }
        return;
      }
    }
#pragma unroll
    for (int s = 0; s < KL; s++) {
      const int pl = pt_ + s * rpp;
      xr[s] = act && pl < npl ? double(s0[long(pl) * kDim]) : 0.0;
    }
  };
  auto put = [&]() {
    if constexpr (sizeof(In) == sizeof(float)) {
      if (two) {
#pragma unroll
        for (int s = 0; s < KL / 2; s++) {
          const int pl = pt_ + s * rpp;
          if (act && pl < PT) {
            ut[pl * ldu + z_] = R(xr[2 * s]);
            ut[pl * ldu + z_ + 1] = R(xr[2 * s + 1]);
          }
        }
        return;
      }
    }
#pragma unroll
    for (int s = 0; s < KL; s++) {
      const int pl = pt_ + s * rpp;
      if (act && pl < PT) ut[pl * ldu + z_] = R(xr[s]);
    }
  };
  int tile = bx;
  if (tile < ntile) load(tile);
  __syncthreads();                                  // ut zeroed before the first put
  put();
  __syncthreads();
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int nzt = bkp / 16;
  for (; tile < ntile; tile += nbx) {
    const int next = tile + nbx;
    if (next < ntile) load(next);
    const long p0 = long(tile) * PT;
    const int npl = int(min(long(PT), long(T.jDim) - p0));
    R* dst = dstv + p0;
    // output tiles straight from the accumulators, modes as rows and points as
    // columns: each store writes 16 consecutive points of four modes
    for (int t = w; t < (PT / 16) * nzt; t += kPtThreads / 64) {
      const int pt = t / nzt, zt = t - pt * nzt;
      v4 o = v4{0, 0, 0, 0};
      for (int k4 = 0; k4 < kzp; k4 += 4) {
        const int z = k4 + (lane >> 4);
        o = Mfma<R>::mma(ft[z * bkp + zt * 16 + (lane & 15)], ut[(pt * 16 + (lane & 15)) * ldu + z], o);
      }
      const int pl = pt * 16 + (lane & 15);
#pragma unroll
      for (int rr = 0; rr < 4; rr++) {
        const int zb = zt * 16 + Mfma<R>::row(lane, rr);
        if (pl < npl && zb < bk) dst[long(zb) * T.jDim + pl] = o[rr];
      }
    }
    __syncthreads();                                // ut consumed
    if (next < ntile) put();
    __syncthreads();                                // ut refilled
  }
}

template <typename In, int KL, int PT>
__global__ void __launch_bounds__(kPtThreads) k_cheb_fwd_pts(DevTables T, const In* __restrict__ phys,
                                                             double* __restrict__ cm, int kzp, int bkp, int nbx) {
  cheb_fwd_pts_body<In, KL, PT>(T, phys, cm, kzp, bkp, nbx, int(blockIdx.x));
}

template <typename In>
__global__ void __launch_bounds__(kPtThreads) k_cheb_fwd_pts_batch(const ChebJob<In>* __restrict__ jobs,
                                                                   const int* __restrict__ start, int njobs) {
  int lin;
  const ChebJob<In>& j = jobs[batch_job(start, njobs, lin)];
  if (lin >= j.nbx * j.T.nvars) return;             // padding blocks of the job
  if (j.T.kDim <= 32) cheb_fwd_pts_body<In, 8, 64>(j.T, j.phys, j.cm, j.kzp, j.bkp, j.nbx, lin);
  else cheb_fwd_pts_body<In, 8, 32>(j.T, j.phys, j.cm, j.kzp, j.bkp, j.nbx, lin);
}

// about four tiles per block: many more blocks than fit at once, so the last
// round of blocks is a small fraction of the launch
static int cheb_fwd_pts_nbx(const DevTables& T) {
  const int pt = cheb_fwd_pt_tile(T.kDim);
  const int ntile = (T.jDim + pt - 1) / pt;
  return std::max(1, (ntile + 3) / 4);
}
int cheb_fwd_pts_blocks(const DevTables& T) { return cheb_fwd_pts_nbx(T) * T.nvars; }
size_t cheb_fwd_pts_lds(const DevTables& T, size_t esize) {
  const int kzp = (T.kDim + 3) / 4 * 4, bkp = (T.nzb + 15) / 16 * 16;
  return esize * (size_t(cheb_fwd_pt_tile(T.kDim)) * (kzp + 1) + size_t(kzp) * bkp);
}

template <typename In>
void launch_cheb_fwd_pts(const DevTables& T, const In* phys, double* cm, hipStream_t st) {
  const int kzp = (T.kDim + 3) / 4 * 4, bkp = (T.nzb + 15) / 16 * 16;
  const int nbx = cheb_fwd_pts_nbx(T);
  if (T.kDim <= 32)
    hipLaunchKernelGGL((k_cheb_fwd_pts<In, 8, 64>), dim3(nbx * T.nvars), dim3(kPtThreads),
                       cheb_fwd_pts_lds(T, sizeof(double)), st, T, phys, cm, kzp, bkp, nbx);
  else
    hipLaunchKernelGGL((k_cheb_fwd_pts<In, 8, 32>), dim3(nbx * T.nvars), dim3(kPtThreads),
                       cheb_fwd_pts_lds(T, sizeof(double)), st, T, phys, cm, kzp, bkp, nbx);
}

template <typename In>
void launch_cheb_fwd_pts_batch(const ChebJob<In>* jobs, const int* start, int njobs, int nblocks, size_t lds,
                               hipStream_t st) {
  hipLaunchKernelGGL(k_cheb_fwd_pts_batch<In>, dim3(nblocks), dim3(kPtThreads), lds, st, jobs, start, njobs);
}
template void launch_cheb_fwd_pts_batch<double>(const ChebJob<double>*, const int*, int, int, size_t, hipStream_t);
template void launch_cheb_fwd_pts_batch<float>(const ChebJob<float>*, const int*, int, int, size_t, hipStream_t);
template void launch_cheb_fwd_pts<double>(const DevTables&, const double*, double*, hipStream_t);
template void launch_cheb_fwd_pts<float>(const DevTables&, const float*, double*, hipStream_t);

// ----------------------------------------------------------------------------
// Ring DFT.  Block = (k-block, ring, variable block); wave w owns row tiles
// rt0 + w + j * kAnaWaves (j < MR) and all NCT column tiles.  Each chunk of 16
// quads is loaded (the next chunk's loads are in flight during the current
// chunk's MFMAs), folded in registers into E+/E-/O+/O-, and stored into one of
// two LDS tiles, so one barrier per chunk suffices.  Rows past kmax compute
// on valid table entries and are never stored.
// ----------------------------------------------------------------------------
template <typename R, typename In, int NCT, int MR>
__device__ __forceinline__ void ring_ana_body(const DevTables& T, const AnaLaunch& P, int lin,
                                              const In* __restrict__ src, R* __restrict__ rc) {
  typedef typename Mfma<R>::v4 v4;
  extern __shared__ __align__(16) unsigned char smraw[];
  R* sm = reinterpret_cast<R*>(smraw);
  const RTab<R> tb = rtab<R>(T);
  // unit = (ring, variable block), largest rings first; the k-blocks of a unit
  // stream the same inputs and run on one XCD
  const int nvb = T.geometry == 3 ? T.nvars : 1;
  int unit, kb;
  xcd_unit_tile(lin, P.nkb_max, unit, kb);
  if (unit >= T.iDim * nvb) return;
  const int r = T.iDim - 1 - unit / nvb;
  const int N = T.ringN[r], ri = T.ringRi[r], N4 = N / 4, N2 = N / 2;
  const bool rlz = T.geometry == 3;
  const int v_blk = unit - (unit / nvb) * nvb;
  int km = 0;
  if (rlz) km = T.kmax[v_blk * T.iDim + r];
  else for (int v = 0; v < T.nvars; v++) km = max(km, T.kmax[v * T.iDim + r]);
  const int ne = (km / 2 + 1 + 15) / 16;            // even-k row tiles (k = 0, 2, ...)
  const int no = ((km + 1) / 2 + 15) / 16;          // odd-k row tiles
  const int nrt = ne + no;
  const int per_block = kAnaWaves * MR;
  const int nkb = (nrt + per_block - 1) / per_block;
  if (kb >= nkb) return;                             // block-uniform
  const int rt0 = kb * nrt / nkb, rt1 = (kb + 1) * nrt / nkb;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int nthr = blockDim.x;
  const int t0 = rlz ? v_blk * T.nzb : 0;
  constexpr int NC = NCT * 16;

  // column (c) and point (p) strides of the input: RL physical slot 1
  // [var][npts], RLZ Chebyshev modes cm[var][zb][point]
  const In* xb = rlz ? src + long(v_blk) * T.jDim * T.nzb + T.ringOff[r] : src + T.ringOff[r];
  const long cs = rlz ? T.jDim : T.npts;
  const long ps = 1;

  // LDS: tw[2N] | EO[2][4][16][ldc]
  R* tw = sm;
  R* EO = tw + 2 * (4 + 4 * T.K);
  const R* tws = tb.twid + T.twOff[r];
  for (int m = threadIdx.x; m < 2 * N; m += nthr) tw[m] = tws[m < N ? m : m - N];

  // staging element of this thread: (ql, col), e = threadIdx.x + s * nthr < 16 * NC
  constexpr int kSlots = (16 * NC + 64 * kAnaWaves - 1) / (64 * kAnaWaves);
  R xv[kSlots][4];
  auto load = [&](int q0) {
#pragma unroll
    for (int s = 0; s < kSlots; s++) {
      const int e = threadIdx.x + s * nthr;
      const int ql = e / NC, col = e - (e / NC) * NC, q = q0 + ql;
      const bool ok = e < 16 * NC && col < P.ncols && q <= N4;
      // points q, N/2 - q, N/2 + q, N - q (mod N)
      const int pa = q, pb = N2 - q, pc = N2 + q, pd = q == 0 ? 0 : N - q;
      const In* c0 = xb + long(col) * cs;
      xv[s][0] = ok ? R(c0[long(pa) * ps]) : R(0);
      xv[s][1] = ok ? R(c0[long(pb) * ps]) : R(0);
      xv[s][2] = ok ? R(c0[long(pc) * ps]) : R(0);
      xv[s][3] = ok ? R(c0[long(pd) * ps]) : R(0);
    }
  };
  auto fold = [&](int q0, R* eo) {
#pragma unroll
    for (int s = 0; s < kSlots; s++) {
      const int e = threadIdx.x + s * nthr;
      if (e >= 16 * NC) continue;
      const int ql = e / NC, col = e - (e / NC) * NC, q = q0 + ql;
      const R xa = xv[s][0], xbb = xv[s][1], xc = xv[s][2], xd = xv[s][3];
      const bool first = q == 0, quarter = q == N4, out = q > N4;
      const R s1 = xa + xd, s2 = xbb + xc, d1 = xa - xd, d2 = xc - xbb;
      // q = 0: x_0 +/- x_{N/2} and no sine part; q = N/4: x_{N/4} +/- x_{3N/4} for both
      const R ep = out ? R(0) : first ? xa + xbb : quarter ? xa + xc : s1 + s2;
      const R em = out ? R(0) : first ? xa - xbb : quarter ? xa - xc : s1 - s2;
      const R op = (out || first) ? R(0) : quarter ? xa + xc : d1 + d2;
      const R om = (out || first) ? R(0) : quarter ? xa - xc : d1 - d2;
      eo[(0 * 16 + ql) * P.ldc + col] = ep;
      eo[(1 * 16 + ql) * P.ldc + col] = em;
      eo[(2 * 16 + ql) * P.ldc + col] = op;
      eo[(3 * 16 + ql) * P.ldc + col] = om;
    }
  };

  // A rows of this lane: k of row tile j, table index (k q) mod N with q the
  // lane's first quad, advancing by 4k mod N per k-step of 4 quads
  auto wrap = [N](int x) { return int(min(unsigned(x), unsigned(x - N))); };
  int idx[MR], inc[MR];
  bool par[MR];
#pragma unroll
  for (int j = 0; j < MR; j++) {
    const int rt = min(rt0 + w + j * kAnaWaves, rt1 - 1);
    const bool odd = rt >= ne;
    const int ti = odd ? rt - ne : rt;
    const int k = ((odd ? 1 : 0) + 2 * (ti * 16 + (lane & 15))) % N;
    par[j] = odd;
    idx[j] = (k * (lane >> 4)) % N;
    inc[j] = (4 * k) % N;
  }
  v4 acr[MR][NCT], aci[MR][NCT];
#pragma unroll
  for (int j = 0; j < MR; j++)
#pragma unroll
    for (int c = 0; c < NCT; c++) { acr[j][c] = v4{0, 0, 0, 0}; aci[j][c] = v4{0, 0, 0, 0}; }
  const int q3 = 3 * N4;
  const int eostride = 4 * 16 * P.ldc;

  load(0);
  fold(0, EO);
  __syncthreads();
  int buf = 0;
  for (int q0 = 0; q0 <= N4; q0 += kQC) {
    const bool more = q0 + kQC <= N4;
    if (more) load(q0 + kQC);
    const R* eo = EO + buf * eostride;
#pragma unroll
    for (int ks = 0; ks < kQC / 4; ks++) {
      const int qrow = ks * 4 + (lane >> 4);         // quad (within the chunk) of this lane's B row
#pragma unroll
      for (int j = 0; j < MR; j++) {
        if (rt0 + w + j * kAnaWaves >= rt1) continue;   // wave-uniform
        const R ca = tw[idx[j]];
        const R sa = tw[idx[j] + q3];                 // sin(x) = cos(x - pi/2); idx + 3N/4 < 2N
        idx[j] = wrap(idx[j] + inc[j]);
        const int ce = par[j] ? 1 : 0;               // E- / O- for odd k
#pragma unroll
        for (int c = 0; c < NCT; c++) {
          const int col = c * 16 + (lane & 15);
          const R be = eo[(ce * 16 + qrow) * P.ldc + col];
          const R bo = eo[((2 + ce) * 16 + qrow) * P.ldc + col];
          acr[j][c] = Mfma<R>::mma(ca, be, acr[j][c]);
          aci[j][c] = Mfma<R>::mma(sa, bo, aci[j][c]);
        }
      }
    }
    if (!more) break;
    buf ^= 1;
    fold(q0 + kQC, EO + buf * eostride);             // the other tile: last read before the previous barrier
    __syncthreads();
  }

  // epilogue: scale, rotate by -k ymin, store in line order (rc's ring layout)
  const R invN = R(1.0 / N);
  const int coef0 = T.coefOff[r];
  const R* pc = tb.phaseC + T.phaseOff[r];
  const R* psn = tb.phaseS + T.phaseOff[r];
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
      R* out = rc + long(t) * T.b_jDim + coef0;
#pragma unroll
      for (int rr = 0; rr < 4; rr++) {
        const int k = (odd ? 1 : 0) + 2 * (ti * 16 + Mfma<R>::row(lane, rr));
        if (k > kmv) continue;
        const R re = acr[j][c][rr] * invN, im = -aci[j][c][rr] * invN;
        if (k == 0) { out[0] = re; continue; }
        const R cc = pc[k - 1], ss = psn[k - 1];
        out[2 * k - 1] = cc * re + ss * im;
        out[2 * k] = -ss * re + cc * im;
      }
    }
  }
  // zero-fill kmax < k <= ri (only the first k-block)
  if (kb == 0) {
    for (int e = threadIdx.x; e < P.ncols * ri; e += nthr) {
      const int col = e / ri, k = 1 + e - col * ri;
      const int v = rlz ? v_blk : col;
      if (k > T.kmax[v * T.iDim + r]) {
        R* out = rc + long(t0 + col) * T.b_jDim + coef0;
        out[2 * k - 1] = R(0);
        out[2 * k] = R(0);
      }
    }
  }
}

template <typename R, typename In, int NCT, int MR>
__global__ void __launch_bounds__(64 * kAnaWaves, 2) k_ring_ana(DevTables T, AnaLaunch P, const In* __restrict__ src,
                                                                R* __restrict__ rc) {
  ring_ana_body<R, In, NCT, MR>(T, P, int(blockIdx.x), src, rc);
}

template <typename R, typename In, int NCT, int MR>
__global__ void __launch_bounds__(64 * kAnaWaves, 2) k_ring_ana_batch(const AnaJob<R, In>* __restrict__ jobs,
                                                                      const int* __restrict__ start, int njobs) {
  int lin;
  const AnaJob<R, In>& j = jobs[batch_job(start, njobs, lin)];
  ring_ana_body<R, In, NCT, MR>(j.T, j.P, lin, j.src, j.rc);
}

AnaLaunch plan_ana(const DevTables& T, size_t esize) {
  AnaLaunch P{};
  const bool rlz = T.geometry == 3;
  P.ncols = rlz ? T.nzb : T.nvars;
  P.nct = (P.ncols + 15) / 16;
  P.mr = P.nct <= 2 ? 2 : 1;
  if (T.tune.ana_mr == 1) P.mr = 1;                 // the tuner may halve the row tiles per wave
  P.ldc = P.nct * 16 + 2;
  const int km = T.K;                          // largest ring, uncapped bound
  const int nrt = (km / 2 + 1 + 15) / 16 + ((km + 1) / 2 + 15) / 16;
  P.nkb_max = (nrt + kAnaWaves * P.mr - 1) / (kAnaWaves * P.mr);
  const size_t n = 2 * size_t(4 + 4 * T.K) + 2 * 4 * 16 * size_t(P.ldc);
  P.lds = n * esize;
  P.nblocks = int(xcd_unit_grid(P.nkb_max, T.iDim * (rlz ? T.nvars : 1)));
  return P;
}

// at most 48 columns per ring transform; RLZ tiles of 64 points hold kDim <= 64
bool ring_ana_supported(const DevTables& T) {
  return plan_ana(T, sizeof(double)).nct <= 3 && (T.geometry != 3 || T.kDim <= 64);
}

template <typename R, typename In>
void launch_ring_ana(const DevTables& T, const In* src, R* rc, hipStream_t st) {
  const AnaLaunch P = plan_ana(T, sizeof(R));
  const dim3 grid(P.nblocks), block(64 * kAnaWaves);
  const int shape = P.nct * 2 + P.mr - 1;           // (nct, mr)
  switch (shape) {
    case 2: hipLaunchKernelGGL((k_ring_ana<R, In, 1, 1>), grid, block, P.lds, st, T, P, src, rc); break;
    case 3: hipLaunchKernelGGL((k_ring_ana<R, In, 1, 2>), grid, block, P.lds, st, T, P, src, rc); break;
    case 4: hipLaunchKernelGGL((k_ring_ana<R, In, 2, 1>), grid, block, P.lds, st, T, P, src, rc); break;
    case 5: hipLaunchKernelGGL((k_ring_ana<R, In, 2, 2>), grid, block, P.lds, st, T, P, src, rc); break;
    default: hipLaunchKernelGGL((k_ring_ana<R, In, 3, 1>), grid, block, P.lds, st, T, P, src, rc); break;
  }
}

template <typename R, typename In>
void launch_ring_ana_batch(const AnaJob<R, In>* jobs, const int* start, int njobs, int nblocks, int nct, int mr,
                           size_t lds, hipStream_t st) {
  const dim3 grid(nblocks), block(64 * kAnaWaves);
  switch (nct * 2 + mr - 1) {
    case 2: hipLaunchKernelGGL((k_ring_ana_batch<R, In, 1, 1>), grid, block, lds, st, jobs, start, njobs); break;
    case 3: hipLaunchKernelGGL((k_ring_ana_batch<R, In, 1, 2>), grid, block, lds, st, jobs, start, njobs); break;
    case 4: hipLaunchKernelGGL((k_ring_ana_batch<R, In, 2, 1>), grid, block, lds, st, jobs, start, njobs); break;
    case 5: hipLaunchKernelGGL((k_ring_ana_batch<R, In, 2, 2>), grid, block, lds, st, jobs, start, njobs); break;
    default: hipLaunchKernelGGL((k_ring_ana_batch<R, In, 3, 1>), grid, block, lds, st, jobs, start, njobs); break;
  }
}
template void launch_ring_ana_batch<double, double>(const AnaJob<double, double>*, const int*, int, int, int, int,
                                                    size_t, hipStream_t);
template void launch_ring_ana_batch<double, float>(const AnaJob<double, float>*, const int*, int, int, int, int,
                                                   size_t, hipStream_t);
template void launch_ring_ana<double, double>(const DevTables&, const double*, double*, hipStream_t);
template void launch_ring_ana<double, float>(const DevTables&, const float*, double*, hipStream_t);

}  // namespace ssg
