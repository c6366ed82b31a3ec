// ssg_kernels.hip -- CDNA4 (gfx950) kernels of the RL / RZ / RLZ transform path, FP64.
//
// Data layouts (per plan, device resident):
//   physical  [slot][var][npts]               (Julia physical[npts, nvars, nslots])
//   spectral  [var][nspec]                    (Julia spectral[nspec, nvars])
//   cb        [var][zb][ncol]                 Chebyshev modes per physical column
//   rc        [var*nzb + zb][b_jDim]          ring coefficients, ring-major, each ring's
//                                             2 ri + 1 values in spectral line order (DC,
//                                             re 1, im 1, re 2, ...): the projection reads
//                                             a run of consecutive lines as one run
//   X         [vb][ring][P|Q][d][tl][k]      phase-rotated ring coefficients of the d0/d1/d2
//                                             radial values (ring synthesis input, see k_sa_si_ring)
//   cbd (RZ)  [d][var][zb][iDim]              Chebyshev modes of d0/d1/d2 radial values
// The ring transforms (ssg_ring_ana.hip, ssg_ring_mfma.hip) write rc and read
// X; for RLZ they also carry the Chebyshev stages, so no (ring, l, zb)
// intermediate is materialised.  A "line" is one B-spline coefficient vector of b_iDim rows in spectral; line
// L = (var * nzb + zb) * nl + p, with p = 0 the k = 0 line and p = 2k-1 / 2k the
// real / imaginary line of wavenumber k (RL and RLZ); nl = 1 for RZ.
#include "ssg_kernels.hpp"

namespace ssg {

static constexpr int kTile = 64;     // physical columns per Chebyshev tile
static constexpr int kBlock = 256;   // threads per block of the tile kernels
static constexpr int kSbChunk = 16;  // spectral rows per chunk (SB output tiles)
static constexpr int kSaChunk = 8;   // rows per unrolled chunk of the ring solve

// A table that no kernel writes, read at a wave-uniform index, through the
// constant address space: such loads become scalar loads, which the texture
// address unit never sees (as generic loads they would be vector loads, the
// compiler being unable to prove that the kernel's stores leave the table alone).
template <typename V>
__device__ __forceinline__ const __attribute__((address_space(4))) V* const_tab(const V* p) {
  return (const __attribute__((address_space(4))) V*)(p);
}

// ----------------------------------------------------------------------------
// Forward Chebyshev: b[zb] = sum_z Fwd[zb][z] u[z] per column (CBtransform!,
// src/Chebyshev.jl:417-426).  A 64-column tile is staged through LDS with a
// coalesced load; wave w produces modes zb = w, w+4, ...; Fwd is wave-uniform.
// ----------------------------------------------------------------------------
__global__ void __launch_bounds__(kBlock) k_cheb_fwd(const double* __restrict__ phys, long npts, int ncol, int kDim,
                                                     int bk, const double* __restrict__ fwd, double* __restrict__ cb) {
  extern __shared__ double sm[];
  const int v = blockIdx.y;
  const long col0 = long(blockIdx.x) * kTile;
  const int ncl = int(min(long(kTile), ncol - col0));
  const double* src = phys + long(v) * npts + col0 * kDim;
  const int ld = kDim + 1;
  for (int e = threadIdx.x; e < ncl * kDim; e += blockDim.x) {
    const int c = e / kDim, z = e - c * kDim;
    sm[c * ld + z] = src[e];
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const double* F = fwd + long(v) * bk * kDim;
  if (lane < ncl) {
    for (int zb = w; zb < bk; zb += kBlock / 64) {
      double acc = 0.0;
      for (int z = 0; z < kDim; z++) acc += F[zb * kDim + z] * sm[lane * ld + z];
      cb[(long(v) * bk + zb) * ncol + col0 + lane] = acc;
    }
  }
}

void launch_cheb_fwd(const DevTables& T, const double* phys, int ncol, double* cb, hipStream_t st) {
  dim3 grid((ncol + kTile - 1) / kTile, T.nvars);
  size_t lds = sizeof(double) * kTile * (T.kDim + 1);
  hipLaunchKernelGGL(k_cheb_fwd, grid, dim3(kBlock), lds, st, phys, T.npts, ncol, T.kDim, T.bk, T.fwd, cb);
}


// value of the line's input at radial point i (RL/RLZ: ring coefficient, zero where k > ri)
template <typename In>
__device__ __forceinline__ In line_input(const DevTables& T, const In* __restrict__ src, long q, int p, int i) {
  if (T.geometry == 2) return src[q * T.iDim + i];   // RZ: cb[v][zb][r]
  const In* ring = src + q * T.b_jDim + T.coefOff[i];
  const int k = (p + 1) >> 1, ri = T.ringRi[i];
  if (k > ri) return In(0);
  return ring[p];
}

// ----------------------------------------------------------------------------
// SB: b = M u with M[m][i] = DX w_mu phi_m(x_i), <= 4 mubar taps per row
// (SBtransform! src/CubicBSpline.jl:1107-1111, matrix 1000-1030).  One lane per
// line streams the radial points cell by cell and keeps the four open rows in
// registers; rows accumulate in the CSC column order of the reference.  The
// inputs of kSbCells cells are loaded together, so that many loads are in
// flight per lane.  FP64 in both forms.
// ----------------------------------------------------------------------------
static constexpr int kSbCells = 4;
// one wave per block: the line count is rarely a multiple of larger blocks
// times the CU count, and one-wave blocks spread the lines evenly over the CUs
static constexpr int kSbBlock = 64;
static constexpr int kSbMaxMu = 5;                   // largest mubar of the Gauss tables

__device__ __forceinline__ void sb_body_at(const DevTables& T, const double* __restrict__ src,
                                           double* __restrict__ spec, long L) {
  typedef double In;
  const int p = int(L % T.nl);
  const long q = L / T.nl;             // var * nzb + zb
  double* out = spec + L * T.b_iDim;   // line L of spectral ([var][zb][p] lines are consecutive)
  const In* sbw = T.sbw;
  const int mb = T.mubar;
  // Ring grids: the input of point i is element p of ring i (line order),
  // formed without branches (every load is issued, at offset 0 for a point
  // past the line or a wavenumber the ring does not carry, and its value
  // discarded), so the loads of a batch of cells are all in flight together.
  const bool ring = T.geometry != 2;
  const int kw = (p + 1) >> 1;
  const In* srcq = src + (ring ? q * T.b_jDim : q * T.iDim);
  auto input = [&](int i, bool valid) -> In {
    const int ic = valid ? i : 0;
    int off = ic;                                    // RZ: cb[v][zb][r]
    bool z = !valid;
    if (ring) {
      const int ri = T.ringRi[ic];
      off = T.coefOff[ic] + p;
      z = z || kw > ri;
    }
    const In v = srcq[z ? 0 : off];
    return z ? In(0) : v;
  };
  In a0 = 0, a1 = 0, a2 = 0, a3 = 0;
  for (int mc0 = 0; mc0 < T.nc; mc0 += kSbCells) {
    In u[kSbCells * kSbMaxMu];
#pragma unroll
    for (int j = 0; j < kSbCells * kSbMaxMu; j++) {
      const int c = j / kSbMaxMu, mu = j - c * kSbMaxMu;
      u[j] = input((mc0 + c) * mb + mu, mc0 + c < T.nc && mu < mb);
    }
#pragma unroll
    for (int c = 0; c < kSbCells; c++) {
      const int mc = mc0 + c;
      if (mc >= T.nc) break;
#pragma unroll
      for (int mu = 0; mu < kSbMaxMu; mu++) {
        if (mu >= mb) break;
        const auto* w = const_tab(sbw) + 4 * (mc * mb + mu);
        const In x = u[c * kSbMaxMu + mu];
        a0 += w[0] * x;
        a1 += w[1] * x;
        a2 += w[2] * x;
        a3 += w[3] * x;
      }
      for (int mu = kSbMaxMu; mu < mb; mu++) {     // :regular rules with more points per cell
        const auto* w = const_tab(sbw) + 4 * (mc * mb + mu);
        const In x = input(mc * mb + mu, true);
        a0 += w[0] * x;
        a1 += w[1] * x;
        a2 += w[2] * x;
        a3 += w[3] * x;
      }
      out[mc] = a0;
      a0 = a1; a1 = a2; a2 = a3; a3 = 0;
    }
  }
  out[T.nc] = a0;
  out[T.nc + 1] = a1;
  out[T.nc + 2] = a2;
}

__global__ void __launch_bounds__(kSbBlock) k_sb(DevTables T, const double* __restrict__ src,
                                                 double* __restrict__ spec) {
  const long L = long(blockIdx.x) * blockDim.x + threadIdx.x;
  if (L < T.nlines) sb_body_at(T, src, spec, L);
}

// ----------------------------------------------------------------------------
// SB of the ring grids, tiled.  A block takes 32 consecutive lines p of one
// transform t (= var * nzb + zb) and 64 consecutive output rows m.  Row m
// gathers the points of cells m-3 .. m (points [(m-3) mb, (m+1) mb)), so the
// block stages the points [(m0-3) mb, (m0+64) mb) of its 32 lines in LDS as
// [point][line]: the lines of one ring point are entries k (real lines,
// ascending) and 2 ri + 1 - k (imaginary lines, descending) of the ring's b
// layout, so a wave's staging loads are two contiguous runs per ring.  Then
// lane = row and wave w computes lines 8w .. 8w + 7: every store instruction
// writes 64 consecutive rows of one line.  Each row sums its points in
// ascending order, the CSC column order of the reference's mul!
// (src/CubicBSpline.jl:1107-1111); taps outside the grid carry weight zero.
// ----------------------------------------------------------------------------
static constexpr int kSbtRows = 64;
static constexpr int kSbtLines = 16;
static constexpr int kSbtThreads = 256;
static constexpr int kSbtLd = kSbtLines + 1;
static constexpr int kSbtBatch = 14;                  // staging loads in flight per thread (one batch for mubar 3)

__host__ __device__ inline int sbt_nlb(const DevTables& T) { return (T.nl + kSbtLines - 1) / kSbtLines; }
__host__ __device__ inline int sbt_nseg(const DevTables& T) { return (T.b_iDim + kSbtRows - 1) / kSbtRows; }
// the [point][line] tile and the ring size and offset of each of its points
static size_t sbt_lds(const DevTables& T) {
  return (sizeof(double) * kSbtLd + 2 * sizeof(int)) * size_t(kSbtRows + 3) * T.mubar;
}
// the tiled kernel takes the Gauss rules (mubar <= 5) of the ring grids
__host__ __device__ inline bool sbt_ok(const DevTables& T) { return T.geometry != 2 && T.mubar <= kSbMaxMu; }

__device__ __forceinline__ void sb_tile_body(const DevTables& T, const double* __restrict__ src,
                                             double* __restrict__ spec, int blk) {
  extern __shared__ double tile[];
  const int mb = T.mubar;
  const int nseg = sbt_nseg(T), nlb = sbt_nlb(T);
  const int seg = blk % nseg, rest = blk / nseg;
  const int lb = rest % nlb, t = rest / nlb;
  if (t >= T.ntr) return;                            // block-uniform (padding blocks of a batched job)
  const int m0 = seg * kSbtRows, p0 = lb * kSbtLines;
  const int i0 = max(0, (m0 - 3) * mb), i1 = min(T.iDim, (m0 + kSbtRows) * mb);
  const double* srct = src + long(t) * T.b_jDim;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int m = m0 + lane;
  // the taps of row m (cell m - 3 + c with tap 3 - c, point mu of the cell) are
  // loaded first, so that their latency overlaps the staging loads
  double wt[4 * kSbMaxMu];
  int ix[4 * kSbMaxMu];
#pragma unroll
  for (int j = 0; j < 4 * kSbMaxMu; j++) {
    const int c = j / kSbMaxMu, mu = j - c * kSbMaxMu;
    const int mc = m - 3 + c;
    const bool ok = m < T.b_iDim && mu < mb && mc >= 0 && mc < T.nc;
    const int i = ok ? mc * mb + mu : i0;
    wt[j] = ok ? T.sbw[4 * i + (3 - c)] : 0.0;
    ix[j] = (i - i0) * kSbtLd;
  }
  // Ring size ri and coefficient offset of each tile point: cylindrical rings
  // have ri = i + 1 + o and offset sum_{j<i} (2 rj + 1) = i (i + 2 o + 2)
  // (o = patchOffsetL); spherical rings read them from the tables, staged in
  // LDS first.  Then the inputs, kSbtBatch independent loads per thread.
  int* rri = reinterpret_cast<int*>(tile + (kSbtRows + 3) * mb * kSbtLd);
  int* roff = rri + (kSbtRows + 3) * mb;
  const int np = i1 - i0;
  const bool tab = T.sph != 0;
  const int o = T.ringRi[0] - 1;
  if (tab) {
    for (int ii = threadIdx.x; ii < np; ii += blockDim.x) {
      rri[ii] = T.ringRi[i0 + ii];
      roff[ii] = T.coefOff[i0 + ii];
    }
    __syncthreads();
  }
  constexpr int lsh = 4;                             // log2(kSbtLines)
  const int ne = np * kSbtLines;
  for (int e0 = threadIdx.x; e0 < ne; e0 += kSbtBatch * kSbtThreads) {
    double v[kSbtBatch];
#pragma unroll
    for (int u = 0; u < kSbtBatch; u++) {
      const int e = e0 + u * kSbtThreads;
      const int ln = e & (kSbtLines - 1), ii = min(e >> lsh, np - 1);
      const int p = p0 + ln, i = i0 + ii;
      const int k = (p + 1) >> 1;
      const int ri = tab ? rri[ii] : i + 1 + o;
      const int cof = tab ? roff[ii] : i * (i + 2 * o + 2);
      const bool on = e < ne && p < T.nl && k <= ri;   // zero where the ring lacks the wavenumber
      const int off = cof + p;
      v[u] = on ? srct[off] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < kSbtBatch; u++) {
      const int e = e0 + u * kSbtThreads;
      if (e < ne) tile[(e >> lsh) * kSbtLd + (e & (kSbtLines - 1))] = v[u];
    }
  }
  __syncthreads();
  if (m >= T.b_iDim) return;                         // no barrier follows
  const long tl0 = long(t) * T.nl;
  const int lines = kSbtLines / (kSbtThreads / 64);
#pragma unroll
  for (int q = 0; q < lines; q++) {
    const int ln = w * lines + q;
    const int p = p0 + ln;
    if (p >= T.nl) break;                            // wave-uniform
    double acc = 0.0;
#pragma unroll
    for (int j = 0; j < 4 * kSbMaxMu; j++) acc += wt[j] * tile[ix[j] + ln];
    spec[(tl0 + p) * T.b_iDim + m] = acc;
  }
}

__global__ void __launch_bounds__(kSbtThreads) k_sb_tile(DevTables T, const double* __restrict__ src,
                                                        double* __restrict__ spec) {
  sb_tile_body(T, src, spec, int(blockIdx.x));
}

// batched: jobs of ring grids with Gauss rules run the tiled body, others the lane-per-line one
__global__ void __launch_bounds__(kSbtThreads) k_sb_batch(const SbJob* __restrict__ jobs, const int* __restrict__ start,
                                                        int njobs) {
  int lin;
  const SbJob& j = jobs[batch_job(start, njobs, lin)];
  if (sbt_ok(j.T)) {
    sb_tile_body(j.T, j.rc, j.spec, lin);
  } else {
    // lane-per-line body in 64-thread slices of the 256-thread block
    const int sub = threadIdx.x >> 6;
    const long blk = long(lin) * (kSbtThreads / kSbBlock) + sub;
    const long L = blk * kSbBlock + (threadIdx.x & 63);
    if (L < j.T.nlines) sb_body_at(j.T, j.rc, j.spec, L);
  }
}

int sb_blocks(const DevTables& T) {
  if (sbt_ok(T)) return T.ntr * sbt_nlb(T) * sbt_nseg(T);
  return int((T.nlines + kSbtThreads - 1) / kSbtThreads);
}
size_t sb_lds(const DevTables& T) { return sbt_ok(T) ? sbt_lds(T) : 0; }

void launch_sb_ring_batch(const SbJob* jobs, const int* start, int njobs, int nblocks, size_t lds, hipStream_t st) {
  hipLaunchKernelGGL(k_sb_batch, dim3(nblocks), dim3(kSbtThreads), lds, st, jobs, start, njobs);
}

void launch_sb(const DevTables& T, const double* src, double* spec, hipStream_t st) {
  hipLaunchKernelGGL(k_sb, dim3(unsigned((T.nlines + kSbBlock - 1) / kSbBlock)), dim3(kSbBlock), 0, st, T, src, spec);
}

void launch_sb_ring(const DevTables& T, const double* rc, double* spec, hipStream_t st) {
  if (sbt_ok(T))
    hipLaunchKernelGGL(k_sb_tile, dim3(unsigned(sb_blocks(T))), dim3(kSbtThreads), sbt_lds(T), st, T, rc, spec);
  else
    hipLaunchKernelGGL(k_sb, dim3(unsigned((T.nlines + kSbBlock - 1) / kSbBlock)), dim3(kSbBlock), 0, st, T, rc, spec);
}

// ----------------------------------------------------------------------------
// SA + SI: a = G^T (L L^T)^{-1} G b, then u^(d)(x_i) = sum_t phi^(d)_{mc-1+t}(x_i)
// a[mc+t] for d = 0, 1, 2 (SAtransform! src/CubicBSpline.jl:1368-1396,
// src/banded_cholesky.jl:202-247; SItransform 1554-1567).  One lane per line;
// the forward-solve intermediate and then x are kept in y[i][line].
// ----------------------------------------------------------------------------
__device__ __forceinline__ void si_store(const DevTables& T, double* __restrict__ dst, long q, int p, int i, int d,
                                         double val) {
  if (T.geometry == 2) {   // RZ: cbd[d][v][zb][r]
    dst[(long(d) * T.nvars * T.nzb + q) * T.iDim + i] = val;
    return;
  }
  double* ring = dst + (long(d) * T.nvars * T.nzb + q) * T.b_jDim + T.coefOff[i];
  const int k = (p + 1) >> 1, ri = T.ringRi[i];
  if (k > ri) return;
  ring[p] = val;
}

__global__ void __launch_bounds__(kBlock) k_sa_si(DevTables T, const double* __restrict__ spec, double* __restrict__ y,
                                                  double* __restrict__ dst, int si) {
  const long L = long(blockIdx.x) * blockDim.x + threadIdx.x;
  if (L >= T.nlines) return;
  const int p = int(L % T.nl);
  const long q = L / T.nl;
  const int zb = int(q % T.nzb), v = int(q / T.nzb);
  const double* b = spec + long(v) * T.nspec + long(zb) * T.nl * T.b_iDim + long(p) * T.b_iDim;
  const int* gi = T.gint + 8 * v;
  const double* gd = T.gdbl + 4 * v;
  const int Mdim = gi[0], Min = gi[1], rL = gi[2], lk = gi[4], rk = gi[5];
  const double aL = gd[0], bL = gd[1], aR = gd[2], bR = gd[3];
  const double* Lb = T.L + long(v) * T.maxMin * 4;
  const double* Li = T.Linv + long(v) * T.maxMin;
  const long ys = T.nlines;
  // forward solve L y = G b
  double y1 = 0.0, y2 = 0.0, y3 = 0.0;   // y[i-1], y[i-2], y[i-3]
  for (int i = 0; i < Min; i++) {
    double g = b[i + rL];
    if (i == 0) { if (lk == 1) g += aL * b[0]; else if (lk == 2) g += aL * b[0] + bL * b[1]; }
    if (i == 1 && lk == 1) g += bL * b[0];
    if (i == Min - 1) { if (rk == 1) g += aR * b[Mdim - 1]; else if (rk == 2) g += aR * b[Mdim - 1] + bR * b[Mdim - 2]; }
    if (i == Min - 2 && rk == 1) g += bR * b[Mdim - 1];
    double s = g;
    if (i >= 3) s -= Lb[(i - 3) * 4 + 3] * y3;
    if (i >= 2) s -= Lb[(i - 2) * 4 + 2] * y2;
    if (i >= 1) s -= Lb[(i - 1) * 4 + 1] * y1;
    const double yi = s * Li[i];
    y[long(i) * ys + L] = yi;
    y3 = y2; y2 = y1; y1 = yi;
  }
  // backward solve L^T x = y (x overwrites y)
  double x1 = 0.0, x2 = 0.0, x3 = 0.0;   // x[i+1], x[i+2], x[i+3]
  for (int i = Min - 1; i >= 0; i--) {
    double s = y[long(i) * ys + L];
    if (i + 1 < Min) s -= Lb[i * 4 + 1] * x1;
    if (i + 2 < Min) s -= Lb[i * 4 + 2] * x2;
    if (i + 3 < Min) s -= Lb[i * 4 + 3] * x3;
    const double xi = s * Li[i];
    y[long(i) * ys + L] = xi;
    x3 = x2; x2 = x1; x1 = xi;
  }
  if (!si) return;                                   // x alone (evaluation elsewhere)
  // a = G^T x, evaluated on the fly; x[0], x[1], x[Min-2], x[Min-1]
  const double xf0 = y[L], xf1 = y[ys + L];
  const double xl1 = y[long(Min - 1) * ys + L], xl2 = y[long(Min - 2) * ys + L];
  // R3X: the background coefficients are added back (SAtransform!, src/CubicBSpline.jl:1386)
  const double* ah = (T.ahat && gi[6]) ? T.ahat + L * T.b_iDim : nullptr;
  auto a_at = [&](int j) -> double {
    double a = 0.0;
    if (j >= rL && j < rL + Min) a = y[long(j - rL) * ys + L];
    if (lk == 1 && j == 0) a = aL * xf0 + bL * xf1;
    if (lk == 2) { if (j == 0) a = aL * xf0; else if (j == 1) a = bL * xf0; }
    if (rk == 1 && j == Mdim - 1) a = aR * xl1 + bR * xl2;
    if (rk == 2) { if (j == Mdim - 2) a = bR * xl1; else if (j == Mdim - 1) a = aR * xl1; }
    if (ah) a += ah[j];
    return a;
  };
  double w0 = a_at(0), w1 = a_at(1), w2 = a_at(2), w3 = a_at(3);
  for (int mc = 0; mc < T.nc; mc++) {
    for (int mu = 0; mu < T.mubar; mu++) {
      const int i = mc * T.mubar + mu;
#pragma unroll
      for (int d = 0; d < 3; d++) {
        const double* w = T.siw + long(d) * T.iDim * 4 + 4 * i;
        double u = 0.0;
        u += w[0] * w0;
        u += w[1] * w1;
        u += w[2] * w2;
        u += w[3] * w3;
        si_store(T, dst, q, p, i, d, u);
      }
    }
    w0 = w1; w1 = w2; w2 = w3;
    w3 = (mc + 4 < Mdim) ? a_at(mc + 4) : 0.0;
  }
}

// R3X branch of SAtransform! (src/CubicBSpline.jl:1377-1380): btilde = b - pq ahat
// for every line of a variable with an R3X side (pq = P + Q, bandwidth 3, the
// products summed in column order), b copied for the other variables.  One
// thread per spectral element.
__global__ void __launch_bounds__(kBlock) k_ahat_btilde(DevTables T, const double* __restrict__ spec,
                                                        double* __restrict__ bt) {
  const long e = long(blockIdx.x) * blockDim.x + threadIdx.x;
  const long n = T.nspec * T.nvars;
  if (e >= n) return;
  const int v = int(e / T.nspec);
  const int M = T.b_iDim;
  const long line0 = e - e % M;
  const int i = int(e - line0);
  double r = spec[e];
  if (T.gint[8 * v + 6]) {
    const double* pq = T.pqb + long(v) * M * 4;
    const double* ah = T.ahat + line0;
    double acc = 0.0;
    for (int j = max(0, i - 3); j <= min(M - 1, i + 3); j++) {
      const int lo = min(i, j), o = abs(i - j);
      acc += pq[lo * 4 + o] * ah[j];
    }
    r -= acc;
  }
  bt[e] = r;
}

void launch_ahat_btilde(const DevTables& T, const double* spec, double* bt, hipStream_t st) {
  const long n = T.nspec * T.nvars;
  hipLaunchKernelGGL(k_ahat_btilde, dim3(unsigned((n + kBlock - 1) / kBlock)), dim3(kBlock), 0, st, T, spec, bt);
}

// ----------------------------------------------------------------------------
// applyFilter! (src/filtering.jl:290-428) on the lines the plan's filter
// changes: one wave per line, 64 consecutive entries per step.  A zero weight
// stores +0.0 without reading (the reference assigns 0.0 to such blocks; a
// product would leave -0.0 or NaN), any other weight multiplies.
// ----------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_apply_filter(double* __restrict__ spec, const long* __restrict__ off,
                                                         const double* __restrict__ w, long nact, int b_iDim) {
  const long a = long(blockIdx.x) * (kBlock / 64) + (threadIdx.x >> 6);
  if (a >= nact) return;
  const double wa = w[a];
  double* __restrict__ q = spec + off[a];
  const int l = threadIdx.x & 63;
  if (wa == 0.0) {
    for (int j = l; j < b_iDim; j += 64) q[j] = 0.0;
  } else {
    for (int j = l; j < b_iDim; j += 64) q[j] *= wa;
  }
}

void launch_apply_filter(double* spec, const long* off, const double* w, long nact, int b_iDim, hipStream_t st) {
  if (nact <= 0) return;
  const long nb = (nact + kBlock / 64 - 1) / (kBlock / 64);
  hipLaunchKernelGGL(k_apply_filter, dim3(unsigned(nb)), dim3(kBlock), 0, st, spec, off, w, nact, b_iDim);
}

void launch_sa_si(const DevTables& T, const double* spec, double* y, double* dst, hipStream_t st) {
  dim3 grid(unsigned((T.nlines + kBlock - 1) / kBlock));
  hipLaunchKernelGGL(k_sa_si, grid, dim3(kBlock), 0, st, T, spec, y, dst, 1);
}

void launch_sa_lines(const DevTables& T, const double* spec, double* y, hipStream_t st) {
  dim3 grid(unsigned((T.nlines + kBlock - 1) / kBlock));
  hipLaunchKernelGGL(k_sa_si, grid, dim3(kBlock), 0, st, T, spec, y, static_cast<double*>(nullptr), 0);
}

// ----------------------------------------------------------------------------
// SA + SI of the ring grids with the inverse phase filter applied on the way
// out, in two kernels: k_sa_ring (Gamma fold and banded solve, x left in y)
// and k_si_ring (Gamma^T, evaluation, rotation, store; parallel over segments
// of cells).  Lanes come in pairs: the real (p = 2k-1) and imaginary (p = 2k) lines
// of wavenumber k of one transform sit in adjacent lanes (k = 0: line p = 0
// in both), run the same Gamma fold, banded solve and evaluation as k_sa_si,
// and swap their SI values, so each pair holds both parts of every ring
// coefficient and rotates them by the ring's phase
// (apply_phasefilter_inverse!, src/phase_filter.jl:154-190) before storing
//   X[vb][ring i][P|Q][d][tl][k] = (a_re, -a_im) of SI order d, k = 0..ri
// (the even lane stores P, the odd lane Q), zero where the wavenumber cap is
// below k, and (value, 0) for k = 0.  Lanes run over (transform, k, part), so
// the stores of a wave are runs of consecutive k; each lane reads its own
// spectral line, a chunk ahead of its recurrence.  y holds the solve
// intermediates, y[i][lane].
//
// The factor entries are the same for every lane of a variable.  As vector
// loads they would be four of every five load instructions of the solve, each
// one occupying the texture address unit like a gather; with FacLds the block
// first copies the factor of its (at most two) variables into LDS, and the
// recurrences read it there.  FacLds needs maxMin <= kSaLdsRows and at least
// kBlock lanes per variable (sa_fac_lds).
// ----------------------------------------------------------------------------
static constexpr int kSaLdsRows = 512;

__host__ __device__ bool sa_fac_lds(const DevTables& T) {
  return T.maxMin <= kSaLdsRows && 2 * long(T.K + 1) * T.nzb >= kBlock;
}

template <bool FacLds>
__device__ __forceinline__ void sa_ring_body(const DevTables& T, const double* __restrict__ spec,
                                             double* __restrict__ y, int blk) {
  const long nlanes = 2 * long(T.ntr) * (T.K + 1);
  const long u0 = long(blk) * blockDim.x;
  const long u = u0 + threadIdx.x;
  auto var_of = [&](long lane) { return int(((lane >> 1) / (T.K + 1)) / T.nzb); };
  const double* Lb;
  const double* Li;
  if constexpr (FacLds) {
    __shared__ double fac[2][5 * kSaLdsRows];
    // a block past the last lane (the batched launch pads each job's block range
    // to a multiple of 8) has no variable: it leaves before staging, as a whole
    if (u0 >= nlanes) return;
    const int vA = var_of(u0), vB = var_of(min(u0 + long(blockDim.x), nlanes) - 1);
    const int n4 = 4 * T.maxMin;
    for (int sl = 0; sl < 2; sl++) {
      const int vs = sl == 0 ? vA : vB;
      if (sl == 1 && vB == vA) break;               // block-uniform
      const double* gL = T.L + long(vs) * n4;
      const double* gI = T.Linv + long(vs) * T.maxMin;
      for (int e = threadIdx.x; e < n4; e += blockDim.x) fac[sl][e] = gL[e];
      for (int e = threadIdx.x; e < T.maxMin; e += blockDim.x) fac[sl][n4 + e] = gI[e];
    }
    __syncthreads();
    if (u >= nlanes) return;
    const int sl = var_of(u) == vA ? 0 : 1;
    Lb = fac[sl];
    Li = fac[sl] + n4;
  } else {
    if (u >= nlanes) return;
    Lb = T.L + long(var_of(u)) * T.maxMin * 4;
    Li = T.Linv + long(var_of(u)) * T.maxMin;
  }
  const int part = int(u & 1);
  const long pu = u >> 1;
  const int k = int(pu % (T.K + 1));
  const int t = int(pu / (T.K + 1));
  const int v = t / T.nzb;
  const long line = long(t) * T.nl + (k == 0 ? 0 : 2 * k - 1 + part);
  const double* b = spec + line * T.b_iDim;
  const int* gi = T.gint + 8 * v;
  const double* gd = T.gdbl + 4 * v;
  const int Mdim = gi[0], Min = gi[1], rL = gi[2], lk = gi[4], rk = gi[5];
  const double aL = gd[0], bL = gd[1], aR = gd[2], bR = gd[3];
  const double bf0 = b[0], bf1 = b[1], bl1 = b[Mdim - 1], bl2 = b[Mdim - 2];
  // forward solve L y = G b, kSaChunk rows at a time.  Each lane reads its own
  // line; the rows of the next chunk are loaded before the current chunk's
  // recurrence, which is fully unrolled over fixed register slots (rows past
  // Min masked: a break would keep it a loop with indexed registers and a
  // branch per boundary test)
  auto row = [&](int i) -> double { return b[min(i, Min - 1) + rL]; };
  // a whole chunk inside the line is read two rows per load (16-byte loads at
  // 8-byte alignment); the last, partial chunk row by row, clamped
  typedef double dv2 __attribute__((ext_vector_type(2), aligned(8)));
  auto chunk = [&](int i1, double* g) {
    if (i1 + kSaChunk <= Min) {
      const dv2* p2 = reinterpret_cast<const dv2*>(b + rL + i1);
#pragma unroll
      for (int c = 0; c < kSaChunk / 2; c++) {
        const dv2 x = p2[c];
        g[2 * c] = x.x;
        g[2 * c + 1] = x.y;
      }
    } else {
#pragma unroll
      for (int c = 0; c < kSaChunk; c++) g[c] = row(i1 + c);
    }
  };
  double gn[kSaChunk];
  chunk(0, gn);
  double y1 = 0.0, y2 = 0.0, y3 = 0.0;
  for (int i0 = 0; i0 < Min; i0 += kSaChunk) {
    double gv[kSaChunk];
#pragma unroll
    for (int c = 0; c < kSaChunk; c++) gv[c] = gn[c];
    if (i0 + kSaChunk < Min) chunk(i0 + kSaChunk, gn);
    double l3[kSaChunk], l2[kSaChunk], l1[kSaChunk], li[kSaChunk];
#pragma unroll
    for (int c = 0; c < kSaChunk; c++) {
      const int i = i0 + c;
      const bool in = i < Min;
      l3[c] = (in && i >= 3) ? Lb[(i - 3) * 4 + 3] : 0.0;
      l2[c] = (in && i >= 2) ? Lb[(i - 2) * 4 + 2] : 0.0;
      l1[c] = (in && i >= 1) ? Lb[(i - 1) * 4 + 1] : 0.0;
      li[c] = in ? Li[i] : 0.0;
    }
    // the fold's boundary terms, as selects
#pragma unroll
    for (int c = 0; c < kSaChunk; c++) {
      const int i = i0 + c;
      double g = gv[c];
      g = (i == 0 && lk == 1) ? g + aL * bf0 : g;
      g = (i == 0 && lk == 2) ? g + (aL * bf0 + bL * bf1) : g;
      g = (i == 1 && lk == 1) ? g + bL * bf0 : g;
      g = (i == Min - 1 && rk == 1) ? g + aR * bl1 : g;
      g = (i == Min - 1 && rk == 2) ? g + (aR * bl1 + bR * bl2) : g;
      g = (i == Min - 2 && rk == 1) ? g + bR * bl1 : g;
      gv[c] = g;
    }
#pragma unroll
    for (int c = 0; c < kSaChunk; c++) {
      const int i = i0 + c;
      double s = gv[c];
      s -= l3[c] * y3;
      s -= l2[c] * y2;
      s -= l1[c] * y1;
      const double yi = s * li[c];
      if (i < Min) {
        y[long(i) * nlanes + u] = yi;
        y3 = y2; y2 = y1; y1 = yi;
      }
    }
  }
  // backward solve L^T x = y (x overwrites y), kSaChunk rows at a time: the y
  // and factor entries of a group are loaded before its recurrence
  double x1 = 0.0, x2 = 0.0, x3 = 0.0;
  for (int i0 = Min - 1; i0 >= 0; i0 -= kSaChunk) {
    double yv[kSaChunk], m1[kSaChunk], m2[kSaChunk], m3[kSaChunk], mi[kSaChunk];
#pragma unroll
    for (int c = 0; c < kSaChunk; c++) {
      const int i = i0 - c;
      const bool in = i >= 0;
      yv[c] = in ? y[long(i) * nlanes + u] : 0.0;
      m1[c] = (in && i + 1 < Min) ? Lb[i * 4 + 1] : 0.0;
      m2[c] = (in && i + 2 < Min) ? Lb[i * 4 + 2] : 0.0;
      m3[c] = (in && i + 3 < Min) ? Lb[i * 4 + 3] : 0.0;
      mi[c] = in ? Li[i] : 0.0;
    }
#pragma unroll
    for (int c = 0; c < kSaChunk; c++) {
      const int i = i0 - c;
      double s = yv[c];
      s -= m1[c] * x1;
      s -= m2[c] * x2;
      s -= m3[c] * x3;
      const double xi = s * mi[c];
      if (i >= 0) {
        y[long(i) * nlanes + u] = xi;
        x3 = x2; x2 = x1; x1 = xi;
      }
    }
  }
}

// SI of the ring grids: lanes (segment of cells, lane of k_sa_ring) with the
// lane fastest; a = G^T x is read from y, the three derivative orders are
// evaluated at the segment's mish points in FP64, and the lane pair rotates
// them and stores them into X of type S (float for the FP32 form's synthesis:
// the evaluation itself must stay FP64, its d/dr and d2/dr2 sums cancel terms
// (L/DX)^2 times larger than the result).
template <typename S>
__device__ __forceinline__ void si_ring_body(const DevTables& T, int nseg, const double* __restrict__ y,
                                             S* __restrict__ X, int blk) {
  typedef double R;
  typedef double Y;
  const RTab<R> tb = rtab<R>(T);
  const long nlanes = 2 * long(T.ntr) * (T.K + 1);
  // each segment's lanes start on a wave: the segment, and so the cell loop
  // and the weight indices, are wave-uniform
  const long nlp = (nlanes + 63) / 64 * 64;
  const long g = long(blk) * blockDim.x + threadIdx.x;
  if (g >= nlp * nseg) return;
  const long u = g % nlp;
  const int seg = __builtin_amdgcn_readfirstlane(int(g / nlp));
  if (u >= nlanes) return;                           // nlanes is even: pairs stay whole
  const int part = int(u & 1);
  const long pu = u >> 1;
  const int k = int(pu % (T.K + 1));
  const int t = int(pu / (T.K + 1));
  const int tl = t % T.ntb, vb = t / T.ntb;
  const int v = t / T.nzb;
  const int* gi = T.gint + 8 * v;
  const double* gd = T.gdbl + 4 * v;
  const int Mdim = gi[0], Min = gi[1], rL = gi[2], lk = gi[4], rk = gi[5];
  const R aL = R(gd[0]), bL = R(gd[1]), aR = R(gd[2]), bR = R(gd[3]);
  // a = G^T x on the fly (selects: lanes of one wave can belong to variables
  // with different folds)
  const R xf0 = R(y[u]), xf1 = R(y[nlanes + u]);
  const R xl1 = R(y[long(Min - 1) * nlanes + u]), xl2 = R(y[long(Min - 2) * nlanes + u]);
  // R3X: the background coefficients of this lane's line are added back
  // (SAtransform!, src/CubicBSpline.jl:1386); both lanes of k = 0 solve line 0
  const double* ah = (T.ahat && gi[6]) ? T.ahat + (long(t) * T.nl + (k == 0 ? 0 : 2 * k - 1 + part)) * T.b_iDim
                                       : nullptr;
  auto a_at = [&](int j) -> R {
    const bool in = j >= rL && j < rL + Min;
    const R yj = R(y[long(in ? j - rL : 0) * nlanes + u]);
    R a = in ? yj : R(0);
    a = (lk == 1 && j == 0) ? aL * xf0 + bL * xf1 : a;
    a = (lk == 2 && j == 0) ? aL * xf0 : a;
    a = (lk == 2 && j == 1) ? bL * xf0 : a;
    a = (rk == 1 && j == Mdim - 1) ? aR * xl1 + bR * xl2 : a;
    a = (rk == 2 && j == Mdim - 2) ? bR * xl1 : a;
    a = (rk == 2 && j == Mdim - 1) ? aR * xl1 : a;
    if (ah) a += R(ah[j]);
    return a;
  };
  // cylindrical rings: ring i has ri = i + 1 + off wavenumbers and starts at row
  // phaseOff[i] + i = i (i + 1) / 2 + i (off + 1) of its variable block; the cap
  // of every ring is min(cap, ri), so k <= ri is on iff k <= the largest ring's
  // cap.  Spherical rings (T.sph) read ri, the row and the cap from the tables.
  const bool tab = T.sph != 0;
  const int off = T.ringRi[0] - 1;
  const int kon = T.kmax[v * T.iDim + T.iDim - 1];
  auto ring_ri = [&](int i) -> int { return tab ? T.ringRi[i] : i + 1 + off; };
  auto ring_row = [&](int i) -> long { return tab ? long(T.phaseOff[i]) + i : long(i) * (i + 1) / 2 + long(i) * (off + 1); };
  auto ring_ph = [&](int i) -> long { return tab ? long(T.phaseOff[i]) : long(i) * (i - 1) / 2 + long(i) * (off + 1); };
  const int col = (part * 3) * T.ntb + tl;           // (P|Q, d = 0, tl) column of the ring block
  S* Xv = X + long(vb) * T.Ktot * 6 * T.ntb;
  const int cs = (T.nc + nseg - 1) / nseg;
  const int c0 = seg * cs, c1 = min(T.nc, c0 + cs);
  if (c0 >= c1) return;                              // the whole pair leaves together
  // the phase factors of a cell's points are loaded one cell ahead
  R pcn[kSbMaxMu], psn[kSbMaxMu];
  auto phases = [&](int mc, R* pc, R* ps) {
#pragma unroll
    for (int mu = 0; mu < kSbMaxMu; mu++) {
      const int i = mc * T.mubar + mu;
      const bool ld = mc < c1 && mu < T.mubar && k > 0 && k <= ring_ri(min(i, T.iDim - 1));
      const long ph = ring_ph(min(i, T.iDim - 1)) + (k > 0 ? k - 1 : 0);   // phaseOff[i] + k - 1
      pc[mu] = ld ? tb.phaseC[ph] : R(1);
      ps[mu] = ld ? tb.phaseS[ph] : R(0);
    }
  };
  phases(c0, pcn, psn);
  R w0 = a_at(c0), w1 = a_at(c0 + 1), w2 = a_at(c0 + 2), w3 = a_at(c0 + 3);
  for (int mc = c0; mc < c1; mc++) {
    R pcc[kSbMaxMu], pss[kSbMaxMu];
#pragma unroll
    for (int mu = 0; mu < kSbMaxMu; mu++) { pcc[mu] = pcn[mu]; pss[mu] = psn[mu]; }
    phases(mc + 1, pcn, psn);
    const R w4 = (mc + 4 < Mdim) ? a_at(mc + 4) : R(0);
    auto point = [&](int mu, R cc, R ss) {
      const int i = mc * T.mubar + mu;
      const int ri = ring_ri(i);
      if (k > ri) return;                            // same for both lanes of a pair
      const bool on = k == 0 || k <= (tab ? T.kmax[v * T.iDim + i] : kon);
      S* xr = Xv + ring_row(i) * 6 * T.ntb + long(col) * (ri + 1) + k;
#pragma unroll
      for (int d = 0; d < 3; d++) {
        const auto* wt = const_tab(tb.siw) + long(d) * T.iDim * 4 + 4 * i;
        R ud = 0;
        ud += wt[0] * w0;
        ud += wt[1] * w1;
        ud += wt[2] * w2;
        ud += wt[3] * w3;
        const R other = __shfl_xor(ud, 1);
        const R br = part ? other : ud, bi = part ? ud : other;
        // k = 0: (value, 0); k >= 1: (c br - s bi, -(s br + c bi)); zero above the cap
        const R P = on ? (k == 0 ? br : cc * br - ss * bi) : R(0);
        const R Q = (on && k > 0) ? -(ss * br + cc * bi) : R(0);
        xr[long(d) * T.ntb * (ri + 1)] = S(part ? Q : P);
      }
    };
#pragma unroll
    for (int mu = 0; mu < kSbMaxMu; mu++) {
      if (mu >= T.mubar) break;
      point(mu, pcc[mu], pss[mu]);
    }
    for (int mu = kSbMaxMu; mu < T.mubar; mu++) {   // :regular rules with more points per cell
      const int i = mc * T.mubar + mu;
      const bool ld = k > 0 && k <= ring_ri(i);
      const long ph = ring_ph(i) + (k > 0 ? k - 1 : 0);
      point(mu, ld ? tb.phaseC[ph] : R(1), ld ? tb.phaseS[ph] : R(0));
    }
    w0 = w1; w1 = w2; w2 = w3; w3 = w4;
  }
}

template <bool FacLds>
__global__ void __launch_bounds__(kBlock) k_sa_ring(DevTables T, const double* __restrict__ spec,
                                                    double* __restrict__ y) {
  sa_ring_body<FacLds>(T, spec, y, int(blockIdx.x));
}

template <typename S>
__global__ void __launch_bounds__(kBlock) k_si_ring(DevTables T, int nseg, const double* __restrict__ y,
                                                    S* __restrict__ X) {
  si_ring_body<S>(T, nseg, y, X, int(blockIdx.x));
}

// batched: y is the job's solve scratch (x in double)
__global__ void __launch_bounds__(kBlock) k_sa_ring_batch(const SaJob* __restrict__ jobs,
                                                          const int* __restrict__ start, int njobs) {
  int lin;
  const SaJob& j = jobs[batch_job(start, njobs, lin)];
  if (sa_fac_lds(j.T)) sa_ring_body<true>(j.T, j.spec, j.y, lin);   // job-uniform
  else sa_ring_body<false>(j.T, j.spec, j.y, lin);
}

template <typename S>
__global__ void __launch_bounds__(kBlock) k_si_ring_batch(const SiJob<S>* __restrict__ jobs,
                                                          const int* __restrict__ start, int njobs) {
  int lin;
  const SiJob<S>& j = jobs[batch_job(start, njobs, lin)];
  si_ring_body<S>(j.T, j.nseg, j.y, j.X, lin);
}

int si_segments(const DevTables& T) {
  if (T.tune.si_seg == 1 || T.tune.si_seg == 2 || T.tune.si_seg == 4 || T.tune.si_seg == 8) return T.tune.si_seg;
  return T.nc >= 128 ? 8 : T.nc >= 64 ? 4 : T.nc >= 32 ? 2 : 1;
}
int sa_ring_blocks(const DevTables& T) { return int((2 * long(T.ntr) * (T.K + 1) + kBlock - 1) / kBlock); }
int si_ring_blocks(const DevTables& T, int nseg) {
  const long nlp = (2 * long(T.ntr) * (T.K + 1) + 63) / 64 * 64;   // lanes per segment, whole waves
  return int((nlp * nseg + kBlock - 1) / kBlock);
}

template <typename S>
void launch_sa_si_ring(const DevTables& T, const double* spec, double* y, S* X, hipStream_t st) {
  if (sasi_supported(T)) {
    launch_sasi<S>(T, spec, X, st);
    return;
  }
  const dim3 g1(sa_ring_blocks(T));
  const int nseg = si_segments(T);
  const dim3 g2(si_ring_blocks(T, nseg));
  if (sa_fac_lds(T)) hipLaunchKernelGGL(k_sa_ring<true>, g1, dim3(kBlock), 0, st, T, spec, y);
  else hipLaunchKernelGGL(k_sa_ring<false>, g1, dim3(kBlock), 0, st, T, spec, y);
  hipLaunchKernelGGL((k_si_ring<S>), g2, dim3(kBlock), 0, st, T, nseg, y, X);
}
template void launch_sa_si_ring<double>(const DevTables&, const double*, double*, double*, hipStream_t);
template void launch_sa_si_ring<float>(const DevTables&, const double*, double*, float*, hipStream_t);

void launch_sa_ring_only(const DevTables& T, const double* spec, double* y, hipStream_t st) {
  if (sa_fac_lds(T)) hipLaunchKernelGGL(k_sa_ring<true>, dim3(sa_ring_blocks(T)), dim3(kBlock), 0, st, T, spec, y);
  else hipLaunchKernelGGL(k_sa_ring<false>, dim3(sa_ring_blocks(T)), dim3(kBlock), 0, st, T, spec, y);
}

void launch_sa_ring_batch(const SaJob* jobs, const int* start, int njobs, int nblocks, hipStream_t st) {
  hipLaunchKernelGGL(k_sa_ring_batch, dim3(nblocks), dim3(kBlock), 0, st, jobs, start, njobs);
}

template <typename S>
void launch_si_ring_batch(const SiJob<S>* jobs, const int* start, int njobs, int nblocks, hipStream_t st) {
  hipLaunchKernelGGL(k_si_ring_batch<S>, dim3(nblocks), dim3(kBlock), 0, st, jobs, start, njobs);
}
template void launch_si_ring_batch<double>(const SiJob<double>*, const int*, int, int, hipStream_t);
template void launch_si_ring_batch<float>(const SiJob<float>*, const int*, int, int, hipStream_t);

// ----------------------------------------------------------------------------
// Inverse Chebyshev: out[z] = sum_zb M[z][zb] in[zb] per column, with M the
// composed CA + CI (+ CIx / CIxx) matrix of the variable (src/Chebyshev.jl:
// 473-535, 636-722).  Tiles of 64 columns; output staged in LDS and written as
// one contiguous span per slot.
// RZ (src/transforms_cartesian.jl:700-720): slot1 = M0 d0, slot4 = M1 d0,
//   slot5 = M2 d0, slot2 = M0 d1, slot3 = M0 d2.
// RLZ (src/transforms_cylindrical.jl:819-844): slot1 = M0 f0, slot6 = M1 f0,
//   slot7 = M2 f0, slot4 = M0 f1, slot5 = M0 f2, slot2 = M0 f3, slot3 = M0 f4.
// ----------------------------------------------------------------------------
__global__ void __launch_bounds__(kBlock) k_cheb_inv(DevTables T, const double* __restrict__ src, int ncol,
                                                     double* __restrict__ phys) {
  extern __shared__ double sm[];
  const int v = blockIdx.y, bk = T.bk, kDim = T.kDim;
  const long col0 = long(blockIdx.x) * kTile;
  const int ncl = int(min(long(kTile), ncol - col0));
  double* in_t = sm;                 // [zb][64]
  double* out_t = sm + bk * kTile;   // [c][kDim + 1]
  const int ld = kDim + 1;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const bool rlz = T.geometry == 3;
  const int nout = rlz ? 7 : 5;
  // (slot, family, matrix)
  const int rz_tab[5][3] = {{0, 0, 0}, {3, 0, 1}, {4, 0, 2}, {1, 1, 0}, {2, 2, 0}};
  const int rlz_tab[7][3] = {{0, 0, 0}, {5, 0, 1}, {6, 0, 2}, {3, 1, 0}, {4, 2, 0}, {1, 3, 0}, {2, 4, 0}};
  int cur_f = -1;
  for (int o = 0; o < nout; o++) {
    const int slot = rlz ? rlz_tab[o][0] : rz_tab[o][0];
    const int f = rlz ? rlz_tab[o][1] : rz_tab[o][1];
    const int mi = rlz ? rlz_tab[o][2] : rz_tab[o][2];
    if (f != cur_f) {
      __syncthreads();
      const double* base = src + ((long(f) * T.nvars + v) * bk) * ncol + col0;
      for (int e = threadIdx.x; e < bk * kTile; e += blockDim.x) {
        const int zb = e / kTile, c = e - zb * kTile;
        in_t[e] = c < ncl ? base[long(zb) * ncol + c] : 0.0;
      }
      cur_f = f;
      __syncthreads();
    }
    const double* M = T.inv + (long(v) * 3 + mi) * kDim * bk;
    for (int z = w; z < kDim; z += kBlock / 64) {
      double acc = 0.0;
      for (int zb = 0; zb < bk; zb++) acc += M[z * bk + zb] * in_t[zb * kTile + lane];
      out_t[lane * ld + z] = acc;
    }
    __syncthreads();
    double* dst = phys + (long(slot) * T.nvars + v) * T.npts + col0 * kDim;
    for (int e = threadIdx.x; e < ncl * kDim; e += blockDim.x) {
      const int c = e / kDim, z = e - c * kDim;
      dst[e] = out_t[c * ld + z];
    }
    __syncthreads();
  }
}

void launch_cheb_inv(const DevTables& T, const double* src, int ncol, double* phys, hipStream_t st) {
  dim3 grid((ncol + kTile - 1) / kTile, T.nvars);
  size_t lds = sizeof(double) * (T.bk * kTile + kTile * (T.kDim + 1));
  hipLaunchKernelGGL(k_cheb_inv, grid, dim3(kBlock), lds, st, T, src, ncol, phys);
}

// ----------------------------------------------------------------------------
// Tiles: sumSpectralTile! / setSpectralTile! (src/tiling.jl:1048-1170, 1204-1327)
// on the ring grids.  A tile's line p of Chebyshev block zb of variable v (lines
// k0, re 1, im 1, ..., as the patch's, only fewer: 1 + 2 K_t of them) goes to the
// patch's same line at radial coefficients siL - 1 .. siL - 2 + b_iDim_t.  One
// thread per tile element; consecutive threads take consecutive coefficients.
// ----------------------------------------------------------------------------
__global__ void k_tile_embed(double* __restrict__ patch, const double* __restrict__ tile, long n, int bt, int nlt,
                             int nzb, int bp, int nlp, long nspec_p, long nspec_t, int off, int add) {
  const long e = long(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const int j = int(e % bt);
  const long r = e / bt;                             // (v, zb, p) of the tile line
  const int pl = int(r % nlt);
  const long vz = r / nlt;
  const int zb = int(vz % nzb), v = int(vz / nzb);
  const double x = tile[v * nspec_t + (long(zb) * nlt + pl) * bt + j];
  double* d = patch + v * nspec_p + (long(zb) * nlp + pl) * bp + off + j;
  *d = add ? *d + x : x;
}

__global__ void k_fill_zero(double* __restrict__ a, long n) {
  const long e = long(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e < n) a[e] = 0.0;
}

void launch_tile_embed(double* patch, const double* tile, int nvars, int nzb, int bp, int nlp, long nspec_p, int bt,
                       int nlt, long nspec_t, int off, bool add, hipStream_t st) {
  if (!add) {
    const long np = long(nvars) * nspec_p;
    hipLaunchKernelGGL(k_fill_zero, dim3(unsigned((np + 255) / 256)), dim3(256), 0, st, patch, np);
  }
  const long n = long(nvars) * nzb * nlt * bt;
  if (n == 0) return;
  hipLaunchKernelGGL(k_tile_embed, dim3(unsigned((n + 255) / 256)), dim3(256), 0, st, patch, tile, n, bt, nlt, nzb,
                     bp, nlp, nspec_p, nspec_t, off, add ? 1 : 0);
}

}  // namespace ssg
