// ssg_kernels.hpp -- device-side table view and the launchers of every stage
// kernel of the transform path.
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

namespace ssg {

// Blocks of a 1-D grid as (tile, unit) pairs, all tiles of one unit on one
// XCD.  The dispatcher deals linear workgroup ids round-robin over the 8 XCDs
// of an MI355X, so units are dealt to XCDs in groups of 8 (unit u on XCD
// u % 8) and every unit keeps its natural place in the dispatch order (the
// blocks of unit group g come before those of group g + 1).  Tiles of a unit
// read the same data, which then stays in one XCD's L2.  Grid size:
// 8 * ntiles * ceil(nunits / 8); blocks with unit >= nunits have no work.
__device__ inline void xcd_unit_tile(int lin, int ntiles, int& unit, int& tile) {
  const int g = lin / (8 * ntiles), rem = lin - g * 8 * ntiles;
  tile = rem >> 3;
  unit = 8 * g + (rem & 7);
}
inline unsigned xcd_unit_grid(int ntiles, int nunits) { return unsigned(8 * ntiles * ((nunits + 7) / 8)); }

// Batched launches: one launch runs the blocks of many jobs (one job = one
// grid's stage); job g owns blocks [start[g], start[g + 1]), and every start
// is a multiple of 8 so that the XCD of a block is that of its local index.
// Returns the job of this block and its local block index.
__device__ inline int batch_job(const int* __restrict__ start, int njobs, int& lblk) {
  const int b = int(blockIdx.x);
  int lo = 0, hi = njobs - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (start[mid] <= b) lo = mid; else hi = mid - 1;
  }
  lblk = b - start[lo];
  return lo;
}

// The tables a kernel of arithmetic type R reads (the double tables; for the
// FP32 form's synthesis, float copies of its twiddles and Chebyshev matrices).
template <typename R>
struct RTab {
  const R* sbw; const R* siw; const R* phaseC; const R* phaseS; const R* twid; const R* fwd; const R* invt;
  const R* fftChirp; const R* fftH; const R* fftTw;   // chirp-z synthesis tables (ssg_ring_fft.hip)
};

// Padded extents of the RLZ inverse Chebyshev contraction in the ring synthesis:
// b_kDim rounded up to the MFMA k-step (4), kDim to the MFMA tile (16).
__host__ __device__ inline int syn_bkp(int nzb) { return (nzb + 3) / 4 * 4; }
__host__ __device__ inline int syn_kzp(int kDim) { return (kDim + 15) / 16 * 16; }

// Launch-shape choices a tuner may force per plan (0 = built-in heuristic).
struct Tuning {
  int syn_ct;     // ring synthesis: 1 = one column tile x two quad tiles per wave, 2 = two x one
  int syn_nwq;    // ring synthesis: largest number of quad groups per block (1, 2, 4)
  int ana_mr;     // ring analysis: row tiles per wave (1, 2)
  int si_seg;     // ring SI: cell segments per lane (1, 2, 4, 8)
  int sa_mode;    // spline solve: 0 or 2 = FP64 substitution, the only one (kept for the ABI's
                  // five tuning ints; an FP32 substitution cannot give the FP64 solution the
                  // evaluation of d/dr and d2/dr2 needs)
};


// Device pointers to the plan tables (uploaded once at plan creation).
struct DevTables {
  int geometry, nvars, nc, mubar, iDim, b_iDim, K, kDim, bk, nzb, nl, jDim, b_jDim, maxMin;
  long npts, nspec, nlines;
  // ring synthesis input X[vb][ring][P|Q][d][tl][k]: ntb transforms per block
  // (RL: the variables, RLZ: the Chebyshev modes of one variable), nvb blocks,
  // Ktot = sum over rings of (ri + 1) wavenumbers; ring r starts at phaseOff[r] + r
  int ntr, ntb, nvb, Ktot;
  // spherical rings: ring sizes follow sin(theta), so ring offsets and caps are
  // read from the tables rather than formed from the ring index
  int sph;
  // R3X background coefficients (spectral layout) or null; the SI stages add
  // them to a = G^T x for variables with an R3X side
  const double* ahat;
  const double* pqb;    // nvars * Mdim * 4: band of the unfolded P + Q
  // B-spline
  const double* sbw;    // iDim * 4
  const double* siw;    // 3 * iDim * 4
  const double* L;      // nvars * maxMin * 4 (column-major band per var)
  const double* Linv;   // nvars * maxMin
  const double* sasiA;  // nvars * 2 * kSasiMaxPart * 9: unforced partition maps of the fused solve (or null)
  const double* sasiR;  // nvars * kSasiMaxPart * kSasiRows * 8: factor rows of the fused solve by coefficient
  const int* gint;      // nvars * 8: Mdim, Min, rankL, rankR, lk, rk, r3x
  const double* gdbl;   // nvars * 4: aL, bL, aR, bR
  // rings
  const int* ringN; const int* ringRi; const int* ringOff; const int* coefOff; const int* phaseOff; const int* twOff;
  const int* kmax;      // nvars * iDim
  const double* phaseC; const double* phaseS; const double* twid;
  const double* phaseCS;  // (phaseC, phaseS) interleaved: one 16-byte load per entry (the fused SA + SI)
  // Chebyshev
  const double* fwd;    // nvars * bk * kDim
  const double* inv;    // nvars * 3 * kDim * bk (RZ)
  const double* invt;   // RLZ: nvars * 3 * bkp * kzp, invt[v][m][zb][z] = M_m[z][zb], zero padded
  RTab<float> f32;      // float copies (FP32 form)
  Tuning tune;
  // chirp-z ring synthesis (ssg_ring_fft.hip): per ring log2 M and the offset of H;
  // chirp (re, im) at ringOff[r] + n; H (re, im) at fftHOff[r] + j; twiddles of kFftMaxM
  const int* fftL; const int* fftHOff;
  const double* fftChirp; const double* fftH; const double* fftTw;
};

template <typename R>
__device__ __forceinline__ RTab<R> rtab(const DevTables& T);
template <>
__device__ __forceinline__ RTab<double> rtab<double>(const DevTables& T) {
  return RTab<double>{T.sbw, T.siw, T.phaseC, T.phaseS, T.twid, T.fwd, T.invt, T.fftChirp, T.fftH, T.fftTw};
}
template <>
__device__ __forceinline__ RTab<float> rtab<float>(const DevTables& T) { return T.f32; }

// Forward Chebyshev stage: physical columns (kDim contiguous) -> cb[v][zb][col] (RZ).
void launch_cheb_fwd(const DevTables& T, const double* phys, int ncol, double* cb, hipStream_t st);
// B-spline Gauss projection of every line into spectral (RZ input cb, FP64).
void launch_sb(const DevTables& T, const double* src, double* spec, hipStream_t st);
// Ring grids: the same from the ring coefficients rc (FP64 in both forms).
void launch_sb_ring(const DevTables& T, const double* rc, double* spec, hipStream_t st);
// Gamma fold + banded LL^T solve + evaluation of d0/d1/d2 at the mish points (RZ).
void launch_sa_si(const DevTables& T, const double* spec, double* y, double* dst, hipStream_t st);
// The same for RL / RLZ with the inverse phase filter: FP64 solve and FP64
// evaluation in both forms, stored into X of type S (double, or float for the
// FP32 form's synthesis).
template <typename S>
void launch_sa_si_ring(const DevTables& T, const double* spec, double* y, S* X, hipStream_t st);
// The same in one kernel (ssg_sasi.hip): the solve partitioned over the waves of a
// block (kSasiRows coefficients per wave), no solve scratch in HBM; for grids with at
// most kSasiMaxPart partitions and at least 64 solve lanes per variable.
static constexpr int kSasiRows = 32;
static constexpr int kSasiMaxPart = 8;
bool sasi_supported(const DevTables& T);
// The fused solve's tables per variable, from the band factor L (nvars * maxMin * 4),
// Linv and each variable's interior size Min and left rank (coefficient j holds row
// j - rankL; partition q holds coefficients q kSasiRows .. q kSasiRows + kSasiRows - 1):
//   maps: the unforced maps A_q of every partition, forward then backward (9 each);
//   rows: per coefficient, L[i][i-1], L[i][i-2], L[i][i-3], 1/L[i][i] (forward) and
//         L[i+1][i], L[i+2][i], L[i+3][i], 1/L[i][i] (backward), zero where the entry
//         or the row i lies outside the interior.
struct SasiTables { std::vector<double> maps, rows; };
SasiTables sasi_tables(const std::vector<double>& L, const std::vector<double>& Linv, const std::vector<int>& Min,
                       const std::vector<int>& rankL, int maxMin);
template <typename S>
void launch_sasi(const DevTables& T, const double* spec, S* X, hipStream_t st);
// The tile pipeline's halves (FP64): SA only, spectral B -> the folded coefficients A
// in the layout of spectral (splineTransform!); and SI from A (no solve, no fold) into X
// (tileTransform!, then launch_ring_syn).
void launch_sasi_to_a(const DevTables& T, const double* spec, double* aout, hipStream_t st);
// sumSpectralTile! (add) / setSpectralTile! (zero the patch, then copy) on ring grids:
// the tile's lines into the patch's at radial coefficient offset off = siL - 1.
void launch_tile_embed(double* patch, const double* tile, int nvars, int nzb, int bp, int nlp, long nspec_p, int bt,
                       int nlt, long nspec_t, int off, bool add, hipStream_t st);
void launch_sasi_from_a(const DevTables& T, const double* a, double* X, hipStream_t st);
// Inverse Chebyshev stage (value, d/dz, d2/dz2 matrices) into the physical slots (RZ).
void launch_cheb_inv(const DevTables& T, const double* src, int ncol, double* phys, hipStream_t st);
// Ring synthesis on MFMA from X in arithmetic R, physical of type R (RL: into
// physical slots; RLZ: fused with the inverse Chebyshev stage into physical slots).
template <typename R>
void launch_ring_syn(const DevTables& T, const R* X, R* phys, hipStream_t st);
// Ring synthesis by chirp-z transforms (ssg_ring_fft.hip).  Two of a variable's
// 5 nzb synthesis columns of a ring share one complex transform of length M = 2^L;
// a block of 256 threads runs 4096 / M of them at once.  The block table holds
// (ring, first pair) per block, largest transforms first, every size class padded
// to whole groups of 8 so that a ring's blocks share an XCD (ring -1: no work);
// a block runs the kernel body of its ring's size.  Output: RL straight into physical slots 1, 4, 5, 2, 3;
// RLZ into Y[v][f * ntb + zb][jDim] for launch_syn_cheb.
struct FftSynBlocks {
  std::vector<int> blk;        // 2 ints per block
  int nblocks = 0;
};
// the table of npair transforms per ring (synthesis: nvars ceil(5 nzb / 2) pairs;
// analysis: RLZ nvars ceil(nzb / 2) pairs, RL nvars single columns)
FftSynBlocks fft_blocks(const std::vector<int>& fftL, int iDim, int npair);
inline int syn_fft_pairs(int nvars, int nzb) { return nvars * ((5 * nzb + 1) / 2); }
inline int ana_fft_pairs(int nvars, int nzb, bool rlz) { return nvars * (rlz ? (nzb + 1) / 2 : 1); }
// Ring analysis by chirp-z transforms (RL: physical slot 1 of type S; RLZ: cm)
// into the ring coefficients rc, the same table layout as the synthesis.
template <typename S>
void launch_ana_fft(const DevTables& T, const int* blk, int nblocks, const S* src, double* rc, hipStream_t st);
template <typename S> struct AnaFftJob { DevTables T; const int* blk; const S* src; double* rc; };
template <typename S>
void launch_ana_fft_batch(const AnaFftJob<S>* jobs, const int* start, int njobs, int nblocks, hipStream_t st);
template <typename S, typename Yt>
void launch_syn_fft(const DevTables& T, const int* blk, int nblocks, const S* X, Yt* out, hipStream_t st);
// RLZ: physical slots from Y, per point (value, d/dz, d2/dz2 of family 0; values of 1..4)
template <typename R>
void launch_syn_cheb(const DevTables& T, const R* Y, R* phys, hipStream_t st);
int syn_cheb_blocks(const DevTables& T);
// kernel class of the plan's b_kDim (1: up to 16 modes, 2: up to 32, 3: up to 48)
int syn_cheb_class(const DevTables& T);
size_t syn_cheb_lds(const DevTables& T, size_t esize);
// batched forms: a job is one grid's blocks of its plan's table (FftJob) or its
// inverse Chebyshev stage
template <typename S, typename Yt> struct FftJob { DevTables T; const int* blk; const S* X; Yt* out; };
template <typename R> struct ChebSynJob { DevTables T; const R* Y; R* phys; };
template <typename S, typename Yt>
void launch_syn_fft_batch(const FftJob<S, Yt>* jobs, const int* start, int njobs, int nblocks, hipStream_t st);
template <typename R>
void launch_syn_cheb_batch(const ChebSynJob<R>* jobs, const int* start, int njobs, int nblocks, size_t lds, int cls,
                           hipStream_t st);
// Forward Chebyshev stage of every RLZ ring point: physical slot 1 (type In) ->
// cm[v][zb][point] (FP64).
template <typename In>
void launch_cheb_fwd_pts(const DevTables& T, const In* phys, double* cm, hipStream_t st);
// Forward ring DFT on MFMA (RL: from physical slot 1 of type In; RLZ: from cm)
// into ring coefficients rc[v*nzb+zb][b_jDim] of type R.
template <typename R, typename In>
void launch_ring_ana(const DevTables& T, const In* src, R* rc, hipStream_t st);
bool ring_ana_supported(const DevTables& T);

// ---------------------------------------------------------------------------
// Launch shapes and batched launches (many grids per launch).  A job is one
// grid's share of a stage; jobs of one launch share the kernel instance, the
// block size and the dynamic LDS size (the largest of the jobs).
// ---------------------------------------------------------------------------
struct SynLaunch {
  int ntb;        // transforms per block: RL nvars, RLZ b_kDim
  int ncols;      // 5 * ntb
  int nctp;       // column tiles of 16, padded to nwc * CT
  int ldp;        // staged row stride (elements)
  int nwq;        // quad groups (QT tiles of 16 quads each) per block
  int nwc;        // column groups (CT tiles each) per block
  int qpb;        // quads per block = 16 * QT * nwq (a power of two)
  int lqpb;       // log2(qpb)
  int bkp;        // b_kDim padded to a multiple of 4 (RLZ)
  int kzp;        // kDim padded to a multiple of 16 (RLZ)
  int ys;         // row stride of the synthesised-value tile (elements)
  int nslot;      // staged elements per thread per chunk
  int ct, qt;     // template shape
  int nqt;        // quad tiles of the largest ring
  int nthr;       // threads per block
  int nblocks;    // blocks of one grid
  bool ok;        // shape fits the kernel's limits
  size_t lds;     // bytes of dynamic LDS
};
SynLaunch plan_syn(const DevTables& T, size_t esize);

struct AnaLaunch {
  int ncols;      // RLZ b_kDim, RL nvars
  int nct;        // column tiles
  int mr;         // row tiles per wave
  int ldc;        // row stride of the folded tiles (elements)
  int nkb_max;    // k-blocks of the largest ring
  int nblocks;    // blocks of one grid
  size_t lds;
};
AnaLaunch plan_ana(const DevTables& T, size_t esize);

template <typename R> struct SynJob { DevTables T; SynLaunch P; const R* X; R* phys; };
template <typename R, typename In> struct AnaJob { DevTables T; AnaLaunch P; const In* src; R* rc; };
template <typename In> struct ChebJob { DevTables T; const In* phys; double* cm; int kzp, bkp, nbx; };
struct SbJob { DevTables T; const double* rc; double* spec; };
struct SaJob { DevTables T; const double* spec; double* y; };
template <typename S> struct SaSiJob { DevTables T; const double* spec; S* X; };
template <typename S> struct SiJob { DevTables T; int nseg; const double* y; S* X; };

// Regular-grid evaluation (regularGridTransform): the banded solve alone (x
// left in y), SI at arbitrary radii, and a batched strided FP64 GEMM.
void launch_sa_ring_only(const DevTables& T, const double* spec, double* y, hipStream_t st);   // RL, RLZ
void launch_sa_lines(const DevTables& T, const double* spec, double* y, hipStream_t st);       // RZ
void launch_si_pts(const DevTables& T, bool ring, const double* y, int npt, const int* pxm, const double* pw,
                   double* S, hipStream_t st);
struct GemmDesc {
  const double* A; const double* B; double* C;
  int M, N, K, batch;
  long ai, ak, bk, bj, ci, cj;   // element (i, k) of A at A[i ai + k ak], and so on
  long sA, sB, sC;               // batch strides
  int msplit;                    // > 0: row i of C at (i % msplit) ci + (i / msplit) ci2
  long ci2;
};
void launch_gemm_f64(const GemmDesc& g, hipStream_t st);
// R3X: btilde = b - pq ahat for the lines of variables with an R3X side, b elsewhere
void launch_ahat_btilde(const DevTables& T, const double* spec, double* bt, hipStream_t st);
// applyFilter!: nact spectral lines of b_iDim entries starting at off[a]
// (device layout), multiplied by w[a], or set to +0.0 where w[a] == 0
void launch_apply_filter(double* spec, const long* off, const double* w, long nact, int b_iDim, hipStream_t st);
// Unstructured evaluation: folded coefficients of every line, the grouped
// point products, the Chebyshev fold and series per point, the fill of points
// outside the domain.
void launch_fold_lines(const DevTables& T, bool ring, const double* y, double* A2, hipStream_t st);
struct UnstrDesc {
  const double* A2;              // [(j * nks + ks) * ntr + t]
  int nks, ntr, b_iDim;
  const int* gstart; const int* gcount; const int* gxm;   // per group: first sorted point, count, first tap
  const double* x;               // [sorted point] radius (x)
  double xmin, DX, DXr; int nc;  // the spline's knots
  const double* lam;             // [sorted point] angles (RL, RLZ)
  double* C; long cbase;         // RZ / RLZ: C[(n - cbase) * ntr + t]
  double* out; const int* perm; long nall;   // RL: out[t * nall + perm[n]]
};
void launch_unstr_gemm(const UnstrDesc& d, int ngroups, hipStream_t st);
struct UnstrCheb {
  const double* C; long npt; int ntr, nzb, kDim, nv;
  const double* ca;              // [v][k][zb]
  const double* z; long cbase;   // z[cbase + n]
  double zmin, zmax;
  double* out; const int* perm; long nall;
};
void launch_unstr_cheb(const UnstrCheb& d, hipStream_t st);
void launch_unstr_fill(const long* idx, long n, int nv, long nall, double fill, double* out, hipStream_t st);
// trig[(dl * (2K + 1) + ks) * nj + j]: row ks = 0 (k = 0), 2k - 1 (cosine line
// of k), 2k (sine line) of the Fourier sum at lambda_j and its dl-th derivative
void launch_trig_table(const double* lam, int nj, int K, double* trig, hipStream_t st);

// one grid's block count of each batched stage
int cheb_fwd_pts_blocks(const DevTables& T);
int sb_blocks(const DevTables& T);
size_t sb_lds(const DevTables& T);
int sa_ring_blocks(const DevTables& T);
int si_ring_blocks(const DevTables& T, int nseg);
int si_segments(const DevTables& T);
size_t cheb_fwd_pts_lds(const DevTables& T, size_t esize);

template <typename R>
void launch_ring_syn_batch(const SynJob<R>* jobs, const int* start, int njobs, int nblocks, int ct, int nthr,
                           size_t lds, hipStream_t st);
template <typename R, typename In>
void launch_ring_ana_batch(const AnaJob<R, In>* jobs, const int* start, int njobs, int nblocks, int nct, int mr,
                           size_t lds, hipStream_t st);
template <typename In>
void launch_cheb_fwd_pts_batch(const ChebJob<In>* jobs, const int* start, int njobs, int nblocks, size_t lds,
                               hipStream_t st);
void launch_sb_ring_batch(const SbJob* jobs, const int* start, int njobs, int nblocks, size_t lds, hipStream_t st);
void launch_sa_ring_batch(const SaJob* jobs, const int* start, int njobs, int nblocks, hipStream_t st);
template <typename S>
void launch_si_ring_batch(const SiJob<S>* jobs, const int* start, int njobs, int nblocks, hipStream_t st);
int sasi_blocks(const DevTables& T);
int sasi_threads(const DevTables& T);
size_t sasi_lds(const DevTables& T);
int sasi_form(const DevTables& T);
template <typename S>
void launch_sasi_batch(const SaSiJob<S>* jobs, const int* start, int njobs, int nblocks, int nthr, size_t lds,
                       int form, hipStream_t st);

}  // namespace ssg
