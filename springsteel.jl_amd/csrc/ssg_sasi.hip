// ssg_sasi.hip -- SA + SI of the ring grids in one kernel: Gamma fold, banded
// LL^T solve, evaluation of d0 / d1 / d2 at the Gauss points and the inverse
// phase filter, from spectral straight to the synthesis input X.
//
// It covers SAtransform! (src/CubicBSpline.jl:1368-1396) with
// _band_forward_solve3! / _band_backward_solve3! (src/banded_cholesky.jl:
// 202-247), SItransform / SIxtransform / SIxxtransform (1554-1567, 1650-1654,
// 1695-1699) and apply_phasefilter_inverse! (src/phase_filter.jl:154-190), as
// k_sa_ring + k_si_ring do, without the solve scratch in HBM.
//
// Lanes.  A solve lane is a line pair (k, re | im) of one Chebyshev block t of one
// variable, ordered as in k_sa_ring; the 2 (K + 1) lanes of each t are padded to
// whole waves (tiles), so that every wave, and every block, belongs to one t and
// one variable: its boundary condition, factor and tables are then wave-uniform
// and read as scalar loads.  The tiles of a t read its spectral lines and write its
// X rows, so they are dealt to one XCD (xcd_unit_tile), whose L2 merges the pieces
// of the rows that the tiles write.
//
// Solve.  A block takes 64 lanes and splits the coefficients of each line into
// partitions of kPR, one wave per partition; every lane keeps its partition's
// coefficients j = p kPR + c in registers (the interior rows i = j - rankL among
// them).  The bandwidth-3 substitutions are chained across partitions through
// their three-row boundary states: for partition p, the state after its rows
// is an affine map S_out = A_p S_in + f_p of the state before them.  A_p (the
// rows run from unit states with no right-hand side) depends on the factor alone
// and is tabulated per variable at plan creation (sasi_tables); each wave runs
// its rows once from the zero state to get f_p, the waves exchange f in LDS,
// each wave composes the maps of the partitions before it (after it, for the
// backward solve) to get its true incoming state, and then runs its rows once
// more from that state.  The factor of an SPD banded matrix damps the
// homogeneous responses, so the composition is well conditioned; the results
// agree with the sequential recurrence to rounding.
//
// Evaluation.  SI then runs per wave on kPR consecutive cells (C0 = p kPR on): the
// taps of cell mc are the coefficients mc .. mc + 3, so a = G^T x comes from a
// register window of coefficients C0 .. C0 + kPR + 2 (the last three are the next
// partition's first, the backward chain's incoming state) and the fold's boundary
// terms (x[0], x[1], x[Min - 2], x[Min - 1], exchanged in LDS).  The lane pair of a wavenumber swaps its
// values (a DPP quad permutation), rotates them by the ring phase and stores
// X[vb][ring][P|Q][d][tl][k] as k_si_ring does: each store of a wave is a run of
// consecutive k.  A ring that no lane of the wave reaches (k > ri on every lane)
// is skipped.
#include <algorithm>
#include <utility>

#include "ssg_kernels.hpp"

namespace ssg {

static constexpr int kPR = kSasiRows;         // coefficients (and cells) per partition, registers per lane
static constexpr int kMaxPart = kSasiMaxPart; // partitions (waves) per block

// partitions of a block: the b_iDim coefficients of a line (nc + 3 > nc cells)
__host__ __device__ inline int sasi_parts(const DevTables& T) { return (T.b_iDim + kPR - 1) / kPR; }
// tiles (waves of lanes) per Chebyshev block t
__host__ __device__ inline int sasi_tiles(const DevTables& T) { return (2 * (T.K + 1) + 63) / 64; }
// the X block of one t must be addressable with 32-bit byte offsets (raw buffer)
bool sasi_supported(const DevTables& T) {
  return T.sasiA != nullptr && T.sasiR != nullptr && T.phaseCS != nullptr && sasi_parts(T) <= kMaxPart &&
         long(T.Ktot) * 6 * T.ntb * 8 < (1L << 31) - 64;
}
// the compiled form of a grid: 0 / 1 for three Gauss points (cylindrical / spherical),
// 2 / 3 for any other rule
int sasi_form(const DevTables& T) { return (T.mubar == 3 ? 0 : 2) + (T.sph ? 1 : 0); }
int sasi_threads(const DevTables& T) { return 64 * sasi_parts(T); }
int sasi_blocks(const DevTables& T) { return int(xcd_unit_grid(sasi_tiles(T), T.ntr)); }
// per partition and lane: the forced state (3); per lane: the four boundary rows of the fold
size_t sasi_lds(const DevTables& T) { return sizeof(double) * (size_t(sasi_parts(T)) * 64 * 3 + 64 * 4); }

SasiTables sasi_tables(const std::vector<double>& L, const std::vector<double>& Linv, const std::vector<int>& Min,
                       const std::vector<int>& rankL, int maxMin) {
  const int nv = int(Min.size());
  const int nj = kMaxPart * kPR;
  SasiTables tb;
  tb.rows.assign(size_t(nv) * nj * 8, 0.0);
  std::vector<double>& A = tb.maps;
  A.assign(size_t(nv) * 2 * kMaxPart * 9, 0.0);
  for (int v = 0; v < nv; v++) {
    const double* Lb = L.data() + size_t(v) * maxMin * 4;   // Lb[i * 4 + m] = L[i + m][i]
    const double* Li = Linv.data() + size_t(v) * maxMin;
    const int M = Min[v];
    for (int j = 0; j < nj; j++) {
      const int i = j - rankL[v];
      if (i < 0 || i >= M) continue;
      double* r = tb.rows.data() + (size_t(v) * nj + j) * 8;
      r[0] = i >= 1 ? Lb[(i - 1) * 4 + 1] : 0.0;
      r[1] = i >= 2 ? Lb[(i - 2) * 4 + 2] : 0.0;
      r[2] = i >= 3 ? Lb[(i - 3) * 4 + 3] : 0.0;
      r[3] = Li[i];
      r[4] = i + 1 < M ? Lb[i * 4 + 1] : 0.0;
      r[5] = i + 2 < M ? Lb[i * 4 + 2] : 0.0;
      r[6] = i + 3 < M ? Lb[i * 4 + 3] : 0.0;
      r[7] = Li[i];
    }
    for (int dir = 0; dir < 2; dir++)
      for (int q = 0; q < kMaxPart; q++) {
        double* a = A.data() + ((size_t(v) * 2 + dir) * kMaxPart + q) * 9;
        // rows of partition q: coefficients q kPR .. q kPR + kPR - 1 less rankL, interior ones
        const int r0 = std::max(q * kPR - rankL[v], 0), r1 = std::min(q * kPR + kPR - rankL[v], M);
        for (int j = 0; j < 3; j++) {
          double s[3] = {0.0, 0.0, 0.0};
          s[j] = 1.0;
          if (dir == 0) {   // state before row i: (y[i-1], y[i-2], y[i-3])
            for (int i = r0; i < r1; i++) {
              const double l1 = i >= 1 ? Lb[(i - 1) * 4 + 1] : 0.0;
              const double l2 = i >= 2 ? Lb[(i - 2) * 4 + 2] : 0.0;
              const double l3 = i >= 3 ? Lb[(i - 3) * 4 + 3] : 0.0;
              const double y = (-l1 * s[0] - l2 * s[1] - l3 * s[2]) * Li[i];
              s[2] = s[1]; s[1] = s[0]; s[0] = y;
            }
          } else {          // going down, state before row i: (x[i+1], x[i+2], x[i+3])
            for (int i = r1 - 1; i >= r0; i--) {
              const double m1 = i + 1 < M ? Lb[i * 4 + 1] : 0.0;
              const double m2 = i + 2 < M ? Lb[i * 4 + 2] : 0.0;
              const double m3 = i + 3 < M ? Lb[i * 4 + 3] : 0.0;
              const double x = (-m1 * s[0] - m2 * s[1] - m3 * s[2]) * Li[i];
              s[2] = s[1]; s[1] = s[0]; s[0] = x;
            }
          }
          for (int m = 0; m < 3; m++) a[3 * m + j] = s[m];   // A[m][j]: row m of the state, unit state j
        }
      }
  }
  return tb;
}

// a read of a table that no kernel writes, through the constant address space: a
// wave-uniform address then becomes a scalar load (the X stores of the kernel would
// otherwise keep the compiler from proving that the table is not clobbered)
template <typename V>
__device__ __forceinline__ V ld_table(const V* p) {
  return *(const __attribute__((address_space(4))) V*)p;
}

// the value of the other lane of the pair (lane ^ 1): DPP quad permutation [1, 0, 3, 2]
__device__ __forceinline__ double pair_swap(double a) {
  const unsigned long long u = __builtin_bit_cast(unsigned long long, a);
  const unsigned lo = unsigned(__builtin_amdgcn_mov_dpp(int(unsigned(u)), 0xB1, 0xF, 0xF, false));
  const unsigned hi = unsigned(__builtin_amdgcn_mov_dpp(int(unsigned(u >> 32)), 0xB1, 0xF, 0xF, false));
  return __builtin_bit_cast(double, (static_cast<unsigned long long>(hi) << 32) | lo);
}

// word 3 of a raw buffer resource on gfx9 parts (no format conversion, no swizzle)
static constexpr int kBufferWord3 = 0x00020000;

// f(integral_constant<int, O>) for O = 0 .. N-1 in order: an unrolled loop whose index is a
// compile-time constant however large the body (register arrays indexed by it stay in
// registers, which a partially unrolled loop would not guarantee)
template <typename F, int... O>
__device__ __forceinline__ void static_for_impl(F&& f, std::integer_sequence<int, O...>) {
  (f(std::integral_constant<int, O>{}), ...);
}
template <int N, typename F>
__device__ __forceinline__ void static_for(F&& f) {
  static_for_impl(f, std::make_integer_sequence<int, N>{});
}

// the row pointer, re-issued after a row's result: the factor reads of the rows that
// follow wait for it, so they stay near their use instead of all being hoisted
__device__ __forceinline__ void hold_rows(const double*& rp, double y) { asm volatile("" : "+s"(rp) : "v"(y)); }

// MODE 0: spectral (B) -> X (the grid transform's SA + SI); MODE 1: B -> the folded
// coefficients A into aout, in the layout of spectral (splineTransform!); MODE 2: A in
// `spec` -> X, no solve and no fold (tileTransform!)
template <typename S, int MB, bool SPH, int MODE>
__device__ __forceinline__ void sasi_body(const DevTables& T, const double* __restrict__ spec, S* __restrict__ X,
                                          int blk, double* __restrict__ aout) {
  extern __shared__ double sh[];
  const int NP = sasi_parts(T);
  int t, tile;                                       // block-uniform
  xcd_unit_tile(blk, sasi_tiles(T), t, tile);
  if (t >= T.ntr) return;                            // a padding block: leaves as a whole
  const int v = t / T.nzb;
  const int lane = threadIdx.x & 63;
  const int p = __builtin_amdgcn_readfirstlane(int(threadIdx.x) >> 6);   // this wave's partition
  const int loc = tile * 64 + lane;
  const bool valid = loc < 2 * (T.K + 1);            // pairs stay whole: 2 (K + 1) and 64 are even
  const int lc = min(loc, 2 * T.K + 1);
  const int part = lc & 1;
  const int k = lc >> 1;
  const int tl = t % T.ntb, vb = t / T.ntb;
  const long line = long(t) * T.nl + (k == 0 ? 0 : 2 * k - 1 + part);
  const double* __restrict__ b = spec + line * T.b_iDim;
  const int Mdim = ld_table(T.gint + 8 * v), Min = ld_table(T.gint + 8 * v + 1), rL = ld_table(T.gint + 8 * v + 2);
  const int lk = ld_table(T.gint + 8 * v + 4), rk = ld_table(T.gint + 8 * v + 5), r3x = ld_table(T.gint + 8 * v + 6);
  const double aL = ld_table(T.gdbl + 4 * v), bL = ld_table(T.gdbl + 4 * v + 1);
  const double aR = ld_table(T.gdbl + 4 * v + 2), bR = ld_table(T.gdbl + 4 * v + 3);
  const int C0 = p * kPR;                            // this wave's first coefficient (and cell)
  const int r0 = C0 - rL;                            // the row of register 0: row r0 + c is interior if in [0, Min)
  // LDS: forced states [NP][64][3] | boundary rows [64][4]
  double* fv = sh;
  double* bnd = fv + long(NP) * 64 * 3;
  // factor rows from coefficient C0 on: row c holds the forward entries of row r0 + c at
  // [8c, 8c + 4) and the backward ones at [8c + 4, 8c + 8), zero off the interior, so the
  // rows off the interior leave x zero without a select
  const double* Rg = T.sasiR + (long(v) * kMaxPart * kPR + C0) * 8;
  const double* Af = T.sasiA + long(v) * 2 * kMaxPart * 9;   // forward maps of the variable
  const double* Ab = Af + kMaxPart * 9;                      // backward maps

  double x[kPR];
  double d0 = 0.0, d1 = 0.0, d2 = 0.0;               // coefficients C0 + kPR .. C0 + kPR + 2 (the next partition's first)
  double xf0 = 0.0, xf1 = 0.0, xl2 = 0.0, xl1 = 0.0; // the fold's boundary rows x[0], x[1], x[Min - 2], x[Min - 1]
  if constexpr (MODE != 2) {
  // right-hand side g = G b on this partition's rows (the fold's boundary terms as selects)
  {
    const double bf0 = b[0], bf1 = b[1], bl1 = b[Mdim - 1], bl2 = b[Mdim - 2];
    typedef double dv2 __attribute__((ext_vector_type(2), aligned(8)));
    if (C0 + kPR <= Mdim) {                          // wave-uniform
      const dv2* p2 = reinterpret_cast<const dv2*>(b + C0);
#pragma unroll
      for (int c = 0; c < kPR / 2; c++) {
        const dv2 z = p2[c];
        x[2 * c] = z.x;
        x[2 * c + 1] = z.y;
      }
    } else {
#pragma unroll
      for (int c = 0; c < kPR; c++) x[c] = b[min(C0 + c, Mdim - 1)];
    }
    // only the edge partitions hold rows off the interior or the fold's rows (the
    // conditions are wave-uniform; selects, since branches per register would be
    // merged into a switch that copies the whole array)
    if (r0 < 0 || r0 + kPR > Min) {
#pragma unroll
      for (int c = 0; c < kPR; c++) x[c] = (r0 + c >= 0 && r0 + c < Min) ? x[c] : 0.0;
    }
    if (r0 <= 1) {
      const double g0 = lk == 1 ? aL * bf0 : lk == 2 ? aL * bf0 + bL * bf1 : 0.0;
      const double g1 = lk == 1 ? bL * bf0 : 0.0;
#pragma unroll
      for (int c = 0; c < kPR; c++) x[c] += r0 + c == 0 ? g0 : r0 + c == 1 ? g1 : 0.0;
    }
    if (r0 + kPR > Min - 2) {
      const double g2 = rk == 1 ? bR * bl1 : 0.0;
      const double g1 = rk == 1 ? aR * bl1 : rk == 2 ? aR * bl1 + bR * bl2 : 0.0;
#pragma unroll
      for (int c = 0; c < kPR; c++) x[c] += r0 + c == Min - 2 ? g2 : r0 + c == Min - 1 ? g1 : 0.0;
    }
  }
  double* fme = fv + (long(p) * 64 + lane) * 3;      // this lane's forced state of partition p

  // ---- forward L y = g.  State before row i: (y[i-1], y[i-2], y[i-3]).  A row off the
  // interior gives zero and shifts it into the state; such rows come only before row 0
  // (from the zero state), after row Min - 1 (no later row is interior) and, going
  // down, above row Min - 1 (from the zero state) or below row 0 (no state is used
  // after them), so the states that matter are those of the interior rows alone.
  {
    double f1 = 0, f2 = 0, f3 = 0;                   // the rows from the zero state
    const double* rp = Rg;
#pragma unroll
    for (int c = 0; c < kPR; c++) {
      const double* r = rp + c * 8;
      const double yf = (x[c] - ld_table(r) * f1 - ld_table(r + 1) * f2 - ld_table(r + 2) * f3) * ld_table(r + 3);
      f3 = f2; f2 = f1; f1 = yf;
      if ((c & 3) == 0) hold_rows(rp, yf);
    }
    fme[0] = f1; fme[1] = f2; fme[2] = f3;
  }
  __syncthreads();
  {
    // incoming state of partition p: compose the maps of partitions 0 .. p-1
    double s0 = 0, s1 = 0, s2 = 0;
#pragma unroll 1
    for (int q = 0; q < p; q++) {
      const double* e = fv + (long(q) * 64 + lane) * 3;
      const double* a = Af + q * 9;
      const double n0 = e[0] + ld_table(a) * s0 + ld_table(a + 1) * s1 + ld_table(a + 2) * s2;
      const double n1 = e[1] + ld_table(a + 3) * s0 + ld_table(a + 4) * s1 + ld_table(a + 5) * s2;
      const double n2 = e[2] + ld_table(a + 6) * s0 + ld_table(a + 7) * s1 + ld_table(a + 8) * s2;
      s0 = n0; s1 = n1; s2 = n2;
    }
    // this partition's rows from the true incoming state
    const double* rp = Rg;
#pragma unroll
    for (int c = 0; c < kPR; c++) {
      const double* r = rp + c * 8;
      const double y = (x[c] - ld_table(r) * s0 - ld_table(r + 1) * s1 - ld_table(r + 2) * s2) * ld_table(r + 3);
      x[c] = y;
      s2 = s1; s1 = s0; s0 = y;
      if ((c & 3) == 0) hold_rows(rp, y);
    }
  }
  __syncthreads();                                   // every wave has read the forward states
  // ---- backward L^T x = y.  State before row i (going down): (x[i+1], x[i+2], x[i+3]).
  {
    double f1 = 0, f2 = 0, f3 = 0;
    const double* rp = Rg;
#pragma unroll
    for (int c = kPR - 1; c >= 0; c--) {
      const double* r = rp + c * 8 + 4;
      const double xf = (x[c] - ld_table(r) * f1 - ld_table(r + 1) * f2 - ld_table(r + 2) * f3) * ld_table(r + 3);
      f3 = f2; f2 = f1; f1 = xf;
      if ((c & 3) == 3) hold_rows(rp, xf);
    }
    fme[0] = f1; fme[1] = f2; fme[2] = f3;
  }
  __syncthreads();
  {
    double s0 = 0, s1 = 0, s2 = 0;
#pragma unroll 1
    for (int q = NP - 1; q > p; q--) {
      const double* e = fv + (long(q) * 64 + lane) * 3;
      const double* a = Ab + q * 9;
      const double n0 = e[0] + ld_table(a) * s0 + ld_table(a + 1) * s1 + ld_table(a + 2) * s2;
      const double n1 = e[1] + ld_table(a + 3) * s0 + ld_table(a + 4) * s1 + ld_table(a + 5) * s2;
      const double n2 = e[2] + ld_table(a + 6) * s0 + ld_table(a + 7) * s1 + ld_table(a + 8) * s2;
      s0 = n0; s1 = n1; s2 = n2;
    }
    d0 = s0; d1 = s1; d2 = s2;
    const double* rp = Rg;
#pragma unroll
    for (int c = kPR - 1; c >= 0; c--) {
      const double* r = rp + c * 8 + 4;
      const double xv = (x[c] - ld_table(r) * s0 - ld_table(r + 1) * s1 - ld_table(r + 2) * s2) * ld_table(r + 3);
      x[c] = xv;                                     // zero off the interior rows
      s2 = s1; s1 = s0; s0 = xv;
      if ((c & 3) == 3) hold_rows(rp, xv);
    }
  }
  // the fold's boundary rows x[0], x[1], x[Min - 2], x[Min - 1]
  {
    double* bme = bnd + lane * 4;
    if (r0 <= 1) {                                   // wave-uniform
      double z0 = 0.0, z1 = 0.0;
#pragma unroll
      for (int c = 0; c < kPR; c++) {
        z0 = r0 + c == 0 ? x[c] : z0;
        z1 = r0 + c == 1 ? x[c] : z1;
      }
      bme[0] = z0; bme[1] = z1;
    }
    if (r0 + kPR > Min - 2) {
      double z2 = 0.0, z3 = 0.0;
#pragma unroll
      for (int c = 0; c < kPR; c++) {
        z2 = r0 + c == Min - 2 ? x[c] : z2;
        z3 = r0 + c == Min - 1 ? x[c] : z3;
      }
      if (r0 <= Min - 2) bme[2] = z2;                // the partition that holds the row writes it
      if (r0 <= Min - 1 && r0 + kPR > Min - 1) bme[3] = z3;
    }
  }
  __syncthreads();
  xf0 = bnd[lane * 4 + 0]; xf1 = bnd[lane * 4 + 1]; xl2 = bnd[lane * 4 + 2]; xl1 = bnd[lane * 4 + 3];
  } else {
    // the line's folded coefficients are given: the window straight from memory
    const int Md = T.b_iDim;
    typedef double dv2 __attribute__((ext_vector_type(2), aligned(8)));
    if (C0 + kPR <= Md) {                            // wave-uniform
      const dv2* p2 = reinterpret_cast<const dv2*>(b + C0);
#pragma unroll
      for (int c = 0; c < kPR / 2; c++) {
        const dv2 z = p2[c];
        x[2 * c] = z.x;
        x[2 * c + 1] = z.y;
      }
    } else {
#pragma unroll
      for (int c = 0; c < kPR; c++) x[c] = C0 + c < Md ? b[min(C0 + c, Md - 1)] : 0.0;
    }
    d0 = C0 + kPR < Md ? b[C0 + kPR] : 0.0;
    d1 = C0 + kPR + 1 < Md ? b[min(C0 + kPR + 1, Md - 1)] : 0.0;
    d2 = C0 + kPR + 2 < Md ? b[min(C0 + kPR + 2, Md - 1)] : 0.0;
  }

  // ---- SI: cells C0 .. C0 + kPR - 1 of this wave, taps j = mc .. mc + 3 on rows j - rL
  if (C0 >= (MODE == 1 ? Mdim : T.nc)) return;       // wave-uniform: no barrier follows
  const RTab<double> tb = rtab<double>(T);
  const double* ah = (MODE != 2 && T.ahat && r3x) ? T.ahat + line * T.b_iDim : nullptr;
  // a = G^T x in place on the window of coefficients C0 .. C0 + kPR + 2 (x[c], then
  // d0 .. d2): the solve left zero off the interior, so only the fold's coefficients
  // 0, 1, Mdim - 2, Mdim - 1 change (wave-uniform branches); R3X adds its background
  // coefficients as the window slides
  auto W = [&](int c) -> double& { return c < kPR ? x[c] : c == kPR ? d0 : c == kPR + 1 ? d1 : d2; };
  if (MODE != 2 && C0 == 0) {
    if (lk == 1) x[0] = aL * xf0 + bL * xf1;
    if (lk == 2) { x[0] = aL * xf0; x[1] = bL * xf0; }
  }
  if (MODE != 2 && C0 + kPR + 2 >= Mdim - 2) {
    const double v2 = rk == 2 ? bR * xl1 : 0.0;      // a[Mdim - 2] (R2; zero for R3)
    const double v1 = rk == 1 ? aR * xl1 + bR * xl2 : rk == 2 ? aR * xl1 : 0.0;   // a[Mdim - 1]
    const bool f2 = rk >= 2, f1 = rk >= 1;           // the coefficients the fold sets
#pragma unroll
    for (int c = 0; c < kPR + 3; c++) {
      W(c) = (f2 && C0 + c == Mdim - 2) ? v2 : W(c);
      W(c) = (f1 && C0 + c == Mdim - 1) ? v1 : W(c);
    }
  }
  // a[j] for the coefficient j = C0 + c, c in [0, kPR + 3) (a constant after unrolling)
  auto a_at = [&](int c) -> double { return (ah && C0 + c < Mdim) ? W(c) + ah[C0 + c] : W(c); };
  if constexpr (MODE == 1) {
    // the wave's own coefficients of the line (SAtransform's a, with the R3X background);
    // the k = 0 line once
    if (valid && !(k == 0 && part == 1)) {
      double* ao = aout + line * T.b_iDim;
#pragma unroll
      for (int c = 0; c < kPR; c++)
        if (C0 + c < Mdim) ao[C0 + c] = a_at(c);
    }
    return;
  }
  const int off = ld_table(T.ringRi) - 1;           // cylindrical rings: ri = i + 1 + off
  const int kon = ld_table(T.kmax + long(v) * T.iDim + T.iDim - 1);
  // X of this t as a raw buffer: element offsets are 32-bit (sasi_supported bounds the
  // block), and a lane with no entry in a ring stores to an offset past the end, which
  // the hardware drops (no branch per store); likewise for the phase-factor loads,
  // which return zero there
  constexpr unsigned OOB = 0x7ffffff0u;             // past the end of every buffer here (< 2^31 bytes)
  const long xblock = long(T.Ktot) * 6 * T.ntb;
  const __amdgpu_buffer_rsrc_t rX = __builtin_amdgcn_make_buffer_rsrc(X + long(vb) * xblock, 0,
                                                                       int(xblock * sizeof(S)), kBufferWord3);
  const int nph = T.Ktot - T.iDim;                   // phase entries: sum of ri
  const __amdgpu_buffer_rsrc_t rCS = __builtin_amdgcn_make_buffer_rsrc((void*)T.phaseCS, 0, nph * 16, kBufferWord3);
  typedef unsigned u32x2 __attribute__((ext_vector_type(2)));
  typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
  typedef double dv2c __attribute__((ext_vector_type(2)));
  // per-lane factors of the rotation, constant over the rings where the cap is (cylindrical
  // grids): sgn = +-1 on the re / im lane that keeps its entry, 0 above the cap; one = 1 on
  // the k = 0 re lane (its value passes unrotated)
  auto sgn_of = [&](int kcap) -> double { return (k == 0 || k > kcap) ? 0.0 : part ? -1.0 : 1.0; };
  const double one = (k == 0 && part == 0) ? 1.0 : 0.0;
  // ring i: its largest wavenumber ri, first phase entry ph0 (its X rows start at ph0 + i)
  // and wavenumber cap, from the tables (spherical rings) or in closed form
  auto ring_ri = [&](int i) -> int { return SPH ? ld_table(T.ringRi + i) : i + 1 + off; };
  auto ring_ph0 = [&](int i) -> int { return SPH ? ld_table(T.phaseOff + i) : i * (i - 1) / 2 + i * (off + 1); };
  // the lane's phase factors of ring i: (cos, sin) of k times the ring's phase shift, or
  // (0, 0) where the lane has no entry (k = 0 or k > ri: the load past the end returns zero)
  auto ring_phase = [&](int i, double& cc, double& ss) {
    const bool ld = valid && k > 0 && k <= ring_ri(i);
    const unsigned vo = ld ? unsigned(ring_ph0(i) + k - 1) * 16u : OOB;
    const dv2c cs = __builtin_bit_cast(dv2c, __builtin_amdgcn_raw_buffer_load_b128(rCS, int(vo), 0, 0));
    cc = cs.x;
    ss = cs.y;
  };
  // one ring i of the cell whose taps are w0 .. w3: d0 / d1 / d2 of the lane's line,
  // rotated by the ring phase (c, s); P (re lane) = c br - s bi and Q (im lane) =
  // -(s br + c bi) of (br, bi) = the pair's values, i.e. own * (+-c) - other * s; k = 0
  // keeps (value, 0); zero above the cap.  X[ring][P|Q][d][tl][k]: the (P|Q, d) rows
  // of the ring are ntb (ri + 1) apart
  const double sgn0 = sgn_of(kon);
  auto ring_out = [&](int i, double cc, double ss, double w0, double w1, double w2, double w3) {
    const int ri = ring_ri(i);
    const double sgn = SPH ? sgn_of(ld_table(T.kmax + long(v) * T.iDim + i)) : sgn0;
    const double cown = sgn * cc + one;
    const double coth = -(sgn * sgn) * ss;
    const int dstr = T.ntb * (ri + 1);               // elements between the d rows
    const int row = (ring_ph0(i) + i) * 6 * T.ntb + tl * (ri + 1);
    // byte offset of the lane's d = 0 entry, or past the end (the d rows add at most
    // 2 ntb (ri + 1) elements, which stays past the end without wrapping)
    const unsigned vo = (valid && k <= ri) ? unsigned(row + (part ? 3 * dstr : 0) + k) * unsigned(sizeof(S)) : OOB;
#pragma unroll
    for (int d = 0; d < 3; d++) {
      const double* wt = tb.siw + long(d) * T.iDim * 4 + 4 * i;
      double ud = 0.0;
      ud += ld_table(wt + 0) * w0;
      ud += ld_table(wt + 1) * w1;
      ud += ld_table(wt + 2) * w2;
      ud += ld_table(wt + 3) * w3;
      const double other = pair_swap(ud);
      const double val = cown * ud + coth * other;
      const int vd = int(vo + unsigned(d * dstr) * unsigned(sizeof(S)));
      if constexpr (sizeof(S) == 8)
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, double(val)), rX, vd, 0, 0);
      else
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, float(val)), rX, vd, 0, 0);
    }
  };
  const int mb = MB > 0 ? MB : T.mubar;
  // wave-uniform: some lane of the wave reaches a ring of cell mc
  auto cell_on = [&](int mc) -> bool {
    int rmax = 0;
    for (int mu = 0; mu < mb; mu++) rmax = max(rmax, ring_ri(mc * mb + mu));
    return __ballot(valid && k <= rmax) != 0;
  };
  double w0 = a_at(0), w1 = a_at(1), w2 = a_at(2);
  if constexpr (MB > 0) {
    // the MB rings of a cell unrolled, their phase factors fetched one cell ahead so that
    // the latency of these loads (L2) overlaps the cell before
    struct CellPh { double c[MB > 0 ? MB : 1], s[MB > 0 ? MB : 1]; };
    auto cell_ph = [&](int mc) -> CellPh {
      CellPh r;
      const int m = min(mc, T.nc - 1);
#pragma unroll
      for (int mu = 0; mu < MB; mu++) ring_phase(m * MB + mu, r.c[mu], r.s[mu]);
      return r;
    };
    CellPh cur = cell_ph(C0);
    static_for<kPR>([&](auto oc) {
      constexpr int o = decltype(oc)::value;
      const double w3 = a_at(o + 3);
      const int mc = C0 + o;
      const CellPh nxt = cell_ph(mc + 1);
      if (mc < T.nc && cell_on(mc)) {
#pragma unroll
        for (int mu = 0; mu < MB; mu++) ring_out(mc * MB + mu, cur.c[mu], cur.s[mu], w0, w1, w2, w3);
      }
      cur = nxt;
      w0 = w1; w1 = w2; w2 = w3;
    });
  } else {
    // any number of rings per cell: the rings as a loop
    static_for<kPR>([&](auto oc) {
      constexpr int o = decltype(oc)::value;
      const double w3 = a_at(o + 3);
      const int mc = C0 + o;
      if (mc < T.nc && cell_on(mc)) {
#pragma unroll 1
        for (int mu = 0; mu < mb; mu++) {
          double cc, ss;
          ring_phase(mc * mb + mu, cc, ss);
          ring_out(mc * mb + mu, cc, ss, w0, w1, w2, w3);
        }
      }
      w0 = w1; w1 = w2; w2 = w3;
    });
  }
}

template <typename S, int MB, bool SPH, int MODE = 0>
__global__ void __launch_bounds__(64 * kMaxPart) k_sasi(DevTables T, const double* __restrict__ spec,
                                                       S* __restrict__ X, double* __restrict__ aout) {
  sasi_body<S, MB, SPH, MODE>(T, spec, X, int(blockIdx.x), aout);
}

// MB: the Gauss points per cell compiled in (3, the default rule), or 0 for any mubar;
// SPH: ring sizes from the tables (spherical grids) or in closed form (cylindrical);
// a batched launch holds jobs of one form (sasi_form; the batch groups them by it)
template <typename S, int MB, bool SPH>
__global__ void __launch_bounds__(64 * kMaxPart) k_sasi_batch(const SaSiJob<S>* __restrict__ jobs,
                                                             const int* __restrict__ start, int njobs) {
  int lin;
  const SaSiJob<S>& j = jobs[batch_job(start, njobs, lin)];
  sasi_body<S, MB, SPH, 0>(j.T, j.spec, j.X, lin, nullptr);
}

template <typename S>
void launch_sasi(const DevTables& T, const double* spec, S* X, hipStream_t st) {
  const dim3 g(sasi_blocks(T)), b(sasi_threads(T));
  const size_t lds = sasi_lds(T);
  const int form = sasi_form(T);
  double* no = nullptr;
  if (form == 0) hipLaunchKernelGGL((k_sasi<S, 3, false>), g, b, lds, st, T, spec, X, no);
  else if (form == 1) hipLaunchKernelGGL((k_sasi<S, 3, true>), g, b, lds, st, T, spec, X, no);
  else if (form == 2) hipLaunchKernelGGL((k_sasi<S, 0, false>), g, b, lds, st, T, spec, X, no);
  else hipLaunchKernelGGL((k_sasi<S, 0, true>), g, b, lds, st, T, spec, X, no);
}
template void launch_sasi<double>(const DevTables&, const double*, double*, hipStream_t);
template void launch_sasi<float>(const DevTables&, const double*, float*, hipStream_t);

void launch_sasi_to_a(const DevTables& T, const double* spec, double* aout, hipStream_t st) {
  hipLaunchKernelGGL((k_sasi<double, 0, false, 1>), dim3(sasi_blocks(T)), dim3(sasi_threads(T)), sasi_lds(T), st, T,
                     spec, static_cast<double*>(nullptr), aout);
}

void launch_sasi_from_a(const DevTables& T, const double* a, double* X, hipStream_t st) {
  const dim3 g(sasi_blocks(T)), b(sasi_threads(T));
  const size_t lds = sasi_lds(T);
  const int form = sasi_form(T);
  double* no = nullptr;
  if (form == 0) hipLaunchKernelGGL((k_sasi<double, 3, false, 2>), g, b, lds, st, T, a, X, no);
  else if (form == 1) hipLaunchKernelGGL((k_sasi<double, 3, true, 2>), g, b, lds, st, T, a, X, no);
  else if (form == 2) hipLaunchKernelGGL((k_sasi<double, 0, false, 2>), g, b, lds, st, T, a, X, no);
  else hipLaunchKernelGGL((k_sasi<double, 0, true, 2>), g, b, lds, st, T, a, X, no);
}

template <typename S>
void launch_sasi_batch(const SaSiJob<S>* jobs, const int* start, int njobs, int nblocks, int nthr, size_t lds,
                       int form, hipStream_t st) {
  const dim3 g(nblocks), b(nthr);
  if (form == 0) hipLaunchKernelGGL((k_sasi_batch<S, 3, false>), g, b, lds, st, jobs, start, njobs);
  else if (form == 1) hipLaunchKernelGGL((k_sasi_batch<S, 3, true>), g, b, lds, st, jobs, start, njobs);
  else if (form == 2) hipLaunchKernelGGL((k_sasi_batch<S, 0, false>), g, b, lds, st, jobs, start, njobs);
  else hipLaunchKernelGGL((k_sasi_batch<S, 0, true>), g, b, lds, st, jobs, start, njobs);
}
template void launch_sasi_batch<double>(const SaSiJob<double>*, const int*, int, int, int, size_t, int, hipStream_t);
template void launch_sasi_batch<float>(const SaSiJob<float>*, const int*, int, int, int, size_t, int, hipStream_t);

}  // namespace ssg
