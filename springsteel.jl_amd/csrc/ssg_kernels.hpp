// ssg_kernels.hpp -- device-side table view and the launchers of every stage
// kernel of the transform path.
#pragma once

#include <hip/hip_runtime.h>

namespace ssg {

// Logical block id in XCD-major order. The dispatcher deals linear workgroup
// ids round-robin over the 8 XCDs of an MI355X; this bijection gives XCD c the
// contiguous logical range [start_c, start_c + count_c), so neighbouring
// logical blocks (which read the same data) share one XCD's L2.
__device__ inline unsigned xcd_major_block() {
  const unsigned nb = gridDim.x * gridDim.y * gridDim.z;
  const unsigned lin = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
  const unsigned xcd = lin & 7u, local = lin >> 3;
  const unsigned per = nb >> 3, rem = nb & 7u;      // XCDs below rem get per + 1
  const unsigned start = xcd * per + (xcd < rem ? xcd : rem);
  return start + local;
}

// Device pointers to the plan tables (uploaded once at plan creation).
struct DevTables {
  int geometry, nvars, nc, mubar, iDim, b_iDim, K, kDim, bk, nzb, nl, jDim, b_jDim, maxMin;
  long npts, nspec, nlines;
  // ring synthesis input X[vb][ring][k][P|Q][d][tl]: ntb transforms per block
  // (RL: the variables, RLZ: the Chebyshev modes of one variable), nvb blocks,
  // Ktot = sum over rings of (ri + 1) rows; ring r starts at row phaseOff[r] + r
  int ntr, ntb, nvb, Ktot;
  // B-spline
  const double* sbw;    // iDim * 4
  const double* siw;    // 3 * iDim * 4
  const double* L;      // nvars * maxMin * 4 (column-major band per var)
  const double* Linv;   // nvars * maxMin
  const int* gint;      // nvars * 8: Mdim, Min, rankL, rankR, lk, rk
  const double* gdbl;   // nvars * 4: aL, bL, aR, bR
  // rings
  const int* ringN; const int* ringRi; const int* ringOff; const int* coefOff; const int* phaseOff; const int* twOff;
  const int* kmax;      // nvars * iDim
  const double* phaseC; const double* phaseS; const double* twid;
  // Chebyshev
  const double* fwd;    // nvars * bk * kDim
  const double* inv;    // nvars * 3 * kDim * bk
};

// Forward Chebyshev stage: physical columns (kDim contiguous) -> cb[v][zb][col].
void launch_cheb_fwd(const DevTables& T, const double* phys, int ncol, double* cb, hipStream_t st);
// B-spline Gauss projection of every line into spectral.
void launch_sb(const DevTables& T, const double* src, double* spec, hipStream_t st);
// Gamma fold + banded LL^T solve + evaluation of d0/d1/d2 at the mish points (RZ).
void launch_sa_si(const DevTables& T, const double* spec, double* y, double* dst, hipStream_t st);
// The same for RL / RLZ, two lines per lane, with the inverse phase filter, into X.
void launch_sa_si_ring(const DevTables& T, const double* spec, double* y, double* X, hipStream_t st);
// Inverse Chebyshev stage (value, d/dz, d2/dz2 matrices) into the physical slots.
void launch_cheb_inv(const DevTables& T, const double* src, int ncol, double* phys, hipStream_t st);
// Ring synthesis on MFMA from X (RL: into physical slots; RLZ: fused with the
// inverse Chebyshev stage into physical slots).
void launch_ring_syn(const DevTables& T, const double* X, double* phys, hipStream_t st);
size_t ring_syn_lds(const DevTables& T);
// Forward Chebyshev stage of every RLZ ring point: physical slot 1 -> cm[v][point][zb].
void launch_cheb_fwd_pts(const DevTables& T, const double* phys, double* cm, hipStream_t st);
// Forward ring DFT on MFMA (RL: from physical slot 1; RLZ: from cm) into ring
// coefficients rc[v*nzb+zb][b_jDim].
void launch_ring_ana(const DevTables& T, const double* src, double* rc, hipStream_t st);
bool ring_ana_supported(const DevTables& T);

}  // namespace ssg
