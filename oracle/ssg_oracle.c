/*
 * ssg_oracle.c -- CPU restatement of the Springsteel.jl transform path for
 * RL, RZ and RLZ grids (spectralTransform! / gridTransform!).
 *
 * TEST INFRASTRUCTURE ONLY.  This is the oracle the HIP library is checked
 * against and the CPU port that bench.py times beside it.  The product never
 * links or calls it.
 *
 * The restatement follows the reference loops literally: the same stage
 * order, the same working-buffer semantics (zero-filled lines where k > ri,
 * never-written derivative slots that stay zero), the same scalar formulas for
 * the B-spline basis, quadrature, P+Q assembly, banded Cholesky, Gamma folds,
 * phase filter and Chebyshev recurrences.  The two places where the reference
 * calls into FFTW (R2HC/HC2R, REDFT00) are restated from their published
 * definitions: a mixed-radix complex FFT (smallest prime factor first, direct
 * DFT of each prime radix) for R2HC/HC2R and a direct O(N^2) sum for DCT-I.
 * Reference paths are relative to the Springsteel.jl repository root.
 */
#include "ssg_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static __thread char g_err[512];
const char* ssgo_error(void) { return g_err; }
static void set_err(const char* m) { snprintf(g_err, sizeof g_err, "%s", m); }

#define ONESIXTH (1.0 / 6.0)
#define FOURSIXTH (4.0 / 6.0)

/* ------------------------------------------------------------------------ */
/* Quadrature -- src/CubicBSpline.jl:128-200 (_GL_NODES, _GL_WEIGHTS,       */
/* _quadrature_rule).                                                         */
/* ------------------------------------------------------------------------ */
void ssgo_quadrature(int mb, int quadrature, double* qpts, double* qwts) {
  if (quadrature == 0) {
    double nodes[5], wts[5];
    if (mb == 1) { nodes[0] = 0.0; wts[0] = 2.0; }
    else if (mb == 2) { nodes[0] = -1.0 / sqrt(3.0); nodes[1] = 1.0 / sqrt(3.0); wts[0] = 1.0; wts[1] = 1.0; }
    else if (mb == 3) {
      nodes[0] = -sqrt(3.0 / 5.0); nodes[1] = 0.0; nodes[2] = sqrt(3.0 / 5.0);
      wts[0] = 5.0 / 9.0; wts[1] = 8.0 / 9.0; wts[2] = 5.0 / 9.0;
    } else if (mb == 4) {
      double a = sqrt(3.0 / 7.0 + 2.0 / 7.0 * sqrt(6.0 / 5.0));
      double b = sqrt(3.0 / 7.0 - 2.0 / 7.0 * sqrt(6.0 / 5.0));
      nodes[0] = -a; nodes[1] = -b; nodes[2] = b; nodes[3] = a;
      wts[0] = (18.0 - sqrt(30.0)) / 36.0; wts[1] = (18.0 + sqrt(30.0)) / 36.0;
      wts[2] = (18.0 + sqrt(30.0)) / 36.0; wts[3] = (18.0 - sqrt(30.0)) / 36.0;
    } else {
      double a = sqrt(5.0 + 2.0 * sqrt(10.0 / 7.0)) / 3.0;
      double b = sqrt(5.0 - 2.0 * sqrt(10.0 / 7.0)) / 3.0;
      nodes[0] = -a; nodes[1] = -b; nodes[2] = 0.0; nodes[3] = b; nodes[4] = a;
      wts[0] = (322.0 - 13.0 * sqrt(70.0)) / 900.0; wts[1] = (322.0 + 13.0 * sqrt(70.0)) / 900.0;
      wts[2] = 128.0 / 225.0;
      wts[3] = (322.0 + 13.0 * sqrt(70.0)) / 900.0; wts[4] = (322.0 - 13.0 * sqrt(70.0)) / 900.0;
    }
    for (int i = 0; i < mb; i++) { qpts[i] = (nodes[i] + 1.0) / 2.0; qwts[i] = wts[i] / 2.0; }
  } else {
    for (int k = 1; k <= mb; k++) { qpts[k - 1] = (double)(2 * k - 1) / (double)(2 * mb); qwts[k - 1] = 1.0 / mb; }
  }
}

/* ------------------------------------------------------------------------ */
/* Cubic B-spline basis -- src/CubicBSpline.jl:585-628 (basis).             */
/* ------------------------------------------------------------------------ */
typedef struct { double xmin, xmax, DX, DXrecip; int num_cells; } splp;

static double basis(const splp* sp, int m, double x, int derivative) {
  double b = 0.0;
  double xm = sp->xmin + (m * sp->DX);
  double delta = (x - xm) * sp->DXrecip;
  double z = fabs(delta);
  if (z < 2.0) {
    if (derivative == 0) {
      z = 2.0 - z;
      b = (z * z * z) * ONESIXTH;
      z -= 1.0;
      if (z > 0) b -= (z * z * z) * FOURSIXTH;
    } else if (derivative == 1) {
      z = 2.0 - z;
      b = (z * z) * ONESIXTH;
      z -= 1.0;
      if (z > 0) b -= (z * z) * FOURSIXTH;
      b *= ((delta > 0) ? -1.0 : 1.0) * 3.0 * sp->DXrecip;
    } else if (derivative == 2) {
      z = 2.0 - z;
      b = z;
      z -= 1.0;
      if (z > 0) b -= z * 4;
      b *= sp->DXrecip * sp->DXrecip;
    } else if (derivative == 3) {
      if (z > 1.0) b = 1.0;
      else if (z < 1.0) b = -3.0;
      b *= ((delta > 0) ? -1.0 : 1.0) * sp->DXrecip * sp->DXrecip * sp->DXrecip;
    }
  }
  return b;
}

double ssgo_basis(double xmin, double xmax, int num_cells, int m, double x, int derivative) {
  splp sp;
  sp.xmin = xmin; sp.xmax = xmax; sp.num_cells = num_cells;
  sp.DX = (xmax - xmin) / num_cells; sp.DXrecip = 1.0 / sp.DX;
  if (x < xmin || x > xmax) { set_err("DomainError: x outside spline domain"); return NAN; }
  return basis(&sp, m, x, derivative);
}

/* ------------------------------------------------------------------------ */
/* Banded LL^T -- src/banded_cholesky.jl:99-112 (copy of the lower band),   */
/* 166-197 (_band_factor3!), 202-231 (forward / backward solves).           */
/* ------------------------------------------------------------------------ */
#define LB(i, j) L[((j) - 1) * 4 + ((i) - 1)]   /* 1-based band storage L[i, j] */

static int band_factor3(double* L, int n) {
  for (int j = 1; j <= n; j++) {
    double d = LB(1, j);
    int ks = j - 3 > 1 ? j - 3 : 1;
    for (int k = ks; k <= j - 1; k++) { double ljk = LB(1 + j - k, k); d -= ljk * ljk; }
    if (d <= 0.0) return -j;
    double ljj = sqrt(d);
    LB(1, j) = ljj;
    double inv_ljj = 1.0 / ljj;
    int ihi = n < j + 3 ? n : j + 3;
    for (int i = j + 1; i <= ihi; i++) {
      double s = LB(1 + i - j, j);
      int kst = i - 3 > 1 ? i - 3 : 1;
      for (int k = kst; k <= j - 1; k++) s -= LB(1 + i - k, k) * LB(1 + j - k, k);
      LB(1 + i - j, j) = s * inv_ljj;
    }
  }
  return 0;
}

/* A is dense n x n column-major; Symmetric(A) (uplo = :U) so A[i, j] for i > j
 * reads the stored upper entry A[j, i] (banded_cholesky.jl:99-112). */
int ssgo_band_factor3(int n, const double* A, double* L) {
  memset(L, 0, sizeof(double) * 4 * n);
  for (int j = 1; j <= n; j++) {
    int ihi = n < j + 3 ? n : j + 3;
    for (int i = j; i <= ihi; i++) LB(1 + i - j, j) = A[(i - 1) * n + (j - 1)];
  }
  int rc = band_factor3(L, n);
  if (rc != 0) set_err("PosDefException");
  return rc;
}

static void band_solve3(int n, const double* L, const double* b, double* y) {
  for (int i = 1; i <= n; i++) {
    double s = b[i - 1];
    int ks = i - 3 > 1 ? i - 3 : 1;
    for (int k = ks; k <= i - 1; k++) s -= LB(1 + i - k, k) * y[k - 1];
    y[i - 1] = s / LB(1, i);
  }
  for (int i = n; i >= 1; i--) {
    double s = y[i - 1];
    int khi = n < i + 3 ? n : i + 3;
    for (int k = i + 1; k <= khi; k++) s -= LB(1 + k - i, i) * y[k - 1];
    y[i - 1] = s / LB(1, i);
  }
}
void ssgo_band_solve3(int n, const double* L, const double* b, double* x) { band_solve3(n, L, b, x); }
#undef LB

/* ------------------------------------------------------------------------ */
/* Spline1D -- src/CubicBSpline.jl:246-259 (SplineParameters), 321-500      */
/* (GammaBC + its mul!s), 733-783 (calcPQfactor), 887-897 (calcMishPoints),  */
/* 1000-1030 (_build_sb_matrix), 1107-1111 (SBtransform!), 1368-1396        */
/* (SAtransform!, non-R3X branch), 1554-1586 (SItransform).                 */
/* ------------------------------------------------------------------------ */
struct ssgo_spline {
  splp sp;
  int mubar, quadrature, Mdim, Minterior, mishDim;
  int rankL, rankR, lk, rk;
  double aL, bL, aR, bR, l_q;
  double qpts[8], qwts[8];
  double* mish;      /* mishDim */
  double* P; double* Q; /* Mdim x Mdim */
  double* L;         /* 4 x Minterior */
  int r3x;           /* an R3X side: SAtransform! takes the ahat branch */
  /* SB matrix in CSC order: for each column i, (row, val) pairs sorted by row */
  int* colptr; int* rowval; double* nzval;
};

static void gamma_mul(const ssgo_spline* s, const double* b, double* out) {
  int Min = s->Minterior, rL = s->rankL, Mdim = s->Mdim;
  for (int i = 1; i <= Min; i++) out[i - 1] = b[i + rL - 1];
  if (s->lk == SSGO_BC_A1) { out[0] += s->aL * b[0]; out[1] += s->bL * b[0]; }
  else if (s->lk == SSGO_BC_A2) { out[0] += s->aL * b[0] + s->bL * b[1]; }
  else if (s->lk == SSGO_BC_PER) { out[Min - 1] += b[0]; }
  if (s->rk == SSGO_BC_A1) { out[Min - 1] += s->aR * b[Mdim - 1]; out[Min - 2] += s->bR * b[Mdim - 1]; }
  else if (s->rk == SSGO_BC_A2) { out[Min - 1] += s->aR * b[Mdim - 1] + s->bR * b[Mdim - 2]; }
  else if (s->rk == SSGO_BC_PER) { out[0] += b[Mdim - 2]; out[1] += b[Mdim - 1]; }
}

static void gamma_tmul(const ssgo_spline* s, const double* x, double* out) {
  int Min = s->Minterior, rL = s->rankL, rR = s->rankR, Mdim = s->Mdim;
  for (int j = 1; j <= rL; j++) out[j - 1] = 0.0;
  for (int j = Mdim - rR + 1; j <= Mdim; j++) out[j - 1] = 0.0;
  for (int i = 1; i <= Min; i++) out[i + rL - 1] = x[i - 1];
  if (s->lk == SSGO_BC_A1) out[0] = s->aL * x[0] + s->bL * x[1];
  else if (s->lk == SSGO_BC_A2) { out[0] = s->aL * x[0]; out[1] = s->bL * x[0]; }
  else if (s->lk == SSGO_BC_PER) out[0] = x[Min - 1];
  if (s->rk == SSGO_BC_A1) out[Mdim - 1] = s->aR * x[Min - 1] + s->bR * x[Min - 2];
  else if (s->rk == SSGO_BC_A2) { out[Mdim - 2] = s->bR * x[Min - 1]; out[Mdim - 1] = s->aR * x[Min - 1]; }
  else if (s->rk == SSGO_BC_PER) { out[Mdim - 2] = x[0]; out[Mdim - 1] = x[1]; }
}

/* dense Gamma (Minterior x Mdim), calcGammaBC CubicBSpline.jl:648-706 */
static double gamma_entry(const ssgo_spline* s, int i, int j) {  /* 1-based */
  double val = (j == i + s->rankL) ? 1.0 : 0.0;
  if (s->lk == SSGO_BC_A1) { if (j == 1) { if (i == 1) val += s->aL; if (i == 2) val += s->bL; } }
  else if (s->lk == SSGO_BC_A2) { if (j == 1 && i == 1) val += s->aL; else if (j == 2 && i == 1) val += s->bL; }
  else if (s->lk == SSGO_BC_PER) { if (j == 1 && i == s->Minterior) val += 1.0; }
  if (s->rk == SSGO_BC_A1) { if (j == s->Mdim) { if (i == s->Minterior) val += s->aR; if (i == s->Minterior - 1) val += s->bR; } }
  else if (s->rk == SSGO_BC_A2) { if (j == s->Mdim && i == s->Minterior) val += s->aR; else if (j == s->Mdim - 1 && i == s->Minterior) val += s->bR; }
  else if (s->rk == SSGO_BC_PER) { if (j == s->Mdim - 1 && i == 1) val += 1.0; else if (j == s->Mdim && i == 2) val += 1.0; }
  return val;
}

ssgo_spline* ssgo_spline_create(double xmin, double xmax, int num_cells, int mubar, int quadrature,
                                double l_q, int bcl_kind, double bcl_a, double bcl_b,
                                int bcr_kind, double bcr_a, double bcr_b) {
  if (mubar < 1 || (quadrature == 0 && mubar > 5)) { set_err("ArgumentError: mubar"); return NULL; }
  if (bcl_kind == SSGO_BC_PER || bcr_kind == SSGO_BC_PER) { set_err("ENOTSUP: periodic spline BC"); return NULL; }
  ssgo_spline* s = (ssgo_spline*)calloc(1, sizeof(ssgo_spline));
  s->sp.xmin = xmin; s->sp.xmax = xmax; s->sp.num_cells = num_cells;
  s->sp.DX = (xmax - xmin) / num_cells; s->sp.DXrecip = 1.0 / s->sp.DX;
  s->mubar = mubar; s->quadrature = quadrature; s->l_q = l_q;
  s->Mdim = num_cells + 3; s->mishDim = num_cells * mubar;
  /* R3X folds like R3 (a pure interior selector, src/CubicBSpline.jl:392, 411) */
  s->r3x = bcl_kind == SSGO_BC_R3X || bcr_kind == SSGO_BC_R3X;
  if (bcl_kind == SSGO_BC_R3X) bcl_kind = SSGO_BC_R3;
  if (bcr_kind == SSGO_BC_R3X) bcr_kind = SSGO_BC_R3;
  s->lk = bcl_kind; s->rk = bcr_kind;
  s->rankL = bcl_kind == SSGO_BC_R0 ? 0 : bcl_kind == SSGO_BC_A1 ? 1 : bcl_kind == SSGO_BC_A2 ? 2 : 3;
  s->rankR = bcr_kind == SSGO_BC_R0 ? 0 : bcr_kind == SSGO_BC_A1 ? 1 : bcr_kind == SSGO_BC_A2 ? 2 : 3;
  s->aL = (bcl_kind == SSGO_BC_A1 || bcl_kind == SSGO_BC_A2) ? bcl_a : 0.0;
  s->bL = (bcl_kind == SSGO_BC_A1 || bcl_kind == SSGO_BC_A2) ? bcl_b : 0.0;
  s->aR = (bcr_kind == SSGO_BC_A1 || bcr_kind == SSGO_BC_A2) ? bcr_a : 0.0;
  s->bR = (bcr_kind == SSGO_BC_A1 || bcr_kind == SSGO_BC_A2) ? bcr_b : 0.0;
  s->Minterior = s->Mdim - s->rankL - s->rankR;
  ssgo_quadrature(mubar, quadrature, s->qpts, s->qwts);
  const splp* sp = &s->sp;
  int Mdim = s->Mdim;

  /* calcMishPoints CubicBSpline.jl:887-897 */
  s->mish = (double*)malloc(sizeof(double) * s->mishDim);
  for (int mc = 0; mc < num_cells; mc++)
    for (int mu = 1; mu <= mubar; mu++)
      s->mish[mu + mubar * mc - 1] = sp->xmin + mc * sp->DX + s->qpts[mu - 1] * sp->DX;

  /* calcPQfactor CubicBSpline.jl:733-783 */
  double eps_q = pow((l_q * sp->DX) / (2 * M_PI), 6);
  s->P = (double*)calloc((size_t)Mdim * Mdim, sizeof(double));
  s->Q = (double*)calloc((size_t)Mdim * Mdim, sizeof(double));
  for (int mi1 = 1; mi1 <= Mdim; mi1++)
    for (int mi2 = 1; mi2 <= Mdim; mi2++) {
      if (abs(mi1 - mi2) > 3) continue;
      int m1 = mi1 - 2, m2 = mi2 - 2;
      for (int mc = 0; mc <= num_cells - 1; mc++) {
        if (mc < m1 - 2 || mc > m1 + 1) continue;
        if (mc < m2 - 2 || mc > m2 + 1) continue;
        for (int mu = 1; mu <= mubar; mu++) {
          double x = sp->xmin + mc * sp->DX + s->qpts[mu - 1] * sp->DX;
          double pm1 = basis(sp, m1, x, 0), qm1 = basis(sp, m1, x, 3);
          double pm2 = basis(sp, m2, x, 0), qm2 = basis(sp, m2, x, 3);
          s->P[(mi2 - 1) * Mdim + (mi1 - 1)] += sp->DX * s->qwts[mu - 1] * pm1 * pm2;
          s->Q[(mi2 - 1) * Mdim + (mi1 - 1)] += sp->DX * s->qwts[mu - 1] * eps_q * qm1 * qm2;
        }
      }
    }
  /* PQopen = (G P G') + (G Q G'), dense products as in the reference */
  int Min = s->Minterior;
  double* G = (double*)malloc(sizeof(double) * Min * Mdim);
  for (int i = 1; i <= Min; i++) for (int j = 1; j <= Mdim; j++) G[(j - 1) * Min + (i - 1)] = gamma_entry(s, i, j);
  double* T = (double*)malloc(sizeof(double) * Min * Mdim);
  double* PQ = (double*)calloc((size_t)Min * Min, sizeof(double));
  for (int pass = 0; pass < 2; pass++) {
    const double* M = pass == 0 ? s->P : s->Q;
    /* T = G * M  (Min x Mdim) */
    for (int i = 0; i < Min; i++) for (int j = 0; j < Mdim; j++) {
      double acc = 0.0;
      for (int k = 0; k < Mdim; k++) acc += G[k * Min + i] * M[j * Mdim + k];
      T[j * Min + i] = acc;
    }
    /* PQ += T * G' */
    for (int i = 0; i < Min; i++) for (int j = 0; j < Min; j++) {
      double acc = 0.0;
      for (int k = 0; k < Mdim; k++) acc += T[k * Min + i] * G[k * Min + j];
      PQ[j * Min + i] += acc;
    }
  }
  s->L = (double*)malloc(sizeof(double) * 4 * Min);
  int rc = ssgo_band_factor3(Min, PQ, s->L);
  free(G); free(T); free(PQ);
  if (rc != 0) { ssgo_spline_destroy(s); return NULL; }

  /* _build_sb_matrix CubicBSpline.jl:1000-1030 -> CSC */
  int nnzmax = 4 * mubar * Mdim;
  int* rows = (int*)malloc(sizeof(int) * nnzmax); int* cols = (int*)malloc(sizeof(int) * nnzmax);
  double* vals = (double*)malloc(sizeof(double) * nnzmax);
  int nnz = 0;
  for (int mi = 1; mi <= Mdim; mi++) {
    int m = mi - 2;
    for (int mc = 0; mc <= num_cells - 1; mc++) {
      if (mc < m - 2 || mc > m + 1) continue;
      for (int mu = 1; mu <= mubar; mu++) {
        int i = mu + mubar * mc;
        double x = sp->xmin + mc * sp->DX + s->qpts[mu - 1] * sp->DX;
        double bm = basis(sp, m, x, 0);
        double v = sp->DX * s->qwts[mu - 1] * bm;
        if (v != 0.0) { rows[nnz] = mi; cols[nnz] = i; vals[nnz] = v; nnz++; }
      }
    }
  }
  s->colptr = (int*)calloc(s->mishDim + 1, sizeof(int));
  s->rowval = (int*)malloc(sizeof(int) * (nnz + 1)); s->nzval = (double*)malloc(sizeof(double) * (nnz + 1));
  for (int e = 0; e < nnz; e++) s->colptr[cols[e]]++;
  for (int c = 1; c <= s->mishDim; c++) s->colptr[c] += s->colptr[c - 1];
  int* fill = (int*)calloc(s->mishDim, sizeof(int));
  /* entries were generated in increasing row order, so each column stays row-sorted */
  for (int e = 0; e < nnz; e++) {
    int c = cols[e] - 1;
    int pos = s->colptr[c] + fill[c]++;
    s->rowval[pos] = rows[e]; s->nzval[pos] = vals[e];
  }
  free(fill); free(rows); free(cols); free(vals);
  return s;
}

void ssgo_spline_destroy(ssgo_spline* s) {
  if (!s) return;
  free(s->mish); free(s->P); free(s->Q); free(s->L); free(s->colptr); free(s->rowval); free(s->nzval);
  free(s);
}
int ssgo_spline_mdim(const ssgo_spline* s) { return s->Mdim; }
int ssgo_spline_minterior(const ssgo_spline* s) { return s->Minterior; }
void ssgo_spline_mishpoints(const ssgo_spline* s, double* x) { memcpy(x, s->mish, sizeof(double) * s->mishDim); }
void ssgo_spline_pq(const ssgo_spline* s, double* P, double* Q) {
  memcpy(P, s->P, sizeof(double) * s->Mdim * s->Mdim); memcpy(Q, s->Q, sizeof(double) * s->Mdim * s->Mdim);
}
void ssgo_spline_band(const ssgo_spline* s, double* L) { memcpy(L, s->L, sizeof(double) * 4 * s->Minterior); }

/* SBtransform!: mul!(b, M, u) with M in CSC (column-order accumulation) */
void ssgo_spline_sb(const ssgo_spline* s, const double* u, double* b) {
  for (int i = 0; i < s->Mdim; i++) b[i] = 0.0;
  for (int c = 0; c < s->mishDim; c++) {
    double uc = u[c];
    for (int e = s->colptr[c]; e < s->colptr[c + 1]; e++) b[s->rowval[e] - 1] += s->nzval[e] * uc;
  }
}

/* SAtransform! (non-R3X): a = G' (L L')^{-1} G b */
static void spline_sa_ws(const ssgo_spline* s, const double* b, double* a, double* wmin, double* wout) {
  gamma_mul(s, b, wmin);
  band_solve3(s->Minterior, s->L, wmin, wout);
  gamma_tmul(s, wout, a);
}
/* SAtransform!, R3X branch (src/CubicBSpline.jl:1375-1386):
 * btilde = b - pq ahat with pq = P + Q (the band |i - j| <= 3 holds every nonzero),
 * a = G' (L L')^{-1} G btilde + ahat */
static void spline_sa_ahat(const ssgo_spline* s, const double* b, const double* ahat, double* a, double* wmin,
                           double* wout, double* bt) {
  int M = s->Mdim;
  for (int i = 0; i < M; i++) {
    double acc = 0.0;
    for (int j = (i - 3 > 0 ? i - 3 : 0); j <= (i + 3 < M - 1 ? i + 3 : M - 1); j++)
      acc += (s->P[(long)j * M + i] + s->Q[(long)j * M + i]) * ahat[j];
    bt[i] = b[i] - acc;
  }
  spline_sa_ws(s, bt, a, wmin, wout);
  for (int i = 0; i < M; i++) a[i] += ahat[i];
}
void ssgo_spline_sa_ahat(const ssgo_spline* s, const double* b, const double* ahat, double* a) {
  double* w1 = (double*)malloc(sizeof(double) * s->Mdim);
  double* w2 = (double*)malloc(sizeof(double) * s->Mdim);
  double* bt = (double*)malloc(sizeof(double) * s->Mdim);
  if (ahat && s->r3x) spline_sa_ahat(s, b, ahat, a, w1, w2, bt);
  else spline_sa_ws(s, b, a, w1, w2);
  free(w1); free(w2); free(bt);
}
void ssgo_spline_sa(const ssgo_spline* s, const double* b, double* a) {
  double* w1 = (double*)malloc(sizeof(double) * s->Minterior);
  double* w2 = (double*)malloc(sizeof(double) * s->Minterior);
  spline_sa_ws(s, b, a, w1, w2);
  free(w1); free(w2);
}

/* SItransform(sp, a, points, u, derivative) CubicBSpline.jl:1554-1567 */
void ssgo_spline_si(const ssgo_spline* s, const double* a, const double* pts, int npts, int derivative, double* u) {
  const splp* sp = &s->sp;
  for (int i = 0; i < npts; i++) {
    u[i] = 0.0;
    int xm = (int)ceil((pts[i] - sp->xmin - (2.0 * sp->DX)) * sp->DXrecip);
    for (int m = xm; m <= xm + 3; m++) {
      if (m >= -1 && m <= sp->num_cells + 1) {
        int mi = m + 2;
        u[i] += basis(sp, m, pts[i], derivative) * a[mi - 1];
      }
    }
  }
}

/* ------------------------------------------------------------------------ */
/* FFT restatement (FFTW R2HC / HC2R definitions).                          */
/* ------------------------------------------------------------------------ */
typedef struct { int n; double* c; double* s; } twid;  /* c[m] = cos(2 pi m / n), s[m] = sin(2 pi m / n) */

static void twid_init(twid* t, int n) {
  t->n = n;
  t->c = (double*)malloc(sizeof(double) * n); t->s = (double*)malloc(sizeof(double) * n);
  for (int m = 0; m < n; m++) { double ang = 2.0 * M_PI * (double)m / (double)n; t->c[m] = cos(ang); t->s[m] = sin(ang); }
}
static void twid_free(twid* t) { free(t->c); free(t->s); }

static int smallest_factor(int n) {
  if (n % 4 == 0) return 4;
  if (n % 2 == 0) return 2;
  for (int p = 3; p * p <= n; p += 2) if (n % p == 0) return p;
  return n;
}

/* out[k] = sum_j x[j*stride] exp(-2 pi i j k / n), n | tw->n */
static void fft_rec(const twid* tw, int n, const double* xr, const double* xi, int stride,
                    double* outr, double* outi, double* tmp) {
  if (n == 1) { outr[0] = xr[0]; outi[0] = xi ? xi[0] : 0.0; return; }
  int p = smallest_factor(n), m = n / p;
  for (int q = 0; q < p; q++)
    fft_rec(tw, m, xr + (size_t)q * stride, xi ? xi + (size_t)q * stride : NULL, stride * p,
            outr + q * m, outi + q * m, tmp);
  int scale = tw->n / n;
  double* tr = tmp; double* ti = tmp + p;
  for (int k = 0; k < m; k++) {
    for (int q = 0; q < p; q++) { tr[q] = outr[q * m + k]; ti[q] = outi[q * m + k]; }
    for (int t = 0; t < p; t++) {
      int kk = k + t * m;
      double ar = 0.0, ai = 0.0;
      for (int q = 0; q < p; q++) {
        int e = (int)(((long)q * kk) % n) * scale;
        double c = tw->c[e], s = -tw->s[e];
        ar += tr[q] * c - ti[q] * s;
        ai += tr[q] * s + ti[q] * c;
      }
      outr[kk] = ar; outi[kk] = ai;
    }
  }
}

static void r2hc_ws(const twid* tw, int n, const double* x, double* out, double* wr, double* wi, double* tmp) {
  fft_rec(tw, n, x, NULL, 1, wr, wi, tmp);
  out[0] = wr[0];
  for (int k = 1; k <= n / 2; k++) out[k] = wr[k];
  for (int k = 1; k < (n + 1) / 2; k++) out[n - k] = wi[k];
}

static void hc2r_ws(const twid* tw, int n, const double* h, double* out, double* xr, double* xi, double* wr,
                    double* wi, double* tmp) {
  /* y_j = sum_k X_k e^{+2 pi i jk/n} = Re FFT(conj X)_j for Hermitian X */
  xr[0] = h[0]; xi[0] = 0.0;
  for (int k = 1; k < (n + 1) / 2; k++) {
    xr[k] = h[k]; xi[k] = -h[n - k];
    xr[n - k] = h[k]; xi[n - k] = h[n - k];
  }
  if (n % 2 == 0) { xr[n / 2] = h[n / 2]; xi[n / 2] = 0.0; }
  fft_rec(tw, n, xr, xi, 1, wr, wi, tmp);
  for (int j = 0; j < n; j++) out[j] = wr[j];
}

void ssgo_r2hc(int n, const double* x, double* out) {
  twid tw; twid_init(&tw, n);
  double* w = (double*)malloc(sizeof(double) * (2 * n + 2 * n + 2));
  r2hc_ws(&tw, n, x, out, w, w + n, w + 2 * n);
  free(w); twid_free(&tw);
}
void ssgo_hc2r(int n, const double* x, double* out) {
  twid tw; twid_init(&tw, n);
  double* w = (double*)malloc(sizeof(double) * (6 * n + 2));
  hc2r_ws(&tw, n, x, out, w, w + n, w + 2 * n, w + 3 * n, w + 4 * n);
  free(w); twid_free(&tw);
}

/* REDFT00 (DCT-I), unnormalised: Y_k = X_0 + (-1)^k X_{N-1} + 2 sum_{j=1}^{N-2} X_j cos(pi j k/(N-1)) */
typedef struct { int n; double* c; } dct1;
static void dct1_init(dct1* d, int n) {
  d->n = n; int M = 2 * (n - 1);
  d->c = (double*)malloc(sizeof(double) * (M > 0 ? M : 1));
  for (int m = 0; m < M; m++) d->c[m] = cos(M_PI * (double)m / (double)(n - 1));
}
static void dct1_apply(const dct1* d, const double* x, double* y) {
  int n = d->n, M = 2 * (n - 1);
  for (int k = 0; k < n; k++) {
    double acc = 0.0;
    for (int j = 1; j <= n - 2; j++) acc += x[j] * d->c[(int)(((long)j * k) % M)];
    y[k] = x[0] + ((k & 1) ? -x[n - 1] : x[n - 1]) + 2.0 * acc;
  }
}
void ssgo_redft00(int n, const double* x, double* out) {
  dct1 d; dct1_init(&d, n); dct1_apply(&d, x, out); free(d.c);
}

/* ------------------------------------------------------------------------ */
/* Fourier ring -- src/Fourier.jl:60-65, 388-395 (FBtransform!), 429-432    */
/* (FAtransform!), 463-468 (FItransform!), 495-503 (FIxcoefficients),       */
/* 540-581 (FIx / FIxx); src/phase_filter.jl:63-190; ring parameters        */
/* src/factory.jl:608-624 (_fill_fourier_rings_cyl!).                        */
/* ------------------------------------------------------------------------ */
typedef struct {
  int yDim, kmax, bDim;
  int klim;            /* wavenumbers a ring passes to and takes from the spline stage */
  double ymin;
  double* cos_k; double* sin_k;  /* [kmax] */
  twid tw;
} ring_t;

static void ring_alloc(ring_t* rg, int kmax_ring, int lpoints) {
  rg->cos_k = (double*)malloc(sizeof(double) * (kmax_ring + 1));
  rg->sin_k = (double*)malloc(sizeof(double) * (kmax_ring + 1));
  for (int k = 1; k <= kmax_ring; k++) { rg->cos_k[k - 1] = cos(k * rg->ymin); rg->sin_k[k - 1] = sin(k * rg->ymin); }
  twid_init(&rg->tw, lpoints);
}
/* _fill_fourier_rings_sph! (src/factory.jl:626-643), _sph_ring_dims (159-163):
 * kmax_ring = max(1, round(sin(theta) max_ri)) with Julia's round (ties to even,
 * as nearbyint in the default rounding mode), lpoints = 4 kmax_ring, the cap
 * applied to kmax and to bDim = 1 + 2 kmax, ymin = dl r / 2 (r 1-based).  The
 * spline stages use the ring's kmax (src/transforms_spherical.jl:153-157). */
static void ring_init_sph(ring_t* rg, double theta, int max_ri, int r, int var_kmax) {
  int kmax_ring = (int)nearbyint(sin(theta) * max_ri);
  if (kmax_ring < 1) kmax_ring = 1;
  int lpoints = 4 * kmax_ring;
  if (var_kmax >= 0 && var_kmax < kmax_ring) kmax_ring = var_kmax;
  double dl = 2 * M_PI / lpoints;
  rg->yDim = lpoints; rg->kmax = kmax_ring; rg->bDim = 1 + 2 * kmax_ring; rg->klim = kmax_ring;
  rg->ymin = 0.5 * dl * r;
  ring_alloc(rg, kmax_ring, lpoints);
}
static void ring_init(ring_t* rg, int ri, int var_kmax) {
  int lpoints = 4 + 4 * ri;
  int kmax_ring = ri;
  if (var_kmax >= 0) kmax_ring = var_kmax < ri ? var_kmax : ri;
  double dl = 2 * M_PI / lpoints;
  double offset = 0.5 * dl * (ri - 1);
  rg->yDim = lpoints; rg->kmax = kmax_ring; rg->bDim = ri * 2 + 1; rg->ymin = offset;
  rg->klim = ri;                                     /* "if k <= r + patchOffsetL" */
  ring_alloc(rg, kmax_ring, lpoints);
}
static void ring_free(ring_t* rg) { free(rg->cos_k); free(rg->sin_k); twid_free(&rg->tw); }

typedef struct { double* a; double* b; double* c; double* d; double* e; double* f; double* tmp; } fws;  /* FFT work */
static void fws_alloc(fws* w, int nmax) {
  double* p = (double*)malloc(sizeof(double) * (6 * (size_t)nmax + 2 * (size_t)nmax + 4));
  w->a = p; w->b = p + nmax; w->c = p + 2 * nmax; w->d = p + 3 * nmax; w->e = p + 4 * nmax; w->f = p + 5 * nmax;
  w->tmp = p + 6 * nmax;
}
static void fws_free(fws* w) { free(w->a); }

/* FBtransform!: scratch = R2HC(u) * (1/yDim); b = phasefilter_forward(scratch) */
static void ring_fb(const ring_t* rg, const double* u, double* b, fws* w) {
  double* sc = w->f;
  r2hc_ws(&rg->tw, rg->yDim, u, sc, w->a, w->b, w->tmp);
  double scale = 1.0 / rg->yDim;
  for (int i = 0; i < rg->yDim; i++) sc[i] *= scale;
  int yDim = rg->yDim, kmax = rg->kmax, bDim = rg->bDim;
  b[0] = sc[0];
  for (int k = 1; k <= kmax; k++) {
    double re = sc[k], im = sc[yDim - k];
    double c = rg->cos_k[k - 1], s = rg->sin_k[k - 1];
    b[k] = c * re + s * im;
    b[bDim - k] = -s * re + c * im;
  }
  if (bDim > 1 + 2 * kmax)
    for (int i = kmax + 2; i <= bDim - kmax; i++) b[i - 1] = 0.0;
}

/* FAtransform!: a = phasefilter_inverse(b) */
static void ring_fa(const ring_t* rg, const double* b, double* a) {
  int yDim = rg->yDim, kmax = rg->kmax, bDim = rg->bDim;
  a[0] = b[0];
  for (int k = 1; k <= kmax; k++) {
    double br = b[k], bi = b[bDim - k];
    double c = rg->cos_k[k - 1], s = rg->sin_k[k - 1];
    a[k] = c * br - s * bi;
    a[yDim - k] = s * br + c * bi;
  }
  for (int i = kmax + 2; i <= yDim - kmax; i++) a[i - 1] = 0.0;
}

/* FIxcoefficients: only 1..kmax entries are written; ax[1] and the middle keep
 * their previous contents, which are zero for the ring's never-written buffer */
static void ring_fix(const ring_t* rg, const double* a, double* ax) {
  for (int k = 1; k <= rg->kmax; k++) {
    ax[k] = -k * a[rg->yDim - k];
    ax[rg->yDim - k] = k * a[k];
  }
}

static void ring_fi(const ring_t* rg, const double* a, double* u, fws* w) {
  hc2r_ws(&rg->tw, rg->yDim, a, u, w->a, w->b, w->c, w->d, w->tmp);
}

/* ------------------------------------------------------------------------ */
/* Chebyshev column -- src/Chebyshev.jl:140-147, 329-343 (calcMishPoints),  */
/* 364-378 (calcFilterMatrix), 417-426 (CBtransform!), 473-499              */
/* (CAtransform!), 530-535 (CItransform!), 636-646 (CIxcoefficients),       */
/* 683-722 (CIx / CIxx), 991-1177 (calcGammaBC).                             */
/* ------------------------------------------------------------------------ */
typedef struct {
  double zmin, zmax; int zDim, bDim;
  int gkind;            /* 0: Vector gamma, 1: Matrix gamma */
  double* gamma;        /* Vector: zDim; Matrix: zDim x zDim column-major */
  double* filter;       /* bDim x zDim column-major */
  dct1 dct;
} cheb_t;

static int cheb_init(cheb_t* c, double zmin, double zmax, int zDim, int bDim, int bcb, int bct) {
  c->zmin = zmin; c->zmax = zmax; c->zDim = zDim; c->bDim = bDim;
  int N = zDim;
  dct1_init(&c->dct, N);
  c->filter = (double*)calloc((size_t)bDim * zDim, sizeof(double));
  if (bDim < zDim) { for (int i = 0; i < bDim; i++) c->filter[i * bDim + i] = 1.0; }
  else { for (int i = 1; i <= zDim; i++) c->filter[(i - 1) * bDim + (i - 1)] = exp(-36.0 * pow((double)i / zDim, 36)); }
#define GM(i, j) c->gamma[((j) - 1) * N + ((i) - 1)]
  if (bcb == SSGO_CBC_R0 && bct == SSGO_CBC_R0) {
    c->gkind = 0; c->gamma = (double*)calloc(N, sizeof(double));
  } else if (bcb == SSGO_CBC_R1T0 && bct == SSGO_CBC_R0) {
    c->gkind = 0; c->gamma = (double*)malloc(sizeof(double) * N);
    for (int i = 0; i < N; i++) c->gamma[i] = 1.0;
    for (int i = 2; i <= N - 1; i++) c->gamma[i - 1] *= 2.0;
    for (int i = 0; i < N; i++) c->gamma[i] *= (-0.5 / (N - 1));
  } else {
    c->gkind = 1; c->gamma = (double*)calloc((size_t)N * N, sizeof(double));
    double* cc = (double*)malloc(sizeof(double) * N);
    for (int i = 0; i < N; i++) cc[i] = 1.0;
    cc[0] *= 2.0; cc[N - 1] *= 2.0;
    /* building blocks */
    double* R1T1B = (double*)calloc((size_t)N * N, sizeof(double));
    double* R1T1T = (double*)calloc((size_t)N * N, sizeof(double));
    double* R1T0T = (double*)malloc(sizeof(double) * N * N);
    double* R1T0Bv = (double*)malloc(sizeof(double) * N);
    {
      double sf = 0.0;
      for (int i = 1; i <= N; i++) { int n = i - 1; sf += -(double)(n * n) / (cc[i - 1] * (N - 1)); }
      for (int i = 1; i <= N; i++) { int n = i - 1; for (int j = 1; j <= N; j++) R1T1B[(j - 1) * N + (i - 1)] = (double)(n * n) / (sf * cc[j - 1] * (N - 1)); }
    }
    {
      double sf = 0.0;
      for (int i = 1; i <= N; i++) { int n = i - 1; sf += -(double)(n * n) * pow(-1.0, n) * pow(-1.0, n + 1) / (cc[i - 1] * (N - 1)); }
      for (int i = 1; i <= N; i++) { int n = i - 1; for (int j = 1; j <= N; j++)
        R1T1T[(j - 1) * N + (i - 1)] = pow(-1.0, j + 1) * pow(-1.0, n + 1) * n * n / (sf * cc[j - 1] * (N - 1)); }
    }
    for (int i = 1; i <= N; i++) for (int j = 1; j <= N; j++)
      R1T0T[(j - 1) * N + (i - 1)] = 1.0 * (-1.0 * pow(-1.0, i - 1) * pow(-1.0, j - 1) / (N - 1));
    for (int j = 1; j <= N; j++) { R1T0T[(j - 1) * N + 0] *= 0.5; R1T0T[(j - 1) * N + (N - 1)] *= 0.5; }
    for (int i = 0; i < N; i++) R1T0Bv[i] = 1.0;
    for (int i = 2; i <= N - 1; i++) R1T0Bv[i - 1] *= 2.0;
    for (int i = 0; i < N; i++) R1T0Bv[i] *= (-0.5 / (N - 1));
    int ok = 1;
    for (int i = 1; i <= N; i++) for (int j = 1; j <= N; j++) {
      double v;
      if (bcb == SSGO_CBC_R1T1 && bct == SSGO_CBC_R0) v = R1T1B[(j - 1) * N + (i - 1)];
      else if (bcb == SSGO_CBC_R0 && bct == SSGO_CBC_R1T0) v = R1T0T[(j - 1) * N + (i - 1)];
      else if (bcb == SSGO_CBC_R0 && bct == SSGO_CBC_R1T1) v = R1T1T[(j - 1) * N + (i - 1)];
      else if (bcb == SSGO_CBC_R1T0 && bct == SSGO_CBC_R1T0) v = R1T0T[(j - 1) * N + (i - 1)] + R1T0Bv[i - 1];
      else if (bcb == SSGO_CBC_R1T1 && bct == SSGO_CBC_R1T1) v = R1T1B[(j - 1) * N + (i - 1)] + R1T1T[(j - 1) * N + (i - 1)];
      else if (bcb == SSGO_CBC_R1T0 && bct == SSGO_CBC_R1T1) v = R1T0Bv[i - 1] + R1T1T[(j - 1) * N + (i - 1)];
      else if (bcb == SSGO_CBC_R1T1 && bct == SSGO_CBC_R1T0) v = R1T1B[(j - 1) * N + (i - 1)] + R1T0T[(j - 1) * N + (i - 1)];
      else { ok = 0; v = 0.0; }
      GM(i, j) = v;
    }
    free(cc); free(R1T1B); free(R1T1T); free(R1T0T); free(R1T0Bv);
    if (!ok) { set_err("DomainError: Chebyshev boundary condition combination not implemented"); return -1; }
  }
#undef GM
  return 0;
}
static void cheb_free(cheb_t* c) { free(c->gamma); free(c->filter); free(c->dct.c); }

/* CBtransform!: scratch = REDFT00(u) * (1/(2(N-1))); b = filter * scratch */
static void cheb_cb(const cheb_t* c, const double* u, double* b, double* sc) {
  dct1_apply(&c->dct, u, sc);
  double scale = 1.0 / (2 * (c->zDim - 1));
  for (int i = 0; i < c->zDim; i++) sc[i] *= scale;
  for (int i = 0; i < c->bDim; i++) {
    double acc = 0.0;
    for (int j = 0; j < c->zDim; j++) acc += c->filter[j * c->bDim + i] * sc[j];
    b[i] = acc;
  }
}

/* CAtransform! */
static void cheb_ca(const cheb_t* c, const double* b, double* a, double* bfill) {
  int N = c->zDim;
  for (int i = 0; i < c->bDim; i++) bfill[i] = b[i];
  for (int i = c->bDim; i < N; i++) bfill[i] = 0.0;
  if (c->gkind == 1) {
    for (int i = 0; i < N; i++) {
      double acc = 0.0;
      for (int j = 0; j < N; j++) acc += c->gamma[i * N + j] * bfill[j];   /* (gamma')[i, j] = gamma[j, i] */
      a[i] = acc + bfill[i];
    }
  } else {
    double dot = 0.0;
    for (int j = 0; j < N; j++) dot += c->gamma[j] * bfill[j];
    for (int i = 0; i < N; i++) a[i] = bfill[i] + dot;
  }
}

/* CIxcoefficients: ax[N] is never written (stays zero in a fresh buffer) */
static void cheb_cix(const cheb_t* c, const double* a, double* ax) {
  int N = c->zDim;
  int k = N;
  ax[k - 2] = (2.0 * (k - 1) * a[k - 1]);
  for (k = N - 1; k >= 2; k--) ax[k - 2] = (2.0 * (k - 1) * a[k - 1]) + ax[k];
  double d = (-0.5 * (c->zmax - c->zmin));
  for (int i = 0; i < N; i++) ax[i] /= d;
}

/* ------------------------------------------------------------------------ */
/* Grid                                                                      */
/* ------------------------------------------------------------------------ */
struct ssgo_grid {
  ssgo_params p;
  int nvars, iDim, b_iDim, jDim, kDim, b_kDim, K, nslots;
  long npts, nspec;
  ssgo_spline** spl;   /* [nvars] (the reference's ibasis rows of one var share parameters) */
  cheb_t* cheb;        /* [nvars] */
  ring_t* rings;       /* RL: [nvars * iDim]; RLZ: [iDim] (shared by vars and z modes) */
  long* ringoff;       /* [iDim+1] physical ring offsets (in ring points) */
  int maxN;
  int dzz_bound;       /* regular transform: d2/dz2 slot holds its rounding scale (cheb_dzz_at) */
  int sph;             /* SL / SLZ: spherical rings; p.geometry holds RL / RLZ for the shared code */
  double* ahat;        /* R3X background coefficients, spectral layout, or NULL */
  int a_input;         /* tileTransform!: the inverse reads A coefficients (no SA) */
};

ssgo_grid* ssgo_grid_create(const ssgo_params* p) {
  if (p->geometry < SSGO_RL || p->geometry > SSGO_SLZ) { set_err("DomainError: geometry"); return NULL; }
  ssgo_grid* g = (ssgo_grid*)calloc(1, sizeof(ssgo_grid));
  g->p = *p;
  /* SL / SLZ run the RL / RLZ transforms (src/transforms_spherical.jl: "identical
   * in structure ... with sin(theta)-dependent ring sizes") on spherical rings */
  g->sph = p->geometry == SSGO_SL || p->geometry == SSGO_SLZ;
  if (p->geometry == SSGO_SL) g->p.geometry = SSGO_RL;
  if (p->geometry == SSGO_SLZ) g->p.geometry = SSGO_RLZ;
  p = &g->p;
  g->nvars = p->nvars;
  g->iDim = p->num_cells * p->mubar;
  g->b_iDim = p->num_cells + 3;
  g->K = g->iDim + p->patchOffsetL;
  g->kDim = p->kDim; g->b_kDim = p->b_kDim;
  g->spl = (ssgo_spline**)calloc(p->nvars, sizeof(ssgo_spline*));
  for (int v = 0; v < p->nvars; v++) {
    g->spl[v] = ssgo_spline_create(p->iMin, p->iMax, p->num_cells, p->mubar, p->quadrature, p->l_q[v],
                                   p->bcl_kind[v], p->bcl_ab[2 * v], p->bcl_ab[2 * v + 1],
                                   p->bcr_kind[v], p->bcr_ab[2 * v], p->bcr_ab[2 * v + 1]);
    if (!g->spl[v]) { ssgo_grid_destroy(g); return NULL; }
  }
  if (p->geometry == SSGO_RZ || p->geometry == SSGO_RLZ) {
    if (p->kDim < 2) { set_err("ArgumentError: zDim must be >= 2"); ssgo_grid_destroy(g); return NULL; }
    if (p->geometry == SSGO_RLZ && p->b_kDim < 3) { set_err("BoundsError: RLZ needs b_kDim >= 3"); ssgo_grid_destroy(g); return NULL; }
    g->cheb = (cheb_t*)calloc(p->nvars, sizeof(cheb_t));
    for (int v = 0; v < p->nvars; v++)
      if (cheb_init(&g->cheb[v], p->kMin, p->kMax, p->kDim, p->b_kDim, p->bcb_kind[v], p->bct_kind[v]) != 0) {
        ssgo_grid_destroy(g); return NULL;
      }
  }
  g->ringoff = (long*)calloc(g->iDim + 1, sizeof(long));
  if (p->geometry == SSGO_RL || p->geometry == SSGO_RLZ) {
    int nr = p->geometry == SSGO_RL ? p->nvars * g->iDim : g->iDim;
    g->rings = (ring_t*)calloc(nr, sizeof(ring_t));
    int max_ri = g->iDim + p->patchOffsetL;
    for (int r = 1; r <= g->iDim; r++) {
      int ri = r + p->patchOffsetL;
      double theta = g->spl[0]->mish[r - 1];
      if (p->geometry == SSGO_RL) {
        for (int v = 0; v < p->nvars; v++) {
          if (g->sph) ring_init_sph(&g->rings[v * g->iDim + r - 1], theta, max_ri, r, p->max_wavenumber[v]);
          else ring_init(&g->rings[v * g->iDim + r - 1], ri, p->max_wavenumber[v]);
        }
      } else {
        if (g->sph) ring_init_sph(&g->rings[r - 1], theta, max_ri, r, p->ring_kmax);
        else ring_init(&g->rings[r - 1], ri, p->ring_kmax);
      }
    }
    long off = 0;
    for (int r = 0; r < g->iDim; r++) {
      int n = g->rings[r].yDim;                      /* ring sizes do not depend on the cap */
      g->ringoff[r] = off; off += n;
      if (n > g->maxN) g->maxN = n;
    }
    g->ringoff[g->iDim] = off;
    g->jDim = (int)off;
  }
  if (p->geometry == SSGO_RL) { g->npts = g->jDim; g->nspec = (long)g->b_iDim * (1 + 2 * g->K); g->nslots = 5; }
  else if (p->geometry == SSGO_RZ) { g->npts = (long)g->iDim * g->kDim; g->nspec = (long)g->b_kDim * g->b_iDim; g->nslots = 5; }
  else { g->npts = (long)g->kDim * g->jDim; g->nspec = (long)g->b_kDim * g->b_iDim * (1 + 2 * g->K); g->nslots = 7; }
  return g;
}

void ssgo_grid_destroy(ssgo_grid* g) {
  if (!g) return;
  if (g->spl) for (int v = 0; v < g->nvars; v++) ssgo_spline_destroy(g->spl[v]);
  free(g->spl);
  if (g->cheb) for (int v = 0; v < g->nvars; v++) cheb_free(&g->cheb[v]);
  free(g->cheb);
  if (g->rings) {
    int nr = g->p.geometry == SSGO_RL ? g->nvars * g->iDim : g->iDim;
    for (int i = 0; i < nr; i++) ring_free(&g->rings[i]);
  }
  free(g->rings); free(g->ringoff); free(g->ahat);
  free(g);
}

void ssgo_dims(const ssgo_grid* g, long* d) {
  d[0] = g->npts; d[1] = g->nspec; d[2] = g->nslots; d[3] = g->iDim; d[4] = g->b_iDim;
  d[5] = g->jDim; d[6] = g->kDim; d[7] = g->b_kDim;
}

/* getGridpoints: RL transforms_cylindrical.jl:75-90, RLZ 454-476, RZ transforms_cartesian.jl:359-373 */
void ssgo_gridpoints(const ssgo_grid* g, double* pts) {
  const ssgo_spline* s = g->spl[0];
  long np = g->npts;
  if (g->p.geometry == SSGO_RZ) {
    const cheb_t* c = &g->cheb[0];
    long q = 0;
    for (int r = 0; r < g->iDim; r++)
      for (int z = 1; z <= g->kDim; z++) {
        double scale = -0.5 * (c->zmax - c->zmin), offset = 0.5 * (c->zmin + c->zmax);
        pts[q] = s->mish[r];
        pts[np + q] = cos((z - 1) * M_PI / (g->kDim - 1)) * scale + offset;
        q++;
      }
    return;
  }
  long q = 0;
  for (int r = 0; r < g->iDim; r++) {
    const ring_t* rg = &g->rings[r];   /* RL: rings of var 1; RLZ: shared rings */
    for (int l = 1; l <= rg->yDim; l++) {
      double lm = rg->ymin + (2 * M_PI * (l - 1) / rg->yDim);
      if (g->p.geometry == SSGO_RL) { pts[q] = s->mish[r]; pts[np + q] = lm; q++; }
      else {
        const cheb_t* c = &g->cheb[0];
        double scale = -0.5 * (c->zmax - c->zmin), offset = 0.5 * (c->zmin + c->zmax);
        for (int z = 1; z <= g->kDim; z++) {
          pts[q] = s->mish[r]; pts[np + q] = lm;
          pts[2 * np + q] = cos((z - 1) * M_PI / (g->kDim - 1)) * scale + offset;
          q++;
        }
      }
    }
  }
}

#define PHYS(buf, i, v, slot) (buf)[((long)(slot) * g->nvars + (v)) * g->npts + (i)]
#define SPEC(buf, row, v) (buf)[(long)(v) * g->nspec + (row)]

/* RL spectralTransform: transforms_cylindrical.jl:159-222 */
static void rl_forward(ssgo_grid* g, const double* physical, double* spectral, int v, double* ringb, fws* w) {
  int iDim = g->iDim, b_iDim = g->b_iDim, K = g->K;
  const ssgo_spline* s = g->spl[v];
  long bo = 0;
  for (int r = 1; r <= iDim; r++) {
    const ring_t* rg = &g->rings[v * iDim + r - 1];
    double* u = w->e;
    for (int l = 0; l < rg->yDim; l++) u[l] = PHYS(physical, g->ringoff[r - 1] + l, v, 0);
    ring_fb(rg, u, ringb + bo, w);
    bo += rg->bDim;
  }
  double* um = (double*)calloc(iDim, sizeof(double));
  double* bb = (double*)malloc(sizeof(double) * b_iDim);
  /* wavenumber 0 */
  bo = 0;
  for (int r = 1; r <= iDim; r++) { um[r - 1] = ringb[bo]; bo += g->rings[v * iDim + r - 1].bDim; }
  ssgo_spline_sb(s, um, bb);
  for (int k = 0; k < b_iDim; k++) SPEC(spectral, k, v) = bb[k];
  double* uR = (double*)malloc(sizeof(double) * iDim);
  double* uI = (double*)malloc(sizeof(double) * iDim);
  for (int k = 1; k <= K; k++) {
    bo = 0;
    for (int r = 1; r <= iDim; r++) {
      const ring_t* rg = &g->rings[v * iDim + r - 1];
      uR[r - 1] = 0.0; uI[r - 1] = 0.0;
      if (k <= rg->klim) { uR[r - 1] = ringb[bo + k]; uI[r - 1] = ringb[bo + rg->bDim - k]; }
      bo += rg->bDim;
    }
    int p = k * 2;
    ssgo_spline_sb(s, uR, bb);
    for (int k2 = 0; k2 < b_iDim; k2++) SPEC(spectral, (long)(p - 1) * b_iDim + k2, v) = bb[k2];
    ssgo_spline_sb(s, uI, bb);
    for (int k2 = 0; k2 < b_iDim; k2++) SPEC(spectral, (long)p * b_iDim + k2, v) = bb[k2];
  }
  free(um); free(bb); free(uR); free(uI);
}

/* SA + SI^(d) of one line at the spline's mish points; with a_in the line holds
 * the A coefficients already (tileTransform!, src/tiling.jl:1540-1658: the
 * spline's .a is set from the array and SI runs without SA) */
static void line_sa_si(const ssgo_spline* s, const double* b, double* out0, double* out1, double* out2,
                       double* a, double* w1, double* w2, const double* ahat, int a_in) {
  if (a_in) {
    memcpy(a, b, sizeof(double) * s->Mdim);
  } else if (ahat && s->r3x) {
    double* bt = (double*)malloc(sizeof(double) * s->Mdim);
    spline_sa_ahat(s, b, ahat, a, w1, w2, bt);
    free(bt);
  } else {
    spline_sa_ws(s, b, a, w1, w2);
  }
  if (out0) ssgo_spline_si(s, a, s->mish, s->mishDim, 0, out0);
  if (out1) ssgo_spline_si(s, a, s->mish, s->mishDim, 1, out1);
  if (out2) ssgo_spline_si(s, a, s->mish, s->mishDim, 2, out2);
}

/* RL gridTransform: transforms_cylindrical.jl:259-428 */
static void rl_inverse(ssgo_grid* g, const double* spectral, double* physical, int v, double* ring3, fws* w) {
  int iDim = g->iDim, b_iDim = g->b_iDim, K = g->K;
  const ssgo_spline* s = g->spl[v];
  long btot = 0;
  for (int r = 1; r <= iDim; r++) btot += g->rings[v * iDim + r - 1].bDim;
  double* rb[3] = {ring3, ring3 + btot, ring3 + 2 * btot};  /* ring b from d0, d1 (spline_r), d2 (spline_rr) */
  double* bb = (double*)malloc(sizeof(double) * b_iDim);
  double* a = (double*)malloc(sizeof(double) * b_iDim);
  double* w1 = (double*)malloc(sizeof(double) * b_iDim);
  double* w2 = (double*)malloc(sizeof(double) * b_iDim);
  double* o[3]; for (int d = 0; d < 3; d++) o[d] = (double*)malloc(sizeof(double) * iDim);
  for (int p = 0; p <= 2 * K; p++) {
    for (int m = 0; m < b_iDim; m++) bb[m] = SPEC(spectral, (long)p * b_iDim + m, v);
    line_sa_si(s, bb, o[0], o[1], o[2], a, w1, w2, g->ahat ? g->ahat + (long)v * g->nspec + (long)p * b_iDim : NULL,
               g->a_input);
    int k = (p + 1) / 2, isim = (p > 0 && (p % 2 == 0));
    long bo = 0;
    for (int r = 1; r <= iDim; r++) {
      const ring_t* rg = &g->rings[v * iDim + r - 1];
      if (p == 0) { for (int d = 0; d < 3; d++) rb[d][bo] = o[d][r - 1]; }
      else if (k <= rg->klim) {
        int idx = isim ? rg->bDim - k : k;
        for (int d = 0; d < 3; d++) rb[d][bo + idx] = o[d][r - 1];
      }
      bo += rg->bDim;
    }
  }
  long bo = 0;
  for (int r = 1; r <= iDim; r++) {
    const ring_t* rg = &g->rings[v * iDim + r - 1];
    int N = rg->yDim;
    double* aa = w->e;
    double* ax = (double*)calloc(N, sizeof(double));
    double* ax2 = (double*)calloc(N, sizeof(double));
    double* uo = (double*)malloc(sizeof(double) * N);
    long po = g->ringoff[r - 1];
    ring_fa(rg, rb[0] + bo, aa);
    ring_fi(rg, aa, uo, w);
    for (int l = 0; l < N; l++) PHYS(physical, po + l, v, 0) = uo[l];
    ring_fix(rg, aa, ax);
    ring_fi(rg, ax, uo, w);
    for (int l = 0; l < N; l++) PHYS(physical, po + l, v, 3) = uo[l];
    ring_fix(rg, ax, ax2);
    ring_fi(rg, ax2, uo, w);
    for (int l = 0; l < N; l++) PHYS(physical, po + l, v, 4) = uo[l];
    for (int d = 1; d <= 2; d++) {
      ring_fa(rg, rb[d] + bo, aa);
      ring_fi(rg, aa, uo, w);
      for (int l = 0; l < N; l++) PHYS(physical, po + l, v, d) = uo[l];
    }
    free(ax); free(ax2); free(uo);
    bo += rg->bDim;
  }
  free(bb); free(a); free(w1); free(w2); for (int d = 0; d < 3; d++) free(o[d]);
}

/* RZ spectralTransform: transforms_cartesian.jl:579-617 */
static void rz_forward(ssgo_grid* g, const double* physical, double* spectral, int v) {
  int iDim = g->iDim, kDim = g->kDim, b_kDim = g->b_kDim, b_iDim = g->b_iDim;
  const cheb_t* c = &g->cheb[v];
  double* tempcb = (double*)malloc(sizeof(double) * b_kDim * iDim);
  double* u = (double*)malloc(sizeof(double) * kDim);
  double* sc = (double*)malloc(sizeof(double) * kDim);
  double* b = (double*)malloc(sizeof(double) * b_kDim);
  for (int r = 0; r < iDim; r++) {
    for (int z = 0; z < kDim; z++) u[z] = PHYS(physical, (long)r * kDim + z, v, 0);
    cheb_cb(c, u, b, sc);
    for (int k = 0; k < b_kDim; k++) tempcb[k + r * b_kDim] = b[k];
  }
  double* um = (double*)malloc(sizeof(double) * iDim);
  double* bb = (double*)malloc(sizeof(double) * b_iDim);
  for (int z = 0; z < b_kDim; z++) {
    for (int r = 0; r < iDim; r++) um[r] = tempcb[z + r * b_kDim];
    ssgo_spline_sb(g->spl[v], um, bb);
    for (int k = 0; k < b_iDim; k++) SPEC(spectral, (long)z * b_iDim + k, v) = bb[k];
  }
  free(tempcb); free(u); free(sc); free(b); free(um); free(bb);
}

/* RZ gridTransform: transforms_cartesian.jl:658-724 */
static void rz_inverse(ssgo_grid* g, const double* spectral, double* physical, int v) {
  int iDim = g->iDim, kDim = g->kDim, b_kDim = g->b_kDim, b_iDim = g->b_iDim;
  const cheb_t* c = &g->cheb[v];
  const ssgo_spline* s = g->spl[v];
  double* sbuf = (double*)malloc(sizeof(double) * 3 * iDim * b_kDim);   /* [d][z][r] */
  double* bb = (double*)malloc(sizeof(double) * b_iDim);
  double* a = (double*)malloc(sizeof(double) * b_iDim);
  double* w1 = (double*)malloc(sizeof(double) * b_iDim);
  double* w2 = (double*)malloc(sizeof(double) * b_iDim);
  for (int z = 0; z < b_kDim; z++) {
    for (int m = 0; m < b_iDim; m++) bb[m] = SPEC(spectral, (long)z * b_iDim + m, v);
    line_sa_si(s, bb, sbuf + (0 * b_kDim + z) * iDim, sbuf + (1 * b_kDim + z) * iDim, sbuf + (2 * b_kDim + z) * iDim, a, w1, w2,
               g->ahat ? g->ahat + (long)v * g->nspec + (long)z * b_iDim : NULL, 0);
  }
  double* kb = (double*)malloc(sizeof(double) * b_kDim);
  double* ka = (double*)malloc(sizeof(double) * kDim);
  double* bfill = (double*)malloc(sizeof(double) * kDim);
  double* ax = (double*)calloc(kDim, sizeof(double));
  double* ax2 = (double*)calloc(kDim, sizeof(double));
  double* uo = (double*)malloc(sizeof(double) * kDim);
  for (int dr = 0; dr <= 2; dr++)
    for (int r = 0; r < iDim; r++) {
      for (int z = 0; z < b_kDim; z++) kb[z] = sbuf[(dr * b_kDim + z) * iDim + r];
      cheb_ca(c, kb, ka, bfill);
      dct1_apply(&c->dct, ka, uo);
      int slot = dr == 0 ? 0 : dr;   /* slot 1 / 2 / 3 (0-based 0 / 1 / 2) */
      for (int z = 0; z < kDim; z++) PHYS(physical, (long)r * kDim + z, v, slot) = uo[z];
      if (dr == 0) {
        cheb_cix(c, ka, ax);
        dct1_apply(&c->dct, ax, uo);
        for (int z = 0; z < kDim; z++) PHYS(physical, (long)r * kDim + z, v, 3) = uo[z];
        cheb_cix(c, ax, ax2);
        dct1_apply(&c->dct, ax2, uo);
        for (int z = 0; z < kDim; z++) PHYS(physical, (long)r * kDim + z, v, 4) = uo[z];
      }
    }
  free(sbuf); free(bb); free(a); free(w1); free(w2); free(kb); free(ka); free(bfill); free(ax); free(ax2); free(uo);
}

/* RLZ spectralTransform: transforms_cylindrical.jl:530-618 */
static void rlz_forward(ssgo_grid* g, const double* physical, double* spectral, int v, int threads) {
  int iDim = g->iDim, kDim = g->kDim, b_kDim = g->b_kDim, b_iDim = g->b_iDim, K = g->K;
  const cheb_t* c = &g->cheb[v];
  long btot = 0;
  for (int r = 0; r < iDim; r++) btot += g->rings[r].bDim;
  long* boff = (long*)malloc(sizeof(long) * (iDim + 1));
  boff[0] = 0; for (int r = 0; r < iDim; r++) boff[r + 1] = boff[r] + g->rings[r].bDim;
  double* ringb = (double*)malloc(sizeof(double) * btot * b_kDim);   /* [z_b][ring coefs] */
#pragma omp parallel num_threads(threads)
  {
    double* u = (double*)malloc(sizeof(double) * kDim);
    double* sc = (double*)malloc(sizeof(double) * kDim);
    double* cb = (double*)malloc(sizeof(double) * b_kDim);
    double* tempcb = (double*)malloc(sizeof(double) * (size_t)b_kDim * g->maxN);
    double* ru = (double*)malloc(sizeof(double) * g->maxN);
    fws w; fws_alloc(&w, g->maxN);
#pragma omp for schedule(dynamic, 1)
    for (int r = 0; r < iDim; r++) {
      const ring_t* rg = &g->rings[r];
      for (int l = 0; l < rg->yDim; l++) {
        long i0 = (g->ringoff[r] + l) * kDim;
        for (int z = 0; z < kDim; z++) u[z] = PHYS(physical, i0 + z, v, 0);
        cheb_cb(c, u, cb, sc);
        for (int k = 0; k < b_kDim; k++) tempcb[k + (long)l * b_kDim] = cb[k];
      }
      for (int zb = 0; zb < b_kDim; zb++) {
        for (int l = 0; l < rg->yDim; l++) ru[l] = tempcb[zb + (long)l * b_kDim];
        ring_fb(rg, ru, ringb + (long)zb * btot + boff[r], &w);
      }
    }
    free(u); free(sc); free(cb); free(tempcb); free(ru); fws_free(&w);
  }
#pragma omp parallel num_threads(threads)
  {
    double* um = (double*)malloc(sizeof(double) * iDim);
    double* bb = (double*)malloc(sizeof(double) * b_iDim);
#pragma omp for schedule(dynamic, 1)
    for (int zb = 0; zb < b_kDim; zb++) {
      const double* rb = ringb + (long)zb * btot;
      long r1 = (long)zb * b_iDim * (1 + K * 2);
      for (int r = 0; r < iDim; r++) um[r] = rb[boff[r]];
      ssgo_spline_sb(g->spl[v], um, bb);
      for (int k = 0; k < b_iDim; k++) SPEC(spectral, r1 + k, v) = bb[k];
      for (int k = 1; k <= K; k++) {
        for (int pass = 0; pass < 2; pass++) {
          for (int r = 0; r < iDim; r++) {
            um[r] = 0.0;
            if (k <= g->rings[r].klim) um[r] = rb[boff[r] + (pass == 0 ? k : g->rings[r].bDim - k)];
          }
          ssgo_spline_sb(g->spl[v], um, bb);
          int p = (k - 1) * 2;
          long p1 = r1 + b_iDim + (long)p * b_iDim + (pass ? b_iDim : 0);
          for (int k2 = 0; k2 < b_iDim; k2++) SPEC(spectral, p1 + k2, v) = bb[k2];
        }
      }
    }
    free(um); free(bb);
  }
  free(ringb); free(boff);
}

/* RLZ gridTransform: transforms_cylindrical.jl:659-853.  The SA solve is the
 * same for dr = 0, 1, 2 (the reference repeats it per dr), so it runs once. */
static void rlz_inverse(ssgo_grid* g, const double* spectral, double* physical, int v, int threads) {
  int iDim = g->iDim, kDim = g->kDim, b_kDim = g->b_kDim, b_iDim = g->b_iDim, K = g->K;
  const cheb_t* c = &g->cheb[v];
  const ssgo_spline* s = g->spl[v];
  long* boff = (long*)malloc(sizeof(long) * (iDim + 1));
  boff[0] = 0; for (int r = 0; r < iDim; r++) boff[r + 1] = boff[r] + g->rings[r].bDim;
  long btot = boff[iDim];
  /* ring b per (dr, z_b) */
  double* rb = (double*)malloc(sizeof(double) * 3 * b_kDim * btot);
#pragma omp parallel num_threads(threads)
  {
    double* bb = (double*)malloc(sizeof(double) * b_iDim);
    double* a = (double*)malloc(sizeof(double) * b_iDim);
    double* w1 = (double*)malloc(sizeof(double) * b_iDim);
    double* w2 = (double*)malloc(sizeof(double) * b_iDim);
    double* o[3]; for (int d = 0; d < 3; d++) o[d] = (double*)malloc(sizeof(double) * iDim);
#pragma omp for schedule(dynamic, 4) collapse(2)
    for (int zb = 0; zb < b_kDim; zb++)
      for (int p = 0; p <= 2 * K; p++) {
        long r1 = (long)zb * b_iDim * (1 + K * 2);
        for (int m = 0; m < b_iDim; m++) bb[m] = SPEC(spectral, r1 + (long)p * b_iDim + m, v);
        line_sa_si(s, bb, o[0], o[1], o[2], a, w1, w2,
                   g->ahat ? g->ahat + (long)v * g->nspec + r1 + (long)p * b_iDim : NULL, g->a_input);
        int k = (p + 1) / 2, isim = (p > 0 && (p % 2 == 0));
        for (int r = 0; r < iDim; r++) {
          if (p == 0) { for (int d = 0; d < 3; d++) rb[(long)(d * b_kDim + zb) * btot + boff[r]] = o[d][r]; }
          else if (k <= g->rings[r].klim) {
            int idx = isim ? g->rings[r].bDim - k : k;
            for (int d = 0; d < 3; d++) rb[(long)(d * b_kDim + zb) * btot + boff[r] + idx] = o[d][r];
          }
        }
      }
    free(bb); free(a); free(w1); free(w2); for (int d = 0; d < 3; d++) free(o[d]);
  }
#pragma omp parallel num_threads(threads)
  {
    fws w; fws_alloc(&w, g->maxN);
    double* rbuf = (double*)malloc(sizeof(double) * (size_t)g->maxN * b_kDim);  /* ringBuffer[l, z_b] */
    double* aa = (double*)malloc(sizeof(double) * g->maxN);
    double* ax = (double*)malloc(sizeof(double) * g->maxN);
    double* ax2 = (double*)malloc(sizeof(double) * g->maxN);
    double* uo = (double*)malloc(sizeof(double) * g->maxN);
    double* kb = (double*)malloc(sizeof(double) * b_kDim);
    double* ka = (double*)malloc(sizeof(double) * kDim);
    double* bfill = (double*)malloc(sizeof(double) * kDim);
    double* cx = (double*)malloc(sizeof(double) * kDim);
    double* cx2 = (double*)malloc(sizeof(double) * kDim);
    double* co = (double*)malloc(sizeof(double) * kDim);
#pragma omp for schedule(dynamic, 1)
    for (int r = 0; r < iDim; r++) {
      const ring_t* rg = &g->rings[r];
      int N = rg->yDim;
      long zi = g->ringoff[r] * kDim;
      for (int dr = 0; dr <= 2; dr++)
        for (int dl = 0; dl <= 2; dl++) {
          if (dr > 0 && dl > 0) continue;
          for (int zb = 0; zb < b_kDim; zb++) {
            ring_fa(rg, rb + (long)(dr * b_kDim + zb) * btot + boff[r], aa);
            if (dl == 0) ring_fi(rg, aa, uo, &w);
            else {
              memset(ax, 0, sizeof(double) * N);
              ring_fix(rg, aa, ax);
              if (dl == 1) ring_fi(rg, ax, uo, &w);
              else { memset(ax2, 0, sizeof(double) * N); ring_fix(rg, ax, ax2); ring_fi(rg, ax2, uo, &w); }
            }
            for (int l = 0; l < N; l++) rbuf[l + (long)zb * N] = uo[l];
          }
          int slot = (dr == 0) ? (dl == 0 ? 0 : 2 + dl) : dr;
          for (int l = 0; l < N; l++) {
            for (int zb = 0; zb < b_kDim; zb++) kb[zb] = rbuf[l + (long)zb * N];
            cheb_ca(c, kb, ka, bfill);
            dct1_apply(&c->dct, ka, co);
            long z1 = zi + (long)l * kDim;
            for (int z = 0; z < kDim; z++) PHYS(physical, z1 + z, v, slot) = co[z];
            if (dr == 0 && dl == 0) {
              memset(cx, 0, sizeof(double) * kDim); memset(cx2, 0, sizeof(double) * kDim);
              cheb_cix(c, ka, cx);
              dct1_apply(&c->dct, cx, co);
              for (int z = 0; z < kDim; z++) PHYS(physical, z1 + z, v, 5) = co[z];
              cheb_cix(c, cx, cx2);
              dct1_apply(&c->dct, cx2, co);
              for (int z = 0; z < kDim; z++) PHYS(physical, z1 + z, v, 6) = co[z];
            }
          }
        }
    }
    fws_free(&w); free(rbuf); free(aa); free(ax); free(ax2); free(uo); free(kb); free(ka); free(bfill);
    free(cx); free(cx2); free(co);
  }
  free(rb); free(boff);
}

int ssgo_spectral_transform(ssgo_grid* g, const double* physical, double* spectral, int threads) {
  if (threads < 1) threads = 1;
  if (g->p.geometry == SSGO_RL) {
    long btot = 0;
    for (int r = 0; r < g->iDim; r++) btot += g->rings[r].bDim;
#pragma omp parallel for num_threads(threads) schedule(dynamic, 1)
    for (int v = 0; v < g->nvars; v++) {
      double* ringb = (double*)malloc(sizeof(double) * btot);
      fws w; fws_alloc(&w, g->maxN);
      rl_forward(g, physical, spectral, v, ringb, &w);
      free(ringb); fws_free(&w);
    }
  } else if (g->p.geometry == SSGO_RZ) {
#pragma omp parallel for num_threads(threads) schedule(dynamic, 1)
    for (int v = 0; v < g->nvars; v++) rz_forward(g, physical, spectral, v);
  } else {
    for (int v = 0; v < g->nvars; v++) rlz_forward(g, physical, spectral, v, threads);
  }
  return 0;
}

int ssgo_grid_transform(ssgo_grid* g, const double* spectral, double* physical, int threads) {
  if (threads < 1) threads = 1;
  if (g->p.geometry == SSGO_RL) {
    long btot = 0;
    for (int r = 0; r < g->iDim; r++) btot += g->rings[r].bDim;
#pragma omp parallel for num_threads(threads) schedule(dynamic, 1)
    for (int v = 0; v < g->nvars; v++) {
      double* ring3 = (double*)calloc(3 * btot, sizeof(double));
      fws w; fws_alloc(&w, g->maxN);
      rl_inverse(g, spectral, physical, v, ring3, &w);
      free(ring3); fws_free(&w);
    }
  } else if (g->p.geometry == SSGO_RZ) {
#pragma omp parallel for num_threads(threads) schedule(dynamic, 1)
    for (int v = 0; v < g->nvars; v++) rz_inverse(g, spectral, physical, v);
  } else {
    for (int v = 0; v < g->nvars; v++) rlz_inverse(g, spectral, physical, v, threads);
  }
  return 0;
}

/* splineTransform!(sharedSpectral, tile) on the ring grids (src/tiling.jl:1490-1516
 * RL, 1675-1700 SL, 1862-1892 RLZ, 2063-2093 SLZ): SAtransform of every line of
 * every Chebyshev block (1 + 2 K lines each) of `shared` into `spectral`. */
int ssgo_spline_transform(ssgo_grid* g, const double* shared, double* spectral, int threads) {
  if (g->p.geometry == SSGO_RZ) { set_err("the tile transforms take the ring grids"); return -95; }
  if (threads < 1) threads = 1;
  const int b_iDim = g->b_iDim, nzb = g->p.geometry == SSGO_RL ? 1 : g->b_kDim;
  const long nlines = (long)nzb * (1 + 2 * g->K);
  for (int v = 0; v < g->nvars; v++) {
    const ssgo_spline* s = g->spl[v];
#pragma omp parallel num_threads(threads)
    {
      double* bb = (double*)malloc(sizeof(double) * b_iDim);
      double* a = (double*)malloc(sizeof(double) * b_iDim);
      double* w1 = (double*)malloc(sizeof(double) * b_iDim);
      double* w2 = (double*)malloc(sizeof(double) * b_iDim);
#pragma omp for schedule(dynamic, 16)
      for (long L = 0; L < nlines; L++) {
        for (int m = 0; m < b_iDim; m++) bb[m] = SPEC(shared, L * b_iDim + m, v);
        line_sa_si(s, bb, NULL, NULL, NULL, a, w1, w2, g->ahat ? g->ahat + (long)v * g->nspec + L * b_iDim : NULL, 0);
        for (int m = 0; m < b_iDim; m++) SPEC(spectral, L * b_iDim + m, v) = a[m];
      }
      free(bb); free(a); free(w1); free(w2);
    }
  }
  return 0;
}

/* tileTransform!(sharedSpectral, tile, physical, spectral) on the ring grids
 * (src/tiling.jl:1540-1658 RL, 1720-1838 SL, 1915-2046 RLZ, 2113-2242 SLZ):
 * gridTransform's body from the A coefficients in `spectral`. */
int ssgo_tile_transform(ssgo_grid* g, const double* spectral, double* physical, int threads) {
  if (g->p.geometry == SSGO_RZ) { set_err("the tile transforms take the ring grids"); return -95; }
  g->a_input = 1;
  const int rc = ssgo_grid_transform(g, spectral, physical, threads);
  g->a_input = 0;
  return rc;
}

/* ------------------------------------------------------------------------ */
/* regularGridTransform (the "next" row: evaluation on a regular grid).     */
/* ------------------------------------------------------------------------ */
/* Chebyshev series at arbitrary z from a = CAtransform!(b):
 * _cheb_eval_pts! / _cheb_dz_pts! / _cheb_dzz_pts!
 * (src/transforms_cylindrical.jl:1030-1094) */
static double cheb_clamp(double x, double lo, double hi) { return x < lo ? lo : (x > hi ? hi : x); }
static double cheb_eval_at(const cheb_t* c, const double* a, double z) {
  int N = c->zDim;
  double scale = -0.5 * (c->zmax - c->zmin), offset = 0.5 * (c->zmin + c->zmax);
  double xi = cheb_clamp((z - offset) / scale, -1.0 + 1e-14, 1.0 - 1e-14);
  double t = acos(xi);
  double val = a[0];
  for (int k = 2; k <= N - 1; k++) val += 2.0 * a[k - 1] * cos((k - 1) * t);
  return val + a[N - 1] * cos((N - 1) * t);
}
static double cheb_dz_at(const cheb_t* c, const double* a, double z) {
  int N = c->zDim;
  double scale = -0.5 * (c->zmax - c->zmin), offset = 0.5 * (c->zmin + c->zmax);
  double xi = cheb_clamp((z - offset) / scale, -1.0 + 1e-10, 1.0 - 1e-10);
  double t = acos(xi), st = sin(t);
  double dfdt = 0.0;
  for (int k = 2; k <= N - 1; k++) dfdt -= 2.0 * (k - 1) * a[k - 1] * sin((k - 1) * t);
  dfdt -= (N - 1) * a[N - 1] * sin((N - 1) * t);
  return dfdt / (-scale * st);
}
/* bound = 0: the reference's value.  bound = 1 (test use): the scale of the
 * rounding error of that FP64 evaluation, (st sum|d2 terms| + |ct| sum|df terms|)
 * / (scale^2 st^3); eps times it bounds the error of the final difference,
 * which at a clamped angle (t ~ 1.4e-5) is far above eps times the value. */
static double cheb_dzz_at(const cheb_t* c, const double* a, double z, int bound) {
  int N = c->zDim;
  double scale = -0.5 * (c->zmax - c->zmin), offset = 0.5 * (c->zmin + c->zmax);
  double xi = cheb_clamp((z - offset) / scale, -1.0 + 1e-10, 1.0 - 1e-10);
  double t = acos(xi), st = sin(t), ct = cos(t);
  double dfdt = 0.0, d2fdt2 = 0.0, m1 = 0.0, m2 = 0.0;
  for (int k = 2; k <= N - 1; k++) {
    int m = k - 1;
    double t1 = 2.0 * m * a[k - 1] * sin(m * t), t2 = 2.0 * (double)(m * m) * a[k - 1] * cos(m * t);
    dfdt -= t1;
    d2fdt2 -= t2;
    m1 += fabs(t1); m2 += fabs(t2);
  }
  double t1 = (N - 1) * a[N - 1] * sin((N - 1) * t), t2 = (double)((N - 1) * (N - 1)) * a[N - 1] * cos((N - 1) * t);
  dfdt -= t1;
  d2fdt2 -= t2;
  m1 += fabs(t1); m2 += fabs(t2);
  if (bound) return (m2 * st + m1 * fabs(ct)) / (scale * scale * (st * st * st));
  return (d2fdt2 * st - dfdt * ct) / (scale * scale * (st * st * st));
}
double ssgo_cheb_eval(const ssgo_grid* g, int v, const double* b, double z, int derivative) {
  const cheb_t* c = &g->cheb[v];
  double* a = (double*)malloc(sizeof(double) * c->zDim);
  double* bf = (double*)malloc(sizeof(double) * c->zDim);
  cheb_ca(c, b, a, bf);
  double r = derivative == 0 ? cheb_eval_at(c, a, z) : derivative == 1 ? cheb_dz_at(c, a, z) : cheb_dzz_at(c, a, z, g->dzz_bound);
  free(a); free(bf);
  return r;
}

/* SA then SI at arbitrary points (SAtransform! + SItransform/SIxtransform/
 * SIxxtransform(spline, points, u), src/CubicBSpline.jl:1368-1396, 1554-1567) */
static void line_sa_pts(const ssgo_spline* s, const double* b, const double* pts, int npts, int derivative,
                        double* u, double* a, double* w1, double* w2) {
  spline_sa_ws(s, b, a, w1, w2);
  ssgo_spline_si(s, a, pts, npts, derivative, u);
}

#define REG(i, v, slot) out[((long)(slot) * g->nvars + (v)) * npr + (i)]

/* RL: src/transforms_cylindrical.jl:929-1020.  ak[d][ks][i]: ks = 0 (k = 0),
 * 2k - 1 (cosine line of wavenumber k), 2k (sine line), d = 0, 1, 2 */
static void rl_regular(ssgo_grid* g, const double* spectral, const double* rp, int nr, const double* lp, int nl,
                       double* out, int v) {
  int b_iDim = g->b_iDim, K = g->K, nks = 1 + 2 * K;
  long npr = (long)nr * nl;
  const ssgo_spline* s = g->spl[v];
  double* ak = (double*)malloc(sizeof(double) * 3 * (size_t)nks * nr);
  double* bb = (double*)malloc(sizeof(double) * b_iDim);
  double* a = (double*)malloc(sizeof(double) * b_iDim);
  double* w1 = (double*)malloc(sizeof(double) * b_iDim);
  double* w2 = (double*)malloc(sizeof(double) * b_iDim);
  for (int ks = 0; ks < nks; ks++) {
    for (int m = 0; m < b_iDim; m++) bb[m] = SPEC(spectral, (long)ks * b_iDim + m, v);
    for (int d = 0; d < 3; d++) line_sa_pts(s, bb, rp, nr, d, ak + ((size_t)d * nks + ks) * nr, a, w1, w2);
  }
#define AK(d, ks, i) ak[((size_t)(d) * nks + (ks)) * nr + (i)]
  long idx = 0;
  for (int i = 0; i < nr; i++)
    for (int j = 0; j < nl; j++) {
      double lam = lp[j];
      double f = AK(0, 0, i), fr = AK(1, 0, i), frr = AK(2, 0, i), fl = 0.0, fll = 0.0;
      for (int k = 1; k <= K; k++) {
        double ck = cos(k * lam), sk = sin(k * lam);
        double rc = AK(0, 2 * k - 1, i), rs = AK(0, 2 * k, i);
        f += 2.0 * (rc * ck + rs * sk);
        fr += 2.0 * (AK(1, 2 * k - 1, i) * ck + AK(1, 2 * k, i) * sk);
        frr += 2.0 * (AK(2, 2 * k - 1, i) * ck + AK(2, 2 * k, i) * sk);
        fl += 2.0 * k * (-rc * sk + rs * ck);
        fll -= 2.0 * (double)(k * k) * (rc * ck + rs * sk);
      }
      REG(idx, v, 0) = f; REG(idx, v, 1) = fr; REG(idx, v, 2) = frr; REG(idx, v, 3) = fl; REG(idx, v, 4) = fll;
      idx++;
    }
#undef AK
  free(ak); free(bb); free(a); free(w1); free(w2);
}

/* RZ: src/transforms_cartesian.jl:946-1001 (_rz_slot 1003-1010) */
static void rz_regular(ssgo_grid* g, const double* spectral, const double* xp, int nx, const double* zp, int nz,
                       double* out, int v) {
  int b_iDim = g->b_iDim, b_kDim = g->b_kDim, N = g->kDim;
  long npr = (long)nx * nz;
  const ssgo_spline* s = g->spl[v];
  const cheb_t* c = &g->cheb[v];
  double* ibuf = (double*)malloc(sizeof(double) * (size_t)nx * b_kDim);
  double* bb = (double*)malloc(sizeof(double) * b_iDim);
  double* a = (double*)malloc(sizeof(double) * b_iDim);
  double* w1 = (double*)malloc(sizeof(double) * b_iDim);
  double* w2 = (double*)malloc(sizeof(double) * b_iDim);
  double* kb = (double*)malloc(sizeof(double) * b_kDim);
  double* ka = (double*)malloc(sizeof(double) * N);
  double* bf = (double*)malloc(sizeof(double) * N);
  double* u = (double*)malloc(sizeof(double) * nx);
  for (int dr = 0; dr <= 2; dr++) {
    for (int z = 0; z < b_kDim; z++) {
      for (int m = 0; m < b_iDim; m++) bb[m] = SPEC(spectral, (long)z * b_iDim + m, v);
      line_sa_pts(s, bb, xp, nx, dr, u, a, w1, w2);
      for (int xi = 0; xi < nx; xi++) ibuf[(size_t)xi * b_kDim + z] = u[xi];
    }
    for (int dk = 0; dk <= (dr == 0 ? 2 : 0); dk++) {
      int slot = dr == 0 ? (dk == 0 ? 0 : 2 + dk) : dr;     /* _rz_slot, 0-based */
      for (int xi = 0; xi < nx; xi++) {
        for (int z = 0; z < b_kDim; z++) kb[z] = ibuf[(size_t)xi * b_kDim + z];
        cheb_ca(c, kb, ka, bf);
        for (int k = 0; k < nz; k++) {
          double r = dk == 0 ? cheb_eval_at(c, ka, zp[k]) : dk == 1 ? cheb_dz_at(c, ka, zp[k]) : cheb_dzz_at(c, ka, zp[k], g->dzz_bound);
          REG((long)xi * nz + k, v, slot) = r;
        }
      }
    }
  }
  free(ibuf); free(bb); free(a); free(w1); free(w2); free(kb); free(ka); free(bf); free(u);
}

/* RLZ: src/transforms_cylindrical.jl:1168-1280 */
static void rlz_regular(ssgo_grid* g, const double* spectral, const double* rp, int nr, const double* lp, int nl,
                        const double* zp, int nz, double* out, int v, int threads) {
  int b_iDim = g->b_iDim, b_kDim = g->b_kDim, K = g->K, N = g->kDim, nks = 1 + 2 * K;
  long npr = (long)nr * nl * nz;
  const ssgo_spline* s = g->spl[v];
  const cheb_t* c = &g->cheb[v];
  double* sv = (double*)malloc(sizeof(double) * (size_t)nr * b_kDim * nks);   /* [zb][ks][i] */
  double* fb = (double*)malloc(sizeof(double) * (size_t)nr * nl * b_kDim);    /* [i][j][zb] */
  for (int dr = 0; dr <= 2; dr++) {
#pragma omp parallel num_threads(threads)
    {
      double* bb = (double*)malloc(sizeof(double) * b_iDim);
      double* a = (double*)malloc(sizeof(double) * b_iDim);
      double* w1 = (double*)malloc(sizeof(double) * b_iDim);
      double* w2 = (double*)malloc(sizeof(double) * b_iDim);
#pragma omp for schedule(dynamic, 4) collapse(2)
      for (int zb = 0; zb < b_kDim; zb++)
        for (int ks = 0; ks < nks; ks++) {
          long r1 = (long)zb * b_iDim * nks;
          for (int m = 0; m < b_iDim; m++) bb[m] = SPEC(spectral, r1 + (long)ks * b_iDim + m, v);
          line_sa_pts(s, bb, rp, nr, dr, sv + ((size_t)zb * nks + ks) * nr, a, w1, w2);
        }
      free(bb); free(a); free(w1); free(w2);
    }
#define SV(zb, ks, i) sv[((size_t)(zb) * nks + (ks)) * nr + (i)]
    for (int dl = 0; dl <= 2; dl++) {
      if (dr > 0 && dl > 0) continue;
#pragma omp parallel for num_threads(threads) schedule(static) collapse(2)
      for (int i = 0; i < nr; i++)
        for (int j = 0; j < nl; j++) {
          double lam = lp[j];
          for (int zb = 0; zb < b_kDim; zb++) {
            double val = dl == 0 ? SV(zb, 0, i) : 0.0;
            for (int k = 1; k <= K; k++) {
              double rc = SV(zb, 2 * k - 1, i), rs = SV(zb, 2 * k, i);
              double ck = cos(k * lam), sk = sin(k * lam);
              if (dl == 0) val += 2.0 * (rc * ck + rs * sk);
              else if (dl == 1) val += 2.0 * k * (-rc * sk + rs * ck);
              else val -= 2.0 * (double)(k * k) * (rc * ck + rs * sk);
            }
            fb[((size_t)i * nl + j) * b_kDim + zb] = val;
          }
        }
#pragma omp parallel num_threads(threads)
      {
        double* ka = (double*)malloc(sizeof(double) * N);
        double* bf = (double*)malloc(sizeof(double) * N);
#pragma omp for schedule(static) collapse(2)
        for (int i = 0; i < nr; i++)
          for (int j = 0; j < nl; j++) {
            cheb_ca(c, fb + ((size_t)i * nl + j) * b_kDim, ka, bf);
            long flat = ((long)i * nl + j) * nz;
            for (int k = 0; k < nz; k++) {
              if (dr == 0 && dl == 0) {
                REG(flat + k, v, 0) = cheb_eval_at(c, ka, zp[k]);
                REG(flat + k, v, 5) = cheb_dz_at(c, ka, zp[k]);
                REG(flat + k, v, 6) = cheb_dzz_at(c, ka, zp[k], g->dzz_bound);
              } else if (dr == 0) {
                REG(flat + k, v, 2 + dl) = cheb_eval_at(c, ka, zp[k]);
              } else {
                REG(flat + k, v, dr) = cheb_eval_at(c, ka, zp[k]);
              }
            }
          }
        free(ka); free(bf);
      }
    }
#undef SV
  }
  free(sv); free(fb);
}
#undef REG

/* ------------------------------------------------------------------------ */
/* evaluate_unstructured (src/interpolation.jl:1735-1822) for RL, RZ, RLZ.   */
/* ------------------------------------------------------------------------ */
/* _cheb_eval_single (src/interpolation.jl:1271-1285): the angle clamped to
 * [-1, 1], not to the 1e-14 margin of _cheb_eval_pts! */
static double cheb_eval_single(const cheb_t* c, const double* a, double z) {
  int N = c->zDim;
  double scale = -0.5 * (c->zmax - c->zmin), offset = 0.5 * (c->zmin + c->zmax);
  double t = acos(cheb_clamp((z - offset) / scale, -1.0, 1.0));
  double val = a[0];
  for (int k = 2; k <= N - 1; k++) val += 2.0 * a[k - 1] * cos((k - 1) * t);
  return val + a[N - 1] * cos((N - 1) * t);
}

/* _check_bounds_nd (src/interpolation.jl:1202-1232): the spline axis and the
 * Chebyshev axis are bounded, the Fourier axis is periodic */
static int unstructured_in_bounds(const ssgo_grid* g, const double* pts, long n, long i) {
  int geo = g->p.geometry;
  double x = pts[i];
  if (!(x >= g->p.iMin && x <= g->p.iMax)) return 0;
  if (geo == SSGO_RZ) { double z = pts[n + i]; return z >= g->p.kMin && z <= g->p.kMax; }
  if (geo == SSGO_RLZ) { double z = pts[2 * n + i]; return z >= g->p.kMin && z <= g->p.kMax; }
  return 1;
}

int ssgo_evaluate_unstructured(ssgo_grid* g, const double* spectral, const double* pts, long n, double* out,
                               double fill, int threads) {
  if (threads < 1) threads = 1;
  int geo = g->p.geometry, b_iDim = g->b_iDim, nzb = geo == SSGO_RL ? 1 : g->b_kDim, N = g->kDim;
  int nks = geo == SSGO_RZ ? 1 : 1 + 2 * g->K;        /* lines per (var, z mode) */
  long nline = (long)nzb * nks;
  for (int v = 0; v < g->nvars; v++) {
    const ssgo_spline* s = g->spl[v];
    /* SAtransform! of every line (_get_ahat_cache_rl / _rlz, 1470-1563): a[line][b_iDim],
     * line = zb * nks + ks, ks = 0 (k = 0), 2k - 1 (cosine), 2k (sine); RZ: line = zb */
    double* acache = (double*)malloc(sizeof(double) * (size_t)nline * b_iDim);
#pragma omp parallel num_threads(threads)
    {
      double* bb = (double*)malloc(sizeof(double) * b_iDim);
      double* w1 = (double*)malloc(sizeof(double) * b_iDim);
      double* w2 = (double*)malloc(sizeof(double) * b_iDim);
#pragma omp for schedule(static)
      for (long l = 0; l < nline; l++) {
        for (int m = 0; m < b_iDim; m++) bb[m] = SPEC(spectral, l * b_iDim + m, v);
        spline_sa_ws(s, bb, acache + (size_t)l * b_iDim, w1, w2);
      }
      free(bb); free(w1); free(w2);
    }
#pragma omp parallel num_threads(threads)
    {
      double* col = (double*)malloc(sizeof(double) * (size_t)nzb);
      double* ka = (double*)malloc(sizeof(double) * N);
      double* bf = (double*)malloc(sizeof(double) * N);
#pragma omp for schedule(static)
      for (long i = 0; i < n; i++) {
        if (!unstructured_in_bounds(g, pts, n, i)) { out[(long)v * n + i] = fill; continue; }
        double r = pts[i];
        double lam = geo == SSGO_RZ ? 0.0 : pts[n + i];
        for (int zb = 0; zb < nzb; zb++) {
          const double* az = acache + (size_t)zb * nks * b_iDim;
          double u;
          /* SItransform at r of every line of this z mode, then the Fourier sum
           * (_eval_unstructured_rl 1628-1668, _eval_unstructured_rlz 1670-1733) */
          ssgo_spline_si(s, az, &r, 1, 0, &u);
          double val = u;
          for (int k = 1; k <= (geo == SSGO_RZ ? 0 : g->K); k++) {
            double rc, rs;
            ssgo_spline_si(s, az + (size_t)(2 * k - 1) * b_iDim, &r, 1, 0, &rc);
            ssgo_spline_si(s, az + (size_t)(2 * k) * b_iDim, &r, 1, 0, &rs);
            val += 2.0 * (rc * cos(k * lam) + rs * sin(k * lam));
          }
          col[zb] = val;
        }
        if (geo == SSGO_RL) out[(long)v * n + i] = col[0];
        else {
          /* CAtransform! then _cheb_eval_single (_eval_single_kdim 1304-1323) */
          cheb_ca(&g->cheb[v], col, ka, bf);
          out[(long)v * n + i] = cheb_eval_single(&g->cheb[v], ka, geo == SSGO_RZ ? pts[n + i] : pts[2 * n + i]);
        }
      }
      free(col); free(ka); free(bf);
    }
    free(acache);
  }
  return 0;
}

void ssgo_set_ahat(ssgo_grid* g, const double* ahat) {
  free(g->ahat);
  g->ahat = NULL;
  if (!ahat) return;
  g->ahat = (double*)malloc(sizeof(double) * (size_t)g->nspec * g->nvars);
  memcpy(g->ahat, ahat, sizeof(double) * (size_t)g->nspec * g->nvars);
}

void ssgo_set_dzz_bound(ssgo_grid* g, int on) { g->dzz_bound = on ? 1 : 0; }

int ssgo_regular_transform(ssgo_grid* g, const double* spectral, const double* ip, int ni, const double* jp, int nj,
                           const double* kp, int nk, double* out, int threads) {
  if (threads < 1) threads = 1;
  for (int v = 0; v < g->nvars; v++) {
    if (g->p.geometry == SSGO_RL) rl_regular(g, spectral, ip, ni, jp, nj, out, v);
    else if (g->p.geometry == SSGO_RZ) rz_regular(g, spectral, ip, ni, kp, nk, out, v);
    else rlz_regular(g, spectral, ip, ni, jp, nj, kp, nk, out, v, threads);
  }
  return 0;
}
