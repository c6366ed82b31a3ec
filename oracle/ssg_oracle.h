/*
 * ssg_oracle.h -- CPU restatement of Springsteel.jl's RL / RZ / RLZ transform
 * path, used ONLY as the test oracle and as the CPU baseline of bench.py.
 *
 * TEST INFRASTRUCTURE: nothing in the product library links, loads or calls
 * this code.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg may use it, and there only as the checker / the timed CPU port.
 *
 * Every function follows the reference file:line named in its comment in
 * ssg_oracle.c.  Layouts are the Julia column-major arrays of the reference:
 *   physical[npts, nvars, nslots]   (nslots = 5 for RL/RZ, 7 for RLZ)
 *   spectral[nspec, nvars]
 */
#ifndef SSG_ORACLE_H
#define SSG_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

enum { SSGO_RL = 1, SSGO_RZ = 2, SSGO_RLZ = 3, SSGO_SL = 4, SSGO_SLZ = 5 };
/* spline BC kinds (CubicBSpline.jl:300-304 GammaBC kinds) */
enum { SSGO_BC_R0 = 0, SSGO_BC_A1 = 1, SSGO_BC_A2 = 2, SSGO_BC_PER = 3, SSGO_BC_R3 = 4, SSGO_BC_R3X = 5 };
/* Chebyshev BC kinds (Chebyshev.jl:949-963 _bc_type) */
enum { SSGO_CBC_R0 = 0, SSGO_CBC_R1T0 = 1, SSGO_CBC_R1T1 = 2 };

typedef struct ssgo_params {
  int geometry;
  int num_cells, mubar, quadrature; /* quadrature: 0 = :gauss, 1 = :regular */
  double iMin, iMax;
  int nvars;
  const double* l_q;      /* [nvars] */
  const int* bcl_kind;    /* [nvars] */
  const double* bcl_ab;   /* [2*nvars] (alpha, beta) */
  const int* bcr_kind;
  const double* bcr_ab;
  double kMin, kMax;
  int kDim, b_kDim;
  const int* bcb_kind;    /* [nvars] Chebyshev bottom */
  const int* bct_kind;    /* [nvars] Chebyshev top */
  const int* max_wavenumber; /* [nvars]; -1 = ring specific (RL per var) */
  int ring_kmax;          /* RLZ rings shared by all vars; -1 = ring specific */
  int patchOffsetL;
} ssgo_params;

typedef struct ssgo_grid ssgo_grid;

ssgo_grid* ssgo_grid_create(const ssgo_params* p);
void ssgo_grid_destroy(ssgo_grid* g);
/* dims[0]=npts, dims[1]=nspec, dims[2]=nslots, dims[3]=iDim, dims[4]=b_iDim,
 * dims[5]=jDim, dims[6]=kDim, dims[7]=b_kDim */
void ssgo_dims(const ssgo_grid* g, long* dims);
const char* ssgo_error(void);
/* threads: number of OpenMP threads to use (1 = the reference's behaviour) */
int ssgo_spectral_transform(ssgo_grid* g, const double* physical, double* spectral, int threads);
int ssgo_grid_transform(ssgo_grid* g, const double* spectral, double* physical, int threads);
/* the tile pipeline (src/tiling.jl): SA of every line (B -> A) and the inverse from A */
int ssgo_spline_transform(ssgo_grid* g, const double* shared, double* spectral, int threads);
int ssgo_tile_transform(ssgo_grid* g, const double* spectral, double* physical, int threads);
/* pts[npts, ndim] column-major, ndim = 2 (RL, RZ) or 3 (RLZ) */
void ssgo_gridpoints(const ssgo_grid* g, double* pts);

/* regularGridTransform (RL: i = r, j = lambda; RZ: i = x, k = z; RLZ: all
 * three).  out[npr, nvars, nslots] column-major with npr = ni * nj * nk over
 * the dimensions the geometry has, the last one fastest. */
int ssgo_regular_transform(ssgo_grid* g, const double* spectral, const double* ip, int ni, const double* jp, int nj,
                           const double* kp, int nk, double* out, int threads);
/* evaluate_unstructured: the value of every variable at n points pts[n, ndim]
 * (column-major; RL (r, lambda), RZ (x, z), RLZ (r, lambda, z)) into
 * out[n, nvars] column-major; points outside the spline (and Chebyshev) domain
 * get `fill` (the reference's NaN default, or its numeric fill value) */
int ssgo_evaluate_unstructured(ssgo_grid* g, const double* spectral, const double* pts, long n, double* out,
                               double fill, int threads);
/* R3X background coefficients (the reference's spline.ahat, set per spectral
 * line through _set_wavenumber_ahat! / set_ahat_r3x!): ahat has the layout of
 * spectral (nspec, nvars), one b_iDim vector per line; NULL clears it.  Only
 * variables with an R3X BC use it, in gridTransform's SAtransform!
 * (src/CubicBSpline.jl:1367-1396): a = G' (PQ \ G (b - pq ahat)) + ahat. */
void ssgo_set_ahat(ssgo_grid* g, const double* ahat);
/* test use: 1 makes the regular transform's d2/dz2 slot hold the scale of the
 * rounding error of the reference's FP64 formula there (eps times it bounds
 * that error) instead of the value, as does ssgo_cheb_eval with derivative 2;
 * 0 restores the value */
void ssgo_set_dzz_bound(ssgo_grid* g, int on);
/* one Chebyshev column: b (b_kDim) -> CAtransform! -> value / d/dz / d2/dz2 at z */
double ssgo_cheb_eval(const ssgo_grid* g, int v, const double* b, double z, int derivative);

/* ---- component level entry points (for pinning the oracle) ---- */
double ssgo_basis(double xmin, double xmax, int num_cells, int m, double x, int derivative);
void ssgo_quadrature(int mubar, int quadrature, double* qpts, double* qwts);
void ssgo_r2hc(int n, const double* x, double* out);
void ssgo_hc2r(int n, const double* x, double* out);
void ssgo_redft00(int n, const double* x, double* out);
/* A: dense n x n column-major (upper triangle used, like Symmetric(A)); L: 4 x n band */
int ssgo_band_factor3(int n, const double* A, double* L);
void ssgo_band_solve3(int n, const double* L, const double* b, double* x);

/* 1-D spline object: same SplineParameters fields as the reference */
typedef struct ssgo_spline ssgo_spline;
/* one line's SAtransform! with ahat (NULL: none); a receives Mdim values */
void ssgo_spline_sa_ahat(const ssgo_spline* s, const double* b, const double* ahat, double* a);
ssgo_spline* ssgo_spline_create(double xmin, double xmax, int num_cells, int mubar, int quadrature,
                                double l_q, int bcl_kind, double bcl_a, double bcl_b,
                                int bcr_kind, double bcr_a, double bcr_b);
void ssgo_spline_destroy(ssgo_spline* s);
int ssgo_spline_mdim(const ssgo_spline* s);
int ssgo_spline_minterior(const ssgo_spline* s);
void ssgo_spline_mishpoints(const ssgo_spline* s, double* x);
void ssgo_spline_pq(const ssgo_spline* s, double* P, double* Q);  /* Mdim x Mdim each */
void ssgo_spline_band(const ssgo_spline* s, double* L);           /* 4 x Minterior */
void ssgo_spline_sb(const ssgo_spline* s, const double* u, double* b);
void ssgo_spline_sa(const ssgo_spline* s, const double* b, double* a);
void ssgo_spline_si(const ssgo_spline* s, const double* a, const double* pts, int npts, int derivative, double* u);

#ifdef __cplusplus
}
#endif
#endif
