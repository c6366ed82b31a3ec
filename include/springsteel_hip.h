/*
 * springsteel_hip.h -- C-ABI of the MI355X transform library for
 * Springsteel.jl RL / RZ / RLZ grids.
 *
 * The library replaces the body of the reference's operator methods
 *   spectralTransform!(grid::_RLGrid)    src/transforms_cylindrical.jl:153-157 (-> 159-222)
 *   gridTransform!(grid::_RLGrid)        src/transforms_cylindrical.jl:254-257 (-> 259-428)
 *   spectralTransform!(grid::_RLZGrid)   src/transforms_cylindrical.jl:524-528 (-> 530-618)
 *   gridTransform!(grid::_RLZGrid)       src/transforms_cylindrical.jl:654-657 (-> 659-853)
 *   spectralTransform!(grid::_2DCartesianRZ)  src/transforms_cartesian.jl:627-631 (-> 579-617)
 *   gridTransform!(grid::_2DCartesianRZ)      src/transforms_cartesian.jl:734-737 (-> 658-724)
 * and the explicit-array forms spectralTransform(grid, physical, spectral) /
 * gridTransform(grid, physical, spectral) those methods call.  A plan is the
 * analogue of createGrid's basis set-up (src/factory.jl:646-740, 461-502).
 *
 * Memory contract: byte-identical to the Julia arrays.
 *   physical: Float64[npts, nvars, nslots] column-major (nslots 5 for RL/RZ, 7 for RLZ)
 *   spectral: Float64[nspec, nvars]        column-major (B-vectors, not coefficients)
 * spectral_transform reads physical[:, :, 1] only and overwrites spectral.
 * grid_transform reads spectral and overwrites every slot of physical.
 *
 * All entry points return SSG_OK (0) or a negative code; ssg_last_error()
 * returns the message of the calling thread's last failure.  One plan is not
 * re-entrant: calls on the same plan must not overlap.  Kernels run on the
 * stream given (NULL = the null stream); host-memory calls synchronise that
 * stream before returning, device-memory calls are asynchronous.
 */
#ifndef SPRINGSTEEL_HIP_H
#define SPRINGSTEEL_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

/* return codes */
#define SSG_OK 0
#define SSG_EINVAL (-22)   /* malformed descriptor or argument (Julia ArgumentError) */
#define SSG_ENOTSUP (-95)  /* feature of the reference this library does not run */
#define SSG_EPOSDEF (-33)  /* (P+Q) not positive definite (LinearAlgebra.PosDefException) */
#define SSG_EDOMAIN (-34)  /* unsupported BC combination (DomainError) */
#define SSG_EBOUNDS (-35)  /* RLZ with b_kDim < 3, a tile that does not fit its patch (BoundsError) */
#define SSG_ENOMEM (-12)
#define SSG_EHIP (-5)      /* HIP runtime error, message in ssg_last_error() */

/* geometries (src/types.jl:366 RZ_Grid, 379 RL_Grid, 416 RLZ_Grid; SL and SLZ:
 * src/transforms_spherical.jl:27-28, the RL / RLZ transforms on rings of
 * 4 round(sin(theta) max_ri) points, src/factory.jl:159-163, 626-643) */
#define SSG_RL 1
#define SSG_RZ 2
#define SSG_RLZ 3
#define SSG_SL 4
#define SSG_SLZ 5

/* spline BC kinds: the GammaBC kinds of src/CubicBSpline.jl:300-304 */
#define SSG_BC_R0 0        /* NaturalBC / CubicBSpline.R0 */
#define SSG_BC_ALPHA1 1    /* rank 1: Dict("α1" => a, "β1" => b) */
#define SSG_BC_ALPHA2 2    /* rank 2: Dict("α2" => a, "β2" => b) */
#define SSG_BC_PERIODIC 3  /* rejected: dense periodic factor (ENOTSUP) */
#define SSG_BC_R3 4        /* homogeneous rank 3 */
#define SSG_BC_R3X 5       /* inhomogeneous rank 3: folds as R3; background coefficients via ssg_plan_set_ahat */

/* Chebyshev BC kinds: src/Chebyshev.jl:949-963 _bc_type */
#define SSG_CBC_R0 0
#define SSG_CBC_R1T0 1     /* key "α0" */
#define SSG_CBC_R1T1 2     /* key "α1" */
#define SSG_CBC_OTHER 3    /* any other key: DomainError in the reference (EDOMAIN) */

/* where the physical / spectral pointers live */
#define SSG_MEM_HOST 0
#define SSG_MEM_DEVICE 1

/* arithmetic form */
#define SSG_FP64 0          /* plain form: every stage in FP64 */
#define SSG_FP32_REFINE 1   /* FP32 storage of physical, FP32 synthesis; FP64 forward path, solve and evaluation */

/* Everything createGrid derives for one RL / RZ / RLZ / SL / SLZ grid
 * (SpringsteelGridParameters src/Springsteel.jl:425-465, resolved per variable).
 * For SL / SLZ, "radial" below reads colatitude theta. */
typedef struct ssg_desc {
  int geometry;            /* SSG_RL / SSG_RZ / SSG_RLZ / SSG_SL / SSG_SLZ */
  int num_cells;           /* spline cells */
  int mubar;               /* quadrature points per cell (1..5 for Gauss) */
  int quadrature;          /* 0 = :gauss, 1 = :regular */
  double iMin, iMax;       /* radial / x domain */
  int nvars;
  const double* l_q;       /* [nvars] spline filter length */
  const int* bcl_kind;     /* [nvars] SSG_BC_* left */
  const double* bcl_ab;    /* [2*nvars] (alpha, beta) for ALPHA1/ALPHA2 */
  const int* bcr_kind;     /* [nvars] SSG_BC_* right */
  const double* bcr_ab;    /* [2*nvars] */
  double kMin, kMax;       /* Chebyshev domain (RZ, RLZ) */
  int kDim, b_kDim;        /* Chebyshev points and modes (RZ, RLZ) */
  const int* bcb_kind;     /* [nvars] SSG_CBC_* bottom (RZ, RLZ) */
  const int* bct_kind;     /* [nvars] SSG_CBC_* top (RZ, RLZ) */
  const int* max_wavenumber; /* [nvars] RL / SL ring cap per variable, -1 = ring specific */
  int ring_kmax;           /* RLZ / SLZ: cap shared by the rings of all variables, -1 = ring specific */
  int patchOffsetL;        /* (spectralIndexL - 1) * mubar */
  int precision;           /* SSG_FP64 / SSG_FP32_REFINE */
  int filter_flags;        /* nonzero -> ENOTSUP: filter weights are passed with ssg_plan_set_filter */
  int wavenumber_ahat;     /* nonzero -> ENOTSUP: background coefficients are passed with ssg_plan_set_ahat */
} ssg_desc;

typedef struct ssg_dims {
  long npts;      /* physical points per variable */
  long nspec;     /* spectral rows per variable */
  int nvars, nslots;
  int iDim, b_iDim, jDim, kDim, b_kDim, kDim_wn;
} ssg_dims;

typedef struct ssg_plan ssg_plan;

/* Host FP64 set-up (quadrature, Gamma fold, P+Q and its banded LL^T factor,
 * ring and phase tables, Chebyshev matrices) and upload to the current device.
 * Replaces createGrid's basis construction (src/factory.jl:646-740, 461-502). */
int ssg_plan_create(const ssg_desc* desc, ssg_plan** out);
int ssg_plan_destroy(ssg_plan* plan);
int ssg_plan_get_dims(const ssg_plan* plan, ssg_dims* out);

/* spectralTransform(grid, physical, spectral):
 *   src/transforms_cylindrical.jl:159-222 (RL), 530-618 (RLZ);
 *   src/transforms_cartesian.jl:579-617 (RZ);
 *   src/transforms_spherical.jl:121-210 (SL), 464-580 (SLZ).
 * Like the reference's explicit-array form it does not filter: spectralTransform!
 * is this call followed by ssg_apply_filter. */
int ssg_spectral_transform(ssg_plan* plan, const double* physical, double* spectral, int mem, void* stream);

/* gridTransform(grid, physical, spectral):
 *   src/transforms_cylindrical.jl:259-428 (RL), 659-853 (RLZ);
 *   src/transforms_cartesian.jl:658-724 (RZ);
 *   src/transforms_spherical.jl:218-390 (SL), 587-800 (SLZ). */
int ssg_grid_transform(ssg_plan* plan, const double* spectral, double* physical, int mem, void* stream);

/* The FP32 form (precision SSG_FP32_REFINE; no counterpart in the reference):
 * the same two transforms with physical stored in Float32, [npts, nvars,
 * nslots] column-major; spectral stays Float64.  The forward path, the banded
 * solve and the radial evaluation run in FP64 (the d/dr and d2/dr2 slots
 * amplify rounding of everything before them by (L/DX)^2); the ring and
 * Chebyshev synthesis run in FP32.  Plans of the other form return SSG_EINVAL,
 * as ssg_spectral_transform / ssg_grid_transform do on an FP32 plan. */
int ssg_spectral_transform_f32(ssg_plan* plan, const float* physical, double* spectral, int mem, void* stream);
int ssg_grid_transform_f32(ssg_plan* plan, const double* spectral, float* physical, int mem, void* stream);

/* ssg_grid_transform without allocation.  No counterpart in the reference: it
 * is for callers that cannot allocate device memory during the call (code
 * under HIP graph capture, tight serving loops, harnesses that forbid it).
 * Device pointers only, as ssg_grid_transform with SSG_MEM_DEVICE (SSG_FP64
 * plans; an FP32 plan returns SSG_EINVAL), with the same kernels and results;
 * its intermediates (the banded-solve lanes, the
 * ring synthesis input, the RZ inverse-Chebyshev input, btilde while ahat is
 * set) live in one static device array of SSG_NOALLOC_SCRATCH_BYTES per GPU,
 * held by the library until the process ends, instead of the plan's buffers.
 * A plan whose intermediates exceed it returns SSG_ENOTSUP.  The call creates
 * no events (stage timing does not record it) and calls no hipMalloc or
 * hipFree.  Every call on one GPU shares that GPU's array and starts by
 * zeroing the part it uses: calls that overlap on one GPU, on any plans and
 * streams, race. */
#define SSG_NOALLOC_SCRATCH_BYTES (64u << 20)
int ssg_grid_transform_noalloc(ssg_plan* plan, const double* spectral, double* physical, void* stream);

/* R3X background coefficients (the reference's spline.ahat: the per-wavenumber
 * registry _set_wavenumber_ahat!, src/multipatch_registry.jl:23-60, or
 * set_ahat_r3x! on a spline, src/CubicBSpline.jl:1437-1467).  ahat has the
 * layout of spectral, [nvars][nspec] column-major: one b_iDim vector per
 * spectral line.  For the variables with an R3X side, gridTransform's solve
 * becomes SAtransform!'s R3X branch (src/CubicBSpline.jl:1367-1396),
 *   a = G^T (PQ \ G (b - pq ahat)) + ahat,  pq = P + Q;
 * other variables ignore it.  The plan keeps a device copy (host or device
 * source per mem); NULL clears it, as does a plan without R3X sides.  The call
 * synchronizes the device.  While set: the batching form rejects the plan
 * (as it rejects any plan with R3X sides), and ssg_regular_transform /
 * ssg_evaluate_unstructured return SSG_ENOTSUP. */
int ssg_plan_set_ahat(ssg_plan* plan, const double* ahat, int mem);

/* Spectral filters: the weights of applyFilter! (src/filtering.jl:262-428),
 * evaluated by the caller with the reference's _filter_weight (SpectralFilter /
 * GaussianFilter, src/filtering.jl:144-187) and its "default" key lookup.
 *   fourier_w:   [nvars][kDim_wn + 1] column-major, the weight of wavenumber k
 *                of variable v at fourier_w[v * (kDim_wn + 1) + k]; 1.0 for a
 *                variable without a Fourier filter; NULL for an empty
 *                fourier_filter dict.
 *   chebyshev_w: [nvars][b_kDim], the weight of Chebyshev mode z_b - 1 at
 *                chebyshev_w[v * b_kDim + z_b - 1]; NULL for an empty dict.
 * RL / SL grids use fourier_w (src/filtering.jl:320-364), RZ grids chebyshev_w
 * (290-314), RLZ / SLZ grids the product c_w * f_w per (z_b, k) block
 * (372-428).  Per spectral line as the reference decides it: a weight that is
 * isapprox 1.0 (Julia's default rtol = sqrt(eps)) leaves the line untouched, a
 * weight of 0.0 stores +0.0, any other weight multiplies.  Both NULL clears
 * the filter.  The plan keeps a device table of the lines to change; the call
 * synchronizes the device. */
int ssg_plan_set_filter(ssg_plan* plan, const double* fourier_w, const double* chebyshev_w);

/* applyFilter!(grid) (src/filtering.jl:262-273) on a spectral array of the plan's
 * layout, in place, host or device memory per mem; a no-op (the reference's
 * early exit) while no filter is set. */
int ssg_apply_filter(ssg_plan* plan, double* spectral, int mem, void* stream);

/* The tile pipeline of src/tiling.jl on the ring grids (RL, RLZ, SL, SLZ; FP64
 * plans whose b_iDim <= 256; otherwise SSG_ENOTSUP / SSG_EINVAL).  Arrays in the
 * layout of spectral / physical, host or device memory per mem.
 *
 * splineTransform!(sharedSpectral, tile) (src/tiling.jl:1490-1516 RL, 1675-1700 SL,
 * 1862-1892 RLZ, 2063-2093 SLZ): SAtransform of every line of `shared` (B coefficients) into `spectral`
 * (A coefficients, the folded solution, with the R3X background where the plan
 * has ahat). */
int ssg_spline_transform(ssg_plan* plan, const double* shared, double* spectral, int mem, void* stream);

/* tileTransform!(sharedSpectral, tile, physical, spectral) (src/tiling.jl:1540-1658
 * RL, 1720-1838 SL, 1915-2046 RLZ, 2113-2242 SLZ): the grid transform from A coefficients, without the solve:
 * SI, SIx, SIxx of every line, the inverse phase filter, the ring synthesis with
 * its azimuthal derivatives and, on RLZ / SLZ, the inverse Chebyshev stage, into
 * every physical slot. */
int ssg_tile_transform(ssg_plan* plan, const double* spectral, double* physical, int mem, void* stream);

/* sumSpectralTile!(patch, tile) (src/tiling.jl:1048-1170): adds every line of the
 * tile's spectral array (its 1 + 2 K_t lines per Chebyshev mode) to the patch's
 * same line at radial coefficients spectral_index_l .. spectral_index_l +
 * b_iDim_t - 1 (1-based, the tile's params.spectralIndexL).  Same geometry,
 * variables and Chebyshev modes; a tile that does not fit is SSG_EBOUNDS. */
int ssg_sum_spectral_tile(ssg_plan* patch, double* patch_spectral, ssg_plan* tile, const double* tile_spectral,
                          int spectral_index_l, int mem, void* stream);

/* setSpectralTile!(patch, tile) (src/tiling.jl:1204-1327): zeroes the patch's
 * spectral array, then writes the tile's lines as sumSpectralTile! adds them. */
int ssg_set_spectral_tile(ssg_plan* patch, double* patch_spectral, ssg_plan* tile, const double* tile_spectral,
                          int spectral_index_l, int mem, void* stream);

/* sumSharedSpectral(sharedSpectral, tile, patchMap, haloMap) for the ring grids,
 * the reference's generic method (src/tiling.jl:1378-1389): per variable column,
 * rows spectral_index_l .. spectral_index_l + b_iDim_t - 4 of `shared` (a
 * column-major (shared_rows, nvars) array) are set to rows 1 .. b_iDim_t - 3 of
 * the tile's spectral column (the maps are not used there). */
int ssg_sum_shared_spectral(double* shared, long shared_rows, ssg_plan* tile, const double* tile_spectral,
                            int spectral_index_l, int mem, void* stream);

/* Per-stage device time, for measurement.  enable != 0 records a HIP event
 * pair around every stage kernel on the call's stream (no extra waits in the
 * calls).  ssg_plan_stage_times waits for the recorded events, writes up to n
 * (stage name, average milliseconds per launch) pairs over every launch since
 * the previous query or set_timing call, and returns the number of stages
 * (or a negative code).  names/ms may be NULL to only count. */
int ssg_plan_set_timing(ssg_plan* plan, int enable);
int ssg_plan_stage_times(ssg_plan* plan, const char** names, double* ms, int n);

const char* ssg_last_error(void);

/* ---- Batching form --------------------------------------------------------
 * No counterpart in the reference, whose grids transform one call at a time;
 * it serves callers that step many independent RL / RLZ grids together (the
 * grids of a multi-patch or ensemble run).  The plans of a batch share one
 * arithmetic form; their device arrays are fixed when the batch is created.
 * Each stage kernel runs once per launch shape for all grids, and the launches
 * of a direction are captured in a hipGraph that each call replays: one
 * hipGraphLaunch per direction.  Per grid, the results are bitwise those of
 * ssg_spectral_transform / ssg_grid_transform on the same plan.  The plans
 * must outlive the batch; calls on one batch must not overlap. */
typedef struct ssg_batch ssg_batch;

/* physical[i] / spectral[i]: device pointers of grid i (layouts as above;
 * physical is Float64, or Float32 for a batch of SSG_FP32_REFINE plans) */
int ssg_batch_create(ssg_plan* const* plans, int nplans, void* const* physical, double* const* spectral,
                     ssg_batch** out);
int ssg_batch_spectral_transform(ssg_batch* batch, void* stream);
int ssg_batch_grid_transform(ssg_batch* batch, void* stream);
/* regularGridTransform(grid, i_pts, j_pts[, k_pts]): src/transforms_cylindrical.jl:
 * 929-1020 (RL: i = r, j = lambda), 1168-1287 (RLZ: r, lambda, z),
 * src/transforms_cartesian.jl:946-1001 (RZ: i = x, k = z),
 * src/transforms_spherical.jl:849-932 (SL: i = theta) and 1010-1124 (SLZ), which
 * restate the RL / RLZ ones.  Evaluates the spectral coefficients (b vectors, as
 * spectralTransform leaves them) on the tensor product of the points; the
 * pointer and count of the dimension a geometry lacks are ignored.
 *   out: [npr][nvars][nslots] column-major, npr = n_i * n_j * n_k over the
 *        dimensions present, the last one fastest; nslots 5 (RL: f, d/dr,
 *        d2/dr2, d/dlambda, d2/dlambda2; RZ: f, d/dx, d2/dx2, d/dz, d2/dz2)
 *        or 7 (RLZ: f, d/dr, d2/dr2, d/dlambda, d2/dlambda2, d/dz, d2/dz2).
 *   The point arrays are host arrays; spectral and out are host or device
 *   memory per mem.  Radii outside [iMin, iMax] return SSG_EDOMAIN (the
 *   reference's basis() DomainError); z is clamped as the reference does.
 *   The evaluation runs in FP64 for plans of either precision. */
int ssg_regular_transform(ssg_plan* plan, const double* spectral, const double* i_pts, int n_i, const double* j_pts,
                          int n_j, const double* k_pts, int n_k, double* out, int mem, void* stream);

/* evaluate_unstructured(grid, pts; out_of_bounds): src/interpolation.jl:1735-1822
 * (RL and SL 1628-1668, RZ 1376-1402, RLZ and SLZ 1670-1733).
 * The value of every variable at npts arbitrary points.
 *   pts: host array [npts][ndim] column-major, ndim 2 (RL: r, lambda; RZ:
 *        x, z) or 3 (RLZ: r, lambda, z).
 *   out: [npts][nvars] column-major, host or device memory per mem (as
 *        spectral).  Points outside [iMin, iMax] (and [kMin, kMax]) get `fill`
 *        (the reference's :nan default passes NaN, a numeric out_of_bounds its
 *        value); with oob_error != 0 any such point returns SSG_EDOMAIN
 *        instead (the reference's out_of_bounds = :error).
 *   The sine lines enter the Fourier sum as +rs sin(k lambda), as in the
 *   reference; the Chebyshev angle is clamped to [-1, 1].  FP64 for plans of
 *   either precision; the call may use the plan's scratch for up to 2^20
 *   points at a time and grows it on demand.  More than 48 Chebyshev modes
 *   (b_kDim) return SSG_ENOTSUP. */
int ssg_evaluate_unstructured(ssg_plan* plan, const double* spectral, const double* pts, long npts, double* out,
                              double fill, int oob_error, int mem, void* stream);

/* Autotuner: times the candidate launch shapes of ring synthesis, ring
 * analysis and SI on every grid's own arrays, keeps the fastest per grid and
 * captures the graphs again.  It runs the transforms, so it overwrites the
 * batch's physical and spectral arrays: tune before loading the fields. */
int ssg_batch_tune(ssg_batch* batch, void* stream);
/* kernel launches in the spectral / grid graph */
int ssg_batch_launches(const ssg_batch* batch, int* spectral_launches, int* grid_launches);
int ssg_batch_destroy(ssg_batch* batch);
/* The same tuner for one plan on its device arrays (no counterpart in the
 * reference; overwrites physical and spectral like a round trip). */
int ssg_plan_tune(ssg_plan* plan, void* physical, double* spectral, void* stream);
/* Tuning of a plan, five ints (0 = built-in choice), no counterpart in the
 * reference:
 *   [0] synthesis shape: 1 = one column tile x two quad tiles per wave,
 *       2 = two x one, 3 = two x two (FP32 form only);
 *   [1] synthesis quad groups per block (1, 2, 4);
 *   [2] analysis row tiles per wave (1, 2);
 *   [3] SI cell segments per lane (1, 2, 4, 8) of the unfused solve path
 *       (ring grids with b_iDim > 256; the fused kernel has no segments);
 *   [4] the spline solve: 0 or 2, the FP64 substitution (the only one: the
 *       FP32 form's evaluation of d/dr and d2/dr2 needs the FP64 solution).
 * Every choice gives the same results bitwise.  ssg_plan_set_tuning pins a choice (for example one
 * the tuner made in an earlier run); it applies to the plan's own calls and to
 * batches created or tuned after it.  Out-of-range values: SSG_EINVAL. */
int ssg_plan_get_tuning(const ssg_plan* plan, int* out5);
int ssg_plan_set_tuning(ssg_plan* plan, const int* in5);

#ifdef __cplusplus
}
#endif
#endif
