"""SpringsteelGrid API mirror for RL / RZ / RLZ grids on the HIP library.

Names and behaviour follow the reference's grid API:
  createGrid            src/factory.jl:1104-1143 (RL 646-685, RLZ 688-740, RZ 461-502)
  spectralTransform_    spectralTransform!(grid)  (RL transforms_cylindrical.jl:153-157,
                        RLZ 524-528, RZ transforms_cartesian.jl:627-631)
  gridTransform_        gridTransform!(grid)      (RL 254-257, RLZ 654-657, RZ 734-737)
  spectralTransform / gridTransform   the explicit-array forms (same files)
  applyFilter_          applyFilter!(grid)        (src/filtering.jl:262-273; spectralTransform_ ends with it)
  getGridpoints         RL transforms_cylindrical.jl:75-90, RLZ 454-476, RZ transforms_cartesian.jl:359-373
Python cannot spell ``!``, so the in-place forms carry a trailing underscore.

Host arrays are NumPy arrays in Julia's column-major layout:
  grid.physical  float64 (npts, nvars, nslots), order="F"
  grid.spectral  float64 (nspec, nvars),        order="F"
Device-resident arrays (``*_device``) are torch tensors with the same bytes:
  physical (nslots, nvars, npts) and spectral (nvars, nspec), C-contiguous.
Every call runs the HIP kernels; there is no CPU path.
"""
from __future__ import annotations

import math

import numpy as np

from . import capi
from .errors import ArgumentError
from .filters import filter_weights
from .params import (SpringsteelGridParameters, compute_derived_params, normalize_geometry, quadpoints, resolve,
                     ring_layout)

def _quadpoints(mubar: int, quadrature: str):
    """Cell-local quadrature positions (src/CubicBSpline.jl:172-200)."""
    return quadpoints(mubar, quadrature)


class SpringsteelGrid:
    """A grid: parameters, host arrays and the device plan (created lazily)."""

    def __init__(self, params: SpringsteelGridParameters, precision: int = capi.SSG_FP64):
        self.params = params
        self.geometry = normalize_geometry(params.geometry)
        self.resolved = resolve(params)
        self.precision = precision
        p = params
        nvars = self.resolved.nvars
        K = p.iDim + p.patchOffsetL
        if self.geometry in ("RL", "SL"):
            self.npts, self.nspec, self.nslots = p.jDim, p.b_iDim * (1 + 2 * K), 5
        elif self.geometry == "RZ":
            self.npts, self.nspec, self.nslots = p.iDim * p.kDim, p.b_kDim * p.b_iDim, 5
        else:
            self.npts, self.nspec, self.nslots = p.kDim * p.jDim, p.b_kDim * p.b_iDim * (1 + 2 * K), 7
        self.spectral = np.zeros((self.nspec, nvars), dtype=np.float64, order="F")
        # the FP32 form stores physical in float32 (spectral stays float64)
        self.phys_dtype = np.float32 if precision == capi.SSG_FP32_REFINE else np.float64
        self.physical = np.zeros((self.npts, nvars, self.nslots), dtype=self.phys_dtype, order="F")
        self._plan = None

    @property
    def plan(self) -> capi.Plan:
        if self._plan is None:
            self._plan = capi.Plan(self.resolved, self.precision)
            d = self._plan.dims
            assert (d.npts, d.nspec, d.nslots) == (self.npts, self.nspec, self.nslots)
            if self.resolved.has_filters:
                names = sorted(self.params.vars, key=lambda k: self.params.vars[k])
                fw, cw = filter_weights(self.params, names, d.kDim_wn, d.b_kDim)
                self._plan.set_filter(fw, cw)
        return self._plan

    def close(self):
        if self._plan is not None:
            self._plan.close()
            self._plan = None


def createGrid(gp: SpringsteelGridParameters, precision: int = capi.SSG_FP64) -> SpringsteelGrid:
    """createGrid (src/factory.jl:1104-1143) for RL / RZ / RLZ; builds the device plan eagerly so
    that set-up errors (PosDefException, DomainError, ...) surface here as in the reference."""
    gp = compute_derived_params(gp)
    g = SpringsteelGrid(gp, precision)
    _ = g.plan
    return g


def _check_host(a: np.ndarray, shape, dtype=np.float64):
    if not (isinstance(a, np.ndarray) and a.dtype == dtype and a.shape == tuple(shape)
            and a.flags.f_contiguous):
        raise ArgumentError(f"expected a Fortran-ordered {np.dtype(dtype).name} array of shape {tuple(shape)}")


def spectralTransform(grid: SpringsteelGrid, physical: np.ndarray, spectral: np.ndarray) -> np.ndarray:
    """Explicit-array forward transform: reads physical[:, :, 1], writes spectral (physical
    float32 for an FP32 grid)."""
    _check_host(physical, (grid.npts, grid.resolved.nvars, grid.nslots), grid.phys_dtype)
    _check_host(spectral, (grid.nspec, grid.resolved.nvars))
    grid.plan.spectral(physical.ctypes.data, spectral.ctypes.data, capi.SSG_MEM_HOST)
    return spectral


def gridTransform(grid: SpringsteelGrid, physical: np.ndarray, spectral: np.ndarray) -> np.ndarray:
    """Explicit-array inverse transform: reads spectral, writes every slot of physical (float32
    for an FP32 grid)."""
    _check_host(physical, (grid.npts, grid.resolved.nvars, grid.nslots), grid.phys_dtype)
    _check_host(spectral, (grid.nspec, grid.resolved.nvars))
    grid.plan.grid(spectral.ctypes.data, physical.ctypes.data, capi.SSG_MEM_HOST)
    return physical


def spectralTransform_(grid: SpringsteelGrid) -> np.ndarray:
    """spectralTransform!(grid): physical[:, :, 1] -> grid.spectral, then applyFilter!."""
    spectralTransform(grid, grid.physical, grid.spectral)
    return applyFilter_(grid)


def applyFilter(grid: SpringsteelGrid, spectral: np.ndarray) -> np.ndarray:
    """The grid's fourier_filter / chebyshev_filter applied to a spectral array in place
    (applyFilter!'s body, src/filtering.jl:262-428); unchanged when both dicts are empty."""
    _check_host(spectral, (grid.nspec, grid.resolved.nvars))
    grid.plan.apply_filter(spectral.ctypes.data, capi.SSG_MEM_HOST)
    return spectral


def applyFilter_(grid: SpringsteelGrid) -> np.ndarray:
    """applyFilter!(grid) on grid.spectral."""
    return applyFilter(grid, grid.spectral)


def applyFilter_device(grid: SpringsteelGrid, spectral):
    """applyFilter! on a device-resident spectral tensor, on torch's current stream."""
    _check_dev(spectral, (grid.resolved.nvars, grid.nspec))
    grid.plan.apply_filter(spectral.data_ptr(), capi.SSG_MEM_DEVICE, _current_stream(spectral))
    return spectral


def gridTransform_(grid: SpringsteelGrid) -> np.ndarray:
    """gridTransform!(grid): grid.spectral -> every slot of grid.physical."""
    return gridTransform(grid, grid.physical, grid.spectral)


def _current_stream(t) -> int:
    import torch
    return torch.cuda.current_stream(t.device).cuda_stream


def _check_dev(t, shape, dtype_name="float64"):
    import torch
    if not (isinstance(t, torch.Tensor) and t.is_cuda and t.is_contiguous() and tuple(t.shape) == tuple(shape)
            and str(t.dtype) == f"torch.{dtype_name}"):
        raise ArgumentError(f"expected a contiguous CUDA {dtype_name} tensor of shape {tuple(shape)}")


def spectralTransform_device(grid: SpringsteelGrid, physical, spectral):
    """Device-resident forward transform on torch's current stream (asynchronous); physical
    float32 for an FP32 grid."""
    _check_dev(physical, (grid.nslots, grid.resolved.nvars, grid.npts), np.dtype(grid.phys_dtype).name)
    _check_dev(spectral, (grid.resolved.nvars, grid.nspec))
    grid.plan.spectral(physical.data_ptr(), spectral.data_ptr(), capi.SSG_MEM_DEVICE, _current_stream(physical))
    return spectral


def gridTransform_device(grid: SpringsteelGrid, spectral, physical):
    """Device-resident inverse transform on torch's current stream (asynchronous); physical
    float32 for an FP32 grid."""
    _check_dev(physical, (grid.nslots, grid.resolved.nvars, grid.npts), np.dtype(grid.phys_dtype).name)
    _check_dev(spectral, (grid.resolved.nvars, grid.nspec))
    grid.plan.grid(spectral.data_ptr(), physical.data_ptr(), capi.SSG_MEM_DEVICE, _current_stream(physical))
    return physical


def getGridpoints(grid: SpringsteelGrid) -> np.ndarray:
    """Physical point coordinates in array order: (r, lambda) for RL, (r, lambda, z) for RLZ,
    (x, z) for RZ, (theta, lambda) for SL and (theta, lambda, z) for SLZ
    (src/transforms_spherical.jl:53-78, 399-420)."""
    p = grid.params
    DX = (p.iMax - p.iMin) / p.num_cells
    q = _quadpoints(p.mubar, p.quadrature)
    xr = np.array([p.iMin + mc * DX + q[mu] * DX for mc in range(p.num_cells) for mu in range(p.mubar)])
    zs = None
    if grid.geometry in ("RZ", "RLZ", "SLZ"):
        scale, offset = -0.5 * (p.kMax - p.kMin), 0.5 * (p.kMin + p.kMax)
        zs = np.array([math.cos((n - 1) * math.pi / (p.kDim - 1)) * scale + offset for n in range(1, p.kDim + 1)])
    if grid.geometry == "RZ":
        return np.column_stack([np.repeat(xr, p.kDim), np.tile(zs, p.iDim)])
    rows = []
    for r, (n, _, ymin) in enumerate(ring_layout(p), start=1):
        lam = np.array([ymin + (2 * math.pi * (l - 1) / n) for l in range(1, n + 1)])
        if grid.geometry in ("RL", "SL"):
            rows.append(np.column_stack([np.full(n, xr[r - 1]), lam]))
        else:
            rows.append(np.column_stack([np.full(n * p.kDim, xr[r - 1]), np.repeat(lam, p.kDim),
                                         np.tile(zs, n)]))
    return np.vstack(rows)


# ----------------------------------------------------------------------------
# Regular-grid output: getRegularGridpoints / regularGridTransform
# (src/transforms_cylindrical.jl:878-893 and 929-1027 (RL), 1117-1137 and
# 1168-1294 (RLZ); src/transforms_cartesian.jl:898-912 and 946-1016 (RZ))
# ----------------------------------------------------------------------------
def _regular_axes(grid: SpringsteelGrid):
    p = grid.params
    r = np.linspace(p.iMin, p.iMax, p.i_regular_out)
    lam = np.array([2 * math.pi * (j - 1) / p.j_regular_out for j in range(1, p.j_regular_out + 1)])
    z = np.linspace(p.kMin, p.kMax, p.k_regular_out) if _eval_geometry(grid) in ("RZ", "RLZ") else None
    return r, lam, z


def _eval_geometry(grid: SpringsteelGrid) -> str:
    """The evaluation paths of SL / SLZ grids are those of RL / RLZ, with colatitude for
    radius: the reference's SL / SLZ regularGridTransform (src/transforms_spherical.jl:
    849-1131) restates the RL / RLZ one line for line, and evaluate_unstructured
    (src/interpolation.jl:1782-1800) sends SL to _eval_unstructured_rl and SLZ to
    _eval_unstructured_rlz; no ring table enters either."""
    return {"SL": "RL", "SLZ": "RLZ"}.get(grid.geometry, grid.geometry)


def getRegularGridpoints(grid: SpringsteelGrid) -> np.ndarray:
    """Uniform (r, lambda) (RL), (r, lambda, z) (RLZ) or (x, z) (RZ) points, the last
    coordinate fastest; sizes from i_regular_out, j_regular_out, k_regular_out."""
    r, lam, z = _regular_axes(grid)
    g = _eval_geometry(grid)
    if g == "RL":
        return np.column_stack([np.repeat(r, len(lam)), np.tile(lam, len(r))])
    if g == "RZ":
        return np.column_stack([np.repeat(r, len(z)), np.tile(z, len(r))])
    return np.column_stack([np.repeat(r, len(lam) * len(z)), np.tile(np.repeat(lam, len(z)), len(r)),
                            np.tile(z, len(r) * len(lam))])


def _regular_points(grid: SpringsteelGrid, args):
    """(i_pts, j_pts, k_pts) from either explicit point vectors or a points matrix, the
    matrix form reduced the way the reference does (RL: sorted unique radii, the first
    n_lambda rows' angles; RZ / RLZ: sorted unique values of every column)."""
    g = _eval_geometry(grid)
    if len(args) == 1:
        pts = np.asarray(args[0], dtype=np.float64)
        if g == "RL":
            r = np.unique(pts[:, 0])
            nl = pts.shape[0] // len(r)
            return r, pts[:nl, 1].copy(), None
        if g == "RZ":
            return np.unique(pts[:, 0]), None, np.unique(pts[:, 1])
        return np.unique(pts[:, 0]), np.unique(pts[:, 1]), np.unique(pts[:, 2])
    if g == "RL" and len(args) == 2:
        return np.asarray(args[0], float), np.asarray(args[1], float), None
    if g == "RZ" and len(args) == 2:
        return np.asarray(args[0], float), None, np.asarray(args[1], float)
    if g == "RLZ" and len(args) == 3:
        return tuple(np.asarray(a, float) for a in args)
    raise ArgumentError("regularGridTransform takes a points matrix or one point vector per dimension")


def _regular_npr(grid: SpringsteelGrid, i, j, k) -> int:
    return len(i) * (len(j) if j is not None else 1) * (len(k) if k is not None else 1)


def regularGridTransform(grid: SpringsteelGrid, *args, spectral: np.ndarray = None) -> np.ndarray:
    """Host form: evaluates grid.spectral (or the given spectral array) at the tensor
    product of the points; returns (npr, nvars, nslots), Fortran order."""
    i, j, k = _regular_points(grid, args)
    nv = grid.resolved.nvars
    spec = grid.spectral
    if spectral is not None:
        _check_host(spectral, (grid.nspec, nv))
        spec = spectral
    out = np.zeros((_regular_npr(grid, i, j, k), nv, grid.nslots), dtype=np.float64, order="F")
    grid.plan.regular(spec.ctypes.data, i, j, k, out.ctypes.data, capi.SSG_MEM_HOST)
    return out


def regularGridTransform_device(grid: SpringsteelGrid, spectral, *args):
    """Device form on torch's current stream: spectral (nvars, nspec) tensor; returns a
    (nslots, nvars, npr) float64 tensor (the same column-major bytes as the host form)."""
    import torch
    _check_dev(spectral, (grid.resolved.nvars, grid.nspec))
    i, j, k = _regular_points(grid, args)
    out = torch.empty((grid.nslots, grid.resolved.nvars, _regular_npr(grid, i, j, k)), dtype=torch.float64,
                      device=spectral.device)
    grid.plan.regular(spectral.data_ptr(), i, j, k, out.data_ptr(), capi.SSG_MEM_DEVICE, _current_stream(spectral))
    return out



def _unstructured_args(grid: SpringsteelGrid, pts, out_of_bounds):
    pts = np.asarray(pts, dtype=np.float64)
    ndim = 3 if _eval_geometry(grid) == "RLZ" else 2
    if pts.ndim != 2 or pts.shape[1] != ndim:
        raise ArgumentError(f"points must be an (N, {ndim}) matrix for a {grid.geometry} grid")
    if isinstance(out_of_bounds, str):
        if out_of_bounds not in ("nan", "error"):
            raise ArgumentError("out_of_bounds must be 'nan', 'error' or a number")
        return pts, math.nan, out_of_bounds == "error"
    return pts, float(out_of_bounds), False


def evaluate_unstructured(grid: SpringsteelGrid, pts, vars=None, out_of_bounds="nan",
                          spectral: np.ndarray = None) -> np.ndarray:
    """evaluate_unstructured (src/interpolation.jl:1735-1822), host form: the value of
    every variable at the rows of pts (N, ndim) -- RL (r, lambda), RZ (x, z), RLZ
    (r, lambda, z) -- as an (N, nvars) Fortran-order array.  Points outside the
    domain get NaN, the numeric out_of_bounds, or raise DomainError ('error');
    variables not named in vars are NaN."""
    pts, fill, err = _unstructured_args(grid, pts, out_of_bounds)
    nv = grid.resolved.nvars
    spec = grid.spectral
    if spectral is not None:
        _check_host(spectral, (grid.nspec, nv))
        spec = spectral
    out = np.empty((pts.shape[0], nv), dtype=np.float64, order="F")
    grid.plan.unstructured(spec.ctypes.data, pts, out.ctypes.data, fill, err, capi.SSG_MEM_HOST)
    if vars is not None:
        keep = {grid.params.vars[n] - 1 for n in vars if n in grid.params.vars}
        for v in range(nv):
            if v not in keep:
                out[:, v] = math.nan
    return out


def evaluate_unstructured_device(grid: SpringsteelGrid, spectral, pts, out_of_bounds="nan"):
    """Device form on torch's current stream: spectral (nvars, nspec) tensor, pts a host
    (N, ndim) array; returns an (nvars, N) float64 tensor (the column-major bytes of the
    host form's (N, nvars) result)."""
    import torch
    _check_dev(spectral, (grid.resolved.nvars, grid.nspec))
    pts, fill, err = _unstructured_args(grid, pts, out_of_bounds)
    out = torch.empty((grid.resolved.nvars, pts.shape[0]), dtype=torch.float64, device=spectral.device)
    grid.plan.unstructured(spectral.data_ptr(), pts, out.data_ptr(), fill, err, capi.SSG_MEM_DEVICE,
                           _current_stream(spectral))
    return out


def set_ahat(grid: SpringsteelGrid, ahat=None):
    """R3X background coefficients (the reference's spline.ahat per spectral line,
    src/CubicBSpline.jl:1367-1396; _set_wavenumber_ahat!, src/multipatch_registry.jl):
    an array of the layout of grid.spectral (nspec, nvars), a torch tensor of the
    device layout (nvars, nspec), or None to clear.  Variables with an R3X side use
    it in gridTransform; the others ignore it."""
    if ahat is None:
        grid.plan.set_ahat(None)
        return
    if isinstance(ahat, np.ndarray):
        a = np.asfortranarray(ahat, dtype=np.float64)
        _check_host(a, (grid.nspec, grid.resolved.nvars))
        grid.plan.set_ahat(a.ctypes.data, capi.SSG_MEM_HOST)
        return
    _check_dev(ahat, (grid.resolved.nvars, grid.nspec))
    grid.plan.set_ahat(ahat.data_ptr(), capi.SSG_MEM_DEVICE)
