"""The tile pipeline of src/tiling.jl on the ring grids (RL, RLZ, SL, SLZ).

  calcTileSizes       src/tiling.jl:185-259 (cylindrical), 265-347 (spherical), with
                      _create_tile_from_patch 50-80
  splineTransform_    splineTransform!(sharedSpectral, tile)   1490-1516 RL, 1675-1700 SL,
                      1862-1892 RLZ, 2063-2093 SLZ
  tileTransform_      tileTransform!(sharedSpectral, tile, physical, spectral)
                      1540-1658 RL, 1720-1838 SL, 1915-2046 RLZ, 2113-2242 SLZ
  sumSpectralTile_    sumSpectralTile!(patch, tile)            1048-1170
  setSpectralTile_    setSpectralTile!(patch, tile)            1204-1327
  sumSharedSpectral   sumSharedSpectral(shared, tile, ...)      1379-1389 (the ring grids' method)

The transforms and the sums run in the HIP library (ssg_spline_transform,
ssg_tile_transform, ssg_sum_spectral_tile, ssg_set_spectral_tile,
ssg_sum_shared_spectral); this module holds the host-side decisions the
reference makes in Julia (the tile sizes) and the argument checks.  Host arrays
are Fortran-ordered as the grid's; ``*_device`` forms take torch tensors in the
device layout of the grid transforms.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

from . import capi
from .errors import ArgumentError, DomainError
from .grid import SpringsteelGrid, _check_dev, _check_host, _current_stream, createGrid
from .params import SpringsteelGridParameters, compute_derived_params, normalize_geometry, ring_layout

_RING = ("RL", "RLZ", "SL", "SLZ")


def _ring_geometry(geometry: str):
    if normalize_geometry(geometry) not in _RING:
        raise ArgumentError(f"the tile pipeline takes ring grids (RL, RLZ, SL, SLZ), not {geometry}")


def _ring_grid(g: SpringsteelGrid):
    _ring_geometry(g.geometry)


def tileParameters(p: SpringsteelGridParameters, num_tiles: int):
    """The parameters of the num_tiles tiles calcTileSizes makes of a patch with parameters
    `p` (derived fields filled in): at least 3 cells each, balanced by ring points from the
    outermost ring inwards (src/tiling.jl:185-259, 265-347); each tile has the patch's
    parameters on [iMin_t, iMax_t] with spectralIndexL (so patchOffsetL) at its first cell,
    and tile_num.  DomainError when a tile would get fewer than 3 cells."""
    _ring_geometry(p.geometry)
    if p.num_cells / num_tiles < 3.0:
        raise DomainError(f"{num_tiles}: Too many tiles for this grid (need at least 3 cells per tile)")
    DX = (p.iMax - p.iMin) / p.num_cells
    iMins = [0.0] * num_tiles
    iMaxs = [0.0] * num_tiles
    nc = [0] * num_tiles
    siL = [1] * num_tiles
    if num_tiles == 1:
        iMins[0], iMaxs[0], nc[0] = p.iMin, p.iMax, p.num_cells
    else:
        lpoints = [n for n, _, _ in ring_layout(p)]          # ring points, ring 1 .. iDim
        q, r_rem = divmod(p.jDim, num_tiles)
        tile_targets = [q + 1 if i <= r_rem else q for i in range(1, num_tiles + 1)]
        target = 1.0
        nc = [0] * num_tiles
        while any(c < 3 for c in nc) and target > 0.1:
            t = num_tiles                                    # 1-based tile being filled
            cell_count = 0
            target -= 0.1
            tile_count = [0] * num_tiles
            nc = [0] * num_tiles
            tile_min = int(math.floor(target * tile_targets[0]))
            for ri in range(p.iDim, 0, -1):
                tile_count[t - 1] += lpoints[ri - 1]
                if ri % 3 == 0:
                    cell_count += 1
                    if cell_count >= 3 and tile_count[t - 1] >= tile_min:
                        nc[t - 1] = cell_count
                        cell_count = 0
                        t -= 1
                if t == 0:
                    break
            nc[0] = p.num_cells - sum(nc[1:])
        if any(c < 3 for c in nc):
            raise DomainError(f"{num_tiles}: Too many tiles for this grid (need at least 3 cells per tile)")
        iMins[0] = p.iMin
        iMaxs[0] = nc[0] * DX + iMins[0]
        for i in range(1, num_tiles - 1):
            iMins[i] = iMaxs[i - 1]
            iMaxs[i] = nc[i] * DX + iMins[i]
            siL[i] = nc[i - 1] + siL[i - 1]
        iMins[-1] = iMaxs[-2]
        iMaxs[-1] = p.iMax
        siL[-1] = nc[-2] + siL[-2]
        nc[-1] = p.num_cells - sum(nc[:-1])
    out = []
    for i in range(num_tiles):
        # _create_tile_from_patch (src/tiling.jl:50-80): the patch's parameters on the
        # tile's interval; the derived dimensions (iDim, b_iDim, patchOffsetL,
        # spectralIndexR, ...) follow from them
        base = {f.name: getattr(p, f.name) for f in dataclasses.fields(p)}
        for k in ("iDim", "b_iDim", "jDim", "b_jDim", "b_kDim", "spectralIndexR", "patchOffsetL", "patchOffsetR",
                  "i_regular_out", "j_regular_out", "k_regular_out"):
            if k in base:
                base[k] = None
        base.update(iMin=iMins[i], iMax=iMaxs[i], num_cells=nc[i], spectralIndexL=siL[i], tile_num=i + 1)
        out.append(compute_derived_params(type(p)(**base)))
    return out


def calcTileSizes(patch: SpringsteelGrid, num_tiles: int):
    """calcTileSizes(patch, num_tiles): the tile grids of tileParameters(patch.params,
    num_tiles), created (with their device plans) in the patch's precision."""
    _ring_grid(patch)
    return [createGrid(gp, precision=patch.precision) for gp in tileParameters(patch.params, num_tiles)]


def _spec_host(g: SpringsteelGrid, a):
    _check_host(a, (g.nspec, g.resolved.nvars))


def splineTransform_(shared: np.ndarray, tile: SpringsteelGrid) -> np.ndarray:
    """splineTransform!(sharedSpectral, tile): SAtransform of every line of the B
    coefficients in `shared` into tile.spectral (A coefficients)."""
    _ring_grid(tile)
    _spec_host(tile, shared)
    tile.plan.spline_transform(shared.ctypes.data, tile.spectral.ctypes.data, capi.SSG_MEM_HOST)
    return tile.spectral


def tileTransform_(shared, tile: SpringsteelGrid, physical: np.ndarray, spectral: np.ndarray) -> np.ndarray:
    """tileTransform!(sharedSpectral, tile, physical, spectral): every physical slot of the
    tile from the A coefficients in `spectral` (no solve); `shared` is not read, as in the
    reference."""
    _ring_grid(tile)
    _spec_host(tile, spectral)
    _check_host(physical, (tile.npts, tile.resolved.nvars, tile.nslots))
    tile.plan.tile_transform(spectral.ctypes.data, physical.ctypes.data, capi.SSG_MEM_HOST)
    return physical


def splineTransform_device(tile: SpringsteelGrid, shared, spectral):
    """splineTransform! on device tensors (nvars, nspec), on torch's current stream."""
    _ring_grid(tile)
    _check_dev(shared, (tile.resolved.nvars, tile.nspec))
    _check_dev(spectral, (tile.resolved.nvars, tile.nspec))
    tile.plan.spline_transform(shared.data_ptr(), spectral.data_ptr(), capi.SSG_MEM_DEVICE, _current_stream(spectral))
    return spectral


def tileTransform_device(tile: SpringsteelGrid, spectral, physical):
    """tileTransform! on device tensors: spectral (nvars, nspec) A coefficients,
    physical (nslots, nvars, npts)."""
    _ring_grid(tile)
    _check_dev(spectral, (tile.resolved.nvars, tile.nspec))
    _check_dev(physical, (tile.nslots, tile.resolved.nvars, tile.npts))
    tile.plan.tile_transform(spectral.data_ptr(), physical.data_ptr(), capi.SSG_MEM_DEVICE, _current_stream(physical))
    return physical


def _embed(patch: SpringsteelGrid, tile: SpringsteelGrid, add: bool, patch_spec=None, tile_spec=None):
    _ring_grid(patch)
    ps = patch.spectral if patch_spec is None else patch_spec
    ts = tile.spectral if tile_spec is None else tile_spec
    if isinstance(ps, np.ndarray):
        _spec_host(patch, ps)
        _spec_host(tile, ts)
        capi.tile_embed(patch.plan, ps.ctypes.data, tile.plan, ts.ctypes.data, tile.params.spectralIndexL, add,
                        capi.SSG_MEM_HOST)
    else:
        _check_dev(ps, (patch.resolved.nvars, patch.nspec))
        _check_dev(ts, (tile.resolved.nvars, tile.nspec))
        capi.tile_embed(patch.plan, ps.data_ptr(), tile.plan, ts.data_ptr(), tile.params.spectralIndexL, add,
                        capi.SSG_MEM_DEVICE, _current_stream(ps))
    return ps


def sumSpectralTile_(patch: SpringsteelGrid, tile: SpringsteelGrid, patch_spectral=None, tile_spectral=None):
    """sumSpectralTile!(patch, tile): the tile's spectral lines added into the patch's at
    radial coefficients spectralIndexL ..; the grids' own arrays, or host arrays / device
    tensors passed in their place."""
    return _embed(patch, tile, True, patch_spectral, tile_spectral)


def setSpectralTile_(patch: SpringsteelGrid, tile: SpringsteelGrid, patch_spectral=None, tile_spectral=None):
    """setSpectralTile!(patch, tile): the patch's spectral zeroed, then the tile's lines
    written where sumSpectralTile! adds them."""
    return _embed(patch, tile, False, patch_spectral, tile_spectral)


def sumSharedSpectral(shared, tile: SpringsteelGrid, patchMap=None, haloMap=None, tile_spectral=None):
    """sumSharedSpectral(sharedSpectral, tile, patchMap, haloMap) for the ring grids (the
    reference's generic method, which does not read the maps): per variable column, rows
    spectralIndexL .. spectralIndexR - 3 of `shared` set to rows 1 .. b_iDim - 3 of the
    tile's spectral column.  `shared` is a Fortran-ordered host array (rows, nvars) or a
    device tensor (nvars, rows)."""
    _ring_grid(tile)
    ts = tile.spectral if tile_spectral is None else tile_spectral
    if isinstance(shared, np.ndarray):
        if not (shared.dtype == np.float64 and shared.flags.f_contiguous and shared.ndim == 2
                and shared.shape[1] == tile.resolved.nvars):
            raise ArgumentError("expected a Fortran-ordered float64 (rows, nvars) array")
        _spec_host(tile, ts)
        capi.sum_shared_spectral(shared.ctypes.data, shared.shape[0], tile.plan, ts.ctypes.data,
                                 tile.params.spectralIndexL, capi.SSG_MEM_HOST)
    else:
        if shared.dim() != 2 or shared.shape[0] != tile.resolved.nvars:
            raise ArgumentError("expected a device tensor (nvars, rows)")
        _check_dev(shared, tuple(shared.shape))
        _check_dev(ts, (tile.resolved.nvars, tile.nspec))
        capi.sum_shared_spectral(shared.data_ptr(), shared.shape[1], tile.plan, ts.data_ptr(),
                                 tile.params.spectralIndexL, capi.SSG_MEM_DEVICE, _current_stream(shared))
    return None
