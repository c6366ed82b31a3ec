"""Pins of the tile pipeline of src/tiling.jl without a GPU.

- calcTileSizes (src/tiling.jl:185-259 cylindrical, 265-347 spherical): the
  properties test/tiling.jl checks for it (count, at least 3 cells and 9 rings
  per tile, the ends, the cell sum, DomainError when a tile would get fewer than
  3 cells), the spectralIndexL progression, and two splits worked out by hand
  from the ring-point rule (the comments give the arithmetic);
- the tiles' geometry (cylindrical): each tile's gridpoints are the patch's
  gridpoints of the tile's rings, and the tile's B coefficients embedded with sumSpectralTile!
  and summed over the tiles are the patch's B coefficients (the projection is
  a sum over the mish points, which the tiles partition);
- sumSpectralTile! / setSpectralTile! (oracle.sum_spectral_tile and
  set_spectral_tile) against the block checks of test/tiling.jl:256-291 and
  395-423, and sumSharedSpectral's generic method against its definition on a
  counting array;
- splineTransform! + tileTransform! in the oracle against its own
  gridTransform (bit for bit: the same SA and SI of every line), and the
  reference's round trips (test/tiling.jl:159-205 RL, 297-333 RLZ, 428-472 SL)
  at their tolerances.
"""
import math

import numpy as np
import pytest

import oracle as O
import springsteel_hip as S
from springsteel_hip import CubicBSpline, Chebyshev
from springsteel_hip.params import compute_derived_params, resolve, ring_layout


class _G:
    """Parameters of a grid and its oracle grid (no device plan on the CPU)."""

    def __init__(self, gp):
        self.params = compute_derived_params(gp)
        self.geometry = self.params.geometry
        self.og = O.OracleGrid(resolve(self.params))
        self.nspec = self.og.nspec


def _tiles(patch, n):
    return [_G(p) for p in S.tileParameters(patch.params, n)]


def _rl(num_cells=12, iMax=100.0, **kw):
    return _G(S.SpringsteelGridParameters(
        geometry="RL", num_cells=num_cells, iMin=0.0, iMax=iMax, vars={"u": 1},
        BCL={"u": CubicBSpline.R0}, BCR={"u": CubicBSpline.R0}, **kw))


def _rlz(num_cells=9, iMax=75.0, kDim=6):
    return _G(S.SpringsteelGridParameters(
        geometry="RLZ", num_cells=num_cells, iMin=0.0, iMax=iMax, kMin=0.0, kMax=10.0, kDim=kDim,
        vars={"u": 1}, BCL={"u": CubicBSpline.R0}, BCR={"u": CubicBSpline.R0},
        BCB={"u": Chebyshev.R0}, BCT={"u": Chebyshev.R0}))


def _sl(num_cells=12):
    return _G(S.SpringsteelGridParameters(
        geometry="SL", num_cells=num_cells, iMin=0.0, iMax=math.pi, vars={"u": 1},
        BCL={"u": CubicBSpline.R0}, BCR={"u": CubicBSpline.R0}))


def _slz(num_cells=9, kDim=6):
    return _G(S.SpringsteelGridParameters(
        geometry="SLZ", num_cells=num_cells, iMin=0.0, iMax=math.pi, kMin=0.0, kMax=10.0, kDim=kDim,
        vars={"u": 1}, BCL={"u": CubicBSpline.R0}, BCR={"u": CubicBSpline.R0},
        BCB={"u": Chebyshev.R0}, BCT={"u": Chebyshev.R0}))


def _dims(g):
    p = g.params
    return (g.geometry, p.b_iDim, p.iDim + p.patchOffsetL, p.b_kDim)


@pytest.mark.parametrize("make,n,too_many", [(_rl, 4, 5), (_rlz, 3, 4), (_sl, 4, 5), (_slz, 3, 4)])
def test_calc_tile_sizes_properties(make, n, too_many):
    patch = make()
    tiles = _tiles(patch, n)
    assert len(tiles) == n
    for i, t in enumerate(tiles):
        assert t.params.num_cells >= 3 and t.params.iDim >= 9
        assert t.geometry == patch.geometry and t.params.tile_num == i + 1
        assert t.params.patchOffsetL == (t.params.spectralIndexL - 1) * 3
        assert t.params.spectralIndexR == t.params.spectralIndexL + t.params.b_iDim - 1
    assert tiles[0].params.iMin == pytest.approx(patch.params.iMin)
    assert tiles[-1].params.iMax == pytest.approx(patch.params.iMax)
    assert sum(t.params.num_cells for t in tiles) == patch.params.num_cells
    assert tiles[0].params.spectralIndexL == 1
    for a, b in zip(tiles, tiles[1:]):
        assert b.params.spectralIndexL == a.params.spectralIndexL + a.params.num_cells
        assert b.params.iMin == a.params.iMax
    with pytest.raises(S.DomainError):
        S.tileParameters(patch.params, too_many)
    one = _tiles(patch, 1)[0]
    assert (one.params.iMin, one.params.iMax, one.params.num_cells, one.params.spectralIndexL) == \
        (patch.params.iMin, patch.params.iMax, patch.params.num_cells, 1)


def test_calc_tile_sizes_hand_splits():
    # The rule walks the rings from the outermost inwards and counts a cell when it
    # reaches a ring index divisible by 3, so the first count comes after one ring
    # (ri = 3 n) and every later one after three more.
    # RL, 12 cells, 4 tiles: jDim = sum(4 + 4 ri, ri = 1..36) = 2808, 702 points per
    # tile.  target 0.9 (min 631): tile 4 holds 952 points at its 3rd count (rings
    # 36..30), tile 3 936 (29..21), tile 2 612 < 631 at its 3rd (20..12) and takes a
    # 4th, leaving 2 cells for tile 1; target 0.8 (min 561) gives 3, 3, 3 and 3.
    t = _tiles(_rl(), 4)
    assert [x.params.num_cells for x in t] == [3, 3, 3, 3]
    assert [x.params.spectralIndexL for x in t] == [1, 4, 7, 10]
    # RL, 20 cells, 3 tiles: jDim = 7560, 2520 per tile, target 0.9 (min 2268):
    # tile 3 holds 244 + 708 + 672 + 636 = 2260 < 2268 at its 4th count (rings
    # 60..51), 2860 at its 5th -> 5 cells; tile 2 564 + 528 + 492 + 456 + 420 = 2460
    # at its 5th (rings 47..33) -> 5 cells; tile 1 keeps 10.
    patch = _rl(num_cells=20, iMax=200.0)
    t = _tiles(patch, 3)
    assert [x.params.num_cells for x in t] == [10, 5, 5]
    assert [x.params.spectralIndexL for x in t] == [1, 11, 16]
    assert [x.params.iMin for x in t] == pytest.approx([0.0, 100.0, 150.0])
    assert [x.params.iMax for x in t] == pytest.approx([100.0, 150.0, 200.0])
    assert [x.params.patchOffsetL for x in t] == [0, 30, 45]


def _ring_rows(patch, tile):
    """Physical rows of the patch that belong to the tile's rings."""
    lp = [n for n, _, _ in ring_layout(patch.params)]
    kz = patch.params.kDim if patch.geometry in ("RLZ", "SLZ") else 1
    r0 = tile.params.patchOffsetL
    a = sum(lp[:r0]) * kz
    return a, a + sum(lp[r0:r0 + tile.params.iDim]) * kz


# Cylindrical tiles only: a spherical tile sizes its rings from sin(theta) times its
# own iDim + patchOffsetL (src/factory.jl:159-163), which is not the patch's, so its
# rings are not the patch's rings.
@pytest.mark.parametrize("make,n", [(_rl, 4), (_rlz, 3), (lambda: _rl(num_cells=20, iMax=200.0), 3)])
def test_tiles_partition_the_patch(make, n):
    patch = make()
    tiles = _tiles(patch, n)
    og = patch.og
    pts = og.gridpoints()
    rng = np.random.default_rng(7)
    phys = np.asfortranarray(rng.standard_normal((og.npts, 1, og.nslots)))
    b_patch = og.spectral_transform(phys)
    b_sum = np.zeros_like(b_patch, order="F")
    for t in tiles:
        ot = t.og
        a, b = _ring_rows(patch, t)
        assert b - a == ot.npts
        np.testing.assert_allclose(ot.gridpoints(), pts[a:b], rtol=0, atol=1e-12)
        pt = np.asfortranarray(phys[a:b])
        b_t = ot.spectral_transform(pt)
        O.sum_spectral_tile(b_sum, _dims(patch), b_t, _dims(t), t.params.spectralIndexL)
    scale = np.abs(b_patch).max()
    np.testing.assert_allclose(b_sum, b_patch, rtol=0, atol=1e-12 * scale)


def test_sum_set_blocks_rlz():
    # test/tiling.jl:256-291
    patch = _rlz()
    t = _tiles(patch, 3)[0]
    ps = np.zeros((patch.nspec, 1), order="F")
    O.sum_spectral_tile(ps, _dims(patch), np.ones((t.nspec, 1), order="F"), _dims(t), t.params.spectralIndexL)
    siL, bt, bp = t.params.spectralIndexL, t.params.b_iDim, patch.params.b_iDim
    wn_p = bp * (1 + 2 * (patch.params.iDim + patch.params.patchOffsetL))
    for r in (siL, bp + siL, 2 * bp + siL, wn_p + siL):
        assert np.all(ps[r - 1:r - 1 + bt] == 1.0)
    ps[:] = 99.0
    O.set_spectral_tile(ps, _dims(patch), np.full((t.nspec, 1), 2.0, order="F"), _dims(t), siL)
    assert np.all(ps[siL - 1:siL - 1 + bt] == 2.0) and np.all(ps[wn_p + siL - 1:wn_p + siL - 1 + bt] == 2.0)
    assert np.all(ps[siL + bt - 1:bp] == 0.0)
    # every patch row is written by at most one block, and by exactly as many
    # blocks as the tile has lines
    cover = np.zeros(patch.nspec, int)
    for pr, tr, m in O._tile_blocks(_dims(patch), _dims(t), siL):
        cover[pr - 1:pr - 1 + m] += 1
    assert cover.max() == 1 and cover.sum() == t.nspec


def test_sum_set_blocks_sl():
    # test/tiling.jl:395-423, on the second tile (siL > 1)
    patch = _sl()
    t = _tiles(patch, 4)[1]
    siL, siR = t.params.spectralIndexL, t.params.spectralIndexR
    ps = np.zeros((patch.nspec, 1), order="F")
    O.sum_spectral_tile(ps, _dims(patch), np.ones((t.nspec, 1), order="F"), _dims(t), siL)
    assert np.all(ps[siL - 1:siR] == 1.0)
    bp = patch.params.b_iDim
    assert np.all(ps[bp + siL - 1:bp + siR] == 1.0)
    assert np.all(ps[:siL - 1] == 0.0)
    ps[:] = 99.0
    O.set_spectral_tile(ps, _dims(patch), np.full((t.nspec, 1), 2.0, order="F"), _dims(t), siL)
    assert np.all(ps[siL - 1:siR] == 2.0)
    assert np.all(ps[:siL - 1] == 0.0) and np.all(ps[siR:bp] == 0.0)
    # the tile's highest wavenumber lines land at lines 2 K_t - 1, 2 K_t of the patch
    Kt = t.params.iDim + t.params.patchOffsetL
    ts = np.asfortranarray(np.arange(1, t.nspec + 1, dtype=float)[:, None])
    O.set_spectral_tile(ps, _dims(patch), ts, _dims(t), siL)
    bt = t.params.b_iDim
    for line in (2 * Kt - 1, 2 * Kt):
        np.testing.assert_array_equal(ps[line * bp + siL - 1:line * bp + siL - 1 + bt, 0], ts[line * bt:(line + 1) * bt, 0])


def test_sum_shared_spectral_generic():
    # src/tiling.jl:1379-1389: rows siL .. siR - 3 set from rows 1 .. b_iDim - 3
    patch = _rl()
    t = _tiles(patch, 4)[2]
    shared = np.full((patch.nspec, 2), -1.0, order="F")
    ts = np.asfortranarray(np.arange(2 * t.nspec, dtype=float).reshape(t.nspec, 2, order="F"))
    O.sum_shared_spectral(shared, ts, t.params.b_iDim, t.params.spectralIndexL)
    siL, n = t.params.spectralIndexL, t.params.b_iDim - 3
    np.testing.assert_array_equal(shared[siL - 1:siL - 1 + n], ts[:n])
    assert np.all(shared[:siL - 1] == -1.0) and np.all(shared[siL - 1 + n:] == -1.0)


@pytest.mark.parametrize("make", [_rl, _rlz, _sl, _slz])
def test_oracle_spline_tile_is_grid_transform(make):
    patch = make()
    for g in (patch, _tiles(patch, 3)[1]):
        og = g.og
        rng = np.random.default_rng(11)
        b = np.asfortranarray(rng.standard_normal((og.nspec, 1)))
        ref = og.grid_transform(b)
        a = og.spline_transform(b)
        got = og.tile_transform(a)
        np.testing.assert_array_equal(got, ref)
        # the A coefficients are the SA solve of every line: re-projecting the
        # spline they define reproduces b
        assert not np.array_equal(a, b)


def _analytic(make, fn, n_slots_check):
    og = make().og
    pts = og.gridpoints()
    phys = np.zeros((og.npts, 1, og.nslots), order="F")
    phys[:, 0, 0] = fn(pts)
    shared = og.spectral_transform(phys)
    a = og.spline_transform(shared)
    return pts, og.tile_transform(a)


def test_round_trip_rl():
    # test/tiling.jl:159-205
    pts, out = _analytic(_rl, lambda p: np.sin(np.pi * p[:, 0] / 100.0), 5)
    assert np.abs(out[:, 0, 0] - np.sin(np.pi * pts[:, 0] / 100.0)).max() < 5e-4
    assert np.abs(out[:, 0, 1] - np.pi / 100.0 * np.cos(np.pi * pts[:, 0] / 100.0)).max() < 0.05
    assert np.abs(out[:, 0, 3]).max() < 1e-10
    assert np.all(np.isfinite(out))


def test_round_trip_rlz():
    # test/tiling.jl:297-333
    f = lambda p: np.sin(np.pi * p[:, 0] / 75.0) * (p[:, 2] / 10.0)
    pts, out = _analytic(_rlz, f, 7)
    assert np.abs(out[:, 0, 0] - f(pts)).max() < 1e-3
    assert np.abs(out[:, 0, 3]).max() < 1e-8
    assert np.all(np.isfinite(out))


def test_round_trip_sl():
    # test/tiling.jl:428-472
    pts, out = _analytic(_sl, lambda p: np.sin(p[:, 0]), 5)
    assert np.abs(out[:, 0, 0] - np.sin(pts[:, 0])).max() < 5e-4
    assert np.abs(out[:, 0, 1] - np.cos(pts[:, 0])).max() < 0.05
    assert np.abs(out[:, 0, 3]).max() < 1e-10
    assert np.all(np.isfinite(out))


def test_tile_symbols_exported():
    from springsteel_hip import capi
    for name in ("ssg_spline_transform", "ssg_tile_transform", "ssg_sum_spectral_tile",
                 "ssg_set_spectral_tile", "ssg_sum_shared_spectral"):
        assert name in capi.EXPORTED
        assert hasattr(capi.load(), name)


def test_tile_argument_checks():
    rz = compute_derived_params(S.SpringsteelGridParameters(geometry="RZ", num_cells=12, iMin=0.0, iMax=1.0,
                                                            kMin=0.0, kMax=1.0, kDim=5, vars={"u": 1}))
    with pytest.raises(S.ArgumentError):
        S.tileParameters(rz, 2)
