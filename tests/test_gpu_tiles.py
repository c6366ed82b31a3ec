"""GPU: the tile pipeline of src/tiling.jl (splineTransform!, tileTransform!,
sumSpectralTile!, setSpectralTile!, sumSharedSpectral) through the C-ABI.

- splineTransform! (B -> A on every line) against the oracle's at the FP64
  spectral bar, and tileTransform! (A -> every physical slot, no solve) against
  the oracle's gridTransform of the same B, which the oracle's spline + tile
  forms reproduce bit for bit (test_tile_pins.py), at the parity tests' slot
  bars; on the four ring geometries, tiles with patchOffsetL > 0, several
  variables and boundary conditions, R3X background coefficients, and a grid of
  b_iDim = 203 (7 partitions of the fused solve);
- the split form against the library's own gridTransform of the same B: bit for
  bit while the fused solve is one partition (b_iDim <= 32); with several, the
  fused kernel's SI takes the next partition's first coefficients from the
  composed backward maps and the split form from that partition's own rows, so
  the two agree to rounding, at the same bars;
- the device-tensor forms bit for bit against the host forms;
- the tile sums against the numpy restatement in the oracle (bit for bit: the
  same additions in the same order), and the distributed projection: every
  tile's HIP spectralTransform of its rings summed into the patch equals the
  patch's HIP spectralTransform to 1e-12 of its scale;
- argument errors: FP32 plans, rows outside the patch, mismatched grids.
"""
import math

import numpy as np
import pytest

import oracle as O
import springsteel_hip as S
from springsteel_hip.errors import ArgumentError, BoundsError, NotSupportedError
from springsteel_hip.params import ring_layout
from parity_util import TOL_FP64, errors, run_both, seeded_field, sensitivity, slot_tolerances

pytestmark = pytest.mark.gpu

CB = S.CubicBSpline


def _gp(geometry, **kw):
    base = dict(geometry=geometry, num_cells=12, iMin=0.0, iMax=100.0, vars={"u": 1})
    if geometry in ("SL", "SLZ"):
        base.update(iMax=math.pi)
    if geometry in ("RLZ", "SLZ"):
        base.update(num_cells=9, kMin=0.0, kMax=10.0, kDim=6)
    base.update(kw)
    return S.SpringsteelGridParameters(**base)


def _tile_params(gp, n, i):
    return S.tileParameters(S.compute_derived_params(gp), n)[i]


CASES = {
    "rl": _gp("RL"),
    "rlz": _gp("RLZ"),
    "sl": _gp("SL"),
    "slz": _gp("SLZ"),
    "rl_tile2of3": _tile_params(_gp("RL", num_cells=20, iMax=200.0), 3, 1),
    "rlz_tile3of3": _tile_params(_gp("RLZ", num_cells=12, kDim=10), 3, 2),
    "slz_tile2of3": _tile_params(_gp("SLZ", num_cells=12, kDim=8), 3, 1),
    "rl_bcs_3vars": _gp("RL", num_cells=16, vars={"u": 1, "v": 2, "w": 3},
                        BCL={"default": CB.R0, "v": CB.R1T0, "w": CB.R1T1},
                        BCR={"default": CB.R0, "v": CB.R1T1, "w": CB.R1T0}),
    "rlz_b203": _gp("RLZ", num_cells=200, iMax=2000.0, kDim=8, vars={"u": 1, "v": 2}),
}


def _split(grid, shared):
    a = S.splineTransform_(shared, grid).copy(order="F")
    phys = np.zeros_like(grid.physical, order="F")
    S.tileTransform_(shared, grid, phys, a)
    return a, phys


@pytest.mark.parametrize("name", list(CASES))
def test_spline_tile_vs_oracle(name):
    gp = CASES[name]
    grid, og, spec, ospec, phys, ophys, phys2, inp = run_both(gp, None, 5)
    shared = np.asfortranarray(ospec)
    print(f"\n[{name}] b_iDim={grid.params.b_iDim} patchOffsetL={grid.params.patchOffsetL}", flush=True)
    a, phys_t = _split(grid, shared)
    oa = og.spline_transform(shared, threads=8)
    ea = errors(a, oa)
    assert max(ea) <= TOL_FP64, ea
    ep = errors(phys_t, ophys, grid.geometry, grid.params)
    tols = slot_tolerances(sensitivity(og, inp, ophys, grid.geometry, grid.params), TOL_FP64)
    es = errors(phys_t, phys2, grid.geometry, grid.params)
    bitwise = np.array_equal(phys_t, phys2)
    print(f"A {max(ea):.1e}; slots {[f'{max(e[s] for e in ep):.1e}' for s in range(phys_t.shape[2])]}; "
          f"vs gridTransform {max(max(e) for e in es):.1e} (bitwise {bitwise})")
    for v in range(len(ep)):
        for s in range(phys_t.shape[2]):
            assert ep[v][s] <= tols[v][s], (v, s, ep[v], tols[v])
            assert es[v][s] <= tols[v][s], (v, s, es[v], tols[v])
    if grid.params.b_iDim <= 32:
        assert bitwise
    grid.close()


@pytest.mark.parametrize("geometry", ["RL", "RLZ"])
def test_spline_tile_r3x_vs_oracle(geometry):
    gp = _gp(geometry, num_cells=10, iMax=50.0, vars={"u": 1, "w": 2},
             BCL={"default": CB.R0, "u": CB.R3X}, BCR={"default": CB.R0, "u": CB.R3X})
    g = S.createGrid(gp)
    og = O.OracleGrid(g.resolved)
    g.physical[:, :, 0] = seeded_field(S.getGridpoints(g), g.resolved.nvars, 3, g.geometry)
    inp = np.asfortranarray(g.physical.copy())
    S.spectralTransform_(g)
    shared = np.asfortranarray(g.spectral.copy())
    ahat = np.asfortranarray(np.random.default_rng(4).normal(size=shared.shape) * 0.3)
    S.set_ahat(g, ahat)
    og.set_ahat(ahat)
    print(f"\n[r3x {geometry}]", flush=True)
    a, phys_t = _split(g, shared)
    oa = og.spline_transform(shared, threads=8)
    assert max(errors(a, oa)) <= TOL_FP64
    ref = og.grid_transform(shared, threads=8)
    tols = slot_tolerances(sensitivity(og, inp, ref, g.geometry, g.params), TOL_FP64)
    ep = errors(phys_t, ref, g.geometry, g.params)
    print([f"{max(e[s] for e in ep):.1e}" for s in range(phys_t.shape[2])])
    for v in range(len(ep)):
        for s in range(phys_t.shape[2]):
            assert ep[v][s] <= tols[v][s], (v, s, ep[v], tols[v])
    g.close()


def test_round_trip_reference_tolerances():
    # test/tiling.jl:159-205 (RL) and 297-333 (RLZ) through the library
    g = S.createGrid(_gp("RL", BCL={"u": CB.R0}, BCR={"u": CB.R0}))
    pts = S.getGridpoints(g)
    g.physical[:] = 0.0
    g.physical[:, 0, 0] = np.sin(np.pi * pts[:, 0] / 100.0)
    S.spectralTransform_(g)
    shared = np.asfortranarray(g.spectral.copy())
    S.splineTransform_(shared, g)
    out = np.zeros_like(g.physical, order="F")
    S.tileTransform_(shared, g, out, g.spectral)
    assert np.abs(out[:, 0, 0] - np.sin(np.pi * pts[:, 0] / 100.0)).max() < 5e-4
    assert np.abs(out[:, 0, 1] - np.pi / 100.0 * np.cos(np.pi * pts[:, 0] / 100.0)).max() < 0.05
    assert np.abs(out[:, 0, 3]).max() < 1e-10 and np.all(np.isfinite(out))
    g.close()
    g = S.createGrid(_gp("RLZ", iMax=75.0))
    pts = S.getGridpoints(g)
    f = np.sin(np.pi * pts[:, 0] / 75.0) * (pts[:, 2] / 10.0)
    g.physical[:] = 0.0
    g.physical[:, 0, 0] = f
    S.spectralTransform_(g)
    shared = np.asfortranarray(g.spectral.copy())
    S.splineTransform_(shared, g)
    out = np.zeros_like(g.physical, order="F")
    S.tileTransform_(shared, g, out, g.spectral)
    assert np.abs(out[:, 0, 0] - f).max() < 1e-3
    assert np.abs(out[:, 0, 3]).max() < 1e-8 and np.all(np.isfinite(out))
    g.close()


def test_device_forms_bitwise_host():
    torch = pytest.importorskip("torch")
    g = S.createGrid(CASES["rlz_tile3of3"])
    nv = g.resolved.nvars
    rng = np.random.default_rng(9)
    shared = np.asfortranarray(rng.normal(size=(g.nspec, nv)))
    a_host, phys_host = _split(g, shared)
    sd = torch.from_numpy(np.ascontiguousarray(shared.T)).cuda()
    ad = torch.zeros((nv, g.nspec), dtype=torch.float64, device="cuda")
    pd = torch.zeros((g.nslots, nv, g.npts), dtype=torch.float64, device="cuda")
    S.splineTransform_device(g, sd, ad)
    S.tileTransform_device(g, ad, pd)
    torch.cuda.synchronize()
    assert np.array_equal(ad.cpu().numpy().T, a_host)
    assert np.array_equal(np.transpose(pd.cpu().numpy(), (2, 1, 0)), phys_host)
    g.close()


def _ring_rows(patch, tile):
    lp = [n for n, _, _ in ring_layout(patch.params)]
    kz = patch.params.kDim if patch.geometry in ("RLZ", "SLZ") else 1
    r0 = tile.params.patchOffsetL
    a = sum(lp[:r0]) * kz
    return a, a + sum(lp[r0:r0 + tile.params.iDim]) * kz


def _dims(g):
    p = g.params
    return (g.geometry, p.b_iDim, p.iDim + p.patchOffsetL, p.b_kDim)


@pytest.mark.parametrize("gp,n", [(_gp("RL", num_cells=20, iMax=200.0, vars={"u": 1, "v": 2}), 3),
                                  (_gp("RLZ", num_cells=12, kDim=10), 3)])
def test_distributed_projection_and_sums(gp, n):
    torch = pytest.importorskip("torch")
    patch = S.createGrid(gp)
    tiles = S.calcTileSizes(patch, n)
    nv = patch.resolved.nvars
    patch.physical[:, :, 0] = seeded_field(S.getGridpoints(patch), nv, 6, patch.geometry)
    S.spectralTransform(patch, patch.physical, patch.spectral)
    b_patch = patch.spectral.copy(order="F")
    host_sum = np.zeros_like(b_patch, order="F")
    ref_sum = np.zeros_like(b_patch, order="F")
    dev_sum = torch.zeros((nv, patch.nspec), dtype=torch.float64, device="cuda")
    tile_b = []
    for t in tiles:
        a, b = _ring_rows(patch, t)
        assert b - a == t.npts
        assert np.allclose(S.getGridpoints(t), S.getGridpoints(patch)[a:b], rtol=0, atol=1e-12)
        ph = np.asfortranarray(patch.physical[a:b])
        S.spectralTransform(t, ph, t.spectral)
        tile_b.append(t.spectral.copy(order="F"))
        S.sumSpectralTile_(patch, t, host_sum, t.spectral)
        S.sumSpectralTile_(patch, t, dev_sum, torch.from_numpy(np.ascontiguousarray(t.spectral.T)).cuda())
        O.sum_spectral_tile(ref_sum, _dims(patch), t.spectral, _dims(t), t.params.spectralIndexL)
    torch.cuda.synchronize()
    assert np.array_equal(host_sum, ref_sum)
    assert np.array_equal(dev_sum.cpu().numpy().T, ref_sum)
    scale = np.abs(b_patch).max()
    print(f"\n{gp.geometry}: tiles {[t.params.num_cells for t in tiles]}, "
          f"sum vs patch {np.abs(host_sum - b_patch).max() / scale:.1e}")
    assert np.abs(host_sum - b_patch).max() <= 1e-12 * scale
    # setSpectralTile! zeroes the patch array, then copies the tile's lines
    for t, tb in zip(tiles, tile_b):
        ref = np.full_like(b_patch, 7.0, order="F")
        O.set_spectral_tile(ref, _dims(patch), tb, _dims(t), t.params.spectralIndexL)
        h = np.full_like(b_patch, 7.0, order="F")
        S.setSpectralTile_(patch, t, h, tb)
        d = torch.full((nv, patch.nspec), 7.0, dtype=torch.float64, device="cuda")
        S.setSpectralTile_(patch, t, d, torch.from_numpy(np.ascontiguousarray(tb.T)).cuda())
        torch.cuda.synchronize()
        assert np.array_equal(h, ref) and np.array_equal(d.cpu().numpy().T, ref)
    # sumSharedSpectral: rows spectralIndexL .. spectralIndexR - 3 of a flat shared array
    shared = np.full((patch.params.b_iDim + 8, nv), -1.0, order="F")
    ref = shared.copy(order="F")
    sd = torch.from_numpy(np.ascontiguousarray(shared.T)).cuda()
    for t, tb in zip(tiles, tile_b):
        S.sumSharedSpectral(shared, t, tile_spectral=tb)
        S.sumSharedSpectral(sd, t, tile_spectral=torch.from_numpy(np.ascontiguousarray(tb.T)).cuda())
        O.sum_shared_spectral(ref, tb, t.params.b_iDim, t.params.spectralIndexL)
    torch.cuda.synchronize()
    assert np.array_equal(shared, ref) and np.array_equal(sd.cpu().numpy().T, ref)
    for t in tiles:
        t.close()
    patch.close()


def test_tile_argument_errors():
    g32 = S.createGrid(_gp("RL"), precision=S.SSG_FP32_REFINE)
    with pytest.raises(ArgumentError):
        S.splineTransform_(np.zeros((g32.nspec, 1), order="F"), g32)
    g32.close()
    patch = S.createGrid(_gp("RL", num_cells=20, iMax=200.0))
    tiles = S.calcTileSizes(patch, 3)
    # a tile whose rows would run past the patch's
    with pytest.raises(BoundsError):
        S.capi.tile_embed(patch.plan, patch.spectral.ctypes.data, tiles[2].plan, tiles[2].spectral.ctypes.data,
                          patch.params.b_iDim, True, S.SSG_MEM_HOST)
    # the patch summed into a tile (more lines than the tile holds)
    with pytest.raises(BoundsError):
        S.sumSpectralTile_(tiles[0], patch)
    rlz = S.createGrid(_gp("RLZ"))
    with pytest.raises(NotSupportedError):
        S.sumSpectralTile_(rlz, S.calcTileSizes(patch, 2)[0])
    with pytest.raises(BoundsError):
        S.sumSharedSpectral(np.zeros((5, 1), order="F"), tiles[1])
    for t in tiles:
        t.close()
    patch.close()
    rlz.close()
