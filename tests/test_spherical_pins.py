"""Pins of the oracle's SL / SLZ restatement (no GPU needed).

The spherical grids run the cylindrical RL / RLZ transforms on rings whose size
follows sin(theta) (src/transforms_spherical.jl:115-400, 458-800; rings
src/factory.jl:159-163, 626-643).  Pinned here by:
- the ring layout against the reference's formulas, and the derived jDim /
  b_jDim (_fourier_j_dims_sph, src/factory.jl:193-203);
- the composed transforms against the independent matrix-form restatement in
  npref.py with spherical rings (the cap applied to kmax and to bDim, the spline
  stages limited to the ring's kmax, ymin = dl r / 2);
- the reference's own testset "Spherical Transforms" (test/transforms.jl:
  1954-2075): round trips of sin(theta) cos(lambda) (and times z for SLZ) and the
  theta derivative of cos(theta) z, at its tolerances, and the grid creation
  checks of test/grids.jl:1517-1568.
"""
import math

import numpy as np
import pytest

import npref
import oracle as O
import springsteel_hip as S
from springsteel_hip.params import compute_derived_params, mish_points, resolve, ring_layout

RTOL = 1e-12


def rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300)


def _grid(**kw):
    gp = compute_derived_params(S.SpringsteelGridParameters(**kw))
    return gp, O.OracleGrid(resolve(gp))


def test_ring_layout_and_dims():
    gp, og = _grid(geometry="SL", num_cells=4, iMin=0.05, iMax=math.pi - 0.05, vars={"u": 1})
    th = mish_points(gp)
    rings = ring_layout(gp)
    max_ri = gp.iDim
    for r, ((n, kr, ymin), t) in enumerate(zip(rings, th), start=1):
        assert kr == max(1, round(math.sin(t) * max_ri)) and n == 4 * kr
        assert abs(ymin - 0.5 * (2 * math.pi / n) * r) < 1e-15
    assert gp.jDim == sum(n for n, _, _ in rings) and gp.b_jDim == sum(1 + 2 * kr for _, kr, _ in rings)
    assert og.npts == gp.jDim and og.nspec == gp.b_iDim * (1 + 2 * gp.iDim)
    pts = og.gridpoints()
    assert pts.shape == (gp.jDim, 2)
    # theta of each ring, lambda = ymin + 2 pi (l - 1) / N
    q = 0
    for (n, kr, ymin), t in zip(rings, th):
        assert np.allclose(pts[q:q + n, 0], t, rtol=0, atol=0)
        assert np.allclose(pts[q:q + n, 1], ymin + 2 * math.pi * np.arange(n) / n, rtol=0, atol=1e-15)
        q += n
    gp, og = _grid(geometry="SLZ", num_cells=3, iMin=0.05, iMax=math.pi - 0.05, kMin=0.0, kMax=10.0, kDim=6,
                   vars={"u": 1})
    assert og.npts == gp.kDim * gp.jDim and og.nslots == 7
    assert og.gridpoints().shape == (gp.kDim * gp.jDim, 3)


def _sph_rings(gp, cap=-1):
    return [npref.sph_ring(t, gp.iDim + gp.patchOffsetL, r, cap) for r, t in enumerate(mish_points(gp), start=1)]


@pytest.mark.parametrize("nc,cap", [(6, -1), (8, 5)])
def test_sl_composed_vs_npref(nc, cap):
    gp, og = _grid(geometry="SL", num_cells=nc, iMin=0.1, iMax=math.pi - 0.1, max_wavenumber={"default": cap},
                   vars={"u": 1})
    pts = og.gridpoints()
    rng = np.random.default_rng(nc)
    phys = og.new_physical()
    th, lam = pts[:, 0], pts[:, 1]
    phys[:, 0, 0] = np.sin(th) * (1 + 0.5 * np.cos(lam) + rng.normal() * np.sin(3 * lam)) + 0.2 * np.cos(2 * th)
    spec = og.spectral_transform(phys)
    sp = npref.NpSpline(0.1, math.pi - 0.1, nc)
    rings = _sph_rings(gp, cap)
    offs = np.cumsum([0] + [rg.N for rg in rings])
    u = [phys[offs[i]:offs[i + 1], 0, 0] for i in range(len(rings))]
    K = gp.iDim + gp.patchOffsetL
    assert rel(spec[:, 0], npref.rl_forward(sp, rings, u, [cap] * len(rings), K=K)) < RTOL
    out = og.grid_transform(spec)
    pref = npref.rl_inverse(sp, rings, spec[:, 0], [cap] * len(rings), K=K)
    for s in range(5):
        assert rel(out[:, 0, s], pref[s]) < 1e-11, s


def test_slz_composed_vs_npref():
    gp, og = _grid(geometry="SLZ", num_cells=4, iMin=0.1, iMax=math.pi - 0.1, kMin=0.0, kMax=10.0, kDim=8,
                   vars={"u": 1})
    pts = og.gridpoints()
    phys = og.new_physical()
    th, lam, z = pts[:, 0], pts[:, 1], pts[:, 2]
    phys[:, 0, 0] = (np.sin(th) + 0.3 * np.sin(th) * np.cos(2 * lam)) * (1 + z / 10.0) + 0.02 * z ** 2
    spec = og.spectral_transform(phys)
    sp = npref.NpSpline(0.1, math.pi - 0.1, 4)
    ch = npref.NpCheb(0.0, 10.0, 8, gp.b_kDim)
    rings = _sph_rings(gp)
    K = gp.iDim + gp.patchOffsetL
    assert rel(spec[:, 0], npref.rlz_forward(sp, ch, rings, phys[:, 0, 0], K=K)) < RTOL
    out = og.grid_transform(spec)
    pref = npref.rlz_inverse(sp, ch, rings, spec[:, 0], K=K)
    for s in range(7):
        assert rel(out[:, 0, s], pref[s]) < 1e-11, s


def test_reference_spherical_testset():
    """test/transforms.jl:1954-2075 at the reference's tolerances."""
    gp, og = _grid(geometry="SL", num_cells=10, iMin=0.0, iMax=math.pi, vars={"u": 1})
    pts = og.gridpoints()
    assert pts.shape == (gp.jDim, 2)
    phys = og.new_physical()
    phys[:, 0, 0] = np.sin(pts[:, 0]) * np.cos(pts[:, 1])
    out = og.grid_transform(og.spectral_transform(phys))
    assert np.max(np.abs(out[:, 0, 0] - phys[:, 0, 0])) < 1e-2

    gp, og = _grid(geometry="SL", num_cells=8, iMin=0.0, iMax=math.pi, vars={"u": 1})
    pts = og.gridpoints()
    phys = og.new_physical()
    phys[:, 0, 0] = np.sin(pts[:, 0]) * np.cos(pts[:, 1])
    out = og.grid_transform(og.spectral_transform(phys))
    for t in (pts[:, 0].min(), pts[:, 0].max()):
        near = np.abs(pts[:, 0] - t) < 0.01
        assert np.max(np.abs(out[near, 0, 0])) < 0.5

    gp, og = _grid(geometry="SLZ", num_cells=10, kDim=8, iMin=0.0, iMax=math.pi, kMin=0.0, kMax=20.0, vars={"u": 1})
    pts = og.gridpoints()
    assert pts.shape[1] == 3
    phys = og.new_physical()
    phys[:, 0, 0] = np.sin(pts[:, 0]) * np.cos(pts[:, 1]) * pts[:, 2]
    out = og.grid_transform(og.spectral_transform(phys))
    assert np.max(np.abs(out[:, 0, 0] - phys[:, 0, 0])) < 1e-2

    gp, og = _grid(geometry="SLZ", num_cells=10, kDim=6, iMin=0.0, iMax=math.pi, kMin=1.0, kMax=10.0, vars={"u": 1})
    pts = og.gridpoints()
    phys = og.new_physical()
    phys[:, 0, 0] = np.cos(pts[:, 0]) * pts[:, 2]
    out = og.grid_transform(og.spectral_transform(phys))
    assert np.max(np.abs(out[:, 0, 1] - (-np.sin(pts[:, 0]) * pts[:, 2]))) < 0.1


def test_rings_capped_per_variable():
    """SL rings are built per variable: a cap changes kmax and bDim of that variable's
    rings only; values above the cap are dropped."""
    gp, og = _grid(geometry="SL", num_cells=6, iMin=0.1, iMax=math.pi - 0.1, vars={"u": 1, "v": 2},
                   max_wavenumber={"default": -1, "v": 2})
    pts = og.gridpoints()
    phys = og.new_physical()
    f = np.sin(pts[:, 0]) * (1 + np.cos(pts[:, 1]) + np.cos(4 * pts[:, 1]))
    phys[:, 0, 0] = f
    phys[:, 1, 0] = f
    spec = og.spectral_transform(phys)
    M = gp.b_iDim
    # wavenumbers 3.. of v are zero; u carries wavenumber 4
    assert np.all(spec[5 * M:, 1] == 0.0)
    assert np.linalg.norm(spec[7 * M:9 * M, 0]) > 1e-3


def _reg_axes(gp):
    from springsteel_hip.grid import SpringsteelGrid, _regular_axes
    return _regular_axes(SpringsteelGrid(gp))


@pytest.mark.parametrize("sph,cyl", [("SL", "RL"), ("SLZ", "RLZ")])
def test_evaluation_restates_the_cylindrical_one(sph, cyl):
    """The reference's SL / SLZ regularGridTransform (src/transforms_spherical.jl:849-1131)
    restates the RL / RLZ one line for line, and evaluate_unstructured sends SL / SLZ to
    _eval_unstructured_rl / _rlz (src/interpolation.jl:1782-1800): no ring table enters.
    So on the same spectral array the spherical grid evaluates to the bytes of the
    cylindrical grid with the same spline (domain, cells, BCs)."""
    kw = dict(num_cells=6, iMin=0.1, iMax=math.pi - 0.1, vars={"u": 1, "v": 2},
              BCL={"default": S.CubicBSpline.R0, "v": S.CubicBSpline.R1T0})
    if sph == "SLZ":
        kw.update(kMin=0.0, kMax=10.0, kDim=8)
    gps, ogs = _grid(geometry=sph, **kw)
    gpc, ogc = _grid(geometry=cyl, **kw)
    assert ogs.nspec == ogc.nspec
    spec = np.asfortranarray(np.random.default_rng(7).normal(size=(ogs.nspec, 2)))
    axes = _reg_axes(gps)
    a = ogs.regular_transform(spec, *axes)
    b = ogc.regular_transform(spec, *axes)
    assert np.array_equal(a, b) and np.all(np.isfinite(a))
    rng = np.random.default_rng(8)
    pts = [rng.uniform(0.0, math.pi, 500), rng.uniform(0.0, 7.0, 500)]
    if sph == "SLZ":
        pts.append(rng.uniform(0.0, 10.0, 500))
    pts = np.column_stack(pts)
    u = ogs.evaluate_unstructured(spec, pts)
    assert np.array_equal(u, ogc.evaluate_unstructured(spec, pts), equal_nan=True)
    assert np.isnan(u[:, 0]).any() and np.isfinite(u[:, 0]).any()      # theta outside [0.1, pi - 0.1]


def test_reference_sl_regular_testsets():
    """test/transforms.jl:1707-1784: sin(theta) and sin(theta) cos(lambda) on the default
    regular points, at the reference's tolerances."""
    gp, og = _grid(geometry="SL", num_cells=10, iMin=0.0, iMax=math.pi, vars={"u": 1})
    pts = og.gridpoints()
    th, lam = _reg_axes(gp)[:2]
    T = np.repeat(th, len(lam))
    L = np.tile(lam, len(th))
    for wave in (0, 1):
        phys = og.new_physical()
        phys[:, 0, 0] = np.sin(pts[:, 0]) * (np.cos(pts[:, 1]) if wave else 1.0)
        out = og.regular_transform(og.spectral_transform(phys), th, lam)
        cl = np.cos(L) if wave else 1.0
        assert np.max(np.abs(out[:, 0, 0] - np.sin(T) * cl)) < 1e-2
        assert np.max(np.abs(out[:, 0, 1] - np.cos(T) * cl)) < 0.1
        if wave:
            assert np.max(np.abs(out[:, 0, 3] + np.sin(T) * np.sin(L))) < 1e-2
            assert np.max(np.abs(out[:, 0, 4] + np.sin(T) * np.cos(L))) < 1e-2
        else:
            assert np.max(np.abs(out[:, 0, 3])) < 1e-10 and np.max(np.abs(out[:, 0, 4])) < 1e-10
