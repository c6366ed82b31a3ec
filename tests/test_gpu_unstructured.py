"""GPU: evaluate_unstructured (ssg_evaluate_unstructured) against the oracle.

Bar: relative L2 per variable <= 1e-12 over the points inside the domain, and
the same NaN / fill pattern outside it.  The oracle's restatement is pinned in
test_unstructured_pins.py (the reference's own testset, closed forms with the
mirrored sine, agreement with the regular transform).
"""
import math

import numpy as np
import pytest

import oracle as O
import springsteel_hip as S
from springsteel_hip.errors import DomainError
from parity_util import seeded_field

pytestmark = pytest.mark.gpu
TOL = 1e-12


def _grid(gp, seed=0):
    g = S.createGrid(gp)
    pts = S.getGridpoints(g)
    g.physical[:, :, 0] = seeded_field(pts, g.resolved.nvars, seed, g.geometry)
    S.spectralTransform_(g)
    return g, O.OracleGrid(g.resolved)


def _points(gp, n, seed, geometry, oob=0.0):
    rng = np.random.default_rng(seed)
    lo, hi = gp.iMin, gp.iMax
    span = hi - lo
    r = rng.uniform(lo - oob * span, hi + oob * span, n)
    cols = [r]
    if geometry in ("RL", "RLZ"):
        cols.append(rng.uniform(-3.0, 10.0, n))
    if geometry in ("RZ", "RLZ"):
        zs = gp.kMax - gp.kMin
        cols.append(rng.uniform(gp.kMin - oob * zs, gp.kMax + oob * zs, n))
    p = np.column_stack(cols)
    # end points of every bounded axis
    p[0, 0], p[1, 0] = lo, hi
    if geometry != "RL":
        p[2, -1], p[3, -1] = gp.kMin, gp.kMax
    return p


def _check(g, og, pts, **kw):
    out = S.evaluate_unstructured(g, pts, **kw)
    spec = np.asfortranarray(g.spectral.copy())
    fill = kw.get("out_of_bounds", "nan")
    ref = og.evaluate_unstructured(spec, pts, fill=float("nan") if fill == "nan" else float(fill), threads=8)
    assert np.array_equal(np.isnan(out), np.isnan(ref))
    ok = ~np.isnan(ref[:, 0])
    worst = 0.0
    for v in range(g.resolved.nvars):
        d = np.linalg.norm(out[ok, v] - ref[ok, v]) / max(np.linalg.norm(ref[ok, v]), 1e-300)
        worst = max(worst, d)
        assert d <= TOL, (v, d)
        assert np.array_equal(out[~ok, v], ref[~ok, v]) or np.all(np.isnan(out[~ok, v]))
    return out, worst


def test_rl_random_points_with_out_of_bounds():
    gp = S.SpringsteelGridParameters(geometry="RL", num_cells=12, iMin=0.0, iMax=60.0,
                                     vars={"u": 1, "v": 2, "w": 3}, BCL={"default": S.CubicBSpline.R0,
                                                                          "v": S.CubicBSpline.R1T0},
                                     BCR={"default": S.CubicBSpline.R0, "w": S.CubicBSpline.R1T1})
    g, og = _grid(gp)
    pts = _points(g.params, 3000, 1, "RL", oob=0.05)
    pts[5, 0] = math.nan
    out, worst = _check(g, og, pts)
    print(f"\nRL unstructured worst rel L2 {worst:.2e}")
    g.close()


def test_rz_random_points_fill_value():
    gp = S.SpringsteelGridParameters(geometry="RZ", num_cells=10, iMin=0.0, iMax=100.0, kMin=0.0, kMax=50.0,
                                     kDim=16, vars={"u": 1, "v": 2}, BCB={"default": S.Chebyshev.R0,
                                                                          "v": S.Chebyshev.R1T0})
    g, og = _grid(gp, seed=1)
    pts = _points(g.params, 2500, 2, "RZ", oob=0.05)
    out, worst = _check(g, og, pts, out_of_bounds=-999.0)
    print(f"\nRZ unstructured worst rel L2 {worst:.2e}")
    g.close()


def test_rlz_random_points_many_per_cell():
    """Several hundred points per cell: groups of 64 with one first tap, several per cell."""
    gp = S.SpringsteelGridParameters(geometry="RLZ", num_cells=6, iMin=0.0, iMax=60.0, kMin=0.0, kMax=10.0,
                                     kDim=12, vars={"u": 1, "v": 2}, BCB={"default": S.Chebyshev.R0,
                                                                          "v": S.Chebyshev.R1T0})
    g, og = _grid(gp, seed=2)
    pts = _points(g.params, 4000, 3, "RLZ", oob=0.02)
    out, worst = _check(g, og, pts)
    print(f"\nRLZ unstructured worst rel L2 {worst:.2e}")
    g.close()


def test_error_mode_vars_and_device_form():
    import torch
    gp = S.SpringsteelGridParameters(geometry="RLZ", num_cells=5, iMin=5.0, iMax=40.0, kMin=-3.0, kMax=7.0,
                                     kDim=10, vars={"u": 1, "v": 2, "w": 3})
    g, og = _grid(gp, seed=3)
    pts = _points(g.params, 500, 4, "RLZ", oob=0.1)
    with pytest.raises(DomainError):
        S.evaluate_unstructured(g, pts, out_of_bounds="error")
    host = S.evaluate_unstructured(g, pts)
    sel = S.evaluate_unstructured(g, pts, vars=["w", "u", "nope"])
    assert np.array_equal(sel[:, 0], host[:, 0], equal_nan=True) and np.all(np.isnan(sel[:, 1]))
    assert np.array_equal(sel[:, 2], host[:, 2], equal_nan=True)
    spec = torch.from_numpy(np.ascontiguousarray(g.spectral.T)).cuda()
    dev = S.evaluate_unstructured_device(g, spec, pts)
    torch.cuda.synchronize()
    assert np.array_equal(dev.cpu().numpy().T, host, equal_nan=True)
    inb = pts[(pts[:, 0] >= 5.0) & (pts[:, 0] <= 40.0) & (pts[:, 2] >= -3.0) & (pts[:, 2] <= 7.0)]
    S.evaluate_unstructured(g, inb, out_of_bounds="error")
    g.close()


def test_large_call_in_chunks_matches_small_calls_and_oracle():
    """1.3 M points: more than one chunk of groups (2^20 points); each point's value
    does not depend on which other points share its call, so slices give the same bytes."""
    gp = S.SpringsteelGridParameters(geometry="RZ", num_cells=8, iMin=0.0, iMax=40.0, kMin=0.0, kMax=20.0, kDim=10,
                                     vars={"u": 1, "v": 2})
    g, og = _grid(gp, seed=5)
    pts = _points(g.params, 1_300_000, 6, "RZ")
    big = S.evaluate_unstructured(g, pts)
    for a, b in ((0, 100_000), (650_000, 700_000), (1_250_000, 1_300_000)):
        assert np.array_equal(S.evaluate_unstructured(g, pts[a:b]), big[a:b])
    spec = np.asfortranarray(g.spectral.copy())
    sub = np.arange(0, 1_300_000, 97)
    ref = og.evaluate_unstructured(spec, pts[sub], threads=8)
    for v in range(2):
        assert np.linalg.norm(big[sub, v] - ref[:, v]) / np.linalg.norm(ref[:, v]) <= TOL
    g.close()


def test_points_on_regular_grid_match_regular_transform():
    """At a tensor product of points, the values equal regularGridTransform's value slot
    (the reference's own testset compares the two, test/interpolation.jl:660-690)."""
    gp = S.SpringsteelGridParameters(geometry="RL", num_cells=8, iMin=5.0, iMax=50.0, vars={"u": 1})
    g, og = _grid(gp, seed=7)
    reg_pts = S.getRegularGridpoints(g)
    reg = S.regularGridTransform(g, reg_pts)
    res = S.evaluate_unstructured(g, reg_pts)
    assert np.linalg.norm(res[:, 0] - reg[:, 0, 0]) / np.linalg.norm(reg[:, 0, 0]) <= TOL
    g.close()
