"""Pins of the oracle's R3X background-coefficient path (no GPU needed).

SAtransform!'s R3X branch (src/CubicBSpline.jl:1367-1396) solves for the
interior coefficients against b - pq ahat and adds ahat back; the grid
transforms take ahat per spectral line (_get_wavenumber_ahat,
src/multipatch_registry.jl:23-33; RL src/transforms_cylindrical.jl:280-325,
RLZ 690-749).  Pinned here by:
- the independent dense restatement in npref.py (scipy B-splines, the interior
  selector for the rank-3 sides), to 1e-12;
- set_ahat_r3x!'s closed form (src/CubicBSpline.jl:1437-1467): border
  coefficients solving the 3 x 3 border system give a spline with exactly the
  prescribed value and first two derivatives at the boundary, whatever b is;
- b = pq ahat returns a = ahat (the interior correction vanishes);
- a zero (or absent) ahat is the homogeneous R3 path, bitwise, through a whole
  grid transform, and a nonzero one changes only R3X variables.
"""
import numpy as np
import pytest

import npref
import oracle as O
import springsteel_hip as S
from springsteel_hip.bcs import SSG_BC_R0, SSG_BC_R3, SSG_BC_R3X
from springsteel_hip.params import compute_derived_params, resolve


def _spl(kind_l, kind_r, nc=10, xmin=0.0, xmax=10.0):
    return O.OracleSpline(xmin, xmax, nc, bcl=(kind_l, 0.0, 0.0), bcr=(kind_r, 0.0, 0.0))


@pytest.mark.parametrize("sides", [(SSG_BC_R3X, SSG_BC_R3X), (SSG_BC_R3X, SSG_BC_R0), (SSG_BC_R0, SSG_BC_R3X)])
def test_sa_ahat_vs_dense_restatement(sides):
    sp = _spl(*sides)
    M = sp.Mdim
    rl = 3 if sides[0] == SSG_BC_R3X else 0
    rr = 3 if sides[1] == SSG_BC_R3X else 0
    ref = npref.NpSpline(0.0, 10.0, 10, gamma=npref.selector_gamma(M, rl, rr))
    rng = np.random.default_rng(sum(sides))
    b, ahat = rng.normal(size=M), rng.normal(size=M)
    a = sp.sa_ahat(b, ahat)
    want = ref.sa_ahat(b, ahat)
    assert np.linalg.norm(a - want) / np.linalg.norm(want) < 1e-12


def test_border_values_closed_form():
    """set_ahat_r3x!(spline, u0, u1, u2, :left/:right): the border coefficients from the
    3 x 3 system of basis values and derivatives at the boundary."""
    nc, xmin, xmax = 12, 0.0, 6.0
    sp = _spl(SSG_BC_R3X, SSG_BC_R3X, nc, xmin, xmax)
    ref = npref.NpSpline(xmin, xmax, nc)
    M = sp.Mdim
    ahat = np.zeros(M)
    for side, x, ms, want in (("left", xmin, (-1, 0, 1), (1.0, 0.5, -2.0)), ("right", xmax, (nc - 1, nc, nc + 1),
                                                                             (2.0, -1.0, 0.25))):
        B = np.array([[ref.node_spline(m).derivative(d)(x) if d else ref.node_spline(m)(x) for m in ms]
                      for d in range(3)])
        B = np.nan_to_num(B, nan=0.0)
        coef = np.linalg.solve(B, np.array(want))
        if side == "left":
            ahat[:3] = coef
        else:
            ahat[-3:] = coef
    rng = np.random.default_rng(7)
    b = sp.sb(np.sin(sp.mishpoints()) + rng.normal(size=sp.mishDim) * 0.1)
    a = sp.sa_ahat(b, ahat)
    for x, want in ((xmin + 1e-12, (1.0, 0.5, -2.0)), (xmax - 1e-12, (2.0, -1.0, 0.25))):
        for d in range(3):
            assert abs(sp.si(a, np.array([x]), d)[0] - want[d]) < 1e-8 * (1 + abs(want[d]) * nc ** d), (x, d)


def test_pq_ahat_returns_ahat():
    sp = _spl(SSG_BC_R3X, SSG_BC_R3X)
    P, Q = sp.pq()
    ahat = np.random.default_rng(3).normal(size=sp.Mdim)
    a = sp.sa_ahat((P + Q) @ ahat, ahat)
    assert np.max(np.abs(a - ahat)) < 1e-12 * np.max(np.abs(ahat)) * 1e2


def _grid(bc, geometry="RL", **kw):
    base = dict(geometry=geometry, num_cells=8, iMin=0.0, iMax=40.0, vars={"u": 1, "w": 2},
                BCL={"default": S.CubicBSpline.R0, "u": bc}, BCR={"default": S.CubicBSpline.R0, "u": bc})
    if geometry != "RL":
        base.update(kMin=0.0, kMax=10.0, kDim=8)
    base.update(kw)
    gp = compute_derived_params(S.SpringsteelGridParameters(**base))
    return gp, O.OracleGrid(resolve(gp))


@pytest.mark.parametrize("geometry", ["RL", "RZ", "RLZ"])
def test_grid_ahat_zero_is_r3_and_nonzero_moves_r3x_only(geometry):
    gp, g3 = _grid(S.CubicBSpline.R3, geometry)
    _, gx = _grid(S.CubicBSpline.R3X, geometry)
    rng = np.random.default_rng(11)
    spec = np.asfortranarray(rng.normal(size=(g3.nspec, g3.nvars)))
    o3 = g3.grid_transform(spec)
    ox = gx.grid_transform(spec)
    assert np.array_equal(o3, ox)
    gx.set_ahat(np.zeros((gx.nspec, gx.nvars)))
    assert np.array_equal(gx.grid_transform(spec), o3)
    ahat = np.asfortranarray(rng.normal(size=(gx.nspec, gx.nvars)))
    gx.set_ahat(ahat)
    o = gx.grid_transform(spec)
    assert np.array_equal(o[:, 1, :], o3[:, 1, :])          # w: R0 both sides, ahat unused
    assert not np.allclose(o[:, 0, 0], o3[:, 0, 0])
    gx.set_ahat(None)
    assert np.array_equal(gx.grid_transform(spec), o3)


def test_grid_ahat_line_mapping_rl():
    """One line's ahat reaches that line only: the grid result moves by the evaluation of
    (sa_ahat(b, ahat) - sa(b)) of that line alone (k = 2, the sine line p = 4)."""
    gp, gx = _grid(S.CubicBSpline.R3X, "RL", vars={"u": 1})
    sp = O.OracleSpline(0.0, 40.0, 8, bcl=(SSG_BC_R3X, 0.0, 0.0), bcr=(SSG_BC_R3X, 0.0, 0.0))
    rng = np.random.default_rng(5)
    spec = np.asfortranarray(rng.normal(size=(gx.nspec, 1)))
    o0 = gx.grid_transform(spec)
    ahat = np.zeros((gx.nspec, 1), order="F")
    M = gp.b_iDim
    line = 4
    ahat[line * M:(line + 1) * M, 0] = rng.normal(size=M)
    gx.set_ahat(ahat)
    o1 = gx.grid_transform(spec)
    b = spec[line * M:(line + 1) * M, 0]
    da = sp.sa_ahat(b, ahat[line * M:(line + 1) * M, 0]) - sp.sa(b)
    # the sine line of k = 2 enters each ring's values as a wavenumber-2 component whose
    # amplitude is the line's SI value at the ring's radius (the phase rotation keeps it)
    du = sp.si(da, sp.mishpoints(), 0)
    assert np.linalg.norm(o1[:, 0, 0] - o0[:, 0, 0]) > 1e-6
    # the change is a pure wavenumber-2 sine in the rotated angle: project it out ring by ring
    ri = np.arange(1, gp.iDim + 1)
    offs = np.cumsum(np.concatenate([[0], 4 + 4 * ri]))
    for r in range(gp.iDim):
        d = o1[offs[r]:offs[r + 1], 0, 0] - o0[offs[r]:offs[r + 1], 0, 0]
        n = len(d)
        c = np.fft.rfft(d) / n
        assert np.allclose(np.abs(c[[0, 1, 3]]), 0.0, atol=1e-12 * (1 + np.abs(du).max()))
        want = abs(du[r]) if ri[r] >= 2 else 0.0          # a ring carries k = 2 only if ri >= 2
        assert abs(abs(c[2]) - want) < 1e-10 * (1 + abs(du[r]))
