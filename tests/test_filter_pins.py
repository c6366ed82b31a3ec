"""Spectral filters (applyFilter!, src/filtering.jl) without a GPU.

Pinned here:
- the host weights (springsteel_hip.filters, the values the library is given)
  against the known answers of the reference's testsets "SpectralFilter
  weights", "SpectralFilter windowed taper", "GaussianFilter weights" and
  "Window functions" (test/filtering.jl:9-132);
- the oracle's restatement of _applyFilter_impl! (oracle.OracleGrid.apply_filter)
  against the reference's grid-level testsets (test/filtering.jl:134-460: RL
  notch + low pass, RZ Chebyshev, RLZ both, Gaussian on RL, SL, several
  variables, the "default" key, empty dicts), each on a spectral array of ones;
- the block <-> mode map against the transform's own layout: a field made of
  one wavenumber / Chebyshev mode transforms to blocks that a zero weight at
  that mode clears and a zero weight at any other mode leaves alone;
- isapprox's rules: a weight within sqrt(eps) of 1 leaves the bytes alone, a
  weight of 0 assigns +0.0 (not a product: -0.0 and NaN do not survive), a NaN
  weight multiplies.
"""
import math

import numpy as np
import pytest

import oracle as O
import springsteel_hip as S
from springsteel_hip.errors import ArgumentError
from springsteel_hip.filters import (GaussianFilter, SpectralFilter, filter_weight, filter_weights, isapprox,
                                     window_weight)
from springsteel_hip.params import compute_derived_params, resolve


def approx(a, b):
    return isapprox(a, b)


def test_spectral_filter_weights_reference_testset():
    """test/filtering.jl:9-51."""
    f = SpectralFilter(low_pass=10)
    assert all(approx(filter_weight(f, k), 1.0) for k in (0, 5, 10))
    assert all(approx(filter_weight(f, k), 0.0) for k in (11, 100))
    f = SpectralFilter(high_pass=5)
    assert approx(filter_weight(f, 0), 0.0) and approx(filter_weight(f, 4), 0.0)
    assert approx(filter_weight(f, 5), 1.0) and approx(filter_weight(f, 10), 1.0)
    f = SpectralFilter(low_pass=20, high_pass=5)
    assert [approx(filter_weight(f, k), 1.0) for k in (0, 4, 5, 15, 20, 21)] == [False, False, True, True, True, False]
    assert approx(filter_weight(f, 21), 0.0) and approx(filter_weight(f, 0), 0.0)
    f = SpectralFilter(notch=[0, 3, 7])
    assert [filter_weight(f, k) for k in (0, 1, 3, 5, 7)] == [0.0, 1.0, 0.0, 1.0, 0.0]
    f = SpectralFilter(low_pass=10, notch=[1])
    assert [filter_weight(f, k) for k in (0, 1, 5, 11)] == [1.0, 0.0, 1.0, 0.0]
    f = SpectralFilter()
    assert all(filter_weight(f, k) == 1.0 for k in range(51))


def test_windowed_taper_reference_testset():
    """test/filtering.jl:53-94."""
    f = SpectralFilter(low_pass=10, window="hann", taper_width=5)
    assert approx(filter_weight(f, 8), 1.0) and approx(filter_weight(f, 10), 1.0)
    assert 0.0 < filter_weight(f, 11) < 1.0
    assert approx(filter_weight(f, 16), 0.0)
    assert all(filter_weight(f, k) >= filter_weight(f, k + 1) for k in range(11, 15))
    f = SpectralFilter(low_pass=10, window="lanczos", taper_width=5)
    assert approx(filter_weight(f, 10), 1.0) and 0.0 < filter_weight(f, 12) < 1.0
    f = SpectralFilter(low_pass=10, window="exponential", taper_width=5)
    assert approx(filter_weight(f, 10), 1.0) and 0.0 < filter_weight(f, 13) < 1.0
    f = SpectralFilter(high_pass=10, window="hann", taper_width=3)
    assert approx(filter_weight(f, 7), 0.0) and 0.0 < filter_weight(f, 8) < 1.0
    assert approx(filter_weight(f, 10), 1.0) and approx(filter_weight(f, 15), 1.0)
    f1 = SpectralFilter(low_pass=10, window="hann", taper_width=0)
    f2 = SpectralFilter(low_pass=10, window="boxcar")
    assert all(approx(filter_weight(f1, k), filter_weight(f2, k)) for k in range(21))


def test_taper_values_from_the_windows():
    """The taper positions: distance d into the stopband over taper_width (t = d / width),
    the window's value there; the windows' closed forms at t = 1/2 (hann 1/2, lanczos
    2/pi, exponential e^-1) and their ends."""
    assert window_weight("hann", 0.5) == pytest.approx(0.5, abs=1e-15)
    assert window_weight("lanczos", 0.5) == pytest.approx(2.0 / math.pi, rel=1e-15)
    assert window_weight("exponential", 0.5) == pytest.approx(math.exp(-1.0), rel=1e-15)
    for w in ("hann", "lanczos", "exponential", "boxcar"):
        assert window_weight(w, 0.0) == 1.0
        assert window_weight(w, -3.0) == 1.0                       # clamped
    assert window_weight("boxcar", 1.0) == 0.0 and window_weight("boxcar", 0.999) == 1.0
    # low pass 10, width 4: k = 12 is t = 1/2 of the window
    for win, val in (("hann", 0.5), ("lanczos", 2.0 / math.pi), ("exponential", math.exp(-1.0))):
        assert filter_weight(SpectralFilter(low_pass=10, window=win, taper_width=4), 12) == pytest.approx(val, rel=1e-15)
        # high pass 10, width 4: k = 8 is t = 1/2 too
        assert filter_weight(SpectralFilter(high_pass=10, window=win, taper_width=4), 8) == pytest.approx(val, rel=1e-15)
    # the notch wins over the taper
    assert filter_weight(SpectralFilter(low_pass=10, window="hann", taper_width=4, notch=[12]), 12) == 0.0


def test_gaussian_and_window_reference_testsets():
    """test/filtering.jl:96-132."""
    f = GaussianFilter(sigma=10.0)
    assert approx(filter_weight(f, 0), 1.0)
    assert approx(filter_weight(f, 10), math.exp(-1.0)) and approx(filter_weight(f, 20), math.exp(-4.0))
    f = GaussianFilter(sigma=10.0, order=2)
    assert approx(filter_weight(f, 10), math.exp(-1.0)) and filter_weight(f, 5) > math.exp(-1.0)
    assert filter_weight(f, 20) == pytest.approx(math.exp(-16.0), rel=1e-14)
    assert all(filter_weight(GaussianFilter(sigma=1000.0), k) > 0.99 for k in range(51))
    for w in ("boxcar", "hann", "lanczos", "exponential"):
        assert abs(window_weight(w, 0.0) - 1.0) <= 0.02
    assert window_weight("hann", 1.0) == 0.0
    assert window_weight("boxcar", 1.0) == 0.0
    assert window_weight("exponential", 1.0) < 0.02
    with pytest.raises(ArgumentError):
        window_weight("invalid", 0.5)


def _grid(**kw):
    gp = compute_derived_params(S.SpringsteelGridParameters(**kw))
    rg = resolve(gp)
    og = O.OracleGrid(rg)
    names = sorted(gp.vars, key=lambda k: gp.vars[k])
    fw, cw = filter_weights(gp, names, gp.iDim + gp.patchOffsetL, gp.b_kDim)
    return gp, og, fw, cw


def _ones(og):
    return np.ones((og.nspec, og.nvars), order="F")


def test_rl_fourier_reference_testset():
    """test/filtering.jl:175-212 and 'Filter integrated' (432-452) layout."""
    gp, og, fw, cw = _grid(geometry="RL", num_cells=10, iMin=0.0, iMax=100.0, vars={"u": 1},
                           fourier_filter={"u": SpectralFilter(low_pass=3, notch=[1])})
    b, K = gp.b_iDim, gp.iDim
    s = og.apply_filter(_ones(og), fw, cw)
    assert np.all(s[:b, 0] == 1.0)
    assert np.all(s[b:3 * b, 0] == 0.0)                               # k = 1 notched
    for k in (2, 3):
        assert np.all(s[(2 * k - 1) * b:(2 * k + 1) * b, 0] == 1.0)
    assert np.all(s[7 * b:(2 * K + 1) * b, 0] == 0.0)                 # k >= 4


def test_rz_chebyshev_reference_testset():
    """test/filtering.jl:246-278."""
    gp, og, fw, cw = _grid(geometry="RZ", num_cells=10, iMin=0.0, iMax=100.0, kMin=0.0, kMax=10.0, kDim=12,
                           vars={"u": 1}, chebyshev_filter={"u": SpectralFilter(low_pass=3)},
                           fourier_filter={"u": SpectralFilter(low_pass=0)})   # ignored on RZ
    b = gp.b_iDim
    s = og.apply_filter(_ones(og), fw, cw)
    assert np.all(s[:4 * b, 0] == 1.0) and np.all(s[4 * b:, 0] == 0.0)


def test_rlz_both_reference_testset():
    """test/filtering.jl:280-320."""
    gp, og, fw, cw = _grid(geometry="RLZ", num_cells=5, iMin=0.0, iMax=50.0, kMin=0.0, kMax=10.0, kDim=6,
                           vars={"u": 1}, fourier_filter={"u": SpectralFilter(low_pass=2)},
                           chebyshev_filter={"u": SpectralFilter(low_pass=2)})
    b, K = gp.b_iDim, gp.iDim
    bs = b * (1 + 2 * K)
    s = og.apply_filter(_ones(og), fw, cw)
    assert np.all(s[:b, 0] == 1.0)
    r1 = b + (3 - 1) * 2 * b
    assert np.all(s[r1:r1 + b, 0] == 0.0)                             # z_b = 1, k = 3
    assert np.all(s[3 * bs:3 * bs + b, 0] == 0.0)                     # z_b = 4, k = 0
    assert np.all(s[b:2 * b, 0] == 1.0)                               # z_b = 1, k = 1
    # every block: kept iff k <= 2 and z_b - 1 <= 2
    for zb in range(gp.b_kDim):
        for p in range(1 + 2 * K):
            k = (p + 1) // 2
            blk = s[zb * bs + p * b: zb * bs + (p + 1) * b, 0]
            assert np.all(blk == (1.0 if (k <= 2 and zb <= 2) else 0.0)), (zb, p)


def test_gaussian_sl_multivar_default_reference_testsets():
    """test/filtering.jl:322-430: Gaussian on RL, SL low pass, one of two variables,
    the "default" key, empty dicts."""
    gp, og, fw, cw = _grid(geometry="RL", num_cells=10, iMin=0.0, iMax=100.0, vars={"u": 1},
                           fourier_filter={"u": GaussianFilter(sigma=5.0)})
    b = gp.b_iDim
    s = og.apply_filter(_ones(og), fw, cw)
    assert np.all(s[:b, 0] == 1.0)
    assert all(approx(x, math.exp(-1.0)) for x in s[9 * b:11 * b, 0])
    assert all(approx(x, math.exp(-4.0)) for x in s[19 * b:21 * b, 0])

    gp, og, fw, cw = _grid(geometry="SL", num_cells=10, iMin=0.0, iMax=math.pi, vars={"u": 1},
                           fourier_filter={"u": SpectralFilter(low_pass=3)})
    b, K = gp.b_iDim, gp.iDim
    s = og.apply_filter(_ones(og), fw, cw)
    assert np.all(s[:7 * b, 0] == 1.0) and np.all(s[7 * b:(2 * K + 1) * b, 0] == 0.0)

    gp, og, fw, cw = _grid(geometry="RL", num_cells=10, iMin=0.0, iMax=100.0, vars={"u": 1, "v": 2},
                           fourier_filter={"u": SpectralFilter(low_pass=3)})
    b = gp.b_iDim
    s = og.apply_filter(_ones(og), fw, cw)
    assert np.all(s[9 * b:11 * b, 0] == 0.0) and np.all(s[:, 1] == 1.0)

    gp, og, fw, cw = _grid(geometry="RL", num_cells=10, iMin=0.0, iMax=100.0, vars={"u": 1, "v": 2},
                           fourier_filter={"default": SpectralFilter(low_pass=3)})
    s = og.apply_filter(_ones(og), fw, cw)
    assert np.all(s[9 * b:11 * b, :] == 0.0)

    gp, og, fw, cw = _grid(geometry="RL", num_cells=10, iMin=0.0, iMax=100.0, vars={"u": 1})
    assert fw is None and cw is None
    s = og.apply_filter(_ones(og), fw, cw)
    assert np.all(s == 1.0)


@pytest.mark.parametrize("geometry", ["RL", "RLZ"])
def test_block_map_against_the_transform_layout(geometry):
    """A field of one wavenumber k (and one Chebyshev mode for RLZ) lands in blocks that
    a zero weight at k clears; zero weights at every other k leave them alone."""
    kw = dict(geometry=geometry, num_cells=6, iMin=0.0, iMax=60.0, vars={"u": 1})
    if geometry == "RLZ":
        kw.update(kMin=0.0, kMax=10.0, kDim=8)
    gp = compute_derived_params(S.SpringsteelGridParameters(**kw))
    og = O.OracleGrid(resolve(gp))
    pts = og.gridpoints()
    K = gp.iDim
    for k in (0, 3):
        phys = og.new_physical()
        r, lam = pts[:, 0], pts[:, 1]
        radial = np.exp(-((r - 30.0) / 15.0) ** 2)
        f = radial * np.cos(k * lam + 0.3)
        if geometry == "RLZ":
            z = pts[:, 2]
            f = f * (1.0 + 0.5 * (z / 5.0 - 1.0))
        phys[:, 0, 0] = f
        spec = og.spectral_transform(phys)
        nz = np.abs(spec[:, 0]) > 1e-12 * np.abs(spec).max()
        only = np.ones((1, K + 1))
        only[0, k] = 0.0
        others = np.zeros((1, K + 1))
        others[0, k] = 1.0
        cleared = og.apply_filter(spec.copy(order="F"), only, None)
        assert not np.any(cleared[nz, 0])
        kept = og.apply_filter(spec.copy(order="F"), others, None)
        assert np.array_equal(kept[nz, 0], spec[nz, 0])
        if geometry == "RLZ":
            # the field is linear in z: Chebyshev modes 0 and 1 only; zero weight at mode >= 2 keeps it
            cw = np.ones((1, gp.b_kDim))
            cw[0, 2:] = 0.0
            out = og.apply_filter(spec.copy(order="F"), None, cw)
            assert np.array_equal(out[nz], spec[nz])
            assert np.all(np.abs(out - spec) <= 1e-12 * np.abs(spec).max())
            cw = np.ones((1, gp.b_kDim))
            cw[0, 1] = 0.0
            assert not np.array_equal(og.apply_filter(spec.copy(order="F"), None, cw), spec)


def test_isapprox_rules():
    gp, og, fw, cw = _grid(geometry="RL", num_cells=4, iMin=0.0, iMax=40.0, vars={"u": 1})
    K, b = gp.iDim, gp.b_iDim
    rng = np.random.default_rng(0)
    spec = np.asfortranarray(rng.normal(size=(og.nspec, 1)))
    w = np.ones((1, K + 1))
    w[0, 1] = 1.0 + 1e-9                   # within sqrt(eps): untouched
    w[0, 2] = 1.0 + 1e-7                   # outside: scaled
    w[0, 3] = 0.0                          # assigned +0.0
    w[0, 4] = -0.0                         # -0.0 == 0.0: assigned +0.0 too
    w[0, 5] = float("nan")                 # neither: multiplies
    out = og.apply_filter(spec.copy(order="F"), w, None)
    assert np.array_equal(out[b:3 * b], spec[b:3 * b])
    assert np.array_equal(out[3 * b:5 * b], spec[3 * b:5 * b] * (1.0 + 1e-7))
    assert np.all(out[5 * b:9 * b] == 0.0) and not np.any(np.signbit(out[5 * b:9 * b]))
    assert np.all(np.isnan(out[9 * b:11 * b]))
    assert np.array_equal(out[11 * b:], spec[11 * b:])
