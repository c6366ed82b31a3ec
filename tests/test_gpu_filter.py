"""GPU: applyFilter! (ssg_plan_set_filter / ssg_apply_filter) against the oracle.

The oracle's restatement of _applyFilter_impl! and the host weights are pinned
in test_filter_pins.py.  Here the library's filter kernel gives the oracle's
bytes on random spectral arrays (negative entries included, so that a zero
weight storing -0.0 would show) for every geometry and filter kind, in host and
device memory; spectralTransform_ is the explicit-array transform followed by
the filter; the reference's own round-trip and "integrated" testsets
(test/filtering.jl:214-244, 432-452) hold.
"""
import math

import numpy as np
import pytest

import oracle as O
import springsteel_hip as S
from springsteel_hip.filters import GaussianFilter, SpectralFilter, filter_weights

pytestmark = pytest.mark.gpu

CASES = {
    "rl_notch_lowpass": dict(geometry="RL", num_cells=10, iMin=0.0, iMax=100.0, vars={"u": 1, "v": 2},
                             fourier_filter={"u": SpectralFilter(low_pass=12, notch=[1, 5]),
                                             "v": SpectralFilter(high_pass=3, window="hann", taper_width=4)}),
    "rz_cheb": dict(geometry="RZ", num_cells=12, iMin=0.0, iMax=100.0, kMin=0.0, kMax=10.0, kDim=16,
                    vars={"u": 1, "w": 2}, chebyshev_filter={"default": SpectralFilter(low_pass=6, window="lanczos",
                                                                                       taper_width=3)}),
    "rlz_both": dict(geometry="RLZ", num_cells=8, iMin=0.0, iMax=80.0, kMin=0.0, kMax=10.0, kDim=12,
                     vars={"u": 1, "v": 2, "w": 3},
                     fourier_filter={"u": GaussianFilter(sigma=6.0, order=2),
                                     "w": SpectralFilter(low_pass=10, window="exponential", taper_width=5)},
                     chebyshev_filter={"v": SpectralFilter(low_pass=4), "w": GaussianFilter(sigma=5.0)}),
    "sl_lowpass": dict(geometry="SL", num_cells=12, iMin=0.0, iMax=math.pi, vars={"u": 1},
                       fourier_filter={"u": SpectralFilter(low_pass=7, window="hann", taper_width=3)}),
    "slz_cheb_only": dict(geometry="SLZ", num_cells=8, iMin=0.0, iMax=math.pi, kMin=0.0, kMax=10.0, kDim=10,
                          vars={"u": 1, "v": 2}, chebyshev_filter={"default": SpectralFilter(high_pass=1)}),
}


def _setup(name):
    gp = S.SpringsteelGridParameters(**CASES[name])
    g = S.createGrid(gp)
    og = O.OracleGrid(g.resolved)
    names = sorted(g.params.vars, key=lambda k: g.params.vars[k])
    fw, cw = filter_weights(g.params, names, g.params.iDim + g.params.patchOffsetL, g.params.b_kDim)
    return g, og, fw, cw


@pytest.mark.parametrize("name", sorted(CASES))
def test_filter_vs_oracle_bitwise(name):
    torch = pytest.importorskip("torch")
    g, og, fw, cw = _setup(name)
    rng = np.random.default_rng(len(name))
    spec = np.asfortranarray(rng.normal(size=(g.nspec, g.resolved.nvars)))
    ref = og.apply_filter(spec.copy(order="F"), fw, cw)
    out = S.applyFilter(g, spec.copy(order="F"))
    changed = int(np.sum(ref != spec))
    print(f"\n{name}: {changed} of {spec.size} entries changed")
    assert changed > 0
    assert np.array_equal(out, ref)
    assert not np.any(np.signbit(out[ref == 0.0]))
    # device form, bitwise the host form
    d = torch.from_numpy(np.ascontiguousarray(spec.T)).cuda()
    S.applyFilter_device(g, d)
    torch.cuda.synchronize()
    assert np.array_equal(d.cpu().numpy().T, ref)
    g.close()


@pytest.mark.parametrize("name", ["rl_notch_lowpass", "rlz_both", "slz_cheb_only"])
def test_spectral_transform_bang_filters(name):
    """spectralTransform!(grid) = spectralTransform + applyFilter!; the explicit-array form
    does not filter (src/transforms_cylindrical.jl:153-157, 524-528)."""
    g, og, fw, cw = _setup(name)
    pts = S.getGridpoints(g)
    rng = np.random.default_rng(3)
    nv = g.resolved.nvars
    g.physical[:, :, 0] = np.cos(pts[:, 1:2] * np.arange(1, nv + 1)) * np.exp(-(pts[:, 0:1] / pts[:, 0].max()) ** 2) \
        + 0.1 * rng.normal(size=(g.npts, nv))
    plain = np.zeros_like(g.spectral)
    S.spectralTransform(g, g.physical, plain)
    S.spectralTransform_(g)
    assert np.array_equal(g.spectral, og.apply_filter(plain.copy(order="F"), fw, cw))
    assert not np.array_equal(g.spectral, plain)
    g.close()


def test_reference_round_trip_and_integrated_testsets():
    """test/filtering.jl:214-244 (low pass lowers the norm and changes the field) and
    432-452 (after spectralTransform! the wavenumbers above the cut are zero)."""
    kw = dict(geometry="RL", num_cells=10, iMin=0.0, iMax=100.0, vars={"u": 1})
    gnf = S.createGrid(S.SpringsteelGridParameters(**kw))
    gf = S.createGrid(S.SpringsteelGridParameters(**kw, fourier_filter={"u": SpectralFilter(low_pass=3)}))
    data = np.random.default_rng(11).normal(size=gnf.npts)
    gnf.physical[:, 0, 0] = data
    gf.physical[:, 0, 0] = data
    S.spectralTransform_(gnf)
    S.spectralTransform_(gf)
    b, K = gf.params.b_iDim, gf.params.iDim
    assert np.all(gf.spectral[7 * b:(2 * K + 1) * b, 0] == 0.0)
    S.gridTransform_(gnf)
    S.gridTransform_(gf)
    assert np.linalg.norm(gf.physical[:, 0, 0]) <= np.linalg.norm(gnf.physical[:, 0, 0])
    assert not np.allclose(gf.physical[:, 0, 0], gnf.physical[:, 0, 0])
    gnf.close()
    gf.close()


def test_clear_and_no_filter_are_noops():
    g, og, fw, cw = _setup("rl_notch_lowpass")
    spec = np.asfortranarray(np.random.default_rng(5).normal(size=(g.nspec, g.resolved.nvars)))
    g.plan.set_filter(None, None)
    assert np.array_equal(S.applyFilter(g, spec.copy(order="F")), spec)
    # weights of exactly 1 and within sqrt(eps) of 1 leave every byte alone
    K = g.params.iDim
    w = np.ones((g.resolved.nvars, K + 1))
    w[:, ::2] = 1.0 + 1e-9
    g.plan.set_filter(w, None)
    assert np.array_equal(S.applyFilter(g, spec.copy(order="F")), spec)
    w[0, 3] = 0.5
    g.plan.set_filter(w, None)
    out = S.applyFilter(g, spec.copy(order="F"))
    assert np.array_equal(out, og.apply_filter(spec.copy(order="F"), w, None))
    g.close()
