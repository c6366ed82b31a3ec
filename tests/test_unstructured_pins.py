"""Pins of the oracle's evaluate_unstructured restatement (no GPU needed).

evaluate_unstructured (src/interpolation.jl:1735-1822; RL 1628-1668, RZ
1376-1402, RLZ 1670-1733) is fixed here by:
- the reference's own testset "evaluate_unstructured" (test/interpolation.jl:
  627-877): agreement with regularGridTransform at the regular points (RL r cos
  lambda, RZ r z, atol 1e-8), the out-of-bounds modes, and the axisymmetric RL /
  RLZ cases (atol 0.05);
- closed forms carried exactly, including a sine term, which comes back mirrored
  (f(r, -lambda)) as in regularGridTransform: the sum adds +rs sin(k lambda);
- agreement to 1e-12 with the regular transform (pinned in test_regular_pins.py)
  at a tensor product of points, for band-limited fields with R1T0 and R0 BCs.
"""
import math

import numpy as np

import oracle as O
import springsteel_hip as S
from springsteel_hip.params import compute_derived_params, resolve


def _grid(**kw):
    gp = compute_derived_params(S.SpringsteelGridParameters(**kw))
    return gp, O.OracleGrid(resolve(gp))


def _regular_points_matrix(gp, geometry):
    r = np.linspace(gp.iMin, gp.iMax, gp.i_regular_out)
    lam = np.array([2 * math.pi * (j - 1) / gp.j_regular_out for j in range(1, gp.j_regular_out + 1)])
    z = np.linspace(gp.kMin, gp.kMax, gp.k_regular_out)
    if geometry == "RL":
        return (r, lam), np.column_stack([np.repeat(r, len(lam)), np.tile(lam, len(r))])
    if geometry == "RZ":
        return (r, z), np.column_stack([np.repeat(r, len(z)), np.tile(z, len(r))])
    return (r, lam, z), np.column_stack([np.repeat(r, len(lam) * len(z)), np.tile(np.repeat(lam, len(z)), len(r)),
                                         np.tile(z, len(r) * len(lam))])


def test_reference_testset_rl_matches_regular():
    gp, og = _grid(geometry="RL", iMin=5.0, iMax=50.0, num_cells=8, vars={"u": 1})
    pts = og.gridpoints()
    phys = og.new_physical()
    phys[:, 0, 0] = pts[:, 0] * np.cos(pts[:, 1])
    spec = og.spectral_transform(phys)
    (r, lam), mat = _regular_points_matrix(gp, "RL")
    reg = og.regular_transform(spec, r, lam)
    res = og.evaluate_unstructured(spec, mat)
    assert res.shape == (mat.shape[0], 1)
    assert np.allclose(res[:, 0], reg[:, 0, 0], rtol=0, atol=1e-8)


def test_reference_testset_rz_matches_regular():
    gp, og = _grid(geometry="RZ", iMin=1.0, iMax=5.0, num_cells=6, kMin=0.0, kMax=10.0, kDim=12, vars={"u": 1})
    pts = og.gridpoints()
    phys = og.new_physical()
    phys[:, 0, 0] = pts[:, 0] * pts[:, 1]
    spec = og.spectral_transform(phys)
    (x, z), mat = _regular_points_matrix(gp, "RZ")
    reg = og.regular_transform(spec, x, None, z)
    res = og.evaluate_unstructured(spec, mat)
    assert np.allclose(res[:, 0], reg[:, 0, 0], rtol=0, atol=1e-8)
    assert np.allclose(res[:, 0], mat[:, 0] * mat[:, 1], rtol=0, atol=1e-10)


def test_reference_testset_axisymmetric_rl_and_rlz():
    gp, og = _grid(geometry="RL", iMin=0.0, iMax=40.0, num_cells=10, vars={"u": 1})
    pts = og.gridpoints()
    phys = og.new_physical()
    phys[:, 0, 0] = np.sin(math.pi * pts[:, 0] / 40.0)
    spec = og.spectral_transform(phys)
    tp = np.array([[10.0, 0.0], [15.0, math.pi / 4], [20.0, math.pi / 2], [25.0, math.pi], [30.0, 3 * math.pi / 2],
                   [5.0, 0.3]])
    res = og.evaluate_unstructured(spec, tp)
    assert np.allclose(res[:, 0], np.sin(math.pi * tp[:, 0] / 40.0), atol=0.05)

    gp, og = _grid(geometry="RLZ", iMin=0.0, iMax=30.0, num_cells=8, kMin=0.0, kMax=10.0, kDim=10, vars={"u": 1})
    pts = og.gridpoints()
    phys = og.new_physical()
    phys[:, 0, 0] = np.sin(math.pi * pts[:, 0] / 30.0)
    spec = og.spectral_transform(phys)
    tp = np.array([[5.0, 0.0, 5.0], [10.0, math.pi / 4, 2.0], [15.0, math.pi / 2, 7.0], [20.0, math.pi, 4.0],
                   [25.0, 3 * math.pi / 2, 8.0]])
    res = og.evaluate_unstructured(spec, tp)
    assert res.shape == (5, 1)
    assert np.allclose(res[:, 0], np.sin(math.pi * tp[:, 0] / 30.0), atol=0.05)


def test_out_of_bounds_fill_modes():
    gp, og = _grid(geometry="RZ", iMin=0.0, iMax=4.0, num_cells=8, kMin=0.0, kMax=2.0, kDim=8, vars={"u": 1})
    phys = og.new_physical()
    phys[:, 0, 0] = 1.0
    spec = og.spectral_transform(phys)
    tp = np.array([[-1.0, 1.0], [2.0, 1.0], [5.0, 1.0], [2.0, 2.5], [2.0, 0.0]])
    res = og.evaluate_unstructured(spec, tp)
    assert np.isnan(res[0, 0]) and np.isnan(res[2, 0]) and np.isnan(res[3, 0])
    assert abs(res[1, 0] - 1.0) < 1e-12 and abs(res[4, 0] - 1.0) < 1e-12
    res = og.evaluate_unstructured(spec, tp, fill=-999.0)
    assert res[0, 0] == -999.0 and res[2, 0] == -999.0 and res[3, 0] == -999.0


def test_closed_form_with_mirrored_sine():
    """f = r cos(lambda) + 0.5 r sin(lambda) + 3 is carried exactly; the reference's
    sum adds the sine lines as +rs sin(k lambda), so the value is f(r, -lambda)."""
    gp, og = _grid(geometry="RL", iMin=5.0, iMax=50.0, num_cells=8, vars={"u": 1})
    pts = og.gridpoints()
    phys = og.new_physical()
    phys[:, 0, 0] = pts[:, 0] * (np.cos(pts[:, 1]) + 0.5 * np.sin(pts[:, 1])) + 3.0
    spec = og.spectral_transform(phys)
    rng = np.random.default_rng(3)
    tp = np.column_stack([rng.uniform(5.0, 50.0, 40), rng.uniform(-2.0, 9.0, 40)])
    res = og.evaluate_unstructured(spec, tp)
    exact = tp[:, 0] * (np.cos(tp[:, 1]) - 0.5 * np.sin(tp[:, 1])) + 3.0
    assert np.max(np.abs(res[:, 0] - exact)) < 1e-11 * np.max(np.abs(exact))


def test_agrees_with_regular_transform_at_tensor_points():
    """Band-limited fields, several variables and BCs (R1T0 in z folds the top
    Chebyshev mode): the values at a tensor product of interior points equal the
    regular transform's value slot.  Interior z: at the z ends the two clamp the
    Chebyshev angle differently (1e-14 against 1)."""
    gp, og = _grid(geometry="RLZ", iMin=0.0, iMax=30.0, num_cells=6, kMin=0.0, kMax=10.0, kDim=10,
                   vars={"u": 1, "v": 2}, BCB={"default": S.Chebyshev.R0, "v": S.Chebyshev.R1T0})
    pts = og.gridpoints()
    phys = og.new_physical()
    r, lam, z = pts[:, 0], pts[:, 1], pts[:, 2]
    phys[:, 0, 0] = np.exp(-(r / 12.0) ** 2) * (1 + 0.3 * np.cos(2 * lam) - 0.2 * np.sin(lam)) * (1 + 0.05 * z)
    phys[:, 1, 0] = np.sin(r / 5.0) * (0.5 + np.cos(3 * lam)) * z
    spec = og.spectral_transform(phys)
    rng = np.random.default_rng(4)
    rr = np.sort(rng.uniform(0.0, 30.0, 5))
    ll = rng.uniform(-1.0, 7.0, 6)
    zz = rng.uniform(0.5, 9.5, 4)
    reg = og.regular_transform(spec, rr, ll, zz)
    mat = np.column_stack([np.repeat(rr, 24), np.tile(np.repeat(ll, 4), 5), np.tile(zz, 30)])
    res = og.evaluate_unstructured(spec, mat, threads=4)
    for v in range(2):
        rel = np.linalg.norm(res[:, v] - reg[:, v, 0]) / np.linalg.norm(reg[:, v, 0])
        assert rel < 1e-12, (v, rel)

    gp, og = _grid(geometry="RZ", iMin=0.0, iMax=40.0, num_cells=8, kMin=0.0, kMax=20.0, kDim=10,
                   vars={"u": 1}, BCB={"u": S.Chebyshev.R1T0})
    pts = og.gridpoints()
    phys = og.new_physical()
    phys[:, 0, 0] = np.cos(pts[:, 0] / 9.0) * np.sin(pts[:, 1] / 6.0)
    spec = og.spectral_transform(phys)
    xx = np.linspace(0.0, 40.0, 7)
    zz = np.linspace(1.0, 19.0, 5)
    reg = og.regular_transform(spec, xx, None, zz)
    res = og.evaluate_unstructured(spec, np.column_stack([np.repeat(xx, 5), np.tile(zz, 7)]))
    assert np.linalg.norm(res[:, 0] - reg[:, 0, 0]) / np.linalg.norm(reg[:, 0, 0]) < 1e-12
