"""GPU: R3X sides and their background coefficients (ssg_plan_set_ahat) against the oracle.

The oracle's R3X branch is pinned in test_r3x_pins.py.  Here the library's
gridTransform with per-line background coefficients matches the oracle at the
FP64 bar on RL, RZ and RLZ grids (and the FP32 form at its bar); without
coefficients an R3X grid is bitwise the R3 grid; clearing them restores the
plain results; the spectral transform never uses them; the evaluation paths and
the batching form decline such grids.
"""
import numpy as np
import pytest

import oracle as O
import springsteel_hip as S
from springsteel_hip import capi
from springsteel_hip.errors import NotSupportedError
from springsteel_hip.metrics import rel_l2_slots
from parity_util import TOL_FP64, seeded_field, sensitivity, slot_tolerances

pytestmark = pytest.mark.gpu

CB = S.CubicBSpline


def _params(geometry, bc=CB.R3X, **kw):
    base = dict(geometry=geometry, num_cells=10, iMin=0.0, iMax=50.0, vars={"u": 1, "w": 2},
                BCL={"default": CB.R0, "u": bc}, BCR={"default": CB.R0, "u": bc})
    if geometry != "RL":
        base.update(kMin=0.0, kMax=10.0, kDim=12)
    base.update(kw)
    return S.SpringsteelGridParameters(**base)


def _spectral(g, seed):
    pts = S.getGridpoints(g)
    g.physical[:, :, 0] = seeded_field(pts, g.resolved.nvars, seed, g.geometry)
    S.spectralTransform_(g)
    return np.asfortranarray(g.spectral.copy())


@pytest.mark.parametrize("geometry", ["RL", "RZ", "RLZ"])
def test_grid_transform_with_ahat_vs_oracle(geometry):
    g = S.createGrid(_params(geometry))
    og = O.OracleGrid(g.resolved)
    spec = _spectral(g, 1)
    rng = np.random.default_rng(2)
    ahat = np.asfortranarray(rng.normal(size=spec.shape) * 0.3)
    S.set_ahat(g, ahat)
    og.set_ahat(ahat)
    phys = np.zeros_like(g.physical, order="F")
    S.gridTransform(g, phys, spec)
    ref = og.grid_transform(spec, threads=8)
    # the input that produced spec, for the rounding sensitivity of the tolerance
    inp = og.new_physical()
    inp[:, :, 0] = g.physical[:, :, 0]
    tols = slot_tolerances(sensitivity(og, inp, ref, g.geometry, g.params), TOL_FP64)
    for v in range(g.resolved.nvars):
        errs = rel_l2_slots(phys[:, v, :], ref[:, v, :], g.geometry, g.params)
        print(f"\n{geometry} R3X var {v}: {[f'{e:.1e}' for e in errs]}")
        for s in range(phys.shape[2]):
            assert errs[s] <= max(tols[v][s], TOL_FP64), (v, s, errs, tols[v])
    # w (R0 sides) ignores ahat; u moved
    S.set_ahat(g, None)
    plain = np.zeros_like(phys, order="F")
    S.gridTransform(g, plain, spec)
    assert np.array_equal(plain[:, 1, :], phys[:, 1, :])
    assert not np.allclose(plain[:, 0, 0], phys[:, 0, 0])
    g.close()


@pytest.mark.parametrize("geometry", ["RL", "RLZ"])
def test_r3x_without_ahat_is_r3_bitwise(geometry):
    gx = S.createGrid(_params(geometry, CB.R3X))
    g3 = S.createGrid(_params(geometry, CB.R3))
    spec = _spectral(g3, 3)
    px = np.zeros_like(gx.physical, order="F")
    p3 = np.zeros_like(g3.physical, order="F")
    S.gridTransform(gx, px, spec)
    S.gridTransform(g3, p3, spec)
    assert np.array_equal(px, p3)
    # a zero ahat goes through the btilde stage and still gives the same bytes
    S.set_ahat(gx, np.zeros((gx.nspec, gx.resolved.nvars), order="F"))
    S.gridTransform(gx, px, spec)
    assert np.array_equal(px, p3)
    gx.close()
    g3.close()


def test_fp32_form_with_ahat():
    gp = _params("RLZ")
    g32 = S.createGrid(gp, precision=S.SSG_FP32_REFINE)
    og = O.OracleGrid(g32.resolved)
    g64 = S.createGrid(gp)
    spec = _spectral(g64, 4)
    ahat = np.asfortranarray(np.random.default_rng(5).normal(size=spec.shape) * 0.3)
    S.set_ahat(g32, ahat)
    og.set_ahat(ahat)
    phys = np.zeros_like(g32.physical, order="F")
    S.gridTransform(g32, phys, spec)
    ref = og.grid_transform(spec, threads=8)
    for v in range(g32.resolved.nvars):
        errs = rel_l2_slots(phys[:, v, :].astype(np.float64), ref[:, v, :], g32.geometry, g32.params)
        print(f"\nFP32 RLZ R3X var {v}: {[f'{e:.1e}' for e in errs]}")
        assert max(errs) <= 1e-6
    g32.close()
    g64.close()


def test_device_form_and_paths_that_decline():
    torch = pytest.importorskip("torch")
    g = S.createGrid(_params("RL"))
    spec = _spectral(g, 6)
    ahat = np.asfortranarray(np.random.default_rng(7).normal(size=spec.shape))
    S.set_ahat(g, ahat)
    host = np.zeros_like(g.physical, order="F")
    S.gridTransform(g, host, spec)
    # the device source gives the same coefficients
    S.set_ahat(g, torch.from_numpy(np.ascontiguousarray(ahat.T)).cuda())
    dspec = torch.from_numpy(np.ascontiguousarray(spec.T)).cuda()
    dphys = torch.zeros((g.nslots, g.resolved.nvars, g.npts), dtype=torch.float64, device="cuda")
    S.gridTransform_device(g, dspec, dphys)
    torch.cuda.synchronize()
    assert np.array_equal(dphys.cpu().numpy().transpose(2, 1, 0), host)
    with pytest.raises(NotSupportedError):
        S.regularGridTransform(g, np.array([10.0]), np.array([0.0]))
    with pytest.raises(NotSupportedError):
        S.evaluate_unstructured(g, np.array([[10.0, 0.0]]))
    with pytest.raises(NotSupportedError):
        capi.Batch([g], [dphys], [dspec])
    S.set_ahat(g, None)
    S.regularGridTransform(g, np.array([10.0]), np.array([0.0]))
    g.close()
