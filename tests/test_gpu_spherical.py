"""GPU: SL and SLZ grids (the RL / RLZ kernels on spherical rings) against the oracle.

Rings of 4 round(sin(theta) max_ri) points: their sizes rise to the equator
and fall again, so ring offsets and caps come from the plan's tables.  Bars as
in test_gpu_parity (FP64) and test_gpu_precision (FP32 form); the oracle's SL /
SLZ restatement is pinned in test_spherical_pins.py.  The batching form runs a
spherical grid next to cylindrical ones, bitwise equal to its single-grid calls.
The regular-grid and unstructured evaluations run the RL / RLZ paths, at their bars.
"""
import math

import numpy as np
import pytest

import springsteel_hip as S
from springsteel_hip import capi
from parity_util import TOL_FP64, errors, run_both, seeded_field, sensitivity, slot_tolerances

pytestmark = pytest.mark.gpu


def _check(gp, seed=0):
    grid, og, spec, ospec, phys, ophys, phys2, inp = run_both(gp, None, seed)
    es = errors(spec, ospec)
    ep = errors(phys, ophys, grid.geometry, grid.params)
    ep2 = errors(phys2, ophys, grid.geometry, grid.params)
    tols = slot_tolerances(sensitivity(og, inp, ophys, grid.geometry, grid.params), TOL_FP64)
    print(f"\n{gp.geometry} nc={gp.num_cells} kDim={gp.kDim} nvars={len(gp.vars)}: spectral {max(es):.2e}, "
          f"slots {[f'{max(e[s] for e in ep):.1e}' for s in range(phys.shape[2])]}")
    assert max(es) <= TOL_FP64, es
    for v in range(len(ep)):
        for s in range(phys.shape[2]):
            assert ep[v][s] <= tols[v][s], (v, s, ep[v], tols[v])
            assert ep2[v][s] <= tols[v][s], (v, s, ep2[v], tols[v])
    grid.close()


def test_sl_reference_grid():
    _check(S.SpringsteelGridParameters(geometry="SL", num_cells=10, iMin=0.0, iMax=math.pi, vars={"u": 1}))


def test_sl_multivar_caps_and_bcs():
    _check(S.SpringsteelGridParameters(geometry="SL", num_cells=16, iMin=0.05, iMax=math.pi - 0.05,
                                       vars={"u": 1, "v": 2, "w": 3}, max_wavenumber={"default": -1, "v": 6},
                                       BCL={"default": S.CubicBSpline.R0, "w": S.CubicBSpline.R1T0},
                                       BCR={"default": S.CubicBSpline.R0, "w": S.CubicBSpline.R1T1}), seed=1)


def test_slz_reference_grid():
    _check(S.SpringsteelGridParameters(geometry="SLZ", num_cells=10, kDim=8, iMin=0.0, iMax=math.pi, kMin=0.0,
                                       kMax=20.0, vars={"u": 1}))


def test_slz_multivar_cap():
    _check(S.SpringsteelGridParameters(geometry="SLZ", num_cells=12, kDim=16, iMin=0.1, iMax=math.pi - 0.1,
                                       kMin=0.0, kMax=10.0, vars={"u": 1, "v": 2}, max_wavenumber={"default": 9},
                                       BCB={"default": S.Chebyshev.R0, "v": S.Chebyshev.R1T0}), seed=2)


def test_slz_larger_grid():
    _check(S.SpringsteelGridParameters(geometry="SLZ", num_cells=60, kDim=24, iMin=0.0, iMax=math.pi, kMin=0.0,
                                       kMax=15000.0, vars={"u": 1, "v": 2, "w": 3}), seed=3)


def test_fp32_form_sl_and_slz():
    from test_gpu_precision import _check as check32
    check32(S.SpringsteelGridParameters(geometry="SL", num_cells=12, iMin=0.0, iMax=math.pi, vars={"u": 1, "v": 2}))
    check32(S.SpringsteelGridParameters(geometry="SLZ", num_cells=10, kDim=12, iMin=0.0, iMax=math.pi, kMin=0.0,
                                        kMax=10.0, vars={"u": 1}))


def test_batch_with_spherical_grid_bitwise():
    torch = pytest.importorskip("torch")
    cases = [dict(geometry="SLZ", num_cells=8, kDim=16, iMin=0.0, iMax=math.pi, kMin=0.0, kMax=10.0,
                  vars={"u": 1, "v": 2}),
             dict(geometry="RL", num_cells=12, iMin=0.0, iMax=120.0, vars={"u": 1}),
             dict(geometry="SL", num_cells=14, iMin=0.0, iMax=math.pi, vars={"a": 1, "b": 2})]
    grids, phys, spec, inputs, ref = [], [], [], [], []
    for n, kw in enumerate(cases):
        g = S.createGrid(S.SpringsteelGridParameters(**kw))
        nv = g.resolved.nvars
        x = torch.from_numpy(np.ascontiguousarray(seeded_field(S.getGridpoints(g), nv, 60 + n, g.geometry).T)).cuda()
        ph = torch.zeros((g.nslots, nv, g.npts), dtype=torch.float64, device="cuda")
        sp = torch.zeros((nv, g.nspec), dtype=torch.float64, device="cuda")
        ph[0].copy_(x)
        S.spectralTransform_device(g, ph, sp)
        S.gridTransform_device(g, sp, ph)
        ref.append((sp.clone(), ph.clone()))
        grids.append(g)
        inputs.append(x)
        phys.append(torch.zeros_like(ph))
        spec.append(torch.zeros_like(sp))
    torch.cuda.synchronize()
    batch = capi.Batch(grids, phys, spec)
    for _ in range(2):
        for ph, x in zip(phys, inputs):
            ph.zero_()
            ph[0].copy_(x)
        for sp in spec:
            sp.zero_()
        torch.cuda.synchronize()
        batch.spectral()
        batch.grid()
        torch.cuda.synchronize()
        for n, ((rs, rp), sp, ph) in enumerate(zip(ref, spec, phys)):
            assert torch.equal(rs, sp), f"grid {n}: spectral differs"
            assert torch.equal(rp, ph), f"grid {n}: physical differs"
    batch.close()
    for g in grids:
        g.close()


@pytest.mark.parametrize("kw", [
    dict(geometry="SL", num_cells=10, iMin=0.0, iMax=math.pi, vars={"u": 1, "v": 2}),
    dict(geometry="SLZ", num_cells=8, kDim=10, iMin=0.05, iMax=math.pi - 0.05, kMin=0.0, kMax=10.0,
         vars={"u": 1}, BCB={"default": S.Chebyshev.R1T0}),
], ids=["SL", "SLZ"])
def test_regular_and_unstructured_vs_oracle(kw):
    """SL / SLZ regularGridTransform (src/transforms_spherical.jl:849-1131) and
    evaluate_unstructured (src/interpolation.jl:1782-1800) run the RL / RLZ paths."""
    from test_gpu_regular import _check as check_regular, _run
    from test_gpu_unstructured import _check as check_unstructured, _points
    gp = S.SpringsteelGridParameters(**kw)
    g, out, ref, axes = _run(gp, seed=4)
    worst = check_regular(g, out, ref, axes)
    og = g._oracle[0]
    pts = _points(g.params, 4000, 5, "RL" if g.geometry == "SL" else "RLZ", oob=0.05)
    _, worst_u = check_unstructured(g, og, pts)
    print(f"\n{g.geometry}: regular worst rel L2 {worst:.2e}, unstructured {worst_u:.2e}")
    g.close()
