"""GPU: the batching form (ssg_batch_*).

A batch of RL and RLZ grids of different sizes, variable counts and Chebyshev
dimensions runs as batched launches replayed from a hipGraph.  Each grid's
spectral and physical results must be bitwise those of its own plan's
single-grid calls (the same kernel bodies with the same launch shapes), on
every replay with eager work in between, and again after the tuner has
changed the launch shapes (every candidate shape keeps the per-element
arithmetic order).  The single-grid path itself is checked against the oracle
by test_gpu_parity / test_gpu_precision; here one grid is checked against the
oracle directly as well.
"""
import numpy as np
import pytest

import oracle as O
import springsteel_hip as S
from springsteel_hip import capi
from springsteel_hip.metrics import rel_l2_slots
from parity_util import TOL_FP64, seeded_field, sensitivity, slot_tolerances

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

CASES = [
    dict(geometry="RL", num_cells=12, iMin=0.0, iMax=120.0, vars={"u": 1}),
    dict(geometry="RLZ", num_cells=8, iMin=0.0, iMax=80.0, kMin=0.0, kMax=10.0, kDim=16, vars={"u": 1, "v": 2}),
    dict(geometry="RL", num_cells=20, iMin=0.0, iMax=50.0, vars={"a": 1, "b": 2, "c": 3}),
    dict(geometry="RLZ", num_cells=5, iMin=0.0, iMax=30.0, kMin=0.0, kMax=5.0, kDim=32, vars={"u": 1}),
    dict(geometry="RLZ", num_cells=10, iMin=0.0, iMax=60.0, kMin=0.0, kMax=8.0, kDim=8,
         vars={"u": 1, "v": 2, "w": 3}),
    dict(geometry="RLZ", num_cells=4, iMin=0.0, iMax=40.0, kMin=0.0, kMax=4.0, kDim=64, vars={"u": 1, "v": 2}),
]


def _dtype(g):
    """physical element type of the grid's form (float32 for the FP32 form)"""
    return torch.float32 if g.precision == S.SSG_FP32_REFINE else torch.float64


def _setup(precision):
    grids, phys, spec, inputs = [], [], [], []
    for n, kw in enumerate(CASES):
        g = S.createGrid(S.SpringsteelGridParameters(**kw), precision=precision)
        nv = g.resolved.nvars
        f = seeded_field(S.getGridpoints(g), nv, 40 + n, g.geometry)
        x = torch.from_numpy(np.ascontiguousarray(f.T)).cuda().to(_dtype(g))
        grids.append(g)
        inputs.append(x)
        phys.append(torch.zeros((g.nslots, nv, g.npts), dtype=_dtype(g), device="cuda"))
        spec.append(torch.zeros((nv, g.nspec), dtype=torch.float64, device="cuda"))
    return grids, phys, spec, inputs


def _reference(grids, inputs):
    """single-grid calls on private arrays"""
    out = []
    for g, x in zip(grids, inputs):
        nv = g.resolved.nvars
        ph = torch.zeros((g.nslots, nv, g.npts), dtype=_dtype(g), device="cuda")
        sp = torch.zeros((nv, g.nspec), dtype=torch.float64, device="cuda")
        ph[0].copy_(x)
        S.spectralTransform_device(g, ph, sp)
        S.gridTransform_device(g, sp, ph)
        out.append((sp.clone(), ph.clone()))
    torch.cuda.synchronize()
    return out


def _run_batch(batch, phys, spec, inputs):
    for ph, x in zip(phys, inputs):
        ph.zero_()
        ph[0].copy_(x)
    for sp in spec:
        sp.zero_()
    torch.cuda.synchronize()
    batch.spectral()
    batch.grid()
    torch.cuda.synchronize()


def _assert_same(ref, phys, spec):
    for n, ((rs, rp), sp, ph) in enumerate(zip(ref, spec, phys)):
        assert torch.equal(rs, sp), f"grid {n}: spectral differs"
        assert torch.equal(rp, ph), f"grid {n}: physical differs"


@pytest.mark.parametrize("precision", [S.SSG_FP64, S.SSG_FP32_REFINE], ids=["fp64", "fp32"])
def test_batch_bitwise_single_grid_and_replays(precision):
    grids, phys, spec, inputs = _setup(precision)
    ref = _reference(grids, inputs)
    batch = capi.Batch(grids, phys, spec)
    ns, ng = batch.launches()
    print(f"\nbatch of {len(grids)} grids: {ns} spectral and {ng} grid launches per graph")
    assert ns <= 8 and ng <= 8
    for rep in range(3):
        _run_batch(batch, phys, spec, inputs)
        _assert_same(ref, phys, spec)
        # eager work between replays
        junk = torch.randn(1 << 20, device="cuda", dtype=torch.float64)
        (junk * 2.0).sum().item()
    batch.close()


def test_batch_tuned_shapes_keep_results():
    grids, phys, spec, inputs = _setup(S.SSG_FP64)
    ref = _reference(grids, inputs)
    batch = capi.Batch(grids, phys, spec)
    batch.tune()
    torch.cuda.synchronize()
    tunes = [capi.plan_tuning(g.plan) for g in grids]
    print("\ntuned shapes:", tunes)
    # the chirp-z ring transforms have no launch shape to tune: their choices stay 0 (built in)
    assert all(t["syn_ct"] == 0 and t["ana_mr"] == 0 and t["si_seg"] in (1, 2, 4, 8) for t in tunes)
    for rep in range(2):
        _run_batch(batch, phys, spec, inputs)
        _assert_same(ref, phys, spec)
    batch.close()


def test_batch_grid_against_oracle():
    grids, phys, spec, inputs = _setup(S.SSG_FP64)
    batch = capi.Batch(grids, phys, spec)
    _run_batch(batch, phys, spec, inputs)
    g = grids[1]
    og = O.OracleGrid(g.resolved)
    hp = og.new_physical()
    hp[:, :, 0] = inputs[1].cpu().numpy().T
    ospec = og.spectral_transform(hp, threads=8)
    ophys = og.grid_transform(ospec, threads=8)
    gphys = phys[1].cpu().numpy()
    tols = slot_tolerances(sensitivity(og, hp, ophys, g.geometry, g.params))
    for v in range(g.resolved.nvars):
        assert np.linalg.norm(spec[1][v].cpu().numpy() - ospec[:, v]) <= TOL_FP64 * np.linalg.norm(ospec[:, v])
        errs = rel_l2_slots(gphys[:, v, :].T, ophys[:, v, :], g.geometry, g.params)
        assert all(e <= t for e, t in zip(errs, tols[v])), (errs, tols[v])
    batch.close()
